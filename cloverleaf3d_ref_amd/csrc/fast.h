// Fused kernels of the resident driver's fast path.
//
// Every value they write is computed with the same FP64 expression, in the
// same order, as the per-call kernels (kernels.h), so a fast run reproduces a
// per-call run bit for bit; they only move fewer bytes:
//  * the mesh is uniform (initialise_chunk_kernel.f90:63-178 gives
//    celldx = dx, xarea = celldy*celldz, volume = dx*dy*dz everywhere), so
//    geometry comes in as scalars instead of arrays;
//  * kernels that run back to back on the same cells are fused;
//  * values the reference writes but never reads are not stored (the
//    predictor's density1/energy1 are overwritten by the corrector before
//    anything reads them, PdV.f90:103-107 + hydro.f90:58; the predictor's
//    soundspeed is recomputed before calc_dt reads it);
//  * intermediates are recomputed where that is cheaper than storing them.
#pragma once

#include <hip/hip_runtime.h>

#include "layout.h"

namespace clover {

// uniform mesh, as initialise_chunk_kernel.f90 computes it
struct Geo {
    double dx, dy, dz;   // celld*, vertexd*
    double vol;          // dx*dy*dz
    double xa, ya, za;   // celldy*celldz, celldx*celldz, celldx*celldy
};

// viscosity_kernel + calc_dt_kernel on the same interior cells: writes
// viscosity, and a block-partial minimum of the cell dt into partial[];
// *nparts receives the number of partials (<= kMaxPartials).
hipError_t launch_fast_visc_dt(const Lay &g, const Geo &m, double dtc_safe, double dtu_safe, double dtv_safe,
                               double dtw_safe, double dtdiv_safe, double g_small, const double *density0,
                               const double *pressure, const double *soundspeed, double *viscosity,
                               const double *xvel0, const double *yvel0, const double *zvel0, double *partial,
                               int *nparts, hipStream_t st);

// PdV predictor (PdV_kernel.f90:76-117) followed by ideal_gas on the
// predicted density/energy (ideal_gas.f90:53-62): writes pressure only.
hipError_t launch_fast_pdv_predict(const Lay &g, const Geo &m, double dt, const double *density0,
                                   const double *energy0, double *pressure, const double *viscosity,
                                   const double *xvel0, const double *yvel0, const double *zvel0, hipStream_t st);

hipError_t launch_fast_accelerate(const Lay &g, const Geo &m, double dt, const double *density0,
                                  const double *pressure, const double *viscosity, const double *xvel0,
                                  const double *yvel0, const double *zvel0, double *xvel1, double *yvel1,
                                  double *zvel1, hipStream_t st);

// PdV corrector (PdV_kernel.f90:124-165) and flux_calc (flux_calc_kernel.f90)
hipError_t launch_fast_pdv_correct_flux(const Lay &g, const Geo &m, double dt, const double *density0,
                                        double *density1, const double *energy0, double *energy1,
                                        const double *pressure, const double *viscosity, const double *xvel0,
                                        const double *xvel1, const double *yvel0, const double *yvel1,
                                        const double *zvel0, const double *zvel1, double *vol_flux_x,
                                        double *vol_flux_y, double *vol_flux_z, hipStream_t st);

// one advec_cell sweep, fluxes and cell update in one kernel
// (cell_kernels.hip): mass_flux_<dir> is written over the reference's face
// range, the new density1/energy1 of interior cells go to rho_out/e_out.
hipError_t launch_fast_advec_cell(const Lay &g, const Geo &m, int advect_x, int dir, int sweep_number,
                                  const double *density1, const double *energy1, double *mass_flux_x,
                                  const double *vol_flux_x, double *mass_flux_y, const double *vol_flux_y,
                                  double *mass_flux_z, const double *vol_flux_z, double *rho_out, double *e_out,
                                  hipStream_t st);

// one advec_mom sweep for all three velocities (advec_mom_driver.f90 calls
// the kernel for xvel1, yvel1, zvel1 with the node quantities of the first
// call), node flux and masses kept on chip (mom_kernels.hip).  The new
// velocities go to vel_out[3]: they read their neighbours' old values, so
// they cannot be updated in place.  Only nodes min..max+1 of vel_out are
// written.
hipError_t launch_fast_advec_mom(const Lay &g, const Geo &m, int advect_x, int sweep_number, int direction,
                                 const double *mass_flux_x, const double *vol_flux_x, const double *mass_flux_y,
                                 const double *vol_flux_y, const double *mass_flux_z, const double *vol_flux_z,
                                 const double *density1, const double *const *vel_in, double *const *vel_out,
                                 hipStream_t st);

hipError_t launch_fast_field_summary(const Lay &g, const Geo &m, const double *density0, const double *energy0,
                                     const double *pressure, const double *xvel0, const double *yvel0,
                                     const double *zvel0, double *partial, int *nparts, hipStream_t st);

}  // namespace clover
