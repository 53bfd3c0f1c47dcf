// Host-side launchers of the per-call hydro kernels.  Each launcher enqueues,
// on `st`, exactly the loop nests of one reference *_kernel call, with the
// reference's iteration domains and FP64 expression order.  Pointers are
// device addresses in the Lay layout (layout.h); 1-D arrays (celldx, ...)
// are dense device arrays indexed from min-2.
#pragma once

#include <hip/hip_runtime.h>

#include "layout.h"

namespace clover {

struct Ptr1D {  // the six 1-D geometry arrays a kernel may need
    const double *celldx, *celldy, *celldz;
    const double *vertexdx, *vertexdy, *vertexdz;
};

hipError_t launch_ideal_gas(const Lay &g, const double *density, const double *energy, double *pressure,
                            double *soundspeed, hipStream_t st);

hipError_t launch_viscosity(const Lay &g, const double *xarea, const double *yarea, const double *zarea,
                            const double *celldx, const double *celldy, const double *celldz,
                            const double *density0, const double *pressure, double *viscosity,
                            const double *xvel0, const double *yvel0, const double *zvel0, hipStream_t st);

// per-cell dt into dt_min (may be null: then only the reduction runs) and a
// block-partial minimum into partial[nblocks]; returns nblocks via *nparts.
hipError_t launch_calc_dt_cells(const Lay &g, double dtc_safe, double dtu_safe, double dtv_safe, double dtw_safe,
                                double dtdiv_safe, double g_small, const double *xarea, const double *yarea,
                                const double *zarea, const double *celldx, const double *celldy,
                                const double *celldz, const double *volume, const double *density0,
                                const double *viscosity_a, const double *soundspeed, const double *xvel0,
                                const double *yvel0, const double *zvel0, double *dt_min, double *partial,
                                int *nparts, hipStream_t st);
// min over partial[0..n) seeded with `seed`, into *out (device scalar)
hipError_t launch_min_final(const double *partial, int n, double seed, double *out, hipStream_t st);

hipError_t launch_pdv(int predict, const Lay &g, double dt, const double *xarea, const double *yarea,
                      const double *zarea, const double *volume, const double *density0, double *density1,
                      const double *energy0, double *energy1, const double *pressure, const double *viscosity,
                      const double *xvel0, const double *xvel1, const double *yvel0, const double *yvel1,
                      const double *zvel0, const double *zvel1, hipStream_t st);

hipError_t launch_revert(const Lay &g, const double *density0, double *density1, const double *energy0,
                         double *energy1, hipStream_t st);

hipError_t launch_accelerate(const Lay &g, double dt, const double *xarea, const double *yarea,
                             const double *zarea, const double *volume, const double *density0,
                             const double *pressure, const double *viscosity, const double *xvel0,
                             const double *yvel0, const double *zvel0, double *xvel1, double *yvel1,
                             double *zvel1, hipStream_t st);

hipError_t launch_flux_calc(const Lay &g, double dt, const double *xarea, const double *yarea,
                            const double *zarea, const double *xvel0, const double *yvel0, const double *zvel0,
                            const double *xvel1, const double *yvel1, const double *zvel1, double *vol_flux_x,
                            double *vol_flux_y, double *vol_flux_z, hipStream_t st);

hipError_t launch_advec_cell(const Lay &g, int advect_x, int dir, int sweep_number, const double *vertexdx,
                             const double *vertexdy, const double *vertexdz, const double *volume,
                             double *density1, double *energy1, double *mass_flux_x, const double *vol_flux_x,
                             double *mass_flux_y, const double *vol_flux_y, double *mass_flux_z,
                             const double *vol_flux_z, double *pre_vol, double *post_vol, double *pre_mass,
                             double *post_mass, double *advec_vol, double *post_ener, double *ener_flux,
                             hipStream_t st);

hipError_t launch_advec_mom(const Lay &g, double *vel1, const double *mass_flux_x, const double *vol_flux_x,
                            const double *mass_flux_y, const double *vol_flux_y, const double *mass_flux_z,
                            const double *vol_flux_z, const double *volume, const double *density1,
                            double *node_flux, double *node_mass_post, double *node_mass_pre, double *mom_flux,
                            double *pre_vol, double *post_vol, const double *celldx, const double *celldy,
                            const double *celldz, int advect_x, int which_vel, int sweep_number, int direction,
                            hipStream_t st);

hipError_t launch_reset_field(const Lay &g, double *density0, const double *density1, double *energy0,
                              const double *energy1, double *xvel0, const double *xvel1, double *yvel0,
                              const double *yvel1, double *zvel0, const double *zvel1, hipStream_t st);

// Reflective boundary on the external faces (update_halo_kernel.f90:115-1425).
// fields[19] are the reference flags; ext[6] = 1 where the face is external
// for both chunk and tile (left, right, bottom, top, back, front).
hipError_t launch_update_halo(const Lay &g, double *const *field_ptrs /* by field id */, const int *fields,
                              const int *ext, int depth, hipStream_t st);

// Tile-to-tile halo copy for one face (update_tile_halo_kernel.f90).
// face: 1 left, 2 right, 3 bottom, 4 top, 5 back, 6 front.
hipError_t launch_tile_halo(int face, const Lay &g, double *const *field_ptrs, const Lay &gn,
                            double *const *nbr_ptrs, const int *fields, int depth, hipStream_t st);

// field_summary partial sums: partial[5*nblocks] (vol, mass, ie, ke, press)
hipError_t launch_field_summary_partial(const Lay &g, const double *volume, const double *density0,
                                        const double *energy0, const double *pressure, const double *xvel0,
                                        const double *yvel0, const double *zvel0, double *partial, int *nparts,
                                        hipStream_t st);
// the same with the vertex velocities in their own layout gv (dense Fortran
// arrays: cell and vertex arrays have different extents)
hipError_t launch_field_summary_partial(const Lay &g, const Lay &gv, const double *volume, const double *density0,
                                        const double *energy0, const double *pressure, const double *xvel0,
                                        const double *yvel0, const double *zvel0, double *partial, int *nparts,
                                        hipStream_t st);
hipError_t launch_sum5_final(const double *partial, int n, double *out, hipStream_t st);

constexpr int kMaxPartials = 16384;
// calc_dt's block minima (one double each) in the same buffers, which hold
// 5 x kMaxPartials doubles for the summary's five sums
constexpr int kMaxDtPartials = 5 * kMaxPartials;

// Optional per-launch timing: while a timer is installed on the calling
// thread, every launcher brackets each kernel with HIP events on its stream
// (no host synchronisation; the owner reads the events later).
struct KernelTimer {
    virtual void begin(const char *kernel, hipStream_t st) = 0;
    virtual void end(hipStream_t st) = 0;
    virtual ~KernelTimer() {}
};
extern thread_local KernelTimer *t_kernel_timer;

}  // namespace clover
