// Fused kernels of the resident driver's fast path.
//
// In the FP64 form every value they write is computed with the same
// expression, in the same order, as the per-call kernels (kernels.h), so a
// fast run reproduces a per-call run bit for bit; they only move fewer bytes:
//  * the mesh is uniform (initialise_chunk_kernel.f90:63-178 gives
//    celldx = dx, xarea = celldy*celldz, volume = dx*dy*dz everywhere), so
//    geometry comes in as scalars instead of arrays;
//  * kernels that run back to back on the same cells are fused;
//  * values the reference writes but never reads are not stored (the
//    predictor's density1/energy1 are overwritten by the corrector before
//    anything reads them, PdV.f90:103-107 + hydro.f90:58; the predictor's
//    soundspeed is recomputed before calc_dt reads it);
//  * intermediates are recomputed where that is cheaper than storing them.
//
// Storage.  Each launcher takes the element type of the field arrays:
// ST_F64 (the drop-in form) or ST_F32 (the FP32-store form: fields held and
// stencils computed in FP32; calc_dt's minimum, advec_cell's mass and energy
// update and the summary sums in FP64).  Array arguments are untyped
// pointers to arrays of that element type.
//
// Shape.  The marching kernels take a launch shape (positions per workgroup
// along the march, block order), which changes only how the work is split,
// never a value; the autotuner (chunk.cpp, ensemble.cpp) picks it.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#include "layout.h"

namespace clover {

enum Store : int { ST_F64 = 0, ST_F32 = 1 };

// The kernels with a launch shape, and the shape: `planes` positions per
// workgroup along the kernel's march (planes for the z marches, faces or
// nodes along the sweep for the y/z advection sweeps; for the x cell sweep,
// which has no march, cell_x_rows below) and `xcd`, the XCD-contiguous
// block order (launch.h hw_blk_xcd) instead of the hardware's.
// (the same indices as CLOVER_SHAPE_* of clover_hip.h)
enum ShapeKernel : int {
    SK_VISC_DT, SK_PDV_PREDICT, SK_ACCEL_PDV, SK_CELL_X, SK_CELL_Y, SK_CELL_Z, SK_MOM_X, SK_MOM_Y, SK_MOM_Z,
    SK_COUNT
};
struct Shape {
    int planes;
    int xcd;
    int bx;   // z marches: threads along x per workgroup (64, 128, 256 or 512; see block_of)
    int wgs;  // batched launches: 0 for one workgroup per tile, else this many
              // persistent workgroups walk the tile list (batch.h tile_walk)
};

// Workgroup of a z march for shape sh: bx threads along x and 256/bx rows
// (one row from 256 up), so wider workgroups keep more x neighbours inside
// one workgroup.  The other kernels have fixed workgroups.
constexpr int kMaxBlock = 512;
inline dim3 block_of(const Shape &sh) { return dim3(sh.bx, sh.bx >= 256 ? 1 : 256 / sh.bx, 1); }
constexpr int kTuneBx[] = {64, 128, 256, 512};
struct Shapes {
    Shape k[SK_COUNT];
};
// the autotuner's candidates for Shape::planes (each with both block orders)
constexpr int kTunePlanes[] = {8, 16, 32, 64, 128};
constexpr int kTuneCount = sizeof(kTunePlanes) / sizeof(kTunePlanes[0]);

// rows of the x cell sweep per wave for `planes`: planes/32, at least one
// (the tuner's candidates give 1, 2 and 4)
__host__ __device__ inline int cell_x_rows(int planes) { return planes >= 64 ? planes / 32 : 1; }

// 64 positions per workgroup of 64 x 4 (one row per wave for the x cell
// sweep); the XCD order for the y/z momentum sweeps
inline Shapes default_shapes()
{
    Shapes s;
    for (int i = 0; i < SK_COUNT; i++) s.k[i] = Shape{i == SK_CELL_X ? 32 : 64, i == SK_MOM_Y || i == SK_MOM_Z, 64};
    return s;
}

inline size_t store_bytes(Store s) { return s == ST_F32 ? sizeof(float) : sizeof(double); }

// A positive divisor b that is constant over the mesh, prepared for division
// without a divide instruction (launch.h by_const):
//   hn = RN(1/b);  hd = the largest double <= 1/b (hn, or its predecessor);
//   ld = RN(1/b - hd) >= 0, so hd + ld is 1/b to about 2^-105.
// q = RN(a*hd + RN(a*ld)) is then within one ulp of a/b, e = q*b - a is exact
// (FMA), and RN(q - e*hn) is the correctly rounded quotient RN(a/b): one
// Newton-Raphson correction of a faithful quotient with the correctly rounded
// reciprocal (P. Markstein, "Computation of elementary functions on the IBM
// RISC System/6000 processor", IBM J. Res. Dev. 34(1), 1990).  With ld >= 0 a
// zero a keeps its sign, so the result equals a/b bit for bit for every
// finite a whose products stay in the normal range.
struct Recip {
    double b, hn, hd, ld;
};

inline Recip recip_of(double b)
{
    Recip r;
    r.b = b;
    r.hn = 1.0 / b;
    r.hd = std::fma(-r.hn, b, 1.0) >= 0.0 ? r.hn : std::nextafter(r.hn, 0.0);
    r.ld = std::fma(-r.hd, b, 1.0) / b;  // 1 - b*hd is exact: hd is within one ulp of 1/b
    return r;
}

// uniform mesh, as initialise_chunk_kernel.f90 computes it
struct Geo {
    double dx, dy, dz;   // celld*, vertexd*
    double vol;          // dx*dy*dz
    double xa, ya, za;   // celldy*celldz, celldx*celldz, celldx*celldy
    Recip rx, ry, rz;    // the spacings as divisors
    Recip r6;            // 6, advec_mom_kernel.f90's limiter divisor
};

// initialise_chunk_kernel.f90:151-178 on a uniform mesh of spacing dx, dy, dz
inline Geo geo_of(double dx, double dy, double dz)
{
    Geo m;
    m.dx = dx;
    m.dy = dy;
    m.dz = dz;
    m.vol = dx * dy * dz;
    m.xa = dy * dz;
    m.ya = dx * dz;
    m.za = dx * dy;
    m.rx = recip_of(dx);
    m.ry = recip_of(dy);
    m.rz = recip_of(dz);
    m.r6 = recip_of(6.0);
    return m;
}

// ideal_gas_kernel.f90:45-60 on interior cells
hipError_t launch_fast_ideal_gas(Store s, const Lay &g, const void *density, const void *energy, void *pressure,
                                 void *soundspeed, hipStream_t st);

// viscosity_kernel + calc_dt_kernel on the same interior cells: writes
// viscosity, and a block-partial minimum (FP64) of the cell dt into
// partial[]; *nparts receives the number of partials (<= kMaxDtPartials).
hipError_t launch_fast_visc_dt(Store s, const Lay &g, const Geo &m, double dtc_safe, double dtu_safe,
                               double dtv_safe, double dtw_safe, double dtdiv_safe, double g_small,
                               const void *density0, const void *pressure, void *viscosity, const void *xvel0,
                               const void *yvel0, const void *zvel0, double *partial, int *nparts, const Shape &sh,
                               hipStream_t st);

// PdV predictor (PdV_kernel.f90:76-117) followed by ideal_gas on the
// predicted density/energy (ideal_gas.f90:53-62): writes pressure only.
hipError_t launch_fast_pdv_predict(Store s, const Lay &g, const Geo &m, double dt, const void *density0,
                                   const void *energy0, void *pressure, const void *viscosity, const void *xvel0,
                                   const void *yvel0, const void *zvel0, const Shape &sh, hipStream_t st);

// accelerate (accelerate_kernel.f90:61-109), then the PdV corrector
// (PdV_kernel.f90:124-165) and flux_calc (flux_calc_kernel.f90:56-90) on its
// level-1 velocities, in one z-march: writes xvel1/yvel1/zvel1, density1,
// energy1 and the three volume fluxes
hipError_t launch_fast_accel_pdv(Store s, const Lay &g, const Geo &m, double dt, const void *density0,
                                 void *density1, const void *energy0, void *energy1, const void *pressure,
                                 const void *viscosity, const void *xvel0, void *xvel1, const void *yvel0,
                                 void *yvel1, const void *zvel0, void *zvel1, void *vol_flux_x, void *vol_flux_y,
                                 void *vol_flux_z, const Shape &sh, hipStream_t st);

// one advec_cell sweep, fluxes and cell update in one kernel
// (cell_kernels.hip): mass_flux_<dir> is written over the reference's face
// range, the new density1/energy1 of interior cells go to rho_out/e_out.
// On the step's third sweep, pressure/soundspeed (both or neither) also get
// ideal_gas of the new state: the next step's ideal_gas results.
hipError_t launch_fast_advec_cell(Store s, const Lay &g, const Geo &m, int advect_x, int dir, int sweep_number,
                                  const void *density1, const void *energy1, void *mass_flux_x,
                                  const void *vol_flux_x, void *mass_flux_y, const void *vol_flux_y,
                                  void *mass_flux_z, const void *vol_flux_z, void *rho_out, void *e_out,
                                  void *pressure, const Shape &sh, hipStream_t st);

// one advec_mom sweep for all three velocities (advec_mom_driver.f90 calls
// the kernel for xvel1, yvel1, zvel1 with the node quantities of the first
// call), node flux and masses kept on chip (mom_kernels.hip).  The new
// velocities go to vel_out[3]: they read their neighbours' old values, so
// they cannot be updated in place.  Only nodes min..max+1 of vel_out are
// written.
hipError_t launch_fast_advec_mom(Store s, const Lay &g, const Geo &m, int advect_x, int sweep_number,
                                 int direction, const void *mass_flux_x, const void *vol_flux_x,
                                 const void *mass_flux_y, const void *vol_flux_y, const void *mass_flux_z,
                                 const void *vol_flux_z, const void *density1, const void *const *vel_in,
                                 void *const *vel_out, const Shape &sh, hipStream_t st);

// field_summary_kernel.f90:51-78: five FP64 partials per block.  The FP64
// form sums as the per-call kernel does (same tree); the FP32 form
// accumulates each thread's cells with compensated (Neumaier) sums.
hipError_t launch_fast_field_summary(Store s, const Lay &g, const Geo &m, const void *density0, const void *energy0,
                                     const void *pressure, const void *xvel0, const void *yvel0, const void *zvel0,
                                     double *partial, int *nparts, hipStream_t st);

// Reflective halo and tile copies on arrays of either element type
// (halo_kernels.hip; same tables and order as launch_update_halo and
// launch_tile_halo of kernels.h).  fields[f] == 1 selects what the
// reference refreshes; axes (null: all) narrows each selected field to a bit
// mask of axes (1 x, 2 y, 4 z).  With poison, the selected (field, axis)
// halos outside the mask get wrong values instead (a test of the mask).
hipError_t launch_update_halo_store(Store s, const Lay &g, void *const *field_ptrs, const int *fields,
                                    const int *ext, int depth, hipStream_t st, const int *axes = nullptr,
                                    bool poison = false);
hipError_t launch_tile_halo_store(Store s, int face, const Lay &g, void *const *field_ptrs, const Lay &gn,
                                  void *const *nbr_ptrs, const int *fields, int depth, hipStream_t st,
                                  const int *axes = nullptr, bool poison = false);

}  // namespace clover
