// Internal pieces shared by the resident drivers (chunk.cpp for one deck,
// ensemble.cpp for many): array ids beyond the halo fields, the tile record,
// the tile decomposition and the host-side mesh/state fill.
#pragma once

#include <initializer_list>
#include <vector>

#include "../../include/clover_hip.h"
#include "layout.h"

namespace clover {

// extra field ids of the resident driver beyond the 19 halo fields
enum : int { X_VOLUME = NUM_FIELDS, X_XAREA, X_YAREA, X_ZAREA, X_W1, X_W2, X_W3, X_W4, X_W5, X_W6, X_W7, NUM_ARRAYS };

inline int class_of(int id)
{
    switch (id) {
    case F_XVEL0: case F_XVEL1: case F_YVEL0: case F_YVEL1: case F_ZVEL0: case F_ZVEL1:
        return AC_VERT;
    case F_VOL_FLUX_X: case F_MASS_FLUX_X: case X_XAREA: return AC_XF;
    case F_VOL_FLUX_Y: case F_MASS_FLUX_Y: case X_YAREA: return AC_YF;
    case F_VOL_FLUX_Z: case F_MASS_FLUX_Z: case X_ZAREA: return AC_ZF;
    default: return id >= X_W1 ? AC_VERT : AC_CELL;
    }
}

struct Tile {
    int nx, ny, nz, left, bottom, back;
    int nbr[6];
    Lay g;
    double *f[NUM_ARRAYS];
    // 1-D geometry, device and host
    double *cellx, *celly, *cellz, *celldx, *celldy, *celldz, *vertexdx, *vertexdy, *vertexdz;
    std::vector<double> h_cellx, h_celly, h_cellz;
};
// tile decomposition of one chunk (clover.f90:258-447)
void decompose(const clover_deck &d, std::vector<Tile> &tiles);

// initialise_chunk.f90:37-44 + initialise_chunk_kernel.f90:63-178 and
// generate_chunk_kernel.f90:82-207, into dense Fortran arrays of tile t
struct HostInit {
    std::vector<double> vertexx, vertexy, vertexz, vertexdx, vertexdy, vertexdz;
    std::vector<double> cellx, celly, cellz, celldx, celldy, celldz;
    std::vector<double> volume, xarea, yarea, zarea, density0, energy0, xvel0, yvel0, zvel0;
};
// fields=false fills only the 1-D arrays (the 3-D fields are painted on the
// device, setup.h)
void host_init(const clover_deck &d, const Tile &t, HostInit &h, bool fields = true);

// The halo exchanges of the fused step, with the fields the reference
// refreshes and the axes of them that a fused kernel reads before the
// field changes or is refreshed again; the rest are dead (the fused kernels
// read the same cells as the reference's loop nests, but the reference
// refreshes every flagged field on every face).
//   FH_TIMESTEP   timestep.f90:69-71 (pressure, energy0, density0, xvel0,
//                 yvel0): viscosity reads pressure at the face neighbours,
//                 accelerate reads density0 at the 8 cells around each node
//                 (corners included); energy0 and the velocities are read
//                 at interior cells and nodes only.
//   FH_VISCOSITY  timestep.f90:80-84: accelerate reads it around nodes.
//   FH_PRESSURE   PdV.f90:99-101 after the predictor: accelerate.
//   FH_CELL1      advection.f90 before the first cell sweep (energy1,
//                 density1, vol_flux_x/y/z, depth 2): the sweep reads
//                 density1/energy1 only along its axis (the rows across it
//                 are interior).  The volume fluxes keep their values for
//                 the whole advection and every momentum sweep reads them
//                 around its nodes on all axes: refreshed here, once.
//   FH_MOM1..3    before each momentum sweep (density1, energy1, x/y/zvel1,
//                 mass_flux_x/y/z): the sweep reads density1 and its own
//                 mass flux on all axes, the velocities along its axis only,
//                 and not energy1.  FH_MOM2 also keeps energy1 along the
//                 third sweep's axis: the third cell sweep follows the
//                 second momentum sweep with no exchange of its own.  The
//                 other two mass fluxes are left out altogether: a momentum
//                 sweep reads face corners (e.g. mass_flux_y at y_max+3 and
//                 z_min-1) that an exchange fills from halo cells written by
//                 the exchange before it, and those hold the same values
//                 whether or not the intermediate exchanges ran.
//   FH_CELL2      before the second (y) cell sweep: density1/energy1 along
//                 y, and density1 along z too: the y sweep also forms
//                 mass_flux_y on the two planes above z_max (its face loop
//                 runs to z_max+2, advec_cell_kernel.f90:215-217), and
//                 FH_MOM2's y faces copy those values into corners the y
//                 momentum sweep reads.  The volume fluxes still hold
//                 FH_CELL1's halos, so they are left out of this exchange
//                 altogether (neither refreshed nor, in the halo-check test
//                 mode, overwritten).
//   mirror_axes (axes whose external faces are mirrored on every tile,
//   Lay::mirror): there the kernels read the interior cell or node that a
//   reflective halo of a cell or vertex field would copy instead of the
//   halo (the value the reflection would write there, sign flip included),
//   so those fields' reflections on those axes are left out of every
//   exchange (raxes; the tile copies of internal faces stay, axes):
//   visc_dt and accel_pdv mirror every axis (level-0 fields, viscosity,
//   pressure), the advection sweeps the x axis (level-1 fields; along y
//   and z the address arithmetic cost them more than the exchanges).  The
//   face fields keep all their exchanges: some of their corner halos hold
//   reflections of values the other passes had not yet refreshed, which
//   only the exchange itself reproduces.
enum FusedHalo : int { FH_TIMESTEP, FH_VISCOSITY, FH_PRESSURE, FH_CELL1, FH_MOM1, FH_CELL2, FH_MOM2, FH_MOM3 };

enum : int { AX_X = 1, AX_Y = 2, AX_Z = 4, AX_ALL = 7 };

struct HaloPlan {
    int depth;
    int fields[NUM_FIELDS];  // 1: refreshed by the reference (and not still valid)
    int axes[NUM_FIELDS];    // axis bits the fused kernels read: tile copies
    int raxes[NUM_FIELDS];   // the same for the reflective faces, less the mirrored ones
};

// The exchange `which` as the reference makes it: every field it flags on
// every axis (timestep.f90:68-74 and 82-84, PdV.f90:97-99, advection.f90:
// 50-57, 71-78, 97-104, 118-125, 156-163; timestep's exchange flags xvel0
// and yvel0 but not zvel0).  A chunk with a tile under 3 cells thick along
// some axis takes these instead of the reduced plan below: there a
// depth-2 reflection reads halo cells of the opposite face, whose values
// then depend on every earlier exchange.
inline HaloPlan reference_halo_plan(int which)
{
    HaloPlan p{};
    auto ref = [&](std::initializer_list<int> ids) {
        for (int f : ids) {
            p.fields[f] = 1;
            p.axes[f] = p.raxes[f] = AX_ALL;
        }
    };
    switch (which) {
    case FH_TIMESTEP: ref({F_PRESSURE, F_ENERGY0, F_DENSITY0, F_XVEL0, F_YVEL0}); p.depth = 1; break;
    case FH_VISCOSITY: ref({F_VISCOSITY}); p.depth = 1; break;
    case FH_PRESSURE: ref({F_PRESSURE}); p.depth = 1; break;
    case FH_CELL1:
    case FH_CELL2: ref({F_ENERGY1, F_DENSITY1, F_VOL_FLUX_X, F_VOL_FLUX_Y, F_VOL_FLUX_Z}); p.depth = 2; break;
    default:
        ref({F_DENSITY1, F_ENERGY1, F_XVEL1, F_YVEL1, F_ZVEL1, F_MASS_FLUX_X, F_MASS_FLUX_Y, F_MASS_FLUX_Z});
        p.depth = 2;
        break;
    }
    return p;
}

// the axes along which both faces of a tile are mirrored (Lay::mirror)
inline int mirror_axes(const Lay &g)
{
    return ((g.mirror & 3) == 3 ? AX_X : 0) | ((g.mirror & 12) == 12 ? AX_Y : 0) | ((g.mirror & 48) == 48 ? AX_Z : 0);
}

// mirror_axes: the axes whose faces the fused kernels mirror (their level-0
// fields everywhere, level-1 ones across x); tan_x: the single-tile FP32
// kernels also mirror vol_flux_y/z and mass_flux_y/z across x (launch.h
// tan_xm), so their x reflections are left out too
inline HaloPlan fused_halo_plan(int which, bool advect_x, int mirror_axes = 0, bool tan_x = false)
{
    HaloPlan p{};
    auto ref = [&](std::initializer_list<int> ids) {
        for (int f : ids) p.fields[f] = 1;
    };
    const int s1 = advect_x ? AX_X : AX_Z, s3 = advect_x ? AX_Z : AX_X;
    auto mass_flux = [](int ax) { return ax == AX_X ? F_MASS_FLUX_X : ax == AX_Y ? F_MASS_FLUX_Y : F_MASS_FLUX_Z; };
    const int fc_ids[] = {F_ENERGY1, F_DENSITY1, F_VOL_FLUX_X, F_VOL_FLUX_Y, F_VOL_FLUX_Z};
    auto mom = [&](int s) {
        ref({F_DENSITY1, F_ENERGY1, F_XVEL1, F_YVEL1, F_ZVEL1, mass_flux(s)});
        p.axes[F_DENSITY1] = AX_ALL;
        p.axes[F_XVEL1] = p.axes[F_YVEL1] = p.axes[F_ZVEL1] = s;
        p.axes[mass_flux(s)] = AX_ALL;
    };
    p.depth = 2;
    switch (which) {
    case FH_TIMESTEP:
        ref({F_PRESSURE, F_ENERGY0, F_DENSITY0, F_XVEL0, F_YVEL0});
        p.axes[F_PRESSURE] = p.axes[F_DENSITY0] = AX_ALL;
        p.depth = 1;
        break;
    case FH_VISCOSITY:
        ref({F_VISCOSITY});
        p.axes[F_VISCOSITY] = AX_ALL;
        p.depth = 1;
        break;
    case FH_PRESSURE:
        ref({F_PRESSURE});
        p.axes[F_PRESSURE] = AX_ALL;
        p.depth = 1;
        break;
    case FH_CELL1:
        for (int f : fc_ids) p.fields[f] = 1;
        p.axes[F_DENSITY1] = p.axes[F_ENERGY1] = s1;
        p.axes[F_VOL_FLUX_X] = p.axes[F_VOL_FLUX_Y] = p.axes[F_VOL_FLUX_Z] = AX_ALL;
        break;
    case FH_CELL2:
        ref({F_ENERGY1, F_DENSITY1});
        p.axes[F_DENSITY1] = AX_Y | AX_Z;
        p.axes[F_ENERGY1] = AX_Y;
        break;
    case FH_MOM1: mom(s1); break;
    case FH_MOM2:
        mom(AX_Y);
        p.axes[F_ENERGY1] = s3;
        break;
    default: mom(s3); break;
    }
    for (int f = 0; f < NUM_FIELDS; f++) p.raxes[f] = p.axes[f];
    for (int f : {F_DENSITY0, F_ENERGY0, F_PRESSURE, F_VISCOSITY, F_SOUNDSPEED, F_XVEL0, F_YVEL0, F_ZVEL0})
        p.raxes[f] &= ~mirror_axes;
    for (int f : {F_DENSITY1, F_ENERGY1, F_XVEL1, F_YVEL1, F_ZVEL1}) p.raxes[f] &= ~(mirror_axes & AX_X);
    if (tan_x)
        for (int f : {F_VOL_FLUX_Y, F_VOL_FLUX_Z, F_MASS_FLUX_Y, F_MASS_FLUX_Z}) p.raxes[f] &= ~(mirror_axes & AX_X);
    return p;
}

}  // namespace clover
