// Internal pieces shared by the resident drivers (chunk.cpp for one deck,
// ensemble.cpp for many): array ids beyond the halo fields, the tile record,
// the tile decomposition and the host-side mesh/state fill.
#pragma once

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

}  // namespace clover
