// Mesh and initial states of one tile painted on the device
// (setup_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/clover_hip.h"
#include "fast.h"
#include "layout.h"

namespace clover {

constexpr int kGenStates = CLOVER_MAX_STATES;

// A deck's states and one tile's mesh origin and spacing
// (initialise_chunk.f90:37-44).
struct GenState {
    int ns;
    double xmin, ymin, zmin, dx, dy, dz;
    double density[kGenStates], energy[kGenStates], xvel[kGenStates], yvel[kGenStates], zvel[kGenStates];
    double sxmin[kGenStates], sxmax[kGenStates], symin[kGenStates], symax[kGenStates], szmin[kGenStates],
        szmax[kGenStates], radius[kGenStates];
    int geometry[kGenStates];
};

// tile offsets left/bottom/back are the global index of its first cell
inline GenState gen_state(const clover_deck &d, int left, int bottom, int back)
{
    GenState s{};
    s.ns = d.number_of_states;
    s.dx = (d.xmax - d.xmin) / (double)(float)d.x_cells;
    s.dy = (d.ymax - d.ymin) / (double)(float)d.y_cells;
    s.dz = (d.zmax - d.zmin) / (double)(float)d.z_cells;
    s.xmin = d.xmin + s.dx * (double)(float)(left - 1);
    s.ymin = d.ymin + s.dy * (double)(float)(bottom - 1);
    s.zmin = d.zmin + s.dz * (double)(float)(back - 1);
    for (int q = 0; q < s.ns && q < kGenStates; q++) {
        s.density[q] = d.state_density[q];
        s.energy[q] = d.state_energy[q];
        s.xvel[q] = d.state_xvel[q];
        s.yvel[q] = d.state_yvel[q];
        s.zvel[q] = d.state_zvel[q];
        s.sxmin[q] = d.state_xmin[q];
        s.sxmax[q] = d.state_xmax[q];
        s.symin[q] = d.state_ymin[q];
        s.symax[q] = d.state_ymax[q];
        s.szmin[q] = d.state_zmin[q];
        s.szmax[q] = d.state_zmax[q];
        s.radius[q] = d.state_radius[q];
        s.geometry[q] = d.state_geometry[q];
    }
    return s;
}

// density0, energy0 and the velocities of every cell/vertex of the tile,
// and volume and areas (cell range; the +3 planes are left as they are;
// null pointers are skipped)
hipError_t launch_generate(Store s, const Lay &g, const GenState &st_in, void *density0, void *energy0, void *xvel0,
                           void *yvel0, void *zvel0, void *volume, void *xarea, void *yarea, void *zarea,
                           hipStream_t st);

}  // namespace clover
