// Mesh and initial states on the device: initialise_chunk_kernel.f90:63-178
// (volume and areas) and generate_chunk_kernel.f90:82-207 (the states painted
// in order, later states over earlier ones), for one tile, straight into the
// device layout.
//
// The reference paints state by state; here each element takes its final
// value directly.  A cell gets the last state whose region contains it (the
// background if none).  A vertex gets the velocity of the last state that
// contains any painted cell around it, because generate_chunk_kernel.f90:
// 164-202 writes the state's velocity to the 8 vertices of each cell it
// paints (vertices outside the background's range min-2..max+2 stay 0
// unless a state paints them).  Vertex and cell coordinates use the expressions of
// initialise_chunk_kernel.f90:63-95, so every value equals the host's.
#include <hip/hip_runtime.h>

#include "launch.h"
#include "setup.h"

namespace clover {

struct GenArgs {
    int ns;  // states, the background first
    double xmin, ymin, zmin, dx, dy, dz;
    double density[kGenStates], energy[kGenStates], xvel[kGenStates], yvel[kGenStates], zvel[kGenStates];
    double sxmin[kGenStates], sxmax[kGenStates], symin[kGenStates], symax[kGenStates], szmin[kGenStates],
        szmax[kGenStates], radius[kGenStates];
    int geometry[kGenStates];
};

// vertexx(j) (initialise_chunk_kernel.f90:63-66), j the tile-local index
__device__ inline double vtx(double lo, double d, int j) { return lo + d * (double)(float)(j - 1); }

// does state s cover cell (j, k, l)?  generate_chunk_kernel.f90:99-162
__device__ inline bool state_hits(const GenArgs &a, int s, int j, int k, int l)
{
    const int geom = a.geometry[s];
    if (geom == 1) {  // cuboid: the cell's vertex bounds against the box
        return vtx(a.xmin, a.dx, j + 1) >= a.sxmin[s] && vtx(a.xmin, a.dx, j) < a.sxmax[s] &&
               vtx(a.ymin, a.dy, k + 1) >= a.symin[s] && vtx(a.ymin, a.dy, k) < a.symax[s] &&
               vtx(a.zmin, a.dz, l + 1) >= a.szmin[s] && vtx(a.zmin, a.dz, l) < a.szmax[s];
    }
    if (geom == 2) {  // sphere about (xmin, ymin, zmin) through cell centres
        const double cx = 0.5 * (vtx(a.xmin, a.dx, j) + vtx(a.xmin, a.dx, j + 1));
        const double cy = 0.5 * (vtx(a.ymin, a.dy, k) + vtx(a.ymin, a.dy, k + 1));
        const double cz = 0.5 * (vtx(a.zmin, a.dz, l) + vtx(a.zmin, a.dz, l + 1));
        const double r = sqrt((cx - a.sxmin[s]) * (cx - a.sxmin[s]) + (cy - a.symin[s]) * (cy - a.symin[s]) +
                              (cz - a.szmin[s]) * (cz - a.szmin[s]));
        return r <= a.radius[s];
    }
    if (geom == 3)  // point: the cell's lower vertex
        return vtx(a.xmin, a.dx, j) == a.sxmin[s] && vtx(a.ymin, a.dy, k) == a.symin[s] &&
               vtx(a.zmin, a.dz, l) == a.szmin[s];
    return false;
}

template <typename S>
__global__ void k_generate(Lay g, GenArgs a, S *__restrict__ density0, S *__restrict__ energy0,
                           S *__restrict__ xvel0, S *__restrict__ yvel0, S *__restrict__ zvel0,
                           S *__restrict__ volume, S *__restrict__ xarea, S *__restrict__ yarea,
                           S *__restrict__ zarea)
{
    // every vertex position (j, k, l) in [min-2, max+3]^3
    const int j = g.x_min - 2 + (int)(blockIdx.x * BX + threadIdx.x);
    const int k = g.y_min - 2 + (int)(blockIdx.y * BY + threadIdx.y);
    const int l = g.z_min - 2 + (int)blockIdx.z;
    if (j > g.x_max + 3 || k > g.y_max + 3) return;
    const int i = idx(g, j, k, l);
    const bool cell = j <= g.x_max + 2 && k <= g.y_max + 2 && l <= g.z_max + 2;
    if (cell) {
        int sc = 0;
        for (int s = a.ns - 1; s >= 1; s--)
            if (state_hits(a, s, j, k, l)) {
                sc = s;
                break;
            }
        density0[i] = a.density[sc];
        energy0[i] = a.energy[sc];
        // initialise_chunk_kernel.f90:151-178: celld* are dx, dy, dz (drivers
        // that hold no geometry arrays pass null)
        if (volume) volume[i] = a.dx * a.dy * a.dz;
        if (xarea) xarea[i] = a.dy * a.dz;
        if (yarea) yarea[i] = a.dx * a.dz;
        if (zarea) zarea[i] = a.dx * a.dy;
    }
    int sv = 0;
    for (int s = a.ns - 1; s >= 1 && !sv; s--)
        for (int dl = -1; dl <= 0 && !sv; dl++)
            for (int dk = -1; dk <= 0 && !sv; dk++)
                for (int dj = -1; dj <= 0 && !sv; dj++) {
                    const int cj = j + dj, ck = k + dk, cl = l + dl;
                    // painted cells: min-2 .. max+2
                    if (cj < g.x_min - 2 || ck < g.y_min - 2 || cl < g.z_min - 2 || cj > g.x_max + 2 ||
                        ck > g.y_max + 2 || cl > g.z_max + 2)
                        continue;
                    if (state_hits(a, s, cj, ck, cl)) sv = s;
                }
    // the background loop covers vertices min-2..max+2 only
    // (generate_chunk_kernel.f90:82-97); the +3 planes keep build_field's 0
    const bool bg = j <= g.x_max + 2 && k <= g.y_max + 2 && l <= g.z_max + 2;
    xvel0[i] = sv || bg ? a.xvel[sv] : 0.0;
    yvel0[i] = sv || bg ? a.yvel[sv] : 0.0;
    zvel0[i] = sv || bg ? a.zvel[sv] : 0.0;
}

hipError_t launch_generate(Store s, const Lay &g, const GenState &st_in, void *density0, void *energy0, void *xvel0,
                           void *yvel0, void *zvel0, void *volume, void *xarea, void *yarea, void *zarea,
                           hipStream_t st)
{
    if (st_in.ns < 1 || st_in.ns > kGenStates) return hipErrorInvalidValue;
    GenArgs a{};
    a.ns = st_in.ns;
    a.xmin = st_in.xmin; a.ymin = st_in.ymin; a.zmin = st_in.zmin;
    a.dx = st_in.dx; a.dy = st_in.dy; a.dz = st_in.dz;
    for (int q = 0; q < a.ns; q++) {
        a.density[q] = st_in.density[q]; a.energy[q] = st_in.energy[q];
        a.xvel[q] = st_in.xvel[q]; a.yvel[q] = st_in.yvel[q]; a.zvel[q] = st_in.zvel[q];
        a.sxmin[q] = st_in.sxmin[q]; a.sxmax[q] = st_in.sxmax[q];
        a.symin[q] = st_in.symin[q]; a.symax[q] = st_in.symax[q];
        a.szmin[q] = st_in.szmin[q]; a.szmax[q] = st_in.szmax[q];
        a.radius[q] = st_in.radius[q]; a.geometry[q] = st_in.geometry[q];
    }
    const int nx = g.x_max - g.x_min + 6, ny = g.y_max - g.y_min + 6, nz = g.z_max - g.z_min + 6;  // vertices
    const dim3 grid((nx + BX - 1) / BX, (ny + BY - 1) / BY, nz);
#define L(T)                                                                                                      \
    TIMED_LAUNCH("k_generate", st, k_generate<T>, grid, dim3(BX, BY), 0, st, g, a, static_cast<T *>(density0),    \
                 static_cast<T *>(energy0), static_cast<T *>(xvel0), static_cast<T *>(yvel0),                     \
                 static_cast<T *>(zvel0), static_cast<T *>(volume), static_cast<T *>(xarea),                      \
                 static_cast<T *>(yarea), static_cast<T *>(zarea))
    if (s == ST_F32) L(float);
    else L(double);
#undef L
    return hipGetLastError();
}

}  // namespace clover
