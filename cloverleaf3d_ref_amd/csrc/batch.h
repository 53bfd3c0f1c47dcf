// Batched launches over many independent decks (the ensemble driver).
//
// Every deck is one tile with its own extents, mesh spacing, dt and time.
// One launch of a batched kernel covers all decks: deck d owns the
// workgroups start[d] .. start[d+1]-1 of the launch (a prefix table built on
// the host from the same per-deck grid the single-deck launcher uses), a
// workgroup finds its deck by binary search and runs the single-deck kernel
// body on that deck's block.  Array pointers, dt and time are read from the
// device-side deck table, so a whole step needs no host involvement and can
// be captured in a graph; the level and sweep exchanges of the fast path are
// done on the device by exchanging the table's pointers.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "driver.h"
#include "fast.h"
#include "launch.h"

namespace clover {

struct DeckDev {
    Lay g;
    Geo m;
    void *f[NUM_ARRAYS];  // field arrays, element type of the ensemble
    double dt, dtold, time;
    double dtmax, dtrise, dtmin;
    double dtc_safe, dtu_safe, dtv_safe, dtw_safe, dtdiv_safe;
    int steps;  // timesteps taken (dt formed)
    int small;  // 0, or the step at which dt first fell below dtmin (timestep.f90:120-122)
};

// A batched launch: total workgroups and the device prefix table (nd+1).
struct BatchGrid {
    const int *start;
    int total;
};

// Workgroups one deck needs in each batched kernel (same shapes as the
// single-deck launches).
enum BatchKernel : int {
    BK_IDEAL_GAS, BK_VISC_DT, BK_PDV_PREDICT, BK_ACCEL_PDV, BK_SUMMARY,
    BK_CELL_X, BK_CELL_Y, BK_CELL_Z, BK_MOM_X, BK_MOM_Y, BK_MOM_Z, BK_COUNT
};
int batch_blocks(BatchKernel k, const Lay &g, int planes, Store s);
// reflective halo along `axis` for `nitems` (field, face) items
int batch_blocks_reflect(const Lay &g, int axis, int depth, int nitems);

// The tiles (workgroup-sized blocks of the decks' grids, numbered through
// the prefix table) that this workgroup runs: t, t+step, ... below end.
// With a workgroup per tile that is one tile, in the hardware's or the
// XCD-contiguous order.  A persistent launch (fewer workgroups than tiles)
// walks the list: with xcd, the workgroups on XCD x (b mod 8: the hardware
// deals workgroups round-robin) share the x-th contiguous eighth of the
// tiles, so neighbouring tiles keep meeting in one L2.
__device__ inline void tile_walk(int total, int xcd, int &t, int &end, int &step)
{
    const int b = (int)blockIdx.x, G = (int)gridDim.x;
    if (G >= total) {
        t = xcd ? xcd_block(b, G) : b;
        end = t + 1;
        step = 1;
    } else if (!xcd || G < kXcds) {
        t = b;
        end = total;
        step = G;
    } else {
        const int x = b % kXcds;
        t = (int)((long long)total * x / kXcds) + b / kXcds;
        end = (int)((long long)total * (x + 1) / kXcds);
        step = (G - x + kXcds - 1) / kXcds;  // workgroups on XCD x
    }
}

// workgroups of a batched launch of bg's tiles with shape sh
inline unsigned batch_wgs(const BatchGrid &bg, const Shape &sh)
{
    return (unsigned)(sh.wgs > 0 && sh.wgs < bg.total ? sh.wgs : bg.total);
}

__device__ inline int deck_of(const int *__restrict__ start, int nd, int b)
{
    int lo = 0, hi = nd - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Launchers.  decks: device table; grid from the matching prefix table.
hipError_t batch_ideal_gas(Store s, const DeckDev *decks, int nd, const BatchGrid &bg, hipStream_t st);
// writes one dt partial per workgroup at partial[workgroup]
hipError_t batch_visc_dt(Store s, const DeckDev *decks, int nd, const BatchGrid &bg, const Shape &sh, double g_small,
                         double *partial, hipStream_t st);
hipError_t batch_pdv_predict(Store s, const DeckDev *decks, int nd, const BatchGrid &bg, const Shape &sh, hipStream_t st);
// accelerate + PdV corrector + flux_calc (launch_fast_accel_pdv)
hipError_t batch_accel_pdv(Store s, const DeckDev *decks, int nd, const BatchGrid &bg, const Shape &sh, hipStream_t st);
// five partials per workgroup at partial[5*workgroup]
hipError_t batch_field_summary(Store s, const DeckDev *decks, int nd, const BatchGrid &bg, double *partial,
                               hipStream_t st);
// advec_cell sweep `dir`; new density1/energy1 into the arrays X_W1/X_W2,
// and with gas (third sweep only) ideal_gas's pressure of them into pressure
hipError_t batch_advec_cell(Store s, const DeckDev *decks, int nd, const BatchGrid &bg, int advect_x, int dir,
                            int sweep_number, bool gas, const Shape &sh, hipStream_t st);
// advec_mom sweep `dir`; new velocities into X_W5..X_W7
hipError_t batch_advec_mom(Store s, const DeckDev *decks, int nd, const BatchGrid &bg, int advect_x, int dir,
                           int sweep_number, const Shape &sh, hipStream_t st);
// reflective halo of the flagged fields along one axis, both faces; axes
// (null: all) as for launch_update_halo_store
hipError_t batch_reflect(Store s, const DeckDev *decks, int nd, const BatchGrid &bg, int axis, int depth,
                         const int *fields, hipStream_t st, const int *axes = nullptr);

// The tiles of one chunk as a batch (chunk.cpp, more than one tile): one
// launch per pass of update_halo over every tile.  An item is one (tile,
// face, field): a copy from tile `nbr` (update_tile_halo_kernel.f90), or the
// reflective halo of an external face (update_halo_kernel.f90, nbr = tile).
struct TileHaloItem {
    int tile, nbr, field, hi, kind, ea, eb, poison;
    double sgn;
};
// blocks one item needs along the face's two tangential axes for tile g
void tile_halo_blocks(const Lay &g, int axis, int depth, int &per_a, int &per_b);
// copy: the tile copies of one pass; else reflective halos along one axis
// (items: device array; per_a x per_b: the largest tile's blocks)
hipError_t batch_tile_halo(Store s, bool copy, const DeckDev *tiles, const TileHaloItem *items, int nitems, int per_a,
                           int per_b, int axis, int depth, hipStream_t st);
// the item lists of one update_halo over the tiles: copy[pass] for the three
// tile passes (y, x, z faces), refl[k][side] for the reflective faces of the
// k-th axis in y, x, z order; nbr6: 6 neighbours per tile (from 1, -1
// external)
void tile_halo_items(const int *nbr6, int ntiles, const Lay *lays, const int *fields, int depth, const int *axes,
                     const int *raxes,
                     bool poison, std::vector<TileHaloItem> (&copy)[3], std::vector<TileHaloItem> (&refl)[3][2]);

// Per-tile minimum of a batched calc_dt's partials (tile t's partials at
// start[t]..start[t+1]-1), seeded with g_big, into out[stride * t]
hipError_t batch_tile_min(int ntiles, const int *start, const double *partial, double g_big, double *out, int stride,
                          hipStream_t st);
// dt of every table entry
hipError_t batch_set_dt(DeckDev *decks, int nd, double dt, hipStream_t st);

// Per-deck steps (one workgroup per deck).
// timestep.f90:86-126 on the device: dt = min over the deck's partials,
// then min(dt, dtold*dtrise, dtmax); small flag; dtold = dt.
hipError_t batch_deck_dt(DeckDev *decks, int nd, const int *visc_start, const double *partial, double g_big,
                         hipStream_t st);
// the five sums of each deck's summary partials, as the single-deck final pass
hipError_t batch_deck_sum5(int nd, const int *sum_start, const double *partial, double *out, hipStream_t st);
// exchange up to 5 pairs of array slots in every deck's table; add dt to
// time when add_time
struct SlotSwap {
    int n;
    int a[5], b[5];
    int add_time;
};
hipError_t batch_swap(DeckDev *decks, int nd, const SlotSwap &sw, hipStream_t st);

}  // namespace clover
