// Halo kernels for gfx950: the reflective boundary of external faces
// (update_halo_kernel.f90:115-1425) and the tile-to-tile copies
// (update_tile_halo_kernel.f90:31-1993), table driven.
//
// One launch handles one axis (both faces when the tile is at least 3 cells
// thick along it, else one face) for every flagged field: blockIdx.z picks
// (field, face), threads cover the two tangential axes, each thread walks
// the 1..depth halo layers.  Exact moves, no arithmetic.
//
// The resident driver refreshes only the (field, axis) halos its kernels
// read (the axis masks below); in its test mode the halos it skips are
// instead filled with wrong values (1.5 x + 0.25 of the value the reference
// would put there), so any read of a skipped halo changes interior results.
#include <hip/hip_runtime.h>

#include "batch.h"
#include "fast.h"
#include "kernels.h"

namespace clover {

// How a field sits on an axis (see the reflective formulas below).
enum { K_CELL = 0, K_STAG = 1, K_FTAN = 2 };

struct HaloDesc {
    unsigned char kind[3];   // per axis x, y, z
    unsigned char neg[3];    // sign flip on the faces normal to this axis
    unsigned char ext[6][2]; // per face L,R,B,T,K,F: +1 on the first/second tangential range
};

// Reflective halo, per field id (data.f90:55-74).  For a face normal to
// axis a, for d = 1..depth:   halo(W) = sgn * f(R)   with
//   CELL  low W=1-d R=d     high W=max+d   R=max+1-d   (cell-centred data)
//   STAG  low W=1-d R=1+d   high W=max+1+d R=max+1-d   (vertex data, face data normal to a)
//   FTAN  low W=1-d R=1+d   high W=max+d   R=max-d     (face data tangential to a)
// and tangential ranges min-depth..max+depth (+ext).  Transcribed from
// update_halo_kernel.f90: cells 115-596, velocities 604-1004, vol/mass_flux_x
// 1018-1142, _y 1156-1280, _z 1291-1417 (vol_flux_z's bottom face stops at
// z_max+depth, mass_flux_z's at z_max+1+depth).
static const HaloDesc kHalo[NUM_FIELDS] = {
    {{0, 0, 0}, {0, 0, 0}, {{0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}}},  // density0
    {{0, 0, 0}, {0, 0, 0}, {{0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}}},  // density1
    {{0, 0, 0}, {0, 0, 0}, {{0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}}},  // energy0
    {{0, 0, 0}, {0, 0, 0}, {{0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}}},  // energy1
    {{0, 0, 0}, {0, 0, 0}, {{0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}}},  // pressure
    {{0, 0, 0}, {0, 0, 0}, {{0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}}},  // viscosity
    {{0, 0, 0}, {0, 0, 0}, {{0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}}},  // soundspeed
    {{1, 1, 1}, {1, 0, 0}, {{1, 1}, {1, 1}, {1, 1}, {1, 1}, {1, 1}, {1, 1}}},  // xvel0
    {{1, 1, 1}, {1, 0, 0}, {{1, 1}, {1, 1}, {1, 1}, {1, 1}, {1, 1}, {1, 1}}},  // xvel1
    {{1, 1, 1}, {0, 1, 0}, {{1, 1}, {1, 1}, {1, 1}, {1, 1}, {1, 1}, {1, 1}}},  // yvel0
    {{1, 1, 1}, {0, 1, 0}, {{1, 1}, {1, 1}, {1, 1}, {1, 1}, {1, 1}, {1, 1}}},  // yvel1
    {{1, 1, 1}, {0, 0, 1}, {{1, 1}, {1, 1}, {1, 1}, {1, 1}, {1, 1}, {1, 1}}},  // zvel0
    {{1, 1, 1}, {0, 0, 1}, {{1, 1}, {1, 1}, {1, 1}, {1, 1}, {1, 1}, {1, 1}}},  // zvel1
    {{1, 2, 2}, {1, 0, 0}, {{0, 0}, {0, 0}, {1, 0}, {1, 0}, {0, 0}, {0, 0}}},  // vol_flux_x
    {{2, 1, 2}, {0, 1, 0}, {{1, 0}, {1, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}}},  // vol_flux_y
    {{2, 2, 1}, {0, 0, 1}, {{0, 1}, {0, 1}, {0, 0}, {0, 1}, {0, 0}, {0, 0}}},  // vol_flux_z
    {{1, 2, 2}, {1, 0, 0}, {{0, 0}, {0, 0}, {1, 0}, {1, 0}, {0, 0}, {0, 0}}},  // mass_flux_x
    {{2, 1, 2}, {0, 1, 0}, {{1, 0}, {1, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}}},  // mass_flux_y
    {{2, 2, 1}, {0, 0, 1}, {{0, 1}, {0, 1}, {0, 1}, {0, 1}, {0, 0}, {0, 0}}},  // mass_flux_z
};

// staggered (vertex-like) along x, y, z, per field id
static const unsigned char kStag[NUM_FIELDS][3] = {
    {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0},
    {1, 1, 1}, {1, 1, 1}, {1, 1, 1}, {1, 1, 1}, {1, 1, 1}, {1, 1, 1},
    {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};

constexpr int kMaxItems = 2 * NUM_FIELDS;

struct HaloItem {
    void *dst;  // arrays of the launch's element type
    const void *src;
    double sgn;
    int kind;  // reflective: K_*; tile copy: staggered on the axis (0/1)
    int hi;    // 0 low face, 1 high face
    int ea, eb;
    int poison;  // write the test-mode wrong value instead
};

__device__ inline double poisoned(double v) { return v * 1.5 + 0.25; }

// does a launch for `axis` refresh field f?  fields[f] == 1 selects the
// fields the reference refreshes; axes (null: every axis) narrows that to
// a bit mask of axes (1 x, 2 y, 4 z) per field.
static inline bool wanted(const int *axes, int f, int axis) { return !axes || ((axes[f] >> axis) & 1); }

struct HaloBatch {
    HaloItem it[kMaxItems];
};

constexpr int HX = 64, HY = 4;

__device__ inline int tan_a(int axis) { return axis == 0 ? 1 : 0; }
__device__ inline int tan_b(int axis) { return axis == 2 ? 1 : 2; }

// reflective halo of one (field, face) item over the tangential block (bx, by)
template <typename S>
__device__ inline void reflect_body(const Lay &g, int axis, int depth, S *dst, double sgn, int kind, int hi, int ea,
                                    int eb, int bx, int by, int poison)
{
    const int mn[3] = {g.x_min, g.y_min, g.z_min};
    const int mx[3] = {g.x_max, g.y_max, g.z_max};
    const int ta = tan_a(axis), tb = tan_b(axis);
    const int c = mn[ta] - depth + (bx * HX + (int)threadIdx.x);
    const int b = mn[tb] - depth + (by * HY + (int)threadIdx.y);
    if (c > mx[ta] + ea + depth || b > mx[tb] + eb + depth) return;
    const int amax = mx[axis];
    // every depth's source lies inside the tile and every target in the halo,
    // so all loads are issued before the first store
    int iw[2], ir[2];
    S v[2];
    for (int d = 1; d <= 2; d++) {
        if (d > depth) break;
        int w, r;
        if (!hi) {
            w = 1 - d;
            r = kind == K_CELL ? d : 1 + d;
        } else if (kind == K_CELL) {
            w = amax + d; r = amax + 1 - d;
        } else if (kind == K_STAG) {
            w = amax + 1 + d; r = amax + 1 - d;
        } else {
            w = amax + d; r = amax - d;
        }
        int jw[3], jr[3];
        jw[axis] = w; jr[axis] = r;
        jw[ta] = jr[ta] = c;
        jw[tb] = jr[tb] = b;
        iw[d - 1] = idx(g, jw[0], jw[1], jw[2]);
        ir[d - 1] = idx(g, jr[0], jr[1], jr[2]);
        v[d - 1] = dst[ir[d - 1]];
    }
    for (int d = 1; d <= 2; d++) {
        if (d > depth) break;
        const double x = sgn * (double)v[d - 1];
        dst[iw[d - 1]] = (S)(poison ? poisoned(x) : x);
    }
}

template <typename S>
__global__ void k_reflect(Lay g, int axis, int depth, HaloBatch batch)
{
    const HaloItem &it = batch.it[blockIdx.z];
    reflect_body<S>(g, axis, depth, static_cast<S *>(it.dst), it.sgn, it.kind, it.hi, it.ea, it.eb, (int)blockIdx.x,
                    (int)blockIdx.y, it.poison);
}

static hipError_t run_batch_reflect(Store s, const Lay &g, int axis, int depth, const HaloBatch &b, int n,
                                    hipStream_t st)
{
    if (n == 0) return hipSuccess;
    if (depth < 1 || depth > 2) return hipErrorInvalidValue;  // reflect_body holds two depths
    const int mn[3] = {g.x_min, g.y_min, g.z_min};
    const int mx[3] = {g.x_max, g.y_max, g.z_max};
    const int ta = axis == 0 ? 1 : 0, tb = axis == 2 ? 1 : 2;
    const int na = mx[ta] - mn[ta] + 2 + 2 * depth, nb = mx[tb] - mn[tb] + 2 + 2 * depth;
    if (t_kernel_timer) t_kernel_timer->begin("k_reflect", st);
    const dim3 grid((na + HX - 1) / HX, (nb + HY - 1) / HY, n);
    if (s == ST_F32) hipLaunchKernelGGL(k_reflect<float>, grid, dim3(HX, HY, 1), 0, st, g, axis, depth, b);
    else hipLaunchKernelGGL(k_reflect<double>, grid, dim3(HX, HY, 1), 0, st, g, axis, depth, b);
    if (t_kernel_timer) t_kernel_timer->end(st);
    return hipGetLastError();
}

hipError_t launch_update_halo_store(Store s, const Lay &g, void *const *fp, const int *fields, const int *ext,
                                    int depth, hipStream_t st, const int *axes, bool poison)
{
    // The reference's order per field is bottom, top, left, right, back,
    // front; fields are independent arrays, so all fields' y faces go first,
    // then x, then z.  The two faces of one axis touch disjoint cells when the
    // tile is at least 3 cells thick along it and then share a launch.
    static const int axis_order[3] = {1, 0, 2};
    const int mn[3] = {g.x_min, g.y_min, g.z_min};
    const int mx[3] = {g.x_max, g.y_max, g.z_max};
    for (int ao = 0; ao < 3; ao++) {
        const int axis = axis_order[ao];
        const bool together = mx[axis] - mn[axis] + 1 >= 3;
        HaloBatch b;
        int n = 0;
        for (int side = 0; side < 2; side++) {
            const int face = 2 * axis + side;  // 0..5 = left,right,bottom,top,back,front
            if (!ext[face]) continue;
            for (int f = 0; f < NUM_FIELDS; f++) {
                if (fields[f] != 1 || !fp[f]) continue;
                const bool want = wanted(axes, f, axis);
                if (!want && !poison) continue;
                const HaloDesc &h = kHalo[f];
                HaloItem &it = b.it[n++];
                it.poison = !want;
                it.dst = fp[f];
                it.src = fp[f];
                it.sgn = h.neg[axis] ? -1.0 : 1.0;
                it.kind = h.kind[axis];
                it.hi = side;
                it.ea = h.ext[face][0];
                it.eb = h.ext[face][1];
            }
            if (!together) {
                hipError_t e = run_batch_reflect(s, g, axis, depth, b, n, st);
                if (e != hipSuccess) return e;
                n = 0;
            }
        }
        hipError_t e = run_batch_reflect(s, g, axis, depth, b, n, st);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// ---------------------------------------------------------------------------
// Tile copies.  Along the copy axis a field is cell-like or staggered;
//   low face   f(min-d)   = N(Nmax+1-d)
//   high face  f(max+d)   = N(Nmin-1+d)   cell-like
//              f(max+1+d) = N(Nmin+d)     staggered
// tangential ranges min-depth..max+depth, +1 where the field is staggered.
template <typename S>
__device__ inline void tile_copy_body(const Lay &g, const Lay &gn, int axis, int depth, S *dst, const S *src, int kind,
                                      int hi, int ea, int eb, int poison, int bx, int by)
{
    const int mn[3] = {g.x_min, g.y_min, g.z_min};
    const int mx[3] = {g.x_max, g.y_max, g.z_max};
    const int nmn[3] = {gn.x_min, gn.y_min, gn.z_min};
    const int nmx[3] = {gn.x_max, gn.y_max, gn.z_max};
    const int ta = tan_a(axis), tb = tan_b(axis);
    const int c = mn[ta] - depth + (bx * HX + (int)threadIdx.x);
    const int b = mn[tb] - depth + (by * HY + (int)threadIdx.y);
    if (c > mx[ta] + ea + depth || b > mx[tb] + eb + depth) return;
    for (int d = 1; d <= depth; d++) {
        int w, r;
        if (!hi) {
            w = mn[axis] - d; r = nmx[axis] + 1 - d;
        } else if (!kind) {
            w = mx[axis] + d; r = nmn[axis] - 1 + d;
        } else {
            w = mx[axis] + 1 + d; r = nmn[axis] + d;
        }
        int iw[3], ir[3];
        iw[axis] = w; ir[axis] = r;
        iw[ta] = ir[ta] = c;
        iw[tb] = ir[tb] = b;
        const S v = src[idx(gn, ir[0], ir[1], ir[2])];
        dst[idx(g, iw[0], iw[1], iw[2])] = poison ? (S)poisoned((double)v) : v;
    }
}

template <typename S>
__global__ void k_tile_copy(Lay g, Lay gn, int axis, int depth, HaloBatch batch)
{
    const HaloItem &it = batch.it[blockIdx.z];
    tile_copy_body<S>(g, gn, axis, depth, static_cast<S *>(it.dst), static_cast<const S *>(it.src), it.kind, it.hi,
                      it.ea, it.eb, it.poison, (int)blockIdx.x, (int)blockIdx.y);
}

hipError_t launch_tile_halo_store(Store s, int face, const Lay &g, void *const *fp, const Lay &gn, void *const *np,
                                  const int *fields, int depth, hipStream_t st, const int *axes, bool poison)
{
    const int axis = (face - 1) / 2;
    const int hi = (face - 1) % 2;
    const int ta = axis == 0 ? 1 : 0, tb = axis == 2 ? 1 : 2;
    HaloBatch b;
    int n = 0;
    for (int f = 0; f < NUM_FIELDS; f++) {
        int gate = f, src = f;
        // quirks of update_tile_halo_kernel.f90: energy1 follows the density1
        // flag on the left, right, bottom and top faces (113-125, 443-455,
        // 771-783, 1103-1115); the right face's zvel0 branch copies yvel0
        // (579-589).
        if (f == F_ENERGY1 && face <= 4) gate = F_DENSITY1;
        if (f == F_ZVEL0 && face == 2) src = F_YVEL0;
        if (fields[gate] != 1) continue;
        if (!fp[src] || !np[src]) continue;
        const bool want = wanted(axes, src, axis);
        if (!want && !poison) continue;
        HaloItem &it = b.it[n++];
        it.poison = !want;
        it.dst = fp[src];
        it.src = np[src];
        it.sgn = 1.0;
        it.kind = kStag[src][axis];
        it.hi = hi;
        it.ea = kStag[src][ta];
        it.eb = kStag[src][tb];
    }
    if (n == 0) return hipSuccess;
    const int mn[3] = {g.x_min, g.y_min, g.z_min};
    const int mx[3] = {g.x_max, g.y_max, g.z_max};
    const int na = mx[ta] - mn[ta] + 2 + 2 * depth, nb = mx[tb] - mn[tb] + 2 + 2 * depth;
    if (t_kernel_timer) t_kernel_timer->begin("k_tile_copy", st);
    const dim3 grid((na + HX - 1) / HX, (nb + HY - 1) / HY, n);
    if (s == ST_F32) hipLaunchKernelGGL(k_tile_copy<float>, grid, dim3(HX, HY, 1), 0, st, g, gn, axis, depth, b);
    else hipLaunchKernelGGL(k_tile_copy<double>, grid, dim3(HX, HY, 1), 0, st, g, gn, axis, depth, b);
    if (t_kernel_timer) t_kernel_timer->end(st);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Batched reflective halo over many decks (batch.h): both faces of one axis
// for the flagged fields; every deck is at least 3 cells thick along every
// axis, so the two faces touch disjoint cells and share the launch.
struct BItem {
    int field, hi, kind, ea, eb;
    double sgn;
};
struct BItems {
    int n;
    BItem it[kMaxItems];
};

__host__ __device__ inline void dims_reflect(const Lay &g, int axis, int depth, int &ga, int &gb)
{
    const int mn[3] = {g.x_min, g.y_min, g.z_min};
    const int mx[3] = {g.x_max, g.y_max, g.z_max};
    const int ta = axis == 0 ? 1 : 0, tb = axis == 2 ? 1 : 2;
    const int na = mx[ta] - mn[ta] + 2 + 2 * depth, nb = mx[tb] - mn[tb] + 2 + 2 * depth;
    ga = (na + HX - 1) / HX;
    gb = (nb + HY - 1) / HY;
}

int batch_blocks_reflect(const Lay &g, int axis, int depth, int nitems)
{
    int ga, gb;
    dims_reflect(g, axis, depth, ga, gb);
    return nitems * ga * gb;
}

template <typename S>
__global__ void kb_reflect(const DeckDev *__restrict__ decks, const int *__restrict__ start, int nd, int axis,
                           int depth, BItems items)
{
    const int d = deck_of(start, nd, (int)blockIdx.x);
    const DeckDev &D = decks[d];
    int ga, gb;
    dims_reflect(D.g, axis, depth, ga, gb);
    const int lb = (int)blockIdx.x - start[d];
    const int item = lb / (ga * gb), r = lb % (ga * gb);
    const BItem &it = items.it[item];
    reflect_body<S>(D.g, axis, depth, static_cast<S *>(D.f[it.field]), it.sgn, it.kind, it.hi, it.ea, it.eb, r % ga,
                    r / ga, 0);
}

hipError_t batch_reflect(Store s, const DeckDev *decks, int nd, const BatchGrid &bg, int axis, int depth,
                         const int *fields, hipStream_t st, const int *axes)
{
    if (depth < 1 || depth > 2) return hipErrorInvalidValue;  // reflect_body holds two depths
    BItems b;
    b.n = 0;
    for (int side = 0; side < 2; side++) {
        const int face = 2 * axis + side;
        for (int f = 0; f < NUM_FIELDS; f++) {
            if (fields[f] != 1 || !wanted(axes, f, axis)) continue;
            const HaloDesc &h = kHalo[f];
            BItem &it = b.it[b.n++];
            it.field = f;
            it.hi = side;
            it.kind = h.kind[axis];
            it.ea = h.ext[face][0];
            it.eb = h.ext[face][1];
            it.sgn = h.neg[axis] ? -1.0 : 1.0;
        }
    }
    if (b.n == 0 || bg.total == 0) return hipSuccess;
    if (t_kernel_timer) t_kernel_timer->begin("kb_reflect", st);
    if (s == ST_F32) hipLaunchKernelGGL(kb_reflect<float>, dim3(bg.total), dim3(HX, HY), 0, st, decks, bg.start, nd, axis, depth, b);
    else hipLaunchKernelGGL(kb_reflect<double>, dim3(bg.total), dim3(HX, HY), 0, st, decks, bg.start, nd, axis, depth, b);
    if (t_kernel_timer) t_kernel_timer->end(st);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// The tiles of one chunk in one launch (batch.h TileHaloItem): every item is
// one (tile, face, field) copy or reflection and owns per_a x per_b blocks
// of the launch, enough for the largest tile's face; blocks past an item's
// own face do nothing.
__host__ __device__ inline void item_block(int b, int per_a, int per_b, int &item, int &bx, int &by)
{
    item = b / (per_a * per_b);
    const int r = b % (per_a * per_b);
    bx = r % per_a;
    by = r / per_a;
}

template <typename S>
__global__ void __launch_bounds__(HX *HY)
kb_tile_copy(const DeckDev *__restrict__ tiles, const TileHaloItem *__restrict__ items, int per_a, int per_b,
             int axis, int depth)
{
    int item, bx, by;
    item_block((int)blockIdx.x, per_a, per_b, item, bx, by);
    const TileHaloItem it = items[item];
    const DeckDev &D = tiles[it.tile];
    const DeckDev &N = tiles[it.nbr];
    tile_copy_body<S>(D.g, N.g, axis, depth, static_cast<S *>(D.f[it.field]), static_cast<const S *>(N.f[it.field]),
                      it.kind, it.hi, it.ea, it.eb, it.poison, bx, by);
}

template <typename S>
__global__ void __launch_bounds__(HX *HY)
kb_tile_reflect(const DeckDev *__restrict__ tiles, const TileHaloItem *__restrict__ items, int per_a, int per_b,
                int axis, int depth)
{
    int item, bx, by;
    item_block((int)blockIdx.x, per_a, per_b, item, bx, by);
    const TileHaloItem it = items[item];
    const DeckDev &D = tiles[it.tile];
    reflect_body<S>(D.g, axis, depth, static_cast<S *>(D.f[it.field]), it.sgn, it.kind, it.hi, it.ea, it.eb, bx, by,
                    it.poison);
}

void tile_halo_blocks(const Lay &g, int axis, int depth, int &per_a, int &per_b)
{
    dims_reflect(g, axis, depth, per_a, per_b);
}

hipError_t batch_tile_halo(Store s, bool copy, const DeckDev *tiles, const TileHaloItem *items, int nitems, int per_a,
                           int per_b, int axis, int depth, hipStream_t st)
{
    if (nitems == 0) return hipSuccess;
    if (depth < 1 || depth > 2) return hipErrorInvalidValue;  // reflect_body holds two depths
    const dim3 grid((unsigned)(nitems * per_a * per_b)), blk(HX, HY);
    if (copy) {
        if (t_kernel_timer) t_kernel_timer->begin("kb_tile_copy", st);
        if (s == ST_F32) hipLaunchKernelGGL(kb_tile_copy<float>, grid, blk, 0, st, tiles, items, per_a, per_b, axis, depth);
        else hipLaunchKernelGGL(kb_tile_copy<double>, grid, blk, 0, st, tiles, items, per_a, per_b, axis, depth);
    } else {
        if (t_kernel_timer) t_kernel_timer->begin("kb_tile_reflect", st);
        if (s == ST_F32) hipLaunchKernelGGL(kb_tile_reflect<float>, grid, blk, 0, st, tiles, items, per_a, per_b, axis, depth);
        else hipLaunchKernelGGL(kb_tile_reflect<double>, grid, blk, 0, st, tiles, items, per_a, per_b, axis, depth);
    }
    if (t_kernel_timer) t_kernel_timer->end(st);
    return hipGetLastError();
}

// The items of the tile passes and reflective passes of one update_halo of
// a tiled chunk, in the order launch_tile_halo_store and
// launch_update_halo_store would touch them tile by tile.
void tile_halo_items(const int *nbr6, int ntiles, const Lay *lays, const int *fields, int depth, const int *axes,
                     const int *raxes, bool poison, std::vector<TileHaloItem> (&copy)[3],
                     std::vector<TileHaloItem> (&refl)[3][2])
{
    (void)depth;
    // update_tile_halo.f90:47-423: top/bottom, then left/right, then front/back
    static const int passes[3][2] = {{4, 3}, {1, 2}, {6, 5}};
    for (int p = 0; p < 3; p++) {
        copy[p].clear();
        for (int t = 0; t < ntiles; t++)
            for (int sd = 0; sd < 2; sd++) {
                const int face = passes[p][sd];
                const int n = nbr6[6 * t + face - 1];
                if (n == -1) continue;
                const int axis = (face - 1) / 2, hi = (face - 1) % 2;
                const int ta = axis == 0 ? 1 : 0, tb = axis == 2 ? 1 : 2;
                for (int f = 0; f < NUM_FIELDS; f++) {
                    int gate = f, src = f;
                    if (f == F_ENERGY1 && face <= 4) gate = F_DENSITY1;
                    if (f == F_ZVEL0 && face == 2) src = F_YVEL0;
                    if (fields[gate] != 1) continue;
                    const bool want = wanted(axes, src, axis);
                    if (!want && !poison) continue;
                    TileHaloItem it{};
                    it.tile = t;
                    it.nbr = n - 1;
                    it.field = src;
                    it.hi = hi;
                    it.kind = kStag[src][axis];
                    it.ea = kStag[src][ta];
                    it.eb = kStag[src][tb];
                    it.poison = !want;
                    it.sgn = 1.0;
                    copy[p].push_back(it);
                }
            }
    }
    // update_halo_kernel.f90 per tile: y faces, then x, then z; the two
    // faces of an axis go to separate lists (refl[axis][side]) so that a
    // tile under 3 cells thick can take them one after the other
    static const int axis_order[3] = {1, 0, 2};
    for (int ao = 0; ao < 3; ao++) {
        const int axis = axis_order[ao];
        for (int sd = 0; sd < 2; sd++) {
            refl[ao][sd].clear();
            const int face = 2 * axis + sd;
            for (int t = 0; t < ntiles; t++) {
                if (nbr6[6 * t + face] != -1) continue;
                (void)lays;
                for (int f = 0; f < NUM_FIELDS; f++) {
                    if (fields[f] != 1) continue;
                    const bool want = wanted(raxes, f, axis);
                    if (!want && !poison) continue;
                    const HaloDesc &h = kHalo[f];
                    TileHaloItem it{};
                    it.tile = t;
                    it.nbr = t;
                    it.field = f;
                    it.hi = sd;
                    it.kind = h.kind[axis];
                    it.ea = h.ext[face][0];
                    it.eb = h.ext[face][1];
                    it.poison = !want;
                    it.sgn = h.neg[axis] ? -1.0 : 1.0;
                    refl[ao][sd].push_back(it);
                }
            }
        }
    }
}

// FP64 entry points used by the drop-in ABI and the per-call driver
hipError_t launch_update_halo(const Lay &g, double *const *fp, const int *fields, const int *ext, int depth,
                              hipStream_t st)
{
    void *p[NUM_FIELDS];
    for (int f = 0; f < NUM_FIELDS; f++) p[f] = fp[f];
    return launch_update_halo_store(ST_F64, g, p, fields, ext, depth, st);
}

hipError_t launch_tile_halo(int face, const Lay &g, double *const *fp, const Lay &gn, double *const *np,
                            const int *fields, int depth, hipStream_t st)
{
    void *p[NUM_FIELDS], *q[NUM_FIELDS];
    for (int f = 0; f < NUM_FIELDS; f++) {
        p[f] = fp[f];
        q[f] = np[f];
    }
    return launch_tile_halo_store(ST_F64, face, g, p, gn, q, fields, depth, st);
}

}  // namespace clover
