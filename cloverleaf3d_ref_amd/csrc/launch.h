// Launch geometry shared by the stencil kernels.
//
// A 64 x 4 workgroup covers 64 consecutive j (one wavefront per row, so every
// array access of a wave is one coalesced 512-byte segment) by 4 rows k, at
// one plane l per grid z index.  Ranges are inclusive, in the reference's
// tile-local indices.
#pragma once

#include <hip/hip_runtime.h>

#include "fast.h"
#include "kernels.h"

namespace clover {

struct Range {
    int j0, j1, k0, k1, l0, l1;  // inclusive
};

constexpr int BX = 64, BY = 4;

static inline dim3 grid_of(const Range &r)
{
    return dim3((unsigned)((r.j1 - r.j0 + 1 + BX - 1) / BX), (unsigned)((r.k1 - r.k0 + 1 + BY - 1) / BY),
                (unsigned)(r.l1 - r.l0 + 1));
}

#define RANGE_THREAD(r)                                              \
    const int j = (r).j0 + (int)(blockIdx.x * BX + threadIdx.x);     \
    const int k = (r).k0 + (int)(blockIdx.y * BY + threadIdx.y);     \
    const int l = (r).l0 + (int)blockIdx.z;                          \
    if (j > (r).j1 || k > (r).k1) return;

// Block coordinates of a kernel body: the hardware's for a launch over one
// tile, or a block of one deck's share of a launch over many decks.
struct Blk {
    int x, y, z;     // block index
    int gx, gy, gz;  // grid dimensions
};

__device__ inline Blk hw_blk()
{
    return Blk{(int)blockIdx.x, (int)blockIdx.y, (int)blockIdx.z, (int)gridDim.x, (int)gridDim.y, (int)gridDim.z};
}

// The hardware deals workgroups round-robin to the 8 XCDs, each with its own
// L2.  hw_blk_xcd() hands XCD x the x-th contiguous eighth of the grid (x
// fastest) instead, so tiles that share cache lines with their neighbours
// along x and y run on the same XCD at about the same time.  It pays in the
// y and z momentum sweeps and costs in the other kernels, which keep the
// plain order.
constexpr int kXcds = 8;

// the logical block that hardware block hw of n runs in the XCD-contiguous
// order (a permutation of 0..n-1)
__host__ __device__ inline int xcd_block(int hw, int n)
{
    const int xcd = hw % kXcds, t = hw / kXcds;
    const int q = n / kXcds, r = n % kXcds;  // XCDs below r take q+1 blocks
    return xcd < r ? xcd * (q + 1) + t : r * (q + 1) + (xcd - r) * q + t;
}

__device__ inline Blk hw_blk_xcd()
{
    const int gx = (int)gridDim.x, gy = (int)gridDim.y, gz = (int)gridDim.z;
    const int hw = (int)blockIdx.x + gx * ((int)blockIdx.y + gy * (int)blockIdx.z);
    const int id = xcd_block(hw, gx * gy * gz);
    return Blk{id % gx, (id / gx) % gy, id / (gx * gy), gx, gy, gz};
}

__device__ inline Blk hw_blk(int xcd) { return xcd ? hw_blk_xcd() : hw_blk(); }

// Marching kernels split the positions they march over into bk.gz chunks of
// at most `planes` (the launch shape, fast.h Shape); chunk z of gz covers
// [c0, c1] of [lo, hi] (empty when c0 > c1, which only a grid larger than
// nchunks() gives).
__host__ __device__ inline int nchunks(int npos, int planes) { return (npos + planes - 1) / planes; }

__device__ inline void chunk_range(const Blk &b, int lo, int hi, int &c0, int &c1)
{
    const int n = hi - lo + 1;
    const int per = (n + b.gz - 1) / b.gz;
    c0 = lo + b.z * per;
    c1 = min(c0 + per - 1, hi);
}

#define TIMED_LAUNCH(name, st, ...)                           \
    do {                                                      \
        if (t_kernel_timer) t_kernel_timer->begin(name, st);  \
        hipLaunchKernelGGL(__VA_ARGS__);                      \
        if (t_kernel_timer) t_kernel_timer->end(st);          \
    } while (0)

#define LAUNCH(kern, r, st, ...)                                                                \
    do {                                                                                        \
        if ((r).j1 >= (r).j0 && (r).k1 >= (r).k0 && (r).l1 >= (r).l0)                           \
            TIMED_LAUNCH(#kern, st, kern, grid_of(r), dim3(BX, BY, 1), 0, st, __VA_ARGS__);      \
    } while (0)

// neighbour access relative to the thread's offset i
#define AT(p, dj, dk, dl) (p)[i + (dj) + (dk) * g.ld + (dl) * g.plane]

__device__ inline double dmin(double a, double b) { return a < b ? a : b; }
__device__ inline double dmax(double a, double b) { return a > b ? a : b; }

// Arithmetic of the two store forms.  The FP64 form evaluates the reference's
// expressions exactly (true divisions, same order); the FP32 form computes in
// FP32 and divides by the uniform mesh's constants through reciprocals formed
// once per thread, and by everything else through the hardware reciprocal.
// min/max: in FP64 the reference's comparison (the operand it keeps on a
// tie, sign of zero included); in FP32 the hardware's min/max instruction
template <typename T>
__device__ inline T tmin(T a, T b)
{
    if constexpr (sizeof(T) == 4) return __builtin_fminf(a, b);
    else return a < b ? a : b;
}
template <typename T>
__device__ inline T tmax(T a, T b)
{
    if constexpr (sizeof(T) == 4) return __builtin_fmaxf(a, b);
    else return a > b ? a : b;
}

// a / b for a mesh constant b (fast.h Recip): in FP64 the correctly rounded
// quotient, equal to a / b, from three FMAs and a multiply instead of the
// divide sequence; in FP32 a times the reciprocal
template <typename T>
__device__ inline T by_const(T a, const Recip &r)
{
    if constexpr (sizeof(T) == 8) {
        const double q = fma(a, r.hd, a * r.ld);  // within one ulp of a / b
        const double e = fma(q, r.b, -a);        // q*b - a, exact
        return fma(-e, r.hn, q);
    } else {
        return a * (T)r.hn;
    }
}

// (s * a) / b for a constant scale s and a mesh constant b: FP64 forms s*a
// and divides as by_const does; FP32 multiplies a by s/b (one product)
template <typename T>
__device__ inline T by_const_scaled(T a, double s, const Recip &r)
{
    if constexpr (sizeof(T) == 8) return by_const<T>((T)s * a, r);
    else return a * (T)(s * r.hn);
}

// a / b
template <typename T>
__device__ inline T quo(T a, T b)
{
    if constexpr (sizeof(T) == 8) return a / b;
    else return a * __builtin_amdgcn_rcpf(b);
}

// a / b for b > 0 where a is mostly zero (the artificial viscosity outside
// shocks): in FP64 a zero a is its own quotient, and the divide sequence is
// skipped when the whole wavefront has a == 0 (the FP32 reciprocal is
// cheaper than the branch)
template <typename T>
__device__ inline T quo_sparse(T a, T b)
{
    if constexpr (sizeof(T) == 4) return quo(a, b);
    T r = a;
    if (a != (T)0) r = quo(a, b);
    return r;
}

// ideal_gas_kernel.f90:50-57 for one cell, in the form's arithmetic (shared
// by the ideal_gas kernels and the advection sweep that also writes the
// next step's equation of state)
template <typename T>
__device__ inline T pressure_of(T density, T energy)
{
    return (T)(1.4 - 1.0) * density * energy;
}

// ideal_gas_kernel.f90:52-57: the soundspeed of a cell from its density and
// its pressure (the pressure ideal_gas forms from density and energy); the
// same operations on the same values as ideal_gas_cell, so the same bits
template <typename T>
__device__ inline T soundspeed_of(T density, T pressure)
{
    const T d = density, p = pressure;
    const T v = quo((T)1, d);
    const T pressurebyenergy = (T)(1.4 - 1.0) * d;
    const T pressurebyvolume = -d * p;
    const T sound_speed_squared = v * v * (p * pressurebyenergy - pressurebyvolume);
    return sqrt(sound_speed_squared);
}

template <typename S>
__device__ inline void ideal_gas_cell(S density, S energy, S &pressure, S &soundspeed)
{
    using T = S;
    const T d = density;
    const T p = pressure_of<T>(d, (T)energy);
    pressure = p;
    soundspeed = soundspeed_of<T>(d, p);
}

// Index a fused kernel reads for index q along axis a (0 x, 1 y, 2 z) of a
// cell field and of a vertex field (Lay::mirror): on a face whose
// reflective halo is not exchanged, the interior index that halo would copy
// (update_halo_kernel.f90: cells W = 1-d <- R = d and W = max+d <- max+1-d;
// vertices W = 1-d <- 1+d and W = max+1+d <- max+1-d), else q itself.  The
// velocity component along a also flips its sign there.
__device__ inline int cell_m(const Lay &g, int a, int q)
{
    const int lo = a == 0 ? g.x_min : a == 1 ? g.y_min : g.z_min;
    const int hi = a == 0 ? g.x_max : a == 1 ? g.y_max : g.z_max;
    const int bits = g.mirror >> (2 * a);
    if ((bits & 1) && q < lo) return 2 * lo - 1 - q;
    if ((bits & 2) && q > hi) return 2 * hi + 1 - q;
    return q;
}

__device__ inline int node_m(const Lay &g, int a, int q)
{
    const int lo = a == 0 ? g.x_min : a == 1 ? g.y_min : g.z_min;
    const int hi = a == 0 ? g.x_max : a == 1 ? g.y_max : g.z_max;
    const int bits = g.mirror >> (2 * a);
    if ((bits & 1) && q < lo) return 2 * lo - q;
    if ((bits & 2) && q > hi + 1) return 2 * hi + 2 - q;
    return q;
}

// The same for vol_flux_y and vol_flux_z across the x faces: update_halo
// reflects them about the edge cell's centre (update_halo_kernel.f90:
// 1173-1193 and 1306-1324: W = 1-d <- 1+d, W = max+d <- max-d), with no
// sign change.  Its later y and z passes reflect the x halos' own values, so
// every corner equals the reflection of the mirrored interior column's
// value.
__device__ inline int tan_xm(const Lay &g, int j)
{
    if ((g.mirror & 1) && j < g.x_min) return 2 * g.x_min - j;
    if ((g.mirror & 2) && j > g.x_max) return 2 * g.x_max - j;
    return j;
}

__device__ inline int cell_xm(const Lay &g, int j) { return cell_m(g, 0, j); }
__device__ inline int node_xm(const Lay &g, int j) { return node_m(g, 0, j); }
__device__ inline bool node_xsign(const Lay &g, int j) { return node_xm(g, j) != j; }

constexpr int RED_THREADS = 256;

__device__ inline double wave_min(double v)
{
    for (int off = 32; off > 0; off >>= 1) v = dmin(v, __shfl_xor(v, off, 64));
    return v;
}

__device__ inline double wave_sum(double v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

}  // namespace clover
