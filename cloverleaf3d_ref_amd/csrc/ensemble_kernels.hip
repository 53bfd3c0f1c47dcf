// Per-deck kernels of the ensemble driver (batch.h): one workgroup (or one
// thread) per deck, so every deck keeps its own dt, time and array roles
// on the device.
#include <hip/hip_runtime.h>

#include "batch.h"
#include "launch.h"

namespace clover {

// timestep.f90:86-126 for one tile per deck: the minimum over the deck's
// calc_dt partials (seeded with g_big, as the single-deck final pass), then
// dt = MIN(dt, dtold*dtrise, dtmax), the small-dt flag and dtold = dt.  The
// reference stops the run where dt falls below dtmin (timestep.f90:120-122);
// a deck of the batch records that step and keeps stepping with the others,
// so its later state is not one the reference produces.
__global__ void __launch_bounds__(RED_THREADS)
kb_deck_dt(DeckDev *__restrict__ decks, const int *__restrict__ start, const double *__restrict__ partial,
           double g_big)
{
    const int d = blockIdx.x;
    const int n = start[d + 1] - start[d];
    const double *p = partial + start[d];
    double m = g_big;
    for (int c = threadIdx.x; c < n; c += RED_THREADS) m = dmin(m, p[c]);
    __shared__ double red[RED_THREADS / 64];
    m = wave_min(m);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double dtlp = red[0];
        for (int w = 1; w < RED_THREADS / 64; w++) dtlp = dmin(dtlp, red[w]);
        DeckDev &D = decks[d];
        double dt = g_big;
        if (dtlp <= dt) dt = dtlp;
        double mm = dt;
        if (D.dtold * D.dtrise < mm) mm = D.dtold * D.dtrise;
        if (D.dtmax < mm) mm = D.dtmax;
        dt = mm;
        D.steps = D.steps + 1;
        if (dt < D.dtmin && D.small == 0) D.small = D.steps;
        D.dt = dt;
        D.dtold = dt;
    }
}

hipError_t batch_deck_dt(DeckDev *decks, int nd, const int *visc_start, const double *partial, double g_big,
                         hipStream_t st)
{
    TIMED_LAUNCH("kb_deck_dt", st, kb_deck_dt, dim3(nd), dim3(RED_THREADS), 0, st, decks, visc_start, partial, g_big);
    return hipGetLastError();
}

// the five sums of each deck's summary partials, in the order of the
// single-deck final pass
__global__ void __launch_bounds__(RED_THREADS)
kb_deck_sum5(const int *__restrict__ start, const double *__restrict__ partial, double *__restrict__ out)
{
    const int d = blockIdx.x;
    const int n = start[d + 1] - start[d];
    const double *p = partial + 5 * start[d];
    double s[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int c = threadIdx.x; c < n; c += RED_THREADS)
        for (int q = 0; q < 5; q++) s[q] += p[5 * c + q];
    __shared__ double red[5][RED_THREADS / 64];
    for (int q = 0; q < 5; q++) {
        const double w = wave_sum(s[q]);
        if ((threadIdx.x & 63) == 0) red[q][threadIdx.x >> 6] = w;
    }
    __syncthreads();
    if (threadIdx.x < 5) {
        double b = 0.0;
        for (int w = 0; w < RED_THREADS / 64; w++) b += red[threadIdx.x][w];
        out[5 * d + threadIdx.x] = b;
    }
}

hipError_t batch_deck_sum5(int nd, const int *sum_start, const double *partial, double *out, hipStream_t st)
{
    TIMED_LAUNCH("kb_deck_sum5", st, kb_deck_sum5, dim3(nd), dim3(RED_THREADS), 0, st, sum_start, partial, out);
    return hipGetLastError();
}

// a tile's partials seeded with g_big: the single-tile final pass's minimum
__global__ void __launch_bounds__(RED_THREADS)
kb_tile_min(const int *__restrict__ start, const double *__restrict__ partial, double g_big, double *__restrict__ out,
            int stride)
{
    const int t = blockIdx.x;
    const int n = start[t + 1] - start[t];
    const double *p = partial + start[t];
    double m = g_big;
    for (int c = threadIdx.x; c < n; c += RED_THREADS) m = dmin(m, p[c]);
    __shared__ double red[RED_THREADS / 64];
    m = wave_min(m);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double b = red[0];
        for (int w = 1; w < RED_THREADS / 64; w++) b = dmin(b, red[w]);
        out[stride * t] = b;
    }
}

hipError_t batch_tile_min(int ntiles, const int *start, const double *partial, double g_big, double *out, int stride,
                          hipStream_t st)
{
    TIMED_LAUNCH("kb_tile_min", st, kb_tile_min, dim3(ntiles), dim3(RED_THREADS), 0, st, start, partial, g_big, out,
                 stride);
    return hipGetLastError();
}

__global__ void kb_set_dt(DeckDev *__restrict__ decks, int nd, double dt)
{
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d < nd) decks[d].dt = dt;
}

hipError_t batch_set_dt(DeckDev *decks, int nd, double dt, hipStream_t st)
{
    TIMED_LAUNCH("kb_set_dt", st, kb_set_dt, dim3((nd + 63) / 64), dim3(64), 0, st, decks, nd, dt);
    return hipGetLastError();
}

__global__ void kb_swap(DeckDev *__restrict__ decks, int nd, SlotSwap sw)
{
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= nd) return;
    DeckDev &D = decks[d];
    for (int q = 0; q < sw.n; q++) {
        void *t = D.f[sw.a[q]];
        D.f[sw.a[q]] = D.f[sw.b[q]];
        D.f[sw.b[q]] = t;
    }
    if (sw.add_time) D.time = D.time + D.dt;
}

hipError_t batch_swap(DeckDev *decks, int nd, const SlotSwap &sw, hipStream_t st)
{
    TIMED_LAUNCH("kb_swap", st, kb_swap, dim3((nd + 63) / 64), dim3(64), 0, st, decks, nd, sw);
    return hipGetLastError();
}

}  // namespace clover
