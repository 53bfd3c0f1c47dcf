// Per-call hydro kernels for gfx950 (FP64).
//
// One HIP kernel per loop nest of the reference *_kernel subroutines, with
// the same iteration domain and the same left-to-right FP64 expression order
// (this file is compiled with -ffp-contract=off, so results match the
// reference bit for bit).  Threads map x-fastest: a 64-lane wavefront covers
// 64 consecutive j of one (k, l) row, so every array access of a wave is one
// coalesced 512-byte segment, and stencil neighbours in j come from the same
// or adjacent cache lines.  All 3-D fields share one stride set (layout.h),
// so one offset per thread addresses every array.
#include <hip/hip_runtime.h>

#include "kernels.h"
#include "launch.h"

namespace clover {

thread_local KernelTimer *t_kernel_timer = nullptr;

// ---------------------------------------------------------------------------
// ideal_gas_kernel.f90:47-58
__global__ void k_ideal_gas(Lay g, Range r, const double *__restrict__ density, const double *__restrict__ energy,
                            double *__restrict__ pressure, double *__restrict__ soundspeed)
{
    RANGE_THREAD(r);
    const int i = idx(g, j, k, l);
    const double d = density[i];
    const double v = 1.0 / d;
    const double p = (1.4 - 1.0) * d * energy[i];
    pressure[i] = p;
    const double pressurebyenergy = (1.4 - 1.0) * d;
    const double pressurebyvolume = -d * p;
    const double sound_speed_squared = v * v * (p * pressurebyenergy - pressurebyvolume);
    soundspeed[i] = sqrt(sound_speed_squared);
}

hipError_t launch_ideal_gas(const Lay &g, const double *density, const double *energy, double *pressure,
                            double *soundspeed, hipStream_t st)
{
    Range r{g.x_min, g.x_max, g.y_min, g.y_max, g.z_min, g.z_max};
    LAUNCH(k_ideal_gas, r, st, g, r, density, energy, pressure, soundspeed);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// viscosity_kernel.f90:71-134
__global__ void k_viscosity(Lay g, Range r, const double *__restrict__ xarea, const double *__restrict__ yarea,
                            const double *__restrict__ zarea, const double *__restrict__ celldx,
                            const double *__restrict__ celldy, const double *__restrict__ celldz,
                            const double *__restrict__ density0, const double *__restrict__ pressure,
                            double *__restrict__ viscosity, const double *__restrict__ xvel0,
                            const double *__restrict__ yvel0, const double *__restrict__ zvel0)
{
    RANGE_THREAD(r);
    const int i = idx(g, j, k, l);
#define U(a, b, c) AT(xvel0, a, b, c)
#define V(a, b, c) AT(yvel0, a, b, c)
#define W(a, b, c) AT(zvel0, a, b, c)
    const double ugradx1 = U(0, 0, 0) + U(0, 1, 0) + U(0, 0, 1) + U(0, 1, 1);
    const double ugradx2 = U(1, 0, 0) + U(1, 1, 0) + U(1, 0, 1) + U(1, 1, 1);
    const double ugrady1 = U(0, 0, 0) + U(1, 0, 0) + U(0, 0, 1) + U(1, 0, 1);
    const double ugrady2 = U(0, 1, 0) + U(1, 1, 0) + U(0, 1, 1) + U(1, 1, 1);
    const double ugradz1 = U(0, 0, 0) + U(1, 0, 0) + U(0, 1, 0) + U(1, 1, 0);
    const double ugradz2 = U(0, 0, 1) + U(1, 0, 1) + U(0, 1, 1) + U(1, 1, 1);

    const double vgradx1 = V(0, 0, 0) + V(0, 1, 0) + V(0, 0, 1) + V(0, 1, 1);
    const double vgradx2 = V(1, 0, 0) + V(1, 1, 0) + V(1, 0, 1) + V(1, 1, 1);
    const double vgrady1 = V(0, 0, 0) + V(1, 0, 0) + V(0, 0, 1) + V(1, 0, 1);
    const double vgrady2 = V(0, 1, 0) + V(1, 1, 0) + V(0, 1, 1) + V(1, 1, 1);
    const double vgradz1 = V(0, 0, 0) + V(1, 0, 0) + V(0, 1, 0) + V(1, 1, 0);
    const double vgradz2 = V(0, 0, 1) + V(1, 0, 1) + V(0, 1, 1) + V(1, 1, 1);

    const double wgradx1 = W(0, 0, 0) + W(0, 1, 0) + W(0, 0, 1) + W(0, 1, 1);
    const double wgradx2 = W(1, 0, 0) + W(1, 1, 0) + W(1, 0, 1) + W(1, 1, 1);
    const double wgrady1 = W(0, 0, 0) + W(1, 0, 0) + W(0, 0, 1) + W(1, 0, 1);
    const double wgrady2 = W(0, 1, 0) + W(1, 1, 0) + W(0, 1, 1) + W(1, 1, 1);
    const double wgradz1 = W(0, 0, 0) + W(1, 0, 0) + W(0, 1, 0) + W(1, 1, 0);
    const double wgradz2 = W(0, 0, 1) + W(1, 0, 1) + W(0, 1, 1) + W(1, 1, 1);
#undef U
#undef V
#undef W
    const double cdx = celldx[j - g.x_min + 2], cdx1 = celldx[j - g.x_min + 3];
    const double cdy = celldy[k - g.y_min + 2], cdy1 = celldy[k - g.y_min + 3];
    const double cdz = celldz[l - g.z_min + 2], cdz1 = celldz[l - g.z_min + 3];

    const double div = (xarea[i] * (ugradx2 - ugradx1) + yarea[i] * (vgrady2 - vgrady1)) + zarea[i] * (wgradz2 - wgradz1);

    const double xx = 0.25 * (ugradx2 - ugradx1) / (cdx);
    const double yy = 0.25 * (vgrady2 - vgrady1) / (cdy);
    const double zz = 0.25 * (wgradz2 - wgradz1) / (cdz);
    const double xy = 0.25 * (ugrady2 - ugrady1) / (cdy) + 0.25 * (vgradx2 - vgradx1) / (cdx);
    const double xz = 0.25 * (ugradz2 - ugradz1) / (cdz) + 0.25 * (wgradx2 - wgradx1) / (cdx);
    const double yz = 0.25 * (vgradz2 - vgradz1) / (cdz) + 0.25 * (wgrady2 - wgrady1) / (cdy);

    double pgradx = (AT(pressure, 1, 0, 0) - AT(pressure, -1, 0, 0)) / (cdx + cdx1);
    double pgrady = (AT(pressure, 0, 1, 0) - AT(pressure, 0, -1, 0)) / (cdy + cdy1);
    double pgradz = (AT(pressure, 0, 0, 1) - AT(pressure, 0, 0, -1)) / (cdz + cdz1);

    const double pgradx2 = pgradx * pgradx;
    const double pgrady2 = pgrady * pgrady;
    const double pgradz2 = pgradz * pgradz;

    const double limiter = (xx * pgradx2 + yy * pgrady2 + zz * pgradz2 + xy * pgradx * pgrady + xz * pgradx * pgradz +
                            yz * pgrady * pgradz) /
                           dmax(pgradx2 + pgrady2 + pgradz2, 1.0e-16);

    if ((limiter > 0.0) || (div >= 0.0)) {
        viscosity[i] = 0.0;
    } else {
        pgradx = copysign(dmax(1.0e-16, fabs(pgradx)), pgradx);
        pgrady = copysign(dmax(1.0e-16, fabs(pgrady)), pgrady);
        pgradz = copysign(dmax(1.0e-16, fabs(pgradz)), pgradz);
        const double pgrad = sqrt(pgradx * pgradx + pgrady * pgrady + pgradz * pgradz);
        const double xgrad = fabs(cdx * pgrad / pgradx);
        const double ygrad = fabs(cdy * pgrad / pgrady);
        const double zgrad = fabs(cdz * pgrad / pgradz);
        const double grad = dmin(dmin(xgrad, ygrad), zgrad);
        const double grad2 = grad * grad;
        viscosity[i] = 2.0 * density0[i] * grad2 * limiter * limiter;
    }
}

hipError_t launch_viscosity(const Lay &g, const double *xarea, const double *yarea, const double *zarea,
                            const double *celldx, const double *celldy, const double *celldz,
                            const double *density0, const double *pressure, double *viscosity,
                            const double *xvel0, const double *yvel0, const double *zvel0, hipStream_t st)
{
    Range r{g.x_min, g.x_max, g.y_min, g.y_max, g.z_min, g.z_max};
    LAUNCH(k_viscosity, r, st, g, r, xarea, yarea, zarea, celldx, celldy, celldz, density0, pressure, viscosity,
           xvel0, yvel0, zvel0);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// calc_dt_kernel.f90:102-149.  Per-cell dt, then a deterministic two-pass
// minimum: each block reduces its cells (wavefront shuffle, then LDS) into
// partial[block]; a one-block kernel reduces the partials in a fixed order.
__global__ void __launch_bounds__(RED_THREADS)
k_calc_dt(Lay g, int nj, int nk, int ncells, double dtc_safe, double dtu_safe, double dtv_safe, double dtw_safe,
          double dtdiv_safe, double g_small, const double *__restrict__ xarea, const double *__restrict__ yarea,
          const double *__restrict__ zarea, const double *__restrict__ celldx, const double *__restrict__ celldy,
          const double *__restrict__ celldz, const double *__restrict__ volume, const double *__restrict__ density0,
          const double *__restrict__ viscosity_a, const double *__restrict__ soundspeed,
          const double *__restrict__ xvel0, const double *__restrict__ yvel0, const double *__restrict__ zvel0,
          double *__restrict__ dt_min, double *__restrict__ partial)
{
    double m = 1.0e300;
    for (int c = blockIdx.x * RED_THREADS + threadIdx.x; c < ncells; c += gridDim.x * RED_THREADS) {
        const int j = g.x_min + c % nj;
        const int k = g.y_min + (c / nj) % nk;
        const int l = g.z_min + c / (nj * nk);
        const int i = idx(g, j, k, l);
        const double cdx = celldx[j - g.x_min + 2], cdy = celldy[k - g.y_min + 2], cdz = celldz[l - g.z_min + 2];
        const double mm = dmin(dmin(cdx, cdy), cdz);
        const double ds = 1.0 / (mm * mm);

        double cc = soundspeed[i] * soundspeed[i];
        cc = cc + 2.0 * viscosity_a[i] / density0[i];

        double dtct = ds * cc;
        dtct = dtc_safe * 1.0 / dmax(sqrt(dtct), g_small);

        const double xa = xarea[i];
        const double du1 = (AT(xvel0, 0, 0, 0) + AT(xvel0, 0, 1, 0) + AT(xvel0, 0, 0, 1) + AT(xvel0, 0, 1, 1)) * xa;
        const double du2 = (AT(xvel0, 1, 0, 0) + AT(xvel0, 1, 1, 0) + AT(xvel0, 1, 0, 1) + AT(xvel0, 1, 1, 1)) * xa;
        const double vol = volume[i];
        const double dtut = dtu_safe * 4.0 * vol / dmax(dmax(fabs(du1), fabs(du2)), 1.0e-5 * vol);

        const double ya = yarea[i];
        const double dv1 = (AT(yvel0, 0, 0, 0) + AT(yvel0, 1, 0, 0) + AT(yvel0, 0, 0, 1) + AT(yvel0, 1, 0, 1)) * ya;
        const double dv2 = (AT(yvel0, 0, 1, 0) + AT(yvel0, 1, 1, 0) + AT(yvel0, 0, 1, 1) + AT(yvel0, 1, 1, 1)) * ya;
        const double dtvt = dtv_safe * 4.0 * vol / dmax(dmax(fabs(dv1), fabs(dv2)), 1.0e-5 * vol);

        const double za = zarea[i];
        const double dw1 = (AT(zvel0, 0, 0, 0) + AT(zvel0, 0, 1, 0) + AT(zvel0, 1, 0, 0) + AT(zvel0, 1, 1, 0)) * za;
        const double dw2 = (AT(zvel0, 0, 0, 1) + AT(zvel0, 0, 1, 1) + AT(zvel0, 1, 0, 1) + AT(zvel0, 1, 1, 1)) * za;
        const double dtwt = dtw_safe * 4.0 * vol / dmax(dmax(fabs(dw1), fabs(dw2)), 1.0e-5 * vol);

        const double div = du2 - du1 + dv2 - dv1 + dw2 - dw1;
        const double dtdivt = dtdiv_safe * 4.0 * (vol) / dmax(vol * 1.0e-05, fabs(div));

        const double cell = dmin(dmin(dmin(dmin(dtct, dtut), dtvt), dtwt), dtdivt);
        if (dt_min) dt_min[i] = cell;
        m = dmin(m, cell);
    }
    __shared__ double red[RED_THREADS / 64];
    m = wave_min(m);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double b = red[0];
        for (int w = 1; w < RED_THREADS / 64; w++) b = dmin(b, red[w]);
        partial[blockIdx.x] = b;
    }
}

static int reduction_blocks(int ncells)
{
    int b = (ncells + RED_THREADS * 4 - 1) / (RED_THREADS * 4);
    if (b < 1) b = 1;
    if (b > kMaxPartials) b = kMaxPartials;
    return b;
}

hipError_t launch_calc_dt_cells(const Lay &g, double dtc_safe, double dtu_safe, double dtv_safe, double dtw_safe,
                                double dtdiv_safe, double g_small, const double *xarea, const double *yarea,
                                const double *zarea, const double *celldx, const double *celldy,
                                const double *celldz, const double *volume, const double *density0,
                                const double *viscosity_a, const double *soundspeed, const double *xvel0,
                                const double *yvel0, const double *zvel0, double *dt_min, double *partial,
                                int *nparts, hipStream_t st)
{
    const int nj = g.x_max - g.x_min + 1, nk = g.y_max - g.y_min + 1, nl = g.z_max - g.z_min + 1;
    const int ncells = nj * nk * nl;
    const int nb = reduction_blocks(ncells);
    *nparts = nb;
    TIMED_LAUNCH("k_calc_dt", st, k_calc_dt, dim3(nb), dim3(RED_THREADS), 0, st, g, nj, nk, ncells, dtc_safe, dtu_safe, dtv_safe,
                       dtw_safe, dtdiv_safe, g_small, xarea, yarea, zarea, celldx, celldy, celldz, volume, density0,
                       viscosity_a, soundspeed, xvel0, yvel0, zvel0, dt_min, partial);
    return hipGetLastError();
}

// One workgroup; each thread keeps four independent minima so that four
// loads are in flight at a time.  The minimum does not depend on the order.
constexpr int MIN_THREADS = 1024;

__global__ void __launch_bounds__(MIN_THREADS) k_min_final(const double *__restrict__ partial, int n, double seed,
                                                           double *__restrict__ out)
{
    double m0 = seed, m1 = seed, m2 = seed, m3 = seed;
    int c = threadIdx.x;
    for (; c + 3 * MIN_THREADS < n; c += 4 * MIN_THREADS) {
        const double a = partial[c], b = partial[c + MIN_THREADS], d = partial[c + 2 * MIN_THREADS],
                     e = partial[c + 3 * MIN_THREADS];
        m0 = dmin(m0, a);
        m1 = dmin(m1, b);
        m2 = dmin(m2, d);
        m3 = dmin(m3, e);
    }
    for (; c < n; c += MIN_THREADS) m0 = dmin(m0, partial[c]);
    double m = dmin(dmin(m0, m1), dmin(m2, m3));
    __shared__ double red[MIN_THREADS / 64];
    m = wave_min(m);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double b = red[0];
        for (int w = 1; w < MIN_THREADS / 64; w++) b = dmin(b, red[w]);
        *out = b;
    }
}

hipError_t launch_min_final(const double *partial, int n, double seed, double *out, hipStream_t st)
{
    TIMED_LAUNCH("k_min_final", st, k_min_final, dim3(1), dim3(MIN_THREADS), 0, st, partial, n, seed, out);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// PdV_kernel.f90:76-165.  The predictor sums the level-0 face velocities
// twice and halves dt; the corrector sums level 0 then level 1.
template <int PREDICT>
__global__ void k_pdv(Lay g, Range r, double dt, const double *__restrict__ xarea, const double *__restrict__ yarea,
                      const double *__restrict__ zarea, const double *__restrict__ volume,
                      const double *__restrict__ density0, double *__restrict__ density1,
                      const double *__restrict__ energy0, double *__restrict__ energy1,
                      const double *__restrict__ pressure, const double *__restrict__ viscosity,
                      const double *__restrict__ xvel0, const double *__restrict__ xb,
                      const double *__restrict__ yvel0, const double *__restrict__ yb,
                      const double *__restrict__ zvel0, const double *__restrict__ zb)
{
    RANGE_THREAD(r);
    const int i = idx(g, j, k, l);
    const double sl = AT(xvel0, 0, 0, 0) + AT(xvel0, 0, 1, 0) + AT(xvel0, 0, 0, 1) + AT(xvel0, 0, 1, 1) +
                      AT(xb, 0, 0, 0) + AT(xb, 0, 1, 0) + AT(xb, 0, 0, 1) + AT(xb, 0, 1, 1);
    const double sr = AT(xvel0, 1, 0, 0) + AT(xvel0, 1, 1, 0) + AT(xvel0, 1, 0, 1) + AT(xvel0, 1, 1, 1) +
                      AT(xb, 1, 0, 0) + AT(xb, 1, 1, 0) + AT(xb, 1, 0, 1) + AT(xb, 1, 1, 1);
    const double sb = AT(yvel0, 0, 0, 0) + AT(yvel0, 1, 0, 0) + AT(yvel0, 0, 0, 1) + AT(yvel0, 1, 0, 1) +
                      AT(yb, 0, 0, 0) + AT(yb, 1, 0, 0) + AT(yb, 0, 0, 1) + AT(yb, 1, 0, 1);
    const double stp = AT(yvel0, 0, 1, 0) + AT(yvel0, 1, 1, 0) + AT(yvel0, 0, 1, 1) + AT(yvel0, 1, 1, 1) +
                       AT(yb, 0, 1, 0) + AT(yb, 1, 1, 0) + AT(yb, 0, 1, 1) + AT(yb, 1, 1, 1);
    const double sk = AT(zvel0, 0, 0, 0) + AT(zvel0, 1, 0, 0) + AT(zvel0, 0, 1, 0) + AT(zvel0, 1, 1, 0) +
                      AT(zb, 0, 0, 0) + AT(zb, 1, 0, 0) + AT(zb, 0, 1, 0) + AT(zb, 1, 1, 0);
    const double sf = AT(zvel0, 0, 0, 1) + AT(zvel0, 1, 0, 1) + AT(zvel0, 0, 1, 1) + AT(zvel0, 1, 1, 1) +
                      AT(zb, 0, 0, 1) + AT(zb, 1, 0, 1) + AT(zb, 0, 1, 1) + AT(zb, 1, 1, 1);
    double left_flux, right_flux, bottom_flux, top_flux, back_flux, front_flux;
    if (PREDICT) {
        left_flux = (AT(xarea, 0, 0, 0) * sl) * 0.125 * dt * 0.5;
        right_flux = (AT(xarea, 1, 0, 0) * sr) * 0.125 * dt * 0.5;
        bottom_flux = (AT(yarea, 0, 0, 0) * sb) * 0.125 * dt * 0.5;
        top_flux = (AT(yarea, 0, 1, 0) * stp) * 0.125 * dt * 0.5;
        back_flux = (AT(zarea, 0, 0, 0) * sk) * 0.125 * dt * 0.5;
        front_flux = (AT(zarea, 0, 0, 1) * sf) * 0.125 * dt * 0.5;
    } else {
        left_flux = (AT(xarea, 0, 0, 0) * sl) * 0.125 * dt;
        right_flux = (AT(xarea, 1, 0, 0) * sr) * 0.125 * dt;
        bottom_flux = (AT(yarea, 0, 0, 0) * sb) * 0.125 * dt;
        top_flux = (AT(yarea, 0, 1, 0) * stp) * 0.125 * dt;
        back_flux = (AT(zarea, 0, 0, 0) * sk) * 0.125 * dt;
        front_flux = (AT(zarea, 0, 0, 1) * sf) * 0.125 * dt;
    }
    const double total_flux = right_flux - left_flux + top_flux - bottom_flux + front_flux - back_flux;
    const double vol = volume[i];
    const double vol_change_s = vol / (vol + total_flux);
    const double recip_volume = 1.0 / vol;
    const double d0 = density0[i];
    const double energy_change = (pressure[i] / d0 + viscosity[i] / d0) * total_flux * recip_volume;
    energy1[i] = energy0[i] - energy_change;
    density1[i] = d0 * vol_change_s;
}

hipError_t launch_pdv(int predict, const Lay &g, double dt, const double *xarea, const double *yarea,
                      const double *zarea, const double *volume, const double *density0, double *density1,
                      const double *energy0, double *energy1, const double *pressure, const double *viscosity,
                      const double *xvel0, const double *xvel1, const double *yvel0, const double *yvel1,
                      const double *zvel0, const double *zvel1, hipStream_t st)
{
    Range r{g.x_min, g.x_max, g.y_min, g.y_max, g.z_min, g.z_max};
    if (predict)
        LAUNCH(k_pdv<1>, r, st, g, r, dt, xarea, yarea, zarea, volume, density0, density1, energy0, energy1, pressure,
               viscosity, xvel0, xvel0, yvel0, yvel0, zvel0, zvel0);
    else
        LAUNCH(k_pdv<0>, r, st, g, r, dt, xarea, yarea, zarea, volume, density0, density1, energy0, energy1, pressure,
               viscosity, xvel0, xvel1, yvel0, yvel1, zvel0, zvel1);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// revert_kernel.f90:44-61
__global__ void k_revert(Lay g, Range r, const double *__restrict__ density0, double *__restrict__ density1,
                         const double *__restrict__ energy0, double *__restrict__ energy1)
{
    RANGE_THREAD(r);
    const int i = idx(g, j, k, l);
    density1[i] = density0[i];
    energy1[i] = energy0[i];
}

hipError_t launch_revert(const Lay &g, const double *density0, double *density1, const double *energy0,
                         double *energy1, hipStream_t st)
{
    Range r{g.x_min, g.x_max, g.y_min, g.y_max, g.z_min, g.z_max};
    LAUNCH(k_revert, r, st, g, r, density0, density1, energy0, energy1);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// accelerate_kernel.f90:61-109 (nodes min..max+1)
__global__ void k_accelerate(Lay g, Range r, double dt, const double *__restrict__ xarea,
                             const double *__restrict__ yarea, const double *__restrict__ zarea,
                             const double *__restrict__ volume, const double *__restrict__ density0,
                             const double *__restrict__ pressure, const double *__restrict__ viscosity,
                             const double *__restrict__ xvel0, const double *__restrict__ yvel0,
                             const double *__restrict__ zvel0, double *__restrict__ xvel1, double *__restrict__ yvel1,
                             double *__restrict__ zvel1)
{
    RANGE_THREAD(r);
    const int i = idx(g, j, k, l);
#define RV(a, b, c) (AT(density0, a, b, c) * AT(volume, a, b, c))
    const double nodal_mass = (RV(-1, -1, 0) + RV(0, -1, 0) + RV(0, 0, 0) + RV(-1, 0, 0) + RV(-1, -1, -1) +
                               RV(0, -1, -1) + RV(0, 0, -1) + RV(-1, 0, -1)) *
                              0.125;
#undef RV
    const double stepbymass_s = 0.25 * dt / nodal_mass;
#define P(a, b, c) AT(pressure, a, b, c)
#define Q(a, b, c) AT(viscosity, a, b, c)
    const double xvel1_s = AT(xvel0, 0, 0, 0) - stepbymass_s * (AT(xarea, 0, 0, 0) * (P(0, 0, 0) - P(-1, 0, 0)) +
                                                                AT(xarea, 0, -1, 0) * (P(0, -1, 0) - P(-1, -1, 0)) +
                                                                AT(xarea, 0, 0, -1) * (P(0, 0, -1) - P(-1, 0, -1)) +
                                                                AT(xarea, 0, -1, -1) * (P(0, -1, -1) - P(-1, -1, -1)));
    const double yvel1_s = AT(yvel0, 0, 0, 0) - stepbymass_s * (AT(yarea, 0, 0, 0) * (P(0, 0, 0) - P(0, -1, 0)) +
                                                                AT(yarea, -1, 0, 0) * (P(-1, 0, 0) - P(-1, -1, 0)) +
                                                                AT(yarea, 0, 0, -1) * (P(0, 0, -1) - P(0, -1, -1)) +
                                                                AT(yarea, -1, 0, -1) * (P(-1, 0, -1) - P(-1, -1, -1)));
    const double zvel1_s = AT(zvel0, 0, 0, 0) - stepbymass_s * (AT(zarea, 0, 0, 0) * (P(0, 0, 0) - P(0, 0, -1)) +
                                                                AT(zarea, 0, -1, 0) * (P(0, -1, 0) - P(0, -1, -1)) +
                                                                AT(zarea, -1, 0, 0) * (P(-1, 0, 0) - P(-1, 0, -1)) +
                                                                AT(zarea, -1, -1, 0) * (P(-1, -1, 0) - P(-1, -1, -1)));
    xvel1[i] = xvel1_s - stepbymass_s * (AT(xarea, 0, 0, 0) * (Q(0, 0, 0) - Q(-1, 0, 0)) +
                                         AT(xarea, 0, -1, 0) * (Q(0, -1, 0) - Q(-1, -1, 0)) +
                                         AT(xarea, 0, 0, -1) * (Q(0, 0, -1) - Q(-1, 0, -1)) +
                                         AT(xarea, 0, -1, -1) * (Q(0, -1, -1) - Q(-1, -1, -1)));
    yvel1[i] = yvel1_s - stepbymass_s * (AT(yarea, 0, 0, 0) * (Q(0, 0, 0) - Q(0, -1, 0)) +
                                         AT(yarea, -1, 0, 0) * (Q(-1, 0, 0) - Q(-1, -1, 0)) +
                                         AT(yarea, 0, 0, -1) * (Q(0, 0, -1) - Q(0, -1, -1)) +
                                         AT(yarea, -1, 0, -1) * (Q(-1, 0, -1) - Q(-1, -1, -1)));
    zvel1[i] = zvel1_s - stepbymass_s * (AT(zarea, 0, 0, 0) * (Q(0, 0, 0) - Q(0, 0, -1)) +
                                         AT(zarea, 0, -1, 0) * (Q(0, -1, 0) - Q(0, -1, -1)) +
                                         AT(zarea, -1, 0, 0) * (Q(-1, 0, 0) - Q(-1, 0, -1)) +
                                         AT(zarea, -1, -1, 0) * (Q(-1, -1, 0) - Q(-1, -1, -1)));
#undef P
#undef Q
}

hipError_t launch_accelerate(const Lay &g, double dt, const double *xarea, const double *yarea,
                             const double *zarea, const double *volume, const double *density0,
                             const double *pressure, const double *viscosity, const double *xvel0,
                             const double *yvel0, const double *zvel0, double *xvel1, double *yvel1,
                             double *zvel1, hipStream_t st)
{
    Range r{g.x_min, g.x_max + 1, g.y_min, g.y_max + 1, g.z_min, g.z_max + 1};
    LAUNCH(k_accelerate, r, st, g, r, dt, xarea, yarea, zarea, volume, density0, pressure, viscosity, xvel0, yvel0,
           zvel0, xvel1, yvel1, zvel1);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// flux_calc_kernel.f90:56-90 over the union of the three face ranges; the
// z flux keeps the reference's repeated zvel0(j+1,k,l) term (line 86).
__global__ void k_flux_calc(Lay g, Range r, double dt, const double *__restrict__ xarea,
                            const double *__restrict__ yarea, const double *__restrict__ zarea,
                            const double *__restrict__ xvel0, const double *__restrict__ yvel0,
                            const double *__restrict__ zvel0, const double *__restrict__ xvel1,
                            const double *__restrict__ yvel1, const double *__restrict__ zvel1,
                            double *__restrict__ vol_flux_x, double *__restrict__ vol_flux_y,
                            double *__restrict__ vol_flux_z)
{
    RANGE_THREAD(r);
    const int i = idx(g, j, k, l);
    const bool jin = j <= g.x_max, kin = k <= g.y_max, lin = l <= g.z_max;
    if (kin && lin)
        vol_flux_x[i] = 0.125 * dt * AT(xarea, 0, 0, 0) *
                        (AT(xvel0, 0, 0, 0) + AT(xvel0, 0, 1, 0) + AT(xvel0, 0, 0, 1) + AT(xvel0, 0, 1, 1) +
                         AT(xvel1, 0, 0, 0) + AT(xvel1, 0, 1, 0) + AT(xvel1, 0, 0, 1) + AT(xvel1, 0, 1, 1));
    if (jin && lin)
        vol_flux_y[i] = 0.125 * dt * AT(yarea, 0, 0, 0) *
                        (AT(yvel0, 0, 0, 0) + AT(yvel0, 1, 0, 0) + AT(yvel0, 0, 0, 1) + AT(yvel0, 1, 0, 1) +
                         AT(yvel1, 0, 0, 0) + AT(yvel1, 1, 0, 0) + AT(yvel1, 0, 0, 1) + AT(yvel1, 1, 0, 1));
    if (jin && kin)
        vol_flux_z[i] = 0.125 * dt * AT(zarea, 0, 0, 0) *
                        (AT(zvel0, 0, 0, 0) + AT(zvel0, 1, 0, 0) + AT(zvel0, 1, 0, 0) + AT(zvel0, 1, 1, 0) +
                         AT(zvel1, 0, 0, 0) + AT(zvel1, 1, 0, 0) + AT(zvel1, 0, 1, 0) + AT(zvel1, 1, 1, 0));
}

hipError_t launch_flux_calc(const Lay &g, double dt, const double *xarea, const double *yarea,
                            const double *zarea, const double *xvel0, const double *yvel0, const double *zvel0,
                            const double *xvel1, const double *yvel1, const double *zvel1, double *vol_flux_x,
                            double *vol_flux_y, double *vol_flux_z, hipStream_t st)
{
    Range r{g.x_min, g.x_max + 1, g.y_min, g.y_max + 1, g.z_min, g.z_max + 1};
    LAUNCH(k_flux_calc, r, st, g, r, dt, xarea, yarea, zarea, xvel0, yvel0, zvel0, xvel1, yvel1, zvel1, vol_flux_x,
           vol_flux_y, vol_flux_z);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// advec_cell_kernel.f90:93-380
//
// (i) pre/post volumes on every cell including the 2-deep halo.  `mode`
// picks the formula: 1 x-sweep 1, 2 x-sweep 3, 3 y-sweep advect_x,
// 4 y-sweep not advect_x, 5 z-sweep 1, 6 z-sweep 3.
__global__ void k_cell_vols(Lay g, Range r, int mode, const double *__restrict__ volume,
                            const double *__restrict__ vol_flux_x, const double *__restrict__ vol_flux_y,
                            const double *__restrict__ vol_flux_z, double *__restrict__ pre_vol,
                            double *__restrict__ post_vol)
{
    RANGE_THREAD(r);
    const int i = idx(g, j, k, l);
    const double V0 = volume[i];
    const double x1 = AT(vol_flux_x, 1, 0, 0), x0 = vol_flux_x[i];
    const double y1 = AT(vol_flux_y, 0, 1, 0), y0 = vol_flux_y[i];
    const double z1 = AT(vol_flux_z, 0, 0, 1), z0 = vol_flux_z[i];
    double pre, post;
    switch (mode) {
    case 1: pre = V0 + (x1 - x0 + y1 - y0 + z1 - z0); post = pre - (x1 - x0); break;
    case 2: pre = V0 + x1 - x0; post = V0; break;
    case 3: pre = V0 + y1 - y0 + z1 - z0; post = pre - (y1 - y0); break;
    case 4: pre = V0 + y1 - y0 + x1 - x0; post = pre - (y1 - y0); break;
    case 5: pre = V0 + (x1 - x0 + y1 - y0 + z1 - z0); post = pre - (z1 - z0); break;
    default: pre = V0 + z1 - z0; post = V0; break;
    }
    pre_vol[i] = pre;
    post_vol[i] = post;
}

// The van Leer limited face value (advec_cell_kernel.f90:146-154, 158-166).
__device__ inline double cell_limiter(double sigma, double sigma3, double sigma4, double diffuw, double diffdw)
{
    const double one_by_six = 1.0 / 6.0;
    if (diffuw * diffdw > 0.0)
        return (1.0 - sigma) * copysign(1.0, diffdw) *
               dmin(dmin(fabs(diffuw), fabs(diffdw)), one_by_six * (sigma3 * fabs(diffuw) + sigma4 * fabs(diffdw)));
    return 0.0;
}

// (ii) mass and energy fluxes through the faces normal to the sweep axis.
// `s` is the element stride along the axis, c/cmax the axis coordinate and
// its max, vd the axis' vertexd* array (indexed from min-2), cmin0 = min-2.
__global__ void k_cell_flux(Lay g, Range r, int dir, const double *__restrict__ vd, const double *__restrict__ vol_flux,
                            const double *__restrict__ pre_vol, const double *__restrict__ density1,
                            const double *__restrict__ energy1, double *__restrict__ mass_flux,
                            double *__restrict__ ener_flux)
{
    RANGE_THREAD(r);
    const int i = idx(g, j, k, l);
    const int s = dir == 1 ? 1 : dir == 2 ? g.ld : g.plane;
    const int c = dir == 1 ? j : dir == 2 ? k : l;
    const int cmax = dir == 1 ? g.x_max : dir == 2 ? g.y_max : g.z_max;
    const int cmin0 = (dir == 1 ? g.x_min : dir == 2 ? g.y_min : g.z_min) - 2;
    const double vf = vol_flux[i];
    int upwind, donor, downwind, dif;  // as axis coordinates
    if (vf > 0.0) {
        upwind = c - 2; donor = c - 1; downwind = c; dif = donor;
    } else {
        upwind = c + 1 < cmax + 2 ? c + 1 : cmax + 2; donor = c; downwind = c - 1; dif = upwind;
    }
    const int iu = i + (upwind - c) * s, id = i + (donor - c) * s, iw = i + (downwind - c) * s;
    const double sigmat = fabs(vf) / pre_vol[id];
    const double sigma3 = (1.0 + sigmat) * (vd[c - cmin0] / vd[dif - cmin0]);
    const double sigma4 = 2.0 - sigmat;
    const double sigmav = sigmat;
    const double rd = density1[id];
    double diffuw = rd - density1[iu];
    double diffdw = density1[iw] - rd;
    double limiter = cell_limiter(sigmav, sigma3, sigma4, diffuw, diffdw);
    const double mf = vf * (rd + limiter);
    mass_flux[i] = mf;
    const double sigmam = fabs(mf) / (rd * pre_vol[id]);
    const double ed = energy1[id];
    diffuw = ed - energy1[iu];
    diffdw = energy1[iw] - ed;
    limiter = cell_limiter(sigmam, sigma3, sigma4, diffuw, diffdw);
    ener_flux[i] = mf * (ed + limiter);
}

// (iii) interior update
__global__ void k_cell_update(Lay g, Range r, int dir, const double *__restrict__ vol_flux,
                              const double *__restrict__ mass_flux, const double *__restrict__ ener_flux,
                              const double *__restrict__ pre_vol, double *__restrict__ density1,
                              double *__restrict__ energy1, double *__restrict__ pre_mass,
                              double *__restrict__ post_mass, double *__restrict__ advec_vol,
                              double *__restrict__ post_ener)
{
    RANGE_THREAD(r);
    const int i = idx(g, j, k, l);
    const int s = dir == 1 ? 1 : dir == 2 ? g.ld : g.plane;
    const double pv = pre_vol[i];
    const double pm = density1[i] * pv;
    const double pom = pm + mass_flux[i] - mass_flux[i + s];
    const double pe = (energy1[i] * pm + ener_flux[i] - ener_flux[i + s]) / pom;
    const double av = pv + vol_flux[i] - vol_flux[i + s];
    pre_mass[i] = pm;
    post_mass[i] = pom;
    post_ener[i] = pe;
    advec_vol[i] = av;
    density1[i] = pom / av;
    energy1[i] = pe;
}

hipError_t launch_advec_cell(const Lay &g, int advect_x, int dir, int sweep_number, const double *vertexdx,
                             const double *vertexdy, const double *vertexdz, const double *volume,
                             double *density1, double *energy1, double *mass_flux_x, const double *vol_flux_x,
                             double *mass_flux_y, const double *vol_flux_y, double *mass_flux_z,
                             const double *vol_flux_z, double *pre_vol, double *post_vol, double *pre_mass,
                             double *post_mass, double *advec_vol, double *post_ener, double *ener_flux,
                             hipStream_t st)
{
    int mode = 0;
    if (dir == 1) mode = sweep_number == 1 ? 1 : sweep_number == 3 ? 2 : 0;
    else if (dir == 2) mode = sweep_number == 2 ? (advect_x ? 3 : 4) : 0;
    else if (dir == 3) mode = sweep_number == 1 ? 5 : sweep_number == 3 ? 6 : 0;
    else return hipSuccess;
    Range all{g.x_min - 2, g.x_max + 2, g.y_min - 2, g.y_max + 2, g.z_min - 2, g.z_max + 2};
    if (mode) LAUNCH(k_cell_vols, all, st, g, all, mode, volume, vol_flux_x, vol_flux_y, vol_flux_z, pre_vol, post_vol);
    Range fr;
    const double *vd, *vf;
    double *mf;
    if (dir == 1) {
        fr = Range{g.x_min, g.x_max + 2, g.y_min, g.y_max, g.z_min, g.z_max};
        vd = vertexdx; vf = vol_flux_x; mf = mass_flux_x;
    } else if (dir == 2) {
        fr = Range{g.x_min, g.x_max, g.y_min, g.y_max + 2, g.z_min, g.z_max + 2};
        vd = vertexdy; vf = vol_flux_y; mf = mass_flux_y;
    } else {
        fr = Range{g.x_min, g.x_max, g.y_min, g.y_max, g.z_min, g.z_max + 2};
        vd = vertexdz; vf = vol_flux_z; mf = mass_flux_z;
    }
    LAUNCH(k_cell_flux, fr, st, g, fr, dir, vd, vf, pre_vol, density1, energy1, mf, ener_flux);
    Range in{g.x_min, g.x_max, g.y_min, g.y_max, g.z_min, g.z_max};
    LAUNCH(k_cell_update, in, st, g, in, dir, vf, mf, ener_flux, pre_vol, density1, energy1, pre_mass, post_mass,
           advec_vol, post_ener);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// advec_mom_kernel.f90:91-424
//
// (i) staggered volumes; mode as advec_mom_kernel.f90:95-163:
// 1 sweep1/x, 2 sweep1/z, 3 sweep2/advect_x, 4 sweep2/not, 5 sweep3/x, 6 sweep3/z
__global__ void k_mom_vols(Lay g, Range r, int mode, const double *__restrict__ volume,
                           const double *__restrict__ vol_flux_x, const double *__restrict__ vol_flux_y,
                           const double *__restrict__ vol_flux_z, double *__restrict__ pre_vol,
                           double *__restrict__ post_vol)
{
    RANGE_THREAD(r);
    const int i = idx(g, j, k, l);
    const double V0 = volume[i];
    const double x1 = AT(vol_flux_x, 1, 0, 0), x0 = vol_flux_x[i];
    const double y1 = AT(vol_flux_y, 0, 1, 0), y0 = vol_flux_y[i];
    const double z1 = AT(vol_flux_z, 0, 0, 1), z0 = vol_flux_z[i];
    double post, pre;
    switch (mode) {
    case 1: post = V0 + y1 - y0 + z1 - z0; pre = post + x1 - x0; break;
    case 2: post = V0 + x1 - x0 + y1 - y0; pre = post + z1 - z0; break;
    case 3: post = V0 + z1 - z0; pre = post + y1 - y0; break;
    case 4: post = V0 + x1 - x0; pre = post + y1 - y0; break;
    case 5: post = V0; pre = post + x1 - x0; break;
    default: post = V0; pre = post + z1 - z0; break;
    }
    post_vol[i] = post;
    pre_vol[i] = pre;
}

// (ii-a) node flux: 8-point average of the mass flux around the staggered face
__global__ void k_node_flux(Lay g, Range r, int dir, const double *__restrict__ mf, double *__restrict__ node_flux)
{
    RANGE_THREAD(r);
    const int i = idx(g, j, k, l);
    double v;
    if (dir == 1)
        v = 0.125 * (AT(mf, 0, -1, 0) + AT(mf, 0, 0, 0) + AT(mf, 1, -1, 0) + AT(mf, 1, 0, 0) + AT(mf, 0, -1, -1) +
                     AT(mf, 0, 0, -1) + AT(mf, 1, -1, -1) + AT(mf, 1, 0, -1));
    else if (dir == 2)
        v = 0.125 * (AT(mf, -1, 0, 0) + AT(mf, 0, 0, 0) + AT(mf, -1, 1, 0) + AT(mf, 0, 1, 0) + AT(mf, -1, 0, -1) +
                     AT(mf, 0, 0, -1) + AT(mf, -1, 1, -1) + AT(mf, 0, 1, -1));
    else
        v = 0.125 * (AT(mf, -1, 0, 0) + AT(mf, 0, 0, 0) + AT(mf, -1, 0, 1) + AT(mf, 0, 0, 1) + AT(mf, -1, -1, 0) +
                     AT(mf, 0, -1, 0) + AT(mf, -1, -1, 1) + AT(mf, 0, -1, 1));
    node_flux[i] = v;
}

// (ii-b) node masses after and before advection
__global__ void k_node_mass(Lay g, Range r, int dir, const double *__restrict__ density1,
                            const double *__restrict__ post_vol, const double *__restrict__ node_flux,
                            double *__restrict__ node_mass_post, double *__restrict__ node_mass_pre)
{
    RANGE_THREAD(r);
    const int i = idx(g, j, k, l);
#define NMP(a, b, c) (AT(density1, a, b, c) * AT(post_vol, a, b, c))
    const double nmp = 0.125 * (NMP(0, -1, 0) + NMP(0, 0, 0) + NMP(-1, -1, 0) + NMP(-1, 0, 0) + NMP(0, -1, -1) +
                                NMP(0, 0, -1) + NMP(-1, -1, -1) + NMP(-1, 0, -1));
#undef NMP
    const int s = dir == 1 ? 1 : dir == 2 ? g.ld : g.plane;
    node_mass_post[i] = nmp;
    node_mass_pre[i] = nmp - node_flux[i - s] + node_flux[i];
}

// (iii) momentum flux with the limited donor velocity
__global__ void k_mom_flux(Lay g, Range r, int dir, const double *__restrict__ cd, const double *__restrict__ node_flux,
                           const double *__restrict__ node_mass_pre, const double *__restrict__ vel1,
                           double *__restrict__ mom_flux)
{
    RANGE_THREAD(r);
    const int i = idx(g, j, k, l);
    const int s = dir == 1 ? 1 : dir == 2 ? g.ld : g.plane;
    const int c = dir == 1 ? j : dir == 2 ? k : l;
    const int cmin0 = (dir == 1 ? g.x_min : dir == 2 ? g.y_min : g.z_min) - 2;
    const double nf = node_flux[i];
    int upwind, donor, downwind, dif;
    if (nf < 0.0) {
        upwind = c + 2; donor = c + 1; downwind = c; dif = donor;
    } else {
        upwind = c - 1; donor = c; downwind = c + 1; dif = upwind;
    }
    const int id = i + (donor - c) * s;
    const double sigma = fabs(nf) / (node_mass_pre[id]);
    const double width = cd[c - cmin0];
    const double vdonor = vel1[id];
    const double vdiffuw = vdonor - vel1[i + (upwind - c) * s];
    const double vdiffdw = vel1[i + (downwind - c) * s] - vdonor;
    double limiter = 0.0;
    if (vdiffuw * vdiffdw > 0.0) {
        const double auw = fabs(vdiffuw);
        const double adw = fabs(vdiffdw);
        double wind = 1.0;
        if (vdiffdw <= 0.0) wind = -1.0;
        limiter = wind * dmin(dmin(width * ((2.0 - sigma) * adw / width + (1.0 + sigma) * auw / cd[dif - cmin0]) / 6.0,
                                   auw),
                              adw);
    }
    const double advec_vel_s = vdonor + (1.0 - sigma) * limiter;
    mom_flux[i] = advec_vel_s * nf;
}

// (iv) velocity update on the interior nodes
__global__ void k_mom_update(Lay g, Range r, int dir, const double *__restrict__ node_mass_pre,
                             const double *__restrict__ node_mass_post, const double *__restrict__ mom_flux,
                             double *__restrict__ vel1)
{
    RANGE_THREAD(r);
    const int i = idx(g, j, k, l);
    const int s = dir == 1 ? 1 : dir == 2 ? g.ld : g.plane;
    vel1[i] = (vel1[i] * node_mass_pre[i] + mom_flux[i - s] - mom_flux[i]) / node_mass_post[i];
}

hipError_t launch_advec_mom(const Lay &g, double *vel1, const double *mass_flux_x, const double *vol_flux_x,
                            const double *mass_flux_y, const double *vol_flux_y, const double *mass_flux_z,
                            const double *vol_flux_z, const double *volume, const double *density1,
                            double *node_flux, double *node_mass_post, double *node_mass_pre, double *mom_flux,
                            double *pre_vol, double *post_vol, const double *celldx, const double *celldy,
                            const double *celldz, int advect_x, int which_vel, int sweep_number, int direction,
                            hipStream_t st)
{
    int mode = 0;
    if (sweep_number == 1 && direction == 1) mode = 1;
    else if (sweep_number == 1 && direction == 3) mode = 2;
    else if (sweep_number == 2 && advect_x) mode = 3;
    else if (sweep_number == 2 && !advect_x) mode = 4;
    else if (sweep_number == 3 && direction == 1) mode = 5;
    else if (sweep_number == 3 && direction == 3) mode = 6;
    Range all{g.x_min - 2, g.x_max + 2, g.y_min - 2, g.y_max + 2, g.z_min - 2, g.z_max + 2};
    if (mode) LAUNCH(k_mom_vols, all, st, g, all, mode, volume, vol_flux_x, vol_flux_y, vol_flux_z, pre_vol, post_vol);
    if (direction < 1 || direction > 3) return hipGetLastError();
    const double *mf = direction == 1 ? mass_flux_x : direction == 2 ? mass_flux_y : mass_flux_z;
    const double *cd = direction == 1 ? celldx : direction == 2 ? celldy : celldz;
    Range rnf, rnm, rmf;
    if (direction == 1) {
        rnf = Range{g.x_min - 2, g.x_max + 2, g.y_min, g.y_max + 1, g.z_min, g.z_max + 1};
        rnm = Range{g.x_min - 1, g.x_max + 2, g.y_min, g.y_max + 1, g.z_min, g.z_max + 1};
        rmf = Range{g.x_min - 1, g.x_max + 1, g.y_min, g.y_max + 1, g.z_min, g.z_max + 1};
    } else if (direction == 2) {
        rnf = Range{g.x_min, g.x_max + 1, g.y_min - 2, g.y_max + 2, g.z_min, g.z_max + 1};
        rnm = Range{g.x_min, g.x_max + 1, g.y_min - 1, g.y_max + 2, g.z_min, g.z_max + 1};
        rmf = Range{g.x_min, g.x_max + 1, g.y_min - 1, g.y_max + 1, g.z_min, g.z_max + 1};
    } else {
        rnf = Range{g.x_min, g.x_max + 1, g.y_min, g.y_max + 1, g.z_min - 2, g.z_max + 2};
        rnm = Range{g.x_min, g.x_max + 1, g.y_min, g.y_max + 1, g.z_min - 1, g.z_max + 2};
        rmf = Range{g.x_min, g.x_max + 1, g.y_min, g.y_max + 1, g.z_min - 1, g.z_max + 1};
    }
    if (which_vel == 1) {
        LAUNCH(k_node_flux, rnf, st, g, rnf, direction, mf, node_flux);
        LAUNCH(k_node_mass, rnm, st, g, rnm, direction, density1, post_vol, node_flux, node_mass_post, node_mass_pre);
    }
    LAUNCH(k_mom_flux, rmf, st, g, rmf, direction, cd, node_flux, node_mass_pre, vel1, mom_flux);
    Range upd{g.x_min, g.x_max + 1, g.y_min, g.y_max + 1, g.z_min, g.z_max + 1};
    LAUNCH(k_mom_update, upd, st, g, upd, direction, node_mass_pre, node_mass_post, mom_flux, vel1);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// reset_field_kernel.f90:51-98 over the node range; cells are the subset
// j<=x_max, k<=y_max, l<=z_max.
__global__ void k_reset_field(Lay g, Range r, double *__restrict__ density0, const double *__restrict__ density1,
                              double *__restrict__ energy0, const double *__restrict__ energy1,
                              double *__restrict__ xvel0, const double *__restrict__ xvel1, double *__restrict__ yvel0,
                              const double *__restrict__ yvel1, double *__restrict__ zvel0,
                              const double *__restrict__ zvel1)
{
    RANGE_THREAD(r);
    const int i = idx(g, j, k, l);
    if (j <= g.x_max && k <= g.y_max && l <= g.z_max) {
        density0[i] = density1[i];
        energy0[i] = energy1[i];
    }
    xvel0[i] = xvel1[i];
    yvel0[i] = yvel1[i];
    zvel0[i] = zvel1[i];
}

hipError_t launch_reset_field(const Lay &g, double *density0, const double *density1, double *energy0,
                              const double *energy1, double *xvel0, const double *xvel1, double *yvel0,
                              const double *yvel1, double *zvel0, const double *zvel1, hipStream_t st)
{
    Range r{g.x_min, g.x_max + 1, g.y_min, g.y_max + 1, g.z_min, g.z_max + 1};
    LAUNCH(k_reset_field, r, st, g, r, density0, density1, energy0, energy1, xvel0, xvel1, yvel0, yvel1, zvel0, zvel1);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// field_summary_kernel.f90:51-78: block partial sums of vol, mass, ie, ke,
// press (two-pass, fixed order, so a run is reproducible bit for bit).  The
// cell arrays are addressed through g, the vertex velocities through gv (the
// same strides in the device layout; the reference's own vertex extents for
// dense Fortran arrays).
__global__ void __launch_bounds__(RED_THREADS)
k_field_summary(Lay g, Lay gv, int nj, int nk, int ncells, const double *__restrict__ volume,
                const double *__restrict__ density0, const double *__restrict__ energy0,
                const double *__restrict__ pressure, const double *__restrict__ xvel0,
                const double *__restrict__ yvel0, const double *__restrict__ zvel0, double *__restrict__ partial)
{
    double s[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int c = blockIdx.x * RED_THREADS + threadIdx.x; c < ncells; c += gridDim.x * RED_THREADS) {
        const int j = g.x_min + c % nj;
        const int k = g.y_min + (c / nj) % nk;
        const int l = g.z_min + c / (nj * nk);
        const int i = idx(g, j, k, l);
        const int iv = idx(gv, j, k, l);
        double vsqrd = 0.0;
        for (int lv = 0; lv <= 1; lv++)
            for (int kv = 0; kv <= 1; kv++)
                for (int jv = 0; jv <= 1; jv++) {
                    const int n = iv + jv + kv * gv.ld + lv * gv.plane;
                    const double u = xvel0[n], v = yvel0[n], w = zvel0[n];
                    vsqrd = vsqrd + 0.125 * (u * u + v * v + w * w);
                }
        const double cell_vol = volume[i];
        const double cell_mass = cell_vol * density0[i];
        s[0] += cell_vol;
        s[1] += cell_mass;
        s[2] += cell_mass * energy0[i];
        s[3] += cell_mass * 0.5 * vsqrd;
        s[4] += cell_vol * pressure[i];
    }
    __shared__ double red[5][RED_THREADS / 64];
    for (int q = 0; q < 5; q++) {
        const double w = wave_sum(s[q]);
        if ((threadIdx.x & 63) == 0) red[q][threadIdx.x >> 6] = w;
    }
    __syncthreads();
    if (threadIdx.x < 5) {
        double b = 0.0;
        for (int w = 0; w < RED_THREADS / 64; w++) b += red[threadIdx.x][w];
        partial[5 * blockIdx.x + threadIdx.x] = b;
    }
}

hipError_t launch_field_summary_partial(const Lay &g, const Lay &gv, const double *volume, const double *density0,
                                        const double *energy0, const double *pressure, const double *xvel0,
                                        const double *yvel0, const double *zvel0, double *partial, int *nparts,
                                        hipStream_t st)
{
    const int nj = g.x_max - g.x_min + 1, nk = g.y_max - g.y_min + 1, nl = g.z_max - g.z_min + 1;
    const int ncells = nj * nk * nl;
    const int nb = reduction_blocks(ncells);
    *nparts = nb;
    TIMED_LAUNCH("k_field_summary", st, k_field_summary, dim3(nb), dim3(RED_THREADS), 0, st, g, gv, nj, nk, ncells,
                 volume, density0, energy0, pressure, xvel0, yvel0, zvel0, partial);
    return hipGetLastError();
}

hipError_t launch_field_summary_partial(const Lay &g, const double *volume, const double *density0,
                                        const double *energy0, const double *pressure, const double *xvel0,
                                        const double *yvel0, const double *zvel0, double *partial, int *nparts,
                                        hipStream_t st)
{
    return launch_field_summary_partial(g, g, volume, density0, energy0, pressure, xvel0, yvel0, zvel0, partial,
                                        nparts, st);
}

__global__ void __launch_bounds__(RED_THREADS) k_sum5_final(const double *__restrict__ partial, int n,
                                                            double *__restrict__ out)
{
    double s[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int c = threadIdx.x; c < n; c += RED_THREADS)
        for (int q = 0; q < 5; q++) s[q] += partial[5 * c + q];
    __shared__ double red[5][RED_THREADS / 64];
    for (int q = 0; q < 5; q++) {
        const double w = wave_sum(s[q]);
        if ((threadIdx.x & 63) == 0) red[q][threadIdx.x >> 6] = w;
    }
    __syncthreads();
    if (threadIdx.x < 5) {
        double b = 0.0;
        for (int w = 0; w < RED_THREADS / 64; w++) b += red[threadIdx.x][w];
        out[threadIdx.x] = b;
    }
}

hipError_t launch_sum5_final(const double *partial, int n, double *out, hipStream_t st)
{
    TIMED_LAUNCH("k_sum5_final", st, k_sum5_final, dim3(1), dim3(RED_THREADS), 0, st, partial, n, out);
    return hipGetLastError();
}

}  // namespace clover
