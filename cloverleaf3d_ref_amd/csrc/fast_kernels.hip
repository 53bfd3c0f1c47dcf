// Fused kernels of the resident driver's fast path (see fast.h), for both
// store forms.
//
// In the FP64 form each kernel evaluates exactly the expressions of the
// reference loop nests it replaces, in the same order and without FMA
// contraction, so its outputs are bit-identical to the per-call kernels'
// (divisions by a mesh constant use launch.h by_const, whose FMA sequence
// returns the correctly rounded quotient, the same bits as the division);
// the savings are in HBM bytes: scalar geometry for the uniform mesh, fused
// neighbours, dead stores dropped, intermediates recomputed in registers.
// The FP32 form runs the same expressions in FP32 (see by_const / quo); the
// dt minimum and the summary sums stay FP64.
//
// The stencils here read two neighbouring planes (l and l+1 for cells and
// faces, l-1 and l for nodes).  Each thread owns one (j, k) column and
// marches up a chunk of planes, keeping the lower plane's values in
// registers (or, in accel_pdv, in an LDS ring shared by the workgroup), so
// every plane is read from memory once per column instead of once per
// plane that uses it.
#include <hip/hip_runtime.h>

#include "batch.h"
#include "fast.h"
#include "launch.h"

namespace clover {

// 2x2 corner values of an array around (j, k) in one plane, in the storage
// type (which is also the arithmetic type of the form):
//   quad_up:   v[a + 2b] = p(j+a, k+b)           (a, b in {0, 1})
//   quad_down: v[(a+1) + 2(b+1)] = p(j+a, k+b)   (a, b in {-1, 0})
template <typename S>
struct QuadS {
    S v[4];
};

template <typename S>
__device__ inline QuadS<S> quad_up_raw(const S *__restrict__ p, int i, int ld)
{
    return QuadS<S>{{p[i], p[i + 1], p[i + ld], p[i + 1 + ld]}};
}

template <typename S>
__device__ inline QuadS<S> quad_down_raw(const S *__restrict__ p, int i, int ld)
{
    return QuadS<S>{{p[i - 1 - ld], p[i - ld], p[i - 1], p[i]}};
}

// A quad_down_raw quad at mirrored faces (Lay::mirror): at the x_min face
// (xlo) the halo column j-1 takes the values of column j, which its
// reflection copies, at the x_max face (xhi) column j those of j-1; the
// same for the rows k-1 and k at the y faces.  The halo values were loaded
// (in bounds) and are replaced where the quad is used, so the loads keep
// their constant offsets and stay in flight.
template <typename S>
__device__ inline QuadS<S> quad_m(QuadS<S> q, bool xlo, bool xhi, bool ylo, bool yhi)
{
    if (xlo) {
        q.v[0] = q.v[1];
        q.v[2] = q.v[3];
    }
    if (xhi) {
        q.v[1] = q.v[0];
        q.v[3] = q.v[2];
    }
    if (ylo) {
        q.v[0] = q.v[2];
        q.v[1] = q.v[3];
    }
    if (yhi) {
        q.v[2] = q.v[0];
        q.v[3] = q.v[1];
    }
    return q;
}

// The FP32 forms of visc_dt and the PdV predictor take two cells (j, j+1,
// j even from x_min) per thread: the node velocities of columns j, j+1 come
// in one 8-byte load per row and column j+2 in a 4-byte one, and the cell
// pairs in 8-byte loads and stores.  quad_a and quad_b are the two cells'
// quad_up_raw quads.  The pairs are 8-byte aligned: x_min sits on a
// 32-element boundary of every row (layout.h) and every field array starts
// on a 256-byte boundary (its own allocation in a chunk, a 256-byte rounded
// slot of the ensemble's arena).
struct NodePair {
    float2 r0, r1;  // nodes j, j+1 of rows k and k+1
    float n0, n1;   // node j+2 of rows k and k+1
};

__device__ inline NodePair load_node_pair(const float *__restrict__ p, int i, int ld)
{
    return NodePair{*reinterpret_cast<const float2 *>(p + i), *reinterpret_cast<const float2 *>(p + i + ld), p[i + 2],
                    p[i + ld + 2]};
}

__device__ inline QuadS<float> quad_a(const NodePair &n) { return QuadS<float>{{n.r0.x, n.r0.y, n.r1.x, n.r1.y}}; }
__device__ inline QuadS<float> quad_b(const NodePair &n) { return QuadS<float>{{n.r0.y, n.n0, n.r1.y, n.n1}}; }

// viscosity_kernel.f90:71-134 and calc_dt_kernel.f90:102-126 at one cell:
// the node velocities of its planes l (ul, vl, wl) and l+1 (uu, vu, wu) as
// quad_up_raw quads, its density and pressure, the pressures of its six
// neighbours -> its viscosity q; raises the running maxima of calc_dt's five
// denominators (mc, mu, mv, mw, mdiv) with the cell's
template <typename T>
__device__ inline T visc_cell(const QuadS<T> &ul, const QuadS<T> &vl, const QuadS<T> &wl, const QuadS<T> &uu,
                              const QuadS<T> &vu, const QuadS<T> &wu, T rho, T p_c, T p_lo, T p_hi, T p_xm, T p_xp,
                              T p_ym, T p_yp, const Geo &m, T ds, T dx, T dy, T dz, T xa, T ya, T za, T &mc, T &mu,
                              T &mv, T &mw, T &mdiv)
{
    const T c = soundspeed_of<T>(rho, p_c);
#define U(a, b, c) ((c) ? uu : ul).v[(a) + 2 * (b)]
#define V(a, b, c) ((c) ? vu : vl).v[(a) + 2 * (b)]
#define W(a, b, c) ((c) ? wu : wl).v[(a) + 2 * (b)]
    const T ugradx1 = U(0, 0, 0) + U(0, 1, 0) + U(0, 0, 1) + U(0, 1, 1);
    const T ugradx2 = U(1, 0, 0) + U(1, 1, 0) + U(1, 0, 1) + U(1, 1, 1);
    const T ugrady1 = U(0, 0, 0) + U(1, 0, 0) + U(0, 0, 1) + U(1, 0, 1);
    const T ugrady2 = U(0, 1, 0) + U(1, 1, 0) + U(0, 1, 1) + U(1, 1, 1);
    const T ugradz1 = U(0, 0, 0) + U(1, 0, 0) + U(0, 1, 0) + U(1, 1, 0);
    const T ugradz2 = U(0, 0, 1) + U(1, 0, 1) + U(0, 1, 1) + U(1, 1, 1);
    const T vgradx1 = V(0, 0, 0) + V(0, 1, 0) + V(0, 0, 1) + V(0, 1, 1);
    const T vgradx2 = V(1, 0, 0) + V(1, 1, 0) + V(1, 0, 1) + V(1, 1, 1);
    const T vgrady1 = V(0, 0, 0) + V(1, 0, 0) + V(0, 0, 1) + V(1, 0, 1);
    const T vgrady2 = V(0, 1, 0) + V(1, 1, 0) + V(0, 1, 1) + V(1, 1, 1);
    const T vgradz1 = V(0, 0, 0) + V(1, 0, 0) + V(0, 1, 0) + V(1, 1, 0);
    const T vgradz2 = V(0, 0, 1) + V(1, 0, 1) + V(0, 1, 1) + V(1, 1, 1);
    const T wgradx1 = W(0, 0, 0) + W(0, 1, 0) + W(0, 0, 1) + W(0, 1, 1);
    const T wgradx2 = W(1, 0, 0) + W(1, 1, 0) + W(1, 0, 1) + W(1, 1, 1);
    const T wgrady1 = W(0, 0, 0) + W(1, 0, 0) + W(0, 0, 1) + W(1, 0, 1);
    const T wgrady2 = W(0, 1, 0) + W(1, 1, 0) + W(0, 1, 1) + W(1, 1, 1);
    const T wgradz1 = W(0, 0, 0) + W(1, 0, 0) + W(0, 1, 0) + W(1, 1, 0);
    const T wgradz2 = W(0, 0, 1) + W(1, 0, 1) + W(0, 1, 1) + W(1, 1, 1);
    // calc_dt_kernel.f90:124-125 sums the z-face nodes in its own order
    const T dw1s = W(0, 0, 0) + W(0, 1, 0) + W(1, 0, 0) + W(1, 1, 0);
    const T dw2s = W(0, 0, 1) + W(0, 1, 1) + W(1, 0, 1) + W(1, 1, 1);
#undef U
#undef V
#undef W
    const T div = (xa * (ugradx2 - ugradx1) + ya * (vgrady2 - vgrady1)) + za * (wgradz2 - wgradz1);
    const T xx = by_const_scaled<T>(ugradx2 - ugradx1, 0.25, m.rx);
    const T yy = by_const_scaled<T>(vgrady2 - vgrady1, 0.25, m.ry);
    const T zz = by_const_scaled<T>(wgradz2 - wgradz1, 0.25, m.rz);
    const T xy = by_const_scaled<T>(ugrady2 - ugrady1, 0.25, m.ry) + by_const_scaled<T>(vgradx2 - vgradx1, 0.25, m.rx);
    const T xz = by_const_scaled<T>(ugradz2 - ugradz1, 0.25, m.rz) + by_const_scaled<T>(wgradx2 - wgradx1, 0.25, m.rx);
    const T yz = by_const_scaled<T>(vgradz2 - vgradz1, 0.25, m.rz) + by_const_scaled<T>(wgrady2 - wgrady1, 0.25, m.ry);
    // p difference / (celld + celld): halving the difference instead
    // of doubling the divisor gives the same quotient (both exact)
    T pgradx = by_const_scaled<T>(p_xp - p_xm, 0.5, m.rx);
    T pgrady = by_const_scaled<T>(p_yp - p_ym, 0.5, m.ry);
    T pgradz = by_const_scaled<T>(p_hi - p_lo, 0.5, m.rz);
    const T pgradx2 = pgradx * pgradx;
    const T pgrady2 = pgrady * pgrady;
    const T pgradz2 = pgradz * pgradz;
    // the limiter's denominator is positive, so its sign is the
    // numerator's, and its value is needed only where q is formed
    // (a positive quotient that underflows to zero gives q = 0 there)
    const T lim_num = xx * pgradx2 + yy * pgrady2 + zz * pgradz2 + xy * pgradx * pgrady +
                      xz * pgradx * pgradz + yz * pgrady * pgradz;
    const T lim_den = tmax(pgradx2 + pgrady2 + pgradz2, (T)1.0e-16);
    T limiter = 0;
    if constexpr (sizeof(T) == 4) limiter = quo(lim_num, lim_den);  // cheaper than a branch
    T q;
    if ((lim_num > (T)0) || (div >= (T)0)) {
        q = (T)0;
    } else {
        if constexpr (sizeof(T) == 8) limiter = quo(lim_num, lim_den);
        pgradx = copysign(tmax((T)1.0e-16, fabs(pgradx)), pgradx);
        pgrady = copysign(tmax((T)1.0e-16, fabs(pgrady)), pgrady);
        pgradz = copysign(tmax((T)1.0e-16, fabs(pgradz)), pgradz);
        const T pgrad = sqrt(pgradx * pgradx + pgrady * pgrady + pgradz * pgradz);
        const T xgrad = fabs(quo(dx * pgrad, pgradx));
        const T ygrad = fabs(quo(dy * pgrad, pgrady));
        const T zgrad = fabs(quo(dz * pgrad, pgradz));
        const T grad = tmin(tmin(xgrad, ygrad), zgrad);
        const T grad2 = grad * grad;
        q = (T)2 * rho * grad2 * limiter * limiter;
    }

    // calc_dt: the denominators of the five candidates (visc_dt_body
    // forms the candidates from their maxima over a thread's cells)
    T cc = c * c;
    // in FP64 the divide is skipped where q is zero (adding 2*0/rho
    // leaves cc as it is); the FP32 reciprocal is cheaper than a branch
    if (sizeof(T) == 4 || q != (T)0) cc = cc + quo((T)2 * q, rho);
    mc = tmax(mc, ds * cc);
    const T du1 = ugradx1 * xa, du2 = ugradx2 * xa;
    mu = tmax(mu, tmax(fabs(du1), fabs(du2)));
    const T dv1 = vgrady1 * ya, dv2 = vgrady2 * ya;
    mv = tmax(mv, tmax(fabs(dv1), fabs(dv2)));
    const T dw1 = dw1s * za, dw2 = dw2s * za;
    mw = tmax(mw, tmax(fabs(dw1), fabs(dw2)));
    const T dv = du2 - du1 + dv2 - dv1 + dw2 - dw1;
    mdiv = tmax(mdiv, fabs(dv));
    return q;
}

// ---------------------------------------------------------------------------
// viscosity_kernel.f90:71-134 + calc_dt_kernel.f90:102-149.  calc_dt reads
// viscosity only at the cell itself, so it runs on the value just computed
// (the halo exchange of viscosity between the two in timestep.f90:80-84
// does not touch interior cells).  Each workgroup leaves one partial of the
// dt minimum, which is FP64 in both forms.
template <typename S>
__device__ __forceinline__ void visc_dt_body(const Blk &bk, Lay g, Geo m, double dtc_safe, double dtu_safe, double dtv_safe, double dtw_safe, double dtdiv_safe,
               double g_small, const S *__restrict__ density0, const S *__restrict__ pressure,
               S *__restrict__ viscosity, const S *__restrict__ xvel0, const S *__restrict__ yvel0,
               const S *__restrict__ zvel0, double *__restrict__ partial)
{
    using T = S;  // arithmetic type of the form
    const int j = g.x_min + bk.x * (int)blockDim.x + (int)threadIdx.x;
    const int k = g.y_min + bk.y * (int)blockDim.y + (int)threadIdx.y;
    int l0, l1;
    chunk_range(bk, g.z_min, g.z_max, l0, l1);
    const bool in = j <= g.x_max && k <= g.y_max && l0 <= l1;  // every thread reaches the reduction
    const double mm = dmin(dmin(m.dx, m.dy), m.dz);
    const T ds = (T)(1.0 / (mm * mm));
    const T dx = (T)m.dx, dy = (T)m.dy, dz = (T)m.dz;
    const T xa = (T)m.xa, ya = (T)m.ya, za = (T)m.za, vol = (T)m.vol;
    const T small = (T)g_small;
    double dtmin_thread = 1.0e300;
    if (in) {
        T mc = 0, mu = 0, mv = 0, mw = 0, mdiv = 0;  // largest dt denominators so far (all are >= 0)
        int i = idx(g, j, k, l0);
        QuadS<S> ul = quad_up_raw(xvel0, i, g.ld), vl = quad_up_raw(yvel0, i, g.ld), wl = quad_up_raw(zvel0, i, g.ld);
        // a mirrored face (Lay::mirror): the neighbour across it is a halo
        // cell, read as the cell itself, which its reflection copies
        const int oxm = cell_m(g, 0, j - 1) - j, oxp = cell_m(g, 0, j + 1) - j;
        const int oym = (cell_m(g, 1, k - 1) - k) * g.ld, oyp = (cell_m(g, 1, k + 1) - k) * g.ld;
        T p_lo = pressure[i + (cell_m(g, 2, l0 - 1) - l0) * g.plane], p_c = pressure[i];
        // The cell's soundspeed is not read: it is ideal_gas's of the cell's
        // density and pressure, formed here with the same operations
        // (soundspeed_of), so it has the same bits.
        for (int l = l0; l <= l1; l++, i += g.plane) {
            // the upper node plane's velocities and pressure, the cell's
            // density and its x/y pressure neighbours, loaded at the top of
            // the plane: holding the next plane's loads in registers as well
            // took 150 VGPRs in FP64 (three waves per SIMD) against 126 (four)
            // and measured 1.55 against 1.49 ms at 512^3
            const QuadS<S> uu = quad_up_raw(xvel0, i + g.plane, g.ld), vu = quad_up_raw(yvel0, i + g.plane, g.ld),
                           wu = quad_up_raw(zvel0, i + g.plane, g.ld);
            const T p_hi = pressure[i + (cell_m(g, 2, l + 1) - l) * g.plane];
            const T rho = density0[i];
            const T p_xm = pressure[i + oxm], p_xp = pressure[i + oxp];
            const T p_ym = pressure[i + oym], p_yp = pressure[i + oyp];
            viscosity[i] = visc_cell<T>(ul, vl, wl, uu, vu, wu, rho, p_c, p_lo, p_hi, p_xm, p_xp, p_ym, p_yp, m, ds,
                                        dx, dy, dz, xa, ya, za, mc, mu, mv, mw, mdiv);
            ul = uu;
            vl = vu;
            wl = wu;
            p_lo = p_c;
            p_c = p_hi;
        }
        // calc_dt_kernel.f90:127-149 per cell is dt = min(C_c / max(sqrt(ds*cc), g_small),
        // C_u / max(|du1|, |du2|, 1e-5*volume), C_v / .., C_w / .., C_div / max(1e-5*volume, |div|))
        // with constant numerators.  Rounded division by a positive
        // denominator and rounded sqrt are monotone, so the minimum over
        // cells of each candidate is its numerator over the largest
        // denominator: the same value, with five divisions per thread
        // instead of per cell.
        const T dtct = quo((T)dtc_safe * (T)1, tmax(sqrt(mc), small));
        const T dtut = quo((T)dtu_safe * (T)4 * vol, tmax(mu, (T)1.0e-5 * vol));
        const T dtvt = quo((T)dtv_safe * (T)4 * vol, tmax(mv, (T)1.0e-5 * vol));
        const T dtwt = quo((T)dtw_safe * (T)4 * vol, tmax(mw, (T)1.0e-5 * vol));
        const T dtdivt = quo((T)dtdiv_safe * (T)4 * (vol), tmax(vol * (T)1.0e-05, mdiv));
        dtmin_thread = (double)tmin(tmin(tmin(tmin(dtct, dtut), dtvt), dtwt), dtdivt);
    }
    __shared__ double red[kMaxBlock / 64];
    const int tid = (int)(threadIdx.y * blockDim.x + threadIdx.x);
    const int nwaves = (int)(blockDim.x * blockDim.y) / 64;
    const double w = wave_min(dtmin_thread);
    if ((tid & 63) == 0) red[tid >> 6] = w;
    __syncthreads();
    if (tid == 0) {
        double b = red[0];
        for (int q = 1; q < nwaves; q++) b = dmin(b, red[q]);
        partial[(bk.z * bk.gy + bk.y) * bk.gx + bk.x] = b;
    }
}

// The FP32 form with two cells per thread (NodePair).  Cell j's x
// neighbours are j-1 (a 4-byte load) and j+1 (the pair's second pressure,
// or cell j itself at a mirrored x_max face), cell j+1's are j (the pair's
// first) and j+2 (a 4-byte load); each cell goes through visc_cell as in
// visc_dt_body, and the dt candidates come from the thread's maxima.  110
// VGPRs against 88; at 512^3: 827 -> 743 us, three rounds on one machine,
// same bits.  The batched kernel (kb_visc_dt) uses it too: 640^3 as 8
// tiles 1.68 -> 1.59 ms (the 512-deck ensemble: unchanged).
__device__ __forceinline__ void visc_dt_pair_body(const Blk &bk, Lay g, Geo m, double dtc_safe, double dtu_safe,
                                                  double dtv_safe, double dtw_safe, double dtdiv_safe, double g_small,
                                                  const float *__restrict__ density0,
                                                  const float *__restrict__ pressure, float *__restrict__ viscosity,
                                                  const float *__restrict__ xvel0, const float *__restrict__ yvel0,
                                                  const float *__restrict__ zvel0, double *__restrict__ partial)
{
    using T = float;
    const int j = g.x_min + 2 * (bk.x * (int)blockDim.x + (int)threadIdx.x);
    const int k = g.y_min + bk.y * (int)blockDim.y + (int)threadIdx.y;
    int l0, l1;
    chunk_range(bk, g.z_min, g.z_max, l0, l1);
    const bool in = j <= g.x_max && k <= g.y_max && l0 <= l1;  // every thread reaches the reduction
    const double mm = dmin(dmin(m.dx, m.dy), m.dz);
    const T ds = (T)(1.0 / (mm * mm));
    const T dx = (T)m.dx, dy = (T)m.dy, dz = (T)m.dz;
    const T xa = (T)m.xa, ya = (T)m.ya, za = (T)m.za, vol = (T)m.vol;
    const T small = (T)g_small;
    double dtmin_thread = 1.0e300;
    if (in) {
        const bool two = j < g.x_max;  // cell j+1 is interior (else the halo cell x_max+1: loaded, not stored)
        T mc = 0, mu = 0, mv = 0, mw = 0, mdiv = 0;
        int i = idx(g, j, k, l0);  // even: x_min sits on a 32-element boundary
        auto pair = [&](const float *__restrict__ p, int ii) { return *reinterpret_cast<const float2 *>(p + ii); };
        NodePair ul = load_node_pair(xvel0, i, g.ld), vl = load_node_pair(yvel0, i, g.ld),
                 wl = load_node_pair(zvel0, i, g.ld);
        const int oxm = cell_m(g, 0, j - 1) - j, oxp2 = cell_m(g, 0, j + 2) - j;
        const bool xp_self = cell_m(g, 0, j + 1) != j + 1;  // j = x_max on a mirrored face
        const int oym = (cell_m(g, 1, k - 1) - k) * g.ld, oyp = (cell_m(g, 1, k + 1) - k) * g.ld;
        float2 p_lo = pair(pressure, i + (cell_m(g, 2, l0 - 1) - l0) * g.plane), p_c = pair(pressure, i);
        for (int l = l0; l <= l1; l++, i += g.plane) {
            const NodePair uu = load_node_pair(xvel0, i + g.plane, g.ld), vu = load_node_pair(yvel0, i + g.plane, g.ld),
                           wu = load_node_pair(zvel0, i + g.plane, g.ld);
            const float2 p_hi = pair(pressure, i + (cell_m(g, 2, l + 1) - l) * g.plane);
            const float2 rho = pair(density0, i);
            const T p_xm = pressure[i + oxm], p_xp2 = pressure[i + oxp2];
            const float2 p_ym = pair(pressure, i + oym), p_yp = pair(pressure, i + oyp);
            const T qa = visc_cell<T>(quad_a(ul), quad_a(vl), quad_a(wl), quad_a(uu), quad_a(vu), quad_a(wu), rho.x,
                                      p_c.x, p_lo.x, p_hi.x, p_xm, xp_self ? p_c.x : p_c.y, p_ym.x, p_yp.x, m, ds, dx,
                                      dy, dz, xa, ya, za, mc, mu, mv, mw, mdiv);
            if (two) {
                const T qb = visc_cell<T>(quad_b(ul), quad_b(vl), quad_b(wl), quad_b(uu), quad_b(vu), quad_b(wu),
                                          rho.y, p_c.y, p_lo.y, p_hi.y, p_c.x, p_xp2, p_ym.y, p_yp.y, m, ds, dx, dy, dz,
                                          xa, ya, za, mc, mu, mv, mw, mdiv);
                *reinterpret_cast<float2 *>(viscosity + i) = make_float2(qa, qb);
            } else {
                viscosity[i] = qa;
            }
            ul = uu;
            vl = vu;
            wl = wu;
            p_lo = p_c;
            p_c = p_hi;
        }
        const T dtct = quo((T)dtc_safe * (T)1, tmax(sqrt(mc), small));
        const T dtut = quo((T)dtu_safe * (T)4 * vol, tmax(mu, (T)1.0e-5 * vol));
        const T dtvt = quo((T)dtv_safe * (T)4 * vol, tmax(mv, (T)1.0e-5 * vol));
        const T dtwt = quo((T)dtw_safe * (T)4 * vol, tmax(mw, (T)1.0e-5 * vol));
        const T dtdivt = quo((T)dtdiv_safe * (T)4 * (vol), tmax(vol * (T)1.0e-05, mdiv));
        dtmin_thread = (double)tmin(tmin(tmin(tmin(dtct, dtut), dtvt), dtwt), dtdivt);
    }
    __shared__ double red[kMaxBlock / 64];
    const int tid = (int)(threadIdx.y * blockDim.x + threadIdx.x);
    const int nwaves = (int)(blockDim.x * blockDim.y) / 64;
    const double w = wave_min(dtmin_thread);
    if ((tid & 63) == 0) red[tid >> 6] = w;
    __syncthreads();
    if (tid == 0) {
        double b = red[0];
        for (int q = 1; q < nwaves; q++) b = dmin(b, red[q]);
        partial[(bk.z * bk.gy + bk.y) * bk.gx + bk.x] = b;
    }
}

template <typename S>
__global__ void __launch_bounds__(kMaxBlock) k_fast_visc_dt(int xcd, Lay g, Geo m, double dtc_safe, double dtu_safe, double dtv_safe, double dtw_safe, double dtdiv_safe,
               double g_small, const S *__restrict__ density0, const S *__restrict__ pressure,
               S *__restrict__ viscosity, const S *__restrict__ xvel0, const S *__restrict__ yvel0,
               const S *__restrict__ zvel0, double *__restrict__ partial)
{
    if constexpr (sizeof(S) == 4)
        visc_dt_pair_body(hw_blk(xcd), g, m, dtc_safe, dtu_safe, dtv_safe, dtw_safe, dtdiv_safe, g_small, density0,
                          pressure, viscosity, xvel0, yvel0, zvel0, partial);
    else
        visc_dt_body<S>(hw_blk(xcd), g, m, dtc_safe, dtu_safe, dtv_safe, dtw_safe, dtdiv_safe, g_small, density0,
                        pressure, viscosity, xvel0, yvel0, zvel0, partial);
}

template <typename S>
static hipError_t fast_visc_dt_t(const Lay &g, const Geo &m, double dtc_safe, double dtu_safe, double dtv_safe,
                               double dtw_safe, double dtdiv_safe, double g_small, const S *density0,
                               const S *pressure, S *viscosity,
                               const S *xvel0, const S *yvel0, const S *zvel0, double *partial,
                               int *nparts, const Shape &sh, hipStream_t st)
{
    const int nx = g.x_max - g.x_min + 1, ny = g.y_max - g.y_min + 1, nz = g.z_max - g.z_min + 1;
    const dim3 blk = block_of(sh);
    const int cpt = sizeof(S) == 4 ? 2 : 1;  // cells per thread along x
    const int bx = (nx + cpt * (int)blk.x - 1) / (cpt * (int)blk.x), by = (ny + (int)blk.y - 1) / (int)blk.y;
    if (bx * by > kMaxDtPartials) return hipErrorInvalidValue;  // caller keeps tiles within the partial buffer
    int gz = nchunks(nz, sh.planes);
    if (bx * by * gz > kMaxDtPartials) gz = kMaxDtPartials / (bx * by);
    *nparts = bx * by * gz;
    TIMED_LAUNCH("k_fast_visc_dt", st, k_fast_visc_dt<S>, dim3(bx, by, gz), blk, 0, st, sh.xcd, g, m, dtc_safe,
                 dtu_safe, dtv_safe, dtw_safe, dtdiv_safe, g_small, density0, pressure, viscosity, xvel0, yvel0,
                 zvel0, partial);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// PdV predictor (PdV_kernel.f90:78-116) then ideal_gas on the predicted
// state (ideal_gas_kernel.f90:50-51): only the new pressure is stored.  The
// predicted density/energy are reverted before anything reads them
// (PdV.f90:103-107) and the predicted soundspeed is recomputed by the next
// ideal_gas before calc_dt reads it.  Fortran sums a+b+c+d+a+b+c+d as a left
// fold, so the 8-term face sums continue from the 4-term partial.
//
// PdV(predict), revert and ideal_gas at one cell: the node velocities of
// its planes l (ul, vl, wl) and l+1 (uu, vu, wu) as quad_up_raw quads, its
// level-0 density, energy and viscosity -> the predicted pressure
template <typename T>
__device__ inline T predict_pressure(const QuadS<T> &ul, const QuadS<T> &vl, const QuadS<T> &wl, const QuadS<T> &uu,
                                     const QuadS<T> &vu, const QuadS<T> &wu, T d0, T e0v, T q0, T xa, T ya, T za,
                                     T vol, T tdt)
{
    const T p0 = pressure_of<T>(d0, e0v);
#define U(a, b, c) ((c) ? uu : ul).v[(a) + 2 * (b)]
#define V(a, b, c) ((c) ? vu : vl).v[(a) + 2 * (b)]
#define W(a, b, c) ((c) ? wu : wl).v[(a) + 2 * (b)]
    const T sl = U(0, 0, 0) + U(0, 1, 0) + U(0, 0, 1) + U(0, 1, 1) + U(0, 0, 0) + U(0, 1, 0) + U(0, 0, 1) + U(0, 1, 1);
    const T sr = U(1, 0, 0) + U(1, 1, 0) + U(1, 0, 1) + U(1, 1, 1) + U(1, 0, 0) + U(1, 1, 0) + U(1, 0, 1) + U(1, 1, 1);
    const T sb = V(0, 0, 0) + V(1, 0, 0) + V(0, 0, 1) + V(1, 0, 1) + V(0, 0, 0) + V(1, 0, 0) + V(0, 0, 1) + V(1, 0, 1);
    const T stp = V(0, 1, 0) + V(1, 1, 0) + V(0, 1, 1) + V(1, 1, 1) + V(0, 1, 0) + V(1, 1, 0) + V(0, 1, 1) + V(1, 1, 1);
    const T sk = W(0, 0, 0) + W(1, 0, 0) + W(0, 1, 0) + W(1, 1, 0) + W(0, 0, 0) + W(1, 0, 0) + W(0, 1, 0) + W(1, 1, 0);
    const T sf = W(0, 0, 1) + W(1, 0, 1) + W(0, 1, 1) + W(1, 1, 1) + W(0, 0, 1) + W(1, 0, 1) + W(0, 1, 1) + W(1, 1, 1);
#undef U
#undef V
#undef W
    const T left_flux = (xa * sl) * (T)0.125 * tdt * (T)0.5;
    const T right_flux = (xa * sr) * (T)0.125 * tdt * (T)0.5;
    const T bottom_flux = (ya * sb) * (T)0.125 * tdt * (T)0.5;
    const T top_flux = (ya * stp) * (T)0.125 * tdt * (T)0.5;
    const T back_flux = (za * sk) * (T)0.125 * tdt * (T)0.5;
    const T front_flux = (za * sf) * (T)0.125 * tdt * (T)0.5;
    const T total_flux = right_flux - left_flux + top_flux - bottom_flux + front_flux - back_flux;
    const T vol_change_s = quo(vol, vol + total_flux);
    const T recip_volume = (T)1 / vol;
    const T energy_change = (quo(p0, d0) + quo_sparse(q0, d0)) * total_flux * recip_volume;
    const T e1 = e0v - energy_change;
    const T d1 = d0 * vol_change_s;
    return (T)(1.4 - 1.0) * d1 * e1;
}

template <typename S>
__device__ __forceinline__ void pdv_predict_body(const Blk &bk, Lay g, Geo m, double dt, const S *__restrict__ density0, const S *__restrict__ energy0,
                   S *__restrict__ pressure, const S *__restrict__ viscosity,
                   const S *__restrict__ xvel0, const S *__restrict__ yvel0,
                   const S *__restrict__ zvel0)
{
    const int j = g.x_min + bk.x * (int)blockDim.x + (int)threadIdx.x;
    const int k = g.y_min + bk.y * (int)blockDim.y + (int)threadIdx.y;
    if (j > g.x_max || k > g.y_max) return;
    int l0, l1;
    chunk_range(bk, g.z_min, g.z_max, l0, l1);
    if (l0 > l1) return;
    using T = S;  // arithmetic type of the form
    const T xa = (T)m.xa, ya = (T)m.ya, za = (T)m.za, vol = (T)m.vol, tdt = (T)dt;
    int i = idx(g, j, k, l0);
    QuadS<S> ul = quad_up_raw(xvel0, i, g.ld), vl = quad_up_raw(yvel0, i, g.ld), wl = quad_up_raw(zvel0, i, g.ld);
    // the next plane's loads run one plane ahead of their use
    QuadS<S> un = quad_up_raw(xvel0, i + g.plane, g.ld), vn = quad_up_raw(yvel0, i + g.plane, g.ld),
             wn = quad_up_raw(zvel0, i + g.plane, g.ld);
    // The cell's pressure is not read: at this point it is ideal_gas's of
    // the cell's level-0 density and energy, formed here with the same
    // expression (pressure_of), so it has the same bits.
    S dn = density0[i], en = energy0[i], qn = viscosity[i];
    for (int l = l0; l <= l1; l++, i += g.plane) {
        const QuadS<S> uu = un, vu = vn, wu = wn;
        const T d0 = dn, e0v = en, q0 = qn;
        if (l < l1) {
            const int i1 = i + g.plane, i2 = i + 2 * g.plane;
            un = quad_up_raw(xvel0, i2, g.ld);
            vn = quad_up_raw(yvel0, i2, g.ld);
            wn = quad_up_raw(zvel0, i2, g.ld);
            dn = density0[i1];
            en = energy0[i1];
            qn = viscosity[i1];
        }
        pressure[i] = predict_pressure<T>(ul, vl, wl, uu, vu, wu, d0, e0v, q0, xa, ya, za, vol, tdt);
        ul = uu;
        vl = vu;
        wl = wu;
    }
}

// The FP32 form with two cells per thread (NodePair): each load carries
// twice the bytes of the one-cell form's (76 VGPRs against 54).  Each
// cell's pressure comes from predict_pressure, as in pdv_predict_body.  At
// 512^3: 834 -> 750 us, three rounds on one machine, same bits.  The batched
// kernel (kb_pdv_predict) uses it too: 640^3 as 8 tiles 1.66 -> 1.53 ms.

__device__ __forceinline__ void pdv_predict_pair_body(const Blk &bk, Lay g, Geo m, double dt,
                                                      const float *__restrict__ density0,
                                                      const float *__restrict__ energy0, float *__restrict__ pressure,
                                                      const float *__restrict__ viscosity,
                                                      const float *__restrict__ xvel0, const float *__restrict__ yvel0,
                                                      const float *__restrict__ zvel0)
{
    const int j = g.x_min + 2 * (bk.x * (int)blockDim.x + (int)threadIdx.x);
    const int k = g.y_min + bk.y * (int)blockDim.y + (int)threadIdx.y;
    if (j > g.x_max || k > g.y_max) return;
    int l0, l1;
    chunk_range(bk, g.z_min, g.z_max, l0, l1);
    if (l0 > l1) return;
    const bool two = j < g.x_max;  // cell j+1 is interior (else it is the halo cell x_max+1, loaded, not stored)
    const float xa = (float)m.xa, ya = (float)m.ya, za = (float)m.za, vol = (float)m.vol, tdt = (float)dt;
    int i = idx(g, j, k, l0);  // even: x_min sits on a 32-element boundary
    NodePair ul = load_node_pair(xvel0, i, g.ld), vl = load_node_pair(yvel0, i, g.ld),
             wl = load_node_pair(zvel0, i, g.ld);
    NodePair un = load_node_pair(xvel0, i + g.plane, g.ld), vn = load_node_pair(yvel0, i + g.plane, g.ld),
             wn = load_node_pair(zvel0, i + g.plane, g.ld);
    auto cells = [&](const float *__restrict__ p, int ii) { return *reinterpret_cast<const float2 *>(p + ii); };
    float2 dn = cells(density0, i), en = cells(energy0, i), qn = cells(viscosity, i);
    for (int l = l0; l <= l1; l++, i += g.plane) {
        const NodePair uu = un, vu = vn, wu = wn;
        const float2 d0 = dn, e0 = en, q0 = qn;
        if (l < l1) {
            const int i1 = i + g.plane, i2 = i + 2 * g.plane;
            un = load_node_pair(xvel0, i2, g.ld);
            vn = load_node_pair(yvel0, i2, g.ld);
            wn = load_node_pair(zvel0, i2, g.ld);
            dn = cells(density0, i1);
            en = cells(energy0, i1);
            qn = cells(viscosity, i1);
        }
        const float pa = predict_pressure<float>(quad_a(ul), quad_a(vl), quad_a(wl), quad_a(uu), quad_a(vu),
                                                 quad_a(wu), d0.x, e0.x, q0.x, xa, ya, za, vol, tdt);
        if (two) {
            const float pb = predict_pressure<float>(quad_b(ul), quad_b(vl), quad_b(wl), quad_b(uu), quad_b(vu),
                                                     quad_b(wu), d0.y, e0.y, q0.y, xa, ya, za, vol, tdt);
            *reinterpret_cast<float2 *>(pressure + i) = make_float2(pa, pb);
        } else {
            pressure[i] = pa;
        }
        ul = uu;
        vl = vu;
        wl = wu;
    }
}

template <typename S>
__global__ void __launch_bounds__(kMaxBlock) k_fast_pdv_predict(int xcd, Lay g, Geo m, double dt, const S *__restrict__ density0, const S *__restrict__ energy0,
                   S *__restrict__ pressure, const S *__restrict__ viscosity,
                   const S *__restrict__ xvel0, const S *__restrict__ yvel0,
                   const S *__restrict__ zvel0)
{
    if constexpr (sizeof(S) == 4)
        pdv_predict_pair_body(hw_blk(xcd), g, m, dt, density0, energy0, pressure, viscosity, xvel0, yvel0, zvel0);
    else
        pdv_predict_body<S>(hw_blk(xcd), g, m, dt, density0, energy0, pressure, viscosity, xvel0, yvel0, zvel0);
}

template <typename S>
static hipError_t fast_pdv_predict_t(const Lay &g, const Geo &m, double dt, const S *density0,
                                   const S *energy0, S *pressure, const S *viscosity,
                                   const S *xvel0, const S *yvel0, const S *zvel0, const Shape &sh, hipStream_t st)
{
    const int nx = g.x_max - g.x_min + 1, ny = g.y_max - g.y_min + 1, nz = g.z_max - g.z_min + 1;
    const dim3 blk = block_of(sh);
    const int cpt = sizeof(S) == 4 ? 2 : 1;  // cells per thread along x
    const dim3 grid((nx + cpt * blk.x - 1) / (cpt * blk.x), (ny + blk.y - 1) / blk.y, nchunks(nz, sh.planes));
    TIMED_LAUNCH("k_fast_pdv_predict", st, k_fast_pdv_predict<S>, grid, blk, 0, st, sh.xcd, g, m, dt, density0,
                 energy0, pressure, viscosity, xvel0, yvel0, zvel0);
    return hipGetLastError();
}

// accelerate_kernel.f90:61-109 at one node: density0, pressure and viscosity
// of the cells (j-1..j, k-1..k) in planes l (du, pu, qu) and l-1 (dl, pl, ql),
// the node's level-0 velocity (u0, v0, w0) -> its level-1 velocity
template <typename S>
__device__ inline void accel_node(const QuadS<S> &du, const QuadS<S> &pu, const QuadS<S> &qu, const QuadS<S> &dl,
                                  const QuadS<S> &pl, const QuadS<S> &ql, S u0, S v0, S w0, S vol, S xa, S ya, S za,
                                  S tdt, S &u1, S &v1, S &w1)
{
    using T = S;
    // (a, b, c) in {-1, 0}^3; c = 0 is plane l, c = -1 plane l-1
#define RV(a, b, c) (((c) ? dl : du).v[((a) + 1) + 2 * ((b) + 1)] * vol)
#define P(a, b, c) ((c) ? pl : pu).v[((a) + 1) + 2 * ((b) + 1)]
#define Q(a, b, c) ((c) ? ql : qu).v[((a) + 1) + 2 * ((b) + 1)]
    const T nodal_mass = (RV(-1, -1, 0) + RV(0, -1, 0) + RV(0, 0, 0) + RV(-1, 0, 0) + RV(-1, -1, -1) +
                          RV(0, -1, -1) + RV(0, 0, -1) + RV(-1, 0, -1)) *
                         (T)0.125;
    const T stepbymass_s = quo((T)0.25 * tdt, nodal_mass);
    const T xvel1_s = u0 - stepbymass_s * (xa * (P(0, 0, 0) - P(-1, 0, 0)) +
                                                      xa * (P(0, -1, 0) - P(-1, -1, 0)) +
                                                      xa * (P(0, 0, -1) - P(-1, 0, -1)) +
                                                      xa * (P(0, -1, -1) - P(-1, -1, -1)));
    const T yvel1_s = v0 - stepbymass_s * (ya * (P(0, 0, 0) - P(0, -1, 0)) +
                                                      ya * (P(-1, 0, 0) - P(-1, -1, 0)) +
                                                      ya * (P(0, 0, -1) - P(0, -1, -1)) +
                                                      ya * (P(-1, 0, -1) - P(-1, -1, -1)));
    const T zvel1_s = w0 - stepbymass_s * (za * (P(0, 0, 0) - P(0, 0, -1)) +
                                                      za * (P(0, -1, 0) - P(0, -1, -1)) +
                                                      za * (P(-1, 0, 0) - P(-1, 0, -1)) +
                                                      za * (P(-1, -1, 0) - P(-1, -1, -1)));
    u1 = xvel1_s - stepbymass_s * (xa * (Q(0, 0, 0) - Q(-1, 0, 0)) + xa * (Q(0, -1, 0) - Q(-1, -1, 0)) +
                                         xa * (Q(0, 0, -1) - Q(-1, 0, -1)) + xa * (Q(0, -1, -1) - Q(-1, -1, -1)));
    v1 = yvel1_s - stepbymass_s * (ya * (Q(0, 0, 0) - Q(0, -1, 0)) + ya * (Q(-1, 0, 0) - Q(-1, -1, 0)) +
                                         ya * (Q(0, 0, -1) - Q(0, -1, -1)) + ya * (Q(-1, 0, -1) - Q(-1, -1, -1)));
    w1 = zvel1_s - stepbymass_s * (za * (Q(0, 0, 0) - Q(0, 0, -1)) + za * (Q(0, -1, 0) - Q(0, -1, -1)) +
                                         za * (Q(-1, 0, 0) - Q(-1, 0, -1)) + za * (Q(-1, -1, 0) - Q(-1, -1, -1)));
#undef RV
#undef P
#undef Q
}

// PdV corrector (PdV_kernel.f90:126-164) and flux_calc (flux_calc_kernel.f90:
// 56-90) at plane l of column (j, k), offset i.  at(c, a, b, p) is the
// velocity component c (0-2 level 0 x/y/z, 3-5 level 1 x/y/z) of node
// (j+a, k+b) in node plane l+p; plane l+1 is read only when lin
// (l <= z_max).  The cell's level-0 state is (d0, e0v, p0, q0).
template <typename S, typename NodeAt>
__device__ inline void corrector_plane(int i, bool jin, bool kin, bool lin, const NodeAt &at, S d0, S e0v, S p0,
                                       S q0, S xa, S ya, S za, S vol, S tdt, S c0, S *__restrict__ density1,
                                       S *__restrict__ energy1, S *__restrict__ vol_flux_x,
                                       S *__restrict__ vol_flux_y, S *__restrict__ vol_flux_z)
{
    using T = S;
#define U0(a, b, c) at(0, a, b, c)
#define V0(a, b, c) at(1, a, b, c)
#define W0(a, b, c) at(2, a, b, c)
#define U1(a, b, c) at(3, a, b, c)
#define V1(a, b, c) at(4, a, b, c)
#define W1(a, b, c) at(5, a, b, c)
    T sl = 0, sb = 0;
    if (kin && lin) {
        sl = U0(0, 0, 0) + U0(0, 1, 0) + U0(0, 0, 1) + U0(0, 1, 1) + U1(0, 0, 0) + U1(0, 1, 0) + U1(0, 0, 1) + U1(0, 1, 1);
        vol_flux_x[i] = c0 * xa * sl;
    }
    if (jin && lin) {
        sb = V0(0, 0, 0) + V0(1, 0, 0) + V0(0, 0, 1) + V0(1, 0, 1) + V1(0, 0, 0) + V1(1, 0, 0) + V1(0, 0, 1) + V1(1, 0, 1);
        vol_flux_y[i] = c0 * ya * sb;
    }
    if (jin && kin)
        vol_flux_z[i] = c0 * za *
                        (W0(0, 0, 0) + W0(1, 0, 0) + W0(1, 0, 0) + W0(1, 1, 0) + W1(0, 0, 0) + W1(1, 0, 0) + W1(0, 1, 0) +
                         W1(1, 1, 0));
    if (jin && kin && lin) {
        const T sr = U0(1, 0, 0) + U0(1, 1, 0) + U0(1, 0, 1) + U0(1, 1, 1) + U1(1, 0, 0) + U1(1, 1, 0) + U1(1, 0, 1) + U1(1, 1, 1);
        const T stp = V0(0, 1, 0) + V0(1, 1, 0) + V0(0, 1, 1) + V0(1, 1, 1) + V1(0, 1, 0) + V1(1, 1, 0) + V1(0, 1, 1) + V1(1, 1, 1);
        const T sk = W0(0, 0, 0) + W0(1, 0, 0) + W0(0, 1, 0) + W0(1, 1, 0) + W1(0, 0, 0) + W1(1, 0, 0) + W1(0, 1, 0) + W1(1, 1, 0);
        const T sf = W0(0, 0, 1) + W0(1, 0, 1) + W0(0, 1, 1) + W0(1, 1, 1) + W1(0, 0, 1) + W1(1, 0, 1) + W1(0, 1, 1) + W1(1, 1, 1);
        const T left_flux = (xa * sl) * (T)0.125 * tdt;
        const T right_flux = (xa * sr) * (T)0.125 * tdt;
        const T bottom_flux = (ya * sb) * (T)0.125 * tdt;
        const T top_flux = (ya * stp) * (T)0.125 * tdt;
        const T back_flux = (za * sk) * (T)0.125 * tdt;
        const T front_flux = (za * sf) * (T)0.125 * tdt;
        const T total_flux = right_flux - left_flux + top_flux - bottom_flux + front_flux - back_flux;
        const T vol_change_s = quo(vol, vol + total_flux);
        const T recip_volume = (T)1 / vol;
        const T energy_change = (quo(p0, d0) + quo_sparse(q0, d0)) * total_flux * recip_volume;
        energy1[i] = e0v - energy_change;
        density1[i] = d0 * vol_change_s;
    }
#undef U0
#undef V0
#undef W0
#undef U1
#undef V1
#undef W1
}

// ---------------------------------------------------------------------------
// field_summary_kernel.f90:51-78 with the scalar cell volume; same
// reduction tree as the per-call kernel.
template <typename S, bool COMP>
__device__ __forceinline__ void field_summary_body(const Blk &bk, Lay g, Geo m, int nj, int nk, int ncells, const S *__restrict__ density0,
                     const S *__restrict__ energy0, const S *__restrict__ pressure,
                     const S *__restrict__ xvel0, const S *__restrict__ yvel0,
                     const S *__restrict__ zvel0, double *__restrict__ partial)
{
    double s[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    double comp[5] = {0.0, 0.0, 0.0, 0.0, 0.0};  // Neumaier compensation (COMP only)
    auto add = [&](int q, double x) {
        if (!COMP) {
            s[q] += x;
            return;
        }
        const double t = s[q] + x;
        if (fabs(s[q]) >= fabs(x)) comp[q] += (s[q] - t) + x;
        else comp[q] += (x - t) + s[q];
        s[q] = t;
    };
    for (int c = bk.x * RED_THREADS + threadIdx.x; c < ncells; c += bk.gx * RED_THREADS) {
        const int j = g.x_min + c % nj;
        const int k = g.y_min + (c / nj) % nk;
        const int l = g.z_min + c / (nj * nk);
        const int i = idx(g, j, k, l);
        double vsqrd = 0.0;
        for (int lv = 0; lv <= 1; lv++)
            for (int kv = 0; kv <= 1; kv++)
                for (int jv = 0; jv <= 1; jv++) {
                    const double u = AT(xvel0, jv, kv, lv), v = AT(yvel0, jv, kv, lv), w = AT(zvel0, jv, kv, lv);
                    vsqrd = vsqrd + 0.125 * (u * u + v * v + w * w);
                }
        const double cell_vol = m.vol;
        const double cell_mass = cell_vol * density0[i];
        add(0, cell_vol);
        add(1, cell_mass);
        add(2, cell_mass * energy0[i]);
        add(3, cell_mass * 0.5 * vsqrd);
        add(4, cell_vol * pressure[i]);
    }
    if (COMP)
        for (int q = 0; q < 5; q++) s[q] += comp[q];
    __shared__ double red[5][RED_THREADS / 64];
    for (int q = 0; q < 5; q++) {
        const double w = wave_sum(s[q]);
        if ((threadIdx.x & 63) == 0) red[q][threadIdx.x >> 6] = w;
    }
    __syncthreads();
    if (threadIdx.x < 5) {
        double b = 0.0;
        for (int w = 0; w < RED_THREADS / 64; w++) b += red[threadIdx.x][w];
        partial[5 * bk.x + threadIdx.x] = b;
    }
}

template <typename S, bool COMP>
__global__ void __launch_bounds__(RED_THREADS) k_fast_field_summary(Lay g, Geo m, int nj, int nk, int ncells, const S *__restrict__ density0,
                     const S *__restrict__ energy0, const S *__restrict__ pressure,
                     const S *__restrict__ xvel0, const S *__restrict__ yvel0,
                     const S *__restrict__ zvel0, double *__restrict__ partial)
{
    field_summary_body<S, COMP>(hw_blk(), g, m, nj, nk, ncells, density0, energy0, pressure, xvel0, yvel0, zvel0, partial);
}

template <typename S>
static hipError_t fast_field_summary_t(const Lay &g, const Geo &m, const S *density0, const S *energy0,
                                     const S *pressure, const S *xvel0, const S *yvel0,
                                     const S *zvel0, double *partial, int *nparts, hipStream_t st)
{
    const int nj = g.x_max - g.x_min + 1, nk = g.y_max - g.y_min + 1, nl = g.z_max - g.z_min + 1;
    const int ncells = nj * nk * nl;
    int nb = (ncells + RED_THREADS * 4 - 1) / (RED_THREADS * 4);
    if (nb < 1) nb = 1;
    if (nb > kMaxPartials) nb = kMaxPartials;
    *nparts = nb;
    TIMED_LAUNCH("k_fast_field_summary", st, (k_fast_field_summary<S, (sizeof(S) < sizeof(double))>), dim3(nb), dim3(RED_THREADS), 0, st, g, m, nj, nk,
                 ncells, density0, energy0, pressure, xvel0, yvel0, zvel0, partial);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// ideal_gas_kernel.f90:45-60 (same expressions as the per-call kernel)
template <typename S>
__device__ __forceinline__ void ideal_gas_body(const Blk &bk, Lay g, Range r, const S *__restrict__ density, const S *__restrict__ energy,
                                 S *__restrict__ pressure, S *__restrict__ soundspeed)
{
    const int j = r.j0 + bk.x * BX + (int)threadIdx.x;
    const int k = r.k0 + bk.y * BY + (int)threadIdx.y;
    const int l = r.l0 + bk.z;
    if (j > r.j1 || k > r.k1) return;
    const int i = idx(g, j, k, l);
    ideal_gas_cell<S>(density[i], energy[i], pressure[i], soundspeed[i]);
}

template <typename S>
__global__ void k_fast_ideal_gas(Lay g, Range r, const S *__restrict__ density, const S *__restrict__ energy,
                                 S *__restrict__ pressure, S *__restrict__ soundspeed)
{
    ideal_gas_body<S>(hw_blk(), g, r, density, energy, pressure, soundspeed);
}

template <typename S>
static hipError_t fast_ideal_gas_t(const Lay &g, const S *density, const S *energy, S *pressure, S *soundspeed,
                                   hipStream_t st)
{
    Range r{g.x_min, g.x_max, g.y_min, g.y_max, g.z_min, g.z_max};
    TIMED_LAUNCH("k_fast_ideal_gas", st, k_fast_ideal_gas<S>, grid_of(r), dim3(BX, BY, 1), 0, st, g, r, density,
                 energy, pressure, soundspeed);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// accelerate (accelerate_kernel.f90:61-109), then the PdV corrector and
// flux_calc on its results (PdV.f90 + flux_calc.f90 follow accelerate in
// hydro.f90:56-60), in one z-march: the level-1 velocities are stored once
// and handed to the corrector on chip instead of being read back.
//
// A workgroup is 64 lanes (x) by AW waves (y); lane 63 and wave AW-1 are a
// halo column and row, so workgroups step by 63 and AW-1.  At node plane lp
// every thread forms the level-1 velocity of its node (accel_node, as the
// accelerate kernel does; owners store it) from its cells' quads (read from
// memory one plane ahead) and its own level-0 velocity.  Both velocity
// levels of the plane go to LDS, into a ring of two planes with two
// barriers per plane, and the owners then run the corrector and fluxes of
// plane lp-1 (corrector_plane) reading the node quads of planes lp-1 and lp
// from the ring; the plane below is read back rather than held in
// registers.  Each chunk also forms the level-1 velocity of the plane above
// it, which the next chunk stores.
// 12 waves: in FP64 the kernel takes 144 VGPRs, three waves per SIMD, so
// one 12-wave workgroup per CU keeps 12 waves in flight where 8-wave ones
// kept 8 (4.14 -> 4.05 ms FP64, 2.26 -> 2.23 ms FP32 at 512^3); 10 waves
// were slower in FP32.  The FP32 kernel is held to 80 VGPRs (six waves per
// SIMD, two workgroups per CU, one value spilled) instead of 81 (five, one
// workgroup): 2.15 -> 2.05 ms at 512^3.  16 waves (FP64 held to 128 VGPRs,
// four waves per SIMD, seven values spilled) ran slower: 4.04 -> 4.55 ms
// FP64, 2.39 -> 2.55 ms FP32.
constexpr int AW = 12;

template <typename S>
__device__ __forceinline__ void accel_pdv_body(const Blk &bk, Lay g, Geo m, double dt,
                                               const S *__restrict__ density0, S *__restrict__ density1,
                                               const S *__restrict__ energy0, S *__restrict__ energy1,
                                               const S *__restrict__ pressure, const S *__restrict__ viscosity,
                                               const S *__restrict__ xvel0, S *__restrict__ xvel1,
                                               const S *__restrict__ yvel0, S *__restrict__ yvel1,
                                               const S *__restrict__ zvel0, S *__restrict__ zvel1,
                                               S *__restrict__ vol_flux_x, S *__restrict__ vol_flux_y,
                                               S *__restrict__ vol_flux_z)
{
    using T = S;  // arithmetic type of the form
    // node velocities (level 0 x/y/z, level 1 x/y/z) of two node planes.
    // Slot s is written at plane t (before the first barrier of t) and read
    // at planes t and t+1 (between their two barriers); it is written again
    // at t+2, past the second barrier of t+1, which every thread reaches
    // only after its reads of t+1.  Two planes rather than three measured
    // faster (8-wave workgroups at 512^3: 4.29 -> 4.16 ms FP64, 2.29 ->
    // 2.24 ms FP32).
    __shared__ S nb[2][6][AW][64];
    const int lane = (int)threadIdx.x, w = (int)threadIdx.y;
    const int j = g.x_min + bk.x * 63 + lane;
    const int k = g.y_min + bk.y * (AW - 1) + w;
    int l0, l1;
    chunk_range(bk, g.z_min, g.z_max + 1, l0, l1);
    if (l0 > l1) return;  // the whole workgroup, before any barrier
    const bool node = j <= g.x_max + 1 && k <= g.y_max + 1;  // j >= x_min, k >= y_min always
    const bool own = node && lane < 63 && w < AW - 1;
    const bool jin = j <= g.x_max, kin = k <= g.y_max;
    const T xa = (T)m.xa, ya = (T)m.ya, za = (T)m.za, vol = (T)m.vol, tdt = (T)dt;
    const T c0 = (T)0.125 * tdt;
    const int ln = min(lane + 1, 63), wn = min(w + 1, AW - 1);
    int i = idx(g, j, k, l0);
    // accelerate's cells of the plane below the node plane
    QuadS<S> dlo{}, plo{}, qlo{};
    // the next plane's loads, issued one plane ahead of their use: its cells
    // (accelerate), the node's level-0 velocity and the cell's energy0
    QuadS<S> dn{}, pn{}, qn{};
    S u0n = 0, v0n = 0, w0n = 0, e0n = 0;
    // mirrored faces: the node's cells j-1 (x_min face) or j (x_max face),
    // k-1 or k (y faces) and the planes below z_min or above z_max are halo
    // cells, read as the interior cells they mirror
    const bool mlo = cell_m(g, 0, j - 1) != j - 1, mhi = cell_m(g, 0, j) != j;
    const bool ylo = cell_m(g, 1, k - 1) != k - 1, yhi = cell_m(g, 1, k) != k;
    auto load = [&](int ii, int lp) {
        dn = quad_down_raw(density0, ii, g.ld);
        pn = quad_down_raw(pressure, ii, g.ld);
        qn = quad_down_raw(viscosity, ii, g.ld);
        u0n = xvel0[ii];
        v0n = yvel0[ii];
        w0n = zvel0[ii];
        e0n = (lp <= g.z_max && jin && kin) ? energy0[ii] : (S)0;
    };
    if (node) {
        dlo = quad_m(quad_down_raw(density0, i - g.plane, g.ld), mlo, mhi, ylo, yhi);
        plo = quad_m(quad_down_raw(pressure, i - g.plane, g.ld), mlo, mhi, ylo, yhi);
        qlo = quad_m(quad_down_raw(viscosity, i - g.plane, g.ld), mlo, mhi, ylo, yhi);
        load(i, l0);
    }
    T e0p = 0;  // energy0 of the cell below
    const int lend = min(l1 + 1, g.z_max + 1);
    int s = 0, sp = 1;  // ring slots of the node plane in hand and the one below
    for (int lp = l0; lp <= lend; lp++, i += g.plane) {
        QuadS<S> dhi = quad_m(dn, mlo, mhi, ylo, yhi), phi = quad_m(pn, mlo, mhi, ylo, yhi),
                 qhi = quad_m(qn, mlo, mhi, ylo, yhi);
        // mirrored z faces: the cell plane below z_min is read as plane
        // z_min, the one above z_max as plane z_max
        if (cell_m(g, 2, lp - 1) != lp - 1) {
            dlo = dhi;
            plo = phi;
            qlo = qhi;
        }
        if (cell_m(g, 2, lp) != lp) {
            dhi = dlo;
            phi = plo;
            qhi = qlo;
        }
        const S u0 = u0n, v0 = v0n, w0 = w0n;
        const T e0q = (T)e0n;
        // the next plane's loads go out before this plane's work, so they
        // have the whole iteration to arrive
        if (node && lp < lend) load(i + g.plane, lp + 1);
        S u1 = 0, v1 = 0, w1 = 0;
        if (node) {
            accel_node<S>(dhi, phi, qhi, dlo, plo, qlo, u0, v0, w0, vol, xa, ya, za, tdt, u1, v1, w1);
            if (own && lp <= l1) {
                xvel1[i] = u1;
                yvel1[i] = v1;
                zvel1[i] = w1;
            }
        }
        nb[s][0][w][lane] = u0;
        nb[s][1][w][lane] = v0;
        nb[s][2][w][lane] = w0;
        nb[s][3][w][lane] = u1;
        nb[s][4][w][lane] = v1;
        nb[s][5][w][lane] = w1;
        __syncthreads();
        const S(*lo)[AW][64] = nb[sp];
        const S(*up)[AW][64] = nb[s];
        if (own && lp > l0) {  // plane lp-1, nodes of planes lp-1 and lp; its cell is the (j, k) corner
            const auto at = [&](int c, int a, int bb, int pl) { return (pl ? up : lo)[c][bb ? wn : w][a ? ln : lane]; };
            corrector_plane<S>(i - g.plane, jin, kin, true, at, dlo.v[3], e0p, plo.v[3], qlo.v[3], xa, ya, za, vol,
                               tdt, c0, density1, energy1, vol_flux_x, vol_flux_y, vol_flux_z);
        }
        // the top z faces (plane z_max+1) use only their own node plane
        if (own && lp == lend && l1 == g.z_max + 1) {
            const auto at = [&](int c, int a, int bb, int) { return up[c][bb ? wn : w][a ? ln : lane]; };
            corrector_plane<S>(i, jin, kin, false, at, (T)0, (T)0, (T)0, (T)0, xa, ya, za, vol, tdt, c0, density1,
                               energy1, vol_flux_x, vol_flux_y, vol_flux_z);
        }
        dlo = dhi;
        plo = phi;
        qlo = qhi;
        e0p = e0q;
        sp = s;
        s ^= 1;
        __syncthreads();  // the reads of slot sp's plane end before plane lp+1 rewrites it
    }
}

template <typename S>
__global__ void __launch_bounds__(64 * AW, sizeof(S) == 4 ? 6 : 1)
    k_fast_accel_pdv(int xcd, Lay g, Geo m, double dt, const S *__restrict__ density0, S *__restrict__ density1,
                     const S *__restrict__ energy0, S *__restrict__ energy1, const S *__restrict__ pressure,
                     const S *__restrict__ viscosity, const S *__restrict__ xvel0, S *__restrict__ xvel1,
                     const S *__restrict__ yvel0, S *__restrict__ yvel1, const S *__restrict__ zvel0,
                     S *__restrict__ zvel1, S *__restrict__ vol_flux_x, S *__restrict__ vol_flux_y,
                     S *__restrict__ vol_flux_z)
{
    accel_pdv_body<S>(hw_blk(xcd), g, m, dt, density0, density1, energy0, energy1, pressure, viscosity, xvel0, xvel1,
                      yvel0, yvel1, zvel0, zvel1, vol_flux_x, vol_flux_y, vol_flux_z);
}

template <typename S>
static hipError_t fast_accel_pdv_t(const Lay &g, const Geo &m, double dt, const S *density0, S *density1,
                                   const S *energy0, S *energy1, const S *pressure, const S *viscosity,
                                   const S *xvel0, S *xvel1, const S *yvel0, S *yvel1, const S *zvel0, S *zvel1,
                                   S *vol_flux_x, S *vol_flux_y, S *vol_flux_z, const Shape &sh, hipStream_t st)
{
    const int nx = g.x_max - g.x_min + 2, ny = g.y_max - g.y_min + 2, nz = g.z_max - g.z_min + 2;  // nodes
    const dim3 grid((nx + 62) / 63, (ny + AW - 2) / (AW - 1), nchunks(nz, sh.planes));
    TIMED_LAUNCH("k_fast_accel_pdv", st, k_fast_accel_pdv<S>, grid, dim3(64, AW, 1), 0, st, sh.xcd, g, m, dt,
                 density0, density1, energy0, energy1, pressure, viscosity, xvel0, xvel1, yvel0, yvel1, zvel0, zvel1,
                 vol_flux_x, vol_flux_y, vol_flux_z);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// storage dispatch
#define CP(T, p) static_cast<const T *>(p)
#define MP(T, p) static_cast<T *>(p)

hipError_t launch_fast_ideal_gas(Store s, const Lay &g, const void *density, const void *energy, void *pressure,
                                 void *soundspeed, hipStream_t st)
{
#define CALL(T) fast_ideal_gas_t<T>(g, CP(T, density), CP(T, energy), MP(T, pressure), MP(T, soundspeed), st)
    return s == ST_F32 ? CALL(float) : CALL(double);
#undef CALL
}

hipError_t launch_fast_visc_dt(Store s, const Lay &g, const Geo &m, double dtc_safe, double dtu_safe,
                               double dtv_safe, double dtw_safe, double dtdiv_safe, double g_small,
                               const void *density0, const void *pressure, void *viscosity, const void *xvel0,
                               const void *yvel0, const void *zvel0, double *partial, int *nparts, const Shape &sh,
                               hipStream_t st)
{
#define CALL(T)                                                                                                 \
    fast_visc_dt_t<T>(g, m, dtc_safe, dtu_safe, dtv_safe, dtw_safe, dtdiv_safe, g_small, CP(T, density0),      \
                      CP(T, pressure), MP(T, viscosity), CP(T, xvel0), CP(T, yvel0), CP(T, zvel0), partial,     \
                      nparts, sh, st)
    return s == ST_F32 ? CALL(float) : CALL(double);
#undef CALL
}

hipError_t launch_fast_pdv_predict(Store s, const Lay &g, const Geo &m, double dt, const void *density0,
                                   const void *energy0, void *pressure, const void *viscosity, const void *xvel0,
                                   const void *yvel0, const void *zvel0, const Shape &sh, hipStream_t st)
{
#define CALL(T)                                                                                                 \
    fast_pdv_predict_t<T>(g, m, dt, CP(T, density0), CP(T, energy0), MP(T, pressure), CP(T, viscosity),         \
                          CP(T, xvel0), CP(T, yvel0), CP(T, zvel0), sh, st)
    return s == ST_F32 ? CALL(float) : CALL(double);
#undef CALL
}

hipError_t launch_fast_accel_pdv(Store s, const Lay &g, const Geo &m, double dt, const void *density0,
                                 void *density1, const void *energy0, void *energy1, const void *pressure,
                                 const void *viscosity, const void *xvel0, void *xvel1, const void *yvel0,
                                 void *yvel1, const void *zvel0, void *zvel1, void *vol_flux_x, void *vol_flux_y,
                                 void *vol_flux_z, const Shape &sh, hipStream_t st)
{
#define CALL(T)                                                                                                 \
    fast_accel_pdv_t<T>(g, m, dt, CP(T, density0), MP(T, density1), CP(T, energy0), MP(T, energy1),             \
                        CP(T, pressure), CP(T, viscosity), CP(T, xvel0), MP(T, xvel1), CP(T, yvel0),             \
                        MP(T, yvel1), CP(T, zvel0), MP(T, zvel1), MP(T, vol_flux_x), MP(T, vol_flux_y),          \
                        MP(T, vol_flux_z), sh, st)
    return s == ST_F32 ? CALL(float) : CALL(double);
#undef CALL
}

hipError_t launch_fast_field_summary(Store s, const Lay &g, const Geo &m, const void *density0, const void *energy0,
                                     const void *pressure, const void *xvel0, const void *yvel0, const void *zvel0,
                                     double *partial, int *nparts, hipStream_t st)
{
#define CALL(T)                                                                                                 \
    fast_field_summary_t<T>(g, m, CP(T, density0), CP(T, energy0), CP(T, pressure), CP(T, xvel0), CP(T, yvel0), \
                            CP(T, zvel0), partial, nparts, st)
    return s == ST_F32 ? CALL(float) : CALL(double);
#undef CALL
}
#undef CP
#undef MP

// ---------------------------------------------------------------------------
// Batched launches over many decks (batch.h): each deck's blocks run the
// bodies above with the deck's extents, mesh, dt and arrays.
__host__ __device__ inline void dims_ideal(const Lay &g, int &gx, int &gy, int &gz)
{
    gx = (g.x_max - g.x_min + BX) / BX;
    gy = (g.y_max - g.y_min + BY) / BY;
    gz = g.z_max - g.z_min + 1;
}

// cpt: cells per thread along x (2 in the FP32 visc_dt and predictor)
__host__ __device__ inline void dims_cells(const Lay &g, int planes, int cpt, int &gx, int &gy, int &gz)
{
    gx = (g.x_max - g.x_min + cpt * BX) / (cpt * BX);
    gy = (g.y_max - g.y_min + BY) / BY;
    gz = nchunks(g.z_max - g.z_min + 1, planes);
}

__host__ __device__ inline void dims_accel_pdv(const Lay &g, int planes, int &gx, int &gy, int &gz)
{
    gx = (g.x_max - g.x_min + 2 + 62) / 63;
    gy = (g.y_max - g.y_min + 2 + AW - 2) / (AW - 1);
    gz = nchunks(g.z_max - g.z_min + 2, planes);
}

__host__ __device__ inline int summary_blocks(const Lay &g)
{
    const int ncells = (g.x_max - g.x_min + 1) * (g.y_max - g.y_min + 1) * (g.z_max - g.z_min + 1);
    int nb = (ncells + RED_THREADS * 4 - 1) / (RED_THREADS * 4);
    if (nb < 1) nb = 1;
    if (nb > kMaxPartials) nb = kMaxPartials;
    return nb;
}

__device__ inline Blk local_blk(int lb, int gx, int gy, int gz)
{
    return Blk{lb % gx, (lb / gx) % gy, lb / (gx * gy), gx, gy, gz};
}

// the deck and the block within it of tile hb of the launch
#define DECK_TILE(hb, dims_call)                                                            \
    const int d = deck_of(start, nd, hb);                                                   \
    const DeckDev &D = decks[d];                                                            \
    int gx, gy, gz;                                                                         \
    dims_call;                                                                              \
    const Blk bk = local_blk((hb) - start[d], gx, gy, gz);
// ... of this workgroup's only tile (a launch with a workgroup per tile)
#define DECK_BLOCK(dims_call, xcd)                                                          \
    const int hb = (xcd) ? xcd_block((int)blockIdx.x, (int)gridDim.x) : (int)blockIdx.x;     \
    DECK_TILE(hb, dims_call)
// the tiles this workgroup walks (tile_walk); each tile's body ends with a
// barrier, since the next one reuses the workgroup's LDS
#define FOR_TILES(xcd)                                                                      \
    int t_, end_, step_;                                                                    \
    tile_walk(start[nd], xcd, t_, end_, step_);                                             \
    for (; t_ < end_; t_ += step_)
#define CF(id) static_cast<const S *>(D.f[id])
#define MF(id) static_cast<S *>(D.f[id])

template <typename S>
__global__ void __launch_bounds__(BX *BY)
kb_ideal_gas(const DeckDev *__restrict__ decks, const int *__restrict__ start, int nd)
{
    DECK_BLOCK(dims_ideal(D.g, gx, gy, gz), 0)
    const Lay g = D.g;
    const Range r{g.x_min, g.x_max, g.y_min, g.y_max, g.z_min, g.z_max};
    ideal_gas_body<S>(bk, g, r, CF(F_DENSITY0), CF(F_ENERGY0), MF(F_PRESSURE), MF(F_SOUNDSPEED));
}

// three waves per SIMD in FP64 (168 VGPRs), which the compiler would miss by one register
template <typename S>
__global__ void __launch_bounds__(BX *BY) __attribute__((amdgpu_waves_per_eu(3)))
kb_visc_dt(const DeckDev *__restrict__ decks, const int *__restrict__ start, int nd, int planes, int xcd,
           double g_small, double *__restrict__ partial)
{
    FOR_TILES(xcd)
    {
        DECK_TILE(t_, dims_cells(D.g, planes, sizeof(S) == 4 ? 2 : 1, gx, gy, gz))
        if constexpr (sizeof(S) == 4)
            visc_dt_pair_body(bk, D.g, D.m, D.dtc_safe, D.dtu_safe, D.dtv_safe, D.dtw_safe, D.dtdiv_safe, g_small,
                              CF(F_DENSITY0), CF(F_PRESSURE), MF(F_VISCOSITY), CF(F_XVEL0), CF(F_YVEL0),
                              CF(F_ZVEL0), partial + start[d]);
        else
            visc_dt_body<S>(bk, D.g, D.m, D.dtc_safe, D.dtu_safe, D.dtv_safe, D.dtw_safe, D.dtdiv_safe, g_small,
                            CF(F_DENSITY0), CF(F_PRESSURE), MF(F_VISCOSITY), CF(F_XVEL0), CF(F_YVEL0), CF(F_ZVEL0),
                            partial + start[d]);
        __syncthreads();
    }
}

template <typename S>
__global__ void __launch_bounds__(BX *BY)
kb_pdv_predict(const DeckDev *__restrict__ decks, const int *__restrict__ start, int nd, int planes, int xcd)
{
    FOR_TILES(xcd)
    {
        DECK_TILE(t_, dims_cells(D.g, planes, sizeof(S) == 4 ? 2 : 1, gx, gy, gz))
        if constexpr (sizeof(S) == 4)
            pdv_predict_pair_body(bk, D.g, D.m, D.dt, CF(F_DENSITY0), CF(F_ENERGY0), MF(F_PRESSURE),
                                  CF(F_VISCOSITY), CF(F_XVEL0), CF(F_YVEL0), CF(F_ZVEL0));
        else
            pdv_predict_body<S>(bk, D.g, D.m, D.dt, CF(F_DENSITY0), CF(F_ENERGY0), MF(F_PRESSURE),
                                CF(F_VISCOSITY), CF(F_XVEL0), CF(F_YVEL0), CF(F_ZVEL0));
        __syncthreads();
    }
}

template <typename S>
// held to six waves per SIMD in FP32 like k_fast_accel_pdv (80 VGPRs, no spill here)
__global__ void __launch_bounds__(64 * AW, sizeof(S) == 4 ? 6 : 1)
kb_accel_pdv(const DeckDev *__restrict__ decks, const int *__restrict__ start, int nd, int planes, int xcd)
{
    FOR_TILES(xcd)
    {
        DECK_TILE(t_, dims_accel_pdv(D.g, planes, gx, gy, gz))
        accel_pdv_body<S>(bk, D.g, D.m, D.dt, CF(F_DENSITY0), MF(F_DENSITY1), CF(F_ENERGY0), MF(F_ENERGY1),
                          CF(F_PRESSURE), CF(F_VISCOSITY), CF(F_XVEL0), MF(F_XVEL1), CF(F_YVEL0), MF(F_YVEL1),
                          CF(F_ZVEL0), MF(F_ZVEL1), MF(F_VOL_FLUX_X), MF(F_VOL_FLUX_Y), MF(F_VOL_FLUX_Z));
        __syncthreads();
    }
}

__host__ __device__ inline void dims_summary(const Lay &g, int &gx, int &gy, int &gz)
{
    gx = summary_blocks(g);
    gy = gz = 1;
}

template <typename S, bool COMP>
__global__ void __launch_bounds__(RED_THREADS)
kb_field_summary(const DeckDev *__restrict__ decks, const int *__restrict__ start, int nd,
                 double *__restrict__ partial)
{
    DECK_BLOCK(dims_summary(D.g, gx, gy, gz), 0)
    const Lay g = D.g;
    const int nj = g.x_max - g.x_min + 1, nk = g.y_max - g.y_min + 1, nl = g.z_max - g.z_min + 1;
    field_summary_body<S, COMP>(bk, g, D.m, nj, nk, nj * nk * nl, CF(F_DENSITY0), CF(F_ENERGY0), CF(F_PRESSURE),
                                CF(F_XVEL0), CF(F_YVEL0), CF(F_ZVEL0), partial + 5 * start[d]);
}
#undef DECK_BLOCK
#undef DECK_TILE
#undef FOR_TILES
#undef CF
#undef MF

int batch_blocks_mom(int dir, const Lay &g, int planes);
int batch_blocks_cell(int dir, const Lay &g, int planes);

int batch_blocks(BatchKernel k, const Lay &g, int planes, Store s)
{
    int gx = 1, gy = 1, gz = 1;
    switch (k) {
    case BK_IDEAL_GAS: dims_ideal(g, gx, gy, gz); break;
    case BK_VISC_DT:
    case BK_PDV_PREDICT: dims_cells(g, planes, s == ST_F32 ? 2 : 1, gx, gy, gz); break;
    case BK_ACCEL_PDV: dims_accel_pdv(g, planes, gx, gy, gz); break;
    case BK_SUMMARY: dims_summary(g, gx, gy, gz); break;
    case BK_CELL_X: return batch_blocks_cell(1, g, planes);
    case BK_CELL_Y: return batch_blocks_cell(2, g, planes);
    case BK_CELL_Z: return batch_blocks_cell(3, g, planes);
    case BK_MOM_X: return batch_blocks_mom(1, g, planes);
    case BK_MOM_Y: return batch_blocks_mom(2, g, planes);
    case BK_MOM_Z: return batch_blocks_mom(3, g, planes);
    default: return 0;
    }
    return gx * gy * gz;
}

#define BLAUNCH(name, kern, block, ...)                                                                       \
    do {                                                                                                     \
        if (bg.total > 0) TIMED_LAUNCH(name, st, kern, dim3(bg.total), block, 0, st, __VA_ARGS__);           \
    } while (0)
// shaped kernels: persistent when the shape asks for fewer workgroups than tiles
#define BLAUNCH_SH(name, kern, block, ...)                                                                    \
    do {                                                                                                     \
        if (bg.total > 0) TIMED_LAUNCH(name, st, kern, dim3(batch_wgs(bg, sh)), block, 0, st, __VA_ARGS__);  \
    } while (0)

hipError_t batch_ideal_gas(Store s, const DeckDev *decks, int nd, const BatchGrid &bg, hipStream_t st)
{
    if (s == ST_F32) BLAUNCH("kb_ideal_gas", kb_ideal_gas<float>, dim3(BX, BY), decks, bg.start, nd);
    else BLAUNCH("kb_ideal_gas", kb_ideal_gas<double>, dim3(BX, BY), decks, bg.start, nd);
    return hipGetLastError();
}

hipError_t batch_visc_dt(Store s, const DeckDev *decks, int nd, const BatchGrid &bg, const Shape &sh, double g_small,
                         double *partial, hipStream_t st)
{
    if (s == ST_F32)
        BLAUNCH_SH("kb_visc_dt", kb_visc_dt<float>, dim3(BX, BY), decks, bg.start, nd, sh.planes, sh.xcd, g_small,
                   partial);
    else
        BLAUNCH_SH("kb_visc_dt", kb_visc_dt<double>, dim3(BX, BY), decks, bg.start, nd, sh.planes, sh.xcd, g_small,
                   partial);
    return hipGetLastError();
}

hipError_t batch_pdv_predict(Store s, const DeckDev *decks, int nd, const BatchGrid &bg, const Shape &sh, hipStream_t st)
{
    if (s == ST_F32) BLAUNCH_SH("kb_pdv_predict", kb_pdv_predict<float>, dim3(BX, BY), decks, bg.start, nd, sh.planes, sh.xcd);
    else BLAUNCH_SH("kb_pdv_predict", kb_pdv_predict<double>, dim3(BX, BY), decks, bg.start, nd, sh.planes, sh.xcd);
    return hipGetLastError();
}

hipError_t batch_accel_pdv(Store s, const DeckDev *decks, int nd, const BatchGrid &bg, const Shape &sh, hipStream_t st)
{
    if (s == ST_F32) BLAUNCH_SH("kb_accel_pdv", kb_accel_pdv<float>, dim3(64, AW), decks, bg.start, nd, sh.planes, sh.xcd);
    else BLAUNCH_SH("kb_accel_pdv", kb_accel_pdv<double>, dim3(64, AW), decks, bg.start, nd, sh.planes, sh.xcd);
    return hipGetLastError();
}

hipError_t batch_field_summary(Store s, const DeckDev *decks, int nd, const BatchGrid &bg, double *partial,
                               hipStream_t st)
{
    if (s == ST_F32)
        BLAUNCH("kb_field_summary", (kb_field_summary<float, true>), dim3(RED_THREADS), decks, bg.start, nd, partial);
    else
        BLAUNCH("kb_field_summary", (kb_field_summary<double, false>), dim3(RED_THREADS), decks, bg.start, nd,
                partial);
    return hipGetLastError();
}
#undef BLAUNCH
#undef BLAUNCH_SH

}  // namespace clover
