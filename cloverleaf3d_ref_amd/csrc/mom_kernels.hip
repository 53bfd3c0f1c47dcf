// advec_mom (advec_mom_kernel.f90:29-428) for one sweep and all three
// velocity components, as one kernel that keeps every intermediate on chip.
//
// Per sweep along axis s the reference forms, node by node,
//   node_flux     nf(n)  = 0.125 * (8 mass fluxes around n)
//   node_mass_post  mp(n) = 0.125 * (8 x density1 * post_vol of the cells around n)
//   node_mass_pre   mq(n) = mp(n) - nf(n-s) + nf(n)
//   mom_flux      F(f)   = advec_vel_s(f) * nf(f)      (face f between n and n+s)
//   vel(n) <- (vel(n) * mq(n) + F(n-s) - F(n)) / mp(n)
// and needs, for node n, nf at n-2s..n+s, mp at n-s..n+s and vel at n-2s..n+2s.
// The per-call path stores nf, mp, mq and mom_flux in work arrays; here
// every value is formed once, in registers, with the same expression and
// operand order (so the FP64 form's results are bit-identical; the FP32 form
// runs the same expressions in FP32), and shared with the neighbours that
// need it:
//   * lanes of a wavefront lie along x (coalesced rows); neighbours along x
//     come from lane shuffles;
//   * the waves of a workgroup lie along the second axis; a neighbour row is
//     read from LDS (double-buffered, one barrier per plane).  Wave 0 is the
//     halo row and stores nothing;
//   * the workgroup marches along the third axis, keeping the previous
//     planes' values in registers.
// The y and z sweeps march along the sweep axis itself, so the +-2 window
// along s is a register pipeline; the x sweep marches along z and takes its
// +-2 window from the lanes (60 of 64 lanes store).
//
// Each face's momentum flux is formed once (the per-call kernel, like the
// reference, forms mom_flux once per face too) and each input element is
// read about once, instead of the 8 density/volume reads per node mass and
// 16 mass-flux reads per pair of node fluxes of a plain per-node kernel.
#include <hip/hip_runtime.h>

#include "batch.h"
#include "fast.h"
#include "launch.h"

namespace clover {

// waves per workgroup (wave 0 is the halo row) of the x (DIR 1), y (2) and
// z (3) sweeps: the y/z marches run faster in 4-wave workgroups despite the
// larger share of halo rows, the x sweep in 8-wave ones
constexpr int mw_of(int dir) { return dir == 1 ? 8 : 4; }

template <typename S>
struct Vel3 {
    const S *in[3];
    S *out[3];
};

// post_vol of advec_mom_kernel.f90:95-163 for the sweep `mode`:
// 1 sweep 1 along x, 2 sweep 1 along z, 3 sweep 2 with advect_x,
// 4 sweep 2 without, 5 and 6 sweep 3
template <typename S>
// (dxt: the offset of the column vol_flux_y and vol_flux_z are read at, on a
// mirrored x face, launch.h tan_xm)
__device__ inline S mom_post_vol(int mode, const Lay &g, int i, S V0, const S *__restrict__ vfx,
                                 const S *__restrict__ vfy, const S *__restrict__ vfz, int dxt)
{
#define D(x) (x)
    const int t = i + dxt;
    switch (mode) {
    case 1: return V0 + D(vfy[t + g.ld]) - D(vfy[t]) + D(vfz[t + g.plane]) - D(vfz[t]);
    case 2: return V0 + D(vfx[i + 1]) - D(vfx[i]) + D(vfy[t + g.ld]) - D(vfy[t]);
    case 3: return V0 + D(vfz[t + g.plane]) - D(vfz[t]);
    case 4: return V0 + D(vfx[i + 1]) - D(vfx[i]);
    default: return V0;
    }
#undef D
}

// momentum flux through the face between nodes n and n+s
// (advec_mom_kernel.f90:213-249 and its y/z twins); vm1..vp2 are the
// velocity at n-s, n, n+s, n+2s; sigma is |nf| / node_mass_pre(donor); d the
// node spacing along the sweep (rd: the same as a divisor), r6: 6 as a divisor
template <typename T>
__device__ inline T mom_face_flux(T nf, T sigma, T d, const Recip &rd, const Recip &r6, T vm1, T v0, T vp1, T vp2)
{
    T upwind, donor, downwind;
    if (nf < (T)0) {
        upwind = vp2; donor = vp1; downwind = v0;
    } else {
        upwind = vm1; donor = v0; downwind = vp1;
    }
    const T width = d;
    const T vdiffuw = donor - upwind;
    const T vdiffdw = downwind - donor;
    T limiter = 0;
    if (vdiffuw * vdiffdw > (T)0) {
        const T auw = fabs(vdiffuw);
        const T adw = fabs(vdiffdw);
        T wind = 1;
        if (vdiffdw <= (T)0) wind = -1;
        limiter = wind * tmin(tmin(by_const<T>(width * (by_const<T>(((T)2 - sigma) * adw, rd) +
                                                        by_const<T>(((T)1 + sigma) * auw, rd)),
                                               r6),
                                   auw),
                              adw);
    }
    const T advec_vel_s = donor + ((T)1 - sigma) * limiter;
    return advec_vel_s * nf;
}

template <typename T>
__device__ inline T face_sigma(T nf, T mq_lo, T mq_hi)
{
    return quo(fabs(nf), nf < (T)0 ? mq_hi : mq_lo);
}

// A value and its neighbours at -1 along the lane axis (or +1, see
// exchange_x) and -1 along the wave axis.
template <typename T>
struct Quad {
    T c, x, w, xw;
};

// Publishes a and p of every thread of the workgroup and returns each
// thread's neighbours: a at lane+dxa, p at lane+dxp, both also one wave
// down.  Lanes and waves without such a neighbour get their own values
// (used only by threads that store nothing).
template <int DXA, int DXP, typename T, int W>
__device__ inline void exchange(T (*buf)[W][64], int lane, int w, T a, T p, Quad<T> &A, Quad<T> &P)
{
    buf[0][w][lane] = a;
    buf[1][w][lane] = p;
    __syncthreads();
    const int la = min(max(lane + DXA, 0), 63), lp = min(max(lane + DXP, 0), 63), wd = w > 0 ? w - 1 : 0;
    A.c = a;
    A.x = buf[0][w][la];
    A.w = buf[0][wd][lane];
    A.xw = buf[0][wd][la];
    P.c = p;
    P.x = buf[1][w][lp];
    P.w = buf[1][wd][lane];
    P.xw = buf[1][wd][lp];
}

// The loads of one plane, kept raw so that they can be issued a plane ahead
// (after the barrier of the plane in hand) and consumed by the next plane:
// the sweep's mass flux, density1 and the volume fluxes post_vol needs.
template <typename S>
struct Raw {
    S a, rho, v0, v1, v2, v3;  // storage type: converted where used, so the loads stay in flight
};

// dxc: offset of the density read (a mirrored x face, launch.h cell_xm);
// dxt: of the vol_flux_y and vol_flux_z reads (tan_xm); dma: of the mass
// flux read (the same in the y and z sweeps, whose mass fluxes follow
// tan_xm too)
template <typename S>
__device__ inline Raw<S> load_raw(int mode, const Lay &g, int i, bool ok_a, bool ok_p, const S *__restrict__ mf,
                                  const S *__restrict__ vfx, const S *__restrict__ vfy, const S *__restrict__ vfz,
                                  const S *__restrict__ density1, int dxc, int dxt, int dma)
{
    Raw<S> r{0, 0, 0, 0, 0, 0};
    if (ok_a) r.a = mf[i + dma];
    if (ok_p) {
        r.rho = density1[i + dxc];
        const int t = i + dxt;
        switch (mode) {
        case 1: r.v0 = vfy[t + g.ld]; r.v1 = vfy[t]; r.v2 = vfz[t + g.plane]; r.v3 = vfz[t]; break;
        case 2: r.v0 = vfx[i + 1]; r.v1 = vfx[i]; r.v2 = vfy[t + g.ld]; r.v3 = vfy[t]; break;
        case 3: r.v0 = vfz[t + g.plane]; r.v1 = vfz[t]; break;
        case 4: r.v0 = vfx[i + 1]; r.v1 = vfx[i]; break;
        default: break;
        }
    }
    return r;
}

// density1 * post_vol with mom_post_vol's expression order
template <typename S>
__device__ inline S raw_cell_mass(int mode, S V0, const Raw<S> &r, bool ok_p)
{
    if (!ok_p) return 0;
    S pv;
    switch (mode) {
    case 1:
    case 2: pv = V0 + r.v0 - r.v1 + r.v2 - r.v3; break;
    case 3:
    case 4: pv = V0 + r.v0 - r.v1; break;
    default: pv = V0; break;
    }
    return r.rho * pv;
}

// exchange() in two halves, so loads can be issued between them
template <typename T, int W>
__device__ inline void publish(T (*buf)[W][64], int lane, int w, T a, T p)
{
    buf[0][w][lane] = a;
    buf[1][w][lane] = p;
    __syncthreads();
}

template <int DXA, int DXP, typename T, int W>
__device__ inline void gather(T (*buf)[W][64], int lane, int w, T a, T p, Quad<T> &A, Quad<T> &P)
{
    const int la = min(max(lane + DXA, 0), 63), lp = min(max(lane + DXP, 0), 63), wd = w > 0 ? w - 1 : 0;
    A.c = a;
    A.x = buf[0][w][la];
    A.w = buf[0][wd][lane];
    A.xw = buf[0][wd][la];
    P.c = p;
    P.x = buf[1][w][lp];
    P.w = buf[1][wd][lane];
    P.xw = buf[1][wd][lp];
}

__device__ inline double shfl_d(double x, int src) { return __shfl(x, src, 64); }
__device__ inline float shfl_d(float x, int src) { return __shfl(x, src, 64); }

// ---------------------------------------------------------------------------
// y and z sweeps: lanes along x on 64 line-aligned columns, all of which
// store; lane 0 also loads and publishes column j-1 (publish65), which the
// first column's node needs.  (A 63-column layout with lane 0 as the j-1
// halo read rows that straddle cache lines: measured at 512^3 the aligned
// one runs FP32 mom_y 1.35 -> 1.28 ms and mom_z 1.33 -> 1.20 ms, FP64 the
// same.)  Waves lie along the other axis; the workgroup marches along the
// sweep axis.
//
// In (x, m = sweep axis, w = wave axis) offsets both sweeps have the same
// node flux (advec_mom_kernel.f90:260-265 for y, 345-350 for z):
//   nf(m) = 0.125*(f(-1,0,0)+f(0,0,0)+f(-1,1,0)+f(0,1,0)+f(-1,0,-1)+f(0,0,-1)+f(-1,1,-1)+f(0,1,-1))
// and node_mass_post (advec_mom_kernel.f90:267-275, 352-360) sums the cells
// (dj,dk,dl) = (0,-1,0),(0,0,0),(-1,-1,0),(-1,0,0),(0,-1,-1),(0,0,-1),(-1,-1,-1),(-1,0,-1),
// which is a different order in (x, m, w) for the two sweeps.
template <int DIR, typename T>
__device__ inline T march_node_flux(const Quad<T> &f0, const Quad<T> &f1)
{
    return (T)0.125 * (f0.x + f0.c + f1.x + f1.c + f0.xw + f0.w + f1.xw + f1.w);
}

template <int DIR, typename T>
__device__ inline T march_node_mass(const Quad<T> &pm, const Quad<T> &p0)  // cell planes m-1, m
{
    if (DIR == 2)  // k = m, l = w
        return (T)0.125 * (pm.c + p0.c + pm.x + p0.x + pm.w + p0.w + pm.xw + p0.xw);
    // DIR == 3: k = w, l = m
    return (T)0.125 * (p0.w + p0.c + p0.xw + p0.x + pm.w + pm.c + pm.xw + pm.x);
}

// The march's x exchange with 64 owned lanes: slot s of a row holds column
// j0-1+s, lane L publishes its own column at slot L+1 and lane 0 also the
// column j0-1 at slot 0.
template <typename T, int W>
__device__ inline void publish65(T (*buf)[W][65], int lane, int w, T a, T p, T am, T pm)
{
    buf[0][w][lane + 1] = a;
    buf[1][w][lane + 1] = p;
    if (lane == 0) {
        buf[0][w][0] = am;
        buf[1][w][0] = pm;
    }
    __syncthreads();
}

template <typename T, int W>
__device__ inline void gather65(T (*buf)[W][65], int lane, int w, T a, T p, Quad<T> &A, Quad<T> &P)
{
    const int wd = w > 0 ? w - 1 : 0;
    A.c = a;
    A.x = buf[0][w][lane];
    A.w = buf[0][wd][lane + 1];
    A.xw = buf[0][wd][lane];
    P.c = p;
    P.x = buf[1][w][lane];
    P.w = buf[1][wd][lane + 1];
    P.xw = buf[1][wd][lane];
}

// The last x tile of a sweep holds the node columns left over by the
// 64-column tiles: one (x_max+1) on a mesh a multiple of 64 cells wide.  In
// the lane-along-x layout its workgroups march every plane with one live
// lane, and at 512^3 they held 1/9 of the launch's workgroup slots (the y and
// z sweeps ran 8 % faster on 511^3, whose nodes fill 8 tiles).  In the FP64
// form a tail of at most kTailCols columns is marched instead with lanes
// along the wave axis: one wave per column and 63 rows (lane 0 is the halo
// row), each lane loading both x neighbours itself and taking the row below
// from lane-1, in the grid's first workgroups (mom_march).  Every value is
// formed by the same expressions from the same operands as in the 64-column
// layout.  Measured at 512^3 FP64: mom_y 2.38 -> 2.32 ms, mom_z 2.26 -> 2.22
// ms; 640^3 as 8 tiles (320^3 tiles, one tail column in six tiles): the
// batched momentum launches 4.41 -> 4.16 ms.  The FP32 form keeps the
// 64-column layout: there the tail's loads (a line per lane, 4 bytes used)
// cost more than the slots save (mom_y 1.29 -> 1.35 ms, mom_z 1.20 -> 1.25
// ms, batched 2.29 -> 2.64 ms).
constexpr int kTailCols = 4;
static_assert(kTailCols <= mw_of(2) && kTailCols <= mw_of(3), "the tail's waves fit in the tile's workgroups");

__host__ __device__ inline int mom_tail_cols(const Lay &g, int gx)
{
    const int tw = g.x_max - g.x_min + 2 - 64 * (gx - 1);  // node columns of the last x tile
    return tw <= kTailCols ? tw : 0;
}

// The tail's waves march at most about kTailPlanes planes of a chunk each
// (nsub pieces of the chunk, as many as the tile's workgroups hold), so
// that they end with the chunk's 64-column workgroups.
constexpr int kTailPlanes = 16;

struct TailSplit {
    int tw, nrg, nsub, tb;  // columns, row groups, pieces per chunk, workgroups per chunk
};

template <int DIR>
__device__ inline TailSplit mom_tail_split(const Lay &g, const Blk &bk, int tw)
{
    constexpr int MW = mw_of(DIR);
    const int nrow = DIR == 3 ? g.y_max - g.y_min + 2 : g.z_max - g.z_min + 2;  // storing rows
    const int npos = DIR == 3 ? g.z_max - g.z_min + 2 : g.y_max - g.y_min + 2;  // march positions
    TailSplit t;
    t.tw = tw;
    t.nrg = (nrow + 62) / 63;
    const int per = (npos + bk.gz - 1) / bk.gz;  // chunk_range's chunk length
    t.nsub = max(1, min((per + kTailPlanes - 1) / kTailPlanes, bk.gy * MW / (tw * t.nrg)));
    t.tb = (tw * t.nrg * t.nsub + MW - 1) / MW;  // <= bk.gy: tw * nrg <= MW * ceil(nrow / (MW-1))
    return t;
}

// TAN: vol_flux_y/z and the sweep's mass flux (y or z) read at the
// mirrored column on mirrored x faces (tan_xm); the single-tile FP32 chunk
// only, whose exchanges leave those x halos out (fused_halo_plan)
template <int DIR, typename S, bool TAIL, bool TAN = false>
__device__ __forceinline__ void mom_march_body(const Blk &bk, S (*xbuf)[2][mw_of(DIR)][65], TailSplit ts, Lay g,
                                               Geo geo, int mode, const S *__restrict__ mf,
                                               const S *__restrict__ vfx, const S *__restrict__ vfy,
                                               const S *__restrict__ vfz, const S *__restrict__ density1, Vel3<S> v)
{
    using T = S;  // arithmetic type of the form
    constexpr int MW = mw_of(DIR);
    const int lane = threadIdx.x, w = threadIdx.y;
    const int sM = DIR == 3 ? g.plane : g.ld;
    const int wmin = DIR == 3 ? g.y_min : g.z_min, wmax = DIR == 3 ? g.y_max : g.z_max;
    const int mmin = DIR == 3 ? g.z_min : g.y_min, mmax = DIR == 3 ? g.z_max : g.y_max;
    const T d = (T)(DIR == 3 ? geo.dz : geo.dy);
    const Recip rd = DIR == 3 ? geo.rz : geo.ry;
    const T V0 = (T)geo.vol;

    int j, wp, m0, m1;
    chunk_range(bk, mmin, mmax + 1, m0, m1);
    if constexpr (TAIL) {
        // wave item: piece fastest, then row group (rows wmin..wmax+1 store), then column
        const int item = bk.y * MW + w;
        if (item >= ts.tw * ts.nrg * ts.nsub) return;  // this wave: no barrier on this path
        const int sub = item % ts.nsub, rest = item / ts.nsub;
        j = g.x_min + 64 * (bk.gx - 1) + rest / ts.nrg;
        wp = wmin - 1 + (rest % ts.nrg) * 63 + lane;
        const int len = (m1 - m0 + 1 + ts.nsub - 1) / ts.nsub;
        m0 += sub * len;
        m1 = min(m0 + len - 1, m1);
    } else {
        j = g.x_min + bk.x * 64 + lane;
        wp = wmin - 1 + bk.y * (MW - 1) + w;
    }
    if (m0 > m1) return;  // the whole workgroup (the tail form: the wave), before any barrier
    // j >= x_min-1 and wp >= wmin-1 always; upper limits of the arrays read:
    // cells and the sweep's mass flux end at +2 across the sweep, vertices at +3
    const bool ok_cell = j <= g.x_max + 2 && wp <= wmax + 2;
    // column j-1: loaded by lane 0 only in the 64-column layout (lane L-1 holds it for lane L)
    const bool ok_cm = (TAIL || lane == 0) && j - 1 <= g.x_max + 2 && wp <= wmax + 2;
    const bool store = (TAIL ? lane : w) >= 1 && j <= g.x_max + 1 && wp <= wmax + 1;
    const int i0 = DIR == 3 ? idx(g, j, wp, mmin) : idx(g, j, mmin, wp);  // offset at march position mmin
#define OFF(q) (i0 + ((q) - mmin) * sM)

    const int dxc = cell_xm(g, j) - j;  // a mirrored x face: the interior cell its halo copies
    const int dxm = cell_xm(g, j - 1) - (j - 1);
    const int dxt = TAN ? tan_xm(g, j) - j : 0, dxtm = TAN ? tan_xm(g, j - 1) - (j - 1) : 0;  // vol_flux_y/z
    int b = 0;
    // the quads of a plane: columns j-1, j of rows wp-1, wp (Quad x, c, xw, w)
    auto xchg = [&](T a, T p, T am, T pm, Quad<T> &F, Quad<T> &P) {
        if constexpr (TAIL) {
            const int src = lane > 0 ? lane - 1 : 0;  // lane 0 (the halo row) stores nothing
            F.c = a;
            F.x = am;
            F.w = shfl_d(a, src);
            F.xw = shfl_d(am, src);
            P.c = p;
            P.x = pm;
            P.w = shfl_d(p, src);
            P.xw = shfl_d(pm, src);
        } else {
            publish65(xbuf[b], lane, w, a, p, am, pm);
            gather65(xbuf[b], lane, w, a, p, F, P);
            b ^= 1;
        }
    };
    auto plane = [&](int q, Quad<T> &F, Quad<T> &P, int &b) {
        const int i = OFF(q);
        const T a = ok_cell ? (T)mf[i + dxt] : (T)0;  // q <= mmax+3: inside the sweep's face extent
        const T p = (ok_cell && q <= mmax + 2)
                        ? (T)density1[i + dxc] * mom_post_vol(mode, g, i, V0, vfx, vfy, vfz, dxt)
                        : (T)0;
        const T am = ok_cm ? (T)mf[i - 1 + dxtm] : (T)0;
        const T pm = (ok_cm && q <= mmax + 2)
                         ? (T)density1[i - 1 + dxm] * mom_post_vol(mode, g, i - 1, V0, vfx, vfy, vfz, dxtm)
                         : (T)0;
        xchg(a, p, am, pm, F, P);
    };
    auto vel_at = [&](int c, int q) { return store ? (T)v.in[c][OFF(q)] : (T)0; };

    Quad<T> Fa, Pa, Fb, Pb;
    // prologue: nf(m0-2..m0), node masses at m0-1, m0, m0+1, velocities m0-2..m0+1
    plane(m0 - 2, Fa, Pa, b);
    plane(m0 - 1, Fb, Pb, b);
    const T nf_m2 = march_node_flux<DIR, T>(Fa, Fb);
    const T mp_m1 = march_node_mass<DIR, T>(Pa, Pb);
    plane(m0, Fa, Pa, b);
    T nf_lo = march_node_flux<DIR, T>(Fb, Fa);  // nf(m-1)
    T mp0 = march_node_mass<DIR, T>(Pb, Pa);    // mp(m)
    const T mq_m1 = mp_m1 - nf_m2 + nf_lo;
    plane(m0 + 1, Fb, Pb, b);
    T nf0 = march_node_flux<DIR, T>(Fa, Fb);     // nf(m)
    T mp1 = march_node_mass<DIR, T>(Pa, Pb);     // mp(m+1)
    T mq0 = mp0 - nf_lo + nf0;                // mq(m)
    // Fb, Pb hold plane m+1

    T vm1[3], v0[3], vp1[3], flo[3];
    {
        const T sig = face_sigma(nf_lo, mq_m1, mq0);
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const T vm2 = vel_at(c, m0 - 2);
            vm1[c] = vel_at(c, m0 - 1);
            v0[c] = vel_at(c, m0);
            vp1[c] = vel_at(c, m0 + 1);
            flo[c] = mom_face_flux(nf_lo, sig, d, rd, geo.r6, vm2, vm1[c], v0[c], vp1[c]);  // F(m0-1)
        }
    }

    // each step loads its own plane m+2 first.  Loading the next plane's
    // values one step ahead instead held them in registers across the step:
    // at 512^3 that measured FP32 mom_y 1.42 against 1.35 ms and mom_z 1.36
    // against 1.31 ms, and FP64 (125 against 95 VGPRs, four waves per SIMD
    // against five) mom_y 2.55 against 2.43 ms and mom_z 2.47 against 2.33.
    for (int m = m0; m <= m1; m++) {
        const Raw<S> rc = load_raw(mode, g, OFF(m + 2), ok_cell, ok_cell && m + 2 <= mmax + 2, mf, vfx, vfy, vfz,
                                   density1, dxc, dxt, dxt);
        const Raw<S> rcm = load_raw(mode, g, OFF(m + 2) - 1, ok_cm, ok_cm && m + 2 <= mmax + 2, mf, vfx, vfy, vfz,
                                    density1, dxm, dxtm, dxtm);
        T vp2[3];
#pragma unroll
        for (int c = 0; c < 3; c++) vp2[c] = store ? (T)v.in[c][OFF(m + 2)] : (T)0;
        const T a = (T)rc.a, p = raw_cell_mass(mode, V0, rc, ok_cell && m + 2 <= mmax + 2);
        const T am = (T)rcm.a, pm = raw_cell_mass(mode, V0, rcm, ok_cm && m + 2 <= mmax + 2);
        Quad<T> Fc, Pc;
        xchg(a, p, am, pm, Fc, Pc);
        const T nf1 = march_node_flux<DIR, T>(Fb, Fc);  // nf(m+1)
        const T mp2 = march_node_mass<DIR, T>(Pb, Pc);  // mp(m+2), garbage at mmax+3 (never used)
        const T mq1 = mp1 - nf0 + nf1;              // mq(m+1)
        const T sig = face_sigma(nf0, mq0, mq1);
        const int i = OFF(m);
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const T fhi = mom_face_flux(nf0, sig, d, rd, geo.r6, vm1[c], v0[c], vp1[c], vp2[c]);  // F(m)
            if (store) v.out[c][i] = quo(v0[c] * mq0 + flo[c] - fhi, mp0);
            flo[c] = fhi;
            vm1[c] = v0[c];
            v0[c] = vp1[c];
            vp1[c] = vp2[c];
        }
        nf0 = nf1;
        mp0 = mp1;
        mp1 = mp2;
        mq0 = mq1;
        Fb = Fc;
        Pb = Pc;
    }
#undef OFF
}

// One workgroup of a y or z sweep.  With a narrow tail the grid's first
// workgroups (in launch order) march the tail, TailSplit::tb per chunk: its
// marches take longest per plane (each load touches a line per lane), and
// dispatched last they ran past the rest of the launch.  The 64-column
// tiles follow in the grid's order with gx-1 tiles to a row, and the
// workgroups left over exit at once.
template <int DIR, typename S, bool TAN = false>
__device__ __forceinline__ void mom_march(const Blk &bk, Lay g, Geo geo, int mode, const S *__restrict__ mf,
                                          const S *__restrict__ vfx, const S *__restrict__ vfy,
                                          const S *__restrict__ vfz, const S *__restrict__ density1, Vel3<S> v)
{
    __shared__ S xbuf[2][2][mw_of(DIR)][65];
    const int tw = sizeof(S) == 8 ? mom_tail_cols(g, bk.gx) : 0;  // the FP64 form only (mom_tail_cols)
    if (!tw) {
        mom_march_body<DIR, S, false, TAN>(bk, xbuf, TailSplit{}, g, geo, mode, mf, vfx, vfy, vfz, density1, v);
        return;
    }
    const TailSplit ts = mom_tail_split<DIR>(g, bk, tw);
    int lin = bk.x + bk.gx * (bk.y + bk.gy * bk.z);
    if (lin < ts.tb * bk.gz) {
        mom_march_body<DIR, S, true, TAN>(Blk{bk.gx - 1, lin % ts.tb, lin / ts.tb, bk.gx, bk.gy, bk.gz}, xbuf, ts, g,
                                     geo, mode, mf, vfx, vfy, vfz, density1, v);
        return;
    }
    lin -= ts.tb * bk.gz;
    const int mx = bk.gx - 1;  // 64-column tiles
    if (lin >= mx * bk.gy * bk.gz) return;
    mom_march_body<DIR, S, false, TAN>(Blk{lin % mx, (lin / mx) % bk.gy, lin / (mx * bk.gy), bk.gx, bk.gy, bk.gz}, xbuf,
                                  TailSplit{}, g, geo, mode, mf, vfx, vfy, vfz, density1, v);
}

template <int DIR, typename S>
__global__ void __launch_bounds__(64 * mw_of(DIR)) k_fast_mom_march(int xcd, Lay g, Geo geo, int mode, const S *__restrict__ mf, const S *__restrict__ vfx,
                 const S *__restrict__ vfy, const S *__restrict__ vfz,
                 const S *__restrict__ density1, Vel3<S> v)
{
    mom_march<DIR, S, sizeof(S) == 4>(hw_blk(xcd), g, geo, mode, mf, vfx, vfy, vfz, density1, v);
}

// ---------------------------------------------------------------------------
// x sweep: lanes along x with lanes 0-1 and 62-63 as halo (60 store), waves
// along y, march along z.  advec_mom_kernel.f90:172-206:
//   nf(j)  = 0.125*(f(0,-1,0)+f(0,0,0)+f(1,-1,0)+f(1,0,0)+f(0,-1,-1)+f(0,0,-1)+f(1,-1,-1)+f(1,0,-1))
//   mp(j)  = cells (0,-1,0),(0,0,0),(-1,-1,0),(-1,0,0),(0,-1,-1),(0,0,-1),(-1,-1,-1),(-1,0,-1)

template <typename S, bool TAN = false>  // TAN: as in mom_march_body
__device__ __forceinline__ void mom_x_body(const Blk &bk, Lay g, Geo geo, int mode, const S *__restrict__ mf, const S *__restrict__ vfx,
             const S *__restrict__ vfy, const S *__restrict__ vfz, const S *__restrict__ density1,
             Vel3<S> v)
{
    using T = S;  // arithmetic type of the form
    constexpr int MW = mw_of(1);
    __shared__ T xbuf[2][2][MW][64];
    const int lane = threadIdx.x, w = threadIdx.y;
    const T d = (T)geo.dx;
    const Recip rd = geo.rx;
    const T V0 = (T)geo.vol;
    const int j = g.x_min - 2 + bk.x * 60 + lane;
    const int k = g.y_min - 1 + bk.y * (MW - 1) + w;
    int l0, l1;
    chunk_range(bk, g.z_min, g.z_max + 1, l0, l1);
    if (l0 > l1) return;  // the whole workgroup: before any barrier
    // j >= x_min-2 and k >= y_min-1 always.  mass_flux_x ends at x_max+3 /
    // y_max+2, cells at x_max+2 / y_max+2, vertices at x_max+3 / y_max+3.
    const bool ok_face = j <= g.x_max + 3 && k <= g.y_max + 2;
    const bool ok_cell = j <= g.x_max + 2 && k <= g.y_max + 2;
    const bool ok_vert = j <= g.x_max + 3 && k <= g.y_max + 3;
    const bool store = lane >= 2 && lane <= 61 && w >= 1 && j <= g.x_max + 1 && k <= g.y_max + 1;
    const int i0 = idx(g, j, k, g.z_min);
#define OFF(q) (i0 + ((q) - g.z_min) * g.plane)
    // a mirrored x face: the interior cell and node its halos copy, and the
    // sign flip of the x velocity there
    const int dxc = cell_xm(g, j) - j, dxn = node_xm(g, j) - j, dxt = TAN ? tan_xm(g, j) - j : 0;
    const bool xneg = node_xsign(g, j);

    auto plane = [&](int q, Quad<T> &F, Quad<T> &P, int &b) {
        const int i = OFF(q);
        const T a = ok_face ? (T)mf[i] : (T)0;
        const T p = ok_cell ? (T)density1[i + dxc] * mom_post_vol(mode, g, i, V0, vfx, vfy, vfz, dxt) : (T)0;
        exchange<+1, -1>(xbuf[b], lane, w, a, p, F, P);  // mass flux at j+1, cell mass at j-1
        b ^= 1;
    };

    int b = 0;
    Quad<T> Fp, Pp;  // plane l-1
    plane(l0 - 1, Fp, Pp, b);
    const int lm = lane > 0 ? lane - 1 : 0, lp = lane < 63 ? lane + 1 : 63, lp2 = lane < 62 ? lane + 2 : 63;
    // FP64: plane l's loads are issued during plane l-1 (after its barrier);
    // FP32: each plane's at the top of its step, which measured faster at
    // 512^3 (1.24 -> 1.20 ms; FP64 2.19 -> 2.29 ms that way)
    constexpr bool ahead = sizeof(S) == 8;
    Raw<S> rn{0, 0, 0, 0, 0, 0};
    S vn[3] = {0, 0, 0};
    auto load_plane = [&](int q) {
        rn = load_raw(mode, g, OFF(q), ok_face, ok_cell, mf, vfx, vfy, vfz, density1, dxc, dxt, 0);
#pragma unroll
        for (int c = 0; c < 3; c++) vn[c] = ok_vert ? v.in[c][OFF(q) + dxn] : (S)0;
    };
    if (ahead) load_plane(l0);
    for (int l = l0; l <= l1; l++) {
        if (!ahead) load_plane(l);
        const Raw<S> rc = rn;
        T vc[3];
#pragma unroll
        for (int c = 0; c < 3; c++) vc[c] = (T)vn[c];
        if (xneg) vc[0] = -vc[0];  // the reflection of the x velocity (update_halo_kernel.f90:604-1004)
        const T a = (T)rc.a, p = raw_cell_mass(mode, V0, rc, ok_cell);
        publish(xbuf[b], lane, w, a, p);
        if (ahead && l < l1) load_plane(l + 1);
        Quad<T> F, P;
        gather<+1, -1>(xbuf[b], lane, w, a, p, F, P);  // mass flux at j+1, cell mass at j-1
        b ^= 1;
        // F.x = f(+1,0), F.w = f(0,-1), F.xw = f(+1,-1); P.x = cell(-1,0), ...
        const T nf = 0.125 * (F.w + F.c + F.xw + F.x + Fp.w + Fp.c + Fp.xw + Fp.x);
        const T mp = 0.125 * (P.w + P.c + P.xw + P.x + Pp.w + Pp.c + Pp.xw + Pp.x);
        const T nf_lo = shfl_d(nf, lm);
        const T mq = mp - nf_lo + nf;
        const T mq_hi = shfl_d(mq, lp);
        const T sig = face_sigma(nf, mq, mq_hi);
        const int i = OFF(l);
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const T v0 = vc[c];
            const T vm1 = shfl_d(v0, lm), vp1 = shfl_d(v0, lp), vp2 = shfl_d(v0, lp2);
            const T fhi = mom_face_flux(nf, sig, d, rd, geo.r6, vm1, v0, vp1, vp2);  // face j | j+1
            const T flo = shfl_d(fhi, lm);                                // face j-1 | j
            if (store) v.out[c][i] = quo(v0 * mq + flo - fhi, mp);
        }
        Fp = F;
        Pp = P;
    }
#undef OFF
}

template <typename S>
__global__ void __launch_bounds__(64 * mw_of(1)) k_fast_mom_x(int xcd, Lay g, Geo geo, int mode, const S *__restrict__ mf, const S *__restrict__ vfx,
             const S *__restrict__ vfy, const S *__restrict__ vfz, const S *__restrict__ density1,
             Vel3<S> v)
{
    mom_x_body<S, sizeof(S) == 4>(hw_blk(xcd), g, geo, mode, mf, vfx, vfy, vfz, density1, v);
}

template <typename S>
static hipError_t fast_advec_mom_t(const Lay &g, const Geo &m, int mode, int direction, const S *mass_flux_x,
                                   const S *vol_flux_x, const S *mass_flux_y, const S *vol_flux_y,
                                   const S *mass_flux_z, const S *vol_flux_z, const S *density1,
                                   const void *const *vel_in, void *const *vel_out, const Shape &sh, hipStream_t st)
{
    Vel3<S> v;
    for (int c = 0; c < 3; c++) {
        v.in[c] = static_cast<const S *>(vel_in[c]);
        v.out[c] = static_cast<S *>(vel_out[c]);
    }
    const int nxn = g.x_max - g.x_min + 2, nyn = g.y_max - g.y_min + 2, nzn = g.z_max - g.z_min + 2;  // nodes
    const int MW = mw_of(direction);
    const dim3 blk(64, MW, 1);
    if (direction == 1) {
        const dim3 grid((nxn + 59) / 60, (nyn + MW - 2) / (MW - 1), nchunks(nzn, sh.planes));
        TIMED_LAUNCH("k_fast_mom_x", st, k_fast_mom_x<S>, grid, blk, 0, st, sh.xcd, g, m, mode, mass_flux_x, vol_flux_x,
                     vol_flux_y, vol_flux_z, density1, v);
    } else if (direction == 2) {
        const dim3 grid((nxn + 63) / 64, (nzn + MW - 2) / (MW - 1), nchunks(nyn, sh.planes));
        TIMED_LAUNCH("k_fast_mom_y", st, (k_fast_mom_march<2, S>), grid, blk, 0, st, sh.xcd, g, m, mode, mass_flux_y,
                     vol_flux_x, vol_flux_y, vol_flux_z, density1, v);
    } else {
        const dim3 grid((nxn + 63) / 64, (nyn + MW - 2) / (MW - 1), nchunks(nzn, sh.planes));
        TIMED_LAUNCH("k_fast_mom_z", st, (k_fast_mom_march<3, S>), grid, blk, 0, st, sh.xcd, g, m, mode, mass_flux_z,
                     vol_flux_x, vol_flux_y, vol_flux_z, density1, v);
    }
    return hipGetLastError();
}

hipError_t launch_fast_advec_mom(Store s, const Lay &g, const Geo &m, int advect_x, int sweep_number,
                                 int direction, const void *mass_flux_x, const void *vol_flux_x,
                                 const void *mass_flux_y, const void *vol_flux_y, const void *mass_flux_z,
                                 const void *vol_flux_z, const void *density1, const void *const *vel_in,
                                 void *const *vel_out, const Shape &sh, hipStream_t st)
{
    int mode = 0;
    if (sweep_number == 1 && direction == 1) mode = 1;
    else if (sweep_number == 1 && direction == 3) mode = 2;
    else if (sweep_number == 2 && advect_x) mode = 3;
    else if (sweep_number == 2 && !advect_x) mode = 4;
    else if (sweep_number == 3 && direction == 1) mode = 5;
    else if (sweep_number == 3 && direction == 3) mode = 6;
    if (!mode || direction < 1 || direction > 3) return hipErrorInvalidValue;
#define CALL(T)                                                                                                 \
    fast_advec_mom_t<T>(g, m, mode, direction, static_cast<const T *>(mass_flux_x),                             \
                        static_cast<const T *>(vol_flux_x), static_cast<const T *>(mass_flux_y),                \
                        static_cast<const T *>(vol_flux_y), static_cast<const T *>(mass_flux_z),                \
                        static_cast<const T *>(vol_flux_z), static_cast<const T *>(density1), vel_in, vel_out, sh, st)
    return s == ST_F32 ? CALL(float) : CALL(double);
#undef CALL
}

// ---------------------------------------------------------------------------
// Batched over many decks (batch.h): the new velocities go to X_W5..X_W7.
__host__ __device__ inline void dims_mom(int dir, const Lay &g, int planes, int &gx, int &gy, int &gz)
{
    const int nxn = g.x_max - g.x_min + 2, nyn = g.y_max - g.y_min + 2, nzn = g.z_max - g.z_min + 2;  // nodes
    const int MW = mw_of(dir);
    if (dir == 1) {
        gx = (nxn + 59) / 60; gy = (nyn + MW - 2) / (MW - 1); gz = nchunks(nzn, planes);
    } else if (dir == 2) {
        gx = (nxn + 63) / 64; gy = (nzn + MW - 2) / (MW - 1); gz = nchunks(nyn, planes);
    } else {
        gx = (nxn + 63) / 64; gy = (nyn + MW - 2) / (MW - 1); gz = nchunks(nzn, planes);
    }
}

int batch_blocks_mom(int dir, const Lay &g, int planes)
{
    int gx, gy, gz;
    dims_mom(dir, g, planes, gx, gy, gz);
    return gx * gy * gz;
}

template <int DIR, typename S>
__device__ __forceinline__ void mom_tile(const DeckDev *__restrict__ decks, const int *__restrict__ start, int nd,
                                         int mode, int planes, int hb)
{
    const int d = deck_of(start, nd, hb);
    const DeckDev &D = decks[d];
    int gx, gy, gz;
    dims_mom(DIR, D.g, planes, gx, gy, gz);
    const int lb = hb - start[d];
    const Blk bk{lb % gx, (lb / gx) % gy, lb / (gx * gy), gx, gy, gz};
    Vel3<S> v;
    v.in[0] = static_cast<const S *>(D.f[F_XVEL1]);
    v.in[1] = static_cast<const S *>(D.f[F_YVEL1]);
    v.in[2] = static_cast<const S *>(D.f[F_ZVEL1]);
    v.out[0] = static_cast<S *>(D.f[X_W5]);
    v.out[1] = static_cast<S *>(D.f[X_W6]);
    v.out[2] = static_cast<S *>(D.f[X_W7]);
    const S *vfx = static_cast<const S *>(D.f[F_VOL_FLUX_X]);
    const S *vfy = static_cast<const S *>(D.f[F_VOL_FLUX_Y]);
    const S *vfz = static_cast<const S *>(D.f[F_VOL_FLUX_Z]);
    const S *rho = static_cast<const S *>(D.f[F_DENSITY1]);
    if constexpr (DIR == 1)
        mom_x_body<S>(bk, D.g, D.m, mode, static_cast<const S *>(D.f[F_MASS_FLUX_X]), vfx, vfy, vfz, rho, v);
    else
        mom_march<DIR, S>(bk, D.g, D.m, mode,
                               static_cast<const S *>(D.f[DIR == 2 ? F_MASS_FLUX_Y : F_MASS_FLUX_Z]), vfx, vfy,
                               vfz, rho, v);
}

template <int DIR, typename S>
// The FP64 y/z sweeps held to five waves per SIMD (96 VGPRs; the tile walk
// takes them to 101, four waves): the batched momentum launches of 640^3 as
// 8 tiles 4.81 -> 4.35 ms with one value spilled.  With the tail form five
// values spill (20 bytes per lane) and the launches still run 4.41 -> 4.16 ms.
__global__ void __launch_bounds__(64 * mw_of(DIR), DIR != 1 && sizeof(S) == 8 ? 5 : 1)
kb_mom(const DeckDev *__restrict__ decks, const int *__restrict__ start, int nd, int mode, int planes, int xcd)
{
    int hb, end, step;  // the tiles of this workgroup (persistent when the grid is short)
    tile_walk(start[nd], xcd, hb, end, step);
    for (; hb < end; hb += step) {
        mom_tile<DIR, S>(decks, start, nd, mode, planes, hb);
        __syncthreads();  // the next tile reuses the workgroup's LDS
    }
}

hipError_t batch_advec_mom(Store s, const DeckDev *decks, int nd, const BatchGrid &bg, int advect_x, int dir,
                           int sweep_number, const Shape &sh, hipStream_t st)
{
    int mode = 0;
    if (sweep_number == 1 && dir == 1) mode = 1;
    else if (sweep_number == 1 && dir == 3) mode = 2;
    else if (sweep_number == 2 && advect_x) mode = 3;
    else if (sweep_number == 2 && !advect_x) mode = 4;
    else if (sweep_number == 3 && dir == 1) mode = 5;
    else if (sweep_number == 3 && dir == 3) mode = 6;
    if (!mode || dir < 1 || dir > 3) return hipErrorInvalidValue;
    if (bg.total == 0) return hipSuccess;
    const dim3 grid(batch_wgs(bg, sh)), blk(64, mw_of(dir));
#define L(D_, T_)                                                                                          \
    TIMED_LAUNCH("kb_mom", st, (kb_mom<D_, T_>), grid, blk, 0, st, decks, bg.start, nd, mode, sh.planes, sh.xcd)
    if (s == ST_F32) {
        if (dir == 1) L(1, float); else if (dir == 2) L(2, float); else L(3, float);
    } else {
        if (dir == 1) L(1, double); else if (dir == 2) L(2, double); else L(3, double);
    }
#undef L
    return hipGetLastError();
}

}  // namespace clover
