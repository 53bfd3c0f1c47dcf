// advec_cell (advec_cell_kernel.f90:30-383) for one sweep as one kernel:
// the mass and energy fluxes of every face along the sweep and the cell
// update that consumes them, with the energy flux and pre_vol kept in
// registers.  The fluxes of a face depend only on cells along the sweep
// (upwind, donor, downwind), so
//   * the x sweep works on rows: lanes along x, each lane forms the flux of
//     its face from its neighbours' values (lane shuffles) and updates its
//     cell from its own and the next lane's flux (60 of 64 lanes store);
//   * the y and z sweeps march each (x, cross) column along the sweep
//     axis with the upwind..downwind window in registers.
// The reference computes every flux with the old density1/energy1 before
// it updates any cell, so the new values go to separate arrays (rho_out,
// e_out) which then take the place of density1/energy1; only interior cells
// of those are written (the halo update that follows every advec_cell
// sweep in advection.f90 refreshes their halos before anything reads them).
// The step's last sweep can also write the pressure of the new state (p_out;
// null otherwise): after reset_field that state is the
// next step's level 0, so these are the values the next ideal_gas call
// (timestep.f90:48 or field_summary.f90:49-60) would compute.
// In the FP64 form every value is formed with the expression and operand
// order of the per-call kernels, so the results are bit-identical.  The FP32
// form computes the face fluxes in FP32 and keeps pre_vol and the cell's
// mass and energy update in FP64.
#include <hip/hip_runtime.h>

#include "batch.h"
#include "fast.h"
#include "launch.h"

namespace clover {

constexpr int CROWS = 4;   // rows (or columns) per workgroup

// pre_vol (advec_cell_kernel.f90:97-116, 193-210, 289-306) for `mode`:
// 1 x sweep 1, 2 x sweep 3, 3 y sweep with advect_x, 4 y without,
// 5 z sweep 1, 6 z sweep 3
template <typename S>
__device__ inline double cell_pre_vol(int mode, const Lay &g, int i, double V0, const S *__restrict__ vfx,
                                      const S *__restrict__ vfy, const S *__restrict__ vfz)
{
#define D(x) ((double)(x))
    switch (mode) {
    case 1:
    case 5: return V0 + (D(vfx[i + 1]) - D(vfx[i]) + D(vfy[i + g.ld]) - D(vfy[i]) + D(vfz[i + g.plane]) - D(vfz[i]));
    case 2: return V0 + D(vfx[i + 1]) - D(vfx[i]);
    case 3: return V0 + D(vfy[i + g.ld]) - D(vfy[i]) + D(vfz[i + g.plane]) - D(vfz[i]);
    case 4: return V0 + D(vfy[i + g.ld]) - D(vfy[i]) + D(vfx[i + 1]) - D(vfx[i]);
    default: return V0 + D(vfz[i + g.plane]) - D(vfz[i]);
    }
#undef D
}

template <typename T>
__device__ inline T cell_limiter(T sigma, T sigma3, T sigma4, T diffuw, T diffdw)
{
    const T one_by_six = (T)(1.0 / 6.0);
    if (diffuw * diffdw > (T)0)
        return ((T)1 - sigma) * copysign((T)1, diffdw) *
               tmin(tmin(fabs(diffuw), fabs(diffdw)), one_by_six * (sigma3 * fabs(diffuw) + sigma4 * fabs(diffdw)));
    return (T)0;
}

// Mass and energy flux through one face (advec_cell_kernel.f90:123-186 and
// twins): ru/rd/rw, eu/ed/ew the density and energy of the upwind, donor and
// downwind cells, pv the donor's pre_vol, vd the vertex spacing; T is the
// form's arithmetic type.
template <typename T>
__device__ inline void cell_face_flux(T vf, T pv, T vd, T ru, T rd, T rw, T eu, T ed, T ew, T &mf, T &ef)
{
    const T sigmat = quo(fabs(vf), pv);
    const T sigma3 = ((T)1 + sigmat) * (vd / vd);  // vertexd*(j)/vertexd*(dif) on the uniform mesh
    const T sigma4 = (T)2 - sigmat;
    T limiter = cell_limiter(sigmat, sigma3, sigma4, rd - ru, rw - rd);
    mf = vf * (rd + limiter);
    const T sigmam = quo(fabs(mf), rd * pv);
    limiter = cell_limiter(sigmam, sigma3, sigma4, ed - eu, ew - ed);
    ef = mf * (ed + limiter);
}

// cell update (advec_cell_kernel.f90:188-191 and twins), FP64 in both forms
__device__ inline void cell_update(double rho, double e, double pv, double mf_lo, double mf_hi, double ef_lo,
                                   double ef_hi, double vf_lo, double vf_hi, double &rho_new, double &e_new)
{
    const double pm = rho * pv;
    const double pom = pm + mf_lo - mf_hi;
    const double pe = (e * pm + ef_lo - ef_hi) / pom;
    const double av = pv + vf_lo - vf_hi;
    rho_new = pom / av;
    e_new = pe;
}

__device__ inline double shfl_c(double x, int src) { return __shfl(x, src, 64); }
__device__ inline float shfl_c(float x, int src) { return __shfl(x, src, 64); }

// ---------------------------------------------------------------------------
// x sweep: faces j in [x_min, x_max+2], rows k, l interior.  Lane i holds
// cell/face j = jt - 2 + i.  A wave walks `rows` consecutive rows (k
// fastest) and issues the next row's loads before the current row's work;
// one row per wave leaves too few bytes in flight in the FP32 form.
template <typename S, bool TAN = false>  // TAN: vol_flux_y/z at tan_xm (mom_kernels.hip mom_march_body)
__device__ __forceinline__ void cell_x_body(const Blk &bk, Lay g, Geo m, int mode, int rows, const S *__restrict__ vfx,
              const S *__restrict__ vfy, const S *__restrict__ vfz, const S *__restrict__ density1,
              const S *__restrict__ energy1, S *__restrict__ mass_flux_x, S *__restrict__ rho_out,
              S *__restrict__ e_out, S *__restrict__ p_out)
{
    const int lane = threadIdx.x;
    const int ny = g.y_max - g.y_min + 1, nz = g.z_max - g.z_min + 1;
    const int nrows = ny * nz;
    const int r0 = (bk.y * CROWS + (int)threadIdx.y) * rows;
    if (r0 >= nrows) return;  // whole wave: no shuffle partner is lost
    const int r1 = min(r0 + rows, nrows) - 1;
    const int j = g.x_min - 2 + bk.x * 60 + lane;
    using T = S;  // arithmetic type of the fluxes
    const bool ok_cell = j <= g.x_max + 2;  // j >= x_min-2 always
    const bool ok_face = j <= g.x_max + 3;
    // pre_vol's volume fluxes (advec_cell_kernel.f90:97-116): all three axes
    // in sweep 1 (mode 1), x only in sweep 3 (mode 2)
    const bool all3 = mode == 1;
    // the next row's values, raw (converted where used, so the loads stay
    // in flight): density, energy, vol_flux_x at j and j+1, vol_flux_y at k
    // and k+1, vol_flux_z at l and l+1
    S rn = 0, en = 0, fx0 = 0, fx1 = 0, fy0 = 0, fy1 = 0, fz0 = 0, fz1 = 0;
    auto row_at = [&](int r) { return idx(g, j, g.y_min + r % ny, g.z_min + r / ny); };
    const int dxc = cell_xm(g, j) - j;  // a mirrored x face: the interior cell its halo copies
    const int dxt = TAN ? tan_xm(g, j) - j : 0;  // the same for vol_flux_y and vol_flux_z
    auto load = [&](int ii) {
        rn = ok_cell ? density1[ii + dxc] : (S)0;
        en = ok_cell ? energy1[ii + dxc] : (S)0;
        fx0 = ok_face ? vfx[ii] : (S)0;
        fx1 = ok_cell ? vfx[ii + 1] : (S)0;
        fy0 = (ok_cell && all3) ? vfy[ii + dxt] : (S)0;
        fy1 = (ok_cell && all3) ? vfy[ii + dxt + g.ld] : (S)0;
        fz0 = (ok_cell && all3) ? vfz[ii + dxt] : (S)0;
        fz1 = (ok_cell && all3) ? vfz[ii + dxt + g.plane] : (S)0;
    };
    int in = row_at(r0);
    load(in);
    for (int r = r0; r <= r1; r++) {
        const int i = in;
        const T rho = rn, e = en;
        const S x0 = fx0, x1 = fx1, y0 = fy0, y1 = fy1, z0 = fz0, z1 = fz1;
        if (r < r1) {
            in = row_at(r + 1);
            load(in);
        }
        double pv = 0.0;
        if (ok_cell)
            pv = all3 ? m.vol + ((double)x1 - (double)x0 + (double)y1 - (double)y0 + (double)z1 - (double)z0)
                      : m.vol + (double)x1 - (double)x0;
        const T vf = x0;
        const int lm2 = max(lane - 2, 0), lm1 = max(lane - 1, 0), lp1 = min(lane + 1, 63);
        const T rm2 = shfl_c(rho, lm2), rm1 = shfl_c(rho, lm1), rp1 = shfl_c(rho, lp1);
        const T em2 = shfl_c(e, lm2), em1 = shfl_c(e, lm1), ep1 = shfl_c(e, lp1);
        const double pvm1 = shfl_c(pv, lm1);
        const T vd = (T)m.dx;
        // face j: upwind/donor/downwind (advec_cell_kernel.f90:123-132); the
        // upwind index is clamped at x_max+2
        T mf, ef;
        if (vf > (T)0) {
            cell_face_flux(vf, (T)pvm1, vd, rm2, rm1, rho, em2, em1, e, mf, ef);
        } else {
            const bool clamp = j + 1 > g.x_max + 2;
            cell_face_flux(vf, (T)pv, vd, clamp ? rho : rp1, rho, rm1, clamp ? e : ep1, e, em1, mf, ef);
        }
        const T mf_hi = shfl_c(mf, lp1), ef_hi = shfl_c(ef, lp1), vf_hi = shfl_c(vf, lp1);
        if (lane < 2 || lane > 61) continue;
        if (j <= g.x_max + 2) mass_flux_x[i] = mf;
        if (j <= g.x_max) {
            double rnew, enew;
            cell_update(rho, e, pv, mf, mf_hi, ef, ef_hi, vf, vf_hi, rnew, enew);
            const S rs = (S)rnew, es = (S)enew;
            rho_out[i] = rs;
            e_out[i] = es;
            if (p_out) p_out[i] = pressure_of<S>(rs, es);
        }
    }
}

template <typename S>
__global__ void __launch_bounds__(64 * CROWS) k_fast_cell_x(int xcd, Lay g, Geo m, int mode, const S *__restrict__ vfx, const S *__restrict__ vfy,
              const S *__restrict__ vfz, const S *__restrict__ density1,
              const S *__restrict__ energy1, S *__restrict__ mass_flux_x, S *__restrict__ rho_out,
              S *__restrict__ e_out, S *__restrict__ p_out, int rows)
{
    cell_x_body<S, sizeof(S) == 4>(hw_blk(xcd), g, m, mode, rows, vfx, vfy, vfz, density1, energy1, mass_flux_x, rho_out, e_out,
                   p_out);
}

// The volume fluxes pre_vol reads for the y and z sweeps (modes 3-6), kept
// raw so a face's loads can be issued a step ahead of their use.
template <typename S>
struct PreRaw {
    S a, b, c, d, e, f;  // storage type: converted where used, so the loads stay in flight
};

template <typename S>
__device__ inline PreRaw<S> load_pre_raw(int mode, const Lay &g, int i, const S *__restrict__ vfx,
                                         const S *__restrict__ vfy, const S *__restrict__ vfz)
{
    PreRaw<S> r{0, 0, 0, 0, 0, 0};
    switch (mode) {
    case 3: r.a = vfy[i + g.ld]; r.b = vfy[i]; r.c = vfz[i + g.plane]; r.d = vfz[i]; break;
    case 4: r.a = vfy[i + g.ld]; r.b = vfy[i]; r.c = vfx[i + 1]; r.d = vfx[i]; break;
    case 5:
        r.a = vfx[i + 1]; r.b = vfx[i]; r.c = vfy[i + g.ld]; r.d = vfy[i]; r.e = vfz[i + g.plane]; r.f = vfz[i];
        break;
    default: r.a = vfz[i + g.plane]; r.b = vfz[i]; break;
    }
    return r;
}

// cell_pre_vol's expressions for modes 3-6 on the raw values
template <typename S>
__device__ inline double pre_of_raw(int mode, double V0, const PreRaw<S> &r)
{
    const double a = r.a, b = r.b, c = r.c, d = r.d, e = r.e, f = r.f;
    switch (mode) {
    case 3:
    case 4: return V0 + a - b + c - d;
    case 5: return V0 + (a - b + c - d + e - f);
    default: return V0 + a - b;
    }
}

// ---------------------------------------------------------------------------
// y and z sweeps: lanes along x (interior j), threadIdx.y along the cross
// axis, march along the sweep axis over faces [smin, smax+2] in chunks.
// Faces of the y sweep run over l in [z_min, z_max+2] (advec_cell_kernel.f90:
// 215-217); the cells updated are interior.
template <int DIR, typename S>
__device__ __forceinline__ void cell_march_body(const Blk &bk, Lay g, Geo m, int mode, const S *__restrict__ vfx, const S *__restrict__ vfy,
                  const S *__restrict__ vfz, const S *__restrict__ density1,
                  const S *__restrict__ energy1, S *__restrict__ mass_flux, S *__restrict__ rho_out,
                  S *__restrict__ e_out, S *__restrict__ p_out)
{
    const int j = g.x_min + (int)(bk.x * 64 + threadIdx.x);
    const int cr = (DIR == 3 ? g.y_min : g.z_min) + (int)(bk.y * CROWS + threadIdx.y);
    const int crmax = DIR == 3 ? g.y_max : g.z_max + 2;  // last cross position with faces
    if (j > g.x_max || cr > crmax) return;                 // no cross-lane traffic in this kernel
    const bool cross_in = DIR == 3 || cr <= g.z_max;       // updates only on interior cells
    const int s = DIR == 3 ? g.plane : g.ld;
    const int smin = DIR == 3 ? g.z_min : g.y_min, smax = DIR == 3 ? g.z_max : g.y_max;
    using T = S;  // arithmetic type of the fluxes
    const T vd = (T)(DIR == 3 ? m.dz : m.dy);
    const S *__restrict__ vfs = DIR == 3 ? vfz : vfy;
    int f0, f1;
    chunk_range(bk, smin, smax + 2, f0, f1);
    if (f0 > f1) return;
    const int ib = DIR == 3 ? idx(g, j, cr, smin) : idx(g, j, smin, cr);  // offset at sweep position smin
#define OFF(q) (ib + ((q) - smin) * s)
    auto pre = [&](int q) { return cell_pre_vol(mode, g, OFF(q), m.vol, vfx, vfy, vfz); };

    // window of old values at f-2, f-1, f, f+1 and pre_vol at f-1, f
    T r_m2 = density1[OFF(f0 - 2)], r_m1 = density1[OFF(f0 - 1)], r_0 = density1[OFF(f0)];
    T e_m2 = energy1[OFF(f0 - 2)], e_m1 = energy1[OFF(f0 - 1)], e_0 = energy1[OFF(f0)];
    double pv_m1 = pre(f0 - 1);
    // flux of face f0-1, needed when cell f0-1 is updated here (f0 > smin)
    T mf_lo = 0, ef_lo = 0, vf_lo = 0;
    if (f0 > smin) {
        vf_lo = vfs[OFF(f0 - 1)];
        const T r_m3 = density1[OFF(f0 - 3)], e_m3 = energy1[OFF(f0 - 3)];
        if (vf_lo > (T)0)
            cell_face_flux(vf_lo, (T)pre(f0 - 2), vd, r_m3, r_m2, r_m1, e_m3, e_m2, e_m1, mf_lo, ef_lo);
        else
            cell_face_flux(vf_lo, (T)pv_m1, vd, r_0, r_m1, r_m2, e_0, e_m1, e_m2, mf_lo, ef_lo);  // f0 <= smax+1
    }
    // face f0's loads; each step issues the next face's before its arithmetic
    S vf_n = vfs[OFF(f0)];
    PreRaw<S> pr_n = load_pre_raw(mode, g, OFF(f0), vfx, vfy, vfz);
    S r_n = (S)0, e_n = (S)0;  // upwind for vf <= 0 is min(f+1, smax+2)
    if (f0 + 1 <= smax + 2) {
        r_n = density1[OFF(f0 + 1)];
        e_n = energy1[OFF(f0 + 1)];
    }
    for (int f = f0; f <= f1; f++) {
        const int i = OFF(f);
        const T vf = vf_n;
        const double pv_0 = pre_of_raw(mode, m.vol, pr_n);
        const T r_p1 = f + 1 <= smax + 2 ? r_n : r_0, e_p1 = f + 1 <= smax + 2 ? e_n : e_0;
        if (f < f1) {
            vf_n = vfs[OFF(f + 1)];
            pr_n = load_pre_raw(mode, g, OFF(f + 1), vfx, vfy, vfz);
            if (f + 2 <= smax + 2) {
                r_n = density1[OFF(f + 2)];
                e_n = energy1[OFF(f + 2)];
            }
        }
        T mf, ef;
        if (vf > (T)0)
            cell_face_flux(vf, (T)pv_m1, vd, r_m2, r_m1, r_0, e_m2, e_m1, e_0, mf, ef);
        else
            cell_face_flux(vf, (T)pv_0, vd, r_p1, r_0, r_m1, e_p1, e_0, e_m1, mf, ef);
        mass_flux[i] = mf;
        const int c = f - 1;  // cell between faces f-1 and f
        if (cross_in && c >= smin && c <= smax) {
            double rn, en;
            cell_update(r_m1, e_m1, pv_m1, mf_lo, mf, ef_lo, ef, vf_lo, vf, rn, en);
            const S rs = (S)rn, es = (S)en;
            rho_out[OFF(c)] = rs;
            e_out[OFF(c)] = es;
            if (p_out) p_out[OFF(c)] = pressure_of<S>(rs, es);
        }
        mf_lo = mf;
        ef_lo = ef;
        vf_lo = vf;
        r_m2 = r_m1;
        r_m1 = r_0;
        r_0 = r_p1;
        e_m2 = e_m1;
        e_m1 = e_0;
        e_0 = e_p1;
        pv_m1 = pv_0;
    }
#undef OFF
}

template <int DIR, typename S>
__global__ void __launch_bounds__(64 * CROWS) k_fast_cell_march(int xcd, Lay g, Geo m, int mode, const S *__restrict__ vfx, const S *__restrict__ vfy,
                  const S *__restrict__ vfz, const S *__restrict__ density1,
                  const S *__restrict__ energy1, S *__restrict__ mass_flux, S *__restrict__ rho_out,
                  S *__restrict__ e_out, S *__restrict__ p_out)
{
    cell_march_body<DIR, S>(hw_blk(xcd), g, m, mode, vfx, vfy, vfz, density1, energy1, mass_flux, rho_out, e_out,
                            p_out);
}

template <typename S>
static hipError_t fast_advec_cell_t(const Lay &g, const Geo &m, int mode, int dir, const S *density1,
                                    const S *energy1, S *mass_flux_x, const S *vol_flux_x, S *mass_flux_y,
                                    const S *vol_flux_y, S *mass_flux_z, const S *vol_flux_z, S *rho_out, S *e_out,
                                    S *p_out, const Shape &sh, hipStream_t st)
{
    const int nx = g.x_max - g.x_min + 1, ny = g.y_max - g.y_min + 1, nz = g.z_max - g.z_min + 1;
    const dim3 blk(64, CROWS, 1);
    if (dir == 1) {
        const int rows = cell_x_rows(sh.planes);
        const dim3 grid((nx + 2 + 59) / 60, ((ny * nz + rows - 1) / rows + CROWS - 1) / CROWS, 1);
        TIMED_LAUNCH("k_fast_cell_x", st, k_fast_cell_x<S>, grid, blk, 0, st, sh.xcd, g, m, mode, vol_flux_x, vol_flux_y,
                     vol_flux_z, density1, energy1, mass_flux_x, rho_out, e_out, p_out, rows);
    } else if (dir == 2) {
        const dim3 grid((nx + 63) / 64, (nz + 2 + CROWS - 1) / CROWS, nchunks(ny + 2, sh.planes));
        TIMED_LAUNCH("k_fast_cell_y", st, (k_fast_cell_march<2, S>), grid, blk, 0, st, sh.xcd, g, m, mode, vol_flux_x,
                     vol_flux_y, vol_flux_z, density1, energy1, mass_flux_y, rho_out, e_out, p_out);
    } else {
        const dim3 grid((nx + 63) / 64, (ny + CROWS - 1) / CROWS, nchunks(nz + 2, sh.planes));
        TIMED_LAUNCH("k_fast_cell_z", st, (k_fast_cell_march<3, S>), grid, blk, 0, st, sh.xcd, g, m, mode, vol_flux_x,
                     vol_flux_y, vol_flux_z, density1, energy1, mass_flux_z, rho_out, e_out, p_out);
    }
    return hipGetLastError();
}

hipError_t launch_fast_advec_cell(Store s, const Lay &g, const Geo &m, int advect_x, int dir, int sweep_number,
                                  const void *density1, const void *energy1, void *mass_flux_x,
                                  const void *vol_flux_x, void *mass_flux_y, const void *vol_flux_y,
                                  void *mass_flux_z, const void *vol_flux_z, void *rho_out, void *e_out,
                                  void *pressure, const Shape &sh, hipStream_t st)
{
    int mode = 0;
    if (dir == 1) mode = sweep_number == 1 ? 1 : sweep_number == 3 ? 2 : 0;
    else if (dir == 2) mode = sweep_number == 2 ? (advect_x ? 3 : 4) : 0;
    else if (dir == 3) mode = sweep_number == 1 ? 5 : sweep_number == 3 ? 6 : 0;
    if (!mode) return hipErrorInvalidValue;  // the resident driver issues only the reference's sweeps
    if (pressure && sweep_number != 3) return hipErrorInvalidValue;  // only the last sweep's state is final
#define CALL(T)                                                                                                 \
    fast_advec_cell_t<T>(g, m, mode, dir, static_cast<const T *>(density1), static_cast<const T *>(energy1),     \
                         static_cast<T *>(mass_flux_x), static_cast<const T *>(vol_flux_x),                      \
                         static_cast<T *>(mass_flux_y), static_cast<const T *>(vol_flux_y),                      \
                         static_cast<T *>(mass_flux_z), static_cast<const T *>(vol_flux_z),                      \
                         static_cast<T *>(rho_out), static_cast<T *>(e_out), static_cast<T *>(pressure), sh, st)
    return s == ST_F32 ? CALL(float) : CALL(double);
#undef CALL
}

// ---------------------------------------------------------------------------
// Batched over many decks (batch.h): new density1/energy1 go to X_W1/X_W2.
__host__ __device__ inline void dims_cell(int dir, const Lay &g, int planes, int &gx, int &gy, int &gz)
{
    const int nx = g.x_max - g.x_min + 1, ny = g.y_max - g.y_min + 1, nz = g.z_max - g.z_min + 1;
    if (dir == 1) {
        const int rows = cell_x_rows(planes);
        gx = (nx + 2 + 59) / 60; gy = ((ny * nz + rows - 1) / rows + CROWS - 1) / CROWS; gz = 1;
    } else if (dir == 2) {
        gx = (nx + 63) / 64; gy = (nz + 2 + CROWS - 1) / CROWS; gz = nchunks(ny + 2, planes);
    } else {
        gx = (nx + 63) / 64; gy = (ny + CROWS - 1) / CROWS; gz = nchunks(nz + 2, planes);
    }
}

int batch_blocks_cell(int dir, const Lay &g, int planes)
{
    int gx, gy, gz;
    dims_cell(dir, g, planes, gx, gy, gz);
    return gx * gy * gz;
}

template <int DIR, typename S>
__global__ void __launch_bounds__(64 * CROWS)
kb_cell(const DeckDev *__restrict__ decks, const int *__restrict__ start, int nd, int mode, int gas, int planes,
        int xcd)
{
    int hb, end, step;  // the tiles of this workgroup (persistent when the grid is short)
    tile_walk(start[nd], xcd, hb, end, step);
    for (; hb < end; hb += step) {
        const int d = deck_of(start, nd, hb);
        const DeckDev &D = decks[d];
        int gx, gy, gz;
        dims_cell(DIR, D.g, planes, gx, gy, gz);
        const int lb = hb - start[d];
        const Blk bk{lb % gx, (lb / gx) % gy, lb / (gx * gy), gx, gy, gz};
#define CF(id) static_cast<const S *>(D.f[id])
#define MF(id) static_cast<S *>(D.f[id])
        if constexpr (DIR == 1)
            cell_x_body<S>(bk, D.g, D.m, mode, cell_x_rows(planes), CF(F_VOL_FLUX_X), CF(F_VOL_FLUX_Y), CF(F_VOL_FLUX_Z), CF(F_DENSITY1),
                           CF(F_ENERGY1), MF(F_MASS_FLUX_X), MF(X_W1), MF(X_W2), gas ? MF(F_PRESSURE) : nullptr);
        else
            cell_march_body<DIR, S>(bk, D.g, D.m, mode, CF(F_VOL_FLUX_X), CF(F_VOL_FLUX_Y), CF(F_VOL_FLUX_Z),
                                    CF(F_DENSITY1), CF(F_ENERGY1), MF(DIR == 2 ? F_MASS_FLUX_Y : F_MASS_FLUX_Z),
                                    MF(X_W1), MF(X_W2), gas ? MF(F_PRESSURE) : nullptr);
#undef CF
#undef MF
        __syncthreads();  // the next tile reuses the workgroup's LDS
    }
}

hipError_t batch_advec_cell(Store s, const DeckDev *decks, int nd, const BatchGrid &bg, int advect_x, int dir,
                            int sweep_number, bool gas, const Shape &sh, hipStream_t st)
{
    int mode = 0;
    if (dir == 1) mode = sweep_number == 1 ? 1 : sweep_number == 3 ? 2 : 0;
    else if (dir == 2) mode = sweep_number == 2 ? (advect_x ? 3 : 4) : 0;
    else if (dir == 3) mode = sweep_number == 1 ? 5 : sweep_number == 3 ? 6 : 0;
    if (!mode) return hipErrorInvalidValue;
    if (bg.total == 0) return hipSuccess;
    const dim3 grid(batch_wgs(bg, sh)), blk(64, CROWS);
    if (gas && sweep_number != 3) return hipErrorInvalidValue;
#define L(D_, T_)                                                                                          \
    TIMED_LAUNCH("kb_cell", st, (kb_cell<D_, T_>), grid, blk, 0, st, decks, bg.start, nd, mode, gas ? 1 : 0,   \
                 sh.planes, sh.xcd)
    if (s == ST_F32) {
        if (dir == 1) L(1, float); else if (dir == 2) L(2, float); else L(3, float);
    } else {
        if (dir == 1) L(1, double); else if (dir == 2) L(2, double); else L(3, double);
    }
#undef L
    return hipGetLastError();
}

}  // namespace clover
