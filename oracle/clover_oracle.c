/*
 * CPU restatement of the CloverLeaf3D hydro kernels (the oracle).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product path, never the
 * thing measured or shipped.  See clover_oracle.h for the conventions.
 *
 * Each function follows the reference subroutine cited above it: same loop
 * domains, same branch structure, same left-to-right evaluation order of
 * every expression (Fortran evaluates a+b+c as (a+b)+c and a*b*c as (a*b)*c,
 * and so does C), so results agree bit for bit with the reference compiled
 * without FMA contraction.  Default-real (single precision) literals of the
 * reference are reproduced as float constants widened to double.
 */
#include "clover_oracle.h"

#include <math.h>
#include <stddef.h>

/* Fortran column-major offset of (j,k,l) in an array whose lower bounds are
 * (x_min-2, y_min-2, z_min-2) and whose first two extents are nj, nk. */
typedef struct {
    long jlo, klo, llo, sk, sl;
} shape_t;

static shape_t mk_shape(int x_min, int x_max, int y_min, int y_max, int z_min, int ex, int ey)
{
    shape_t s;
    long nj = (long)(x_max - x_min + 5 + ex);
    long nk = (long)(y_max - y_min + 5 + ey);
    s.jlo = x_min - 2;
    s.klo = y_min - 2;
    s.llo = z_min - 2;
    s.sk = nj;
    s.sl = nj * nk;
    return s;
}

#define OFF(s, j, k, l) (((long)(j) - (s).jlo) + (s).sk * ((long)(k) - (s).klo) + (s).sl * ((long)(l) - (s).llo))

/* Per-function shapes: cell (C), vertex (V), x/y/z faces (FX, FY, FZ). */
#define SHAPES                                                        \
    const shape_t SC = mk_shape(x_min, x_max, y_min, y_max, z_min, 0, 0); \
    const shape_t SV = mk_shape(x_min, x_max, y_min, y_max, z_min, 1, 1); \
    const shape_t SX = mk_shape(x_min, x_max, y_min, y_max, z_min, 1, 0); \
    const shape_t SY = mk_shape(x_min, x_max, y_min, y_max, z_min, 0, 1); \
    (void)SC; (void)SV; (void)SX; (void)SY

#define C_(a, j, k, l) (a)[OFF(SC, j, k, l)]
#define V_(a, j, k, l) (a)[OFF(SV, j, k, l)]
#define X_(a, j, k, l) (a)[OFF(SX, j, k, l)]
#define Y_(a, j, k, l) (a)[OFF(SY, j, k, l)]
#define Z_(a, j, k, l) (a)[OFF(SC, j, k, l)] /* z-face strides equal the cell strides */
#define D1X(a, j) (a)[(long)(j) - (x_min - 2)]
#define D1Y(a, k) (a)[(long)(k) - (y_min - 2)]
#define D1Z(a, l) (a)[(long)(l) - (z_min - 2)]

/* Fortran intrinsics on REAL(8). */
static inline double fmin2(double a, double b) { return a < b ? a : b; }
static inline double fmax2(double a, double b) { return a > b ? a : b; }
static inline double fsign(double a, double b) { return copysign(fabs(a), b); }

/* ------------------------------------------------------------------------ */
/* ideal_gas_kernel.f90:47-58 */
void oracle_ideal_gas_kernel(OR_DIMS, const double *density, const double *energy,
                             double *pressure, double *soundspeed)
{
    SHAPES;
#pragma omp parallel for
    for (int l = z_min; l <= z_max; l++)
        for (int k = y_min; k <= y_max; k++)
            for (int j = x_min; j <= x_max; j++) {
                double v = 1.0 / C_(density, j, k, l);
                C_(pressure, j, k, l) = (1.4 - 1.0) * C_(density, j, k, l) * C_(energy, j, k, l);
                double pressurebyenergy = (1.4 - 1.0) * C_(density, j, k, l);
                double pressurebyvolume = -C_(density, j, k, l) * C_(pressure, j, k, l);
                double sound_speed_squared =
                    v * v * (C_(pressure, j, k, l) * pressurebyenergy - pressurebyvolume);
                C_(soundspeed, j, k, l) = sqrt(sound_speed_squared);
            }
}

/* ------------------------------------------------------------------------ */
/* viscosity_kernel.f90:71-134 */
void oracle_viscosity_kernel(OR_DIMS, const double *xarea, const double *yarea, const double *zarea,
                             const double *celldx, const double *celldy, const double *celldz,
                             const double *density0, const double *pressure, double *viscosity,
                             const double *xvel0, const double *yvel0, const double *zvel0)
{
    SHAPES;
#pragma omp parallel for
    for (int l = z_min; l <= z_max; l++)
        for (int k = y_min; k <= y_max; k++)
            for (int j = x_min; j <= x_max; j++) {
#define U(a, b, c) V_(xvel0, a, b, c)
#define W(a, b, c) V_(zvel0, a, b, c)
#define VV(a, b, c) V_(yvel0, a, b, c)
                double ugradx1 = U(j, k, l) + U(j, k + 1, l) + U(j, k, l + 1) + U(j, k + 1, l + 1);
                double ugradx2 = U(j + 1, k, l) + U(j + 1, k + 1, l) + U(j + 1, k, l + 1) + U(j + 1, k + 1, l + 1);
                double ugrady1 = U(j, k, l) + U(j + 1, k, l) + U(j, k, l + 1) + U(j + 1, k, l + 1);
                double ugrady2 = U(j, k + 1, l) + U(j + 1, k + 1, l) + U(j, k + 1, l + 1) + U(j + 1, k + 1, l + 1);
                double ugradz1 = U(j, k, l) + U(j + 1, k, l) + U(j, k + 1, l) + U(j + 1, k + 1, l);
                double ugradz2 = U(j, k, l + 1) + U(j + 1, k, l + 1) + U(j, k + 1, l + 1) + U(j + 1, k + 1, l + 1);

                double vgradx1 = VV(j, k, l) + VV(j, k + 1, l) + VV(j, k, l + 1) + VV(j, k + 1, l + 1);
                double vgradx2 = VV(j + 1, k, l) + VV(j + 1, k + 1, l) + VV(j + 1, k, l + 1) + VV(j + 1, k + 1, l + 1);
                double vgrady1 = VV(j, k, l) + VV(j + 1, k, l) + VV(j, k, l + 1) + VV(j + 1, k, l + 1);
                double vgrady2 = VV(j, k + 1, l) + VV(j + 1, k + 1, l) + VV(j, k + 1, l + 1) + VV(j + 1, k + 1, l + 1);
                double vgradz1 = VV(j, k, l) + VV(j + 1, k, l) + VV(j, k + 1, l) + VV(j + 1, k + 1, l);
                double vgradz2 = VV(j, k, l + 1) + VV(j + 1, k, l + 1) + VV(j, k + 1, l + 1) + VV(j + 1, k + 1, l + 1);

                double wgradx1 = W(j, k, l) + W(j, k + 1, l) + W(j, k, l + 1) + W(j, k + 1, l + 1);
                double wgradx2 = W(j + 1, k, l) + W(j + 1, k + 1, l) + W(j + 1, k, l + 1) + W(j + 1, k + 1, l + 1);
                double wgrady1 = W(j, k, l) + W(j + 1, k, l) + W(j, k, l + 1) + W(j + 1, k, l + 1);
                double wgrady2 = W(j, k + 1, l) + W(j + 1, k + 1, l) + W(j, k + 1, l + 1) + W(j + 1, k + 1, l + 1);
                double wgradz1 = W(j, k, l) + W(j + 1, k, l) + W(j, k + 1, l) + W(j + 1, k + 1, l);
                double wgradz2 = W(j, k, l + 1) + W(j + 1, k, l + 1) + W(j, k + 1, l + 1) + W(j + 1, k + 1, l + 1);
#undef U
#undef W
#undef VV
                double div = (X_(xarea, j, k, l) * (ugradx2 - ugradx1) + Y_(yarea, j, k, l) * (vgrady2 - vgrady1)) +
                             Z_(zarea, j, k, l) * (wgradz2 - wgradz1);
                (void)div;

                double xx = 0.25 * (ugradx2 - ugradx1) / (D1X(celldx, j));
                double yy = 0.25 * (vgrady2 - vgrady1) / (D1Y(celldy, k));
                double zz = 0.25 * (wgradz2 - wgradz1) / (D1Z(celldz, l));
                double xy = 0.25 * (ugrady2 - ugrady1) / (D1Y(celldy, k)) + 0.25 * (vgradx2 - vgradx1) / (D1X(celldx, j));
                double xz = 0.25 * (ugradz2 - ugradz1) / (D1Z(celldz, l)) + 0.25 * (wgradx2 - wgradx1) / (D1X(celldx, j));
                double yz = 0.25 * (vgradz2 - vgradz1) / (D1Z(celldz, l)) + 0.25 * (wgrady2 - wgrady1) / (D1Y(celldy, k));

                double pgradx = (C_(pressure, j + 1, k, l) - C_(pressure, j - 1, k, l)) / (D1X(celldx, j) + D1X(celldx, j + 1));
                double pgrady = (C_(pressure, j, k + 1, l) - C_(pressure, j, k - 1, l)) / (D1Y(celldy, k) + D1Y(celldy, k + 1));
                double pgradz = (C_(pressure, j, k, l + 1) - C_(pressure, j, k, l - 1)) / (D1Z(celldz, l) + D1Z(celldz, l + 1));

                double pgradx2 = pgradx * pgradx;
                double pgrady2 = pgrady * pgrady;
                double pgradz2 = pgradz * pgradz;

                double limiter = (xx * pgradx2 + yy * pgrady2 + zz * pgradz2 + xy * pgradx * pgrady +
                                  xz * pgradx * pgradz + yz * pgrady * pgradz) /
                                 fmax2(pgradx2 + pgrady2 + pgradz2, 1.0e-16);

                if ((limiter > 0.0) || (div >= 0.0)) {
                    C_(viscosity, j, k, l) = 0.0;
                } else {
                    pgradx = fsign(fmax2(1.0e-16, fabs(pgradx)), pgradx);
                    pgrady = fsign(fmax2(1.0e-16, fabs(pgrady)), pgrady);
                    pgradz = fsign(fmax2(1.0e-16, fabs(pgradz)), pgradz);
                    double pgrad = sqrt(pgradx * pgradx + pgrady * pgrady + pgradz * pgradz);
                    double xgrad = fabs(D1X(celldx, j) * pgrad / pgradx);
                    double ygrad = fabs(D1Y(celldy, k) * pgrad / pgrady);
                    double zgrad = fabs(D1Z(celldz, l) * pgrad / pgradz);
                    double grad = fmin2(fmin2(xgrad, ygrad), zgrad);
                    double grad2 = grad * grad;
                    C_(viscosity, j, k, l) = 2.0 * C_(density0, j, k, l) * grad2 * limiter * limiter;
                }
            }
}

/* ------------------------------------------------------------------------ */
/* calc_dt_kernel.f90:96-181 */
void oracle_calc_dt_kernel(OR_DIMS, double g_small, double g_big, double dtmin,
                           double dtc_safe, double dtu_safe, double dtv_safe, double dtw_safe,
                           double dtdiv_safe,
                           const double *xarea, const double *yarea, const double *zarea,
                           const double *cellx, const double *celly, const double *cellz,
                           const double *celldx, const double *celldy, const double *celldz,
                           const double *volume, const double *density0, const double *energy0,
                           const double *pressure, const double *viscosity_a, const double *soundspeed,
                           const double *xvel0, const double *yvel0, const double *zvel0,
                           double *dt_min, double *dt_min_val, int *dtl_control,
                           double *xl_pos, double *yl_pos, double *zl_pos,
                           int *jldt, int *kldt, int *lldt, int *small)
{
    SHAPES;
    (void)energy0;
    (void)pressure;
    *small = 0;
    double dmin = g_big;
    double jkl_control = (double)1.1f; /* default-real literal (calc_dt_kernel.f90:97) */

#pragma omp parallel for
    for (int l = z_min; l <= z_max; l++)
        for (int k = y_min; k <= y_max; k++)
            for (int j = x_min; j <= x_max; j++) {
                double m = fmin2(fmin2(D1X(celldx, j), D1Y(celldy, k)), D1Z(celldz, l));
                double ds = 1.0 / (m * m);

                double cc = C_(soundspeed, j, k, l) * C_(soundspeed, j, k, l);
                cc = cc + 2.0 * C_(viscosity_a, j, k, l) / C_(density0, j, k, l);

                double dtct = ds * cc;
                dtct = dtc_safe * 1.0 / fmax2(sqrt(dtct), g_small);

                double du1 = (V_(xvel0, j, k, l) + V_(xvel0, j, k + 1, l) + V_(xvel0, j, k, l + 1) + V_(xvel0, j, k + 1, l + 1)) *
                             X_(xarea, j, k, l);
                double du2 = (V_(xvel0, j + 1, k, l) + V_(xvel0, j + 1, k + 1, l) + V_(xvel0, j + 1, k, l + 1) +
                              V_(xvel0, j + 1, k + 1, l + 1)) *
                             X_(xarea, j, k, l);

                double vol = C_(volume, j, k, l);
                double dtut = dtu_safe * 4.0 * vol / fmax2(fmax2(fabs(du1), fabs(du2)), 1.0e-5 * vol);

                double dv1 = (V_(yvel0, j, k, l) + V_(yvel0, j + 1, k, l) + V_(yvel0, j, k, l + 1) + V_(yvel0, j + 1, k, l + 1)) *
                             Y_(yarea, j, k, l);
                double dv2 = (V_(yvel0, j, k + 1, l) + V_(yvel0, j + 1, k + 1, l) + V_(yvel0, j, k + 1, l + 1) +
                              V_(yvel0, j + 1, k + 1, l + 1)) *
                             Y_(yarea, j, k, l);

                double dtvt = dtv_safe * 4.0 * vol / fmax2(fmax2(fabs(dv1), fabs(dv2)), 1.0e-5 * vol);

                double dw1 = (V_(zvel0, j, k, l) + V_(zvel0, j, k + 1, l) + V_(zvel0, j + 1, k, l) + V_(zvel0, j + 1, k + 1, l)) *
                             Z_(zarea, j, k, l);
                double dw2 = (V_(zvel0, j, k, l + 1) + V_(zvel0, j, k + 1, l + 1) + V_(zvel0, j + 1, k, l + 1) +
                              V_(zvel0, j + 1, k + 1, l + 1)) *
                             Z_(zarea, j, k, l);

                double dtwt = dtw_safe * 4.0 * vol / fmax2(fmax2(fabs(dw1), fabs(dw2)), 1.0e-5 * vol);

                double div = du2 - du1 + dv2 - dv1 + dw2 - dw1;

                double dtdivt = dtdiv_safe * 4.0 * (vol) / fmax2(vol * 1.0e-05, fabs(div));

                V_(dt_min, j, k, l) = fmin2(fmin2(fmin2(fmin2(dtct, dtut), dtvt), dtwt), dtdivt);
            }

    for (int l = z_min; l <= z_max; l++)
        for (int k = y_min; k <= y_max; k++)
            for (int j = x_min; j <= x_max; j++)
                if (V_(dt_min, j, k, l) < dmin) dmin = V_(dt_min, j, k, l);
    *dt_min_val = dmin;

    /* calc_dt_kernel.f90:152-161: location outputs derived from the constant
     * jkl_control; 10.01 is a default-real literal. */
    *dtl_control = (int)((double)10.01f * (jkl_control - (double)(int)jkl_control));
    jkl_control = jkl_control - (jkl_control - (double)(int)jkl_control);
    *jldt = ((int)jkl_control) % x_max;
    *kldt = (int)(1.0 + (jkl_control / (double)x_max));
    *lldt = (int)(1.0 + (jkl_control / (double)x_max));
    *xl_pos = D1X(cellx, *jldt);
    *yl_pos = D1Y(celly, *kldt);
    *zl_pos = D1Z(cellz, *lldt);

    if (dmin < dtmin) *small = 1;
}

/* ------------------------------------------------------------------------ */
/* PdV_kernel.f90:74-165 */
void oracle_PdV_kernel(int predict, OR_DIMS, double dt,
                       const double *xarea, const double *yarea, const double *zarea, const double *volume,
                       const double *density0, double *density1, const double *energy0, double *energy1,
                       const double *pressure, const double *viscosity,
                       const double *xvel0, const double *xvel1, const double *yvel0, const double *yvel1,
                       const double *zvel0, const double *zvel1)
{
    SHAPES;
    /* The predictor sums the level-0 face velocities twice and halves the
     * time step (80-97); the corrector sums level 0 and level 1 (128-145). */
    const double *xb = predict ? xvel0 : xvel1;
    const double *yb = predict ? yvel0 : yvel1;
    const double *zb = predict ? zvel0 : zvel1;
#pragma omp parallel for
    for (int l = z_min; l <= z_max; l++)
        for (int k = y_min; k <= y_max; k++)
            for (int j = x_min; j <= x_max; j++) {
#define SX4(a, jj) (V_(a, jj, k, l) + V_(a, jj, k + 1, l) + V_(a, jj, k, l + 1) + V_(a, jj, k + 1, l + 1))
#define SY4(a, kk) (V_(a, j, kk, l) + V_(a, j + 1, kk, l) + V_(a, j, kk, l + 1) + V_(a, j + 1, kk, l + 1))
#define SZ4(a, ll) (V_(a, j, k, ll) + V_(a, j + 1, k, ll) + V_(a, j, k + 1, ll) + V_(a, j + 1, k + 1, ll))
                /* 8-term sums written out left to right as in the source */
                double sl = V_(xvel0, j, k, l) + V_(xvel0, j, k + 1, l) + V_(xvel0, j, k, l + 1) + V_(xvel0, j, k + 1, l + 1) +
                            V_(xb, j, k, l) + V_(xb, j, k + 1, l) + V_(xb, j, k, l + 1) + V_(xb, j, k + 1, l + 1);
                double sr = V_(xvel0, j + 1, k, l) + V_(xvel0, j + 1, k + 1, l) + V_(xvel0, j + 1, k, l + 1) +
                            V_(xvel0, j + 1, k + 1, l + 1) + V_(xb, j + 1, k, l) + V_(xb, j + 1, k + 1, l) +
                            V_(xb, j + 1, k, l + 1) + V_(xb, j + 1, k + 1, l + 1);
                double sb = V_(yvel0, j, k, l) + V_(yvel0, j + 1, k, l) + V_(yvel0, j, k, l + 1) + V_(yvel0, j + 1, k, l + 1) +
                            V_(yb, j, k, l) + V_(yb, j + 1, k, l) + V_(yb, j, k, l + 1) + V_(yb, j + 1, k, l + 1);
                double st = V_(yvel0, j, k + 1, l) + V_(yvel0, j + 1, k + 1, l) + V_(yvel0, j, k + 1, l + 1) +
                            V_(yvel0, j + 1, k + 1, l + 1) + V_(yb, j, k + 1, l) + V_(yb, j + 1, k + 1, l) +
                            V_(yb, j, k + 1, l + 1) + V_(yb, j + 1, k + 1, l + 1);
                double sk = V_(zvel0, j, k, l) + V_(zvel0, j + 1, k, l) + V_(zvel0, j, k + 1, l) + V_(zvel0, j + 1, k + 1, l) +
                            V_(zb, j, k, l) + V_(zb, j + 1, k, l) + V_(zb, j, k + 1, l) + V_(zb, j + 1, k + 1, l);
                double sf = V_(zvel0, j, k, l + 1) + V_(zvel0, j + 1, k, l + 1) + V_(zvel0, j, k + 1, l + 1) +
                            V_(zvel0, j + 1, k + 1, l + 1) + V_(zb, j, k, l + 1) + V_(zb, j + 1, k, l + 1) +
                            V_(zb, j, k + 1, l + 1) + V_(zb, j + 1, k + 1, l + 1);
#undef SX4
#undef SY4
#undef SZ4
                double left_flux, right_flux, bottom_flux, top_flux, back_flux, front_flux;
                if (predict) {
                    left_flux = (X_(xarea, j, k, l) * sl) * 0.125 * dt * 0.5;
                    right_flux = (X_(xarea, j + 1, k, l) * sr) * 0.125 * dt * 0.5;
                    bottom_flux = (Y_(yarea, j, k, l) * sb) * 0.125 * dt * 0.5;
                    top_flux = (Y_(yarea, j, k + 1, l) * st) * 0.125 * dt * 0.5;
                    back_flux = (Z_(zarea, j, k, l) * sk) * 0.125 * dt * 0.5;
                    front_flux = (Z_(zarea, j, k, l + 1) * sf) * 0.125 * dt * 0.5;
                } else {
                    left_flux = (X_(xarea, j, k, l) * sl) * 0.125 * dt;
                    right_flux = (X_(xarea, j + 1, k, l) * sr) * 0.125 * dt;
                    bottom_flux = (Y_(yarea, j, k, l) * sb) * 0.125 * dt;
                    top_flux = (Y_(yarea, j, k + 1, l) * st) * 0.125 * dt;
                    back_flux = (Z_(zarea, j, k, l) * sk) * 0.125 * dt;
                    front_flux = (Z_(zarea, j, k, l + 1) * sf) * 0.125 * dt;
                }
                double total_flux = right_flux - left_flux + top_flux - bottom_flux + front_flux - back_flux;

                double vol_change_s = C_(volume, j, k, l) / (C_(volume, j, k, l) + total_flux);

                /* min_cell_volume (102-105) is computed but never used. */

                double recip_volume = 1.0 / C_(volume, j, k, l);

                double energy_change = (C_(pressure, j, k, l) / C_(density0, j, k, l) +
                                        C_(viscosity, j, k, l) / C_(density0, j, k, l)) *
                                       total_flux * recip_volume;

                C_(energy1, j, k, l) = C_(energy0, j, k, l) - energy_change;

                C_(density1, j, k, l) = C_(density0, j, k, l) * vol_change_s;
            }
}

/* ------------------------------------------------------------------------ */
/* revert_kernel.f90:44-61 */
void oracle_revert_kernel(OR_DIMS, const double *density0, double *density1,
                          const double *energy0, double *energy1)
{
    SHAPES;
#pragma omp parallel for
    for (int l = z_min; l <= z_max; l++)
        for (int k = y_min; k <= y_max; k++)
            for (int j = x_min; j <= x_max; j++) {
                C_(density1, j, k, l) = C_(density0, j, k, l);
                C_(energy1, j, k, l) = C_(energy0, j, k, l);
            }
}

/* ------------------------------------------------------------------------ */
/* accelerate_kernel.f90:61-109 */
void oracle_accelerate_kernel(OR_DIMS, double dt,
                              const double *xarea, const double *yarea, const double *zarea,
                              const double *volume, const double *density0, const double *pressure,
                              const double *viscosity,
                              const double *xvel0, const double *yvel0, const double *zvel0,
                              double *xvel1, double *yvel1, double *zvel1)
{
    SHAPES;
#pragma omp parallel for
    for (int l = z_min; l <= z_max + 1; l++)
        for (int k = y_min; k <= y_max + 1; k++)
            for (int j = x_min; j <= x_max + 1; j++) {
#define RV(a, b, c) (C_(density0, a, b, c) * C_(volume, a, b, c))
                double nodal_mass = (RV(j - 1, k - 1, l) + RV(j, k - 1, l) + RV(j, k, l) + RV(j - 1, k, l) +
                                     RV(j - 1, k - 1, l - 1) + RV(j, k - 1, l - 1) + RV(j, k, l - 1) + RV(j - 1, k, l - 1)) *
                                    0.125;
#undef RV
                double stepbymass_s = 0.25 * dt / nodal_mass;
#define P(a, b, c) C_(pressure, a, b, c)
#define Q(a, b, c) C_(viscosity, a, b, c)
                double xvel1_s = V_(xvel0, j, k, l) -
                                 stepbymass_s * (X_(xarea, j, k, l) * (P(j, k, l) - P(j - 1, k, l)) +
                                                 X_(xarea, j, k - 1, l) * (P(j, k - 1, l) - P(j - 1, k - 1, l)) +
                                                 X_(xarea, j, k, l - 1) * (P(j, k, l - 1) - P(j - 1, k, l - 1)) +
                                                 X_(xarea, j, k - 1, l - 1) * (P(j, k - 1, l - 1) - P(j - 1, k - 1, l - 1)));
                double yvel1_s = V_(yvel0, j, k, l) -
                                 stepbymass_s * (Y_(yarea, j, k, l) * (P(j, k, l) - P(j, k - 1, l)) +
                                                 Y_(yarea, j - 1, k, l) * (P(j - 1, k, l) - P(j - 1, k - 1, l)) +
                                                 Y_(yarea, j, k, l - 1) * (P(j, k, l - 1) - P(j, k - 1, l - 1)) +
                                                 Y_(yarea, j - 1, k, l - 1) * (P(j - 1, k, l - 1) - P(j - 1, k - 1, l - 1)));
                double zvel1_s = V_(zvel0, j, k, l) -
                                 stepbymass_s * (Z_(zarea, j, k, l) * (P(j, k, l) - P(j, k, l - 1)) +
                                                 Z_(zarea, j, k - 1, l) * (P(j, k - 1, l) - P(j, k - 1, l - 1)) +
                                                 Z_(zarea, j - 1, k, l) * (P(j - 1, k, l) - P(j - 1, k, l - 1)) +
                                                 Z_(zarea, j - 1, k - 1, l) * (P(j - 1, k - 1, l) - P(j - 1, k - 1, l - 1)));
                V_(xvel1, j, k, l) = xvel1_s -
                                     stepbymass_s * (X_(xarea, j, k, l) * (Q(j, k, l) - Q(j - 1, k, l)) +
                                                     X_(xarea, j, k - 1, l) * (Q(j, k - 1, l) - Q(j - 1, k - 1, l)) +
                                                     X_(xarea, j, k, l - 1) * (Q(j, k, l - 1) - Q(j - 1, k, l - 1)) +
                                                     X_(xarea, j, k - 1, l - 1) * (Q(j, k - 1, l - 1) - Q(j - 1, k - 1, l - 1)));
                V_(yvel1, j, k, l) = yvel1_s -
                                     stepbymass_s * (Y_(yarea, j, k, l) * (Q(j, k, l) - Q(j, k - 1, l)) +
                                                     Y_(yarea, j - 1, k, l) * (Q(j - 1, k, l) - Q(j - 1, k - 1, l)) +
                                                     Y_(yarea, j, k, l - 1) * (Q(j, k, l - 1) - Q(j, k - 1, l - 1)) +
                                                     Y_(yarea, j - 1, k, l - 1) * (Q(j - 1, k, l - 1) - Q(j - 1, k - 1, l - 1)));
                V_(zvel1, j, k, l) = zvel1_s -
                                     stepbymass_s * (Z_(zarea, j, k, l) * (Q(j, k, l) - Q(j, k, l - 1)) +
                                                     Z_(zarea, j, k - 1, l) * (Q(j, k - 1, l) - Q(j, k - 1, l - 1)) +
                                                     Z_(zarea, j - 1, k, l) * (Q(j - 1, k, l) - Q(j - 1, k, l - 1)) +
                                                     Z_(zarea, j - 1, k - 1, l) * (Q(j - 1, k - 1, l) - Q(j - 1, k - 1, l - 1)));
#undef P
#undef Q
            }
}

/* ------------------------------------------------------------------------ */
/* flux_calc_kernel.f90:56-90 (vol_flux_z keeps the reference's repeated
 * zvel0(j+1,k,l) term, flux_calc_kernel.f90:86) */
void oracle_flux_calc_kernel(OR_DIMS, double dt,
                             const double *xarea, const double *yarea, const double *zarea,
                             const double *xvel0, const double *yvel0, const double *zvel0,
                             const double *xvel1, const double *yvel1, const double *zvel1,
                             double *vol_flux_x, double *vol_flux_y, double *vol_flux_z)
{
    SHAPES;
#pragma omp parallel for
    for (int l = z_min; l <= z_max; l++)
        for (int k = y_min; k <= y_max; k++)
            for (int j = x_min; j <= x_max + 1; j++)
                X_(vol_flux_x, j, k, l) =
                    0.125 * dt * X_(xarea, j, k, l) *
                    (V_(xvel0, j, k, l) + V_(xvel0, j, k + 1, l) + V_(xvel0, j, k, l + 1) + V_(xvel0, j, k + 1, l + 1) +
                     V_(xvel1, j, k, l) + V_(xvel1, j, k + 1, l) + V_(xvel1, j, k, l + 1) + V_(xvel1, j, k + 1, l + 1));
#pragma omp parallel for
    for (int l = z_min; l <= z_max; l++)
        for (int k = y_min; k <= y_max + 1; k++)
            for (int j = x_min; j <= x_max; j++)
                Y_(vol_flux_y, j, k, l) =
                    0.125 * dt * Y_(yarea, j, k, l) *
                    (V_(yvel0, j, k, l) + V_(yvel0, j + 1, k, l) + V_(yvel0, j, k, l + 1) + V_(yvel0, j + 1, k, l + 1) +
                     V_(yvel1, j, k, l) + V_(yvel1, j + 1, k, l) + V_(yvel1, j, k, l + 1) + V_(yvel1, j + 1, k, l + 1));
#pragma omp parallel for
    for (int l = z_min; l <= z_max + 1; l++)
        for (int k = y_min; k <= y_max; k++)
            for (int j = x_min; j <= x_max; j++)
                Z_(vol_flux_z, j, k, l) =
                    0.125 * dt * Z_(zarea, j, k, l) *
                    (V_(zvel0, j, k, l) + V_(zvel0, j + 1, k, l) + V_(zvel0, j + 1, k, l) + V_(zvel0, j + 1, k + 1, l) +
                     V_(zvel1, j, k, l) + V_(zvel1, j + 1, k, l) + V_(zvel1, j, k + 1, l) + V_(zvel1, j + 1, k + 1, l));
}

/* ------------------------------------------------------------------------ */
/* The van Leer limited face value shared by the three directions of
 * advec_cell_kernel.f90 (146-154 for density, 158-166 for energy). */
static inline double cell_limiter(double sigma, double sigma3, double sigma4, double diffuw, double diffdw)
{
    const double one_by_six = 1.0 / 6.0;
    if (diffuw * diffdw > 0.0)
        return (1.0 - sigma) * fsign(1.0, diffdw) *
               fmin2(fmin2(fabs(diffuw), fabs(diffdw)), one_by_six * (sigma3 * fabs(diffuw) + sigma4 * fabs(diffdw)));
    return 0.0;
}

/* advec_cell_kernel.f90:93-380 */
void oracle_advec_cell_kernel(OR_DIMS, int advect_x, int dir, int sweep_number,
                              const double *vertexdx, const double *vertexdy, const double *vertexdz,
                              const double *volume, double *density1, double *energy1,
                              double *mass_flux_x, const double *vol_flux_x,
                              double *mass_flux_y, const double *vol_flux_y,
                              double *mass_flux_z, const double *vol_flux_z,
                              double *pre_vol, double *post_vol, double *pre_mass, double *post_mass,
                              double *advec_vol, double *post_ener, double *ener_flux)
{
    SHAPES;
    if (dir == 1) {
        if (sweep_number == 1) {
#pragma omp parallel for
            for (int l = z_min - 2; l <= z_max + 2; l++)
                for (int k = y_min - 2; k <= y_max + 2; k++)
                    for (int j = x_min - 2; j <= x_max + 2; j++) {
                        V_(pre_vol, j, k, l) = C_(volume, j, k, l) +
                                               (X_(vol_flux_x, j + 1, k, l) - X_(vol_flux_x, j, k, l) + Y_(vol_flux_y, j, k + 1, l) -
                                                Y_(vol_flux_y, j, k, l) + Z_(vol_flux_z, j, k, l + 1) - Z_(vol_flux_z, j, k, l));
                        V_(post_vol, j, k, l) = V_(pre_vol, j, k, l) - (X_(vol_flux_x, j + 1, k, l) - X_(vol_flux_x, j, k, l));
                    }
        } else if (sweep_number == 3) {
#pragma omp parallel for
            for (int l = z_min - 2; l <= z_max + 2; l++)
                for (int k = y_min - 2; k <= y_max + 2; k++)
                    for (int j = x_min - 2; j <= x_max + 2; j++) {
                        V_(pre_vol, j, k, l) = C_(volume, j, k, l) + X_(vol_flux_x, j + 1, k, l) - X_(vol_flux_x, j, k, l);
                        V_(post_vol, j, k, l) = C_(volume, j, k, l);
                    }
        }
#pragma omp parallel for
        for (int l = z_min; l <= z_max; l++)
            for (int k = y_min; k <= y_max; k++)
                for (int j = x_min; j <= x_max + 2; j++) {
                    int upwind, donor, downwind, dif;
                    if (X_(vol_flux_x, j, k, l) > 0.0) {
                        upwind = j - 2; donor = j - 1; downwind = j; dif = donor;
                    } else {
                        upwind = (j + 1 < x_max + 2) ? j + 1 : x_max + 2; donor = j; downwind = j - 1; dif = upwind;
                    }
                    double sigmat = fabs(X_(vol_flux_x, j, k, l)) / V_(pre_vol, donor, k, l);
                    double sigma3 = (1.0 + sigmat) * (D1X(vertexdx, j) / D1X(vertexdx, dif));
                    double sigma4 = 2.0 - sigmat;
                    double sigmav = sigmat;
                    double diffuw = C_(density1, donor, k, l) - C_(density1, upwind, k, l);
                    double diffdw = C_(density1, downwind, k, l) - C_(density1, donor, k, l);
                    double limiter = cell_limiter(sigmav, sigma3, sigma4, diffuw, diffdw);
                    X_(mass_flux_x, j, k, l) = X_(vol_flux_x, j, k, l) * (C_(density1, donor, k, l) + limiter);

                    double sigmam = fabs(X_(mass_flux_x, j, k, l)) / (C_(density1, donor, k, l) * V_(pre_vol, donor, k, l));
                    diffuw = C_(energy1, donor, k, l) - C_(energy1, upwind, k, l);
                    diffdw = C_(energy1, downwind, k, l) - C_(energy1, donor, k, l);
                    limiter = cell_limiter(sigmam, sigma3, sigma4, diffuw, diffdw);
                    V_(ener_flux, j, k, l) = X_(mass_flux_x, j, k, l) * (C_(energy1, donor, k, l) + limiter);
                }
#pragma omp parallel for
        for (int l = z_min; l <= z_max; l++)
            for (int k = y_min; k <= y_max; k++)
                for (int j = x_min; j <= x_max; j++) {
                    V_(pre_mass, j, k, l) = C_(density1, j, k, l) * V_(pre_vol, j, k, l);
                    V_(post_mass, j, k, l) = V_(pre_mass, j, k, l) + X_(mass_flux_x, j, k, l) - X_(mass_flux_x, j + 1, k, l);
                    V_(post_ener, j, k, l) = (C_(energy1, j, k, l) * V_(pre_mass, j, k, l) + V_(ener_flux, j, k, l) -
                                              V_(ener_flux, j + 1, k, l)) /
                                             V_(post_mass, j, k, l);
                    V_(advec_vol, j, k, l) = V_(pre_vol, j, k, l) + X_(vol_flux_x, j, k, l) - X_(vol_flux_x, j + 1, k, l);
                    C_(density1, j, k, l) = V_(post_mass, j, k, l) / V_(advec_vol, j, k, l);
                    C_(energy1, j, k, l) = V_(post_ener, j, k, l);
                }
    } else if (dir == 2) {
        if (sweep_number == 2) {
            if (advect_x) {
#pragma omp parallel for
                for (int l = z_min - 2; l <= z_max + 2; l++)
                    for (int k = y_min - 2; k <= y_max + 2; k++)
                        for (int j = x_min - 2; j <= x_max + 2; j++) {
                            V_(pre_vol, j, k, l) = C_(volume, j, k, l) + Y_(vol_flux_y, j, k + 1, l) - Y_(vol_flux_y, j, k, l) +
                                                   Z_(vol_flux_z, j, k, l + 1) - Z_(vol_flux_z, j, k, l);
                            V_(post_vol, j, k, l) = V_(pre_vol, j, k, l) - (Y_(vol_flux_y, j, k + 1, l) - Y_(vol_flux_y, j, k, l));
                        }
            } else {
#pragma omp parallel for
                for (int l = z_min - 2; l <= z_max + 2; l++)
                    for (int k = y_min - 2; k <= y_max + 2; k++)
                        for (int j = x_min - 2; j <= x_max + 2; j++) {
                            V_(pre_vol, j, k, l) = C_(volume, j, k, l) + Y_(vol_flux_y, j, k + 1, l) - Y_(vol_flux_y, j, k, l) +
                                                   X_(vol_flux_x, j + 1, k, l) - X_(vol_flux_x, j, k, l);
                            V_(post_vol, j, k, l) = V_(pre_vol, j, k, l) - (Y_(vol_flux_y, j, k + 1, l) - Y_(vol_flux_y, j, k, l));
                        }
            }
        }
#pragma omp parallel for
        for (int l = z_min; l <= z_max + 2; l++)
            for (int k = y_min; k <= y_max + 2; k++)
                for (int j = x_min; j <= x_max; j++) {
                    int upwind, donor, downwind, dif;
                    if (Y_(vol_flux_y, j, k, l) > 0.0) {
                        upwind = k - 2; donor = k - 1; downwind = k; dif = donor;
                    } else {
                        upwind = (k + 1 < y_max + 2) ? k + 1 : y_max + 2; donor = k; downwind = k - 1; dif = upwind;
                    }
                    double sigmat = fabs(Y_(vol_flux_y, j, k, l)) / V_(pre_vol, j, donor, l);
                    double sigma3 = (1.0 + sigmat) * (D1Y(vertexdy, k) / D1Y(vertexdy, dif));
                    double sigma4 = 2.0 - sigmat;
                    double sigmav = sigmat;
                    double diffuw = C_(density1, j, donor, l) - C_(density1, j, upwind, l);
                    double diffdw = C_(density1, j, downwind, l) - C_(density1, j, donor, l);
                    double limiter = cell_limiter(sigmav, sigma3, sigma4, diffuw, diffdw);
                    Y_(mass_flux_y, j, k, l) = Y_(vol_flux_y, j, k, l) * (C_(density1, j, donor, l) + limiter);

                    double sigmam = fabs(Y_(mass_flux_y, j, k, l)) / (C_(density1, j, donor, l) * V_(pre_vol, j, donor, l));
                    diffuw = C_(energy1, j, donor, l) - C_(energy1, j, upwind, l);
                    diffdw = C_(energy1, j, downwind, l) - C_(energy1, j, donor, l);
                    limiter = cell_limiter(sigmam, sigma3, sigma4, diffuw, diffdw);
                    V_(ener_flux, j, k, l) = Y_(mass_flux_y, j, k, l) * (C_(energy1, j, donor, l) + limiter);
                }
#pragma omp parallel for
        for (int l = z_min; l <= z_max; l++)
            for (int k = y_min; k <= y_max; k++)
                for (int j = x_min; j <= x_max; j++) {
                    V_(pre_mass, j, k, l) = C_(density1, j, k, l) * V_(pre_vol, j, k, l);
                    V_(post_mass, j, k, l) = V_(pre_mass, j, k, l) + Y_(mass_flux_y, j, k, l) - Y_(mass_flux_y, j, k + 1, l);
                    V_(post_ener, j, k, l) = (C_(energy1, j, k, l) * V_(pre_mass, j, k, l) + V_(ener_flux, j, k, l) -
                                              V_(ener_flux, j, k + 1, l)) /
                                             V_(post_mass, j, k, l);
                    V_(advec_vol, j, k, l) = V_(pre_vol, j, k, l) + Y_(vol_flux_y, j, k, l) - Y_(vol_flux_y, j, k + 1, l);
                    C_(density1, j, k, l) = V_(post_mass, j, k, l) / V_(advec_vol, j, k, l);
                    C_(energy1, j, k, l) = V_(post_ener, j, k, l);
                }
    } else if (dir == 3) {
        if (sweep_number == 1) {
#pragma omp parallel for
            for (int l = z_min - 2; l <= z_max + 2; l++)
                for (int k = y_min - 2; k <= y_max + 2; k++)
                    for (int j = x_min - 2; j <= x_max + 2; j++) {
                        V_(pre_vol, j, k, l) = C_(volume, j, k, l) +
                                               (X_(vol_flux_x, j + 1, k, l) - X_(vol_flux_x, j, k, l) + Y_(vol_flux_y, j, k + 1, l) -
                                                Y_(vol_flux_y, j, k, l) + Z_(vol_flux_z, j, k, l + 1) - Z_(vol_flux_z, j, k, l));
                        V_(post_vol, j, k, l) = V_(pre_vol, j, k, l) - (Z_(vol_flux_z, j, k, l + 1) - Z_(vol_flux_z, j, k, l));
                    }
        } else if (sweep_number == 3) {
#pragma omp parallel for
            for (int l = z_min - 2; l <= z_max + 2; l++)
                for (int k = y_min - 2; k <= y_max + 2; k++)
                    for (int j = x_min - 2; j <= x_max + 2; j++) {
                        V_(pre_vol, j, k, l) = C_(volume, j, k, l) + Z_(vol_flux_z, j, k, l + 1) - Z_(vol_flux_z, j, k, l);
                        V_(post_vol, j, k, l) = C_(volume, j, k, l);
                    }
        }
#pragma omp parallel for
        for (int l = z_min; l <= z_max + 2; l++)
            for (int k = y_min; k <= y_max; k++)
                for (int j = x_min; j <= x_max; j++) {
                    int upwind, donor, downwind, dif;
                    if (Z_(vol_flux_z, j, k, l) > 0.0) {
                        upwind = l - 2; donor = l - 1; downwind = l; dif = donor;
                    } else {
                        upwind = (l + 1 < z_max + 2) ? l + 1 : z_max + 2; donor = l; downwind = l - 1; dif = upwind;
                    }
                    double sigmat = fabs(Z_(vol_flux_z, j, k, l)) / V_(pre_vol, j, k, donor);
                    double sigma3 = (1.0 + sigmat) * (D1Z(vertexdz, l) / D1Z(vertexdz, dif));
                    double sigma4 = 2.0 - sigmat;
                    double sigmav = sigmat;
                    double diffuw = C_(density1, j, k, donor) - C_(density1, j, k, upwind);
                    double diffdw = C_(density1, j, k, downwind) - C_(density1, j, k, donor);
                    double limiter = cell_limiter(sigmav, sigma3, sigma4, diffuw, diffdw);
                    Z_(mass_flux_z, j, k, l) = Z_(vol_flux_z, j, k, l) * (C_(density1, j, k, donor) + limiter);

                    double sigmam = fabs(Z_(mass_flux_z, j, k, l)) / (C_(density1, j, k, donor) * V_(pre_vol, j, k, donor));
                    diffuw = C_(energy1, j, k, donor) - C_(energy1, j, k, upwind);
                    diffdw = C_(energy1, j, k, downwind) - C_(energy1, j, k, donor);
                    limiter = cell_limiter(sigmam, sigma3, sigma4, diffuw, diffdw);
                    V_(ener_flux, j, k, l) = Z_(mass_flux_z, j, k, l) * (C_(energy1, j, k, donor) + limiter);
                }
#pragma omp parallel for
        for (int l = z_min; l <= z_max; l++)
            for (int k = y_min; k <= y_max; k++)
                for (int j = x_min; j <= x_max; j++) {
                    V_(pre_mass, j, k, l) = C_(density1, j, k, l) * V_(pre_vol, j, k, l);
                    V_(post_mass, j, k, l) = V_(pre_mass, j, k, l) + Z_(mass_flux_z, j, k, l) - Z_(mass_flux_z, j, k, l + 1);
                    V_(post_ener, j, k, l) = (C_(energy1, j, k, l) * V_(pre_mass, j, k, l) + V_(ener_flux, j, k, l) -
                                              V_(ener_flux, j, k, l + 1)) /
                                             V_(post_mass, j, k, l);
                    V_(advec_vol, j, k, l) = V_(pre_vol, j, k, l) + Z_(vol_flux_z, j, k, l) - Z_(vol_flux_z, j, k, l + 1);
                    C_(density1, j, k, l) = V_(post_mass, j, k, l) / V_(advec_vol, j, k, l);
                    C_(energy1, j, k, l) = V_(post_ener, j, k, l);
                }
    }
}

/* ------------------------------------------------------------------------ */
/* Limited momentum flux shared by the three directions of
 * advec_mom_kernel.f90 (223-240, 308-325, 394-411). */
static inline double mom_face(double vdonor, double vupwind, double vdownwind, double sigma, double width,
                              double celld_dif)
{
    double vdiffuw = vdonor - vupwind;
    double vdiffdw = vdownwind - vdonor;
    double limiter = 0.0;
    if (vdiffuw * vdiffdw > 0.0) {
        double auw = fabs(vdiffuw);
        double adw = fabs(vdiffdw);
        double wind = 1.0;
        if (vdiffdw <= 0.0) wind = -1.0;
        limiter = wind * fmin2(fmin2(width * ((2.0 - sigma) * adw / width + (1.0 + sigma) * auw / celld_dif) / 6.0, auw), adw);
    }
    return vdonor + (1.0 - sigma) * limiter;
}

/* advec_mom_kernel.f90:91-424 */
void oracle_advec_mom_kernel(OR_DIMS, double *vel1,
                             const double *mass_flux_x, const double *vol_flux_x,
                             const double *mass_flux_y, const double *vol_flux_y,
                             const double *mass_flux_z, const double *vol_flux_z,
                             const double *volume, const double *density1,
                             double *node_flux, double *node_mass_post, double *node_mass_pre,
                             double *advec_vel, double *mom_flux, double *pre_vol, double *post_vol,
                             const double *celldx, const double *celldy, const double *celldz,
                             int advect_x, int which_vel, int sweep_number, int direction)
{
    SHAPES;
    (void)advec_vel;
    /* (i) staggered volumes, all cells including the 2-deep halo (95-163) */
    int mode = 0;
    if (sweep_number == 1 && direction == 1) mode = 1;
    else if (sweep_number == 1 && direction == 3) mode = 2;
    else if (sweep_number == 2 && advect_x) mode = 3;
    else if (sweep_number == 2 && !advect_x) mode = 4;
    else if (sweep_number == 3 && direction == 1) mode = 5;
    else if (sweep_number == 3 && direction == 3) mode = 6;
    if (mode) {
#pragma omp parallel for
        for (int l = z_min - 2; l <= z_max + 2; l++)
            for (int k = y_min - 2; k <= y_max + 2; k++)
                for (int j = x_min - 2; j <= x_max + 2; j++) {
                    double V0 = C_(volume, j, k, l);
                    double dx_ = X_(vol_flux_x, j + 1, k, l), dx0 = X_(vol_flux_x, j, k, l);
                    double dy_ = Y_(vol_flux_y, j, k + 1, l), dy0 = Y_(vol_flux_y, j, k, l);
                    double dz_ = Z_(vol_flux_z, j, k, l + 1), dz0 = Z_(vol_flux_z, j, k, l);
                    double post, pre;
                    switch (mode) {
                    case 1: post = V0 + dy_ - dy0 + dz_ - dz0; pre = post + dx_ - dx0; break;
                    case 2: post = V0 + dx_ - dx0 + dy_ - dy0; pre = post + dz_ - dz0; break;
                    case 3: post = V0 + dz_ - dz0; pre = post + dy_ - dy0; break;
                    case 4: post = V0 + dx_ - dx0; pre = post + dy_ - dy0; break;
                    case 5: post = V0; pre = post + dx_ - dx0; break;
                    default: post = V0; pre = post + dz_ - dz0; break;
                    }
                    V_(post_vol, j, k, l) = post;
                    V_(pre_vol, j, k, l) = pre;
                }
    }

#define NMP(j_, k_, l_) (C_(density1, j_, k_, l_) * V_(post_vol, j_, k_, l_))
    if (direction == 1) {
        if (which_vel == 1) {
#pragma omp parallel for
            for (int l = z_min; l <= z_max + 1; l++)
                for (int k = y_min; k <= y_max + 1; k++)
                    for (int j = x_min - 2; j <= x_max + 2; j++)
                        V_(node_flux, j, k, l) =
                            0.125 * (X_(mass_flux_x, j, k - 1, l) + X_(mass_flux_x, j, k, l) + X_(mass_flux_x, j + 1, k - 1, l) +
                                     X_(mass_flux_x, j + 1, k, l) + X_(mass_flux_x, j, k - 1, l - 1) + X_(mass_flux_x, j, k, l - 1) +
                                     X_(mass_flux_x, j + 1, k - 1, l - 1) + X_(mass_flux_x, j + 1, k, l - 1));
#pragma omp parallel for
            for (int l = z_min; l <= z_max + 1; l++)
                for (int k = y_min; k <= y_max + 1; k++)
                    for (int j = x_min - 1; j <= x_max + 2; j++)
                        V_(node_mass_post, j, k, l) =
                            0.125 * (NMP(j, k - 1, l) + NMP(j, k, l) + NMP(j - 1, k - 1, l) + NMP(j - 1, k, l) +
                                     NMP(j, k - 1, l - 1) + NMP(j, k, l - 1) + NMP(j - 1, k - 1, l - 1) + NMP(j - 1, k, l - 1));
#pragma omp parallel for
            for (int l = z_min; l <= z_max + 1; l++)
                for (int k = y_min; k <= y_max + 1; k++)
                    for (int j = x_min - 1; j <= x_max + 2; j++)
                        V_(node_mass_pre, j, k, l) = V_(node_mass_post, j, k, l) - V_(node_flux, j - 1, k, l) + V_(node_flux, j, k, l);
        }
#pragma omp parallel for
        for (int l = z_min; l <= z_max + 1; l++)
            for (int k = y_min; k <= y_max + 1; k++)
                for (int j = x_min - 1; j <= x_max + 1; j++) {
                    int upwind, donor, downwind, dif;
                    if (V_(node_flux, j, k, l) < 0.0) {
                        upwind = j + 2; donor = j + 1; downwind = j; dif = donor;
                    } else {
                        upwind = j - 1; donor = j; downwind = j + 1; dif = upwind;
                    }
                    double sigma = fabs(V_(node_flux, j, k, l)) / (V_(node_mass_pre, donor, k, l));
                    double advec_vel_s = mom_face(V_(vel1, donor, k, l), V_(vel1, upwind, k, l), V_(vel1, downwind, k, l), sigma,
                                                  D1X(celldx, j), D1X(celldx, dif));
                    V_(mom_flux, j, k, l) = advec_vel_s * V_(node_flux, j, k, l);
                }
#pragma omp parallel for
        for (int l = z_min; l <= z_max + 1; l++)
            for (int k = y_min; k <= y_max + 1; k++)
                for (int j = x_min; j <= x_max + 1; j++)
                    V_(vel1, j, k, l) = (V_(vel1, j, k, l) * V_(node_mass_pre, j, k, l) + V_(mom_flux, j - 1, k, l) -
                                         V_(mom_flux, j, k, l)) /
                                        V_(node_mass_post, j, k, l);
    } else if (direction == 2) {
        if (which_vel == 1) {
#pragma omp parallel for
            for (int l = z_min; l <= z_max + 1; l++)
                for (int k = y_min - 2; k <= y_max + 2; k++)
                    for (int j = x_min; j <= x_max + 1; j++)
                        V_(node_flux, j, k, l) =
                            0.125 * (Y_(mass_flux_y, j - 1, k, l) + Y_(mass_flux_y, j, k, l) + Y_(mass_flux_y, j - 1, k + 1, l) +
                                     Y_(mass_flux_y, j, k + 1, l) + Y_(mass_flux_y, j - 1, k, l - 1) + Y_(mass_flux_y, j, k, l - 1) +
                                     Y_(mass_flux_y, j - 1, k + 1, l - 1) + Y_(mass_flux_y, j, k + 1, l - 1));
#pragma omp parallel for
            for (int l = z_min; l <= z_max + 1; l++)
                for (int k = y_min - 1; k <= y_max + 2; k++)
                    for (int j = x_min; j <= x_max + 1; j++)
                        V_(node_mass_post, j, k, l) =
                            0.125 * (NMP(j, k - 1, l) + NMP(j, k, l) + NMP(j - 1, k - 1, l) + NMP(j - 1, k, l) +
                                     NMP(j, k - 1, l - 1) + NMP(j, k, l - 1) + NMP(j - 1, k - 1, l - 1) + NMP(j - 1, k, l - 1));
#pragma omp parallel for
            for (int l = z_min; l <= z_max + 1; l++)
                for (int k = y_min - 1; k <= y_max + 2; k++)
                    for (int j = x_min; j <= x_max + 1; j++)
                        V_(node_mass_pre, j, k, l) = V_(node_mass_post, j, k, l) - V_(node_flux, j, k - 1, l) + V_(node_flux, j, k, l);
        }
#pragma omp parallel for
        for (int l = z_min; l <= z_max + 1; l++)
            for (int k = y_min - 1; k <= y_max + 1; k++)
                for (int j = x_min; j <= x_max + 1; j++) {
                    int upwind, donor, downwind, dif;
                    if (V_(node_flux, j, k, l) < 0.0) {
                        upwind = k + 2; donor = k + 1; downwind = k; dif = donor;
                    } else {
                        upwind = k - 1; donor = k; downwind = k + 1; dif = upwind;
                    }
                    double sigma = fabs(V_(node_flux, j, k, l)) / (V_(node_mass_pre, j, donor, l));
                    double advec_vel_s = mom_face(V_(vel1, j, donor, l), V_(vel1, j, upwind, l), V_(vel1, j, downwind, l), sigma,
                                                  D1Y(celldy, k), D1Y(celldy, dif));
                    V_(mom_flux, j, k, l) = advec_vel_s * V_(node_flux, j, k, l);
                }
#pragma omp parallel for
        for (int l = z_min; l <= z_max + 1; l++)
            for (int k = y_min; k <= y_max + 1; k++)
                for (int j = x_min; j <= x_max + 1; j++)
                    V_(vel1, j, k, l) = (V_(vel1, j, k, l) * V_(node_mass_pre, j, k, l) + V_(mom_flux, j, k - 1, l) -
                                         V_(mom_flux, j, k, l)) /
                                        V_(node_mass_post, j, k, l);
    } else if (direction == 3) {
        if (which_vel == 1) {
#pragma omp parallel for
            for (int l = z_min - 2; l <= z_max + 2; l++)
                for (int k = y_min; k <= y_max + 1; k++)
                    for (int j = x_min; j <= x_max + 1; j++)
                        V_(node_flux, j, k, l) =
                            0.125 * (Z_(mass_flux_z, j - 1, k, l) + Z_(mass_flux_z, j, k, l) + Z_(mass_flux_z, j - 1, k, l + 1) +
                                     Z_(mass_flux_z, j, k, l + 1) + Z_(mass_flux_z, j - 1, k - 1, l) + Z_(mass_flux_z, j, k - 1, l) +
                                     Z_(mass_flux_z, j - 1, k - 1, l + 1) + Z_(mass_flux_z, j, k - 1, l + 1));
#pragma omp parallel for
            for (int l = z_min - 1; l <= z_max + 2; l++)
                for (int k = y_min; k <= y_max + 1; k++)
                    for (int j = x_min; j <= x_max + 1; j++)
                        V_(node_mass_post, j, k, l) =
                            0.125 * (NMP(j, k - 1, l) + NMP(j, k, l) + NMP(j - 1, k - 1, l) + NMP(j - 1, k, l) +
                                     NMP(j, k - 1, l - 1) + NMP(j, k, l - 1) + NMP(j - 1, k - 1, l - 1) + NMP(j - 1, k, l - 1));
#pragma omp parallel for
            for (int l = z_min - 1; l <= z_max + 2; l++)
                for (int k = y_min; k <= y_max + 1; k++)
                    for (int j = x_min; j <= x_max + 1; j++)
                        V_(node_mass_pre, j, k, l) = V_(node_mass_post, j, k, l) - V_(node_flux, j, k, l - 1) + V_(node_flux, j, k, l);
        }
#pragma omp parallel for
        for (int l = z_min - 1; l <= z_max + 1; l++)
            for (int k = y_min; k <= y_max + 1; k++)
                for (int j = x_min; j <= x_max + 1; j++) {
                    int upwind, donor, downwind, dif;
                    if (V_(node_flux, j, k, l) < 0.0) {
                        upwind = l + 2; donor = l + 1; downwind = l; dif = donor;
                    } else {
                        upwind = l - 1; donor = l; downwind = l + 1; dif = upwind;
                    }
                    double sigma = fabs(V_(node_flux, j, k, l)) / (V_(node_mass_pre, j, k, donor));
                    double advec_vel_s = mom_face(V_(vel1, j, k, donor), V_(vel1, j, k, upwind), V_(vel1, j, k, downwind), sigma,
                                                  D1Z(celldz, l), D1Z(celldz, dif));
                    V_(mom_flux, j, k, l) = advec_vel_s * V_(node_flux, j, k, l);
                }
#pragma omp parallel for
        for (int l = z_min; l <= z_max + 1; l++)
            for (int k = y_min; k <= y_max + 1; k++)
                for (int j = x_min; j <= x_max + 1; j++)
                    V_(vel1, j, k, l) = (V_(vel1, j, k, l) * V_(node_mass_pre, j, k, l) + V_(mom_flux, j, k, l - 1) -
                                         V_(mom_flux, j, k, l)) /
                                        V_(node_mass_post, j, k, l);
    }
#undef NMP
}

/* ------------------------------------------------------------------------ */
/* reset_field_kernel.f90:51-98 */
void oracle_reset_field_kernel(OR_DIMS, double *density0, const double *density1,
                               double *energy0, const double *energy1,
                               double *xvel0, const double *xvel1, double *yvel0, const double *yvel1,
                               double *zvel0, const double *zvel1)
{
    SHAPES;
#pragma omp parallel for
    for (int l = z_min; l <= z_max; l++)
        for (int k = y_min; k <= y_max; k++)
            for (int j = x_min; j <= x_max; j++) {
                C_(density0, j, k, l) = C_(density1, j, k, l);
                C_(energy0, j, k, l) = C_(energy1, j, k, l);
            }
#pragma omp parallel for
    for (int l = z_min; l <= z_max + 1; l++)
        for (int k = y_min; k <= y_max + 1; k++)
            for (int j = x_min; j <= x_max + 1; j++) {
                V_(xvel0, j, k, l) = V_(xvel1, j, k, l);
                V_(yvel0, j, k, l) = V_(yvel1, j, k, l);
                V_(zvel0, j, k, l) = V_(zvel1, j, k, l);
            }
}

/* ------------------------------------------------------------------------ */
/* Reflective boundary (update_halo_kernel.f90:115-1425), table driven.
 *
 * For each field and each external face the reference writes, for
 * d = 1..depth, halo(W(d)) = sign * f(R(d)) over a rectangle of the two
 * tangential axes whose upper end is max+depth (+1 where marked).  The
 * index pair (W, R) depends on how the field sits on the face's axis:
 *   CELL  cell-centred on the axis:      lo W=1-d  R=d      hi W=max+d    R=max+1-d
 *   STAG  staggered on the axis:         lo W=1-d  R=1+d    hi W=max+1+d  R=max+1-d
 *   FTAN  face data, tangential axis:    lo W=1-d  R=1+d    hi W=max+d    R=max-d
 * The lower index is the literal 1 of the reference (tiles are 1-based). */
enum { K_CELL = 0, K_STAG = 1, K_FTAN = 2 };

typedef struct {
    int kind[3];      /* per axis x, y, z */
    int neg[3];       /* sign flip on the faces normal to this axis */
    int ext[6][2];    /* per face (left,right,bottom,top,back,front): +1 on the upper end of
                         the first / second tangential axis (x,y order, skipping the normal) */
} halo_desc_t;

/* field order = FIELD_* ids 1..19 (data.f90:55-74) */
static const halo_desc_t HALO_DESC[19] = {
    /* density0  */ {{0, 0, 0}, {0, 0, 0}, {{0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}}},
    /* density1  */ {{0, 0, 0}, {0, 0, 0}, {{0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}}},
    /* energy0   */ {{0, 0, 0}, {0, 0, 0}, {{0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}}},
    /* energy1   */ {{0, 0, 0}, {0, 0, 0}, {{0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}}},
    /* pressure  */ {{0, 0, 0}, {0, 0, 0}, {{0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}}},
    /* viscosity */ {{0, 0, 0}, {0, 0, 0}, {{0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}}},
    /* soundspeed*/ {{0, 0, 0}, {0, 0, 0}, {{0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}}},
    /* xvel0     */ {{1, 1, 1}, {1, 0, 0}, {{1, 1}, {1, 1}, {1, 1}, {1, 1}, {1, 1}, {1, 1}}},
    /* xvel1     */ {{1, 1, 1}, {1, 0, 0}, {{1, 1}, {1, 1}, {1, 1}, {1, 1}, {1, 1}, {1, 1}}},
    /* yvel0     */ {{1, 1, 1}, {0, 1, 0}, {{1, 1}, {1, 1}, {1, 1}, {1, 1}, {1, 1}, {1, 1}}},
    /* yvel1     */ {{1, 1, 1}, {0, 1, 0}, {{1, 1}, {1, 1}, {1, 1}, {1, 1}, {1, 1}, {1, 1}}},
    /* zvel0     */ {{1, 1, 1}, {0, 0, 1}, {{1, 1}, {1, 1}, {1, 1}, {1, 1}, {1, 1}, {1, 1}}},
    /* zvel1     */ {{1, 1, 1}, {0, 0, 1}, {{1, 1}, {1, 1}, {1, 1}, {1, 1}, {1, 1}, {1, 1}}},
    /* vol_flux_x  (1018-1142)  left/right: (y,z)  bottom/top: (x,z)  back/front: (x,y) */
    /* vol_flux_x */ {{1, 2, 2}, {1, 0, 0}, {{0, 0}, {0, 0}, {1, 0}, {1, 0}, {0, 0}, {0, 0}}},
    /* vol_flux_y (1156-1280) */ {{2, 1, 2}, {0, 1, 0}, {{1, 0}, {1, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}}},
    /* vol_flux_z (1291-1350): bottom l to z_max+d, top l to z_max+1+d */
    /* vol_flux_z */ {{2, 2, 1}, {0, 0, 1}, {{0, 1}, {0, 1}, {0, 0}, {0, 1}, {0, 0}, {0, 0}}},
    /* mass_flux_x */ {{1, 2, 2}, {1, 0, 0}, {{0, 0}, {0, 0}, {1, 0}, {1, 0}, {0, 0}, {0, 0}}},
    /* mass_flux_y */ {{2, 1, 2}, {0, 1, 0}, {{1, 0}, {1, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}}},
    /* mass_flux_z */ {{2, 2, 1}, {0, 0, 1}, {{0, 1}, {0, 1}, {0, 1}, {0, 1}, {0, 0}, {0, 0}}},
};

/* Array class extents (+0/+1 on x, y) per field id, for the strides. */
static const int FIELD_EX[19] = {0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1, 1, 1, 0, 0, 1, 0, 0};
static const int FIELD_EY[19] = {0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1, 1, 0, 1, 0, 0, 1, 0};

static void halo_face(OR_DIMS, double *a, int fid, int face, int depth)
{
    const halo_desc_t *h = &HALO_DESC[fid];
    const shape_t s = mk_shape(x_min, x_max, y_min, y_max, z_min, FIELD_EX[fid], FIELD_EY[fid]);
    const int axis = (face - 1) / 2;  /* 0 x, 1 y, 2 z */
    const int hi = (face - 1) % 2;    /* 0 low face, 1 high face */
    /* face order in the table: left,right,bottom,top,back,front = ids 1..6 */
    const int *ext = h->ext[face - 1];
    const int kind = h->kind[axis];
    const double sgn = h->neg[axis] ? -1.0 : 1.0;
    const int amax = axis == 0 ? x_max : axis == 1 ? y_max : z_max;
    /* tangential axes in (x,y,z) order skipping `axis` */
    const int ta = axis == 0 ? 1 : 0;
    const int tb = axis == 2 ? 1 : 2;
    const int tmin[3] = {x_min, y_min, z_min};
    const int tmax[3] = {x_max, y_max, z_max};
    for (int b = tmin[tb] - depth; b <= tmax[tb] + ext[1] + depth; b++)
        for (int c = tmin[ta] - depth; c <= tmax[ta] + ext[0] + depth; c++)
            for (int d = 1; d <= depth; d++) {
                int w, r;
                if (!hi) {
                    w = 1 - d;
                    r = (kind == K_CELL) ? d : 1 + d;
                } else if (kind == K_CELL) {
                    w = amax + d;
                    r = amax + 1 - d;
                } else if (kind == K_STAG) {
                    w = amax + 1 + d;
                    r = amax + 1 - d;
                } else {
                    w = amax + d;
                    r = amax - d;
                }
                int iw[3], ir[3];
                iw[axis] = w;
                ir[axis] = r;
                iw[ta] = ir[ta] = c;
                iw[tb] = ir[tb] = b;
                a[OFF(s, iw[0], iw[1], iw[2])] = sgn * a[OFF(s, ir[0], ir[1], ir[2])];
            }
}

void oracle_update_halo_kernel(OR_DIMS, const int *chunk_neighbours, const int *tile_neighbours,
                               double *density0, double *energy0, double *pressure, double *viscosity,
                               double *soundspeed, double *density1, double *energy1,
                               double *xvel0, double *yvel0, double *zvel0,
                               double *xvel1, double *yvel1, double *zvel1,
                               double *vol_flux_x, double *vol_flux_y, double *vol_flux_z,
                               double *mass_flux_x, double *mass_flux_y, double *mass_flux_z,
                               const int *fields, int depth)
{
    double *arr[19] = {density0, density1, energy0, energy1, pressure, viscosity, soundspeed,
                       xvel0, xvel1, yvel0, yvel1, zvel0, zvel1,
                       vol_flux_x, vol_flux_y, vol_flux_z, mass_flux_x, mass_flux_y, mass_flux_z};
    /* per field: bottom, top, left, right, back, front (face ids 3,4,1,2,5,6) */
    static const int order[6] = {3, 4, 1, 2, 5, 6};
    for (int f = 0; f < 19; f++) {
        if (fields[f] != 1) continue;
        for (int o = 0; o < 6; o++) {
            int face = order[o];
            if (chunk_neighbours[face - 1] == -1 && tile_neighbours[face - 1] == -1)
                halo_face(x_min, x_max, y_min, y_max, z_min, z_max, arr[f], f, face, depth);
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Tile-to-tile halo copies (update_tile_halo_kernel.f90).  Along the copy
 * axis a field is either cell-like or staggered (vertex data and the face
 * data normal to that axis).  Tangential ranges run min-depth..max+depth,
 * +1 on axes where the array is staggered.
 *   left   f(min-d)   = L(Lmax+1-d)             (all data)
 *   right  f(max+d)   = R(Rmin-1+d)   cell-like;  f(max+1+d) = R(Rmin+d) staggered
 *   bottom f(min-d)   = B(Bmax+1-d)             (all data)
 *   top    f(max+d)   = T(Tmin-1+d)   cell-like;  f(max+1+d) = T(Tmin+d) staggered
 *   back   f(min-d)   = K(Kmax+1-d)             (all data)
 *   front  f(max+d)   = F(Fmin-1+d)   cell-like;  f(max+1+d) = F(Fmin+d) staggered
 * Quirks kept: energy1 is gated by the density1 flag on the left, right,
 * bottom and top faces (e.g. 113-125, 443-455, 771-783, 1103-1115); on the
 * right face the zvel0 flag copies yvel0 (579-589). */

/* tile-halo argument order -> field id (data.f90:55-74) */
static const int TILE_ARG_FID[19] = {0, 2, 4, 5, 6, 1, 3, 7, 9, 11, 8, 10, 12, 13, 14, 15, 16, 17, 18};
/* staggered on x, y, z per field id */
static const int STAG_AX[19][3] = {
    {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0},
    {1, 1, 1}, {1, 1, 1}, {1, 1, 1}, {1, 1, 1}, {1, 1, 1}, {1, 1, 1},
    {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};

void oracle_update_tile_halo_face(int face, OR_DIMS, double *const *arr,
                                  int n_xmin, int n_xmax, int n_ymin, int n_ymax, int n_zmin, int n_zmax,
                                  double *const *nbr, const int *fields, int depth)
{
    (void)n_zmax;
    double *byfid[19], *nbyfid[19];
    for (int a = 0; a < 19; a++) {
        byfid[TILE_ARG_FID[a]] = arr[a];
        nbyfid[TILE_ARG_FID[a]] = nbr[a];
    }
    const int axis = (face - 1) / 2;
    const int hi = (face - 1) % 2;
    const int ta = axis == 0 ? 1 : 0;
    const int tb = axis == 2 ? 1 : 2;
    const int tmin[3] = {x_min, y_min, z_min};
    const int tmax[3] = {x_max, y_max, z_max};
    const int nmin[3] = {n_xmin, n_ymin, n_zmin};
    const int nmax[3] = {n_xmax, n_ymax, n_zmax};
    for (int f = 0; f < 19; f++) {
        int gate = f;
        int src = f, dst = f;
        if (f == 3 && face <= 4) gate = 1;          /* energy1 gated by density1 */
        if (f == 11 && face == 2) src = dst = 9;    /* right face: zvel0 flag copies yvel0 */
        if (fields[gate] != 1) continue;
        const shape_t s = mk_shape(x_min, x_max, y_min, y_max, z_min, FIELD_EX[dst], FIELD_EY[dst]);
        const shape_t sn = mk_shape(n_xmin, n_xmax, n_ymin, n_ymax, n_zmin, FIELD_EX[dst], FIELD_EY[dst]);
        const int stag = STAG_AX[dst][axis];
        double *a = byfid[dst];
        const double *n = nbyfid[src];
        for (int b = tmin[tb] - depth; b <= tmax[tb] + STAG_AX[dst][tb] + depth; b++)
            for (int c = tmin[ta] - depth; c <= tmax[ta] + STAG_AX[dst][ta] + depth; c++)
                for (int d = 1; d <= depth; d++) {
                    int w, r;
                    if (!hi) {
                        w = tmin[axis] - d;
                        r = nmax[axis] + 1 - d;
                    } else if (!stag) {
                        w = tmax[axis] + d;
                        r = nmin[axis] - 1 + d;
                    } else {
                        w = tmax[axis] + 1 + d;
                        r = nmin[axis] + d;
                    }
                    int iw[3], ir[3];
                    iw[axis] = w;
                    ir[axis] = r;
                    iw[ta] = ir[ta] = c;
                    iw[tb] = ir[tb] = b;
                    a[OFF(s, iw[0], iw[1], iw[2])] = n[OFF(sn, ir[0], ir[1], ir[2])];
                }
    }
}

/* ------------------------------------------------------------------------ */
/* field_summary_kernel.f90:51-78, summed serially in l,k,j order */
void oracle_field_summary_kernel(OR_DIMS, const double *volume, const double *density0,
                                 const double *energy0, const double *pressure,
                                 const double *xvel0, const double *yvel0, const double *zvel0,
                                 double *vol, double *mass, double *ie, double *ke, double *press)
{
    SHAPES;
    double svol = 0.0, smass = 0.0, sie = 0.0, ske = 0.0, spress = 0.0;
    for (int l = z_min; l <= z_max; l++)
        for (int k = y_min; k <= y_max; k++)
            for (int j = x_min; j <= x_max; j++) {
                double vsqrd = 0.0;
                for (int lv = l; lv <= l + 1; lv++)
                    for (int kv = k; kv <= k + 1; kv++)
                        for (int jv = j; jv <= j + 1; jv++)
                            vsqrd = vsqrd + 0.125 * (V_(xvel0, jv, kv, lv) * V_(xvel0, jv, kv, lv) +
                                                     V_(yvel0, jv, kv, lv) * V_(yvel0, jv, kv, lv) +
                                                     V_(zvel0, jv, kv, lv) * V_(zvel0, jv, kv, lv));
                double cell_vol = C_(volume, j, k, l);
                double cell_mass = cell_vol * C_(density0, j, k, l);
                svol = svol + cell_vol;
                smass = smass + cell_mass;
                sie = sie + cell_mass * C_(energy0, j, k, l);
                ske = ske + cell_mass * 0.5 * vsqrd;
                spress = spress + cell_vol * C_(pressure, j, k, l);
            }
    *vol = svol;
    *mass = smass;
    *ie = sie;
    *ke = ske;
    *press = spress;
}

/* ------------------------------------------------------------------------ */
/* initialise_chunk_kernel.f90:63-178; float(j-x_min) is a default-real
 * conversion, exact for these integers. */
void oracle_initialise_chunk_kernel(OR_DIMS, double xmin, double ymin, double zmin,
                                    double dx, double dy, double dz,
                                    double *vertexx, double *vertexdx, double *vertexy, double *vertexdy,
                                    double *vertexz, double *vertexdz,
                                    double *cellx, double *celldx, double *celly, double *celldy,
                                    double *cellz, double *celldz,
                                    double *volume, double *xarea, double *yarea, double *zarea)
{
    SHAPES;
    for (int j = x_min - 2; j <= x_max + 3; j++) {
        D1X(vertexx, j) = xmin + dx * (double)(float)(j - x_min);
        D1X(vertexdx, j) = dx;
    }
    for (int k = y_min - 2; k <= y_max + 3; k++) {
        D1Y(vertexy, k) = ymin + dy * (double)(float)(k - y_min);
        D1Y(vertexdy, k) = dy;
    }
    for (int l = z_min - 2; l <= z_max + 3; l++) {
        D1Z(vertexz, l) = zmin + dz * (double)(float)(l - z_min);
        D1Z(vertexdz, l) = dz;
    }
    for (int j = x_min - 2; j <= x_max + 2; j++) {
        D1X(cellx, j) = 0.5 * (D1X(vertexx, j) + D1X(vertexx, j + 1));
        D1X(celldx, j) = dx;
    }
    for (int k = y_min - 2; k <= y_max + 2; k++) {
        D1Y(celly, k) = 0.5 * (D1Y(vertexy, k) + D1Y(vertexy, k + 1));
        D1Y(celldy, k) = dy;
    }
    for (int l = z_min - 2; l <= z_max + 2; l++) {
        D1Z(cellz, l) = 0.5 * (D1Z(vertexz, l) + D1Z(vertexz, l + 1));
        D1Z(celldz, l) = dz;
    }
    for (int l = z_min - 2; l <= z_max + 2; l++)
        for (int k = y_min - 2; k <= y_max + 2; k++)
            for (int j = x_min - 2; j <= x_max + 2; j++) {
                C_(volume, j, k, l) = dx * dy * dz;
                X_(xarea, j, k, l) = D1Y(celldy, k) * D1Z(celldz, l);
                Y_(yarea, j, k, l) = D1X(celldx, j) * D1Z(celldz, l);
                Z_(zarea, j, k, l) = D1X(celldx, j) * D1Y(celldy, k);
            }
}

/* ------------------------------------------------------------------------ */
/* generate_chunk_kernel.f90:82-207 (g_rect=1, g_circ=2, g_point=3, data.f90:90-92) */
void oracle_generate_chunk_kernel(OR_DIMS,
                                  const double *vertexx, const double *vertexy, const double *vertexz,
                                  const double *cellx, const double *celly, const double *cellz,
                                  double *density0, double *energy0,
                                  double *xvel0, double *yvel0, double *zvel0,
                                  int number_of_states,
                                  const double *state_density, const double *state_energy,
                                  const double *state_xvel, const double *state_yvel, const double *state_zvel,
                                  const double *state_xmin, const double *state_xmax,
                                  const double *state_ymin, const double *state_ymax,
                                  const double *state_zmin, const double *state_zmax,
                                  const double *state_radius, const int *state_geometry)
{
    SHAPES;
    for (int l = z_min - 2; l <= z_max + 2; l++)
        for (int k = y_min - 2; k <= y_max + 2; k++)
            for (int j = x_min - 2; j <= x_max + 2; j++) {
                C_(energy0, j, k, l) = state_energy[0];
                C_(density0, j, k, l) = state_density[0];
                V_(xvel0, j, k, l) = state_xvel[0];
                V_(yvel0, j, k, l) = state_yvel[0];
                V_(zvel0, j, k, l) = state_zvel[0];
            }
    for (int s = 1; s < number_of_states; s++) {
        double x_cent = state_xmin[s], y_cent = state_ymin[s], z_cent = state_zmin[s];
        for (int l = z_min - 2; l <= z_max + 2; l++)
            for (int k = y_min - 2; k <= y_max + 2; k++)
                for (int j = x_min - 2; j <= x_max + 2; j++) {
                    int hit = 0;
                    if (state_geometry[s] == 1) {
                        hit = D1X(vertexx, j + 1) >= state_xmin[s] && D1X(vertexx, j) < state_xmax[s] &&
                              D1Y(vertexy, k + 1) >= state_ymin[s] && D1Y(vertexy, k) < state_ymax[s] &&
                              D1Z(vertexz, l + 1) >= state_zmin[s] && D1Z(vertexz, l) < state_zmax[s];
                    } else if (state_geometry[s] == 2) {
                        double radius = sqrt((D1X(cellx, j) - x_cent) * (D1X(cellx, j) - x_cent) +
                                             (D1Y(celly, k) - y_cent) * (D1Y(celly, k) - y_cent) +
                                             (D1Z(cellz, l) - z_cent) * (D1Z(cellz, l) - z_cent));
                        hit = radius <= state_radius[s];
                    } else if (state_geometry[s] == 3) {
                        hit = D1X(vertexx, j) == x_cent && D1Y(vertexy, k) == y_cent && D1Z(vertexz, l) == z_cent;
                    }
                    if (!hit) continue;
                    C_(energy0, j, k, l) = state_energy[s];
                    C_(density0, j, k, l) = state_density[s];
                    for (int lt = l; lt <= l + 1; lt++)
                        for (int kt = k; kt <= k + 1; kt++)
                            for (int jt = j; jt <= j + 1; jt++) {
                                V_(xvel0, jt, kt, lt) = state_xvel[s];
                                V_(yvel0, jt, kt, lt) = state_yvel[s];
                                V_(zvel0, jt, kt, lt) = state_zvel[s];
                            }
                }
    }
}

/* ------------------------------------------------------------------------ */
/* The six per-face entry points with the reference argument lists
 * (update_tile_halo_kernel.f90:31, 361, 690, 1022, 1351, 1674). */
#define TILE_ARGS                                                                                              \
    OR_DIMS, double *a0, double *a1, double *a2, double *a3, double *a4, double *a5, double *a6, double *a7,     \
        double *a8, double *a9, double *a10, double *a11, double *a12, double *a13, double *a14, double *a15,  \
        double *a16, double *a17, double *a18, int n_xmin, int n_xmax, int n_ymin, int n_ymax, int n_zmin,     \
        int n_zmax, double *b0, double *b1, double *b2, double *b3, double *b4, double *b5, double *b6,        \
        double *b7, double *b8, double *b9, double *b10, double *b11, double *b12, double *b13, double *b14,   \
        double *b15, double *b16, double *b17, double *b18, const int *fields, int depth
#define TILE_FWD(face)                                                                                         \
    {                                                                                                          \
        double *a[19] = {a0, a1, a2, a3, a4, a5, a6, a7, a8, a9, a10, a11, a12, a13, a14, a15, a16, a17, a18}; \
        double *b[19] = {b0, b1, b2, b3, b4, b5, b6, b7, b8, b9, b10, b11, b12, b13, b14, b15, b16, b17, b18}; \
        oracle_update_tile_halo_face(face, x_min, x_max, y_min, y_max, z_min, z_max, a, n_xmin, n_xmax,      \
                                     n_ymin, n_ymax, n_zmin, n_zmax, b, fields, depth);                        \
    }
void oracle_update_tile_halo_left_kernel(TILE_ARGS) TILE_FWD(1)
void oracle_update_tile_halo_right_kernel(TILE_ARGS) TILE_FWD(2)
void oracle_update_tile_halo_bottom_kernel(TILE_ARGS) TILE_FWD(3)
void oracle_update_tile_halo_top_kernel(TILE_ARGS) TILE_FWD(4)
void oracle_update_tile_halo_back_kernel(TILE_ARGS) TILE_FWD(5)
void oracle_update_tile_halo_front_kernel(TILE_ARGS) TILE_FWD(6)
