/*
 * Whole-run driver of the oracle: start (start.f90:25-144) and the hydro
 * loop (hydro.f90:46-87) with the phase drivers timestep.f90, PdV.f90,
 * advection.f90, update_halo.f90, update_tile_halo.f90 and
 * field_summary.f90, restated in C.
 *
 * TEST INFRASTRUCTURE ONLY (see clover_oracle.h).  The kernels it calls come
 * from a table: either the C restatement in clover_oracle.c, or the
 * reference's own Fortran kernels compiled into oracle/_ref/libclover_ref.so
 * (loaded with dlopen, symbols ref_<kernel>).
 */
#include "clover_oracle.h"

#include <dlfcn.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

typedef void (*fn_ideal_gas)(OR_DIMS, const double *, const double *, double *, double *);
typedef void (*fn_viscosity)(OR_DIMS, const double *, const double *, const double *, const double *,
                             const double *, const double *, const double *, const double *, double *,
                             const double *, const double *, const double *);
typedef void (*fn_calc_dt)(OR_DIMS, double, double, double, double, double, double, double, double,
                           const double *, const double *, const double *, const double *, const double *,
                           const double *, const double *, const double *, const double *, const double *,
                           const double *, const double *, const double *, const double *, const double *,
                           const double *, const double *, const double *, double *, double *, int *, double *,
                           double *, double *, int *, int *, int *, int *);
typedef void (*fn_pdv)(int, OR_DIMS, double, const double *, const double *, const double *, const double *,
                       const double *, double *, const double *, double *, const double *, const double *,
                       const double *, const double *, const double *, const double *, const double *,
                       const double *);
typedef void (*fn_revert)(OR_DIMS, const double *, double *, const double *, double *);
typedef void (*fn_accelerate)(OR_DIMS, double, const double *, const double *, const double *, const double *,
                              const double *, const double *, const double *, const double *, const double *,
                              const double *, double *, double *, double *);
typedef void (*fn_flux_calc)(OR_DIMS, double, const double *, const double *, const double *, const double *,
                             const double *, const double *, const double *, const double *, const double *,
                             double *, double *, double *);
typedef void (*fn_advec_cell)(OR_DIMS, int, int, int, const double *, const double *, const double *,
                              const double *, double *, double *, double *, const double *, double *,
                              const double *, double *, const double *, double *, double *, double *, double *,
                              double *, double *, double *);
typedef void (*fn_advec_mom)(OR_DIMS, double *, const double *, const double *, const double *, const double *,
                             const double *, const double *, const double *, const double *, double *, double *,
                             double *, double *, double *, double *, double *, const double *, const double *,
                             const double *, int, int, int, int);
typedef void (*fn_reset)(OR_DIMS, double *, const double *, double *, const double *, double *, const double *,
                         double *, const double *, double *, const double *);
typedef void (*fn_update_halo)(OR_DIMS, const int *, const int *, double *, double *, double *, double *, double *,
                               double *, double *, double *, double *, double *, double *, double *, double *,
                               double *, double *, double *, double *, double *, double *, const int *, int);
typedef void (*fn_tile_halo)(int, OR_DIMS, double *const *, int, int, int, int, int, int, double *const *,
                             const int *, int);
typedef void (*fn_summary)(OR_DIMS, const double *, const double *, const double *, const double *,
                           const double *, const double *, const double *, double *, double *, double *, double *,
                           double *);
typedef void (*fn_init_chunk)(OR_DIMS, double, double, double, double, double, double, double *, double *,
                              double *, double *, double *, double *, double *, double *, double *, double *,
                              double *, double *, double *, double *, double *, double *);
typedef void (*fn_gen_chunk)(OR_DIMS, const double *, const double *, const double *, const double *,
                             const double *, const double *, double *, double *, double *, double *, double *, int,
                             const double *, const double *, const double *, const double *, const double *,
                             const double *, const double *, const double *, const double *, const double *,
                             const double *, const double *, const int *);

typedef struct {
    fn_ideal_gas ideal_gas;
    fn_viscosity viscosity;
    fn_calc_dt calc_dt;
    fn_pdv pdv;
    fn_revert revert;
    fn_accelerate accelerate;
    fn_flux_calc flux_calc;
    fn_advec_cell advec_cell;
    fn_advec_mom advec_mom;
    fn_reset reset;
    fn_update_halo update_halo;
    fn_tile_halo tile_halo;
    fn_summary summary;
    fn_init_chunk init_chunk;
    fn_gen_chunk gen_chunk;
} ktab;

enum { F_D0, F_D1, F_E0, F_E1, F_P, F_Q, F_C, F_U0, F_U1, F_V0, F_V1, F_W0, F_W1, F_VFX, F_VFY, F_VFZ, F_MFX, F_MFY, F_MFZ };

typedef struct {
    int nx, ny, nz;
    int left, bottom, back;
    int nbr[6];
    double *d0, *d1, *e0, *e1, *p, *q, *c, *vol;
    double *u0, *u1, *v0, *v1, *w0, *w1;
    double *w[8]; /* work_array1..7 at w[1..7] */
    double *vfx, *mfx, *xarea, *vfy, *mfy, *yarea, *vfz, *mfz, *zarea;
    double *cellx, *celly, *cellz, *celldx, *celldy, *celldz;
    double *vertexx, *vertexy, *vertexz, *vertexdx, *vertexdy, *vertexdz;
} otile;

#define TD(t) 1, (t)->nx, 1, (t)->ny, 1, (t)->nz

static double *zalloc(size_t n)
{
    double *a = (double *)calloc(n, sizeof(double));
    if (!a) {
        fprintf(stderr, "oracle: out of memory\n");
        exit(1);
    }
    return a;
}

/* build_field.f90:41-209: allocate and zero every array of a tile */
static void build_field(otile *t)
{
    size_t cx = t->nx + 4, cy = t->ny + 4, cz = t->nz + 4;
    size_t vx = t->nx + 5, vy = t->ny + 5, vz = t->nz + 5;
    size_t nc = cx * cy * cz, nv = vx * vy * vz;
    t->d0 = zalloc(nc); t->d1 = zalloc(nc); t->e0 = zalloc(nc); t->e1 = zalloc(nc);
    t->p = zalloc(nc); t->q = zalloc(nc); t->c = zalloc(nc); t->vol = zalloc(nc);
    t->u0 = zalloc(nv); t->u1 = zalloc(nv); t->v0 = zalloc(nv); t->v1 = zalloc(nv);
    t->w0 = zalloc(nv); t->w1 = zalloc(nv);
    for (int i = 1; i <= 7; i++) t->w[i] = zalloc(nv);
    t->vfx = zalloc(vx * cy * cz); t->mfx = zalloc(vx * cy * cz); t->xarea = zalloc(vx * cy * cz);
    t->vfy = zalloc(cx * vy * cz); t->mfy = zalloc(cx * vy * cz); t->yarea = zalloc(cx * vy * cz);
    t->vfz = zalloc(cx * cy * vz); t->mfz = zalloc(cx * cy * vz); t->zarea = zalloc(cx * cy * vz);
    t->cellx = zalloc(cx); t->celly = zalloc(cy); t->cellz = zalloc(cz);
    t->celldx = zalloc(cx); t->celldy = zalloc(cy); t->celldz = zalloc(cz);
    t->vertexx = zalloc(vx); t->vertexy = zalloc(vy); t->vertexz = zalloc(vz);
    t->vertexdx = zalloc(vx); t->vertexdy = zalloc(vy); t->vertexdz = zalloc(vz);
}

static void free_field(otile *t)
{
    double *all[] = {t->d0, t->d1, t->e0, t->e1, t->p, t->q, t->c, t->vol, t->u0, t->u1, t->v0, t->v1, t->w0,
                     t->w1, t->vfx, t->mfx, t->xarea, t->vfy, t->mfy, t->yarea, t->vfz, t->mfz, t->zarea,
                     t->cellx, t->celly, t->cellz, t->celldx, t->celldy, t->celldz, t->vertexx, t->vertexy,
                     t->vertexz, t->vertexdx, t->vertexdy, t->vertexdz};
    for (size_t i = 0; i < sizeof(all) / sizeof(all[0]); i++) free(all[i]);
    for (int i = 1; i <= 7; i++) free(t->w[i]);
}

/* clover.f90:282-285 and 323-326: surface and volume are default-REAL
 * (single-precision) expressions, `1.0*x_cells` being a REAL(4) product;
 * only their values are stored in the REAL(8) variables, and the metric is
 * the REAL(8) quotient.  C's float arithmetic (SSE, no excess precision)
 * evaluates the same operations in the same order. */
static double tile_metric(int x_cells, int y_cells, int z_cells, int cx, int cy, int cz)
{
    float surface = (((1.0f * (float)x_cells) / (float)cx) * ((1.0f * (float)y_cells) / (float)cy) * 2.0f) +
                    (((1.0f * (float)x_cells) / (float)cx) * ((1.0f * (float)z_cells) / (float)cz) * 2.0f) +
                    (((1.0f * (float)y_cells) / (float)cy) * ((1.0f * (float)z_cells) / (float)cz) * 2.0f);
    float volume = ((1.0f * (float)x_cells) / (float)cx) * ((1.0f * (float)y_cells) / (float)cy) *
                   ((1.0f * (float)z_cells) / (float)cz);
    return (double)surface / (double)volume;
}

/* clover.f90:258-447 */
int oracle_decompose_tiles(int x_cells, int y_cells, int z_cells, int tiles_per_chunk, int tile_dims,
                           int *t_nx, int *t_ny, int *t_nz, int *t_left, int *t_bottom, int *t_back,
                           int *neighbours)
{
    int current_x = 1, current_y = 1, current_z = tiles_per_chunk;
    double best_metric = tile_metric(x_cells, y_cells, z_cells, current_x, current_y, current_z);
    int chunk_x = current_x, chunk_y = current_y, chunk_z = current_z;
    int x_bound = tiles_per_chunk, y_bound = tiles_per_chunk;
    if (tile_dims == 2) { x_bound = 1; y_bound = tiles_per_chunk; }
    if (tile_dims == 1) { x_bound = 1; y_bound = 1; }
    for (int t = 1; t <= x_bound; t++) {
        if (tiles_per_chunk % t != 0) continue;
        current_x = t;
        int div1 = tiles_per_chunk / t;
        if (y_bound < div1) div1 = y_bound;
        for (int j = 1; j <= div1; j++) {
            if (div1 % j != 0) continue;
            current_y = j;
            if (tiles_per_chunk % (t * j) != 0) continue;
            current_z = tiles_per_chunk / (t * j);
            double current_metric = tile_metric(x_cells, y_cells, z_cells, current_x, current_y, current_z);
            if (current_metric < best_metric) {
                chunk_x = current_x; chunk_y = current_y; chunk_z = current_z;
                best_metric = current_metric;
            }
        }
    }
    int delta_x = x_cells / chunk_x, delta_y = y_cells / chunk_y, delta_z = z_cells / chunk_z;
    int mod_x = x_cells % chunk_x, mod_y = y_cells % chunk_y, mod_z = z_cells % chunk_z;
    int add_x_prev = 0, add_y_prev = 0, add_z_prev = 0;
    int tile = 0;
    for (int cz = 1; cz <= chunk_z; cz++) {
        for (int cy = 1; cy <= chunk_y; cy++) {
            for (int cx = 1; cx <= chunk_x; cx++) {
                int add_x = cx <= mod_x, add_y = cy <= mod_y, add_z = cz <= mod_z;
                int left = (cx - 1) * delta_x + 1 + add_x_prev;
                int right = left + delta_x - 1 + add_x;
                int bottom = (cy - 1) * delta_y + 1 + add_y_prev;
                int top = bottom + delta_y - 1 + add_y;
                int back = (cz - 1) * delta_z + 1 + add_z_prev;
                int front = back + delta_z - 1 + add_z;
                int id = tile + 1;
                int *nb = neighbours + 6 * tile;
                nb[0] = cx == 1 ? -1 : id - 1;
                nb[1] = cx == chunk_x ? -1 : id + 1;
                nb[2] = cy == 1 ? -1 : id - chunk_x;
                nb[3] = cy == chunk_y ? -1 : id + chunk_x;
                nb[4] = cz == 1 ? -1 : id - chunk_x * chunk_y;
                nb[5] = cz == chunk_z ? -1 : id + chunk_x * chunk_y;
                t_nx[tile] = right - left + 1;
                t_ny[tile] = top - bottom + 1;
                t_nz[tile] = front - back + 1;
                t_left[tile] = left;
                t_bottom[tile] = bottom;
                t_back[tile] = back;
                if (cx <= mod_x) add_x_prev++;
                tile++;
            }
            add_x_prev = 0;
            if (cy <= mod_y) add_y_prev++;
        }
        add_x_prev = 0;
        add_y_prev = 0;
        if (cz <= mod_z) add_z_prev++;
    }
    return tile;
}

typedef struct {
    const oracle_deck *deck;
    const ktab *k;
    int ntiles;
    otile *tiles;
    double dt, dtold, time;
    int step, advect_x;
    double g_small, g_big;
} ostate;

static const int EXT6[6] = {-1, -1, -1, -1, -1, -1}; /* one chunk: every chunk face is external */

static double *const *tile_arrays(otile *t, double **buf)
{
    /* update_tile_halo argument order */
    double *a[19] = {t->d0, t->e0, t->p, t->q, t->c, t->d1, t->e1, t->u0, t->v0, t->w0, t->u1, t->v1, t->w1,
                     t->vfx, t->vfy, t->vfz, t->mfx, t->mfy, t->mfz};
    memcpy(buf, a, sizeof(a));
    return buf;
}

/* update_tile_halo.f90:47-423: top/bottom for all tiles, then left/right,
 * then front/back. */
static void update_tile_halo(ostate *s, const int *fields, int depth)
{
    static const int passes[3][2] = {{4, 3}, {1, 2}, {6, 5}};
    for (int pass = 0; pass < 3; pass++)
        for (int ti = 0; ti < s->ntiles; ti++)
            for (int f = 0; f < 2; f++) {
                int face = passes[pass][f];
                otile *t = &s->tiles[ti];
                int n = t->nbr[face - 1];
                if (n == -1) continue;
                otile *nt = &s->tiles[n - 1];
                double *a[19], *b[19];
                tile_arrays(t, a);
                tile_arrays(nt, b);
                s->k->tile_halo(face, TD(t), a, TD(nt), b, fields, depth);
            }
}

/* update_halo.f90:27-88 */
static void update_halo(ostate *s, const int *fields, int depth)
{
    update_tile_halo(s, fields, depth);
    for (int ti = 0; ti < s->ntiles; ti++) {
        otile *t = &s->tiles[ti];
        s->k->update_halo(TD(t), EXT6, t->nbr, t->d0, t->e0, t->p, t->q, t->c, t->d1, t->e1, t->u0, t->v0, t->w0,
                          t->u1, t->v1, t->w1, t->vfx, t->vfy, t->vfz, t->mfx, t->mfy, t->mfz, fields, depth);
    }
}

static void set_fields(int *fields, const int *ids, int n)
{
    memset(fields, 0, 19 * sizeof(int));
    for (int i = 0; i < n; i++) fields[ids[i]] = 1;
}

static void ideal_gas(ostate *s, int predict)
{
    for (int ti = 0; ti < s->ntiles; ti++) {
        otile *t = &s->tiles[ti];
        if (!predict) s->k->ideal_gas(TD(t), t->d0, t->e0, t->p, t->c);
        else s->k->ideal_gas(TD(t), t->d1, t->e1, t->p, t->c);
    }
}

/* field_summary.f90:27-139 (tile sums added in tile order) */
static void field_summary(ostate *s, oracle_summary *row)
{
    ideal_gas(s, 0);
    double tv = 0, tm = 0, ti_ = 0, tk = 0, tp = 0;
    for (int ti = 0; ti < s->ntiles; ti++) {
        otile *t = &s->tiles[ti];
        double vol, mass, ie, ke, press;
        s->k->summary(TD(t), t->vol, t->d0, t->e0, t->p, t->u0, t->v0, t->w0, &vol, &mass, &ie, &ke, &press);
        tv = tv + vol; tm = tm + mass; ti_ = ti_ + ie; tk = tk + ke; tp = tp + press;
    }
    row->step = s->step;
    row->time = s->time;
    row->vol = tv;
    row->mass = tm;
    row->ie = ti_;
    row->ke = tk;
    row->press = tp;
}

/* timestep.f90:28-128; returns 1 on a small time step */
static int timestep(ostate *s)
{
    const oracle_deck *dk = s->deck;
    int fields[19];
    s->dt = s->g_big;
    ideal_gas(s, 0);
    static const int f1[] = {F_P, F_E0, F_D0, F_U0, F_V0};
    set_fields(fields, f1, 5);
    update_halo(s, fields, 1);
    for (int ti = 0; ti < s->ntiles; ti++) {
        otile *t = &s->tiles[ti];
        s->k->viscosity(TD(t), t->xarea, t->yarea, t->zarea, t->celldx, t->celldy, t->celldz, t->d0, t->p, t->q,
                        t->u0, t->v0, t->w0);
    }
    static const int f2[] = {F_Q};
    set_fields(fields, f2, 1);
    update_halo(s, fields, 1);
    for (int ti = 0; ti < s->ntiles; ti++) {
        otile *t = &s->tiles[ti];
        double dtlp, xl, yl, zl;
        int ctl, jl, kl, ll, small;
        s->k->calc_dt(TD(t), s->g_small, s->g_big, dk->dtmin, dk->dtc_safe, dk->dtu_safe, dk->dtv_safe,
                      dk->dtw_safe, dk->dtdiv_safe, t->xarea, t->yarea, t->zarea, t->cellx, t->celly, t->cellz,
                      t->celldx, t->celldy, t->celldz, t->vol, t->d0, t->e0, t->p, t->q, t->c, t->u0, t->v0, t->w0,
                      t->w[1], &dtlp, &ctl, &xl, &yl, &zl, &jl, &kl, &ll, &small);
        if (dtlp <= s->dt) s->dt = dtlp;
    }
    double m = s->dt;
    if (s->dtold * dk->dtrise < m) m = s->dtold * dk->dtrise;
    if (dk->dtmax < m) m = dk->dtmax;
    s->dt = m;
    int small = s->dt < dk->dtmin;
    s->dtold = s->dt;
    return small;
}

/* PdV.f90:26-109 */
static void pdv(ostate *s, int predict)
{
    for (int ti = 0; ti < s->ntiles; ti++) {
        otile *t = &s->tiles[ti];
        s->k->pdv(predict, TD(t), s->dt, t->xarea, t->yarea, t->zarea, t->vol, t->d0, t->d1, t->e0, t->e1, t->p,
                  t->q, t->u0, t->u1, t->v0, t->v1, t->w0, t->w1);
    }
    if (predict) {
        int fields[19];
        ideal_gas(s, 1);
        static const int f[] = {F_P};
        set_fields(fields, f, 1);
        update_halo(s, fields, 1);
        for (int ti = 0; ti < s->ntiles; ti++) {
            otile *t = &s->tiles[ti];
            s->k->revert(TD(t), t->d0, t->d1, t->e0, t->e1);
        }
    }
}

static void advec_cell(ostate *s, int sweep, int dir)
{
    for (int ti = 0; ti < s->ntiles; ti++) {
        otile *t = &s->tiles[ti];
        s->k->advec_cell(TD(t), s->advect_x, dir, sweep, t->vertexdx, t->vertexdy, t->vertexdz, t->vol, t->d1, t->e1,
                         t->mfx, t->vfx, t->mfy, t->vfy, t->mfz, t->vfz, t->w[1], t->w[2], t->w[3], t->w[4], t->w[5],
                         t->w[6], t->w[7]);
    }
}

static void advec_mom(ostate *s, int sweep, int dir)
{
    for (int ti = 0; ti < s->ntiles; ti++) {
        otile *t = &s->tiles[ti];
        double *vel[3] = {t->u1, t->v1, t->w1};
        for (int which = 1; which <= 3; which++)
            s->k->advec_mom(TD(t), vel[which - 1], t->mfx, t->vfx, t->mfy, t->vfy, t->mfz, t->vfz, t->vol, t->d1,
                            t->w[1], t->w[2], t->w[3], t->w[4], t->w[5], t->w[6], t->w[7], t->celldx, t->celldy,
                            t->celldz, s->advect_x, which, sweep, dir);
    }
}

/* advection.f90:26-181 */
static void advection(ostate *s)
{
    int fields[19];
    static const int fc[] = {F_E1, F_D1, F_VFX, F_VFY, F_VFZ};
    static const int fm[] = {F_D1, F_E1, F_U1, F_V1, F_W1, F_MFX, F_MFY, F_MFZ};
    int dir = s->advect_x ? 1 : 3;

    set_fields(fields, fc, 5);
    update_halo(s, fields, 2);
    advec_cell(s, 1, dir);
    set_fields(fields, fm, 8);
    update_halo(s, fields, 2);
    advec_mom(s, 1, dir);

    set_fields(fields, fc, 5);
    update_halo(s, fields, 2);
    advec_cell(s, 2, 2);
    set_fields(fields, fm, 8);
    update_halo(s, fields, 2);
    advec_mom(s, 2, 2);

    dir = s->advect_x ? 3 : 1;
    advec_cell(s, 3, dir);
    set_fields(fields, fm, 8);
    update_halo(s, fields, 2);
    advec_mom(s, 3, dir);
}

/* The reference has one subroutine per face with the 2 x 19 arrays spelled
 * out; this adapter expands the oracle's (face, array list) form. */
typedef void (*fn_ref_tile)(OR_DIMS, double *, double *, double *, double *, double *, double *, double *,
                            double *, double *, double *, double *, double *, double *, double *, double *,
                            double *, double *, double *, double *, int, int, int, int, int, int, double *,
                            double *, double *, double *, double *, double *, double *, double *, double *,
                            double *, double *, double *, double *, double *, double *, double *, double *,
                            double *, double *, const int *, int);
static fn_ref_tile ref_tile_fns[6];

static void ref_tile_adapter(int face, OR_DIMS, double *const *a, int n_xmin, int n_xmax, int n_ymin, int n_ymax,
                             int n_zmin, int n_zmax, double *const *b, const int *fields, int depth)
{
    ref_tile_fns[face - 1](x_min, x_max, y_min, y_max, z_min, z_max, a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7],
                           a[8], a[9], a[10], a[11], a[12], a[13], a[14], a[15], a[16], a[17], a[18], n_xmin, n_xmax,
                           n_ymin, n_ymax, n_zmin, n_zmax, b[0], b[1], b[2], b[3], b[4], b[5], b[6], b[7], b[8], b[9],
                           b[10], b[11], b[12], b[13], b[14], b[15], b[16], b[17], b[18], fields, depth);
}

/* Loads a kernel table from a shared library whose entry points follow the
 * reference argument lists: the reference's own kernels (prefix "ref_") or
 * any other implementation of the same C signatures, e.g. the product's
 * drop-in ABI (prefix "clover_hip_"), so that a library can be driven in
 * exactly the reference's call order. */
static int load_lib(const char *path, const char *prefix, ktab *k)
{
    void *h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
    if (!h) {
        fprintf(stderr, "oracle: cannot load %s: %s\n", path, dlerror());
        return -1;
    }
    char name[128];
#define SYM(field, kname)                                                    \
    do {                                                                     \
        snprintf(name, sizeof name, "%s%s", prefix, kname);                  \
        *(void **)(&k->field) = dlsym(h, name);                              \
        if (!k->field) {                                                     \
            fprintf(stderr, "oracle: %s lacks %s\n", path, name);           \
            return -1;                                                       \
        }                                                                    \
    } while (0)
    SYM(ideal_gas, "ideal_gas_kernel");
    SYM(viscosity, "viscosity_kernel");
    SYM(calc_dt, "calc_dt_kernel");
    snprintf(name, sizeof name, "%sPdV_kernel", prefix);
    *(void **)(&k->pdv) = dlsym(h, name);
    if (!k->pdv) SYM(pdv, "pdv_kernel");
    SYM(revert, "revert_kernel");
    SYM(accelerate, "accelerate_kernel");
    SYM(flux_calc, "flux_calc_kernel");
    SYM(advec_cell, "advec_cell_kernel");
    SYM(advec_mom, "advec_mom_kernel");
    SYM(reset, "reset_field_kernel");
    SYM(update_halo, "update_halo_kernel");
    SYM(summary, "field_summary_kernel");
#undef SYM
    /* The set-up kernels are optional: the reference's own are in _ref, a
     * drop-in kernel library need not have them (the C ones run then). */
    snprintf(name, sizeof name, "%sinitialise_chunk_kernel", prefix);
    *(void **)(&k->init_chunk) = dlsym(h, name);
    if (!k->init_chunk) k->init_chunk = oracle_initialise_chunk_kernel;
    snprintf(name, sizeof name, "%sgenerate_chunk_kernel", prefix);
    *(void **)(&k->gen_chunk) = dlsym(h, name);
    if (!k->gen_chunk) k->gen_chunk = oracle_generate_chunk_kernel;
    static const char *tile_names[6] = {"update_tile_halo_left_kernel", "update_tile_halo_right_kernel",
                                        "update_tile_halo_bottom_kernel", "update_tile_halo_top_kernel",
                                        "update_tile_halo_back_kernel", "update_tile_halo_front_kernel"};
    for (int f = 0; f < 6; f++) {
        snprintf(name, sizeof name, "%s%s", prefix, tile_names[f]);
        *(void **)(&ref_tile_fns[f]) = dlsym(h, name);
        if (!ref_tile_fns[f]) {
            fprintf(stderr, "oracle: %s lacks %s\n", path, name);
            return -1;
        }
    }
    k->tile_halo = ref_tile_adapter;
    return 0;
}

static double last_hydro_seconds = 0.0;

double oracle_last_hydro_seconds(void) { return last_hydro_seconds; }

static char kernel_prefix[64] = "ref_";

void oracle_set_kernel_prefix(const char *prefix)
{
    snprintf(kernel_prefix, sizeof kernel_prefix, "%s", prefix ? prefix : "ref_");
}

/* The kernel table: the C restatement (lib NULL) or a library's kernels. */
static int get_ktab(const char *lib, ktab *k)
{
    if (lib) return load_lib(lib, kernel_prefix, k);
    k->ideal_gas = oracle_ideal_gas_kernel;
    k->viscosity = oracle_viscosity_kernel;
    k->calc_dt = oracle_calc_dt_kernel;
    k->pdv = oracle_PdV_kernel;
    k->revert = oracle_revert_kernel;
    k->accelerate = oracle_accelerate_kernel;
    k->flux_calc = oracle_flux_calc_kernel;
    k->advec_cell = oracle_advec_cell_kernel;
    k->advec_mom = oracle_advec_mom_kernel;
    k->reset = oracle_reset_field_kernel;
    k->update_halo = oracle_update_halo_kernel;
    k->tile_halo = oracle_update_tile_halo_face;
    k->summary = oracle_field_summary_kernel;
    k->init_chunk = oracle_initialise_chunk_kernel;
    k->gen_chunk = oracle_generate_chunk_kernel;
    return 0;
}

/* Test hook: update_halo.f90:27-88 (the tile passes of update_tile_halo.f90
 * in their order, then each tile's reflective pass) over caller-owned
 * tiles, through the same driver code oracle_run uses.  dims: 3 per tile;
 * nbr: 6 per tile (tile ids from 1, -1 external); arrays: 19 per tile in
 * update_tile_halo argument order.  Returns 0, or -1 if lib cannot load. */
int oracle_update_halo_tiles(int ntiles, const int *dims, const int *nbr, double *const *arrays, const int *fields,
                             int depth, const char *lib)
{
    ktab k;
    if (get_ktab(lib, &k)) return -1;
    ostate s;
    memset(&s, 0, sizeof(s));
    s.k = &k;
    s.ntiles = ntiles;
    s.tiles = calloc(ntiles, sizeof(otile));
    for (int ti = 0; ti < ntiles; ti++) {
        otile *t = &s.tiles[ti];
        double *const *a = arrays + 19 * ti;
        t->nx = dims[3 * ti]; t->ny = dims[3 * ti + 1]; t->nz = dims[3 * ti + 2];
        memcpy(t->nbr, nbr + 6 * ti, 6 * sizeof(int));
        t->d0 = a[0]; t->e0 = a[1]; t->p = a[2]; t->q = a[3]; t->c = a[4]; t->d1 = a[5]; t->e1 = a[6];
        t->u0 = a[7]; t->v0 = a[8]; t->w0 = a[9]; t->u1 = a[10]; t->v1 = a[11]; t->w1 = a[12];
        t->vfx = a[13]; t->vfy = a[14]; t->vfz = a[15]; t->mfx = a[16]; t->mfy = a[17]; t->mfz = a[18];
    }
    update_halo(&s, fields, depth);
    free(s.tiles);
    return 0;
}

/* Final-field dump of the next oracle_run: interior values gathered from
 * every tile into whole-chunk arrays (Fortran order; cells X*Y*Z, nodes
 * (X+1)*(Y+1)*(Z+1); a node on a tile boundary takes the last tile's
 * value).  dst: 13 pointers, density0, energy0, pressure, viscosity,
 * soundspeed, density1, energy1, xvel0, yvel0, zvel0, xvel1, yvel1, zvel1;
 * NULL entries are skipped.  The setting holds for one run. */
static double *dump_dst[13];
static int dump_on = 0, dump_visit = 0;

void oracle_set_dump(double *const *dst, int visit)
{
    dump_on = dst != NULL;
    dump_visit = dst != NULL && visit;
    for (int i = 0; i < 13; i++) dump_dst[i] = dst ? dst[i] : NULL;
}

/* visit.f90:65-81: ideal_gas on level 0, the pressure and velocity halos
 * (depth 1), then viscosity -- the values a visit dump writes */
static void visit_prepare(ostate *s)
{
    ideal_gas(s, 0);
    int fields[19];
    static const int fv[] = {F_P, F_U0, F_V0, F_W0};
    set_fields(fields, fv, 4);
    update_halo(s, fields, 1);
    for (int ti = 0; ti < s->ntiles; ti++) {
        otile *t = &s->tiles[ti];
        s->k->viscosity(TD(t), t->xarea, t->yarea, t->zarea, t->celldx, t->celldy, t->celldz, t->d0, t->p, t->q,
                        t->u0, t->v0, t->w0);
    }
}

static void dump_fields(ostate *s)
{
    if (dump_visit) visit_prepare(s);
    const oracle_deck *d = s->deck;
    const size_t X = d->x_cells, Y = d->y_cells;
    for (int ti = 0; ti < s->ntiles; ti++) {
        const otile *t = &s->tiles[ti];
        const double *src[13] = {t->d0, t->e0, t->p, t->q, t->c, t->d1, t->e1,
                                 t->u0, t->v0, t->w0, t->u1, t->v1, t->w1};
        for (int f = 0; f < 13; f++) {
            if (!dump_dst[f]) continue;
            const int node = f >= 7;
            const size_t ex = t->nx + 4 + node, ey = t->ny + 4 + node; /* tile extents */
            const size_t gx = X + node, gy = Y + node;                 /* dump extents */
            for (int l = 1; l <= t->nz + node; l++)
                for (int kk = 1; kk <= t->ny + node; kk++)
                    for (int j = 1; j <= t->nx + node; j++) {
                        size_t si = (size_t)(j + 1) + ex * ((size_t)(kk + 1) + ey * (size_t)(l + 1));
                        size_t gj = t->left - 1 + j - 1, gk = t->bottom - 1 + kk - 1, gl = t->back - 1 + l - 1;
                        dump_dst[f][gj + gx * (gk + gy * gl)] = src[f][si];
                    }
        }
    }
}

int oracle_run(const oracle_deck *deck, const char *ref_lib, oracle_summary *rows, int max_rows,
               double *dt_hist, int max_steps, int *steps_done)
{
    ktab k;
    if (get_ktab(ref_lib, &k)) return -1;
    ostate s;
    memset(&s, 0, sizeof(s));
    s.deck = deck;
    s.k = &k;
    s.g_small = (double)1.0e-16f; /* default-real literals, data.f90:30-31 */
    s.g_big = (double)1.0e+21f;

    int T = deck->tiles_per_chunk;
    int *nx = calloc(T, sizeof(int)), *ny = calloc(T, sizeof(int)), *nz = calloc(T, sizeof(int));
    int *lf = calloc(T, sizeof(int)), *bt = calloc(T, sizeof(int)), *bk = calloc(T, sizeof(int));
    int *nb = calloc(6 * T, sizeof(int));
    s.ntiles = oracle_decompose_tiles(deck->x_cells, deck->y_cells, deck->z_cells, T, deck->tile_dims, nx, ny, nz,
                                      lf, bt, bk, nb);
    s.tiles = calloc(s.ntiles, sizeof(otile));

    /* start.f90:25-144 */
    s.time = 0.0;
    s.step = 0;
    s.dtold = deck->dtinit;
    s.dt = deck->dtinit;
    double dx = (deck->xmax - deck->xmin) / (double)(float)deck->x_cells;
    double dy = (deck->ymax - deck->ymin) / (double)(float)deck->y_cells;
    double dz = (deck->zmax - deck->zmin) / (double)(float)deck->z_cells;
    for (int ti = 0; ti < s.ntiles; ti++) {
        otile *t = &s.tiles[ti];
        t->nx = nx[ti]; t->ny = ny[ti]; t->nz = nz[ti];
        t->left = lf[ti]; t->bottom = bt[ti]; t->back = bk[ti];
        memcpy(t->nbr, nb + 6 * ti, 6 * sizeof(int));
        build_field(t);
        /* initialise_chunk.f90:37-44 */
        double xmin = deck->xmin + dx * (double)(float)(t->left - 1);
        double ymin = deck->ymin + dy * (double)(float)(t->bottom - 1);
        double zmin = deck->zmin + dz * (double)(float)(t->back - 1);
        k.init_chunk(TD(t), xmin, ymin, zmin, dx, dy, dz, t->vertexx, t->vertexdx, t->vertexy, t->vertexdy,
                     t->vertexz, t->vertexdz, t->cellx, t->celldx, t->celly, t->celldy, t->cellz, t->celldz, t->vol,
                     t->xarea, t->yarea, t->zarea);
        k.gen_chunk(TD(t), t->vertexx, t->vertexy, t->vertexz, t->cellx, t->celly, t->cellz, t->d0, t->e0, t->u0,
                    t->v0, t->w0, deck->number_of_states, deck->state_density, deck->state_energy, deck->state_xvel,
                    deck->state_yvel, deck->state_zvel, deck->state_xmin, deck->state_xmax, deck->state_ymin,
                    deck->state_ymax, deck->state_zmin, deck->state_zmax, deck->state_radius,
                    deck->state_geometry);
    }
    free(nx); free(ny); free(nz); free(lf); free(bt); free(bk); free(nb);
    s.advect_x = 1;
    ideal_gas(&s, 0);
    int fields[19];
    static const int fstart[] = {F_D0, F_E0, F_P, F_Q, F_D1, F_E1, F_U0, F_V0, F_W0, F_U1, F_V1, F_W1};
    set_fields(fields, fstart, 12);
    update_halo(&s, fields, 2);
    int nrows = 0;
    if (nrows < max_rows) field_summary(&s, &rows[nrows++]);

    /* hydro.f90:46-87 (timed like the reference's Wall clock, hydro.f90:44,90) */
    int rc = 0;
    struct timespec ts0, ts1;
    clock_gettime(CLOCK_MONOTONIC, &ts0);
    for (;;) {
        s.step = s.step + 1;
        if (timestep(&s)) {
            fprintf(stderr, "oracle: small timestep at step %d\n", s.step);
            rc = -2;
            break;
        }
        if (dt_hist && s.step <= max_steps) dt_hist[s.step - 1] = s.dt;
        pdv(&s, 1);
        for (int ti = 0; ti < s.ntiles; ti++) {
            otile *t = &s.tiles[ti];
            k.accelerate(TD(t), s.dt, t->xarea, t->yarea, t->zarea, t->vol, t->d0, t->p, t->q, t->u0, t->v0, t->w0,
                         t->u1, t->v1, t->w1);
        }
        pdv(&s, 0);
        for (int ti = 0; ti < s.ntiles; ti++) {
            otile *t = &s.tiles[ti];
            k.flux_calc(TD(t), s.dt, t->xarea, t->yarea, t->zarea, t->u0, t->v0, t->w0, t->u1, t->v1, t->w1, t->vfx,
                        t->vfy, t->vfz);
        }
        advection(&s);
        for (int ti = 0; ti < s.ntiles; ti++) {
            otile *t = &s.tiles[ti];
            k.reset(TD(t), t->d0, t->d1, t->e0, t->e1, t->u0, t->u1, t->v0, t->v1, t->w0, t->w1);
        }
        s.advect_x = !s.advect_x;
        s.time = s.time + s.dt;
        if (deck->summary_frequency != 0 && s.step % deck->summary_frequency == 0)
            if (nrows < max_rows) field_summary(&s, &rows[nrows++]);
        if (s.time + s.g_small > deck->end_time || s.step >= deck->end_step) {
            if (nrows < max_rows) field_summary(&s, &rows[nrows++]);
            break;
        }
    }
    clock_gettime(CLOCK_MONOTONIC, &ts1);
    last_hydro_seconds = (double)(ts1.tv_sec - ts0.tv_sec) + 1e-9 * (double)(ts1.tv_nsec - ts0.tv_nsec);
    if (steps_done) *steps_done = s.step;
    if (dump_on) dump_fields(&s);
    oracle_set_dump(NULL, 0);
    for (int ti = 0; ti < s.ntiles; ti++) free_field(&s.tiles[ti]);
    free(s.tiles);
    return rc ? rc : nrows;
}
