/*
 * CPU restatement of the CloverLeaf3D hydro kernels (the oracle).
 *
 * TEST INFRASTRUCTURE ONLY.  This library exists to check the HIP product
 * path; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it, and only as the checker.  Nothing in cloverleaf3d_ref_amd/
 * links, loads or calls it.
 *
 * Every function restates one reference subroutine with the same argument
 * list (scalars by value, arrays as base pointers to Fortran column-major
 * storage with lower bound min-2 on every axis), the same loop domains and the
 * same left-to-right expression order, so that it reproduces the reference's
 * FP64 results bit for bit (build with -ffp-contract=off, no fast-math).
 *
 * Array classes (extents per axis, nx = x_max-x_min+1):
 *   cell   (nx+4, ny+4, nz+4)        density, energy, pressure, viscosity,
 *                                    soundspeed, volume
 *   vertex (nx+5, ny+5, nz+5)        x/y/zvel0/1, work arrays
 *   xface  (nx+5, ny+4, nz+4)        xarea, vol_flux_x, mass_flux_x
 *   yface  (nx+4, ny+5, nz+4)        yarea, vol_flux_y, mass_flux_y
 *   zface  (nx+4, ny+4, nz+5)        zarea, vol_flux_z, mass_flux_z
 * (build_field.f90:41-147).
 */
#ifndef CLOVER_ORACLE_H
#define CLOVER_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define OR_DIMS int x_min, int x_max, int y_min, int y_max, int z_min, int z_max

/* ideal_gas_kernel.f90:27-62 */
void oracle_ideal_gas_kernel(OR_DIMS, const double *density, const double *energy,
                             double *pressure, double *soundspeed);

/* viscosity_kernel.f90:32-139 */
void oracle_viscosity_kernel(OR_DIMS, const double *xarea, const double *yarea, const double *zarea,
                             const double *celldx, const double *celldy, const double *celldz,
                             const double *density0, const double *pressure, double *viscosity,
                             const double *xvel0, const double *yvel0, const double *zvel0);

/* calc_dt_kernel.f90:28-183 */
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
                           int *jldt, int *kldt, int *lldt, int *small);

/* PdV_kernel.f90:32-172 */
void oracle_PdV_kernel(int predict, OR_DIMS, double dt,
                       const double *xarea, const double *yarea, const double *zarea, const double *volume,
                       const double *density0, double *density1, const double *energy0, double *energy1,
                       const double *pressure, const double *viscosity,
                       const double *xvel0, const double *xvel1, const double *yvel0, const double *yvel1,
                       const double *zvel0, const double *zvel1);

/* revert_kernel.f90:29-65 */
void oracle_revert_kernel(OR_DIMS, const double *density0, double *density1,
                          const double *energy0, double *energy1);

/* accelerate_kernel.f90:27-115 */
void oracle_accelerate_kernel(OR_DIMS, double dt,
                              const double *xarea, const double *yarea, const double *zarea,
                              const double *volume, const double *density0, const double *pressure,
                              const double *viscosity,
                              const double *xvel0, const double *yvel0, const double *zvel0,
                              double *xvel1, double *yvel1, double *zvel1);

/* flux_calc_kernel.f90:26-95 */
void oracle_flux_calc_kernel(OR_DIMS, double dt,
                             const double *xarea, const double *yarea, const double *zarea,
                             const double *xvel0, const double *yvel0, const double *zvel0,
                             const double *xvel1, const double *yvel1, const double *zvel1,
                             double *vol_flux_x, double *vol_flux_y, double *vol_flux_z);

/* advec_cell_kernel.f90:30-383 */
void oracle_advec_cell_kernel(OR_DIMS, int advect_x, int dir, int sweep_number,
                              const double *vertexdx, const double *vertexdy, const double *vertexdz,
                              const double *volume, double *density1, double *energy1,
                              double *mass_flux_x, const double *vol_flux_x,
                              double *mass_flux_y, const double *vol_flux_y,
                              double *mass_flux_z, const double *vol_flux_z,
                              double *pre_vol, double *post_vol, double *pre_mass, double *post_mass,
                              double *advec_vol, double *post_ener, double *ener_flux);

/* advec_mom_kernel.f90:29-428 */
void oracle_advec_mom_kernel(OR_DIMS, double *vel1,
                             const double *mass_flux_x, const double *vol_flux_x,
                             const double *mass_flux_y, const double *vol_flux_y,
                             const double *mass_flux_z, const double *vol_flux_z,
                             const double *volume, const double *density1,
                             double *node_flux, double *node_mass_post, double *node_mass_pre,
                             double *advec_vel, double *mom_flux, double *pre_vol, double *post_vol,
                             const double *celldx, const double *celldy, const double *celldz,
                             int advect_x, int which_vel, int sweep_number, int direction);

/* reset_field_kernel.f90:27-101 */
void oracle_reset_field_kernel(OR_DIMS, double *density0, const double *density1,
                               double *energy0, const double *energy1,
                               double *xvel0, const double *xvel1, double *yvel0, const double *yvel1,
                               double *zvel0, const double *zvel1);

/* update_halo_kernel.f90:32-1427 */
void oracle_update_halo_kernel(OR_DIMS, const int *chunk_neighbours, const int *tile_neighbours,
                               double *density0, double *energy0, double *pressure, double *viscosity,
                               double *soundspeed, double *density1, double *energy1,
                               double *xvel0, double *yvel0, double *zvel0,
                               double *xvel1, double *yvel1, double *zvel1,
                               double *vol_flux_x, double *vol_flux_y, double *vol_flux_z,
                               double *mass_flux_x, double *mass_flux_y, double *mass_flux_z,
                               const int *fields, int depth);

/* update_tile_halo_kernel.f90:31-1993.  `arr` and `nbr` each hold the 19
 * field base pointers in the reference argument order: density0, energy0,
 * pressure, viscosity, soundspeed, density1, energy1, xvel0, yvel0, zvel0,
 * xvel1, yvel1, zvel1, vol_flux_x, vol_flux_y, vol_flux_z, mass_flux_x,
 * mass_flux_y, mass_flux_z.  face: 1 left, 2 right, 3 bottom, 4 top,
 * 5 back, 6 front (data.f90:46-52). */
void oracle_update_tile_halo_face(int face, OR_DIMS, double *const *arr,
                                  int n_xmin, int n_xmax, int n_ymin, int n_ymax, int n_zmin, int n_zmax,
                                  double *const *nbr, const int *fields, int depth);

/* field_summary_kernel.f90:27-81 */
void oracle_field_summary_kernel(OR_DIMS, const double *volume, const double *density0,
                                 const double *energy0, const double *pressure,
                                 const double *xvel0, const double *yvel0, const double *zvel0,
                                 double *vol, double *mass, double *ie, double *ke, double *press);

/* initialise_chunk_kernel.f90:26-182 */
void oracle_initialise_chunk_kernel(OR_DIMS, double xmin, double ymin, double zmin,
                                    double dx, double dy, double dz,
                                    double *vertexx, double *vertexdx, double *vertexy, double *vertexdy,
                                    double *vertexz, double *vertexdz,
                                    double *cellx, double *celldx, double *celly, double *celldy,
                                    double *cellz, double *celldz,
                                    double *volume, double *xarea, double *yarea, double *zarea);

/* generate_chunk_kernel.f90:30-211 */
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
                                  const double *state_radius, const int *state_geometry);

/* ---- whole-run driver (hydro.f90, timestep.f90, PdV.f90, advection.f90,
 *      update_halo.f90, update_tile_halo.f90, field_summary.f90, start.f90) ---- */

#define ORACLE_MAX_STATES 16

typedef struct {
    int x_cells, y_cells, z_cells;
    double xmin, ymin, zmin, xmax, ymax, zmax;
    double dtinit, dtmax, dtrise, dtmin;
    double dtc_safe, dtu_safe, dtv_safe, dtw_safe, dtdiv_safe;
    double end_time;
    int end_step;
    int summary_frequency;
    int tiles_per_chunk;
    int tile_dims;              /* 1, 2 or 3 (x_bound/y_bound choice, clover.f90:286-299) */
    int number_of_states;
    double state_density[ORACLE_MAX_STATES], state_energy[ORACLE_MAX_STATES];
    double state_xvel[ORACLE_MAX_STATES], state_yvel[ORACLE_MAX_STATES], state_zvel[ORACLE_MAX_STATES];
    double state_xmin[ORACLE_MAX_STATES], state_xmax[ORACLE_MAX_STATES];
    double state_ymin[ORACLE_MAX_STATES], state_ymax[ORACLE_MAX_STATES];
    double state_zmin[ORACLE_MAX_STATES], state_zmax[ORACLE_MAX_STATES];
    double state_radius[ORACLE_MAX_STATES];
    int state_geometry[ORACLE_MAX_STATES];
} oracle_deck;

/* One field_summary row: step, time, vol, mass, press, ie, ke. */
typedef struct {
    int step;
    double time, vol, mass, press, ie, ke;
} oracle_summary;

/* Runs start + the hydro loop.  ref_lib: NULL for the C kernels above, or
 * the path of oracle/_ref/libclover_ref.so to drive the reference's own
 * compiled kernels through the same driver.  Writes up to max_rows summary
 * rows and returns how many, or a negative value on error.  *dt_hist (if
 * non-NULL, length >= max_steps) receives the dt of every step. */
int oracle_run(const oracle_deck *deck, const char *ref_lib, oracle_summary *rows, int max_rows,
               double *dt_hist, int max_steps, int *steps_done);

/* Tile decomposition (clover.f90:258-447): fills per-tile local extents,
 * global offsets and neighbours (6 per tile, -1 external).  Returns tiles. */
/* Symbol prefix oracle_run uses to find kernels in ref_lib ("ref_" by
 * default; "clover_hip_" drives the product's drop-in entry points in the
 * reference's call order, for the drop-in boundary test). */
void oracle_set_kernel_prefix(const char *prefix);

/* Wall time of the hydro loop of the last oracle_run (start excluded). */
double oracle_last_hydro_seconds(void);

int oracle_decompose_tiles(int x_cells, int y_cells, int z_cells, int tiles_per_chunk, int tile_dims,
                           int *t_nx, int *t_ny, int *t_nz, int *t_left, int *t_bottom, int *t_back,
                           int *neighbours);

/* update_halo.f90:27-88 over caller-owned tiles (tile passes in
 * update_tile_halo.f90's order, then each tile's reflective pass), with the
 * kernels of lib (NULL: the C restatement; else symbols under the prefix
 * set by oracle_set_kernel_prefix).  dims: 3 per tile; nbr: 6 per tile;
 * arrays: 19 per tile in update_tile_halo argument order. */
int oracle_update_halo_tiles(int ntiles, const int *dims, const int *nbr, double *const *arrays, const int *fields,
                             int depth, const char *lib);

/* Gather the final interior fields of the next oracle_run into whole-chunk
 * arrays: 13 pointers (density0, energy0, pressure, viscosity, soundspeed,
 * density1, energy1, xvel0, yvel0, zvel0, xvel1, yvel1, zvel1), cells
 * X*Y*Z and nodes (X+1)*(Y+1)*(Z+1), Fortran order; NULL entries skipped.
 * visit != 0 first runs visit.f90:65-81's ideal_gas, halo update and
 * viscosity, so the arrays hold what a visit dump at the last step writes. */
void oracle_set_dump(double *const *dst, int visit);

#ifdef __cplusplus
}
#endif
#endif
