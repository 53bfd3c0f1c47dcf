/*
 * C ABI of the MI355X (gfx950) CloverLeaf3D hydro kernels.
 *
 * Two layers, one shared library (cloverleaf3d_ref_amd/libclover_hip.so):
 *
 * 1. Drop-in kernels.  One entry point per reference *_kernel subroutine,
 *    same arguments in the same order (scalars by value, arrays as the base
 *    address of the caller's Fortran column-major arrays with lower bound
 *    min-2, LOGICALs as int, flags by pointer).  A Fortran shim module of the
 *    same name as the reference module calls these through ISO_C_BINDING.
 *    Arrays are identified by their host base address: the first time an
 *    address is seen the library allocates a device copy in HBM and uploads
 *    it; from then on the device copy is authoritative and no field data
 *    crosses PCIe, except dt (calc_dt) and the five summary sums
 *    (field_summary), which are returned synchronously.  clover_hip_fetch()
 *    copies a device array back into its host array on request.  Calls are
 *    enqueued in program order on one HIP stream.
 *
 * 2. Device-resident driver.  clover_chunk_* runs start + the hydro loop of
 *    the reference (start.f90, hydro.f90, timestep.f90, PdV.f90,
 *    advection.f90, update_halo.f90, field_summary.f90) with every field in
 *    HBM for the whole run.
 *
 * Every function returns 0 on success and a negative code on error;
 * clover_hip_last_error() describes the last error of the calling thread.
 */
#ifndef CLOVER_HIP_H
#define CLOVER_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define CLOVER_DIMS int x_min, int x_max, int y_min, int y_max, int z_min, int z_max

/* ---------------------------------------------------------------- lifecycle */
int clover_hip_init(int device);           /* select device, create the stream */
int clover_hip_finalize(void);             /* free every device array (clover_finalize) */
int clover_hip_sync(void);                 /* wait for all enqueued work */
int clover_hip_fetch(const void *host);    /* device copy -> host array */
int clover_hip_push(const void *host);     /* host array -> device copy */
int clover_hip_forget(const void *host);   /* drop one device copy */
int clover_hip_resident_count(void);       /* arrays currently resident */
const char *clover_hip_last_error(void);

/* ------------------------------------------------------- drop-in kernels */

/* ideal_gas_kernel.f90:27-62 */
int clover_hip_ideal_gas_kernel(CLOVER_DIMS, double *density, double *energy, double *pressure,
                                double *soundspeed);

/* viscosity_kernel.f90:32-139 */
int clover_hip_viscosity_kernel(CLOVER_DIMS, double *xarea, double *yarea, double *zarea, double *celldx,
                                double *celldy, double *celldz, double *density0, double *pressure,
                                double *viscosity, double *xvel0, double *yvel0, double *zvel0);

/* calc_dt_kernel.f90:28-183; dt_min (work_array1) receives the per-cell dt */
int clover_hip_calc_dt_kernel(CLOVER_DIMS, double g_small, double g_big, double dtmin, double dtc_safe,
                              double dtu_safe, double dtv_safe, double dtw_safe, double dtdiv_safe,
                              double *xarea, double *yarea, double *zarea, double *cellx, double *celly,
                              double *cellz, double *celldx, double *celldy, double *celldz, double *volume,
                              double *density0, double *energy0, double *pressure, double *viscosity_a,
                              double *soundspeed, double *xvel0, double *yvel0, double *zvel0, double *dt_min,
                              double *dt_min_val, int *dtl_control, double *xl_pos, double *yl_pos,
                              double *zl_pos, int *jldt, int *kldt, int *lldt, int *small);

/* PdV_kernel.f90:32-172 (predict: LOGICAL as int) */
int clover_hip_PdV_kernel(int predict, CLOVER_DIMS, double dt, double *xarea, double *yarea, double *zarea,
                          double *volume, double *density0, double *density1, double *energy0, double *energy1,
                          double *pressure, double *viscosity, double *xvel0, double *xvel1, double *yvel0,
                          double *yvel1, double *zvel0, double *zvel1);

/* revert_kernel.f90:29-65 */
int clover_hip_revert_kernel(CLOVER_DIMS, double *density0, double *density1, double *energy0, double *energy1);

/* accelerate_kernel.f90:27-115 */
int clover_hip_accelerate_kernel(CLOVER_DIMS, double dt, double *xarea, double *yarea, double *zarea,
                                 double *volume, double *density0, double *pressure, double *viscosity,
                                 double *xvel0, double *yvel0, double *zvel0, double *xvel1, double *yvel1,
                                 double *zvel1);

/* flux_calc_kernel.f90:26-95 */
int clover_hip_flux_calc_kernel(CLOVER_DIMS, double dt, double *xarea, double *yarea, double *zarea,
                                double *xvel0, double *yvel0, double *zvel0, double *xvel1, double *yvel1,
                                double *zvel1, double *vol_flux_x, double *vol_flux_y, double *vol_flux_z);

/* advec_cell_kernel.f90:30-383 (advect_x: LOGICAL as int) */
int clover_hip_advec_cell_kernel(CLOVER_DIMS, int advect_x, int dir, int sweep_number, double *vertexdx,
                                 double *vertexdy, double *vertexdz, double *volume, double *density1,
                                 double *energy1, double *mass_flux_x, double *vol_flux_x, double *mass_flux_y,
                                 double *vol_flux_y, double *mass_flux_z, double *vol_flux_z, double *pre_vol,
                                 double *post_vol, double *pre_mass, double *post_mass, double *advec_vol,
                                 double *post_ener, double *ener_flux);

/* advec_mom_kernel.f90:29-428; node_flux, node_mass_post and node_mass_pre
 * computed by the which_vel=1 call are reused by the which_vel=2,3 calls */
int clover_hip_advec_mom_kernel(CLOVER_DIMS, double *vel1, double *mass_flux_x, double *vol_flux_x,
                                double *mass_flux_y, double *vol_flux_y, double *mass_flux_z, double *vol_flux_z,
                                double *volume, double *density1, double *node_flux, double *node_mass_post,
                                double *node_mass_pre, double *advec_vel, double *mom_flux, double *pre_vol,
                                double *post_vol, double *celldx, double *celldy, double *celldz, int advect_x,
                                int which_vel, int sweep_number, int direction);

/* reset_field_kernel.f90:27-101 */
int clover_hip_reset_field_kernel(CLOVER_DIMS, double *density0, double *density1, double *energy0,
                                  double *energy1, double *xvel0, double *xvel1, double *yvel0, double *yvel1,
                                  double *zvel0, double *zvel1);

/* update_halo_kernel.f90:32-1427; fields[19], neighbours[6] */
int clover_hip_update_halo_kernel(CLOVER_DIMS, const int *chunk_neighbours, const int *tile_neighbours,
                                  double *density0, double *energy0, double *pressure, double *viscosity,
                                  double *soundspeed, double *density1, double *energy1, double *xvel0,
                                  double *yvel0, double *zvel0, double *xvel1, double *yvel1, double *zvel1,
                                  double *vol_flux_x, double *vol_flux_y, double *vol_flux_z, double *mass_flux_x,
                                  double *mass_flux_y, double *mass_flux_z, const int *fields, int depth);

/* update_tile_halo_kernel.f90:31, 361, 690, 1022, 1351, 1674: the tile's
 * bounds and 19 arrays, the neighbour's bounds and 19 arrays, flags, depth */
#define CLOVER_TILE_ARGS                                                                                        \
    CLOVER_DIMS, double *density0, double *energy0, double *pressure, double *viscosity, double *soundspeed,     \
        double *density1, double *energy1, double *xvel0, double *yvel0, double *zvel0, double *xvel1,          \
        double *yvel1, double *zvel1, double *vol_flux_x, double *vol_flux_y, double *vol_flux_z,              \
        double *mass_flux_x, double *mass_flux_y, double *mass_flux_z, int n_xmin, int n_xmax, int n_ymin,      \
        int n_ymax, int n_zmin, int n_zmax, double *n_density0, double *n_energy0, double *n_pressure,        \
        double *n_viscosity, double *n_soundspeed, double *n_density1, double *n_energy1, double *n_xvel0,      \
        double *n_yvel0, double *n_zvel0, double *n_xvel1, double *n_yvel1, double *n_zvel1,                   \
        double *n_vol_flux_x, double *n_vol_flux_y, double *n_vol_flux_z, double *n_mass_flux_x,               \
        double *n_mass_flux_y, double *n_mass_flux_z, const int *fields, int depth
int clover_hip_update_tile_halo_left_kernel(CLOVER_TILE_ARGS);
int clover_hip_update_tile_halo_right_kernel(CLOVER_TILE_ARGS);
int clover_hip_update_tile_halo_top_kernel(CLOVER_TILE_ARGS);
int clover_hip_update_tile_halo_bottom_kernel(CLOVER_TILE_ARGS);
int clover_hip_update_tile_halo_front_kernel(CLOVER_TILE_ARGS);
int clover_hip_update_tile_halo_back_kernel(CLOVER_TILE_ARGS);

/* field_summary_kernel.f90:27-81 */
int clover_hip_field_summary_kernel(CLOVER_DIMS, double *volume, double *density0, double *energy0,
                                    double *pressure, double *xvel0, double *yvel0, double *zvel0, double *vol,
                                    double *mass, double *ie, double *ke, double *press);

/* No counterpart in the reference: field_summary_kernel.f90:27-81 for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The arrays are the
 * caller's DEVICE arrays in the reference's own shapes (build_field.f90:
 * 41-147: volume, density0, energy0, pressure (x_min-2:x_max+2, ...),
 * xvel0, yvel0, zvel0 (x_min-2:x_max+3, ...)); sums receives vol, mass, ie,
 * ke, press (5 doubles of device memory).  Enqueued on `stream` (a
 * hipStream_t; NULL for the default stream) without waiting and without any
 * allocation: the block partial sums go to a fixed 640 KiB device array of
 * the library, one per GPU, held until the process ends and shared by every
 * call on that GPU, so calls that overlap on one GPU race.  Tiles whose
 * vertex arrays exceed 2^31-1 elements are rejected (E_ARG).  The sums equal
 * clover_hip_field_summary_kernel's bit for bit. */
int clover_hip_field_summary_kernel_noalloc(CLOVER_DIMS, const double *volume, const double *density0,
                                            const double *energy0, const double *pressure, const double *xvel0,
                                            const double *yvel0, const double *zvel0, double *sums, void *stream);

/* ------------------------------------------------ device-resident driver */

#define CLOVER_MAX_STATES 16

/* A parsed input deck (read_input.f90:24-305), state extents already
 * shrunk by a hundredth of a cell (read_input.f90:293-303). */
typedef struct {
    int x_cells, y_cells, z_cells;
    double xmin, ymin, zmin, xmax, ymax, zmax;
    double dtinit, dtmax, dtrise, dtmin;
    double dtc_safe, dtu_safe, dtv_safe, dtw_safe, dtdiv_safe;
    double end_time;
    int end_step;
    int summary_frequency;
    int tiles_per_chunk;
    int tile_dims; /* 1, 2 or 3: 1d/2d/3d_tile_decomposition */
    int number_of_states;
    double state_density[CLOVER_MAX_STATES], state_energy[CLOVER_MAX_STATES];
    double state_xvel[CLOVER_MAX_STATES], state_yvel[CLOVER_MAX_STATES], state_zvel[CLOVER_MAX_STATES];
    double state_xmin[CLOVER_MAX_STATES], state_xmax[CLOVER_MAX_STATES];
    double state_ymin[CLOVER_MAX_STATES], state_ymax[CLOVER_MAX_STATES];
    double state_zmin[CLOVER_MAX_STATES], state_zmax[CLOVER_MAX_STATES];
    double state_radius[CLOVER_MAX_STATES];
    int state_geometry[CLOVER_MAX_STATES]; /* 1 cuboid, 2 sphere, 3 point */
} clover_deck;

typedef struct {
    int step;
    double time, vol, mass, press, ie, ke;
} clover_summary;

typedef struct clover_chunk clover_chunk;

/* Allocates every tile's fields in HBM, builds the mesh and paints the
 * states (initialise_chunk, generate_chunk), then primes pressure and the
 * halos as start.f90:103-129 does. */
int clover_chunk_create(const clover_deck *deck, int device, clover_chunk **out);
int clover_chunk_destroy(clover_chunk *c);

/* As clover_chunk_create, with flags.  By default the chunk runs the fused
 * kernels, which write bit-identical values with fewer HBM passes (uniform
 * mesh geometry as scalars, neighbouring loop nests fused, values the
 * reference never reads not stored, reset_field as an exchange of the level
 * 0/1 arrays); CLOVER_CHUNK_PER_CALL runs one kernel per reference loop nest
 * instead, every intermediate stored as the reference stores it. */
#define CLOVER_CHUNK_PER_CALL 1
/* The FP32-store form: every field array held in FP32 (half the HBM bytes)
 * and the stencils computed in FP32, while calc_dt's minimum, the
 * advection's mass and energy accumulation and the field_summary sums
 * (compensated) run in FP64.  Fused kernels only. */
#define CLOVER_CHUNK_FP32 2
/* Test mode of the fused kernels: the fused step refreshes only the halos its
 * kernels read; with this flag the halos the reference would refresh and the
 * fused step skips are filled with wrong values instead, so that a read of
 * one changes the results (which must then still equal the per-call
 * kernels'). */
#define CLOVER_CHUNK_HALO_CHECK 4
int clover_chunk_create_ex(const clover_deck *deck, int device, int flags, clover_chunk **out);

/* One step of hydro.f90:48-68: timestep, PdV predictor/corrector,
 * accelerate, flux_calc, advection, reset_field; time += dt.  *dt gets the
 * step's dt; returns 1 (and does nothing else) on a small time step. */
int clover_chunk_step(clover_chunk *c, double *dt);

/* field_summary.f90:27-139: ideal_gas, then the five sums (tiles summed in
 * tile order).  out: vol, mass, ie, ke, press. */
int clover_chunk_field_summary(clover_chunk *c, double *out);

/* The device work of visit.f90:65-87 before a VTK dump: ideal_gas on every
 * tile, the pressure and level-0 velocity halos at depth 1, viscosity.
 * Afterwards clover_chunk_fetch_field returns the values visit.f90:127-198
 * writes.  The run continues unchanged: the next timestep recomputes the
 * same values. */
int clover_chunk_visit_prepare(clover_chunk *c);

/* The hydro loop until end_step / end_time (hydro.f90:46-87) or max_steps
 * more steps, with a summary row every summary_frequency steps and at the
 * end.  Row 0 (start) is not included; call clover_chunk_field_summary. */
int clover_chunk_run(clover_chunk *c, int max_steps, clover_summary *rows, int max_rows, int *nrows);

/* step count, simulated time, last dt */
int clover_chunk_state(const clover_chunk *c, int *step, double *time, double *dt);
/* The rest of timestep.f90:115-116's step line for the last step: dt_control
 * (1 'sound'), jdt, kdt, x_pos and y_pos as calc_dt_kernel.f90:154-160
 * reports them, from the tile whose dt was taken (timestep.f90:93-100). */
int clover_chunk_dt_control(const clover_chunk *c, int *control, int *jdt, int *kdt, double *x_pos, double *y_pos);
int clover_chunk_tiles(const clover_chunk *c, int *ntiles);
/* tile t: interior extents, global offsets, neighbours[6] */
int clover_chunk_tile_info(const clover_chunk *c, int t, int *dims3, int *offs3, int *neighbours6);
/* The tile split clover_chunk_create makes, on the host only (no device
 * call): clover_decompose_tile, clover.f90:258-447, with its surface and
 * volume in default REAL (clover.f90:282-285, 323-326).  Fills 3 extents,
 * 3 global offsets (first cell, from 1) and 6 neighbours (tile ids from 1,
 * -1 external) per tile; returns the tile count or E_ARG. */
int clover_decompose_tiles(int x_cells, int y_cells, int z_cells, int tiles_per_chunk, int tile_dims, int max_tiles,
                           int *dims3, int *offs3, int *neighbours6);

/* Copy one field of tile t into a dense Fortran array of the field's own
 * class.  field: 0..18 the FIELD_* ids minus one (data.f90:55-74); 19 volume,
 * 20 xarea, 21 yarea, 22 zarea.  Between steps the fused kernels' interiors
 * equal the reference's, with three differences that nothing reads: the
 * level-1 arrays hold the previous level 0 (reset_field is an exchange),
 * halo cells no kernel reads are not refreshed, and pressure/soundspeed
 * already hold ideal_gas of the new level 0 (the reference computes that at
 * the next timestep or field_summary; after a field_summary both agree).
 * The fused kernels form soundspeed where they read it, so fetching it
 * first runs that ideal_gas (the pressure it writes has the bits it had). */
int clover_chunk_fetch_field(clover_chunk *c, int t, int field, double *host);

/* Per-phase device time (ms) accumulated with HIP events when profiling is
 * on: the reference profiler's categories (hydro.f90:147-167). */
int clover_chunk_set_profiling(clover_chunk *c, int on);
int clover_chunk_profile(const clover_chunk *c, double *ms, int *calls, int n);
const char *clover_chunk_profile_name(int i);

/* HBM bytes held by the chunk */
long long clover_chunk_device_bytes(const clover_chunk *c);

/* Per-kernel device time: with timing on, every kernel launch of the
 * chunk's calls is bracketed by HIP events on the chunk's stream (no extra
 * host synchronisation).  clover_chunk_kernel_times writes one line per
 * kernel, "name total_ms launches", into buf.  Turning timing on resets it. */
int clover_chunk_set_kernel_timing(clover_chunk *c, int on);
int clover_chunk_kernel_times(clover_chunk *c, char *buf, int buflen);

/* Launch shapes of the fused marching kernels (no counterpart in the
 * reference, whose loops have no launch geometry).  A shape is `planes`,
 * the positions a workgroup marches over (planes of the z marches, faces or
 * nodes along the y/z advection sweeps; the x cell sweep, which has no
 * march, walks planes/32 rows per wave, at least one),
 * `xcd`: 1 gives each of the 8 XCDs a contiguous eighth of the grid instead
 * of the hardware's round-robin order, and `block_x`: the workgroup width
 * of the visc/dt and PdV-predictor z marches (64, 128, 256 or 512 threads
 * along x, with 256/64, 256/128 or 1 rows; the other kernels' workgroups are
 * fixed, block_x 64).
 * A shape changes how the work is split, never a value.  Kernel indices: */
#define CLOVER_SHAPE_VISC_DT 0
#define CLOVER_SHAPE_PDV_PREDICT 1
#define CLOVER_SHAPE_ACCEL_PDV 2 /* accelerate + PdV corrector + flux_calc */
#define CLOVER_SHAPE_CELL_X 3
#define CLOVER_SHAPE_CELL_Y 4
#define CLOVER_SHAPE_CELL_Z 5
#define CLOVER_SHAPE_MOM_X 6
#define CLOVER_SHAPE_MOM_Y 7
#define CLOVER_SHAPE_MOM_Z 8
#define CLOVER_SHAPE_COUNT 9
/* The autotuner: times every shaped kernel with every candidate shape (planes
 * 8, 16, 32, 64, 128 x both block orders x, for the z marches, the four
 * workgroup widths) on this chunk's own arrays and keeps
 * the fastest.  Candidates write only spare work arrays, so every field
 * keeps its value.  ms_before/ms_after (may be null) get one pass of all
 * shaped kernels with the old and the chosen shapes.  Fused kernels only. */
int clover_chunk_autotune(clover_chunk *c, double *ms_before, double *ms_after);
int clover_chunk_get_shape(const clover_chunk *c, int kernel, int *planes, int *xcd, int *block_x);
int clover_chunk_set_shape(clover_chunk *c, int kernel, int planes, int xcd, int block_x);

/* ------------------------------------------------------------ ensembles */

/* Many independent decks advanced together on one device: each deck is one
 * tile with its own mesh, states, dt and time (a per-deck minimum of the
 * calc_dt cells on the device), and every kernel launch covers all decks.
 * The fields evolve exactly as clover_chunk_* evolves each deck alone with
 * the fused kernels.  Decks need tiles_per_chunk 1 and at least 3 cells
 * along every axis.  flags: 0 or CLOVER_CHUNK_FP32. */
typedef struct clover_ensemble clover_ensemble;
int clover_ensemble_create(const clover_deck *decks, int ndecks, int device, int flags, clover_ensemble **out);
int clover_ensemble_destroy(clover_ensemble *e);
/* Advance every deck `steps` steps (hydro.f90:48-68, no summaries); with
 * use_graphs the step of each advect_x parity is captured in a graph once
 * and replayed.  Returns 1 if a deck's dt has fallen below its dtmin (the
 * small time step on which the reference stops, timestep.f90:120-122): that
 * deck keeps stepping with the others, so from that step on its state is
 * not one the reference produces; clover_ensemble_state says which deck and
 * from which step. */
int clover_ensemble_run(clover_ensemble *e, int steps, int use_graphs);
/* field_summary of every deck: out[5*d + (vol, mass, ie, ke, press)] */
int clover_ensemble_summaries(clover_ensemble *e, double *out);
/* steps taken; per deck: simulated time, last dt, and small: 0, or the step
 * at which the deck's dt first fell below dtmin */
int clover_ensemble_state(clover_ensemble *e, int *step, double *time, double *dt, int *small);
/* one field (0..18) of deck d as a dense Fortran array of its class */
int clover_ensemble_fetch_field(clover_ensemble *e, int deck, int field, double *host);
long long clover_ensemble_device_bytes(const clover_ensemble *e);
/* The autotuner of the batched kernels: as clover_chunk_autotune, over all
 * decks at once, with outputs sent to spare arrays by exchanging deck-table
 * slots around the timed launches.  Drops the captured step graphs, which
 * are captured again with the chosen shapes on the next run. */
int clover_ensemble_autotune(clover_ensemble *e, double *ms_before, double *ms_after);
int clover_ensemble_get_shape(const clover_ensemble *e, int kernel, int *planes, int *xcd, int *block_x);
/* Workgroups of a batched kernel's launch: 0 for one per tile (a tile is
 * one workgroup-sized block of one deck), else that many persistent
 * workgroups walk the ensemble's tile list.  No counterpart in the
 * reference (it has one deck); chosen by clover_ensemble_autotune. */
int clover_ensemble_get_workgroups(const clover_ensemble *e, int kernel, int *workgroups);
/* Set a batched kernel's launch shape by hand (CLOVER_SHAPE_* kernel, planes
 * >= 8, xcd 0/1, workgroups as above).  Changes how the work is split,
 * never a value; captured steps are dropped and captured again. */
int clover_ensemble_set_shape(clover_ensemble *e, int kernel, int planes, int xcd, int workgroups);

#ifdef __cplusplus
}
#endif
#endif
