// Ensemble driver: many independent decks advanced together.
//
// Each deck is one tile with its own mesh, states, dt and time.  A step is
// the fast path of the single-deck driver (chunk.cpp) with every launch
// covering all decks (batch.h): timestep (halos, viscosity + calc_dt, a
// per-deck dt on the device; its ideal_gas results come from the previous
// step's third cell sweep), PdV predictor, accelerate + PdV corrector +
// flux_calc, the three advection sweeps and reset_field, the array
// exchanges done in the device-side deck table.  No step reads
// anything back to the host, so the step for each advect_x parity is
// captured once in a graph and replayed.
#include <hip/hip_runtime.h>

#include <cstring>
#include <map>
#include <tuple>
#include <vector>

#include "../../include/clover_hip.h"
#include "batch.h"
#include "kernels.h"
#include "runtime.h"
#include "setup.h"

using namespace clover;

struct clover_ensemble {
    int nd = 0;
    int device = 0;
    Store store = ST_F64;
    hipStream_t st = nullptr;
    std::vector<clover_deck> decks;
    std::vector<DeckDev> host;  // the start table (device copy is authoritative afterwards)
    DeckDev *dev = nullptr;
    char *arena = nullptr;
    long long bytes = 0;
    std::vector<int *> tables;  // device prefix tables
    std::map<std::tuple<int, int, int>, BatchGrid> grids;
    double *visc_partial = nullptr;
    double *sum_partial = nullptr;
    double *sums = nullptr;
    double g_small = 0.0, g_big = 0.0;
    int step = 0;
    int advect_x = 1;
    hipGraphExec_t graph[2] = {nullptr, nullptr};  // index: advect_x
    Shapes shapes = default_shapes();              // launch shapes (autotuner)
};

namespace {

// prefix table of per-deck block counts, uploaded once
BatchGrid make_grid(clover_ensemble *e, const std::vector<int> &counts)
{
    std::vector<int> start(e->nd + 1, 0);
    for (int d = 0; d < e->nd; d++) start[d + 1] = start[d] + counts[d];
    int *p = nullptr;
    hip_check(hipMalloc(&p, sizeof(int) * (e->nd + 1)), "hipMalloc");
    hip_check(hipMemcpy(p, start.data(), sizeof(int) * (e->nd + 1), hipMemcpyHostToDevice), "H2D");
    e->tables.push_back(p);
    e->bytes += (long long)sizeof(int) * (e->nd + 1);
    return BatchGrid{p, start[e->nd]};
}

// the launch shape a batched kernel takes (none: -1)
int shape_index(int bk)
{
    switch (bk) {
    case BK_VISC_DT: return SK_VISC_DT;
    case BK_PDV_PREDICT: return SK_PDV_PREDICT;
    case BK_ACCEL_PDV: return SK_ACCEL_PDV;
    case BK_CELL_X: return SK_CELL_X;
    case BK_CELL_Y: return SK_CELL_Y;
    case BK_CELL_Z: return SK_CELL_Z;
    case BK_MOM_X: return SK_MOM_X;
    case BK_MOM_Y: return SK_MOM_Y;
    case BK_MOM_Z: return SK_MOM_Z;
    default: return -1;
    }
}

const Shape &shape(clover_ensemble *e, int bk) { return e->shapes.k[shape_index(bk)]; }

// kind: 0..BK_COUNT-1 kernels (a: planes of the launch shape, b unused);
// 100 + axis reflective halos (a: depth, b: items)
const BatchGrid &grid(clover_ensemble *e, int kind, int a, int b = 0)
{
    auto key = std::make_tuple(kind, a, b);
    auto it = e->grids.find(key);
    if (it != e->grids.end()) return it->second;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    hip_check(hipStreamIsCapturing(e->st, &cs), "capture status");
    if (cs != hipStreamCaptureStatusNone) throw Error(E_STATE, "ensemble: launch table missing during capture");
    std::vector<int> counts(e->nd);
    for (int d = 0; d < e->nd; d++)
        counts[d] = kind < 100 ? batch_blocks((BatchKernel)kind, e->host[d].g, a, e->store)
                               : batch_blocks_reflect(e->host[d].g, kind - 100, a, b);
    return e->grids[key] = make_grid(e, counts);
}

// a kernel's grid for its current launch shape
const BatchGrid &kgrid(clover_ensemble *e, int bk)
{
    const int si = shape_index(bk);
    return grid(e, bk, si < 0 ? 0 : e->shapes.k[si].planes);
}

void set_fields(int *fields, std::initializer_list<int> ids)
{
    for (int f = 0; f < NUM_FIELDS; f++) fields[f] = 0;
    for (int id : ids) fields[id] = 1;
}

// update_halo for single-tile decks: the reflective boundary on every face,
// y faces, then x, then z (as launch_update_halo orders them), reduced to
// the halos the fused kernels read (fused_halo_plan, driver.h)
int items_on_axis(const HaloPlan &p, int axis)
{
    int n = 0;
    for (int f = 0; f < NUM_FIELDS; f++) n += p.fields[f] == 1 && ((p.raxes[f] >> axis) & 1);
    return 2 * n;  // both faces
}

// every flagged field on every axis (start.f90's exchange)
void halo(clover_ensemble *e, const int *fields, int depth)
{
    int n = 0;
    for (int f = 0; f < NUM_FIELDS; f++) n += fields[f] == 1;
    static const int axis_order[3] = {1, 0, 2};
    for (int ao = 0; ao < 3; ao++) {
        const int axis = axis_order[ao];
        hip_check(batch_reflect(e->store, e->dev, e->nd, grid(e, 100 + axis, depth, 2 * n), axis, depth, fields,
                                e->st),
                  "halo");
    }
}

void halo(clover_ensemble *e, int which)
{
    const HaloPlan p = fused_halo_plan(which, e->advect_x != 0, AX_ALL);
    static const int axis_order[3] = {1, 0, 2};
    for (int ao = 0; ao < 3; ao++) {
        const int axis = axis_order[ao];
        const int n = items_on_axis(p, axis);
        if (n == 0) continue;
        hip_check(batch_reflect(e->store, e->dev, e->nd, grid(e, 100 + axis, p.depth, n), axis, p.depth, p.fields,
                                e->st, p.raxes),
                  "halo");
    }
}

void ideal_gas(clover_ensemble *e)
{
    hip_check(batch_ideal_gas(e->store, e->dev, e->nd, kgrid(e, BK_IDEAL_GAS), e->st), "ideal_gas");
}

void swap(clover_ensemble *e, std::initializer_list<std::pair<int, int>> pairs, bool add_time)
{
    SlotSwap sw{};
    for (auto &p : pairs) {
        sw.a[sw.n] = p.first;
        sw.b[sw.n] = p.second;
        sw.n++;
    }
    sw.add_time = add_time ? 1 : 0;
    hip_check(batch_swap(e->dev, e->nd, sw, e->st), "swap");
}

void advec_cell(clover_ensemble *e, int sweep, int dir)
{
    const BatchKernel k = dir == 1 ? BK_CELL_X : dir == 2 ? BK_CELL_Y : BK_CELL_Z;
    // the third sweep also writes ideal_gas of the new state: the next step's
    // (or a summary's) equation of state
    hip_check(batch_advec_cell(e->store, e->dev, e->nd, kgrid(e, k), e->advect_x, dir, sweep, sweep == 3,
                               shape(e, k), e->st),
              "advec_cell");
    swap(e, {{F_DENSITY1, X_W1}, {F_ENERGY1, X_W2}}, false);
}

void advec_mom(clover_ensemble *e, int sweep, int dir)
{
    const BatchKernel k = dir == 1 ? BK_MOM_X : dir == 2 ? BK_MOM_Y : BK_MOM_Z;
    hip_check(batch_advec_mom(e->store, e->dev, e->nd, kgrid(e, k), e->advect_x, dir, sweep, shape(e, k), e->st),
              "advec_mom");
    swap(e, {{F_XVEL1, X_W5}, {F_YVEL1, X_W6}, {F_ZVEL1, X_W7}}, false);
}

// hydro.f90:48-68 for every deck, device only
void enqueue_step(clover_ensemble *e)
{
    // timestep.f90:28-128; its ideal_gas results were written by the previous
    // step's third cell sweep (or by the start sequence)
    halo(e, FH_TIMESTEP);
    const BatchGrid &vg = kgrid(e, BK_VISC_DT);
    hip_check(batch_visc_dt(e->store, e->dev, e->nd, vg, shape(e, BK_VISC_DT), e->g_small, e->visc_partial, e->st),
              "visc_dt");
    hip_check(batch_deck_dt(e->dev, e->nd, vg.start, e->visc_partial, e->g_big, e->st), "deck dt");
    halo(e, FH_VISCOSITY);
    // PdV predictor (+ ideal_gas on the predicted state), pressure halo
    hip_check(batch_pdv_predict(e->store, e->dev, e->nd, kgrid(e, BK_PDV_PREDICT), shape(e, BK_PDV_PREDICT), e->st),
              "PdV predict");
    halo(e, FH_PRESSURE);
    hip_check(batch_accel_pdv(e->store, e->dev, e->nd, kgrid(e, BK_ACCEL_PDV), shape(e, BK_ACCEL_PDV), e->st),
              "accelerate + PdV correct");
    // advection.f90:26-181
    int dir = e->advect_x ? 1 : 3;
    halo(e, FH_CELL1);
    advec_cell(e, 1, dir);
    halo(e, FH_MOM1);
    advec_mom(e, 1, dir);
    halo(e, FH_CELL2);
    advec_cell(e, 2, 2);
    halo(e, FH_MOM2);
    advec_mom(e, 2, 2);
    dir = e->advect_x ? 3 : 1;
    advec_cell(e, 3, dir);
    halo(e, FH_MOM3);
    advec_mom(e, 3, dir);
    // reset_field as an exchange of levels, then time = time + dt
    swap(e, {{F_DENSITY0, F_DENSITY1}, {F_ENERGY0, F_ENERGY1}, {F_XVEL0, F_XVEL1}, {F_YVEL0, F_YVEL1},
             {F_ZVEL0, F_ZVEL1}},
         true);
}

// every prefix table a step uses, built before any graph capture (building
// one allocates and copies, which a capture cannot contain)
void prepare_grids(clover_ensemble *e)
{
    for (int k = 0; k < BK_COUNT; k++) kgrid(e, k);
    // the step's halo updates, for both sweep orders
    for (int ax = 0; ax < 2; ax++)
        for (int which = FH_TIMESTEP; which <= FH_MOM3; which++) {
            const HaloPlan p = fused_halo_plan(which, ax != 0, AX_ALL);
            for (int axis = 0; axis < 3; axis++) {
                const int n = items_on_axis(p, axis);
                if (n) grid(e, 100 + axis, p.depth, n);
            }
        }
}

// ---------------------------------------------------------------- autotuner
// The batched twin of the single-deck tuner (chunk.cpp): each shaped kernel
// runs over all decks with every candidate shape and the fastest is kept.
// The kernels take their arrays from the deck table, so outputs are sent to
// spare arrays by exchanging table slots around the timed launches (the
// pressure pdv_predict updates in place becomes a spare array too; its
// contents do not matter for timing).
SlotSwap tune_slots(int bk)
{
    SlotSwap sw{};
    auto pair = [&](int a, int b) {
        sw.a[sw.n] = a;
        sw.b[sw.n] = b;
        sw.n++;
    };
    switch (bk) {
    case BK_VISC_DT: pair(F_VISCOSITY, X_W1); break;
    case BK_PDV_PREDICT: pair(F_PRESSURE, X_W1); break;
    case BK_ACCEL_PDV:  // 8 outputs, 5 spare arrays: the volume fluxes keep their slots
        pair(F_DENSITY1, X_W1); pair(F_ENERGY1, X_W2);
        pair(F_XVEL1, X_W5); pair(F_YVEL1, X_W6); pair(F_ZVEL1, X_W7);
        break;
    case BK_CELL_X: pair(F_MASS_FLUX_X, X_W5); break;  // the new density/energy go to W1/W2 anyway
    case BK_CELL_Y: pair(F_MASS_FLUX_Y, X_W5); break;
    case BK_CELL_Z: pair(F_MASS_FLUX_Z, X_W5); break;
    default: break;  // momentum sweeps write W5..W7 only
    }
    return sw;
}

void tune_launch(clover_ensemble *e, int bk, const Shape &sh)
{
    const BatchGrid &g = grid(e, bk, sh.planes);
    hipError_t r = hipSuccess;
    switch (bk) {
    case BK_VISC_DT: r = batch_visc_dt(e->store, e->dev, e->nd, g, sh, e->g_small, e->visc_partial, e->st); break;
    case BK_PDV_PREDICT: r = batch_pdv_predict(e->store, e->dev, e->nd, g, sh, e->st); break;
    case BK_ACCEL_PDV: r = batch_accel_pdv(e->store, e->dev, e->nd, g, sh, e->st); break;
    case BK_CELL_X:
    case BK_CELL_Y:
    case BK_CELL_Z: {
        const int dir = bk - BK_CELL_X + 1;
        r = batch_advec_cell(e->store, e->dev, e->nd, g, 1, dir, dir == 2 ? 2 : 1, false, sh, e->st);
        break;
    }
    default: {
        const int dir = bk - BK_MOM_X + 1;
        r = batch_advec_mom(e->store, e->dev, e->nd, g, 1, dir, dir == 2 ? 2 : 1, sh, e->st);
        break;
    }
    }
    hip_check(r, "autotune launch");
}

double tune_time(clover_ensemble *e, int bk, const Shape &sh, hipEvent_t a, hipEvent_t b)
{
    const int reps = 3;
    grid(e, bk, sh.planes);  // any table allocation happens before the timing
    const SlotSwap sw = tune_slots(bk);
    if (sw.n) hip_check(batch_swap(e->dev, e->nd, sw, e->st), "swap");
    tune_launch(e, bk, sh);
    hip_check(hipEventRecord(a, e->st), "event");
    for (int r = 0; r < reps; r++) tune_launch(e, bk, sh);
    hip_check(hipEventRecord(b, e->st), "event");
    if (sw.n) hip_check(batch_swap(e->dev, e->nd, sw, e->st), "swap back");
    hip_check(hipEventSynchronize(b), "event sync");
    float ms = 0;
    hip_check(hipEventElapsedTime(&ms, a, b), "event time");
    return ms / reps;
}

// captured steps hold the old shapes: drop them (the next step captures
// again) and build the launch tables of the new ones
void shapes_changed(clover_ensemble *e)
{
    for (auto &g : e->graph)
        if (g) {
            (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
    prepare_grids(e);
}

void autotune(clover_ensemble *e, double *before, double *after)
{
    int cus = 0;
    hip_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, e->device), "CU count");
    hipEvent_t a = nullptr, b = nullptr;
    hip_check(hipEventCreate(&a), "event");
    hip_check(hipEventCreate(&b), "event");
    double t0 = 0.0, t1 = 0.0;
    for (int bk = 0; bk < BK_COUNT; bk++) {
        const int si = shape_index(bk);
        if (si < 0) continue;
        const Shape start = e->shapes.k[si];
        double best = tune_time(e, bk, start, a, b);
        Shape best_shape = start;
        t0 += best;
        for (int planes : kTunePlanes)
            for (int xcd = 0; xcd < 2; xcd++) {
                const Shape cand{planes, xcd, start.bx, 0};  // the batched workgroups are fixed at 64 x 4
                if (cand.planes == start.planes && cand.xcd == start.xcd && start.wgs == 0) continue;
                const double ms = tune_time(e, bk, cand, a, b);
                if (ms < best) {
                    best = ms;
                    best_shape = cand;
                }
            }
        // then a persistent grid of 1, 2, 4 or 8 workgroups per CU walking
        // the tiles, where that is fewer workgroups than tiles
        const int tiles = grid(e, bk, best_shape.planes).total;
        Shape pick = best_shape;
        for (int per_cu : {1, 2, 4, 8}) {
            Shape cand = best_shape;
            cand.wgs = per_cu * cus;
            if (cand.wgs >= tiles) break;
            const double ms = tune_time(e, bk, cand, a, b);
            if (ms < best) {
                best = ms;
                pick = cand;
            }
        }
        e->shapes.k[si] = pick;
        t1 += best;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    shapes_changed(e);
    if (before) *before = t0;
    if (after) *after = t1;
}

void destroy(clover_ensemble *e)
{
    if (!e) return;
    if (e->st) (void)hipStreamSynchronize(e->st);
    for (auto &g : e->graph)
        if (g) (void)hipGraphExecDestroy(g);
    for (int *p : e->tables) (void)hipFree(p);
    if (e->dev) (void)hipFree(e->dev);
    if (e->arena) (void)hipFree(e->arena);
    if (e->visc_partial) (void)hipFree(e->visc_partial);
    if (e->sum_partial) (void)hipFree(e->sum_partial);
    if (e->sums) (void)hipFree(e->sums);
    if (e->st) (void)hipStreamDestroy(e->st);
    delete e;
}

// the arrays the fused kernels use (volume, areas and W3/W4 stay unallocated)
bool allocated(int a) { return a < NUM_FIELDS || a == X_W1 || a == X_W2 || a >= X_W5; }

}  // namespace

extern "C" {

int clover_ensemble_create(const clover_deck *decks, int ndecks, int device, int flags, clover_ensemble **out)
{
    clover_ensemble *e = nullptr;
    try {
        if (!decks || !out || ndecks < 1) throw Error(E_ARG, "bad argument");
        if (flags & ~CLOVER_CHUNK_FP32) throw Error(E_ARG, "unknown flags");
        e = new clover_ensemble();
        e->nd = ndecks;
        e->device = device;
        e->store = (flags & CLOVER_CHUNK_FP32) ? ST_F32 : ST_F64;
        e->decks.assign(decks, decks + ndecks);
        e->g_small = (double)1.0e-16f;  // default-real literals (data.f90:30-31)
        e->g_big = (double)1.0e+21f;
        hip_check(hipSetDevice(device), "hipSetDevice");
        hip_check(hipStreamCreateWithFlags(&e->st, hipStreamNonBlocking), "hipStreamCreate");
        const size_t es = store_bytes(e->store);
        // layouts and one arena for every array of every deck
        std::vector<Tile> tiles(ndecks);
        std::vector<size_t> offs(ndecks);
        size_t total = 0;
        e->host.resize(ndecks);
        for (int d = 0; d < ndecks; d++) {
            const clover_deck &dk = decks[d];
            if (dk.tiles_per_chunk != 1) throw Error(E_ARG, "ensemble decks run as one tile each");
            if (dk.x_cells < 3 || dk.y_cells < 3 || dk.z_cells < 3)
                throw Error(E_ARG, "ensemble decks need at least 3 cells along every axis");
            if (dk.number_of_states < 1 || dk.number_of_states > CLOVER_MAX_STATES)
                throw Error(E_ARG, "bad state count");
            std::vector<Tile> one;
            decompose(dk, one);
            tiles[d] = one[0];
            const Lay g = make_lay(1, dk.x_cells, 1, dk.y_cells, 1, dk.z_cells);
            const int64_t n = lay_elems(g);
            if (n >= (int64_t(1) << 31)) throw Error(E_ARG, "deck too large for 32-bit element offsets");
            const size_t arr = ((size_t)n * es + 255) / 256 * 256;
            offs[d] = total;
            int na = 0;
            for (int a = 0; a < NUM_ARRAYS; a++) na += allocated(a);
            total += arr * na;
            DeckDev &D = e->host[d];
            memset(&D, 0, sizeof D);
            D.g = g;
            // every deck is one tile at least 3 cells thick with external
            // faces: the fused kernels read the mirrored interior instead of
            // the cell and vertex fields' reflective halos, and the
            // exchanges leave those out
            D.g.mirror = 63;
            // the uniform mesh as initialise_chunk.f90:37-39 computes it
            D.m = geo_of((dk.xmax - dk.xmin) / (double)(float)dk.x_cells,
                         (dk.ymax - dk.ymin) / (double)(float)dk.y_cells,
                         (dk.zmax - dk.zmin) / (double)(float)dk.z_cells);
            D.dt = dk.dtinit;
            D.dtold = dk.dtinit;
            D.time = 0.0;
            D.dtmax = dk.dtmax;
            D.dtrise = dk.dtrise;
            D.dtmin = dk.dtmin;
            D.dtc_safe = dk.dtc_safe;
            D.dtu_safe = dk.dtu_safe;
            D.dtv_safe = dk.dtv_safe;
            D.dtw_safe = dk.dtw_safe;
            D.dtdiv_safe = dk.dtdiv_safe;
        }
        hip_check(hipMalloc(&e->arena, total), "hipMalloc arena");
        hip_check(hipMemsetAsync(e->arena, 0, total, e->st), "hipMemset arena");
        e->bytes += (long long)total;
        for (int d = 0; d < ndecks; d++) {
            DeckDev &D = e->host[d];
            const size_t arr = ((size_t)lay_elems(D.g) * es + 255) / 256 * 256;
            char *p = e->arena + offs[d];
            for (int a = 0; a < NUM_ARRAYS; a++) {
                if (!allocated(a)) continue;
                D.f[a] = p;
                p += arr;
            }
            // mesh and states painted on the device (generate_chunk)
            hip_check(launch_generate(e->store, D.g, gen_state(decks[d], 1, 1, 1), D.f[F_DENSITY0], D.f[F_ENERGY0],
                                      D.f[F_XVEL0], D.f[F_YVEL0], D.f[F_ZVEL0], nullptr, nullptr, nullptr, nullptr,
                                      e->st),
                      "generate");
        }
        hip_check(hipMalloc(&e->dev, sizeof(DeckDev) * ndecks), "hipMalloc decks");
        hip_check(hipMemcpy(e->dev, e->host.data(), sizeof(DeckDev) * ndecks, hipMemcpyHostToDevice), "H2D decks");
        prepare_grids(e);
        // calc_dt's partials: room for the visc/dt grid of every shape the
        // autotuner tries (the most blocks at its fewest planes)
        const BatchGrid &vg = grid(e, BK_VISC_DT, kTunePlanes[0]);
        const BatchGrid &sg = kgrid(e, BK_SUMMARY);
        hip_check(hipMalloc(&e->visc_partial, sizeof(double) * (vg.total + 1)), "hipMalloc");
        hip_check(hipMalloc(&e->sum_partial, sizeof(double) * 5 * (sg.total + 1)), "hipMalloc");
        hip_check(hipMalloc(&e->sums, sizeof(double) * 5 * ndecks), "hipMalloc");
        // start.f90:58-129
        ideal_gas(e);
        int f[NUM_FIELDS];
        set_fields(f, {F_DENSITY0, F_ENERGY0, F_PRESSURE, F_VISCOSITY, F_DENSITY1, F_ENERGY1, F_XVEL0, F_YVEL0,
                       F_ZVEL0, F_XVEL1, F_YVEL1, F_ZVEL1});
        halo(e, f, 2);
        hip_check(hipStreamSynchronize(e->st), "sync");
        *out = e;
        return 0;
    } catch (const Error &err) {
        set_last_error(err.what());
        destroy(e);
        return err.code;
    } catch (const std::exception &err) {
        set_last_error(err.what());
        destroy(e);
        return E_ARG;
    }
}

int clover_ensemble_destroy(clover_ensemble *e)
{
    destroy(e);
    return 0;
}

int clover_ensemble_run(clover_ensemble *e, int steps, int use_graphs)
{
    CLOVER_API_BEGIN
    if (!e || steps < 0) throw Error(E_ARG, "bad argument");
    hip_check(hipSetDevice(e->device), "hipSetDevice");
    for (int s = 0; s < steps; s++) {
        if (use_graphs) {
            hipGraphExec_t &ge = e->graph[e->advect_x];
            if (!ge) {
                hipGraph_t g = nullptr;
                hip_check(hipStreamBeginCapture(e->st, hipStreamCaptureModeThreadLocal), "begin capture");
                try {
                    enqueue_step(e);
                } catch (...) {
                    (void)hipStreamEndCapture(e->st, &g);
                    if (g) (void)hipGraphDestroy(g);
                    throw;
                }
                hip_check(hipStreamEndCapture(e->st, &g), "end capture");
                hip_check(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0), "instantiate");
                (void)hipGraphDestroy(g);
            }
            hip_check(hipGraphLaunch(ge, e->st), "graph launch");
        } else {
            enqueue_step(e);
        }
        e->advect_x = !e->advect_x;
        e->step++;
    }
    // the reference aborts on a small time step (timestep.f90:120-122): say
    // so as clover_chunk_step does, with return code 1
    std::vector<DeckDev> h(e->nd);
    hip_check(hipMemcpyAsync(h.data(), e->dev, sizeof(DeckDev) * e->nd, hipMemcpyDeviceToHost, e->st), "D2H");
    hip_check(hipStreamSynchronize(e->st), "sync");
    for (int d = 0; d < e->nd; d++)
        if (h[d].small) return 1;
    CLOVER_API_END
}

int clover_ensemble_summaries(clover_ensemble *e, double *out)
{
    CLOVER_API_BEGIN
    if (!e || !out) throw Error(E_ARG, "bad argument");
    hip_check(hipSetDevice(e->device), "hipSetDevice");
    ideal_gas(e);  // field_summary.f90:49-60 recomputes pressure first
    const BatchGrid &sg = kgrid(e, BK_SUMMARY);
    hip_check(batch_field_summary(e->store, e->dev, e->nd, sg, e->sum_partial, e->st), "field_summary");
    hip_check(batch_deck_sum5(e->nd, sg.start, e->sum_partial, e->sums, e->st), "summary sums");
    hip_check(hipMemcpyAsync(out, e->sums, sizeof(double) * 5 * e->nd, hipMemcpyDeviceToHost, e->st), "D2H");
    hip_check(hipStreamSynchronize(e->st), "sync");
    CLOVER_API_END
}

int clover_ensemble_state(clover_ensemble *e, int *step, double *time, double *dt, int *small)
{
    CLOVER_API_BEGIN
    if (!e) throw Error(E_ARG, "bad argument");
    std::vector<DeckDev> h(e->nd);
    hip_check(hipMemcpyAsync(h.data(), e->dev, sizeof(DeckDev) * e->nd, hipMemcpyDeviceToHost, e->st), "D2H");
    hip_check(hipStreamSynchronize(e->st), "sync");
    if (step) *step = e->step;
    for (int d = 0; d < e->nd; d++) {
        if (time) time[d] = h[d].time;
        if (dt) dt[d] = h[d].dt;
        if (small) small[d] = h[d].small;
    }
    CLOVER_API_END
}

int clover_ensemble_fetch_field(clover_ensemble *e, int deck, int field, double *host)
{
    CLOVER_API_BEGIN
    if (!e || !host || deck < 0 || deck >= e->nd) throw Error(E_ARG, "bad argument");
    if (field < 0 || field >= NUM_FIELDS) throw Error(E_ARG, "bad field id");
    // the step forms soundspeed where it reads it instead of storing it:
    // ideal_gas of the current level 0 (its pressure comes out with the same
    // bits the step left)
    if (field == F_SOUNDSPEED) ideal_gas(e);
    DeckDev D;
    hip_check(hipMemcpyAsync(&D, e->dev + deck, sizeof D, hipMemcpyDeviceToHost, e->st), "D2H");
    hip_check(hipStreamSynchronize(e->st), "sync");
    const int cls = class_of(field);
    if (e->store == ST_F64) {
        download_dense(D.g, cls, (const double *)D.f[field], host, e->st);
        hip_check(hipStreamSynchronize(e->st), "sync");
    } else {
        std::vector<float> fl((size_t)dense_elems(D.g, cls));
        download_dense_any(D.g, cls, D.f[field], fl.data(), sizeof(float), e->st);
        hip_check(hipStreamSynchronize(e->st), "sync");
        for (size_t q = 0; q < fl.size(); q++) host[q] = fl[q];
    }
    CLOVER_API_END
}

long long clover_ensemble_device_bytes(const clover_ensemble *e) { return e ? e->bytes : 0; }

int clover_ensemble_autotune(clover_ensemble *e, double *ms_before, double *ms_after)
{
    CLOVER_API_BEGIN
    if (!e) throw Error(E_ARG, "null ensemble");
    hip_check(hipSetDevice(e->device), "hipSetDevice");
    autotune(e, ms_before, ms_after);
    CLOVER_API_END
}

int clover_ensemble_get_shape(const clover_ensemble *e, int kernel, int *planes, int *xcd, int *block_x)
{
    CLOVER_API_BEGIN
    if (!e || !planes || !xcd || !block_x) throw Error(E_ARG, "null argument");
    if (kernel < 0 || kernel >= SK_COUNT) throw Error(E_ARG, "kernel out of range");
    *planes = e->shapes.k[kernel].planes;
    *xcd = e->shapes.k[kernel].xcd;
    *block_x = e->shapes.k[kernel].bx;
    CLOVER_API_END
}

int clover_ensemble_get_workgroups(const clover_ensemble *e, int kernel, int *workgroups)
{
    CLOVER_API_BEGIN
    if (!e || !workgroups) throw Error(E_ARG, "null argument");
    if (kernel < 0 || kernel >= SK_COUNT) throw Error(E_ARG, "kernel out of range");
    *workgroups = e->shapes.k[kernel].wgs;
    CLOVER_API_END
}

int clover_ensemble_set_shape(clover_ensemble *e, int kernel, int planes, int xcd, int workgroups)
{
    CLOVER_API_BEGIN
    if (!e) throw Error(E_ARG, "null ensemble");
    if (kernel < 0 || kernel >= SK_COUNT) throw Error(E_ARG, "kernel out of range");
    // calc_dt's partial buffer holds the visc/dt tiles of the fewest planes the tuner tries
    if (planes < kTunePlanes[0]) throw Error(E_ARG, "planes below the smallest tuned value");
    if (xcd != 0 && xcd != 1) throw Error(E_ARG, "xcd must be 0 or 1");
    if (workgroups < 0) throw Error(E_ARG, "workgroups must be >= 0");
    hip_check(hipSetDevice(e->device), "hipSetDevice");
    hip_check(hipStreamSynchronize(e->st), "sync");
    Shape &sh = e->shapes.k[kernel];
    sh.planes = planes;
    sh.xcd = xcd;
    sh.wgs = workgroups;
    shapes_changed(e);
    CLOVER_API_END
}

}  // extern "C"
