// Device-resident driver: start + the hydro loop of the reference with every
// field of every tile in HBM for the whole run.
//
// Phase order follows hydro.f90:48-68, timestep.f90:28-128, PdV.f90:26-109,
// advection.f90:26-181, update_halo.f90:27-88 (tile exchange in three passes,
// update_tile_halo.f90:47-423, then the reflective boundary per tile) and
// field_summary.f90:27-139.  Only dt (8 bytes per step) and the summary sums
// cross to the host.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <map>
#include <string>
#include <vector>

#include "../../include/clover_hip.h"
#include "fast.h"
#include "kernels.h"
#include "runtime.h"
#include "driver.h"
#include "setup.h"
#include "batch.h"

namespace clover {

enum Prof : int {
    P_IDEAL_GAS, P_VISCOSITY, P_TIMESTEP, P_PDV, P_REVERT, P_ACCEL, P_FLUX, P_CELL_ADVEC, P_MOM_ADVEC, P_RESET,
    P_TILE_HALO, P_SELF_HALO, P_SUMMARY, P_COUNT
};
static const char *kProfNames[P_COUNT] = {"ideal_gas", "viscosity", "timestep", "PdV", "revert",
                                          "acceleration", "fluxes", "cell_advection", "momentum_advection",
                                          "reset", "tile_halo_exchange", "self_halo_exchange", "summary"};

// Per-kernel device time from HIP events recorded around every launch.
// Events come from a pool and are only read back when the pool is full or
// the owner asks, so timing adds no host synchronisation to a step.
class EventTimer : public KernelTimer {
  public:
    struct Stat {
        double ms = 0.0;
        long long launches = 0;
    };
    ~EventTimer() override
    {
        for (auto &e : pool_) {
            (void)hipEventDestroy(e.a);
            (void)hipEventDestroy(e.b);
        }
    }
    void begin(const char *kernel, hipStream_t st) override
    {
        if (used_ == pool_.size()) {
            if (pool_.size() >= kPool) drain();
            if (used_ == pool_.size()) {
                Pair p;
                hip_check(hipEventCreate(&p.a), "hipEventCreate");
                hip_check(hipEventCreate(&p.b), "hipEventCreate");
                pool_.push_back(p);
            }
        }
        Pair &p = pool_[used_];
        p.name = kernel;
        hip_check(hipEventRecord(p.a, st), "hipEventRecord");
    }
    void end(hipStream_t st) override { hip_check(hipEventRecord(pool_[used_++].b, st), "hipEventRecord"); }
    void drain()
    {
        for (size_t i = 0; i < used_; i++) {
            hip_check(hipEventSynchronize(pool_[i].b), "hipEventSynchronize");
            float ms = 0.f;
            hip_check(hipEventElapsedTime(&ms, pool_[i].a, pool_[i].b), "hipEventElapsedTime");
            Stat &s = stats[pool_[i].name];
            s.ms += ms;
            s.launches++;
        }
        used_ = 0;
    }
    void reset()
    {
        drain();
        stats.clear();
    }
    std::map<std::string, Stat> stats;

  private:
    struct Pair {
        hipEvent_t a = nullptr, b = nullptr;
        const char *name = nullptr;
    };
    static constexpr size_t kPool = 8192;
    std::vector<Pair> pool_;
    size_t used_ = 0;
};

struct TimerScope {
    TimerScope(EventTimer *t) : prev(t_kernel_timer) { t_kernel_timer = t; }
    ~TimerScope() { t_kernel_timer = prev; }
    KernelTimer *prev;
};


}  // namespace clover

using namespace clover;

struct clover_chunk {
    clover_deck deck;
    int device;
    hipStream_t st;
    std::vector<Tile> tiles;
    double *partial;  // per tile: kMaxPartials * 5
    double *scalar;   // per tile: 8
    double *h_scalar; // pinned
    double dt, dtold, time;
    int step;
    int advect_x;
    double g_small, g_big;
    long long bytes;
    bool profiling;
    hipEvent_t ev0, ev1;
    double prof_ms[P_COUNT];
    int prof_calls[P_COUNT];
    EventTimer *ktimer;  // non-null while per-kernel timing is on
    bool fast;           // fused kernels (fast.h) instead of the per-call ones
    Store store;         // element type of the field arrays (ST_F32 only with the fused kernels)
    bool halo_check;     // CLOVER_CHUNK_HALO_CHECK
    bool gas_fresh;      // pressure interiors hold ideal_gas(density0, energy0)
    bool sound_fresh;    // and soundspeed interiors too (the fused step forms soundspeed where it reads it)
    Shapes shapes;       // launch shapes of the fused kernels (autotuner)
    int dt_tile;         // the tile whose dt the last timestep took (-1: none yet)
    Geo geo;             // the uniform mesh, for the fused kernels
    // More than one tile with the fused kernels: each stage is one launch over
    // every tile (batch.h, the ensemble's batched kernels with the tiles as
    // decks) and each update_halo one launch per pass.  The tiles' arrays
    // are read from a device table that every array exchange keeps in step
    // with `tiles`.
    bool batched;
    DeckDev *tdev;
    std::map<std::pair<int, int>, BatchGrid> grids;  // (batched kernel, planes) -> prefix table
    std::vector<int *> tables;
    struct HaloTabs {
        TileHaloItem *dev;
        int n_copy[3], off_copy[3], pa_copy[3], pb_copy[3];
        int n_refl[3][2], off_refl[3][2], pa_refl[3], pb_refl[3];
        bool split[3];  // the faces of that reflective axis in two launches (a tile under 3 cells thick)
    };
    std::map<std::string, HaloTabs> halo_tabs;  // per update_halo signature
};

// installs the chunk's kernel timer (if any) for the duration of an API call
#define KTIMER_SCOPE(c) TimerScope _kts((c)->ktimer)

namespace clover {

// ------------------------------------------------------------ profiling
struct Phase {
    clover_chunk *c;
    int p;
    Phase(clover_chunk *c_, int p_) : c(c_), p(p_)
    {
        if (c->profiling) hip_check(hipEventRecord(c->ev0, c->st), "event");
    }
    ~Phase() noexcept(false)
    {
        if (!c->profiling) return;
        hip_check(hipEventRecord(c->ev1, c->st), "event");
        hip_check(hipEventSynchronize(c->ev1), "event sync");
        float ms = 0.f;
        hip_check(hipEventElapsedTime(&ms, c->ev0, c->ev1), "elapsed");
        c->prof_ms[p] += ms;
        c->prof_calls[p]++;
    }
};

// ------------------------------------------------------------ tile decomposition (clover.f90:258-447)
void decompose(const clover_deck &d, std::vector<Tile> &tiles)
{
    const int T = d.tiles_per_chunk;
    const int xc = d.x_cells, yc = d.y_cells, zc = d.z_cells;
    int cur_x = 1, cur_y = 1, cur_z = T;
    // clover.f90:282-285 and 323-326 form surface and volume in default REAL
    // (single precision: `1.0*x_cells` is a REAL(4) product) and only widen
    // the two results to REAL(8) before dividing.  Where two splits tie in
    // exact arithmetic, the single-precision rounding decides which wins.
    auto metric = [&](int cx, int cy, int cz) {
        const float fx = (1.0f * (float)xc) / (float)cx, fy = (1.0f * (float)yc) / (float)cy,
                    fz = (1.0f * (float)zc) / (float)cz;
        const float s = ((fx * fy) * 2.0f + (fx * fz) * 2.0f) + (fy * fz) * 2.0f;
        const float v = (fx * fy) * fz;
        return (double)s / (double)v;
    };
    double best = metric(cur_x, cur_y, cur_z);
    int chunk_x = cur_x, chunk_y = cur_y, chunk_z = cur_z;
    int x_bound = T, y_bound = T;
    if (d.tile_dims == 2) { x_bound = 1; y_bound = T; }
    if (d.tile_dims == 1) { x_bound = 1; y_bound = 1; }
    for (int t = 1; t <= x_bound; t++) {
        if (T % t) continue;
        int div1 = T / t;
        if (y_bound < div1) div1 = y_bound;
        for (int j = 1; j <= div1; j++) {
            if (div1 % j) continue;
            if (T % (t * j)) continue;
            const double m = metric(t, j, T / (t * j));
            if (m < best) { chunk_x = t; chunk_y = j; chunk_z = T / (t * j); best = m; }
        }
    }
    const int dx = xc / chunk_x, dy = yc / chunk_y, dz = zc / chunk_z;
    const int mx = xc % chunk_x, my = yc % chunk_y, mz = zc % chunk_z;
    int ax_prev = 0, ay_prev = 0, az_prev = 0;
    tiles.clear();
    for (int cz = 1; cz <= chunk_z; cz++) {
        for (int cy = 1; cy <= chunk_y; cy++) {
            for (int cx = 1; cx <= chunk_x; cx++) {
                const int ax = cx <= mx, ay = cy <= my, az = cz <= mz;
                Tile t{};
                const int left = (cx - 1) * dx + 1 + ax_prev, right = left + dx - 1 + ax;
                const int bottom = (cy - 1) * dy + 1 + ay_prev, top = bottom + dy - 1 + ay;
                const int back = (cz - 1) * dz + 1 + az_prev, front = back + dz - 1 + az;
                const int id = (int)tiles.size() + 1;
                t.nbr[0] = cx == 1 ? -1 : id - 1;
                t.nbr[1] = cx == chunk_x ? -1 : id + 1;
                t.nbr[2] = cy == 1 ? -1 : id - chunk_x;
                t.nbr[3] = cy == chunk_y ? -1 : id + chunk_x;
                t.nbr[4] = cz == 1 ? -1 : id - chunk_x * chunk_y;
                t.nbr[5] = cz == chunk_z ? -1 : id + chunk_x * chunk_y;
                t.nx = right - left + 1; t.ny = top - bottom + 1; t.nz = front - back + 1;
                t.left = left; t.bottom = bottom; t.back = back;
                tiles.push_back(t);
                if (cx <= mx) ax_prev++;
            }
            ax_prev = 0;
            if (cy <= my) ay_prev++;
        }
        ax_prev = 0;
        ay_prev = 0;
        if (cz <= mz) az_prev++;
    }
}

// ------------------------------------------------------------ mesh and states on the host
// initialise_chunk.f90:37-44 + initialise_chunk_kernel.f90:63-178 and
// generate_chunk_kernel.f90:82-207, into dense Fortran arrays, then uploaded.

void host_init(const clover_deck &d, const Tile &t, HostInit &h, bool fields)
{
    const int nx = t.nx, ny = t.ny, nz = t.nz;
    const double dx = (d.xmax - d.xmin) / (double)(float)d.x_cells;
    const double dy = (d.ymax - d.ymin) / (double)(float)d.y_cells;
    const double dz = (d.zmax - d.zmin) / (double)(float)d.z_cells;
    const double xmin = d.xmin + dx * (double)(float)(t.left - 1);
    const double ymin = d.ymin + dy * (double)(float)(t.bottom - 1);
    const double zmin = d.zmin + dz * (double)(float)(t.back - 1);
    // 1-D arrays are indexed from min-2 = -1 (offset +2 from index 1)
    h.vertexx.assign(nx + 5, 0.0); h.vertexdx.assign(nx + 5, 0.0);
    h.vertexy.assign(ny + 5, 0.0); h.vertexdy.assign(ny + 5, 0.0);
    h.vertexz.assign(nz + 5, 0.0); h.vertexdz.assign(nz + 5, 0.0);
    h.cellx.assign(nx + 4, 0.0); h.celldx.assign(nx + 4, 0.0);
    h.celly.assign(ny + 4, 0.0); h.celldy.assign(ny + 4, 0.0);
    h.cellz.assign(nz + 4, 0.0); h.celldz.assign(nz + 4, 0.0);
    for (int j = -1; j <= nx + 3; j++) { h.vertexx[j + 1] = xmin + dx * (double)(float)(j - 1); h.vertexdx[j + 1] = dx; }
    for (int k = -1; k <= ny + 3; k++) { h.vertexy[k + 1] = ymin + dy * (double)(float)(k - 1); h.vertexdy[k + 1] = dy; }
    for (int l = -1; l <= nz + 3; l++) { h.vertexz[l + 1] = zmin + dz * (double)(float)(l - 1); h.vertexdz[l + 1] = dz; }
    for (int j = -1; j <= nx + 2; j++) { h.cellx[j + 1] = 0.5 * (h.vertexx[j + 1] + h.vertexx[j + 2]); h.celldx[j + 1] = dx; }
    for (int k = -1; k <= ny + 2; k++) { h.celly[k + 1] = 0.5 * (h.vertexy[k + 1] + h.vertexy[k + 2]); h.celldy[k + 1] = dy; }
    for (int l = -1; l <= nz + 2; l++) { h.cellz[l + 1] = 0.5 * (h.vertexz[l + 1] + h.vertexz[l + 2]); h.celldz[l + 1] = dz; }

    if (!fields) return;
    const size_t cx = nx + 4, cy = ny + 4, cz = nz + 4, vx = nx + 5, vy = ny + 5, vz = nz + 5;
    h.volume.assign(cx * cy * cz, 0.0);
    h.xarea.assign(vx * cy * cz, 0.0);
    h.yarea.assign(cx * vy * cz, 0.0);
    h.zarea.assign(cx * cy * vz, 0.0);
    h.density0.assign(cx * cy * cz, 0.0);
    h.energy0.assign(cx * cy * cz, 0.0);
    h.xvel0.assign(vx * vy * vz, 0.0);
    h.yvel0.assign(vx * vy * vz, 0.0);
    h.zvel0.assign(vx * vy * vz, 0.0);
    auto C = [&](int j, int k, int l) { return (size_t)(j + 1) + cx * ((size_t)(k + 1) + cy * (size_t)(l + 1)); };
    auto V = [&](int j, int k, int l) { return (size_t)(j + 1) + vx * ((size_t)(k + 1) + vy * (size_t)(l + 1)); };
    auto XF = [&](int j, int k, int l) { return (size_t)(j + 1) + vx * ((size_t)(k + 1) + cy * (size_t)(l + 1)); };
    auto YF = [&](int j, int k, int l) { return (size_t)(j + 1) + cx * ((size_t)(k + 1) + vy * (size_t)(l + 1)); };
    for (int l = -1; l <= nz + 2; l++)
        for (int k = -1; k <= ny + 2; k++)
            for (int j = -1; j <= nx + 2; j++) {
                h.volume[C(j, k, l)] = dx * dy * dz;
                h.xarea[XF(j, k, l)] = h.celldy[k + 1] * h.celldz[l + 1];
                h.yarea[YF(j, k, l)] = h.celldx[j + 1] * h.celldz[l + 1];
                h.zarea[C(j, k, l)] = h.celldx[j + 1] * h.celldy[k + 1];
            }
    // generate_chunk_kernel.f90: background state over cells, then states 2..n
    for (int l = -1; l <= nz + 2; l++)
        for (int k = -1; k <= ny + 2; k++)
            for (int j = -1; j <= nx + 2; j++) {
                h.energy0[C(j, k, l)] = d.state_energy[0];
                h.density0[C(j, k, l)] = d.state_density[0];
                h.xvel0[V(j, k, l)] = d.state_xvel[0];
                h.yvel0[V(j, k, l)] = d.state_yvel[0];
                h.zvel0[V(j, k, l)] = d.state_zvel[0];
            }
    for (int s = 1; s < d.number_of_states; s++) {
        const double xc = d.state_xmin[s], yc = d.state_ymin[s], zc = d.state_zmin[s];
        for (int l = -1; l <= nz + 2; l++)
            for (int k = -1; k <= ny + 2; k++)
                for (int j = -1; j <= nx + 2; j++) {
                    bool hit = false;
                    if (d.state_geometry[s] == 1) {
                        hit = h.vertexx[j + 2] >= d.state_xmin[s] && h.vertexx[j + 1] < d.state_xmax[s] &&
                              h.vertexy[k + 2] >= d.state_ymin[s] && h.vertexy[k + 1] < d.state_ymax[s] &&
                              h.vertexz[l + 2] >= d.state_zmin[s] && h.vertexz[l + 1] < d.state_zmax[s];
                    } else if (d.state_geometry[s] == 2) {
                        const double r = std::sqrt((h.cellx[j + 1] - xc) * (h.cellx[j + 1] - xc) +
                                                   (h.celly[k + 1] - yc) * (h.celly[k + 1] - yc) +
                                                   (h.cellz[l + 1] - zc) * (h.cellz[l + 1] - zc));
                        hit = r <= d.state_radius[s];
                    } else if (d.state_geometry[s] == 3) {
                        hit = h.vertexx[j + 1] == xc && h.vertexy[k + 1] == yc && h.vertexz[l + 1] == zc;
                    }
                    if (!hit) continue;
                    h.energy0[C(j, k, l)] = d.state_energy[s];
                    h.density0[C(j, k, l)] = d.state_density[s];
                    for (int lt = l; lt <= l + 1; lt++)
                        for (int kt = k; kt <= k + 1; kt++)
                            for (int jt = j; jt <= j + 1; jt++) {
                                h.xvel0[V(jt, kt, lt)] = d.state_xvel[s];
                                h.yvel0[V(jt, kt, lt)] = d.state_yvel[s];
                                h.zvel0[V(jt, kt, lt)] = d.state_zvel[s];
                            }
                }
    }
}

static double *dev_alloc_bytes(clover_chunk *c, size_t bytes)
{
    double *p = nullptr;
    hip_check(hipMalloc(&p, bytes), "hipMalloc");
    hip_check(hipMemsetAsync(p, 0, bytes, c->st), "hipMemset");
    c->bytes += (long long)bytes;
    return p;
}

static double *dev_alloc(clover_chunk *c, size_t n) { return dev_alloc_bytes(c, n * sizeof(double)); }


static double *dev_upload_1d(clover_chunk *c, const std::vector<double> &v)
{
    double *p = dev_alloc(c, v.size());
    hip_check(hipMemcpyAsync(p, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice, c->st), "H2D");
    return p;
}

// ------------------------------------------------------------ tiles as a batch
// the tile table on the device, from the tiles' current arrays
// tile t's table entry from the host's view of its arrays (which every
// exchange keeps equal to the device table's)
static DeckDev tile_entry(const clover_chunk *c, int t, double dt)
{
    DeckDev D;
    std::memset(&D, 0, sizeof(D));
    D.g = c->tiles[t].g;
    D.m = c->geo;
    for (int a = 0; a < NUM_ARRAYS; a++) D.f[a] = c->tiles[t].f[a];
    D.dt = dt;
    D.dtc_safe = c->deck.dtc_safe;
    D.dtu_safe = c->deck.dtu_safe;
    D.dtv_safe = c->deck.dtv_safe;
    D.dtw_safe = c->deck.dtw_safe;
    D.dtdiv_safe = c->deck.dtdiv_safe;
    return D;
}

static void tiles_table_init(clover_chunk *c)
{
    const int T = (int)c->tiles.size();
    std::vector<DeckDev> h(T);
    for (int t = 0; t < T; t++) h[t] = tile_entry(c, t, 0.0);
    c->tdev = reinterpret_cast<DeckDev *>(dev_alloc_bytes(c, sizeof(DeckDev) * T));
    hip_check(hipMemcpyAsync(c->tdev, h.data(), sizeof(DeckDev) * T, hipMemcpyHostToDevice, c->st), "tile table");
    hip_check(hipStreamSynchronize(c->st), "sync");  // h dies here
}

// prefix table of batched kernel k with `planes` over the tiles (built once)
static const BatchGrid &tgrid(clover_chunk *c, int k, int planes)
{
    auto key = std::make_pair(k, planes);
    auto it = c->grids.find(key);
    if (it != c->grids.end()) return it->second;
    const int T = (int)c->tiles.size();
    std::vector<int> start(T + 1, 0);
    for (int t = 0; t < T; t++) start[t + 1] = start[t] + batch_blocks((BatchKernel)k, c->tiles[t].g, planes, c->store);
    int *p = nullptr;
    hip_check(hipMalloc(&p, sizeof(int) * (T + 1)), "hipMalloc");
    c->tables.push_back(p);
    c->bytes += (long long)sizeof(int) * (T + 1);
    hip_check(hipMemcpyAsync(p, start.data(), sizeof(int) * (T + 1), hipMemcpyHostToDevice, c->st), "prefix table");
    hip_check(hipStreamSynchronize(c->st), "sync");  // `start` dies here
    return c->grids[key] = BatchGrid{p, start[T]};
}

// the shape index of a batched kernel (none: -1)
static int shape_of(int bk)
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

static const BatchGrid &kgrid(clover_chunk *c, int bk)
{
    const int si = shape_of(bk);
    return tgrid(c, bk, si < 0 ? 0 : c->shapes.k[si].planes);
}

// exchange array roles in every tile, on the host and in the device table
static void tswap(clover_chunk *c, std::initializer_list<std::pair<int, int>> pairs)
{
    SlotSwap sw{};
    for (auto &pr : pairs) {
        for (auto &t : c->tiles) std::swap(t.f[pr.first], t.f[pr.second]);
        sw.a[sw.n] = pr.first;
        sw.b[sw.n] = pr.second;
        sw.n++;
    }
    if (c->batched) hip_check(batch_swap(c->tdev, (int)c->tiles.size(), sw, c->st), "swap");
}

// update_halo over the tiles in one launch per pass: the three tile passes,
// then the reflective faces axis by axis (y, x, z)
static void batched_update_halo(clover_chunk *c, const int *fields, int depth, const int *axes, const int *raxes)
{
    const bool poison = axes && c->halo_check;
    std::string key(reinterpret_cast<const char *>(fields), sizeof(int) * NUM_FIELDS);
    key += std::string(1, (char)depth) + std::string(1, (char)poison);
    if (axes) key += std::string(reinterpret_cast<const char *>(axes), sizeof(int) * NUM_FIELDS);
    if (raxes) key += std::string(reinterpret_cast<const char *>(raxes), sizeof(int) * NUM_FIELDS);
    auto it = c->halo_tabs.find(key);
    if (it == c->halo_tabs.end()) {
        const int T = (int)c->tiles.size();
        std::vector<int> nbr(6 * T);
        std::vector<Lay> lays(T);
        for (int t = 0; t < T; t++) {
            for (int f = 0; f < 6; f++) nbr[6 * t + f] = c->tiles[t].nbr[f];
            lays[t] = c->tiles[t].g;
        }
        std::vector<TileHaloItem> copy[3], refl[3][2];
        tile_halo_items(nbr.data(), T, lays.data(), fields, depth, axes, raxes ? raxes : axes, poison, copy, refl);
        clover_chunk::HaloTabs tab{};
        std::vector<TileHaloItem> all;
        static const int axis_order[3] = {1, 0, 2};
        for (int p = 0; p < 3; p++) {
            const int axis = axis_order[p];  // the passes go y, x, z too
            tab.off_copy[p] = (int)all.size();
            tab.n_copy[p] = (int)copy[p].size();
            all.insert(all.end(), copy[p].begin(), copy[p].end());
            int pa = 1, pb = 1;
            for (auto &x : copy[p]) {
                int a, b;
                tile_halo_blocks(lays[x.tile], axis, depth, a, b);
                pa = std::max(pa, a);
                pb = std::max(pb, b);
            }
            tab.pa_copy[p] = pa;
            tab.pb_copy[p] = pb;
            pa = pb = 1;
            bool split = false;
            for (int sd = 0; sd < 2; sd++) {
                tab.off_refl[p][sd] = (int)all.size();
                tab.n_refl[p][sd] = (int)refl[p][sd].size();
                all.insert(all.end(), refl[p][sd].begin(), refl[p][sd].end());
                for (auto &x : refl[p][sd]) {
                    int a, b;
                    tile_halo_blocks(lays[x.tile], axis, depth, a, b);
                    pa = std::max(pa, a);
                    pb = std::max(pb, b);
                    const Lay &g = lays[x.tile];
                    const int mn[3] = {g.x_min, g.y_min, g.z_min}, mx[3] = {g.x_max, g.y_max, g.z_max};
                    if (mx[axis] - mn[axis] + 1 < 3) split = true;
                }
            }
            tab.pa_refl[p] = pa;
            tab.pb_refl[p] = pb;
            tab.split[p] = split;
        }
        tab.dev = nullptr;
        if (!all.empty()) {
            tab.dev = reinterpret_cast<TileHaloItem *>(dev_alloc_bytes(c, sizeof(TileHaloItem) * all.size()));
            hip_check(hipMemcpyAsync(tab.dev, all.data(), sizeof(TileHaloItem) * all.size(), hipMemcpyHostToDevice,
                                     c->st),
                      "halo items");
            hip_check(hipStreamSynchronize(c->st), "sync");  // `all` dies here
        }
        it = c->halo_tabs.emplace(key, tab).first;
    }
    const clover_chunk::HaloTabs &tab = it->second;
    static const int axis_order[3] = {1, 0, 2};
    {
        Phase ph(c, P_TILE_HALO);
        for (int p = 0; p < 3; p++)
            hip_check(batch_tile_halo(c->store, true, c->tdev, tab.dev + tab.off_copy[p], tab.n_copy[p],
                                      tab.pa_copy[p], tab.pb_copy[p], axis_order[p], depth, c->st),
                      "tile halo");
    }
    Phase ph(c, P_SELF_HALO);
    for (int p = 0; p < 3; p++) {
        if (tab.split[p]) {
            for (int sd = 0; sd < 2; sd++)
                hip_check(batch_tile_halo(c->store, false, c->tdev, tab.dev + tab.off_refl[p][sd], tab.n_refl[p][sd],
                                          tab.pa_refl[p], tab.pb_refl[p], axis_order[p], depth, c->st),
                          "update_halo");
        } else {
            // the two faces' items are adjacent in the table
            hip_check(batch_tile_halo(c->store, false, c->tdev, tab.dev + tab.off_refl[p][0],
                                      tab.n_refl[p][0] + tab.n_refl[p][1], tab.pa_refl[p], tab.pb_refl[p],
                                      axis_order[p], depth, c->st),
                      "update_halo");
        }
    }
}

// ------------------------------------------------------------ phases
static void set_fields(int *fields, std::initializer_list<int> ids)
{
    for (int f = 0; f < NUM_FIELDS; f++) fields[f] = 0;
    for (int id : ids) fields[id] = 1;
}

// axes: null, or the per-field axis masks of a fused-step plan (driver.h);
// raxes: the masks for the reflective faces (null: axes)
static void update_halo(clover_chunk *c, const int *fields, int depth, const int *axes = nullptr,
                        const int *raxes = nullptr)
{
    if (!raxes) raxes = axes;
    if (c->batched) {
        batched_update_halo(c, fields, depth, axes, raxes);
        return;
    }
    {
        Phase ph(c, P_TILE_HALO);
        // update_tile_halo.f90:47-423: top/bottom for all tiles, then
        // left/right, then front/back
        static const int passes[3][2] = {{4, 3}, {1, 2}, {6, 5}};
        for (int p = 0; p < 3; p++)
            for (size_t t = 0; t < c->tiles.size(); t++)
                for (int s = 0; s < 2; s++) {
                    const int face = passes[p][s];
                    Tile &tt = c->tiles[t];
                    const int n = tt.nbr[face - 1];
                    if (n == -1) continue;
                    Tile &nt = c->tiles[n - 1];
                    void *fp[NUM_FIELDS], *np[NUM_FIELDS];
                    for (int f = 0; f < NUM_FIELDS; f++) {
                        fp[f] = tt.f[f];
                        np[f] = nt.f[f];
                    }
                    hip_check(launch_tile_halo_store(c->store, face, tt.g, fp, nt.g, np, fields, depth, c->st, axes,
                                                     axes && c->halo_check),
                              "tile_halo");
                }
    }
    Phase ph(c, P_SELF_HALO);
    for (auto &t : c->tiles) {
        int ext[6];
        for (int f = 0; f < 6; f++) ext[f] = t.nbr[f] == -1;  // one chunk: every chunk face is external
        void *fp[NUM_FIELDS];
        for (int f = 0; f < NUM_FIELDS; f++) fp[f] = t.f[f];
        hip_check(launch_update_halo_store(c->store, t.g, fp, fields, ext, depth, c->st, raxes,
                                           axes && c->halo_check),
                  "update_halo");
    }
}

// an exchange of the fused step, reduced to the halos its kernels read
// (the reference's own exchange where a tile is under 3 cells thick)
static void fused_halo(clover_chunk *c, int which)
{
    bool thin = false;
    for (auto &t : c->tiles) thin = thin || t.nx < 3 || t.ny < 3 || t.nz < 3;
    if (thin) {
        const HaloPlan p = reference_halo_plan(which);
        update_halo(c, p.fields, p.depth, nullptr);
        return;
    }
    // every external face of every tile is mirrored (Lay::mirror, set
    // where no tile is thin)
    bool mirrored = false;
    for (auto &t : c->tiles) mirrored = mirrored || t.g.mirror != 0;
    // the single-tile FP32 kernels read vol_flux_y/z's x halos as the mirrored columns (launch.h tan_xm)
    const HaloPlan p =
        fused_halo_plan(which, c->advect_x != 0, mirrored ? AX_ALL : 0, c->store == ST_F32 && !c->batched);
    update_halo(c, p.fields, p.depth, p.axes, p.raxes);
}

// need_sound: the caller reads soundspeed too (a fetch); the fused kernels
// form it from density and pressure where they need it
static void ideal_gas(clover_chunk *c, bool predict, bool need_sound = false)
{
    // the fused path's last cell sweep leaves the new level-0 fields'
    // pressure behind; computing it again would write the same bits
    if (c->fast && !predict && c->gas_fresh && (c->sound_fresh || !need_sound)) return;
    c->gas_fresh = !predict;
    c->sound_fresh = !predict;
    Phase ph(c, P_IDEAL_GAS);
    for (auto &t : c->tiles) {
        if (c->fast && !predict)
            hip_check(launch_fast_ideal_gas(c->store, t.g, t.f[F_DENSITY0], t.f[F_ENERGY0], t.f[F_PRESSURE],
                                            t.f[F_SOUNDSPEED], c->st),
                      "ideal_gas");
        else if (!predict)
            hip_check(launch_ideal_gas(t.g, t.f[F_DENSITY0], t.f[F_ENERGY0], t.f[F_PRESSURE], t.f[F_SOUNDSPEED], c->st),
                      "ideal_gas");
        else
            hip_check(launch_ideal_gas(t.g, t.f[F_DENSITY1], t.f[F_ENERGY1], t.f[F_PRESSURE], t.f[F_SOUNDSPEED], c->st),
                      "ideal_gas");
    }
}

// timestep.f90:28-128.  Returns 1 on a small time step.
static int timestep(clover_chunk *c)
{
    int fields[NUM_FIELDS];
    const clover_deck &d = c->deck;
    const int T = (int)c->tiles.size();
    c->dt = c->g_big;
    ideal_gas(c, false);
    if (c->fast) {
        fused_halo(c, FH_TIMESTEP);
    } else {
        set_fields(fields, {F_PRESSURE, F_ENERGY0, F_DENSITY0, F_XVEL0, F_YVEL0});
        update_halo(c, fields, 1);
    }
    int nparts[64] = {0};
    // every tile in one launch (batched) when the tiles' calc_dt partials at
    // the current shape fit the partial buffer; each tile's partials are
    // reduced to its own minimum, as the per-tile final pass does
    if (c->batched && kgrid(c, BK_VISC_DT).total <= T * kMaxPartials * 5) {
        Phase ph(c, P_VISCOSITY);
        const BatchGrid &vg = kgrid(c, BK_VISC_DT);
        hip_check(batch_visc_dt(c->store, c->tdev, T, vg, c->shapes.k[SK_VISC_DT], c->g_small, c->partial, c->st),
                  "visc_dt");
        hip_check(batch_tile_min(T, vg.start, c->partial, c->g_big, c->scalar, 8, c->st), "calc_dt min");
    } else if (c->fast) {
        // viscosity and calc_dt in one pass: calc_dt reads viscosity only at
        // the cell itself, which the viscosity halo exchange does not change
        Phase ph(c, P_VISCOSITY);
        for (int ti = 0; ti < T; ti++) {
            Tile &t = c->tiles[ti];
            hip_check(launch_fast_visc_dt(c->store, t.g, c->geo, d.dtc_safe, d.dtu_safe, d.dtv_safe, d.dtw_safe, d.dtdiv_safe,
                                          c->g_small, t.f[F_DENSITY0], t.f[F_PRESSURE],
                                          t.f[F_VISCOSITY], t.f[F_XVEL0], t.f[F_YVEL0], t.f[F_ZVEL0],
                                          c->partial + (size_t)ti * kMaxPartials * 5, &nparts[ti % 64],
                                          c->shapes.k[SK_VISC_DT], c->st),
                      "visc_dt");
            hip_check(launch_min_final(c->partial + (size_t)ti * kMaxPartials * 5, nparts[ti % 64], c->g_big,
                                       c->scalar + 8 * ti, c->st),
                      "calc_dt min");
        }
    } else {
        Phase ph(c, P_VISCOSITY);
        for (auto &t : c->tiles)
            hip_check(launch_viscosity(t.g, t.f[X_XAREA], t.f[X_YAREA], t.f[X_ZAREA], t.celldx, t.celldy, t.celldz,
                                       t.f[F_DENSITY0], t.f[F_PRESSURE], t.f[F_VISCOSITY], t.f[F_XVEL0],
                                       t.f[F_YVEL0], t.f[F_ZVEL0], c->st),
                      "viscosity");
    }
    if (c->fast) {
        fused_halo(c, FH_VISCOSITY);
    } else {
        set_fields(fields, {F_VISCOSITY});
        update_halo(c, fields, 1);
    }
    {
        Phase ph(c, P_TIMESTEP);
        if (!c->fast) {
            for (int ti = 0; ti < T; ti++) {
                Tile &t = c->tiles[ti];
                int np = 0;
                double *part = c->partial + (size_t)ti * kMaxPartials * 5;
                hip_check(launch_calc_dt_cells(t.g, d.dtc_safe, d.dtu_safe, d.dtv_safe, d.dtw_safe, d.dtdiv_safe,
                                               c->g_small, t.f[X_XAREA], t.f[X_YAREA], t.f[X_ZAREA], t.celldx,
                                               t.celldy, t.celldz, t.f[X_VOLUME], t.f[F_DENSITY0], t.f[F_VISCOSITY],
                                               t.f[F_SOUNDSPEED], t.f[F_XVEL0], t.f[F_YVEL0], t.f[F_ZVEL0], nullptr,
                                               part, &np, c->st),
                          "calc_dt");
                hip_check(launch_min_final(part, np, c->g_big, c->scalar + 8 * ti, c->st), "calc_dt min");
            }
        }
        hip_check(hipMemcpyAsync(c->h_scalar, c->scalar, sizeof(double) * 8 * T, hipMemcpyDeviceToHost, c->st),
                  "dt D2H");
        hip_check(hipStreamSynchronize(c->st), "sync");
        // timestep.f90:93-100: the minimum over tiles, ties to the last tile
        for (int ti = 0; ti < T; ti++) {
            const double dtlp = c->h_scalar[8 * ti];
            if (dtlp <= c->dt) {
                c->dt = dtlp;
                c->dt_tile = ti;
            }
        }
    }
    double m = c->dt;
    if (c->dtold * d.dtrise < m) m = c->dtold * d.dtrise;
    if (d.dtmax < m) m = d.dtmax;
    c->dt = m;
    if (c->dt < d.dtmin) return 1;
    c->dtold = c->dt;
    if (c->batched) hip_check(batch_set_dt(c->tdev, T, c->dt, c->st), "set dt");
    return 0;
}

static void pdv(clover_chunk *c, bool predict)
{
    if (c->fast && predict) {
        // PdV predictor + ideal_gas on the predicted state; revert is a no-op
        // here because the predicted density/energy are never stored
        {
            Phase ph(c, P_PDV);
            if (c->batched) {
                hip_check(batch_pdv_predict(c->store, c->tdev, (int)c->tiles.size(), kgrid(c, BK_PDV_PREDICT),
                                            c->shapes.k[SK_PDV_PREDICT], c->st),
                          "PdV predict");
            } else {
                for (auto &t : c->tiles)
                    hip_check(launch_fast_pdv_predict(c->store, t.g, c->geo, c->dt, t.f[F_DENSITY0],
                                                      t.f[F_ENERGY0], t.f[F_PRESSURE], t.f[F_VISCOSITY],
                                                      t.f[F_XVEL0], t.f[F_YVEL0], t.f[F_ZVEL0],
                                                      c->shapes.k[SK_PDV_PREDICT], c->st),
                              "PdV predict");
            }
            c->gas_fresh = false;  // pressure now holds the predicted state's
        }
        fused_halo(c, FH_PRESSURE);
        return;
    }
    {
        Phase ph(c, P_PDV);
        for (auto &t : c->tiles)
            hip_check(launch_pdv(predict ? 1 : 0, t.g, c->dt, t.f[X_XAREA], t.f[X_YAREA], t.f[X_ZAREA],
                                 t.f[X_VOLUME], t.f[F_DENSITY0], t.f[F_DENSITY1], t.f[F_ENERGY0], t.f[F_ENERGY1],
                                 t.f[F_PRESSURE], t.f[F_VISCOSITY], t.f[F_XVEL0], t.f[F_XVEL1], t.f[F_YVEL0],
                                 t.f[F_YVEL1], t.f[F_ZVEL0], t.f[F_ZVEL1], c->st),
                      "PdV");
    }
    if (predict) {
        int fields[NUM_FIELDS];
        ideal_gas(c, true);
        set_fields(fields, {F_PRESSURE});
        update_halo(c, fields, 1);
        Phase ph(c, P_REVERT);
        for (auto &t : c->tiles)
            hip_check(launch_revert(t.g, t.f[F_DENSITY0], t.f[F_DENSITY1], t.f[F_ENERGY0], t.f[F_ENERGY1], c->st),
                      "revert");
    }
}

static void advec_cell(clover_chunk *c, int sweep, int dir)
{
    Phase ph(c, P_CELL_ADVEC);
    if (c->batched) {
        const int k = dir == 1 ? BK_CELL_X : dir == 2 ? BK_CELL_Y : BK_CELL_Z;
        hip_check(batch_advec_cell(c->store, c->tdev, (int)c->tiles.size(), kgrid(c, k), c->advect_x, dir, sweep,
                                   sweep == 3, c->shapes.k[shape_of(k)], c->st),
                  "advec_cell");
        if (sweep == 3) {
            c->gas_fresh = true;
            c->sound_fresh = false;
        }
        tswap(c, {{F_DENSITY1, X_W1}, {F_ENERGY1, X_W2}});
        return;
    }
    for (auto &t : c->tiles) {
        if (c->fast) {
            // new density1/energy1 land in spare arrays, which then take their
            // place (interior cells only; the halo update that follows every
            // advec_cell sweep refreshes the halos before anything reads them)
            // the third sweep also writes the next step's ideal_gas pressure
            const bool last = sweep == 3;
            hip_check(launch_fast_advec_cell(c->store, t.g, c->geo, c->advect_x, dir, sweep, t.f[F_DENSITY1], t.f[F_ENERGY1],
                                             t.f[F_MASS_FLUX_X], t.f[F_VOL_FLUX_X], t.f[F_MASS_FLUX_Y],
                                             t.f[F_VOL_FLUX_Y], t.f[F_MASS_FLUX_Z], t.f[F_VOL_FLUX_Z], t.f[X_W1],
                                             t.f[X_W2], last ? t.f[F_PRESSURE] : nullptr,
                                             c->shapes.k[SK_CELL_X + dir - 1], c->st),
                      "advec_cell");
            if (last) {  // for the level 0 that reset_field makes of these
                c->gas_fresh = true;
                c->sound_fresh = false;
            }
            std::swap(t.f[F_DENSITY1], t.f[X_W1]);
            std::swap(t.f[F_ENERGY1], t.f[X_W2]);
        } else
            hip_check(launch_advec_cell(t.g, c->advect_x, dir, sweep, t.vertexdx, t.vertexdy, t.vertexdz,
                                        t.f[X_VOLUME], t.f[F_DENSITY1], t.f[F_ENERGY1], t.f[F_MASS_FLUX_X],
                                        t.f[F_VOL_FLUX_X], t.f[F_MASS_FLUX_Y], t.f[F_VOL_FLUX_Y], t.f[F_MASS_FLUX_Z],
                                        t.f[F_VOL_FLUX_Z], t.f[X_W1], t.f[X_W2], t.f[X_W3], t.f[X_W4], t.f[X_W5],
                                        t.f[X_W6], t.f[X_W7], c->st),
                      "advec_cell");
    }
}

static void advec_mom(clover_chunk *c, int sweep, int dir)
{
    Phase ph(c, P_MOM_ADVEC);
    if (c->batched) {
        const int k = dir == 1 ? BK_MOM_X : dir == 2 ? BK_MOM_Y : BK_MOM_Z;
        hip_check(batch_advec_mom(c->store, c->tdev, (int)c->tiles.size(), kgrid(c, k), c->advect_x, dir, sweep,
                                  c->shapes.k[shape_of(k)], c->st),
                  "advec_mom");
        tswap(c, {{F_XVEL1, X_W5}, {F_YVEL1, X_W6}, {F_ZVEL1, X_W7}});
        return;
    }
    for (auto &t : c->tiles) {
        if (c->fast) {
            // the three velocity calls of advec_mom_driver.f90 in one pass; the
            // new velocities land in the spare arrays, which then take the
            // velocities' place (only nodes min..max+1 are new; the halos are
            // refreshed by the next update_halo before anything reads them)
            const void *vin[3] = {t.f[F_XVEL1], t.f[F_YVEL1], t.f[F_ZVEL1]};
            void *vout[3] = {t.f[X_W5], t.f[X_W6], t.f[X_W7]};
            hip_check(launch_fast_advec_mom(c->store, t.g, c->geo, c->advect_x, sweep, dir, t.f[F_MASS_FLUX_X],
                                            t.f[F_VOL_FLUX_X], t.f[F_MASS_FLUX_Y], t.f[F_VOL_FLUX_Y],
                                            t.f[F_MASS_FLUX_Z], t.f[F_VOL_FLUX_Z], t.f[F_DENSITY1], vin, vout,
                                            c->shapes.k[SK_MOM_X + dir - 1], c->st),
                      "advec_mom");
            std::swap(t.f[F_XVEL1], t.f[X_W5]);
            std::swap(t.f[F_YVEL1], t.f[X_W6]);
            std::swap(t.f[F_ZVEL1], t.f[X_W7]);
            continue;
        }
        double *vel[3] = {t.f[F_XVEL1], t.f[F_YVEL1], t.f[F_ZVEL1]};
        for (int which = 1; which <= 3; which++)
            hip_check(launch_advec_mom(t.g, vel[which - 1], t.f[F_MASS_FLUX_X], t.f[F_VOL_FLUX_X],
                                       t.f[F_MASS_FLUX_Y], t.f[F_VOL_FLUX_Y], t.f[F_MASS_FLUX_Z], t.f[F_VOL_FLUX_Z],
                                       t.f[X_VOLUME], t.f[F_DENSITY1], t.f[X_W1], t.f[X_W2], t.f[X_W3], t.f[X_W5],
                                       t.f[X_W6], t.f[X_W7], t.celldx, t.celldy, t.celldz, c->advect_x, which, sweep,
                                       dir, c->st),
                      "advec_mom");
    }
}

// advection.f90:26-181
static void advection(clover_chunk *c)
{
    int fc[NUM_FIELDS], fm[NUM_FIELDS];
    set_fields(fc, {F_ENERGY1, F_DENSITY1, F_VOL_FLUX_X, F_VOL_FLUX_Y, F_VOL_FLUX_Z});
    set_fields(fm, {F_DENSITY1, F_ENERGY1, F_XVEL1, F_YVEL1, F_ZVEL1, F_MASS_FLUX_X, F_MASS_FLUX_Y, F_MASS_FLUX_Z});
    // the fused step refreshes only what its kernels read (driver.h)
    auto exchange = [&](const int *fields, int which) {
        if (c->fast) fused_halo(c, which);
        else update_halo(c, fields, 2);
    };
    int dir = c->advect_x ? 1 : 3;
    exchange(fc, FH_CELL1);
    advec_cell(c, 1, dir);
    exchange(fm, FH_MOM1);
    advec_mom(c, 1, dir);
    exchange(fc, FH_CELL2);
    advec_cell(c, 2, 2);
    exchange(fm, FH_MOM2);
    advec_mom(c, 2, 2);
    dir = c->advect_x ? 3 : 1;
    advec_cell(c, 3, dir);
    exchange(fm, FH_MOM3);
    advec_mom(c, 3, dir);
}

static void field_summary(clover_chunk *c, double *out)
{
    ideal_gas(c, false);
    Phase ph(c, P_SUMMARY);
    const int T = (int)c->tiles.size();
    for (int ti = 0; ti < T; ti++) {
        Tile &t = c->tiles[ti];
        int np = 0;
        double *part = c->partial + (size_t)ti * kMaxPartials * 5;
        if (c->fast)
            hip_check(launch_fast_field_summary(c->store, t.g, c->geo, t.f[F_DENSITY0], t.f[F_ENERGY0], t.f[F_PRESSURE],
                                                t.f[F_XVEL0], t.f[F_YVEL0], t.f[F_ZVEL0], part, &np, c->st),
                      "field_summary");
        else
            hip_check(launch_field_summary_partial(t.g, t.f[X_VOLUME], t.f[F_DENSITY0], t.f[F_ENERGY0],
                                                   t.f[F_PRESSURE], t.f[F_XVEL0], t.f[F_YVEL0], t.f[F_ZVEL0], part,
                                                   &np, c->st),
                      "field_summary");
        hip_check(launch_sum5_final(part, np, c->scalar + 8 * ti, c->st), "field_summary sum");
    }
    hip_check(hipMemcpyAsync(c->h_scalar, c->scalar, sizeof(double) * 8 * T, hipMemcpyDeviceToHost, c->st), "D2H");
    hip_check(hipStreamSynchronize(c->st), "sync");
    double s[5] = {0, 0, 0, 0, 0};
    for (int ti = 0; ti < T; ti++)
        for (int q = 0; q < 5; q++) s[q] = s[q] + c->h_scalar[8 * ti + q];
    for (int q = 0; q < 5; q++) out[q] = s[q];
}

// visit.f90:65-87: the device work before the dump: ideal_gas on every tile,
// the pressure and level-0 velocity halos at depth 1, then viscosity.  The
// next timestep recomputes the same values from the same inputs, so the run
// continues unchanged.
static void visit_prepare(clover_chunk *c)
{
    ideal_gas(c, false);
    int fields[NUM_FIELDS];
    set_fields(fields, {F_PRESSURE, F_XVEL0, F_YVEL0, F_ZVEL0});
    update_halo(c, fields, 1);
    const clover_deck &d = c->deck;
    Phase ph(c, P_VISCOSITY);
    for (size_t ti = 0; ti < c->tiles.size(); ti++) {
        Tile &t = c->tiles[ti];
        if (c->fast) {
            // the fused kernel's viscosity is bit-identical to the per-call
            // kernel's; its dt partials are overwritten by the next timestep
            int np = 0;
            hip_check(launch_fast_visc_dt(c->store, t.g, c->geo, d.dtc_safe, d.dtu_safe, d.dtv_safe, d.dtw_safe,
                                          d.dtdiv_safe, c->g_small, t.f[F_DENSITY0], t.f[F_PRESSURE],
                                          t.f[F_VISCOSITY], t.f[F_XVEL0], t.f[F_YVEL0],
                                          t.f[F_ZVEL0], c->partial + ti * kMaxPartials * 5, &np,
                                          c->shapes.k[SK_VISC_DT], c->st),
                      "visc_dt");
        } else {
            hip_check(launch_viscosity(t.g, t.f[X_XAREA], t.f[X_YAREA], t.f[X_ZAREA], t.celldx, t.celldy, t.celldz,
                                       t.f[F_DENSITY0], t.f[F_PRESSURE], t.f[F_VISCOSITY], t.f[F_XVEL0],
                                       t.f[F_YVEL0], t.f[F_ZVEL0], c->st),
                      "viscosity");
        }
    }
    hip_check(hipStreamSynchronize(c->st), "sync");
}

// ---------------------------------------------------------------- autotuner
// Launches shaped kernel k over every tile with its current shape, writing
// only the spare work arrays: pdv_predict updates pressure in place, so it
// runs on a copy of pressure in work array 3 (the copy is made by the caller,
// outside the timing); every other output goes to W1..W7.  The inputs are
// the chunk's own fields as a step leaves them.
static void tune_launch(clover_chunk *c, int k)
{
    for (size_t ti = 0; ti < c->tiles.size(); ti++) {
        Tile &t = c->tiles[ti];
        const Shape &sh = c->shapes.k[k];
        const clover_deck &d = c->deck;
        hipError_t e = hipSuccess;
        switch (k) {
        case SK_VISC_DT: {
            int np = 0;
            e = launch_fast_visc_dt(c->store, t.g, c->geo, d.dtc_safe, d.dtu_safe, d.dtv_safe, d.dtw_safe,
                                    d.dtdiv_safe, c->g_small, t.f[F_DENSITY0], t.f[F_PRESSURE],
                                    t.f[X_W3], t.f[F_XVEL0], t.f[F_YVEL0], t.f[F_ZVEL0],
                                    c->partial + ti * kMaxPartials * 5, &np, sh, c->st);
            // the shape sets the number of partials, so their minimum is
            // part of its cost (the value lands where the next timestep
            // writes its own)
            if (e == hipSuccess)
                e = launch_min_final(c->partial + ti * kMaxPartials * 5, np, c->g_big, c->scalar + 8 * ti, c->st);
            break;
        }
        case SK_PDV_PREDICT:
            e = launch_fast_pdv_predict(c->store, t.g, c->geo, c->dtold, t.f[F_DENSITY0], t.f[F_ENERGY0], t.f[X_W3],
                                        t.f[F_VISCOSITY], t.f[F_XVEL0], t.f[F_YVEL0], t.f[F_ZVEL0], sh, c->st);
            break;
        case SK_ACCEL_PDV:
            e = launch_fast_accel_pdv(c->store, t.g, c->geo, c->dtold, t.f[F_DENSITY0], t.f[X_W1], t.f[F_ENERGY0],
                                      t.f[X_W2], t.f[F_PRESSURE], t.f[F_VISCOSITY], t.f[F_XVEL0], t.f[X_W5],
                                      t.f[F_YVEL0], t.f[X_W6], t.f[F_ZVEL0], t.f[X_W7], t.f[X_W3], t.f[X_W4],
                                      t.f[X_W3], sh, c->st);
            break;
        case SK_CELL_X:
        case SK_CELL_Y:
        case SK_CELL_Z: {
            const int dir = k - SK_CELL_X + 1, sweep = dir == 2 ? 2 : 1;
            e = launch_fast_advec_cell(c->store, t.g, c->geo, 1, dir, sweep, t.f[F_DENSITY1], t.f[F_ENERGY1],
                                       t.f[X_W3], t.f[F_VOL_FLUX_X], t.f[X_W3], t.f[F_VOL_FLUX_Y], t.f[X_W3],
                                       t.f[F_VOL_FLUX_Z], t.f[X_W1], t.f[X_W2], nullptr, sh, c->st);
            break;
        }
        default: {
            const int dir = k - SK_MOM_X + 1, sweep = dir == 2 ? 2 : 1;
            const void *vin[3] = {t.f[F_XVEL1], t.f[F_YVEL1], t.f[F_ZVEL1]};
            void *vout[3] = {t.f[X_W5], t.f[X_W6], t.f[X_W7]};
            e = launch_fast_advec_mom(c->store, t.g, c->geo, 1, sweep, dir, t.f[F_MASS_FLUX_X], t.f[F_VOL_FLUX_X],
                                      t.f[F_MASS_FLUX_Y], t.f[F_VOL_FLUX_Y], t.f[F_MASS_FLUX_Z], t.f[F_VOL_FLUX_Z],
                                      t.f[F_DENSITY1], vin, vout, sh, c->st);
            break;
        }
        }
        hip_check(e, "autotune launch");
    }
}

// The batched twin of tune_launch for a chunk whose stages run as batched
// launches: the kernel runs over every tile with a copy of the tile table
// whose output slots point at spare arrays (the same ones tune_launch
// writes), so the timed launch is the one a step makes and every field keeps
// its value.  `ttab` is that table (device, one entry per tile).
static void tune_launch_batched(clover_chunk *c, int k, DeckDev *ttab)
{
    const int T = (int)c->tiles.size();
    const Shape &sh = c->shapes.k[k];
    const int bk = k == SK_VISC_DT ? BK_VISC_DT : k == SK_PDV_PREDICT ? BK_PDV_PREDICT
                 : k == SK_ACCEL_PDV ? BK_ACCEL_PDV : k <= SK_CELL_Z ? BK_CELL_X + (k - SK_CELL_X)
                 : BK_MOM_X + (k - SK_MOM_X);
    const BatchGrid &g = tgrid(c, bk, sh.planes);
    hipError_t e = hipSuccess;
    switch (k) {
    case SK_VISC_DT:
        if (g.total > T * kMaxPartials * 5) {  // the step takes the per-tile launches then
            tune_launch(c, k);
            return;
        }
        e = batch_visc_dt(c->store, ttab, T, g, sh, c->g_small, c->partial, c->st);
        if (e == hipSuccess) e = batch_tile_min(T, g.start, c->partial, c->g_big, c->scalar, 8, c->st);
        break;
    case SK_PDV_PREDICT: e = batch_pdv_predict(c->store, ttab, T, g, sh, c->st); break;
    case SK_ACCEL_PDV: e = batch_accel_pdv(c->store, ttab, T, g, sh, c->st); break;
    case SK_CELL_X:
    case SK_CELL_Y:
    case SK_CELL_Z: {
        const int dir = k - SK_CELL_X + 1;
        e = batch_advec_cell(c->store, ttab, T, g, 1, dir, dir == 2 ? 2 : 1, false, sh, c->st);
        break;
    }
    default: {
        const int dir = k - SK_MOM_X + 1;
        e = batch_advec_mom(c->store, ttab, T, g, 1, dir, dir == 2 ? 2 : 1, sh, c->st);
        break;
    }
    }
    hip_check(e, "autotune launch");
}

// the tile table with kernel k's outputs sent to the spare arrays
// tune_launch uses (the batched cell and momentum sweeps write their new
// fields to W1/W2 and W5..W7 anyway)
static std::vector<DeckDev> tune_table(clover_chunk *c, int k)
{
    std::vector<DeckDev> h;
    for (int t = 0; t < (int)c->tiles.size(); t++) h.push_back(tile_entry(c, t, c->dtold));
    for (auto &D : h) {
        switch (k) {
        case SK_VISC_DT: D.f[F_VISCOSITY] = D.f[X_W3]; break;
        case SK_PDV_PREDICT: D.f[F_PRESSURE] = D.f[X_W3]; break;  // a copy of pressure (tune_time)
        case SK_ACCEL_PDV:
            D.f[F_DENSITY1] = D.f[X_W1];
            D.f[F_ENERGY1] = D.f[X_W2];
            D.f[F_XVEL1] = D.f[X_W5];
            D.f[F_YVEL1] = D.f[X_W6];
            D.f[F_ZVEL1] = D.f[X_W7];
            D.f[F_VOL_FLUX_X] = D.f[X_W3];
            D.f[F_VOL_FLUX_Y] = D.f[X_W4];
            D.f[F_VOL_FLUX_Z] = D.f[X_W3];
            break;
        case SK_CELL_X: D.f[F_MASS_FLUX_X] = D.f[X_W3]; break;
        case SK_CELL_Y: D.f[F_MASS_FLUX_Y] = D.f[X_W3]; break;
        case SK_CELL_Z: D.f[F_MASS_FLUX_Z] = D.f[X_W3]; break;
        default: break;
        }
    }
    return h;
}

// average milliseconds of one tune_launch of kernel k (after one untimed run)
static double tune_time(clover_chunk *c, int k, hipEvent_t a, hipEvent_t b, DeckDev *ttab = nullptr)
{
    const int reps = 3;
    if (k == SK_PDV_PREDICT)
        for (auto &t : c->tiles)
            hip_check(hipMemcpyAsync(t.f[X_W3], t.f[F_PRESSURE], (size_t)lay_elems(t.g) * store_bytes(c->store),
                                     hipMemcpyDeviceToDevice, c->st),
                      "autotune copy");
    auto launch = [&]() {
        if (ttab) tune_launch_batched(c, k, ttab);
        else tune_launch(c, k);
    };
    launch();
    hip_check(hipEventRecord(a, c->st), "event");
    for (int r = 0; r < reps; r++) launch();
    hip_check(hipEventRecord(b, c->st), "event");
    hip_check(hipEventSynchronize(b), "event sync");
    float ms = 0;
    hip_check(hipEventElapsedTime(&ms, a, b), "event time");
    return ms / reps;
}

// Every shaped kernel, every candidate shape: positions per workgroup along
// its march from kTunePlanes (the x cell sweep: rows per wave, cell_x_rows) times the two
// block orders, and for the z marches times the workgroup widths kTuneBx;
// the fastest is kept.  Shapes split the work differently but never change a
// value, so later steps give the same bits whatever is chosen.
static bool z_march(int k) { return k == SK_VISC_DT || k == SK_PDV_PREDICT; }

static void autotune(clover_chunk *c, double *before, double *after)
{
    hipEvent_t a = nullptr, b = nullptr;
    hip_check(hipEventCreate(&a), "event");
    hip_check(hipEventCreate(&b), "event");
    double t0 = 0.0, t1 = 0.0;
    // batched chunks time the batched launches, through a tile table whose
    // output slots point at spare arrays (their z marches have one
    // workgroup width)
    DeckDev *ttab = nullptr;
    if (c->batched) hip_check(hipMalloc(&ttab, sizeof(DeckDev) * c->tiles.size()), "hipMalloc");
    for (int k = 0; k < SK_COUNT; k++) {
        if (ttab) {
            const std::vector<DeckDev> h = tune_table(c, k);
            hip_check(hipMemcpyAsync(ttab, h.data(), sizeof(DeckDev) * h.size(), hipMemcpyHostToDevice, c->st),
                      "tune table");
            hip_check(hipStreamSynchronize(c->st), "sync");
        }
        const Shape start = c->shapes.k[k];
        double best = tune_time(c, k, a, b, ttab);
        Shape best_shape = start;
        t0 += best;
        for (int planes : kTunePlanes)
            for (int xcd = 0; xcd < 2; xcd++)
                for (int bx : kTuneBx) {
                    if ((ttab || !z_march(k)) && bx != start.bx) continue;
                    const Shape cand{planes, xcd, bx};
                    if (cand.planes == start.planes && cand.xcd == start.xcd && cand.bx == start.bx) continue;
                    c->shapes.k[k] = cand;
                    const double ms = tune_time(c, k, a, b, ttab);
                    if (ms < best) {
                        best = ms;
                        best_shape = cand;
                    }
                }
        c->shapes.k[k] = best_shape;
        t1 += best;
    }
    if (ttab) (void)hipFree(ttab);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (before) *before = t0;
    if (after) *after = t1;
}

static void do_step(clover_chunk *c)
{
    pdv(c, true);
    if (c->fast) {
        // accelerate, the PdV corrector and flux_calc in one march
        Phase ph(c, P_ACCEL);
        if (c->batched) {
            hip_check(batch_accel_pdv(c->store, c->tdev, (int)c->tiles.size(), kgrid(c, BK_ACCEL_PDV),
                                      c->shapes.k[SK_ACCEL_PDV], c->st),
                      "accelerate + PdV correct + flux_calc");
        } else {
            for (auto &t : c->tiles)
                hip_check(launch_fast_accel_pdv(c->store, t.g, c->geo, c->dt, t.f[F_DENSITY0], t.f[F_DENSITY1],
                                                t.f[F_ENERGY0], t.f[F_ENERGY1], t.f[F_PRESSURE], t.f[F_VISCOSITY],
                                                t.f[F_XVEL0], t.f[F_XVEL1], t.f[F_YVEL0], t.f[F_YVEL1],
                                                t.f[F_ZVEL0], t.f[F_ZVEL1], t.f[F_VOL_FLUX_X], t.f[F_VOL_FLUX_Y],
                                                t.f[F_VOL_FLUX_Z], c->shapes.k[SK_ACCEL_PDV], c->st),
                          "accelerate + PdV correct + flux_calc");
        }
    } else {
        {
            Phase ph(c, P_ACCEL);
            for (auto &t : c->tiles)
                hip_check(launch_accelerate(t.g, c->dt, t.f[X_XAREA], t.f[X_YAREA], t.f[X_ZAREA], t.f[X_VOLUME],
                                            t.f[F_DENSITY0], t.f[F_PRESSURE], t.f[F_VISCOSITY], t.f[F_XVEL0],
                                            t.f[F_YVEL0], t.f[F_ZVEL0], t.f[F_XVEL1], t.f[F_YVEL1], t.f[F_ZVEL1],
                                            c->st),
                          "accelerate");
        }
        pdv(c, false);
    }
    if (!c->fast) {
        Phase ph(c, P_FLUX);
        for (auto &t : c->tiles)
            hip_check(launch_flux_calc(t.g, c->dt, t.f[X_XAREA], t.f[X_YAREA], t.f[X_ZAREA], t.f[F_XVEL0],
                                       t.f[F_YVEL0], t.f[F_ZVEL0], t.f[F_XVEL1], t.f[F_YVEL1], t.f[F_ZVEL1],
                                       t.f[F_VOL_FLUX_X], t.f[F_VOL_FLUX_Y], t.f[F_VOL_FLUX_Z], c->st),
                      "flux_calc");
    }
    advection(c);
    if (c->batched) {
        // reset_field: the level 0/1 arrays exchange roles in every tile and
        // in the device table, as the per-tile fused path below does
        tswap(c, {{F_DENSITY0, F_DENSITY1}, {F_ENERGY0, F_ENERGY1}, {F_XVEL0, F_XVEL1}, {F_YVEL0, F_YVEL1},
                  {F_ZVEL0, F_ZVEL1}});
    } else {
        Phase ph(c, P_RESET);
        for (auto &t : c->tiles) {
            if (c->fast) {
                // reset_field: level 1 becomes level 0 by exchanging the
                // arrays' roles.  Interiors equal the reference's copy; halo
                // cells differ but are refreshed before any kernel reads them
                std::swap(t.f[F_DENSITY0], t.f[F_DENSITY1]);
                std::swap(t.f[F_ENERGY0], t.f[F_ENERGY1]);
                std::swap(t.f[F_XVEL0], t.f[F_XVEL1]);
                std::swap(t.f[F_YVEL0], t.f[F_YVEL1]);
                std::swap(t.f[F_ZVEL0], t.f[F_ZVEL1]);
                continue;
            }
            hip_check(launch_reset_field(t.g, t.f[F_DENSITY0], t.f[F_DENSITY1], t.f[F_ENERGY0], t.f[F_ENERGY1],
                                         t.f[F_XVEL0], t.f[F_XVEL1], t.f[F_YVEL0], t.f[F_YVEL1], t.f[F_ZVEL0],
                                         t.f[F_ZVEL1], c->st),
                      "reset_field");
        }
    }
    c->advect_x = !c->advect_x;
    c->time = c->time + c->dt;
}

static void destroy(clover_chunk *c)
{
    if (!c) return;
    if (c->st) (void)hipStreamSynchronize(c->st);
    for (auto &t : c->tiles) {
        for (int a = 0; a < NUM_ARRAYS; a++)
            if (t.f[a]) (void)hipFree(t.f[a]);
        double *oned[] = {t.cellx, t.celly, t.cellz, t.celldx, t.celldy, t.celldz, t.vertexdx, t.vertexdy, t.vertexdz};
        for (double *p : oned)
            if (p) (void)hipFree(p);
    }
    for (int *p : c->tables) (void)hipFree(p);
    for (auto &kv : c->halo_tabs)
        if (kv.second.dev) (void)hipFree(kv.second.dev);
    if (c->tdev) (void)hipFree(c->tdev);
    if (c->partial) (void)hipFree(c->partial);
    if (c->scalar) (void)hipFree(c->scalar);
    if (c->h_scalar) (void)hipHostFree(c->h_scalar);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    delete c->ktimer;
    if (c->st) (void)hipStreamDestroy(c->st);
    delete c;
}

}  // namespace clover

extern "C" {

int clover_chunk_create(const clover_deck *deck, int device, clover_chunk **out)
{
    return clover_chunk_create_ex(deck, device, 0, out);
}

int clover_chunk_create_ex(const clover_deck *deck, int device, int flags, clover_chunk **out)
{
    clover_chunk *c = nullptr;
    try {
        if (!deck || !out) throw Error(E_ARG, "null argument");
        const clover_deck &d = *deck;
        if (d.x_cells < 1 || d.y_cells < 1 || d.z_cells < 1) throw Error(E_ARG, "cells must be positive");
        if (d.number_of_states < 1 || d.number_of_states > CLOVER_MAX_STATES) throw Error(E_ARG, "bad state count");
        if (d.tiles_per_chunk < 1) throw Error(E_ARG, "tiles_per_chunk must be positive");
        if (flags & ~(CLOVER_CHUNK_PER_CALL | CLOVER_CHUNK_FP32 | CLOVER_CHUNK_HALO_CHECK))
            throw Error(E_ARG, "unknown flags");
        if ((flags & CLOVER_CHUNK_PER_CALL) && (flags & CLOVER_CHUNK_FP32))
            throw Error(E_ARG, "the FP32-store form runs the fused kernels only");
        if ((flags & CLOVER_CHUNK_PER_CALL) && (flags & CLOVER_CHUNK_HALO_CHECK))
            throw Error(E_ARG, "the halo check applies to the fused kernels only");
        c = new clover_chunk();  // value-initialised: every scalar member starts at zero
        c->deck = d;
        c->device = device;
        c->fast = !(flags & CLOVER_CHUNK_PER_CALL);
        c->store = (flags & CLOVER_CHUNK_FP32) ? ST_F32 : ST_F64;
        c->halo_check = (flags & CLOVER_CHUNK_HALO_CHECK) != 0;
        c->shapes = default_shapes();
        c->dt_tile = -1;
        // the uniform mesh as initialise_chunk.f90:37-39 and
        // initialise_chunk_kernel.f90:151-178 compute it
        c->geo = geo_of((d.xmax - d.xmin) / (double)(float)d.x_cells, (d.ymax - d.ymin) / (double)(float)d.y_cells,
                        (d.zmax - d.zmin) / (double)(float)d.z_cells);
        hip_check(hipSetDevice(device), "hipSetDevice");
        hip_check(hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking), "hipStreamCreate");
        hip_check(hipEventCreate(&c->ev0), "event");
        hip_check(hipEventCreate(&c->ev1), "event");
        c->g_small = (double)1.0e-16f;  // default-real literals (data.f90:30-31)
        c->g_big = (double)1.0e+21f;
        decompose(d, c->tiles);
        const int T = (int)c->tiles.size();
        c->partial = dev_alloc(c, (size_t)T * kMaxPartials * 5);
        c->scalar = dev_alloc(c, (size_t)T * 8);
        hip_check(hipHostMalloc(&c->h_scalar, sizeof(double) * 8 * T, hipHostMallocDefault), "hipHostMalloc");
        bool thin = false;
        for (auto &t : c->tiles) thin = thin || t.nx < 3 || t.ny < 3 || t.nz < 3;
        for (auto &t : c->tiles) {
            t.g = make_lay(1, t.nx, 1, t.ny, 1, t.nz);
            // tiles at least 3 cells thick: on their external faces the
            // fused kernels read the mirrored interior instead of the cell
            // and vertex fields' reflective halos, which the exchanges then
            // leave out (a chunk with a thinner tile takes the reference's
            // exchanges, fused_halo)
            if (c->fast && !thin)
                for (int f = 0; f < 6; f++)
                    if (t.nbr[f] == -1) t.g.mirror |= 1 << f;
            const int64_t n = lay_elems(t.g);
            if (n >= (int64_t(1) << 31)) throw Error(E_ARG, "tile too large for 32-bit element offsets");
            for (int a = 0; a < NUM_ARRAYS; a++) t.f[a] = dev_alloc_bytes(c, (size_t)n * store_bytes(c->store));
            // 1-D mesh on the host, the 3-D fields painted on the device
            HostInit h;
            host_init(d, t, h, false);
            hip_check(launch_generate(c->store, t.g, gen_state(d, t.left, t.bottom, t.back), t.f[F_DENSITY0],
                                      t.f[F_ENERGY0], t.f[F_XVEL0], t.f[F_YVEL0], t.f[F_ZVEL0], t.f[X_VOLUME],
                                      t.f[X_XAREA], t.f[X_YAREA], t.f[X_ZAREA], c->st),
                      "generate");
            t.cellx = dev_upload_1d(c, h.cellx);
            t.celly = dev_upload_1d(c, h.celly);
            t.cellz = dev_upload_1d(c, h.cellz);
            t.celldx = dev_upload_1d(c, h.celldx);
            t.celldy = dev_upload_1d(c, h.celldy);
            t.celldz = dev_upload_1d(c, h.celldz);
            t.vertexdx = dev_upload_1d(c, h.vertexdx);
            t.vertexdy = dev_upload_1d(c, h.vertexdy);
            t.vertexdz = dev_upload_1d(c, h.vertexdz);
            t.h_cellx = h.cellx;
            t.h_celly = h.celly;
            t.h_cellz = h.cellz;
            hip_check(hipStreamSynchronize(c->st), "sync");  // host buffers die here
        }
        // more than one tile: the fused step as batched launches, if every
        // tile's calc_dt partials fit the partial buffer at the shortest
        // march chunk the tuner may pick
        if (c->fast && T > 1) {
            long long need = 0;
            for (auto &t : c->tiles) need += batch_blocks(BK_VISC_DT, t.g, kTunePlanes[0], c->store);
            if (need <= (long long)T * kMaxPartials * 5) {
                c->batched = true;
                tiles_table_init(c);
            }
        }
        // start.f90:58-129
        c->time = 0.0;
        c->step = 0;
        c->dtold = d.dtinit;
        c->dt = d.dtinit;
        c->advect_x = 1;
        ideal_gas(c, false);
        int fields[NUM_FIELDS];
        set_fields(fields, {F_DENSITY0, F_ENERGY0, F_PRESSURE, F_VISCOSITY, F_DENSITY1, F_ENERGY1, F_XVEL0, F_YVEL0,
                            F_ZVEL0, F_XVEL1, F_YVEL1, F_ZVEL1});
        update_halo(c, fields, 2);
        hip_check(hipStreamSynchronize(c->st), "sync");
        *out = c;
        return 0;
    } catch (const Error &e) {
        set_last_error(e.what());
        destroy(c);
        return e.code;
    } catch (const std::exception &e) {
        set_last_error(e.what());
        destroy(c);
        return E_ARG;
    }
}

int clover_chunk_destroy(clover_chunk *c)
{
    destroy(c);
    return 0;
}

int clover_chunk_step(clover_chunk *c, double *dt)
{
    CLOVER_API_BEGIN
    if (!c) throw Error(E_ARG, "null chunk");
    KTIMER_SCOPE(c);
    c->step = c->step + 1;
    if (timestep(c)) {
        if (dt) *dt = c->dt;
        set_last_error("timestep: small timestep");
        return 1;
    }
    do_step(c);
    if (dt) *dt = c->dt;
    CLOVER_API_END
}

int clover_chunk_field_summary(clover_chunk *c, double *out)
{
    CLOVER_API_BEGIN
    if (!c || !out) throw Error(E_ARG, "null argument");
    KTIMER_SCOPE(c);
    field_summary(c, out);
    CLOVER_API_END
}

int clover_chunk_visit_prepare(clover_chunk *c)
{
    CLOVER_API_BEGIN
    if (!c) throw Error(E_ARG, "null chunk");
    KTIMER_SCOPE(c);
    visit_prepare(c);
    CLOVER_API_END
}

int clover_chunk_run(clover_chunk *c, int max_steps, clover_summary *rows, int max_rows, int *nrows)
{
    CLOVER_API_BEGIN
    if (!c) throw Error(E_ARG, "null chunk");
    KTIMER_SCOPE(c);
    int n = 0;
    const clover_deck &d = c->deck;
    auto summary_row = [&]() {
        double s[5];
        field_summary(c, s);
        if (rows && n < max_rows) {
            rows[n].step = c->step;
            rows[n].time = c->time;
            rows[n].vol = s[0];
            rows[n].mass = s[1];
            rows[n].ie = s[2];
            rows[n].ke = s[3];
            rows[n].press = s[4];
        }
        n++;
    };
    for (int it = 0; it < max_steps; it++) {
        c->step = c->step + 1;
        if (timestep(c)) {
            if (nrows) *nrows = n < max_rows ? n : max_rows;
            set_last_error("timestep: small timestep");
            return 1;
        }
        do_step(c);
        if (d.summary_frequency != 0 && c->step % d.summary_frequency == 0) summary_row();
        if (c->time + c->g_small > d.end_time || c->step >= d.end_step) {
            summary_row();
            break;
        }
    }
    hip_check(hipStreamSynchronize(c->st), "sync");
    if (nrows) *nrows = n < max_rows ? n : max_rows;
    CLOVER_API_END
}

int clover_chunk_state(const clover_chunk *c, int *step, double *time, double *dt)
{
    if (!c) return E_ARG;
    if (step) *step = c->step;
    if (time) *time = c->time;
    if (dt) *dt = c->dt;
    return 0;
}

int clover_chunk_tiles(const clover_chunk *c, int *ntiles)
{
    if (!c || !ntiles) return E_ARG;
    *ntiles = (int)c->tiles.size();
    return 0;
}

int clover_chunk_tile_info(const clover_chunk *c, int t, int *dims3, int *offs3, int *neighbours6)
{
    if (!c || t < 0 || t >= (int)c->tiles.size()) return E_ARG;
    const Tile &tt = c->tiles[t];
    if (dims3) { dims3[0] = tt.nx; dims3[1] = tt.ny; dims3[2] = tt.nz; }
    if (offs3) { offs3[0] = tt.left; offs3[1] = tt.bottom; offs3[2] = tt.back; }
    if (neighbours6) memcpy(neighbours6, tt.nbr, sizeof(tt.nbr));
    return 0;
}

int clover_decompose_tiles(int x_cells, int y_cells, int z_cells, int tiles_per_chunk, int tile_dims, int max_tiles,
                           int *dims3, int *offs3, int *neighbours6)
{
    if (x_cells < 1 || y_cells < 1 || z_cells < 1 || tiles_per_chunk < 1 || tile_dims < 1 || tile_dims > 3)
        return E_ARG;
    if (tiles_per_chunk > max_tiles) return E_ARG;
    clover_deck d{};
    d.x_cells = x_cells;
    d.y_cells = y_cells;
    d.z_cells = z_cells;
    d.tiles_per_chunk = tiles_per_chunk;
    d.tile_dims = tile_dims;
    std::vector<Tile> tiles;
    decompose(d, tiles);
    for (size_t t = 0; t < tiles.size(); t++) {
        const Tile &tt = tiles[t];
        if (dims3) { dims3[3 * t] = tt.nx; dims3[3 * t + 1] = tt.ny; dims3[3 * t + 2] = tt.nz; }
        if (offs3) { offs3[3 * t] = tt.left; offs3[3 * t + 1] = tt.bottom; offs3[3 * t + 2] = tt.back; }
        if (neighbours6) memcpy(neighbours6 + 6 * t, tt.nbr, sizeof(tt.nbr));
    }
    return (int)tiles.size();
}

int clover_chunk_fetch_field(clover_chunk *c, int t, int field, double *host)
{
    CLOVER_API_BEGIN
    if (!c || !host || t < 0 || t >= (int)c->tiles.size()) throw Error(E_ARG, "bad argument");
    if (field < 0 || field > X_ZAREA) throw Error(E_ARG, "bad field id");
    KTIMER_SCOPE(c);
    // the fused step forms soundspeed where it reads it instead of storing it
    if (field == F_SOUNDSPEED && c->fast) ideal_gas(c, false, true);
    Tile &tt = c->tiles[t];
    if (c->store == ST_F64) {
        download_dense(tt.g, class_of(field), tt.f[field], host, c->st);
        hip_check(hipStreamSynchronize(c->st), "sync");
    } else {
        std::vector<float> f((size_t)dense_elems(tt.g, class_of(field)));
        download_dense_any(tt.g, class_of(field), tt.f[field], f.data(), sizeof(float), c->st);
        hip_check(hipStreamSynchronize(c->st), "sync");
        for (size_t q = 0; q < f.size(); q++) host[q] = f[q];
    }
    CLOVER_API_END
}

int clover_chunk_set_profiling(clover_chunk *c, int on)
{
    if (!c) return E_ARG;
    c->profiling = on != 0;
    for (int p = 0; p < P_COUNT; p++) { c->prof_ms[p] = 0.0; c->prof_calls[p] = 0; }
    return 0;
}

int clover_chunk_profile(const clover_chunk *c, double *ms, int *calls, int n)
{
    if (!c) return E_ARG;
    for (int p = 0; p < n && p < P_COUNT; p++) {
        if (ms) ms[p] = c->prof_ms[p];
        if (calls) calls[p] = c->prof_calls[p];
    }
    return P_COUNT;
}

const char *clover_chunk_profile_name(int i) { return (i >= 0 && i < P_COUNT) ? kProfNames[i] : ""; }

long long clover_chunk_device_bytes(const clover_chunk *c) { return c ? c->bytes : 0; }

int clover_chunk_set_kernel_timing(clover_chunk *c, int on)
{
    CLOVER_API_BEGIN
    if (!c) throw Error(E_ARG, "null chunk");
    if (on) {
        if (!c->ktimer) c->ktimer = new EventTimer();
        c->ktimer->reset();
    } else {
        delete c->ktimer;
        c->ktimer = nullptr;
    }
    CLOVER_API_END
}

int clover_chunk_kernel_times(clover_chunk *c, char *buf, int buflen)
{
    CLOVER_API_BEGIN
    if (!c || !buf || buflen < 1) throw Error(E_ARG, "bad argument");
    if (!c->ktimer) throw Error(E_STATE, "kernel timing is off");
    c->ktimer->drain();
    std::string out;
    char line[256];
    for (const auto &kv : c->ktimer->stats) {
        snprintf(line, sizeof line, "%s %.6f %lld\n", kv.first.c_str(), kv.second.ms, kv.second.launches);
        out += line;
    }
    if ((int)out.size() + 1 > buflen) throw Error(E_ARG, "buffer too small");
    memcpy(buf, out.c_str(), out.size() + 1);
    CLOVER_API_END
}

int clover_chunk_autotune(clover_chunk *c, double *ms_before, double *ms_after)
{
    CLOVER_API_BEGIN
    if (!c) throw Error(E_ARG, "null chunk");
    if (!c->fast) throw Error(E_STATE, "the per-call kernels have no launch shapes");
    autotune(c, ms_before, ms_after);
    CLOVER_API_END
}

int clover_chunk_get_shape(const clover_chunk *c, int kernel, int *planes, int *xcd, int *block_x)
{
    CLOVER_API_BEGIN
    if (!c || !planes || !xcd || !block_x) throw Error(E_ARG, "null argument");
    if (kernel < 0 || kernel >= SK_COUNT) throw Error(E_ARG, "kernel out of range");
    *planes = c->shapes.k[kernel].planes;
    *xcd = c->shapes.k[kernel].xcd;
    *block_x = c->shapes.k[kernel].bx;
    CLOVER_API_END
}

int clover_chunk_set_shape(clover_chunk *c, int kernel, int planes, int xcd, int block_x)
{
    CLOVER_API_BEGIN
    if (!c) throw Error(E_ARG, "null chunk");
    if (kernel < 0 || kernel >= SK_COUNT) throw Error(E_ARG, "kernel out of range");
    if (planes < 1 || (xcd != 0 && xcd != 1)) throw Error(E_ARG, "planes must be positive, xcd 0 or 1");
    if (block_x != 64 && block_x != 128 && block_x != 256 && block_x != 512)
        throw Error(E_ARG, "block_x must be 64, 128, 256 or 512");
    if (block_x != 64 && !z_march(kernel)) throw Error(E_ARG, "only the z marches take a workgroup width");
    c->shapes.k[kernel] = Shape{planes, xcd, block_x};
    CLOVER_API_END
}

int clover_chunk_dt_control(const clover_chunk *c, int *control, int *jdt, int *kdt, double *x_pos, double *y_pos)
{
    CLOVER_API_BEGIN
    if (!c || !control || !jdt || !kdt || !x_pos || !y_pos) throw Error(E_ARG, "null argument");
    if (c->dt_tile < 0) throw Error(E_STATE, "no timestep taken yet");
    const Tile &t = c->tiles[c->dt_tile];
    // calc_dt_kernel.f90:97 and 154-160: jkl_control = 1.1 (a default-real
    // literal), so the control is INT(10.01 * 0.1...) = 1 ('sound'),
    // j = MOD(1, x_max) and k = INT(1 + 1.0/x_max), positions cellx(j) and
    // celly(k) of the tile (tile-local x_max; arrays from x_min-2 = -1)
    const double jkl = (double)1.1f;
    *control = (int)((double)10.01f * (jkl - (double)(int)jkl));
    const int xm = t.nx;
    const int j = ((int)jkl) % xm;
    const int k = (int)(1.0 + ((double)(int)jkl) / (double)xm);
    *jdt = j;
    *kdt = k;
    *x_pos = t.h_cellx.at((size_t)(j + 1));
    *y_pos = t.h_celly.at((size_t)(k + 1));
    CLOVER_API_END
}

}  // extern "C"
