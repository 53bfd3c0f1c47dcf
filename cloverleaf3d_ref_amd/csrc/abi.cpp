// Drop-in C ABI: one entry point per reference *_kernel subroutine.
//
// Host arrays are mapped to device copies by their base address (the
// residency map).  First sight allocates the copy in HBM and uploads the host
// contents; later calls use the device copy, which stays authoritative until
// clover_hip_fetch() copies it back.  Each entry point enqueues the kernels of
// one reference call on the library stream, in program order; only calc_dt
// (dt and the location outputs) and field_summary (five sums) wait for the
// device and return data to the host.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/clover_hip.h"
#include "kernels.h"
#include "runtime.h"

namespace clover {

static thread_local std::string t_last_error;
void set_last_error(const std::string &msg) { t_last_error = msg; }

int64_t dense_elems(const Lay &g, int cls)
{
    int ex, ey, ez;
    class_extents(cls, g.x_max - g.x_min + 1, g.y_max - g.y_min + 1, g.z_max - g.z_min + 1, ex, ey, ez);
    return (int64_t)ex * ey * ez;
}

void upload_dense_any(const Lay &g, int cls, const void *host, void *dev, size_t es, hipStream_t st)
{
    int ex, ey, ez;
    class_extents(cls, g.x_max - g.x_min + 1, g.y_max - g.y_min + 1, g.z_max - g.z_min + 1, ex, ey, ez);
    if (class_is_1d(cls)) {
        hip_check(hipMemcpyAsync(dev, host, es * ex, hipMemcpyHostToDevice, st), "upload 1-D");
        return;
    }
    for (int l = 0; l < ez; l++)
        hip_check(hipMemcpy2DAsync((char *)dev + es * (g.base + (size_t)g.plane * l), es * g.ld,
                                   (const char *)host + es * ((size_t)ex * ey * l), es * ex, es * ex, ey,
                                   hipMemcpyHostToDevice, st),
                  "upload plane");
}

void download_dense_any(const Lay &g, int cls, const void *dev, void *host, size_t es, hipStream_t st)
{
    int ex, ey, ez;
    class_extents(cls, g.x_max - g.x_min + 1, g.y_max - g.y_min + 1, g.z_max - g.z_min + 1, ex, ey, ez);
    if (class_is_1d(cls)) {
        hip_check(hipMemcpyAsync(host, dev, es * ex, hipMemcpyDeviceToHost, st), "download 1-D");
        return;
    }
    for (int l = 0; l < ez; l++)
        hip_check(hipMemcpy2DAsync((char *)host + es * ((size_t)ex * ey * l), es * ex,
                                   (const char *)dev + es * (g.base + (size_t)g.plane * l), es * g.ld, es * ex, ey,
                                   hipMemcpyDeviceToHost, st),
                  "download plane");
}

void upload_dense(const Lay &g, int cls, const double *host, double *dev, hipStream_t st)
{
    upload_dense_any(g, cls, host, dev, sizeof(double), st);
}

void download_dense(const Lay &g, int cls, const double *dev, double *host, hipStream_t st)
{
    download_dense_any(g, cls, dev, host, sizeof(double), st);
}

namespace {

struct Resident {
    double *dev = nullptr;
    int cls = 0;
    Lay g{};
    size_t bytes = 0;
};

struct Ctx {
    bool init = false;
    int device = 0;
    hipStream_t stream = nullptr;
    double *partial = nullptr;  // reduction partials (device)
    double *scalar = nullptr;   // reduction results (device)
    double *h_scalar = nullptr; // pinned host mirror of `scalar`
    std::unordered_map<const void *, Resident> map;
    std::mutex mu;
};

Ctx &ctx()
{
    static Ctx c;
    return c;
}

void ensure_init()
{
    Ctx &c = ctx();
    if (c.init) return;
    hip_check(hipSetDevice(c.device), "hipSetDevice");
    hip_check(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking), "hipStreamCreate");
    hip_check(hipMalloc(&c.partial, sizeof(double) * 5 * kMaxPartials), "hipMalloc partials");
    hip_check(hipMalloc(&c.scalar, sizeof(double) * 8), "hipMalloc scalar");
    hip_check(hipHostMalloc(&c.h_scalar, sizeof(double) * 8, hipHostMallocDefault), "hipHostMalloc");
    noalloc_scratch();  // with the device current
    c.init = true;
}

bool same_lay(const Lay &a, const Lay &b)
{
    return a.x_min == b.x_min && a.x_max == b.x_max && a.y_min == b.y_min && a.y_max == b.y_max &&
           a.z_min == b.z_min && a.z_max == b.z_max;
}

// Device copy of the host array `h` of class `cls` on a tile with layout g.
double *res(void *h, int cls, const Lay &g)
{
    if (!h) throw Error(E_ARG, "null array argument");
    Ctx &c = ctx();
    auto it = c.map.find(h);
    if (it != c.map.end()) {
        const Resident &r = it->second;
        if (r.cls != cls || !same_lay(r.g, g))
            throw Error(E_ARG, "array at a known address passed with a different shape or class");
        return r.dev;
    }
    Resident r;
    r.cls = cls;
    r.g = g;
    if (class_is_1d(cls)) {
        r.bytes = sizeof(double) * dense_elems(g, cls);
    } else {
        const int64_t n = lay_elems(g);
        if (n >= (int64_t(1) << 31)) throw Error(E_ARG, "tile too large for 32-bit element offsets");
        r.bytes = sizeof(double) * n;
    }
    hip_check(hipMalloc(&r.dev, r.bytes), "hipMalloc field");
    hip_check(hipMemsetAsync(r.dev, 0, r.bytes, c.stream), "hipMemset field");
    upload_dense(g, cls, (const double *)h, r.dev, c.stream);
    c.map.emplace(h, r);
    return r.dev;
}

void check_launch(hipError_t e, const char *what) { hip_check(e, what); }

Lay lay_of(int x_min, int x_max, int y_min, int y_max, int z_min, int z_max)
{
    if (x_max < x_min || y_max < y_min || z_max < z_min) throw Error(E_ARG, "empty tile bounds");
    return make_lay(x_min, x_max, y_min, y_max, z_min, z_max);
}

}  // namespace
}  // namespace clover

using namespace clover;

#define PREP()                                                  \
    ensure_init();                                              \
    Ctx &C = ctx();                                             \
    std::lock_guard<std::mutex> lock(C.mu);                     \
    const Lay g = lay_of(x_min, x_max, y_min, y_max, z_min, z_max)

extern "C" {

const char *clover_hip_last_error(void) { return t_last_error.c_str(); }

int clover_hip_init(int device)
{
    CLOVER_API_BEGIN
    Ctx &c = ctx();
    if (c.init && c.device != device) throw Error(E_STATE, "already initialised on another device");
    c.device = device;
    ensure_init();
    CLOVER_API_END
}

int clover_hip_finalize(void)
{
    CLOVER_API_BEGIN
    Ctx &c = ctx();
    if (!c.init) return 0;
    std::lock_guard<std::mutex> lock(c.mu);
    hip_check(hipStreamSynchronize(c.stream), "sync");
    for (auto &kv : c.map) hip_check(hipFree(kv.second.dev), "hipFree");
    c.map.clear();
    hip_check(hipFree(c.partial), "hipFree");
    hip_check(hipFree(c.scalar), "hipFree");
    hip_check(hipHostFree(c.h_scalar), "hipHostFree");
    hip_check(hipStreamDestroy(c.stream), "hipStreamDestroy");
    c.init = false;
    CLOVER_API_END
}

int clover_hip_sync(void)
{
    CLOVER_API_BEGIN
    ensure_init();
    hip_check(hipStreamSynchronize(ctx().stream), "sync");
    CLOVER_API_END
}

int clover_hip_fetch(const void *host)
{
    CLOVER_API_BEGIN
    ensure_init();
    Ctx &c = ctx();
    std::lock_guard<std::mutex> lock(c.mu);
    auto it = c.map.find(host);
    if (it == c.map.end()) throw Error(E_ARG, "clover_hip_fetch: address not resident");
    download_dense(it->second.g, it->second.cls, it->second.dev, (double *)host, c.stream);
    hip_check(hipStreamSynchronize(c.stream), "sync");
    CLOVER_API_END
}

int clover_hip_push(const void *host)
{
    CLOVER_API_BEGIN
    ensure_init();
    Ctx &c = ctx();
    std::lock_guard<std::mutex> lock(c.mu);
    auto it = c.map.find(host);
    if (it == c.map.end()) throw Error(E_ARG, "clover_hip_push: address not resident");
    upload_dense(it->second.g, it->second.cls, (const double *)host, it->second.dev, c.stream);
    hip_check(hipStreamSynchronize(c.stream), "sync");
    CLOVER_API_END
}

int clover_hip_forget(const void *host)
{
    CLOVER_API_BEGIN
    ensure_init();
    Ctx &c = ctx();
    std::lock_guard<std::mutex> lock(c.mu);
    auto it = c.map.find(host);
    if (it == c.map.end()) return 0;
    hip_check(hipStreamSynchronize(c.stream), "sync");
    hip_check(hipFree(it->second.dev), "hipFree");
    c.map.erase(it);
    CLOVER_API_END
}

int clover_hip_resident_count(void) { return (int)ctx().map.size(); }

int clover_hip_ideal_gas_kernel(CLOVER_DIMS, double *density, double *energy, double *pressure, double *soundspeed)
{
    CLOVER_API_BEGIN
    PREP();
    check_launch(launch_ideal_gas(g, res(density, AC_CELL, g), res(energy, AC_CELL, g), res(pressure, AC_CELL, g),
                                  res(soundspeed, AC_CELL, g), C.stream),
                 "ideal_gas");
    CLOVER_API_END
}

int clover_hip_viscosity_kernel(CLOVER_DIMS, double *xarea, double *yarea, double *zarea, double *celldx,
                                double *celldy, double *celldz, double *density0, double *pressure,
                                double *viscosity, double *xvel0, double *yvel0, double *zvel0)
{
    CLOVER_API_BEGIN
    PREP();
    check_launch(launch_viscosity(g, res(xarea, AC_XF, g), res(yarea, AC_YF, g), res(zarea, AC_ZF, g),
                                  res(celldx, AC_C1X, g), res(celldy, AC_C1Y, g), res(celldz, AC_C1Z, g),
                                  res(density0, AC_CELL, g), res(pressure, AC_CELL, g), res(viscosity, AC_CELL, g),
                                  res(xvel0, AC_VERT, g), res(yvel0, AC_VERT, g), res(zvel0, AC_VERT, g), C.stream),
                 "viscosity");
    CLOVER_API_END
}

int clover_hip_calc_dt_kernel(CLOVER_DIMS, double g_small, double g_big, double dtmin, double dtc_safe,
                              double dtu_safe, double dtv_safe, double dtw_safe, double dtdiv_safe, double *xarea,
                              double *yarea, double *zarea, double *cellx, double *celly, double *cellz,
                              double *celldx, double *celldy, double *celldz, double *volume, double *density0,
                              double *energy0, double *pressure, double *viscosity_a, double *soundspeed,
                              double *xvel0, double *yvel0, double *zvel0, double *dt_min, double *dt_min_val,
                              int *dtl_control, double *xl_pos, double *yl_pos, double *zl_pos, int *jldt, int *kldt,
                              int *lldt, int *small)
{
    CLOVER_API_BEGIN
    PREP();
    if (!dt_min_val || !dtl_control || !xl_pos || !yl_pos || !zl_pos || !jldt || !kldt || !lldt || !small)
        throw Error(E_ARG, "calc_dt: null output");
    res(energy0, AC_CELL, g);  // read by the reference only in its small-dt diagnostic
    res(pressure, AC_CELL, g);
    res(cellx, AC_C1X, g);     // location outputs, read below from the (unchanging) host copies
    res(celly, AC_C1Y, g);
    res(cellz, AC_C1Z, g);
    int nparts = 0;
    check_launch(launch_calc_dt_cells(g, dtc_safe, dtu_safe, dtv_safe, dtw_safe, dtdiv_safe, g_small,
                                      res(xarea, AC_XF, g), res(yarea, AC_YF, g), res(zarea, AC_ZF, g),
                                      res(celldx, AC_C1X, g), res(celldy, AC_C1Y, g), res(celldz, AC_C1Z, g),
                                      res(volume, AC_CELL, g), res(density0, AC_CELL, g),
                                      res(viscosity_a, AC_CELL, g), res(soundspeed, AC_CELL, g),
                                      res(xvel0, AC_VERT, g), res(yvel0, AC_VERT, g), res(zvel0, AC_VERT, g),
                                      res(dt_min, AC_VERT, g), C.partial, &nparts, C.stream),
                 "calc_dt");
    check_launch(launch_min_final(C.partial, nparts, g_big, C.scalar, C.stream), "calc_dt min");
    hip_check(hipMemcpyAsync(C.h_scalar, C.scalar, sizeof(double), hipMemcpyDeviceToHost, C.stream), "dt D2H");
    hip_check(hipStreamSynchronize(C.stream), "sync");
    const double dmin = C.h_scalar[0];
    *dt_min_val = dmin;
    // calc_dt_kernel.f90:152-161: jkl_control is the default-real 1.1 and
    // never updated, so the location outputs are constants of the tile.
    double jkl = (double)1.1f;
    *dtl_control = (int)((double)10.01f * (jkl - (double)(int)jkl));
    jkl = jkl - (jkl - (double)(int)jkl);
    *jldt = ((int)jkl) % x_max;
    *kldt = (int)(1.0 + (jkl / (double)x_max));
    *lldt = (int)(1.0 + (jkl / (double)x_max));
    // geometry is never written by a kernel: the host copies are current
    *xl_pos = cellx[*jldt - (x_min - 2)];
    *yl_pos = celly[*kldt - (y_min - 2)];
    *zl_pos = cellz[*lldt - (z_min - 2)];
    *small = dmin < dtmin ? 1 : 0;
    if (*small) fprintf(stderr, "Timestep information:\n timestep : %.16e\n", dmin);
    CLOVER_API_END
}

int clover_hip_PdV_kernel(int predict, CLOVER_DIMS, double dt, double *xarea, double *yarea, double *zarea,
                          double *volume, double *density0, double *density1, double *energy0, double *energy1,
                          double *pressure, double *viscosity, double *xvel0, double *xvel1, double *yvel0,
                          double *yvel1, double *zvel0, double *zvel1)
{
    CLOVER_API_BEGIN
    PREP();
    check_launch(launch_pdv(predict, g, dt, res(xarea, AC_XF, g), res(yarea, AC_YF, g), res(zarea, AC_ZF, g),
                            res(volume, AC_CELL, g), res(density0, AC_CELL, g), res(density1, AC_CELL, g),
                            res(energy0, AC_CELL, g), res(energy1, AC_CELL, g), res(pressure, AC_CELL, g),
                            res(viscosity, AC_CELL, g), res(xvel0, AC_VERT, g), res(xvel1, AC_VERT, g),
                            res(yvel0, AC_VERT, g), res(yvel1, AC_VERT, g), res(zvel0, AC_VERT, g),
                            res(zvel1, AC_VERT, g), C.stream),
                 "PdV");
    CLOVER_API_END
}

int clover_hip_revert_kernel(CLOVER_DIMS, double *density0, double *density1, double *energy0, double *energy1)
{
    CLOVER_API_BEGIN
    PREP();
    check_launch(launch_revert(g, res(density0, AC_CELL, g), res(density1, AC_CELL, g), res(energy0, AC_CELL, g),
                               res(energy1, AC_CELL, g), C.stream),
                 "revert");
    CLOVER_API_END
}

int clover_hip_accelerate_kernel(CLOVER_DIMS, double dt, double *xarea, double *yarea, double *zarea,
                                 double *volume, double *density0, double *pressure, double *viscosity,
                                 double *xvel0, double *yvel0, double *zvel0, double *xvel1, double *yvel1,
                                 double *zvel1)
{
    CLOVER_API_BEGIN
    PREP();
    check_launch(launch_accelerate(g, dt, res(xarea, AC_XF, g), res(yarea, AC_YF, g), res(zarea, AC_ZF, g),
                                   res(volume, AC_CELL, g), res(density0, AC_CELL, g), res(pressure, AC_CELL, g),
                                   res(viscosity, AC_CELL, g), res(xvel0, AC_VERT, g), res(yvel0, AC_VERT, g),
                                   res(zvel0, AC_VERT, g), res(xvel1, AC_VERT, g), res(yvel1, AC_VERT, g),
                                   res(zvel1, AC_VERT, g), C.stream),
                 "accelerate");
    CLOVER_API_END
}

int clover_hip_flux_calc_kernel(CLOVER_DIMS, double dt, double *xarea, double *yarea, double *zarea, double *xvel0,
                                double *yvel0, double *zvel0, double *xvel1, double *yvel1, double *zvel1,
                                double *vol_flux_x, double *vol_flux_y, double *vol_flux_z)
{
    CLOVER_API_BEGIN
    PREP();
    check_launch(launch_flux_calc(g, dt, res(xarea, AC_XF, g), res(yarea, AC_YF, g), res(zarea, AC_ZF, g),
                                  res(xvel0, AC_VERT, g), res(yvel0, AC_VERT, g), res(zvel0, AC_VERT, g),
                                  res(xvel1, AC_VERT, g), res(yvel1, AC_VERT, g), res(zvel1, AC_VERT, g),
                                  res(vol_flux_x, AC_XF, g), res(vol_flux_y, AC_YF, g), res(vol_flux_z, AC_ZF, g),
                                  C.stream),
                 "flux_calc");
    CLOVER_API_END
}

int clover_hip_advec_cell_kernel(CLOVER_DIMS, int advect_x, int dir, int sweep_number, double *vertexdx,
                                 double *vertexdy, double *vertexdz, double *volume, double *density1,
                                 double *energy1, double *mass_flux_x, double *vol_flux_x, double *mass_flux_y,
                                 double *vol_flux_y, double *mass_flux_z, double *vol_flux_z, double *pre_vol,
                                 double *post_vol, double *pre_mass, double *post_mass, double *advec_vol,
                                 double *post_ener, double *ener_flux)
{
    CLOVER_API_BEGIN
    PREP();
    check_launch(launch_advec_cell(g, advect_x, dir, sweep_number, res(vertexdx, AC_V1X, g),
                                   res(vertexdy, AC_V1Y, g), res(vertexdz, AC_V1Z, g), res(volume, AC_CELL, g),
                                   res(density1, AC_CELL, g), res(energy1, AC_CELL, g), res(mass_flux_x, AC_XF, g),
                                   res(vol_flux_x, AC_XF, g), res(mass_flux_y, AC_YF, g), res(vol_flux_y, AC_YF, g),
                                   res(mass_flux_z, AC_ZF, g), res(vol_flux_z, AC_ZF, g), res(pre_vol, AC_VERT, g),
                                   res(post_vol, AC_VERT, g), res(pre_mass, AC_VERT, g), res(post_mass, AC_VERT, g),
                                   res(advec_vol, AC_VERT, g), res(post_ener, AC_VERT, g),
                                   res(ener_flux, AC_VERT, g), C.stream),
                 "advec_cell");
    CLOVER_API_END
}

int clover_hip_advec_mom_kernel(CLOVER_DIMS, double *vel1, double *mass_flux_x, double *vol_flux_x,
                                double *mass_flux_y, double *vol_flux_y, double *mass_flux_z, double *vol_flux_z,
                                double *volume, double *density1, double *node_flux, double *node_mass_post,
                                double *node_mass_pre, double *advec_vel, double *mom_flux, double *pre_vol,
                                double *post_vol, double *celldx, double *celldy, double *celldz, int advect_x,
                                int which_vel, int sweep_number, int direction)
{
    CLOVER_API_BEGIN
    PREP();
    res(advec_vel, AC_VERT, g);  // unused by the reference kernel (work_array4)
    check_launch(launch_advec_mom(g, res(vel1, AC_VERT, g), res(mass_flux_x, AC_XF, g), res(vol_flux_x, AC_XF, g),
                                  res(mass_flux_y, AC_YF, g), res(vol_flux_y, AC_YF, g), res(mass_flux_z, AC_ZF, g),
                                  res(vol_flux_z, AC_ZF, g), res(volume, AC_CELL, g), res(density1, AC_CELL, g),
                                  res(node_flux, AC_VERT, g), res(node_mass_post, AC_VERT, g),
                                  res(node_mass_pre, AC_VERT, g), res(mom_flux, AC_VERT, g), res(pre_vol, AC_VERT, g),
                                  res(post_vol, AC_VERT, g), res(celldx, AC_C1X, g), res(celldy, AC_C1Y, g),
                                  res(celldz, AC_C1Z, g), advect_x, which_vel, sweep_number, direction, C.stream),
                 "advec_mom");
    CLOVER_API_END
}

int clover_hip_reset_field_kernel(CLOVER_DIMS, double *density0, double *density1, double *energy0, double *energy1,
                                  double *xvel0, double *xvel1, double *yvel0, double *yvel1, double *zvel0,
                                  double *zvel1)
{
    CLOVER_API_BEGIN
    PREP();
    check_launch(launch_reset_field(g, res(density0, AC_CELL, g), res(density1, AC_CELL, g), res(energy0, AC_CELL, g),
                                    res(energy1, AC_CELL, g), res(xvel0, AC_VERT, g), res(xvel1, AC_VERT, g),
                                    res(yvel0, AC_VERT, g), res(yvel1, AC_VERT, g), res(zvel0, AC_VERT, g),
                                    res(zvel1, AC_VERT, g), C.stream),
                 "reset_field");
    CLOVER_API_END
}

// the 19 halo fields by FIELD id order (data.f90:55-74)
#define HALO_PTRS                                                                                               \
    double *fp[NUM_FIELDS] = {                                                                                  \
        res(density0, AC_CELL, g),   res(density1, AC_CELL, g),   res(energy0, AC_CELL, g),                     \
        res(energy1, AC_CELL, g),    res(pressure, AC_CELL, g),   res(viscosity, AC_CELL, g),                   \
        res(soundspeed, AC_CELL, g), res(xvel0, AC_VERT, g),      res(xvel1, AC_VERT, g),                       \
        res(yvel0, AC_VERT, g),      res(yvel1, AC_VERT, g),      res(zvel0, AC_VERT, g),                       \
        res(zvel1, AC_VERT, g),      res(vol_flux_x, AC_XF, g),   res(vol_flux_y, AC_YF, g),                    \
        res(vol_flux_z, AC_ZF, g),   res(mass_flux_x, AC_XF, g),  res(mass_flux_y, AC_YF, g),                   \
        res(mass_flux_z, AC_ZF, g)}

int clover_hip_update_halo_kernel(CLOVER_DIMS, const int *chunk_neighbours, const int *tile_neighbours,
                                  double *density0, double *energy0, double *pressure, double *viscosity,
                                  double *soundspeed, double *density1, double *energy1, double *xvel0, double *yvel0,
                                  double *zvel0, double *xvel1, double *yvel1, double *zvel1, double *vol_flux_x,
                                  double *vol_flux_y, double *vol_flux_z, double *mass_flux_x, double *mass_flux_y,
                                  double *mass_flux_z, const int *fields, int depth)
{
    CLOVER_API_BEGIN
    PREP();
    if (!chunk_neighbours || !tile_neighbours || !fields) throw Error(E_ARG, "update_halo: null flags");
    if (depth < 1 || depth > 2) throw Error(E_ARG, "update_halo: depth must be 1 or 2");
    HALO_PTRS;
    int ext[6];
    for (int f = 0; f < 6; f++) ext[f] = chunk_neighbours[f] == -1 && tile_neighbours[f] == -1;
    check_launch(launch_update_halo(g, fp, fields, ext, depth, C.stream), "update_halo");
    CLOVER_API_END
}

static int tile_halo(int face, CLOVER_TILE_ARGS)
{
    CLOVER_API_BEGIN
    PREP();
    if (!fields) throw Error(E_ARG, "update_tile_halo: null flags");
    if (depth < 1 || depth > 2) throw Error(E_ARG, "update_tile_halo: depth must be 1 or 2");
    HALO_PTRS;
    const Lay gn = lay_of(n_xmin, n_xmax, n_ymin, n_ymax, n_zmin, n_zmax);
    double *np[NUM_FIELDS] = {
        res(n_density0, AC_CELL, gn),   res(n_density1, AC_CELL, gn),  res(n_energy0, AC_CELL, gn),
        res(n_energy1, AC_CELL, gn),    res(n_pressure, AC_CELL, gn),  res(n_viscosity, AC_CELL, gn),
        res(n_soundspeed, AC_CELL, gn), res(n_xvel0, AC_VERT, gn),     res(n_xvel1, AC_VERT, gn),
        res(n_yvel0, AC_VERT, gn),      res(n_yvel1, AC_VERT, gn),     res(n_zvel0, AC_VERT, gn),
        res(n_zvel1, AC_VERT, gn),      res(n_vol_flux_x, AC_XF, gn),  res(n_vol_flux_y, AC_YF, gn),
        res(n_vol_flux_z, AC_ZF, gn),   res(n_mass_flux_x, AC_XF, gn), res(n_mass_flux_y, AC_YF, gn),
        res(n_mass_flux_z, AC_ZF, gn)};
    check_launch(launch_tile_halo(face, g, fp, gn, np, fields, depth, C.stream), "update_tile_halo");
    CLOVER_API_END
}

#define TILE_FWD(face)                                                                                          \
    tile_halo(face, x_min, x_max, y_min, y_max, z_min, z_max, density0, energy0, pressure, viscosity, soundspeed, \
              density1, energy1, xvel0, yvel0, zvel0, xvel1, yvel1, zvel1, vol_flux_x, vol_flux_y, vol_flux_z,   \
              mass_flux_x, mass_flux_y, mass_flux_z, n_xmin, n_xmax, n_ymin, n_ymax, n_zmin, n_zmax, n_density0,  \
              n_energy0, n_pressure, n_viscosity, n_soundspeed, n_density1, n_energy1, n_xvel0, n_yvel0, n_zvel0, \
              n_xvel1, n_yvel1, n_zvel1, n_vol_flux_x, n_vol_flux_y, n_vol_flux_z, n_mass_flux_x, n_mass_flux_y, \
              n_mass_flux_z, fields, depth)

int clover_hip_update_tile_halo_left_kernel(CLOVER_TILE_ARGS) { return TILE_FWD(1); }
int clover_hip_update_tile_halo_right_kernel(CLOVER_TILE_ARGS) { return TILE_FWD(2); }
int clover_hip_update_tile_halo_bottom_kernel(CLOVER_TILE_ARGS) { return TILE_FWD(3); }
int clover_hip_update_tile_halo_top_kernel(CLOVER_TILE_ARGS) { return TILE_FWD(4); }
int clover_hip_update_tile_halo_back_kernel(CLOVER_TILE_ARGS) { return TILE_FWD(5); }
int clover_hip_update_tile_halo_front_kernel(CLOVER_TILE_ARGS) { return TILE_FWD(6); }

int clover_hip_field_summary_kernel(CLOVER_DIMS, double *volume, double *density0, double *energy0,
                                    double *pressure, double *xvel0, double *yvel0, double *zvel0, double *vol,
                                    double *mass, double *ie, double *ke, double *press)
{
    CLOVER_API_BEGIN
    PREP();
    if (!vol || !mass || !ie || !ke || !press) throw Error(E_ARG, "field_summary: null output");
    int nparts = 0;
    // xvel0..zvel0 are vertex arrays (build_field.f90:62-67), whatever
    // extent the summary kernel declares for them
    check_launch(launch_field_summary_partial(g, res(volume, AC_CELL, g), res(density0, AC_CELL, g),
                                              res(energy0, AC_CELL, g), res(pressure, AC_CELL, g),
                                              res(xvel0, AC_VERT, g), res(yvel0, AC_VERT, g), res(zvel0, AC_VERT, g),
                                              C.partial, &nparts, C.stream),
                 "field_summary");
    check_launch(launch_sum5_final(C.partial, nparts, C.scalar, C.stream), "field_summary sum");
    hip_check(hipMemcpyAsync(C.h_scalar, C.scalar, sizeof(double) * 5, hipMemcpyDeviceToHost, C.stream), "D2H");
    hip_check(hipStreamSynchronize(C.stream), "sync");
    *vol = C.h_scalar[0];
    *mass = C.h_scalar[1];
    *ie = C.h_scalar[2];
    *ke = C.h_scalar[3];
    *press = C.h_scalar[4];
    CLOVER_API_END
}

}  // extern "C"
