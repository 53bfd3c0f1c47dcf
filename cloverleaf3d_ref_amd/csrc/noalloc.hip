// field_summary without device allocation, for callers that cannot call
// hipMalloc: code under HIP graph capture, tight serving loops, harnesses
// that forbid allocation.  The caller's arrays are device arrays in the
// reference's own Fortran shapes (build_field.f90:41-147) and the five sums
// go to device memory, asynchronously on the caller's stream.  The block
// partial sums live in `scratch`, a fixed device array of this library; each
// GPU holds its own copy from the first call on it until the process ends.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>

#include "../../include/clover_hip.h"
#include "kernels.h"
#include "runtime.h"

namespace clover {
namespace {

// Five partial sums per block of k_field_summary, whose block count is
// capped at kMaxPartials: 640 KiB, enough for a tile of any size.
constexpr int kScratchDoubles = 5 * kMaxPartials;
static __device__ __align__(256) double scratch[kScratchDoubles];

// Device addresses of `scratch`, one per GPU.  In host code the array's
// name is only a stand-in for the symbol; the kernels get these addresses.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch[kMaxDevices] = {};

// Layout of a dense Fortran array with extents ex x ey x . starting at
// (x_min-2, y_min-2, z_min-2): the layout.h indexing with no padding.
Lay dense_lay(int x_min, int x_max, int y_min, int y_max, int z_min, int z_max, int ex, int ey)
{
    Lay g;
    g.x_min = x_min; g.x_max = x_max; g.y_min = y_min; g.y_max = y_max; g.z_min = z_min; g.z_max = z_max;
    g.ld = ex;
    g.plane = ex * ey;
    g.base = 0;
    g.mirror = 0;
    return g;
}

}  // namespace

// The current device's address of `scratch`, looked up once per device with
// that device current (the library's init call, else the first call here).
double *noalloc_scratch()
{
    int dev = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) throw Error(E_ARG, "device ordinal beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        if (!p) throw Error(E_HIP, "hipGetSymbolAddress(scratch) returned null");
        g_scratch[dev] = static_cast<double *>(p);
    }
    return g_scratch[dev];
}

}  // namespace clover

using namespace clover;

extern "C" int clover_hip_field_summary_kernel_noalloc(CLOVER_DIMS, const double *volume, const double *density0,
                                                       const double *energy0, const double *pressure,
                                                       const double *xvel0, const double *yvel0,
                                                       const double *zvel0, double *sums, void *stream)
{
    CLOVER_API_BEGIN
    if (x_max < x_min || y_max < y_min || z_max < z_min) throw Error(E_ARG, "field_summary_noalloc: empty tile");
    if (!volume || !density0 || !energy0 || !pressure || !xvel0 || !yvel0 || !zvel0 || !sums)
        throw Error(E_ARG, "field_summary_noalloc: null array");
    const int64_t nx = (int64_t)x_max - x_min + 1, ny = (int64_t)y_max - y_min + 1, nz = (int64_t)z_max - z_min + 1;
    // the largest array, a vertex array, is indexed with 32-bit offsets
    if ((nx + 5) * (ny + 5) * (nz + 5) > INT32_MAX)
        throw Error(E_ARG, "field_summary_noalloc: tile of " + std::to_string(nx) + "x" + std::to_string(ny) + "x" +
                               std::to_string(nz) + " cells exceeds the maximum of 2^31-1 vertex elements");
    double *part = noalloc_scratch();
    const Lay gc = dense_lay(x_min, x_max, y_min, y_max, z_min, z_max, (int)nx + 4, (int)ny + 4);
    const Lay gv = dense_lay(x_min, x_max, y_min, y_max, z_min, z_max, (int)nx + 5, (int)ny + 5);
    hipStream_t st = static_cast<hipStream_t>(stream);
    int nparts = 0;
    // every partial the final sum reads is written by this call's first
    // kernel, so nothing in scratch needs clearing
    hip_check(launch_field_summary_partial(gc, gv, volume, density0, energy0, pressure, xvel0, yvel0, zvel0, part,
                                           &nparts, st),
              "field_summary_noalloc");
    hip_check(launch_sum5_final(part, nparts, sums, st), "field_summary_noalloc sum");
    CLOVER_API_END
}
