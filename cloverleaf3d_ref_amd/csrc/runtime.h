// Internal runtime helpers shared by the drop-in ABI and the resident driver:
// error propagation across the C boundary and host<->device copies between
// dense Fortran arrays and the device layout.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

#include "layout.h"

namespace clover {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

enum : int {
    E_HIP = -1,     // a HIP runtime call failed
    E_ARG = -2,     // invalid argument (shape, class, pointer)
    E_STATE = -3,   // library not initialised / wrong state
    E_NOMEM = -4,   // device allocation failed
};

void set_last_error(const std::string &msg);

inline void hip_check(hipError_t e, const char *what)
{
    if (e != hipSuccess)
        throw Error(e == hipErrorOutOfMemory ? E_NOMEM : E_HIP,
                    std::string(what) + ": " + hipGetErrorString(e));
}

#define CLOVER_API_BEGIN try {
#define CLOVER_API_END                                    \
    return 0;                                             \
    }                                                     \
    catch (const clover::Error &e) {                      \
        clover::set_last_error(e.what());                 \
        return e.code;                                    \
    }                                                     \
    catch (const std::exception &e) {                     \
        clover::set_last_error(e.what());                 \
        return clover::E_ARG;                             \
    }

// Copy a dense Fortran array of class `cls` (host) into / out of the device
// layout g.  1-D classes are plain contiguous copies.
void upload_dense(const Lay &g, int cls, const double *host, double *dev, hipStream_t st);
void download_dense(const Lay &g, int cls, const double *dev, double *host, hipStream_t st);
// the same for arrays of es-byte elements
void upload_dense_any(const Lay &g, int cls, const void *host, void *dev, size_t es, hipStream_t st);
void download_dense_any(const Lay &g, int cls, const void *dev, void *host, size_t es, hipStream_t st);

// element count of the dense host array of class cls
int64_t dense_elems(const Lay &g, int cls);

// the current device's address of the allocation-free entry point's fixed
// scratch array (looked up on first use per device)
double *noalloc_scratch();

}  // namespace clover
