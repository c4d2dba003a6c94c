// Error plumbing of the C ABI: a per-thread message plus status codes.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <string>

#include "gmvs_amd.h"

namespace gmvs {

void set_error(const char* fmt, ...) __attribute__((format(printf, 1, 2)));
void clear_error();

}  // namespace gmvs

#define GMVS_HIP_CHECK(expr)                                                                  \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess) {                                                               \
            gmvs::set_error("%s:%d %s failed: %s", __FILE__, __LINE__, #expr,                 \
                            hipGetErrorString(e_));                                           \
            return GMVS_ERR_HIP;                                                              \
        }                                                                                     \
    } while (0)

#define GMVS_ARG_CHECK(cond, ...)          \
    do {                                   \
        if (!(cond)) {                     \
            gmvs::set_error(__VA_ARGS__);  \
            return GMVS_ERR_ARG;           \
        }                                  \
    } while (0)

// Launch check: a bad launch configuration surfaces here, not at the next sync.
#define GMVS_LAUNCH_CHECK() GMVS_HIP_CHECK(hipGetLastError())
