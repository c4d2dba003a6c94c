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

// Opt-in kernel timing: when enabled, every launch of the library is
// bracketed by a pair of hipEvents recorded on its stream; collecting waits
// for them and sums the elapsed time per kernel name.
enum KernelId {
    K_SETUP_VIEW = 0,
    K_PACK,
    K_INIT,
    K_CANDIDATES,
    K_SELECT,
    K_REFINE,
    K_DEPTH_NORMAL,
    K_FILTER,
    K_EXPORT,
    K_COST_BATCH,
    K_FUSE_FILL,
    K_FUSE_CHECK,
    K_FUSE_SCATTER,
    K_FUSE_VIEWDESC,
    K_POINTS,
    K_FORMAT,
    K_SWN,
    K_SEG_SORT,
    K_SEG_STATS,
    K_NUM
};

bool profiling_enabled();
void profile_begin(hipStream_t st, int kid);
void profile_end(hipStream_t st, int kid);
// Device counters [K_NUM][2] (NCCs, geometric terms) while profiling is on,
// else nullptr.
unsigned long long* profile_work_counters();

struct ProfScope {
    hipStream_t st;
    int kid;
    ProfScope(hipStream_t s, int k) : st(s), kid(k) { profile_begin(st, kid); }
    ~ProfScope() { profile_end(st, kid); }
};

}  // namespace gmvs

#define GMVS_PROF(st, kid) gmvs::ProfScope gmvs_prof_scope_(st, kid)

// hipLaunchKernelGGL bracketed by the (opt-in) timing events.
#define GMVS_LAUNCH(kid, st, ...)                    \
    do {                                             \
        gmvs::ProfScope gmvs_prof_scope_(st, kid);   \
        hipLaunchKernelGGL(__VA_ARGS__);             \
    } while (0)

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
