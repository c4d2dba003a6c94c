// C-ABI entry points that need no device code: error string, version,
// load_camera / load_problem and the host Philox block.
#include "abi_common.h"

#include <cstring>
#include <mutex>
#include <vector>

#include "pm_device.h"

namespace gmvs {

namespace {
thread_local std::string g_error;
}

void set_error(const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_error = buf;
}

void clear_error() { g_error.clear(); }

namespace {

const char* const kKernelNames[K_NUM] = {
    "pm_setup_view", "pm_pack_quads", "pm_init",      "pm_candidates",  "pm_select",  "pm_refine",
    "pm_depth_normal", "pm_filter",   "pm_export",    "pm_cost_batch",  "fuse_fill",
    "fuse_check",     "fuse_scatter", "fuse_write_views", "points_gather", "format_rows",
    "sliding_window_normal", "segment_sort", "segment_stats"};

struct Pending {
    int kid;
    hipEvent_t a, b;
};

struct Profiler {
    std::mutex mu;
    bool on = false;
    std::vector<Pending> pending;
    std::vector<hipEvent_t> pool;
    double ms[K_NUM] = {};
    long long count[K_NUM] = {};
    unsigned long long* d_work = nullptr;  // [K_NUM][2], device

    hipEvent_t take() {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        return e;
    }
};

Profiler& prof() {
    static Profiler p;
    return p;
}

thread_local hipEvent_t t_start = nullptr;

}  // namespace

bool profiling_enabled() { return prof().on; }

unsigned long long* profile_work_counters() {
    Profiler& p = prof();
    return p.on ? p.d_work : nullptr;
}

void profile_begin(hipStream_t st, int kid) {
    Profiler& p = prof();
    if (!p.on) return;
    std::lock_guard<std::mutex> lk(p.mu);
    t_start = p.take();
    if (t_start && hipEventRecord(t_start, st) != hipSuccess) {
        p.pool.push_back(t_start);
        t_start = nullptr;
    }
}

void profile_end(hipStream_t st, int kid) {
    Profiler& p = prof();
    if (!p.on || !t_start) return;
    std::lock_guard<std::mutex> lk(p.mu);
    hipEvent_t e = p.take();
    if (e && hipEventRecord(e, st) == hipSuccess) {
        p.pending.push_back(Pending{kid, t_start, e});
    } else {
        p.pool.push_back(t_start);
        if (e) p.pool.push_back(e);
    }
    t_start = nullptr;
}

}  // namespace gmvs

extern "C" int gmvs_profile_enable(int32_t on) {
    gmvs::clear_error();
    gmvs::Profiler& p = gmvs::prof();
    std::lock_guard<std::mutex> lk(p.mu);
    if (on && !p.d_work) {
        const size_t bytes = sizeof(unsigned long long) * 2 * gmvs::K_NUM;
        GMVS_HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&p.d_work), bytes));
        GMVS_HIP_CHECK(hipMemset(p.d_work, 0, bytes));
    }
    p.on = on != 0;
    return GMVS_OK;
}

extern "C" int gmvs_profile_collect(gmvs_kernel_stat* out, int32_t max_out, int32_t* num_out) {
    gmvs::clear_error();
    gmvs::Profiler& p = gmvs::prof();
    std::lock_guard<std::mutex> lk(p.mu);
    for (const gmvs::Pending& q : p.pending) {
        float ms = 0.f;
        GMVS_HIP_CHECK(hipEventSynchronize(q.b));
        GMVS_HIP_CHECK(hipEventElapsedTime(&ms, q.a, q.b));
        p.ms[q.kid] += ms;
        p.count[q.kid] += 1;
        p.pool.push_back(q.a);
        p.pool.push_back(q.b);
    }
    p.pending.clear();
    unsigned long long work[2 * gmvs::K_NUM] = {};
    if (p.d_work)
        GMVS_HIP_CHECK(hipMemcpy(work, p.d_work, sizeof(work), hipMemcpyDeviceToHost));
    int n = 0;
    for (int k = 0; k < gmvs::K_NUM; ++k) {
        if (p.count[k] == 0) continue;
        if (out && n < max_out) {
            std::snprintf(out[n].name, sizeof(out[n].name), "%s", gmvs::kKernelNames[k]);
            out[n].launches = p.count[k];
            out[n].total_ms = p.ms[k];
            out[n].ncc = work[2 * k];
            out[n].geom = work[2 * k + 1];
        }
        ++n;
    }
    if (num_out) *num_out = n;
    return GMVS_OK;
}

extern "C" int gmvs_profile_reset(void) {
    gmvs::clear_error();
    gmvs::Profiler& p = gmvs::prof();
    std::lock_guard<std::mutex> lk(p.mu);
    for (const gmvs::Pending& q : p.pending) {
        (void)hipEventSynchronize(q.b);
        p.pool.push_back(q.a);
        p.pool.push_back(q.b);
    }
    p.pending.clear();
    for (int k = 0; k < gmvs::K_NUM; ++k) {
        p.ms[k] = 0.0;
        p.count[k] = 0;
    }
    if (p.d_work) {
        GMVS_HIP_CHECK(hipDeviceSynchronize());
        GMVS_HIP_CHECK(hipMemset(p.d_work, 0, sizeof(unsigned long long) * 2 * gmvs::K_NUM));
    }
    return GMVS_OK;
}

extern "C" const char* gmvs_last_error(void) { return gmvs::g_error.c_str(); }

extern "C" const char* gmvs_version(void) { return "gmvs_amd 0.1.0 (gfx950)"; }

// mvs::load_camera (gmvs/src/data_structure.cpp:4-29).
extern "C" int gmvs_load_camera(const float K[9], const float R[9], const float t[3],
                                int32_t height, int32_t width, float depth_min, float depth_max,
                                gmvs_camera* out) {
    gmvs::clear_error();
    GMVS_ARG_CHECK(K && R && t && out, "load_camera: NULL argument");
    std::memcpy(out->K, K, sizeof(float) * 9);
    std::memcpy(out->R, R, sizeof(float) * 9);
    std::memcpy(out->t, t, sizeof(float) * 3);
    out->height = height;
    out->width = width;
    out->depth_min = depth_min;
    out->depth_max = depth_max;
    return GMVS_OK;
}

// mvs::load_problem (gmvs/src/data_structure.cpp:31-39); the reference
// asserts num_src <= MAX_NGB and silently overflows under NDEBUG.
extern "C" int gmvs_load_problem(int32_t ref_id, const int32_t* src_ids, int32_t num_src,
                                 gmvs_problem* out) {
    gmvs::clear_error();
    GMVS_ARG_CHECK(out, "load_problem: out is NULL");
    GMVS_ARG_CHECK(num_src >= 0 && num_src <= GMVS_MAX_NGB,
                   "load_problem: %d source views exceed MAX_NGB=%d", num_src, GMVS_MAX_NGB);
    GMVS_ARG_CHECK(num_src == 0 || src_ids, "load_problem: src_ids is NULL");
    std::memset(out, 0, sizeof(*out));
    out->ref_image_id = ref_id;
    out->num_ngb = num_src;
    for (int i = 0; i < num_src; ++i) out->src_image_ids[i] = src_ids[i];
    return GMVS_OK;
}

extern "C" void gmvs_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    gmvs::philox_block(ctr[0], ctr[1], ctr[2], ctr[3], key[0], key[1], out);
}
