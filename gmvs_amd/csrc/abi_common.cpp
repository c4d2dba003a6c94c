// C-ABI entry points that need no device code: error string, version,
// load_camera / load_problem and the host Philox block.
#include "abi_common.h"

#include <cstring>

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

}  // namespace gmvs

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
