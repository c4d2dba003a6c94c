// Multi-view geometric-consistency fusion of depth/normal maps (gfx950).
//
// Replaces batch_fusion_kernel + Fuser::run_fusion (gmvs/src/fusion.cu:61-182,
// 211-275). The arithmetic is the reference's FP32 with every product and
// sum rounded as written (this file is compiled with -ffp-contract=off and
// correctly rounded division/sqrt), so masks and fused counts are bit-exact
// with a sequential CPU loop over the same expressions. Two fixed rules make
// the result deterministic where the reference races:
//   * scatter collisions (several reference pixels reprojecting to one output
//     pixel, fusion.cu:164-171): the highest reference pixel index wins, via
//     an atomicMax on the pixel index followed by a second pass that lets
//     only the winner write;
//   * masks[ref] is read from a copy taken before any mask of the view is
//     written.
// The angle test acos(dot) < theta is evaluated as dot > cos(theta), with a
// dot outside [-1, 1] (or NaN) passing as acosf's NaN does in get_angle_cu
// (fusion.cu:50-58). gmvs_fuse_view_fp64 runs the same check with the world
// points, projections and tests in FP64 (not in the reference), bit-exact
// with the oracle's FP64 instantiation of the same loop.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

#include "abi_common.h"

#pragma clang fp contract(off)

namespace gmvs {
namespace {

struct FuseView {
    gmvs_camera cam;
    const float* depth;
    const float* normal;
    int32_t global_id;
};

// The geometry is evaluated in T: float (the reference's FP32, the default)
// or double (gmvs_fuse_view_fp64: the world points and projections in FP64,
// for scenes whose world coordinates are large against their depths, where
// FP32 world points carry reprojection errors near the 2 px threshold). The
// camera parameters, depths and normals are FP32 inputs widened exactly.
template <class T>
struct F3 {
    T x, y, z;
};

template <class T>
__device__ inline F3<T> fz_world_point(T x, T y, T depth, const gmvs_camera& c) {
    F3<T> p;
    p.x = depth * (x - (T)c.K[2]) / (T)c.K[0];
    p.y = depth * (y - (T)c.K[5]) / (T)c.K[4];
    p.z = depth;
    F3<T> t;
    t.x = (T)c.R[0] * p.x + (T)c.R[3] * p.y + (T)c.R[6] * p.z;
    t.y = (T)c.R[1] * p.x + (T)c.R[4] * p.y + (T)c.R[7] * p.z;
    t.z = (T)c.R[2] * p.x + (T)c.R[5] * p.y + (T)c.R[8] * p.z;
    F3<T> C;
    C.x = -((T)c.R[0] * (T)c.t[0] + (T)c.R[3] * (T)c.t[1] + (T)c.R[6] * (T)c.t[2]);
    C.y = -((T)c.R[1] * (T)c.t[0] + (T)c.R[4] * (T)c.t[1] + (T)c.R[7] * (T)c.t[2]);
    C.z = -((T)c.R[2] * (T)c.t[0] + (T)c.R[5] * (T)c.t[1] + (T)c.R[8] * (T)c.t[2]);
    F3<T> o;
    o.x = t.x + C.x;
    o.y = t.y + C.y;
    o.z = t.z + C.z;
    return o;
}

template <class T>
__device__ inline void fz_project(const F3<T>& X, const gmvs_camera& c, T& u, T& v, T& d) {
    const T t0 = (T)c.R[0] * X.x + (T)c.R[1] * X.y + (T)c.R[2] * X.z + (T)c.t[0];
    const T t1 = (T)c.R[3] * X.x + (T)c.R[4] * X.y + (T)c.R[5] * X.z + (T)c.t[1];
    const T t2 = (T)c.R[6] * X.x + (T)c.R[7] * X.y + (T)c.R[8] * X.z + (T)c.t[2];
    d = (T)c.K[6] * t0 + (T)c.K[7] * t1 + (T)c.K[8] * t2;
    u = ((T)c.K[0] * t0 + (T)c.K[1] * t1 + (T)c.K[2] * t2) / d;
    v = ((T)c.K[3] * t0 + (T)c.K[4] * t1 + (T)c.K[5] * t2) / d;
}

// int(x + 0.5) in [0, n): truncation toward zero (fusion.cu:112-113).
template <class T>
__device__ inline bool fz_round_in(T x, int n, int* out) {
    const T c = x + (T)0.5;
    if (!(c > (T)-1 && c < (T)n)) return false;
    *out = (int)c;
    return true;
}

// The view table goes to the device as kernel arguments, up to kViewChunk
// views per launch (2 KB of arguments), so a call needs no host copy.
constexpr int kViewChunk = 16;
struct FuseViewChunk {
    FuseView v[kViewChunk];
};

__global__ void fuse_write_views(FuseViewChunk c, int n, FuseView* __restrict__ dst) {
    const int i = threadIdx.x;
    if (blockIdx.x == 0 && i < n) dst[i] = c.v[i];
}

__global__ void fuse_fill(int hw, float* __restrict__ depth, float* __restrict__ normal,
                          int32_t* __restrict__ winner, const uint8_t* __restrict__ mask_src,
                          uint8_t* __restrict__ mask_copy) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= hw) return;
    depth[i] = 0.f;
    normal[3 * i] = 0.f;
    normal[3 * i + 1] = 0.f;
    normal[3 * i + 2] = 0.f;
    winner[i] = -1;
    mask_copy[i] = mask_src[i];
}

// Pass 1: per reference pixel, the consistency check against every source,
// the averaged point/normal, its reprojection and the mask updates.
template <class T>
__global__ __launch_bounds__(256) void fuse_check(const FuseView* __restrict__ views, int nsrc,
                                                  int h, int w, int geom_consistent,
                                                  float reproj_px, float rel_depth,
                                                  float cos_angle,
                                                  const uint8_t* __restrict__ ref_mask,
                                                  uint8_t* __restrict__ masks,
                                                  int32_t* __restrict__ winner,
                                                  int32_t* __restrict__ rec_idx,
                                                  float4* __restrict__ rec_val) {
    const int ri = blockIdx.x * blockDim.x + threadIdx.x;
    if (ri >= h * w) return;
    rec_idx[ri] = -1;
    if (ref_mask[ri]) return;
    const gmvs_camera& rc = views[0].cam;
    const T rd = views[0].depth[ri];
    const F3<T> rn{views[0].normal[3 * ri], views[0].normal[3 * ri + 1],
                   views[0].normal[3 * ri + 2]};
    if (rd <= (T)0) return;
    const int r = ri / w, cc = ri - r * w;
    const F3<T> rp = fz_world_point((T)cc, (T)r, rd, rc);
    F3<T> cp = rp, cn = rn;
    int num = 0;
    uint32_t used_mask = 0;
    for (int j = 0; j < nsrc; ++j) {
        const FuseView& sv = views[j + 1];
        T u, v, pd;
        fz_project(rp, sv.cam, u, v, pd);
        int sr, scol;
        if (!fz_round_in(v, h, &sr) || !fz_round_in(u, w, &scol)) continue;
        const int si = sr * w + scol;
        const T sd = sv.depth[si];
        const F3<T> sn{sv.normal[3 * si], sv.normal[3 * si + 1], sv.normal[3 * si + 2]};
        if (sd <= (T)0) continue;
        const F3<T> sp = fz_world_point((T)scol, (T)sr, sd, sv.cam);
        T tu, tv;
        fz_project(sp, rc, tu, tv, pd);
        const T dx = (T)cc - tu;
        const T dy = (T)r - tv;
        const T err = sqrt(dx * dx + dy * dy);
        const T rel = fabs(pd - rd) / rd;
        const T dot = rn.x * sn.x + rn.y * sn.y + rn.z * sn.z;
        const bool ang_ok = !(dot >= (T)-1 && dot <= (T)1) || dot > (T)cos_angle;
        if (err < (T)reproj_px && rel < (T)rel_depth && ang_ok) {
            cp.x += sp.x;
            cp.y += sp.y;
            cp.z += sp.z;
            cn.x += sn.x;
            cn.y += sn.y;
            cn.z += sn.z;
            used_mask |= 1u << j;
            ++num;
        }
    }
    if (num < geom_consistent) return;
    const T den = (T)num + (T)1;
    cp.x /= den;
    cp.y /= den;
    cp.z /= den;
    cn.x /= den;
    cn.y /= den;
    cn.z /= den;
    T pu, pv, pdep;
    fz_project(cp, rc, pu, pv, pdep);
    int pr, pc;
    if (fz_round_in(pv, h, &pr) && fz_round_in(pu, w, &pc) && pdep > (T)0.001f) {
        const int pi = pr * w + pc;
        atomicMax(&winner[pi], ri);
        rec_idx[ri] = pi;
        rec_val[ri] = make_float4((float)pdep, (float)cn.x, (float)cn.y, (float)cn.z);
    }
    // mark the used source pixels (recomputed: the same arithmetic as above)
    const size_t hw = (size_t)h * w;
    for (int j = 0; j < nsrc; ++j) {
        if (!((used_mask >> j) & 1u)) continue;
        const FuseView& sv = views[j + 1];
        T u, v, pd;
        fz_project(rp, sv.cam, u, v, pd);
        int sr, scol;
        fz_round_in(v, h, &sr);
        fz_round_in(u, w, &scol);
        masks[(size_t)sv.global_id * hw + (size_t)sr * w + scol] = 1;
    }
}

// Pass 2: the winning (highest-index) reference pixel writes each output pixel.
__global__ void fuse_scatter(int hw, const int32_t* __restrict__ winner,
                             const int32_t* __restrict__ rec_idx,
                             const float4* __restrict__ rec_val, float* __restrict__ depth,
                             float* __restrict__ normal) {
    const int ri = blockIdx.x * blockDim.x + threadIdx.x;
    if (ri >= hw) return;
    const int pi = rec_idx[ri];
    if (pi < 0 || winner[pi] != ri) return;
    const float4 v = rec_val[ri];
    depth[pi] = v.x;
    normal[3 * pi] = v.y;
    normal[3 * pi + 1] = v.z;
    normal[3 * pi + 2] = v.w;
}

constexpr size_t align_up(size_t x) { return (x + 255) & ~size_t(255); }

struct FuseLayout {
    size_t views, mask_copy, winner, rec_idx, rec_val, total;
};

constexpr FuseLayout fuse_layout(size_t hw) {
    FuseLayout L{};
    size_t o = 0;
    L.views = o;
    o = align_up(o + sizeof(FuseView) * (GMVS_MAX_NGB + 1));
    L.mask_copy = o;
    o = align_up(o + hw);
    L.winner = o;
    o = align_up(o + 4 * hw);
    L.rec_idx = o;
    o = align_up(o + 4 * hw);
    L.rec_val = o;
    o = align_up(o + 16 * hw);
    L.total = o;
    return L;
}

}  // namespace
}  // namespace gmvs

using namespace gmvs;

extern "C" size_t gmvs_fuse_workspace_bytes(int32_t height, int32_t width) {
    if (height <= 0 || width <= 0) return 0;
    return fuse_layout(size_t(height) * width).total;
}

namespace gmvs {
namespace {

int fuse_view_impl(bool fp64, int32_t num_images, const gmvs_camera* cameras,
                   const float* const* depths, const float* const* normals, uint8_t* masks,
                   int32_t num_mask_views, int32_t ref_id, const int32_t* src_ids,
                   int32_t geom_consistent, float reproj_px, float rel_depth, float cos_angle,
                   float* out_depth, float* out_normal, void* workspace, size_t workspace_bytes,
                   void* stream) {
    GMVS_ARG_CHECK(cameras && depths && normals && masks && out_depth && out_normal,
                   "fuse_view: NULL argument");
    GMVS_ARG_CHECK(num_images >= 1 && num_images <= GMVS_MAX_NGB + 1,
                   "fuse_view: num_images must be in [1, %d] (got %d)", GMVS_MAX_NGB + 1,
                   num_images);
    GMVS_ARG_CHECK(num_images == 1 || src_ids, "fuse_view: src_ids is NULL");
    GMVS_ARG_CHECK(ref_id >= 0 && ref_id < num_mask_views,
                   "fuse_view: ref_id %d outside the %d mask views", ref_id, num_mask_views);
    const int h = cameras[0].height, w = cameras[0].width;
    GMVS_ARG_CHECK(h > 0 && w > 0, "fuse_view: bad image size %dx%d", w, h);
    for (int i = 0; i < num_images; ++i) {
        GMVS_ARG_CHECK(depths[i] && normals[i], "fuse_view: depths/normals[%d] is NULL", i);
        GMVS_ARG_CHECK(cameras[i].height == h && cameras[i].width == w,
                       "fuse_view: view %d is %dx%d, reference %dx%d", i, cameras[i].width,
                       cameras[i].height, w, h);
        if (i)
            GMVS_ARG_CHECK(src_ids[i - 1] >= 0 && src_ids[i - 1] < num_mask_views,
                           "fuse_view: src id %d outside the %d mask views", src_ids[i - 1],
                           num_mask_views);
    }
    const FuseLayout L = fuse_layout(size_t(h) * w);
    GMVS_ARG_CHECK(workspace, "fuse_view: workspace is NULL");
    if (workspace_bytes < L.total) {
        set_error("fuse_view: workspace too small: %zu < %zu bytes", workspace_bytes, L.total);
        return GMVS_ERR_NOSPACE;
    }
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    char* ws = reinterpret_cast<char*>(workspace);
    FuseView* vtab = reinterpret_cast<FuseView*>(ws + L.views);
    for (int i0 = 0; i0 < num_images; i0 += kViewChunk) {
        FuseViewChunk chunk{};
        const int n = std::min(kViewChunk, num_images - i0);
        for (int k = 0; k < n; ++k) {
            const int i = i0 + k;
            FuseView& fv = chunk.v[k];
            fv.cam = cameras[i];
            fv.depth = depths[i];
            fv.normal = normals[i];
            fv.global_id = i ? src_ids[i - 1] : ref_id;
        }
        GMVS_LAUNCH(K_FUSE_VIEWDESC, st, fuse_write_views, dim3(1), dim3(kViewChunk), 0, st, chunk,
                    n, vtab + i0);
        GMVS_LAUNCH_CHECK();
    }
    const int hw = h * w;
    const dim3 grid((hw + 255) / 256), block(256);
    uint8_t* mask_copy = reinterpret_cast<uint8_t*>(ws + L.mask_copy);
    int32_t* winner = reinterpret_cast<int32_t*>(ws + L.winner);
    int32_t* rec_idx = reinterpret_cast<int32_t*>(ws + L.rec_idx);
    float4* rec_val = reinterpret_cast<float4*>(ws + L.rec_val);
    GMVS_LAUNCH(K_FUSE_FILL, st, fuse_fill, grid, block, 0, st, hw, out_depth, out_normal, winner,
                       masks + (size_t)ref_id * hw, mask_copy);
    GMVS_LAUNCH_CHECK();
    auto check = fp64 ? fuse_check<double> : fuse_check<float>;
    GMVS_LAUNCH(K_FUSE_CHECK, st, check, grid, block, 0, st,
                       reinterpret_cast<const FuseView*>(ws + L.views), num_images - 1, h, w,
                       geom_consistent, reproj_px, rel_depth, cos_angle, mask_copy, masks, winner,
                       rec_idx, rec_val);
    GMVS_LAUNCH_CHECK();
    GMVS_LAUNCH(K_FUSE_SCATTER, st, fuse_scatter, grid, block, 0, st, hw, winner, rec_idx, rec_val, out_depth,
                       out_normal);
    GMVS_LAUNCH_CHECK();
    return GMVS_OK;
}

}  // namespace
}  // namespace gmvs

extern "C" int gmvs_fuse_view(int32_t num_images, const gmvs_camera* cameras,
                              const float* const* depths, const float* const* normals,
                              uint8_t* masks, int32_t num_mask_views, int32_t ref_id,
                              const int32_t* src_ids,
                              int32_t geom_consistent, float reproj_px, float rel_depth,
                              float cos_angle, float* out_depth, float* out_normal,
                              void* workspace, size_t workspace_bytes, void* stream) {
    clear_error();
    return fuse_view_impl(false, num_images, cameras, depths, normals, masks, num_mask_views,
                          ref_id, src_ids, geom_consistent, reproj_px, rel_depth, cos_angle,
                          out_depth, out_normal, workspace, workspace_bytes, stream);
}

extern "C" int gmvs_fuse_view_fp64(int32_t num_images, const gmvs_camera* cameras,
                                   const float* const* depths, const float* const* normals,
                                   uint8_t* masks, int32_t num_mask_views, int32_t ref_id,
                                   const int32_t* src_ids, int32_t geom_consistent,
                                   float reproj_px, float rel_depth, float cos_angle,
                                   float* out_depth, float* out_normal, void* workspace,
                                   size_t workspace_bytes, void* stream) {
    clear_error();
    return fuse_view_impl(true, num_images, cameras, depths, normals, masks, num_mask_views,
                          ref_id, src_ids, geom_consistent, reproj_px, rel_depth, cos_angle,
                          out_depth, out_normal, workspace, workspace_bytes, stream);
}

// ---- no-alloc variant ------------------------------------------------------
// gmvs_fuse_view with its scratch (view table, reference-mask snapshot, winner
// indices, fused records: 25 B per pixel) in a file-scope device array instead
// of a caller workspace, for callers that cannot allocate. One copy of the
// array per device; every call on a device reuses it (each call rewrites all
// of it that it reads), so calls that overlap on one device race.
namespace gmvs {
namespace {

constexpr int64_t kStaticFusePixels = 2048 * 1536;
constexpr size_t kStaticFuseBytes = fuse_layout(size_t(kStaticFusePixels)).total;
constexpr int kMaxDevices = 64;

}  // namespace
}  // namespace gmvs

static __device__ __align__(256) unsigned char scratch[gmvs::kStaticFuseBytes];

namespace gmvs {
namespace {

std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device, looked up once per
// device (the array's host-side name is not a device address).
int static_scratch(void** out) {
    int dev = 0;
    GMVS_HIP_CHECK(hipGetDevice(&dev));
    GMVS_ARG_CHECK(dev >= 0 && dev < kMaxDevices, "fuse_view_noalloc: device %d beyond %d", dev,
                   kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        GMVS_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        GMVS_ARG_CHECK(p != nullptr, "fuse_view_noalloc: no device address for the scratch array");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return GMVS_OK;
}

}  // namespace
}  // namespace gmvs

extern "C" int64_t gmvs_fuse_noalloc_max_pixels(void) { return kStaticFusePixels; }

extern "C" int gmvs_fuse_view_noalloc(int32_t num_images, const gmvs_camera* cameras,
                                      const float* const* depths, const float* const* normals,
                                      uint8_t* masks, int32_t num_mask_views, int32_t ref_id,
                                      const int32_t* src_ids, int32_t geom_consistent,
                                      float reproj_px, float rel_depth, float cos_angle,
                                      float* out_depth, float* out_normal, void* stream) {
    clear_error();
    GMVS_ARG_CHECK(cameras, "fuse_view_noalloc: cameras is NULL");
    const int64_t hw = int64_t(cameras[0].height) * cameras[0].width;
    GMVS_ARG_CHECK(hw <= kStaticFusePixels,
                   "fuse_view_noalloc: %dx%d = %lld pixels exceeds the static scratch limit of "
                   "%lld pixels (use gmvs_fuse_view with a workspace)",
                   cameras[0].width, cameras[0].height, (long long)hw,
                   (long long)kStaticFusePixels);
    GMVS_ARG_CHECK(!profiling_enabled(),
                   "fuse_view_noalloc: kernel profiling creates timing events; disable it first");
    void* ws = nullptr;
    const int rc = static_scratch(&ws);
    if (rc != GMVS_OK) return rc;
    return gmvs_fuse_view(num_images, cameras, depths, normals, masks, num_mask_views, ref_id,
                          src_ids, geom_consistent, reproj_px, rel_depth, cos_angle, out_depth,
                          out_normal, ws, kStaticFuseBytes, stream);
}
