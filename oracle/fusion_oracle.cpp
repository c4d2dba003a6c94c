// Oracle for the gmvs fusion step -- TEST INFRASTRUCTURE ONLY.
//
// FP32 restatement of batch_fusion_kernel
// (gmvs/src/fusion.cu:61-182), the bit-exact reference that the HIP fusion
// kernel's masks and fused point counts are compared with. Built with
// -ffp-contract=off so every product and sum is rounded exactly as written;
// the HIP kernel is written with the same expression trees.
//
// Fixed rules where the reference is racy or hard-coded (the product uses the
// same rules):
//   * thresholds are parameters (reference: 2 px, 0.01, 0.174533 rad,
//     fusion.cu:133);
//   * the angle test acos(dot) < theta is evaluated as dot > cos(theta), and
//     a dot product outside [-1, 1] (or NaN) passes, as acosf's NaN does in
//     get_angle_cu (fusion.cu:50-58);
//   * pixel rounding int(x + 0.5f) is taken as in range when
//     -1 < x + 0.5f < n (truncation toward zero, fusion.cu:112-113);
//   * scatter collisions: the highest reference pixel index wins, which is
//     what the in-order write phase's last writer is (fusion.cu:164-171);
//   * masks[ref] is read for the whole view before any mask of this view is
//     written (the reference only writes masks of source views).
//
// The loop is written once over the geometry type T: float is the
// reference's FP32 (oracle_fuse_view), double the FP64 check of
// gmvs_fuse_view_fp64 (oracle_fuse_view64), the inputs widened exactly.
#include <cmath>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "oracle_common.h"

namespace {

template <class T>
struct P3 {
    T x, y, z;
};

// get_3D_point_on_world_cu (fusion.cu:26-40)
template <class T>
P3<T> world_point(T x, T y, T depth, const OCamera& c) {
    P3<T> p;
    p.x = depth * (x - T(c.K[2])) / T(c.K[0]);
    p.y = depth * (y - T(c.K[5])) / T(c.K[4]);
    p.z = depth;
    P3<T> t;
    t.x = T(c.R[0]) * p.x + T(c.R[3]) * p.y + T(c.R[6]) * p.z;
    t.y = T(c.R[1]) * p.x + T(c.R[4]) * p.y + T(c.R[7]) * p.z;
    t.z = T(c.R[2]) * p.x + T(c.R[5]) * p.y + T(c.R[8]) * p.z;
    P3<T> C;
    C.x = -(T(c.R[0]) * T(c.t[0]) + T(c.R[3]) * T(c.t[1]) + T(c.R[6]) * T(c.t[2]));
    C.y = -(T(c.R[1]) * T(c.t[0]) + T(c.R[4]) * T(c.t[1]) + T(c.R[7]) * T(c.t[2]));
    C.z = -(T(c.R[2]) * T(c.t[0]) + T(c.R[5]) * T(c.t[1]) + T(c.R[8]) * T(c.t[2]));
    P3<T> o;
    o.x = t.x + C.x;
    o.y = t.y + C.y;
    o.z = t.z + C.z;
    return o;
}

// project_on_camera_cu (fusion.cu:42-48)
template <class T>
void project(const P3<T>& X, const OCamera& c, T& u, T& v, T& d) {
    const T t0 = T(c.R[0]) * X.x + T(c.R[1]) * X.y + T(c.R[2]) * X.z + T(c.t[0]);
    const T t1 = T(c.R[3]) * X.x + T(c.R[4]) * X.y + T(c.R[5]) * X.z + T(c.t[1]);
    const T t2 = T(c.R[6]) * X.x + T(c.R[7]) * X.y + T(c.R[8]) * X.z + T(c.t[2]);
    d = T(c.K[6]) * t0 + T(c.K[7]) * t1 + T(c.K[8]) * t2;
    u = (T(c.K[0]) * t0 + T(c.K[1]) * t1 + T(c.K[2]) * t2) / d;
    v = (T(c.K[3]) * t0 + T(c.K[4]) * t1 + T(c.K[5]) * t2) / d;
}

// int(x + 0.5) in [0, n) ?
template <class T>
bool round_in(T x, int n, int* out) {
    const T c = x + T(0.5);
    if (!(c > T(-1) && c < T(n))) return false;
    *out = int(c);
    return true;
}

template <class T>
int64_t fuse_view(const OCamera* cams, int num_images, const float* const* depths,
                  const float* const* normals, uint8_t* masks, int ref_id, const int32_t* src_ids,
                  int geom_consistent, float reproj_px, float rel_depth, float cos_angle,
                  float* out_depth, float* out_normal) {
    const int h = cams[0].height, w = cams[0].width;
    const size_t hw = size_t(h) * w;
    const int nsrc = num_images - 1;
    const OCamera& rc = cams[0];
    std::vector<uint8_t> ref_mask(masks + size_t(ref_id) * hw, masks + size_t(ref_id) * hw + hw);
    // Phase 1, every reference pixel on its own (it reads only the inputs and
    // this view's mask as it was on entry): the consistency checks, the
    // fused point and the source pixels it used. Phase 2, in pixel order: the
    // scatter (so the highest pixel index is the last writer) and the source
    // masks. The arithmetic of each pixel is the sequential loop's.
    std::vector<uint8_t> ok(hw, 0);
    std::vector<int32_t> dest(hw, -1);
    std::vector<float> pdep(hw), pnrm(hw * 3);
    std::vector<int16_t> used(hw * size_t(nsrc) * 2);
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 4)
#endif
    for (int r = 0; r < h; ++r) {
        for (int cc = 0; cc < w; ++cc) {
            const size_t ri = size_t(r) * w + cc;
            int16_t* u_ri = &used[ri * size_t(nsrc) * 2];
            if (ref_mask[ri]) continue;
            const T rd = depths[0][ri];
            const P3<T> rn{normals[0][3 * ri], normals[0][3 * ri + 1], normals[0][3 * ri + 2]};
            if (rd <= T(0)) continue;
            const P3<T> rp = world_point(T(cc), T(r), rd, rc);
            P3<T> cp = rp, cn = rn;
            int num = 0;
            for (int j = 0; j < nsrc; ++j) {
                u_ri[2 * j] = -1;
                u_ri[2 * j + 1] = -1;
                const OCamera& sc = cams[j + 1];
                T u, v, pd;
                project(rp, sc, u, v, pd);
                int sr, scol;
                if (!round_in(v, h, &sr) || !round_in(u, w, &scol)) continue;
                const size_t si = size_t(sr) * w + scol;
                const T sd = depths[j + 1][si];
                const P3<T> sn{normals[j + 1][3 * si], normals[j + 1][3 * si + 1],
                               normals[j + 1][3 * si + 2]};
                if (sd <= T(0)) continue;
                const P3<T> sp = world_point(T(scol), T(sr), sd, sc);
                T tu, tv;
                project(sp, rc, tu, tv, pd);
                const T dx = T(cc) - tu;
                const T dy = T(r) - tv;
                const T err = std::sqrt(dx * dx + dy * dy);
                const T rel = std::fabs(pd - rd) / rd;
                const T dot = rn.x * sn.x + rn.y * sn.y + rn.z * sn.z;
                const bool ang_ok = !(dot >= T(-1) && dot <= T(1)) || dot > T(cos_angle);
                if (err < T(reproj_px) && rel < T(rel_depth) && ang_ok) {
                    cp.x += sp.x;
                    cp.y += sp.y;
                    cp.z += sp.z;
                    cn.x += sn.x;
                    cn.y += sn.y;
                    cn.z += sn.z;
                    u_ri[2 * j] = int16_t(scol);
                    u_ri[2 * j + 1] = int16_t(sr);
                    ++num;
                }
            }
            if (num < geom_consistent) continue;
            ok[ri] = 1;
            const T den = T(num) + T(1);
            cp.x /= den;
            cp.y /= den;
            cp.z /= den;
            cn.x /= den;
            cn.y /= den;
            cn.z /= den;
            T pu, pv, pdp;
            project(cp, rc, pu, pv, pdp);
            int pr, pc;
            if (round_in(pv, h, &pr) && round_in(pu, w, &pc) && pdp > T(0.001f)) {
                dest[ri] = int32_t(size_t(pr) * w + pc);
                pdep[ri] = float(pdp);
                pnrm[3 * ri] = float(cn.x);
                pnrm[3 * ri + 1] = float(cn.y);
                pnrm[3 * ri + 2] = float(cn.z);
            }
        }
    }
    int64_t consistent_pixels = 0;
    for (size_t ri = 0; ri < hw; ++ri) {
        if (!ok[ri]) continue;
        ++consistent_pixels;
        if (dest[ri] >= 0) {
            const size_t pi = size_t(dest[ri]);
            out_depth[pi] = pdep[ri];
            out_normal[3 * pi] = pnrm[3 * ri];
            out_normal[3 * pi + 1] = pnrm[3 * ri + 1];
            out_normal[3 * pi + 2] = pnrm[3 * ri + 2];
        }
        const int16_t* u_ri = &used[ri * size_t(nsrc) * 2];
        for (int j = 0; j < nsrc; ++j) {
            if (u_ri[2 * j] == -1) continue;
            masks[size_t(src_ids[j]) * hw + size_t(u_ri[2 * j + 1]) * w + u_ri[2 * j]] = 1;
        }
    }
    return consistent_pixels;
}

}  // namespace

extern "C" {

// Fuses one reference view. cams/depths/normals hold the reference first and
// then its num_images-1 sources; masks is the global [N, H, W] byte mask,
// indexed with ref_id and src_ids (global view ids). out_depth [H*W] and
// out_normal [H*W*3] must be zeroed by the caller. Returns the number of
// consistent reference pixels (before the scatter).
int64_t oracle_fuse_view(const OCamera* cams, int num_images, const float* const* depths,
                         const float* const* normals, uint8_t* masks, int ref_id,
                         const int32_t* src_ids, int geom_consistent, float reproj_px,
                         float rel_depth, float cos_angle, float* out_depth, float* out_normal) {
    return fuse_view<float>(cams, num_images, depths, normals, masks, ref_id, src_ids,
                            geom_consistent, reproj_px, rel_depth, cos_angle, out_depth,
                            out_normal);
}

// The same with the geometry in FP64 (gmvs_fuse_view_fp64).
int64_t oracle_fuse_view64(const OCamera* cams, int num_images, const float* const* depths,
                           const float* const* normals, uint8_t* masks, int ref_id,
                           const int32_t* src_ids, int geom_consistent, float reproj_px,
                           float rel_depth, float cos_angle, float* out_depth, float* out_normal) {
    return fuse_view<double>(cams, num_images, depths, normals, masks, ref_id, src_ids,
                             geom_consistent, reproj_px, rel_depth, cos_angle, out_depth,
                             out_normal);
}

}  // extern "C"
