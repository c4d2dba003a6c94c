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
#include <cmath>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "oracle_common.h"

namespace {

struct P3 {
    float x, y, z;
};

P3 world_point(float x, float y, float depth, const OCamera& c) {
    P3 p;
    p.x = depth * (x - c.K[2]) / c.K[0];
    p.y = depth * (y - c.K[5]) / c.K[4];
    p.z = depth;
    P3 t;
    t.x = c.R[0] * p.x + c.R[3] * p.y + c.R[6] * p.z;
    t.y = c.R[1] * p.x + c.R[4] * p.y + c.R[7] * p.z;
    t.z = c.R[2] * p.x + c.R[5] * p.y + c.R[8] * p.z;
    P3 C;
    C.x = -(c.R[0] * c.t[0] + c.R[3] * c.t[1] + c.R[6] * c.t[2]);
    C.y = -(c.R[1] * c.t[0] + c.R[4] * c.t[1] + c.R[7] * c.t[2]);
    C.z = -(c.R[2] * c.t[0] + c.R[5] * c.t[1] + c.R[8] * c.t[2]);
    P3 o;
    o.x = t.x + C.x;
    o.y = t.y + C.y;
    o.z = t.z + C.z;
    return o;
}

void project(const P3& X, const OCamera& c, float& u, float& v, float& d) {
    const float t0 = c.R[0] * X.x + c.R[1] * X.y + c.R[2] * X.z + c.t[0];
    const float t1 = c.R[3] * X.x + c.R[4] * X.y + c.R[5] * X.z + c.t[1];
    const float t2 = c.R[6] * X.x + c.R[7] * X.y + c.R[8] * X.z + c.t[2];
    d = c.K[6] * t0 + c.K[7] * t1 + c.K[8] * t2;
    u = (c.K[0] * t0 + c.K[1] * t1 + c.K[2] * t2) / d;
    v = (c.K[3] * t0 + c.K[4] * t1 + c.K[5] * t2) / d;
}

// int(x + 0.5f) in [0, n) ?
bool round_in(float x, int n, int* out) {
    const float c = x + 0.5f;
    if (!(c > -1.0f && c < float(n))) return false;
    *out = int(c);
    return true;
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
            const float rd = depths[0][ri];
            const P3 rn{normals[0][3 * ri], normals[0][3 * ri + 1], normals[0][3 * ri + 2]};
            if (rd <= 0.0f) continue;
            const P3 rp = world_point(float(cc), float(r), rd, rc);
            P3 cp = rp, cn = rn;
            int num = 0;
            for (int j = 0; j < nsrc; ++j) {
                u_ri[2 * j] = -1;
                u_ri[2 * j + 1] = -1;
                const OCamera& sc = cams[j + 1];
                float u, v, pd;
                project(rp, sc, u, v, pd);
                int sr, scol;
                if (!round_in(v, h, &sr) || !round_in(u, w, &scol)) continue;
                const size_t si = size_t(sr) * w + scol;
                const float sd = depths[j + 1][si];
                const P3 sn{normals[j + 1][3 * si], normals[j + 1][3 * si + 1],
                            normals[j + 1][3 * si + 2]};
                if (sd <= 0.0f) continue;
                const P3 sp = world_point(float(scol), float(sr), sd, sc);
                float tu, tv;
                project(sp, rc, tu, tv, pd);
                const float dx = float(cc) - tu;
                const float dy = float(r) - tv;
                const float err = std::sqrt(dx * dx + dy * dy);
                const float rel = std::fabs(pd - rd) / rd;
                const float dot = rn.x * sn.x + rn.y * sn.y + rn.z * sn.z;
                const bool ang_ok = !(dot >= -1.0f && dot <= 1.0f) || dot > cos_angle;
                if (err < reproj_px && rel < rel_depth && ang_ok) {
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
            const float den = float(num) + 1.0f;
            cp.x /= den;
            cp.y /= den;
            cp.z /= den;
            cn.x /= den;
            cn.y /= den;
            cn.z /= den;
            float pu, pv, pdp;
            project(cp, rc, pu, pv, pdp);
            int pr, pc;
            if (round_in(pv, h, &pr) && round_in(pu, w, &pc) && pdp > 0.001f) {
                dest[ri] = int32_t(size_t(pr) * w + pc);
                pdep[ri] = pdp;
                pnrm[3 * ri] = cn.x;
                pnrm[3 * ri + 1] = cn.y;
                pnrm[3 * ri + 2] = cn.z;
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

}  // extern "C"
