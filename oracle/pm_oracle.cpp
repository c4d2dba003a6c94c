// Oracle for the gmvs PatchMatch MVS hot path -- TEST INFRASTRUCTURE ONLY.
//
// A CPU restatement of the reference's PatchMatch arithmetic, templated over
// the scalar type T. T = double is the FP64 oracle every GPU result is
// checked against; T = float (with OpenMP over the same-colour pixels of each
// half-sweep) is the host-CPU baseline that bench.py times. Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
//
// Semantics follow gmvs/src/patch_match.cu ("src") for the NCC (software
// bilinear, clamped reference taps), the geometric term and the init from
// depth/normal, and gmvs/lib/src/patch_match_kernel.cu ("lib") for random
// init, the planar prior and the plain first stage. Where the reference is
// racy or draws from curand, the oracle uses the fixed rules the product
// also uses:
//   * snapshot reads: every neighbour cost/plane read in a half-sweep sees
//     the state at the start of that half-sweep (lib 978-1104 reads
//     same-colour pixels that are being rewritten in the same launch);
//   * Philox4x32-10 with fixed draw slots (oracle_common.h) replaces
//     curand_init(3407, y, x) / curand_uniform (lib 1491, src 1069);
//   * uniform draws are ((x >> 8) + 1) * 2^-24, in (0, 1] like curand_uniform.
// Reference quirks that are mirrored on purpose (each also mirrored by the
// product): cost_array[8][32] = {2.0f} leaves all entries but [0][0] at zero
// (lib 854, src 717); the right_far probe keeps the MAX cost (lib 954,
// src 794); the do/while perturbed-depth loop runs once (lib 729-734); the
// prior pass reuses world-frame normals as camera-frame (lib 1546-1550); the
// prior path leaves selected_views alone for unmasked adoptions
// (lib 1306-1318); the restricted cost starts at 0 when the masked branch
// adopts nothing (lib 1240, consumed at lib 784-801).
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <vector>

#include "oracle_common.h"

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr double kPi = 3.14159265358979323846;

template <class T>
struct Cam {
    T K[9], R[9], t[3];
    int h, w;
};

template <class T>
Cam<T> to_cam(const OCamera& c) {
    Cam<T> o;
    for (int i = 0; i < 9; ++i) {
        o.K[i] = T(c.K[i]);
        o.R[i] = T(c.R[i]);
    }
    for (int i = 0; i < 3; ++i) o.t[i] = T(c.t[i]);
    o.h = c.height;
    o.w = c.width;
    return o;
}

template <class T>
struct V4 {
    T x, y, z, w;
};

// ------------------------------------------------------------ RNG slots ---
struct Rng {
    uint32_t key[2];
    uint32_t pixel;
    uint32_t phase;
    // Returns the uniform in (0, 1] of one draw slot.
    double u(int slot) const {
        uint32_t ctr[4] = {pixel, phase, (uint32_t)(slot / 4), 0u};
        uint32_t out[4];
        oracle_philox4x32_10_block(ctr, key, out);
        return double((out[slot % 4] >> 8) + 1u) * (1.0 / 16777216.0);
    }
};

// ------------------------------------------------- camera-frame geometry ---
// get_3D_point / get_distance_to_origin / compute_depth_from_plane_hypothesis
// (src 94-134; lib 85-135).
template <class T>
T dist_to_origin(const Cam<T>& c, int px, int py, T depth, const V4<T>& n) {
    const T X0 = depth * (T(px) - c.K[2]) / c.K[0];
    const T X1 = depth * (T(py) - c.K[5]) / c.K[4];
    const T X2 = depth;
    return -(n.x * X0 + n.y * X1 + n.z * X2);
}

template <class T>
T depth_from_plane(const Cam<T>& c, const V4<T>& pl, int px, int py) {
    return -pl.w * c.K[0] /
           ((T(px) - c.K[2]) * pl.x + (c.K[0] / c.K[4]) * (T(py) - c.K[5]) * pl.y + c.K[0] * pl.z);
}

template <class T>
void view_dir(const Cam<T>& c, int px, int py, T depth, T out[3]) {
    const T X0 = depth * (T(px) - c.K[2]) / c.K[0];
    const T X1 = depth * (T(py) - c.K[5]) / c.K[4];
    const T X2 = depth;
    const T nrm = std::sqrt(X0 * X0 + X1 * X1 + X2 * X2);
    out[0] = X0 / nrm;
    out[1] = X1 / nrm;
    out[2] = X2 / nrm;
}

template <class T>
void normalize3(V4<T>& v) {
    const T s = T(1) / std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
    v.x *= s;
    v.y *= s;
    v.z *= s;
}

// generate_random_normal (src 136-168; lib 137-172): Marsaglia's method with
// a fixed number of draw pairs; if every pair is rejected the first pair is
// rescaled onto s = 1/4 (probability ~2e-11 per call).
template <class T>
V4<T> random_normal(const Cam<T>& c, int px, int py, const Rng& rng, int slot0, T depth) {
    T q1 = 0, q2 = 0, s = 2;
    bool ok = false;
    for (int a = 0; a < ORNG_REJECT_TRIES && !ok; ++a) {
        q1 = T(2) * T(rng.u(slot0 + 2 * a)) - T(1);
        q2 = T(2) * T(rng.u(slot0 + 2 * a + 1)) - T(1);
        s = q1 * q1 + q2 * q2;
        ok = s < T(1);
    }
    if (!ok) {
        q1 = T(2) * T(rng.u(slot0)) - T(1);
        q2 = T(2) * T(rng.u(slot0 + 1)) - T(1);
        const T r = std::sqrt(q1 * q1 + q2 * q2);
        if (r > T(0)) {
            q1 = q1 / r * T(0.5);
            q2 = q2 / r * T(0.5);
        } else {
            q1 = T(0.5);
            q2 = T(0);
        }
        s = q1 * q1 + q2 * q2;
    }
    const T sq = std::sqrt(T(1) - s);
    V4<T> n{T(2) * q1 * sq, T(2) * q2 * sq, T(1) - T(2) * s, T(0)};
    T vd[3];
    view_dir(c, px, py, depth, vd);
    if (n.x * vd[0] + n.y * vd[1] + n.z * vd[2] > T(0)) {
        n.x = -n.x;
        n.y = -n.y;
        n.z = -n.z;
    }
    normalize3(n);
    return n;
}

// generate_perturbed_normal (src 170-209; lib 174-213).
template <class T>
V4<T> perturbed_normal(const Cam<T>& c, int px, int py, const V4<T>& n, const Rng& rng,
                       int slot0, T pert) {
    T vd[3];
    view_dir(c, px, py, T(1), vd);
    const T a1 = (T(rng.u(slot0 + 0)) - T(0.5)) * pert;
    const T a2 = (T(rng.u(slot0 + 1)) - T(0.5)) * pert;
    const T a3 = (T(rng.u(slot0 + 2)) - T(0.5)) * pert;
    const T s1 = std::sin(a1), s2 = std::sin(a2), s3 = std::sin(a3);
    const T c1 = std::cos(a1), c2 = std::cos(a2), c3 = std::cos(a3);
    const T R[9] = {c2 * c3, c3 * s1 * s2 - c1 * s3, s1 * s3 + c1 * c3 * s2,
                    c2 * s3, c1 * c3 + s1 * s2 * s3, c1 * s2 * s3 - c3 * s1,
                    -s2,     c2 * s1,                c1 * c2};
    V4<T> o{R[0] * n.x + R[1] * n.y + R[2] * n.z, R[3] * n.x + R[4] * n.y + R[5] * n.z,
            R[6] * n.x + R[7] * n.y + R[8] * n.z, n.w};
    if (o.x * vd[0] + o.y * vd[1] + o.z * vd[2] >= T(0)) o = n;
    normalize3(o);
    return o;
}

// ------------------------------------------------------------ homography ---
// compute_homography (src 211-296): H = K_s (R_s R_r^T - t_rel n^T / d) K_r^-1
// with t_rel = R_s (C_r - C_s), C = -R^T t. As in the reference, K_r^-1 uses
// fx, fy, cx, cy only and K_s contributes K0, K2, K4, K5, K8.
template <class T>
void homography(const Cam<T>& r, const Cam<T>& s, const V4<T>& pl, T H[9]) {
    T Cr[3], Cs[3];
    for (int i = 0; i < 3; ++i) {
        Cr[i] = -(r.R[i] * r.t[0] + r.R[3 + i] * r.t[1] + r.R[6 + i] * r.t[2]);
        Cs[i] = -(s.R[i] * s.t[0] + s.R[3 + i] * s.t[1] + s.R[6 + i] * s.t[2]);
    }
    T Rrel[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            Rrel[3 * i + j] = s.R[3 * i] * r.R[3 * j] + s.R[3 * i + 1] * r.R[3 * j + 1] +
                              s.R[3 * i + 2] * r.R[3 * j + 2];
    const T Crel[3] = {Cr[0] - Cs[0], Cr[1] - Cs[1], Cr[2] - Cs[2]};
    T trel[3];
    for (int i = 0; i < 3; ++i)
        trel[i] = s.R[3 * i] * Crel[0] + s.R[3 * i + 1] * Crel[1] + s.R[3 * i + 2] * Crel[2];
    T M[9];
    const T nv[3] = {pl.x, pl.y, pl.z};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) M[3 * i + j] = Rrel[3 * i + j] - trel[i] * nv[j] / pl.w;
    T tmp[9];
    for (int i = 0; i < 3; ++i) {
        tmp[3 * i + 0] = M[3 * i + 0] / r.K[0];
        tmp[3 * i + 1] = M[3 * i + 1] / r.K[4];
        tmp[3 * i + 2] = -M[3 * i + 0] * r.K[2] / r.K[0] - M[3 * i + 1] * r.K[5] / r.K[4] +
                         M[3 * i + 2];
    }
    for (int j = 0; j < 3; ++j) {
        H[0 + j] = s.K[0] * tmp[0 + j] + s.K[2] * tmp[6 + j];
        H[3 + j] = s.K[4] * tmp[3 + j] + s.K[5] * tmp[6 + j];
        H[6 + j] = s.K[8] * tmp[6 + j];
    }
}

// Software bilinear sampling (src 10-33): index = clamp(trunc(coord), 0, n-1)
// (coordinates below 1, NaN included, give index 0 and so extrapolate with a
// negative or NaN fraction); the upper neighbour is clamped to n-1.
inline int clamp_index(double c, int n) {
    if (!(c >= 1.0)) return 0;
    if (c >= double(n - 1)) return n - 1;
    return int(c);
}

template <class T>
T bilinear(const float* img, T x, T y, int h, int w) {
    const int hl = clamp_index(double(y), h);
    const int wl = clamp_index(double(x), w);
    const T fy = y - T(hl);
    const T fx = x - T(wl);
    const int hu = (hl == h - 1) ? hl : hl + 1;
    const int wu = (wl == w - 1) ? wl : wl + 1;
    const T v0 = T(img[hl * w + wl]), v1 = T(img[hl * w + wu]);
    const T v2 = T(img[hu * w + wl]), v3 = T(img[hu * w + wu]);
    const T a = v0 + fx * (v1 - v0);
    const T b = v2 + fx * (v3 - v2);
    return a + fy * (b - a);
}

// Diagnostics of one NCC evaluation: distance of the warped centre to the
// source-image bounds (px) and the smaller of the two weighted variances.
struct NccDiag {
    double edge;
    double minvar;
};

// compute_bilateral_NCC (src 362-445).
template <class T>
T bilateral_ncc(const float* ref, const Cam<T>& rc, const float* src, const Cam<T>& sc, int px,
                int py, const V4<T>& pl, const OParams& prm, NccDiag* diag) {
    const T cost_max = T(2);
    const int radius = prm.patch_size / 2;
    T H[9];
    homography(rc, sc, pl, H);
    const T cz = H[6] * px + H[7] * py + H[8];
    const T cx = (H[0] * px + H[1] * py + H[2]) / cz;
    const T cy = (H[3] * px + H[4] * py + H[5]) / cz;
    if (diag) {
        diag->edge = std::min(std::min(std::fabs(double(cx)), std::fabs(double(sc.w) - double(cx))),
                              std::min(std::fabs(double(cy)), std::fabs(double(sc.h) - double(cy))));
        diag->minvar = 1e30;
    }
    if (cx >= T(sc.w) || cx < T(0) || cy >= T(sc.h) || cy < T(0)) return cost_max;
    const int h = rc.h, w = rc.w;
    const T ss = T(2) * T(prm.sigma_spatial) * T(prm.sigma_spatial);
    const T sc2 = T(2) * T(prm.sigma_color) * T(prm.sigma_color);
    // The weighted moments are accumulated about the centre intensity
    // (r - c, s - c). Variance and covariance are shift-invariant, so this is
    // the reference's E[x^2] - E[x]^2 exactly in real arithmetic, without its
    // cancellation when T = float.
    T sr = 0, srr = 0, ssrc = 0, sss = 0, srs = 0, sw = 0;
    const T rcen = T(ref[py * w + px]);
    for (int i = -radius; i < radius + 1; i += prm.radius_increment) {
        for (int j = -radius; j < radius + 1; j += prm.radius_increment) {
            const int tx = std::min(std::max(px + i, 0), w - 1);
            const int ty = std::min(std::max(py + j, 0), h - 1);
            const T r = T(ref[ty * w + tx]) - rcen;
            const T uz = H[6] * tx + H[7] * ty + H[8];
            const T ux = (H[0] * tx + H[1] * ty + H[2]) / uz;
            const T uy = (H[3] * tx + H[4] * ty + H[5]) / uz;
            const T s = bilinear<T>(src, ux, uy, h, w) - rcen;
            const T spatial = std::sqrt(T(i) * T(i) + T(j) * T(j));
            const T wt = std::exp(-spatial / ss - std::fabs(r) / sc2);
            sr += wt * r;
            srr += wt * r * r;
            ssrc += wt * s;
            sss += wt * s * s;
            srs += wt * r * s;
            sw += wt;
        }
    }
    const T inv = T(1) / sw;
    sr *= inv;
    srr *= inv;
    ssrc *= inv;
    sss *= inv;
    srs *= inv;
    const T var_r = srr - sr * sr;
    const T var_s = sss - ssrc * ssrc;
    if (diag) diag->minvar = std::min(double(var_r), double(var_s));
    if (var_r < T(1e-5) || var_s < T(1e-5)) return cost_max;
    const T cov = srs - sr * ssrc;
    return std::fmax(T(0), std::fmin(cost_max, T(1) - cov / std::sqrt(var_r * var_s)));
}

// get_3D_point_on_World / project_on_camera (src 524-560).
template <class T>
void world_point(T x, T y, T depth, const Cam<T>& c, T out[3]) {
    const T p0 = depth * (x - c.K[2]) / c.K[0];
    const T p1 = depth * (y - c.K[5]) / c.K[4];
    const T p2 = depth;
    const T t0 = c.R[0] * p0 + c.R[3] * p1 + c.R[6] * p2;
    const T t1 = c.R[1] * p0 + c.R[4] * p1 + c.R[7] * p2;
    const T t2 = c.R[2] * p0 + c.R[5] * p1 + c.R[8] * p2;
    const T C0 = -(c.R[0] * c.t[0] + c.R[3] * c.t[1] + c.R[6] * c.t[2]);
    const T C1 = -(c.R[1] * c.t[0] + c.R[4] * c.t[1] + c.R[7] * c.t[2]);
    const T C2 = -(c.R[2] * c.t[0] + c.R[5] * c.t[1] + c.R[8] * c.t[2]);
    out[0] = t0 + C0;
    out[1] = t1 + C1;
    out[2] = t2 + C2;
}

template <class T>
void project(const T X[3], const Cam<T>& c, T& u, T& v, T& d) {
    const T t0 = c.R[0] * X[0] + c.R[1] * X[1] + c.R[2] * X[2] + c.t[0];
    const T t1 = c.R[3] * X[0] + c.R[4] * X[1] + c.R[5] * X[2] + c.t[1];
    const T t2 = c.R[6] * X[0] + c.R[7] * X[1] + c.R[8] * X[2] + c.t[2];
    d = c.K[6] * t0 + c.K[7] * t1 + c.K[8] * t2;
    u = (c.K[0] * t0 + c.K[1] * t1 + c.K[2] * t2) / d;
    v = (c.K[3] * t0 + c.K[4] * t1 + c.K[5] * t2) / d;
}

// compute_geom_consistency_cost (src 562-593).
template <class T>
T geom_cost(const float* src_depth, const Cam<T>& rc, const Cam<T>& sc, const V4<T>& pl, int px,
            int py) {
    const T maxc = T(5);
    const T depth = depth_from_plane(rc, pl, px, py);
    T X[3];
    world_point(T(px), T(py), depth, rc, X);
    T su, sv, sd;
    project(X, sc, su, sv, sd);
    const T sdep = bilinear<T>(src_depth, su, sv, rc.h, rc.w);
    if (sdep == T(0)) return maxc;
    T Y[3];
    world_point(su, sv, sdep, sc, Y);
    T bu, bv, bd;
    project(Y, rc, bu, bv, bd);
    const T dc = T(px) - bu, dr = T(py) - bv;
    return std::fmin(maxc, std::sqrt(dc * dc + dr * dr));
}

// FindMinCostIndex / FindMaxCostIndex (lib 15-37): last index on ties.
template <class T>
int argmin_last(const T* c, int n) {
    T m = c[0];
    int k = 0;
    for (int i = 1; i < n; ++i)
        if (c[i] <= m) {
            m = c[i];
            k = i;
        }
    return k;
}

template <class T>
int argmax_last(const T* c, int n) {
    T m = c[0];
    int k = 0;
    for (int i = 1; i < n; ++i)
        if (c[i] >= m) {
            m = c[i];
            k = i;
        }
    return k;
}

// ------------------------------------------------ decision steps of a sweep ---
// The per-pixel update of a half-sweep is split into the reference's steps so
// that each step can be checked on its own with hand-built inputs (the
// oracle_t_* hooks at the end of this file call exactly these functions).

// Candidate search of checkerboard_propagation (lib 862-1104, src 720-913)
// over the cost grid cs (the half-sweep snapshot). flag[k] and pos[k] use the
// reference's order: 0 up_near, 1 up_far, 2 down_near, 3 down_far,
// 4 left_near, 5 left_far, 6 right_near, 7 right_far. Far probes: offset 3,
// then 3 + 2i (i = 1..10) while inside the image; near probes: offset 1, then
// the V pattern (1 + i) further out and i to either side (i = 0..2). Every
// branch keeps the first strict minimum except right_far, which keeps the
// first strict maximum (lib 954, src 794) unless OCOMPAT_FIX_RIGHT_FAR.
template <class T>
void candidate_search(const T* cs, int h, int w, int px, int py, uint32_t compat, bool flag[8],
                      int pos[8]) {
    const int c = py * w + px;
    for (int k = 0; k < 8; ++k) {
        flag[k] = false;
        pos[k] = -1;
    }
    // Far probes along one axis: step = +-w (vertical) or +-1 (horizontal);
    // room(i) says whether the probe 3 + 2i cells out is inside the image.
    auto far = [&](int k, int step, auto room, bool keep_max) {
        flag[k] = true;
        int best = c + 3 * step;
        T m = cs[best];
        for (int i = 1; i < 11; ++i)
            if (room(i)) {
                const int q = c + (3 + 2 * i) * step;
                const bool better = keep_max ? (m < cs[q]) : (cs[q] < m);
                if (better) {
                    m = cs[q];
                    best = q;
                }
            }
        pos[k] = best;
    };
    if (py > 2) far(1, -w, [&](int i) { return py > 2 + 2 * i; }, false);
    if (py < h - 3) far(3, w, [&](int i) { return py < h - 3 - 2 * i; }, false);
    if (px > 2) far(5, -1, [&](int i) { return px > 2 + 2 * i; }, false);
    if (px < w - 3)
        far(7, 1, [&](int i) { return px < w - 3 - 2 * i; }, !(compat & OCOMPAT_FIX_RIGHT_FAR));
    // Near probes: out = the outward step, side = the sideways step; ok_out(i)
    // and ok_lo(i) / ok_hi(i) are the reference's bounds tests.
    auto near = [&](int k, int out, int side, auto ok_lo, auto ok_hi) {
        flag[k] = true;
        const int base = c + out;
        int best = base;
        T m = cs[best];
        for (int i = 0; i < 3; ++i) {
            if (ok_lo(i)) {
                const int q = base + (1 + i) * out - i * side;
                if (cs[q] < m) {
                    m = cs[q];
                    best = q;
                }
            }
            if (ok_hi(i)) {
                const int q = base + (1 + i) * out + i * side;
                if (cs[q] < m) {
                    m = cs[q];
                    best = q;
                }
            }
        }
        pos[k] = best;
    };
    if (py > 0)
        near(0, -w, 1, [&](int i) { return py > 1 + i && px > i; },
             [&](int i) { return py > 1 + i && px < w - 1 - i; });
    if (py < h - 1)
        near(2, w, 1, [&](int i) { return py < h - 2 - i && px > i; },
             [&](int i) { return py < h - 2 - i && px < w - 1 - i; });
    if (px > 0)
        near(4, -1, w, [&](int i) { return px > 1 + i && py > i; },
             [&](int i) { return px > 1 + i && py < h - 1 - i; });
    if (px < w - 1)
        near(6, 1, w, [&](int i) { return px < w - 2 - i && py > i; },
             [&](int i) { return px < w - 2 - i && py < h - 1 - i; });
}

// Joint view selection (lib 1114-1183, src 915-981) and transform_PDF_to_CDF
// (src 79-92). ca is the 8 x nv cost table (candidate-major), nbsel the
// selected-view masks of the up, down, left and right neighbours (used only
// where the near flags 0, 2, 4, 6 are set), draws the 15 uniforms in (0, 1].
// Outputs: cdf[nv], wv[nv] (draw counts), *tsel (views with a count), *wn
// (the sum of the counts).
template <class T>
void select_views(const T* ca, int nv, const bool flag[8], const uint32_t nbsel[4], int iter,
                  const T draws[15], T* cdf, T* wv, uint32_t* tsel, T* wn) {
    T prior_v[32], prob[32];
    for (int v = 0; v < nv; ++v) prior_v[v] = T(0);
    for (int i = 0; i < 4; ++i)
        if (flag[2 * i])
            for (int v = 0; v < nv; ++v) prior_v[v] += ((nbsel[i] >> v) & 1u) ? T(0.9) : T(0.1);
    const T tau = T(0.8) * std::exp(T(iter) * T(iter) / T(-90));
    for (int v = 0; v < nv; ++v) {
        T count = 0, tmpw = 0;
        int cf = 0;
        for (int k = 0; k < 8; ++k) {
            const T x = ca[size_t(k) * nv + v];
            if (x < tau) {
                tmpw += std::exp(x * x / T(-0.18));
                count += T(1);
            }
            if (x > T(1.2)) ++cf;
        }
        prob[v] = T(0);
        if (count > T(2) && cf < 3)
            prob[v] = tmpw / count;
        else if (cf < 3)
            prob[v] = std::exp(tau * tau / T(-0.32));
        prob[v] = prob[v] * prior_v[v];
    }
    T sum = 0;
    for (int v = 0; v < nv; ++v) sum += prob[v];
    const T inv = T(1) / sum;
    T cum = 0;
    for (int v = 0; v < nv; ++v) {
        cum += prob[v] * inv;
        cdf[v] = cum;
        wv[v] = T(0);
    }
    for (int d = 0; d < 15; ++d) {
        const T rp = draws[d] - T(FLT_EPSILON);
        for (int v = 0; v < nv; ++v)
            if (cdf[v] > rp) {
                wv[v] += T(1);
                break;
            }
    }
    *tsel = 0;
    *wn = T(0);
    for (int v = 0; v < nv; ++v)
        if (wv[v] > T(0)) {
            *tsel |= 1u << v;
            *wn += wv[v];
        }
}

// Final costs of the 8 candidates (src 983-1002): the draw-count-weighted
// mean over the drawn views; with the geometric term (gt non-null, 8 x nv)
// flagged candidates add 0.1 * geom per view, unflagged ones add nothing.
template <class T>
void final_costs(const T* ca, int nv, const bool flag[8], const T* wv, T wn, const T* gt,
                 T fc[8]) {
    for (int k = 0; k < 8; ++k) {
        T s = 0;
        for (int v = 0; v < nv; ++v)
            if (wv[v] > T(0)) {
                const T x = ca[size_t(k) * nv + v];
                if (gt && flag[k])
                    s += wv[v] * (x + T(0.1) * gt[size_t(k) * nv + v]);
                else
                    s += wv[v] * x;
            }
        fc[k] = s / wn;
    }
}

// Adoption of the cheapest candidate (lib 1306-1334, src 1023-1039): the
// last index of the minimum final cost; adopted only if that candidate is
// flagged, its depth dk lies in [dmin, dmax] and it is strictly cheaper than
// the current plane. Returns the candidate or -1.
template <class T>
int adopt_min(const T fc[8], const bool flag[8], const T dk[8], T cnow, T dmin, T dmax) {
    const int k = argmin_last(fc, 8);
    if (!flag[k]) return -1;
    return (dk[k] >= dmin && dk[k] <= dmax && fc[k] < cnow) ? k : -1;
}

// Adoption under the planar prior (lib 1278-1305): the last index of the
// maximum restricted score rs (unflagged candidates score 0); adopted only if
// flagged, in range and strictly above the current plane's score rnow.
template <class T>
int adopt_max(const T rs[8], const bool flag[8], const T dk[8], T rnow, T dmin, T dmax) {
    const int k = argmax_last(rs, 8);
    if (!flag[k]) return -1;
    return (dk[k] >= dmin && dk[k] <= dmax && rs[k] > rnow) ? k : -1;
}

// The five refinement hypotheses in the reference's order (lib 740-749):
// depths {d_rand, d_cur, d_rand, d_cur, d_pert} with normals {n_cur, n_rand,
// n_rand, n_pert, n_cur}; n_cur and d_cur are the plane after the
// propagation step (captured once, before any refinement adoption).
template <class T>
void refine_hypotheses(T drand, const V4<T>& nrand, T dcur, const V4<T>& ncur, T dpert,
                       const V4<T>& npert, T depths[5], V4<T> normals[5]) {
    depths[0] = drand;
    depths[1] = dcur;
    depths[2] = drand;
    depths[3] = dcur;
    depths[4] = dpert;
    normals[0] = ncur;
    normals[1] = nrand;
    normals[2] = nrand;
    normals[3] = npert;
    normals[4] = ncur;
}

// Sequential adoption of the refinement hypotheses (lib 751-810): each
// hypothesis i with cost tc[i] and depth db[i] is compared with the state as
// left by the hypotheses before it. Plain: adopt if db in [dmin, dmax] and
// tc < *cost. Prior (pm): adopt if in range and its score sc[i] > *restricted
// (the restricted score is 0 when the propagation step adopted nothing under
// the prior, lib 1240). Updates *cost (and *restricted) and returns the last
// adopted hypothesis or -1.
template <class T>
int refine_adopt(bool pm, const T tc[5], const T db[5], const T sc[5], T dmin, T dmax, T* cost,
                 T* restricted) {
    int last = -1;
    for (int i = 0; i < 5; ++i) {
        const bool in = db[i] >= dmin && db[i] <= dmax;
        if (pm) {
            if (in && sc[i] > *restricted) {
                *cost = tc[i];
                *restricted = sc[i];
                last = i;
            }
        } else if (in && tc[i] < *cost) {
            *cost = tc[i];
            last = i;
        }
    }
    return last;
}

// Restricted (planar-prior) score (lib 1268-1277, 784-793):
// exp(-c^2 / beta) * (gamma + exp(-dd^2 / 2 sd^2) exp(-a^2 / 2 sa^2)) with
// gamma = 0.5, beta = 0.18, sd = (dmax - dmin) / 64, sa = 5 degrees,
// dd = depth - depth_prior, a = acos(n . n_prior).
template <class T>
T prior_score(T cost_v, T depth_v, T depth_prior, const V4<T>& pp, const V4<T>& pl, T dmin,
              T dmax) {
    const T gamma = T(0.5), beta = T(0.18);
    const T dsig = (dmax - dmin) / T(64);
    const T two_ds2 = T(2) * dsig * dsig;
    const T asig = T(kPi) * (T(5) / T(180));
    const T two_as2 = T(2) * asig * asig;
    const T dd = depth_v - depth_prior;
    // The dot product of two unit normals is clamped into [-1, 1] (NaN
    // stays NaN): the reference's acos of the raw dot (lib 1266, 1276,
    // 789) returns NaN for identical normals whenever rounding lands just
    // above 1, which silently drops the candidate in one precision and
    // not in another.
    T dot = pp.x * pl.x + pp.y * pl.y + pp.z * pl.z;
    dot = dot > T(1) ? T(1) : (dot < T(-1) ? T(-1) : dot);
    const T ang = std::acos(dot);
    const T pr = gamma + std::exp(-dd * dd / two_ds2) * std::exp(-ang * ang / two_as2);
    return std::exp(-cost_v * cost_v / beta) * pr;
}

// Decisions of one pixel update and their margins (distance of the deciding
// value from the value it was compared with), recorded only by the lockstep
// comparison of the FP32 and FP64 instantiations.
struct Decision {
    uint32_t tsel;    // drawn views
    int32_t kext;     // argmin (plain) / argmax (prior) candidate
    int32_t kprop;    // adopted candidate or -1
    int32_t kref;     // last adopted refinement hypothesis or -1
    int32_t pm;       // masked pixel under the prior
    int32_t n2;       // cost-table entries equal to 2 (invalid view or flat patch)
    double m_draw;    // min of |cdf[v] - (u - FLT_EPSILON)| and of |cost - tau|, |cost - 1.2|
    double m_ext;     // gap between the best and the runner-up candidate value
    double m_adopt;   // |best candidate value - current plane's value|
    double m_refine;  // min over hypotheses of |value - running threshold|
    double val[8];    // the 8 candidate values compared (final costs or restricted scores)
    double fc[8];     // the 8 final costs
    double now;       // the current plane's value
};

// --------------------------------------------------------------- the run ---
template <class T>
struct Run {
    const OPMArgs& a;
    int h, w, nv;  // nv = number of source views
    std::vector<Cam<T>> cams;
    std::vector<V4<T>> plane, plane0;  // live state and half-sweep snapshot
    std::vector<T> cost, cost0;
    std::vector<uint32_t> sel;
    T dmin, dmax;
    Decision* rec = nullptr;  // per-pixel decision records (lockstep comparison only)

    explicit Run(const OPMArgs& args) : a(args) {
        h = a.cameras[0].height;
        w = a.cameras[0].width;
        nv = a.num_images - 1;
        for (int i = 0; i < a.num_images; ++i) cams.push_back(to_cam<T>(a.cameras[i]));
        plane.resize(size_t(h) * w);
        cost.resize(size_t(h) * w);
        sel.assign(size_t(h) * w, 0u);
        dmin = T(a.params.depth_min);
        dmax = T(a.params.depth_max);
    }

    Rng rng(int px, int py, uint32_t phase) const {
        Rng r;
        r.key[0] = a.seed;
        r.key[1] = a.rng_view;
        r.pixel = uint32_t(py * w + px);
        r.phase = phase;
        return r;
    }

    T ncc(int v, int px, int py, const V4<T>& pl) const {
        return bilateral_ncc<T>(a.images[0], cams[0], a.images[v + 1], cams[v + 1], px, py, pl,
                                a.params, nullptr);
    }

    T geom(int v, int px, int py, const V4<T>& pl) const {
        return geom_cost<T>(a.depths[v + 1], cams[0], cams[v + 1], pl, px, py);
    }

    bool masked(int idx) const {
        return a.params.planar_prior && a.plane_masks && a.plane_masks[idx] > 0u;
    }

    V4<T> prior_plane(int idx) const {
        const float* p = a.prior_planes + size_t(idx) * 4;
        return V4<T>{T(p[0]), T(p[1]), T(p[2]), T(p[3])};
    }

    // compute_multi_view_initial_cost_and_selected_views (src 447-501).
    T initial_cost(int px, int py, const V4<T>& pl, uint32_t* bits) const {
        std::vector<T> c(nv), sorted;
        int valid = 0;
        for (int v = 0; v < nv; ++v) {
            c[v] = ncc(v, px, py, pl);
            if (c[v] < T(2)) ++valid;
        }
        sorted = c;
        // sort_small (src 35-42) is an insertion sort; std::stable_sort
        // orders the same values identically.
        std::stable_sort(sorted.begin(), sorted.end(), [](T x, T y) { return x < y; });
        *bits = 0;
        const int k = std::min(valid, a.params.top_k);
        if (k <= 0) return T(2);
        T s = 0;
        for (int i = 0; i < k; ++i) s += sorted[i];
        const T thr = sorted[k - 1];
        for (int v = 0; v < nv; ++v)
            if (c[v] <= thr) *bits |= (1u << v);
        return s / T(k);
    }

    void init_pixel(int px, int py) {
        const int idx = py * w + px;
        const Cam<T>& rc = cams[0];
        const Rng r = rng(px, py, ORNG_PHASE_INIT);
        V4<T> pl;
        if (a.init_mode == 1) {
            // initialization (src 1046-1082): world normal -> camera frame,
            // then the plane distance from the input depth.
            const float* nw = a.ref_normal + size_t(idx) * 3;
            const T n0 = T(nw[0]), n1 = T(nw[1]), n2 = T(nw[2]);
            pl.x = rc.R[0] * n0 + rc.R[1] * n1 + rc.R[2] * n2;
            pl.y = rc.R[3] * n0 + rc.R[4] * n1 + rc.R[5] * n2;
            pl.z = rc.R[6] * n0 + rc.R[7] * n1 + rc.R[8] * n2;
            pl.w = dist_to_origin(rc, px, py, T(a.depths[0][idx]), pl);
        } else if (a.init_mode == 2) {
            // RandomInitialization prior branch (lib 1513-1551).
            const float* st = a.planes + size_t(idx) * 4;
            if (masked(idx) && a.costs[idx] >= 0.1f) {
                const V4<T> pp = prior_plane(idx);
                const T pert = T(0.02);
                const T dlo = (T(1) - T(3) * pert) * pp.w;
                const T dhi = (T(1) + T(3) * pert) * pp.w;
                const T dp = T(r.u(ORNG_INIT_DEPTH)) * (dhi - dlo) + dlo;
                pl = perturbed_normal(rc, px, py, pp, r, ORNG_INIT_ANGLE, T(3) * pert * T(kPi));
                pl.w = dp;
            } else {
                pl = V4<T>{T(st[0]), T(st[1]), T(st[2]), T(st[3])};
                if (a.compat & OCOMPAT_FIX_PRIOR_NORMALS) {
                    // world -> camera frame, as the depth/normal init does
                    const T n0 = pl.x, n1 = pl.y, n2 = pl.z;
                    pl.x = rc.R[0] * n0 + rc.R[1] * n1 + rc.R[2] * n2;
                    pl.y = rc.R[3] * n0 + rc.R[4] * n1 + rc.R[5] * n2;
                    pl.z = rc.R[6] * n0 + rc.R[7] * n1 + rc.R[8] * n2;
                }
                pl.w = dist_to_origin(rc, px, py, T(st[3]), pl);
            }
        } else {
            // GenerateRandomPlaneHypothesis (lib 215-230).
            const T depth = T(r.u(ORNG_INIT_DEPTH)) * (dmax - dmin) + dmin;
            pl = random_normal(rc, px, py, r, ORNG_INIT_NORMAL, depth);
            pl.w = dist_to_origin(rc, px, py, depth, pl);
        }
        plane[idx] = pl;
        uint32_t bits = 0;
        cost[idx] = initial_cost(px, py, pl, &bits);
        sel[idx] = bits;
    }

    // One pixel of one half-sweep: propagation (src 678-1041, lib 813-1351)
    // then refinement (src 595-676, lib 684-811), as the decision steps above
    // applied to this pixel's NCCs.
    void update_pixel(int px, int py, int iter) {
        const int c = py * w + px;
        const Cam<T>& rc = cams[0];
        const bool geo = a.geom != 0;
        const bool prior = a.params.planar_prior;
        const Rng r = rng(px, py, uint32_t(iter));

        bool flag[8];
        int pos[8];
        candidate_search<T>(cost0.data(), h, w, px, py, a.compat, flag, pos);

        // cost_array[8][32] = {2.0f}: only [0][0] is 2, the rest are 0
        // (all 2 with OCOMPAT_FIX_COST_DEFAULT).
        std::vector<T> ca(size_t(8) * nv, (a.compat & OCOMPAT_FIX_COST_DEFAULT) ? T(2) : T(0));
        ca[0] = T(2);
        for (int k = 0; k < 8; ++k)
            if (flag[k])
                for (int v = 0; v < nv; ++v) ca[size_t(k) * nv + v] = ncc(v, px, py, plane0[pos[k]]);

        const int nb[4] = {c - w, c + w, c - 1, c + 1};
        uint32_t nbsel[4];
        for (int i = 0; i < 4; ++i) nbsel[i] = flag[2 * i] ? sel[nb[i]] : 0u;
        T draws[15];
        for (int d = 0; d < 15; ++d) draws[d] = T(r.u(ORNG_VIEW + d));
        T cdf[32], wv[32], wn;
        uint32_t tsel;
        select_views<T>(ca.data(), nv, flag, nbsel, iter, draws, cdf, wv, &tsel, &wn);
        Decision* dr = rec ? rec + c : nullptr;
        if (dr) {
            dr->tsel = tsel;
            double m = 1e30;
            for (int d = 0; d < 15; ++d)
                for (int v = 0; v < nv; ++v)
                    m = std::min(m, std::fabs(double(cdf[v]) - (double(draws[d]) - double(FLT_EPSILON))));
            // and the cost-table entries' distances to the counting thresholds
            const double tau = double(T(0.8) * std::exp(T(iter) * T(iter) / T(-90)));
            for (size_t i = 0; i < ca.size(); ++i)
                m = std::min(m, std::min(std::fabs(double(ca[i]) - tau), std::fabs(double(ca[i]) - 1.2)));
            dr->m_draw = m;
            int n2 = 0;
            for (size_t i = 0; i < ca.size(); ++i) n2 += ca[i] == T(2);
            dr->n2 = n2;
            dr->kprop = -1;
            dr->kref = -1;
        }

        std::vector<T> gt;
        if (geo) {
            gt.assign(size_t(8) * nv, T(0));
            for (int k = 0; k < 8; ++k)
                if (flag[k])
                    for (int v = 0; v < nv; ++v)
                        if (wv[v] > T(0)) gt[size_t(k) * nv + v] = geom(v, px, py, plane0[pos[k]]);
        }
        T fc[8];
        final_costs<T>(ca.data(), nv, flag, wv, wn, geo ? gt.data() : nullptr, fc);

        // Current hypothesis with the new view weights; views with zero
        // weight add exactly zero and are skipped.
        const V4<T> cur = plane[c];
        T cnow = 0;
        for (int v = 0; v < nv; ++v)
            if (wv[v] > T(0)) {
                const T x = ncc(v, px, py, cur);
                cnow += wv[v] * (geo ? x + T(0.1) * geom(v, px, py, cur) : x);
            }
        cnow /= wn;
        cost[c] = cnow;
        T depth_now = depth_from_plane(rc, cur, px, py);

        T dk[8];
        for (int k = 0; k < 8; ++k)
            dk[k] = flag[k] ? depth_from_plane(rc, plane0[pos[k]], px, py) : T(0);
        const bool pm = prior && masked(c);
        T dprior = T(0);
        V4<T> pp{0, 0, 0, 0};
        if (prior && a.prior_planes) {
            pp = prior_plane(c);
            dprior = depth_from_plane(rc, pp, px, py);
        }
        // best / runner-up gap of the 8 candidate values
        auto gap = [](const T* v, int k) {
            double g = 1e30;
            for (int i = 0; i < 8; ++i)
                if (i != k) g = std::min(g, std::fabs(double(v[i]) - double(v[k])));
            return g;
        };
        if (dr) dr->pm = pm ? 1 : 0;
        T restricted = T(0);
        if (pm) {
            T rs[8];
            for (int k = 0; k < 8; ++k)
                rs[k] = flag[k] ? prior_score(fc[k], dk[k], dprior, pp, plane0[pos[k]], dmin, dmax)
                                : T(0);
            const T rnow = prior_score(cnow, depth_now, dprior, pp, cur, dmin, dmax);
            const int k = adopt_max<T>(rs, flag, dk, rnow, dmin, dmax);
            if (dr) {
                dr->kext = argmax_last(rs, 8);
                dr->kprop = k;
                dr->m_ext = gap(rs, dr->kext);
                dr->m_adopt = std::fabs(double(rs[dr->kext]) - double(rnow));
                for (int i = 0; i < 8; ++i) {
                    dr->val[i] = double(rs[i]);
                    dr->fc[i] = double(fc[i]);
                }
                dr->now = double(rnow);
            }
            if (k >= 0) {
                depth_now = dk[k];
                plane[c] = plane0[pos[k]];
                cost[c] = fc[k];
                restricted = rs[k];
                sel[c] = tsel;
            }
        } else {
            const int k = adopt_min<T>(fc, flag, dk, cnow, dmin, dmax);
            if (dr) {
                dr->kext = argmin_last(fc, 8);
                dr->kprop = k;
                dr->m_ext = gap(fc, dr->kext);
                dr->m_adopt = std::fabs(double(fc[dr->kext]) - double(cnow));
                for (int i = 0; i < 8; ++i) dr->val[i] = dr->fc[i] = double(fc[i]);
                dr->now = double(cnow);
            }
            if (k >= 0) {
                depth_now = dk[k];
                plane[c] = plane0[pos[k]];
                cost[c] = fc[k];
                // the prior path leaves selected_views alone (lib 1306-1318)
                if (!prior || (a.compat & OCOMPAT_FIX_PRIOR_SELECTION)) sel[c] = tsel;
            }
        }

        // ---- refinement
        const T pert = T(0.02);
        T drand;
        V4<T> nrand;
        const T dsig = (dmax - dmin) / T(64);
        if (pm) {
            drand = T(r.u(ORNG_DEPTH_RAND)) * T(6) * dsig + (dprior - T(3) * dsig);
            nrand = perturbed_normal(rc, px, py, pp, r, ORNG_PRIOR_ANGLE,
                                     T(kPi) * (T(5) / T(180)));
        } else {
            drand = T(r.u(ORNG_DEPTH_RAND)) * (dmax - dmin) + dmin;
            nrand = random_normal(rc, px, py, r, ORNG_NORMAL, depth_now);
        }
        // the reference's do { draw } while (d < dmin && d > dmax) runs once
        const T dlo = (T(1) - pert) * depth_now;
        const T dhi = (T(1) + pert) * depth_now;
        const T dpert = T(r.u(ORNG_DEPTH_PERT)) * (dhi - dlo) + dlo;
        const V4<T> ncur = plane[c];
        const V4<T> npert = perturbed_normal(rc, px, py, ncur, r, ORNG_PERT_ANGLE, pert * T(kPi));
        T depths[5];
        V4<T> normals[5];
        refine_hypotheses<T>(drand, nrand, depth_now, ncur, dpert, npert, depths, normals);
        V4<T> tp[5];
        T tc[5], db[5], sc[5];
        for (int i = 0; i < 5; ++i) {
            tp[i] = normals[i];
            tp[i].w = dist_to_origin(rc, px, py, depths[i], tp[i]);
            T s = 0;
            for (int v = 0; v < nv; ++v)
                if (wv[v] > T(0)) {
                    const T x = ncc(v, px, py, tp[i]);
                    s += wv[v] * (geo ? x + T(0.1) * geom(v, px, py, tp[i]) : x);
                }
            tc[i] = s / wn;
            db[i] = depth_from_plane(rc, tp[i], px, py);
            sc[i] = pm ? prior_score(tc[i], depths[i], dprior, pp, tp[i], dmin, dmax) : T(0);
        }
        T cst = cost[c];
        if (dr) {
            double m = 1e30;
            T c2 = cst, r2 = restricted;
            for (int i = 0; i < 5; ++i) {
                m = std::min(m, pm ? std::fabs(double(sc[i]) - double(r2))
                                   : std::fabs(double(tc[i]) - double(c2)));
                const bool in = db[i] >= dmin && db[i] <= dmax;
                if (pm ? (in && sc[i] > r2) : (in && tc[i] < c2)) {
                    c2 = tc[i];
                    r2 = sc[i];
                }
            }
            dr->m_refine = m;
        }
        const int k = refine_adopt<T>(pm, tc, db, sc, dmin, dmax, &cst, &restricted);
        if (dr) dr->kref = k;
        if (k >= 0) {
            plane[c] = tp[k];
            cost[c] = cst;
        }
    }

    void half_sweep(int iter, int color) {
        plane0 = plane;
        cost0 = cost;
#pragma omp parallel for schedule(dynamic, 4)
        for (int py = 0; py < h; ++py)
            for (int px = (py + color) & 1; px < w; px += 2) update_pixel(px, py, iter);
    }

    // GetDepthandNormal (lib 1655-1674).
    void depth_normal() {
        const Cam<T>& rc = cams[0];
#pragma omp parallel for
        for (int py = 0; py < h; ++py)
            for (int px = 0; px < w; ++px) {
                V4<T>& p = plane[py * w + px];
                p.w = depth_from_plane(rc, p, px, py);
                const T x = rc.R[0] * p.x + rc.R[3] * p.y + rc.R[6] * p.z;
                const T y = rc.R[1] * p.x + rc.R[4] * p.y + rc.R[7] * p.z;
                const T z = rc.R[2] * p.x + rc.R[5] * p.y + rc.R[8] * p.z;
                p.x = x;
                p.y = y;
                p.z = z;
            }
    }

    // CheckerboardFilter (src 1165-1281, lib 1355-1463).
    void filter_pass(int color) {
#pragma omp parallel for
        for (int py = 0; py < h; ++py)
            for (int px = (py + color) & 1; px < w; px += 2) {
                const int c = py * w + px;
                if (cost[c] < T(0.001)) continue;
                T f[21];
                int n = 0;
                auto d = [&](int q) { return plane[q].w; };
                f[n++] = d(c);
                const int up = c - w, dn = c + w, lf = c - 1, rt = c + 1;
                if (py > 0) f[n++] = d(up);
                if (py > 2) f[n++] = d(c - 3 * w);
                if (py > 4) f[n++] = d(c - 5 * w);
                if (py < h - 1) f[n++] = d(dn);
                if (py < h - 3) f[n++] = d(c + 3 * w);
                if (py < h - 5) f[n++] = d(c + 5 * w);
                if (px > 0) f[n++] = d(lf);
                if (px > 2) f[n++] = d(c - 3);
                if (px > 4) f[n++] = d(c - 5);
                if (px < w - 1) f[n++] = d(rt);
                if (px < w - 3) f[n++] = d(c + 3);
                if (px < w - 5) f[n++] = d(c + 5);
                if (py > 0 && px < w - 2) f[n++] = d(up + 2);
                if (py < h - 1 && px < w - 2) f[n++] = d(dn + 2);
                if (py > 0 && px > 1) f[n++] = d(up - 2);
                if (py < h - 1 && px > 1) f[n++] = d(dn - 2);
                if (px > 0 && py > 2) f[n++] = d(lf - 2 * w);
                if (px < w - 1 && py > 2) f[n++] = d(rt - 2 * w);
                if (px > 0 && py < h - 2) f[n++] = d(lf + 2 * w);
                if (px < w - 1 && py < h - 2) f[n++] = d(rt + 2 * w);
                // insertion sort, as sort_small (src 35-42)
                for (int i = 1; i < n; ++i) {
                    const T t = f[i];
                    int j = i;
                    for (; j >= 1 && t < f[j - 1]; --j) f[j] = f[j - 1];
                    f[j] = t;
                }
                const int m = n / 2;
                plane[c].w = (n % 2 == 0) ? (f[m - 1] + f[m]) / T(2) : f[m];
            }
    }

    void run() {
        const int iters = a.params.max_iterations;
#pragma omp parallel for
        for (int py = 0; py < h; ++py)
            for (int px = 0; px < w; ++px) init_pixel(px, py);
        for (int it = 0; it < iters; ++it) {
            half_sweep(it, 0);
            half_sweep(it, 1);
        }
        depth_normal();
        if (!a.skip_filter) {
            filter_pass(0);
            filter_pass(1);
        }
        const size_t n = size_t(h) * w;
        for (size_t i = 0; i < n; ++i) {
            a.planes[4 * i + 0] = float(plane[i].x);
            a.planes[4 * i + 1] = float(plane[i].y);
            a.planes[4 * i + 2] = float(plane[i].z);
            a.planes[4 * i + 3] = float(plane[i].w);
            a.costs[i] = float(cost[i]);
            if (a.selected_views) a.selected_views[i] = sel[i];
            if (a.out_depth) a.out_depth[i] = float(plane[i].w);
            if (a.out_normal) {
                a.out_normal[3 * i + 0] = float(plane[i].x);
                a.out_normal[3 * i + 1] = float(plane[i].y);
                a.out_normal[3 * i + 2] = float(plane[i].z);
            }
            if (a.out_cost) a.out_cost[i] = float(cost[i]);
        }
    }
};

int check_args(const OPMArgs* a) {
    if (!a || !a->cameras || !a->images || !a->planes || !a->costs) return 1;
    if (a->num_images < 2 || a->num_images > 32) return 2;
    if (a->init_mode == 1 && (!a->depths || !a->depths[0] || !a->ref_normal)) return 3;
    if (a->geom && !a->depths) return 4;
    if (a->params.planar_prior && (!a->prior_planes || !a->plane_masks)) return 5;
    return 0;
}

// Geometric-consistency cost for n (pixel, plane) queries, FP64 (precision
// 64, the oracle) or FP32 (precision 32, the CPU port's arithmetic).
template <class T>
void geom_batch_t(const OCamera* cams, int num_images, const float* const* depths, int n,
                  const int32_t* xy, const float* planes, double* out) {
    const int nv = num_images - 1;
    std::vector<Cam<T>> c;
    for (int i = 0; i < num_images; ++i) c.push_back(to_cam<T>(cams[i]));
    for (int q = 0; q < n; ++q) {
        const V4<T> pl{T(planes[4 * q]), T(planes[4 * q + 1]), T(planes[4 * q + 2]),
                       T(planes[4 * q + 3])};
        for (int v = 0; v < nv; ++v)
            out[size_t(q) * nv + v] =
                double(geom_cost<T>(depths[v + 1], c[0], c[v + 1], pl, xy[2 * q], xy[2 * q + 1]));
    }
}

}  // namespace

extern "C" {

void oracle_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    oracle_philox4x32_10_block(ctr, key, out);
}

// FP64 bilateral NCC for n (pixel, plane) queries against every source view.
// xy: [n*2] int, planes: [n*4] float (camera frame, (n, d)).
// costs/edge/minvar: [n*(num_images-1)] double.
int oracle_cost_batch(const OCamera* cams, int num_images, const float* const* images,
                      const OParams* params, int n, const int32_t* xy, const float* planes,
                      double* costs, double* edge, double* minvar) {
    if (!cams || !images || !params || num_images < 2) return 1;
    const int nv = num_images - 1;
    std::vector<Cam<double>> c;
    for (int i = 0; i < num_images; ++i) c.push_back(to_cam<double>(cams[i]));
#pragma omp parallel for schedule(dynamic, 64)
    for (int q = 0; q < n; ++q) {
        const V4<double> pl{planes[4 * q], planes[4 * q + 1], planes[4 * q + 2], planes[4 * q + 3]};
        for (int v = 0; v < nv; ++v) {
            NccDiag d;
            costs[size_t(q) * nv + v] = bilateral_ncc<double>(
                images[0], c[0], images[v + 1], c[v + 1], xy[2 * q], xy[2 * q + 1], pl, *params, &d);
            if (edge) edge[size_t(q) * nv + v] = d.edge;
            if (minvar) minvar[size_t(q) * nv + v] = d.minvar;
        }
    }
    return 0;
}

int oracle_geom_batch(const OCamera* cams, int num_images, const float* const* depths, int n,
                      const int32_t* xy, const float* planes, double* out) {
    if (!cams || !depths || num_images < 2) return 1;
    geom_batch_t<double>(cams, num_images, depths, n, xy, planes, out);
    return 0;
}

int oracle_geom_batch32(const OCamera* cams, int num_images, const float* const* depths, int n,
                        const int32_t* xy, const float* planes, double* out) {
    if (!cams || !depths || num_images < 2) return 1;
    geom_batch_t<float>(cams, num_images, depths, n, xy, planes, out);
    return 0;
}

// ---- FP64 hooks of the decision steps, for known-answer tests on hand-built
// inputs. Each calls the same template the pass uses.

// Candidate search over an h x w cost grid at (px, py): flag[8], pos[8].
int oracle_t_candidates(const double* costs, int h, int w, int px, int py, uint32_t compat,
                        int32_t* flag, int32_t* pos) {
    if (!costs || px < 0 || py < 0 || px >= w || py >= h) return 1;
    bool f[8];
    int p[8];
    candidate_search<double>(costs, h, w, px, py, compat, f, p);
    for (int k = 0; k < 8; ++k) {
        flag[k] = f[k] ? 1 : 0;
        pos[k] = p[k];
    }
    return 0;
}

// View selection from an 8 x nv cost table, the four neighbour masks and 15
// given draws.
int oracle_t_select_views(const double* ca, int nv, const int32_t* flag, const uint32_t* nbsel,
                          int iter, const double* draws, double* cdf, double* wv, uint32_t* tsel,
                          double* wn) {
    if (!ca || nv < 1 || nv > 32) return 1;
    bool f[8];
    for (int k = 0; k < 8; ++k) f[k] = flag[k] != 0;
    select_views<double>(ca, nv, f, nbsel, iter, draws, cdf, wv, tsel, wn);
    return 0;
}

// Final costs of the 8 candidates; gt (8 x nv geometric terms) may be NULL.
int oracle_t_final_costs(const double* ca, int nv, const int32_t* flag, const double* wv,
                         double wn, const double* gt, double* fc) {
    if (!ca || nv < 1 || nv > 32) return 1;
    bool f[8];
    for (int k = 0; k < 8; ++k) f[k] = flag[k] != 0;
    final_costs<double>(ca, nv, f, wv, wn, gt, fc);
    return 0;
}

// Propagation adoption: prior = 0 takes final costs and the current cost
// (adopt_min), prior = 1 restricted scores and the current score
// (adopt_max). Returns the adopted candidate or -1.
int oracle_t_adopt(int prior, const double* score, const int32_t* flag, const double* dk,
                   double now, double dmin, double dmax) {
    bool f[8];
    for (int k = 0; k < 8; ++k) f[k] = flag[k] != 0;
    return prior ? adopt_max<double>(score, f, dk, now, dmin, dmax)
                 : adopt_min<double>(score, f, dk, now, dmin, dmax);
}

// The five refinement hypotheses: depths[5], normals[5 * 3].
void oracle_t_refine_hypotheses(double drand, const double* nrand, double dcur,
                                const double* ncur, double dpert, const double* npert,
                                double* depths, double* normals) {
    const V4<double> nr{nrand[0], nrand[1], nrand[2], 0};
    const V4<double> nc{ncur[0], ncur[1], ncur[2], 0};
    const V4<double> np{npert[0], npert[1], npert[2], 0};
    V4<double> n[5];
    refine_hypotheses<double>(drand, nr, dcur, nc, dpert, np, depths, n);
    for (int i = 0; i < 5; ++i) {
        normals[3 * i] = n[i].x;
        normals[3 * i + 1] = n[i].y;
        normals[3 * i + 2] = n[i].z;
    }
}

// Sequential refinement adoption; *cost and *restricted are in/out.
int oracle_t_refine_adopt(int pm, const double* tc, const double* db, const double* sc,
                          double dmin, double dmax, double* cost, double* restricted) {
    return refine_adopt<double>(pm != 0, tc, db, sc, dmin, dmax, cost, restricted);
}

// Restricted score of one (cost, depth, normal) against a prior plane.
double oracle_t_prior_score(double cost, double depth, double depth_prior, const double* pp,
                            const double* pl, double dmin, double dmax) {
    const V4<double> a{pp[0], pp[1], pp[2], 0}, b{pl[0], pl[1], pl[2], 0};
    return prior_score<double>(cost, depth, depth_prior, a, b, dmin, dmax);
}

// Lockstep comparison of the FP32 and FP64 instantiations (diagnostics for
// the tests): the FP64 pass runs as usual; before each half-sweep its state
// is copied into an FP32 run, both update every pixel of the colour from
// that same state, and each pixel whose decisions differ is recorded:
// rows of (iteration, colour, pixel, what, masked) in ids[5 * n] with what =
// 1 (drawn views differ) | 2 (argmin / argmax differs) | 4 (propagation
// adoption differs) | 8 (refinement adoption differs) | 16 (an NCC of the
// cost table is 2 in one precision only: a variance at the 1e-5 floor or a
// warped centre on the image border), and the FP64 run's
// margins (m_draw, m_ext, m_adopt, m_refine) in margins[4 * n]. Only the
// first divergence of each kind in the pixel's update counts: a pixel whose
// drawn views differ has different costs after that. Returns the number of
// rows (at most cap) or a negative status; *decisions counts pixel updates.
int oracle_pm_lockstep(const OPMArgs* args, int32_t cap, int32_t* ids, double* margins,
                       int64_t* decisions, double* values) {
    const int rc = check_args(args);
    if (rc) return -rc;
    Run<double> rd(*args);
    Run<float> rf(*args);
    const int h = rd.h, w = rd.w;
    const size_t n = size_t(h) * w;
#pragma omp parallel for
    for (int py = 0; py < h; ++py)
        for (int px = 0; px < w; ++px) rd.init_pixel(px, py);
    std::vector<Decision> recd(n), recf(n);
    rd.rec = recd.data();
    rf.rec = recf.data();
    int32_t rows = 0;
    int64_t count = 0;
    for (int it = 0; it < args->params.max_iterations; ++it)
        for (int color = 0; color < 2; ++color) {
            rd.plane0 = rd.plane;
            rd.cost0 = rd.cost;
            for (size_t i = 0; i < n; ++i) {
                const V4<double>& p = rd.plane[i];
                rf.plane[i] = V4<float>{float(p.x), float(p.y), float(p.z), float(p.w)};
                rf.cost[i] = float(rd.cost[i]);
                rf.sel[i] = rd.sel[i];
            }
            rf.plane0 = rf.plane;
            rf.cost0 = rf.cost;
#pragma omp parallel for schedule(dynamic, 4)
            for (int py = 0; py < h; ++py)
                for (int px = (py + color) & 1; px < w; px += 2) {
                    rd.update_pixel(px, py, it);
                    rf.update_pixel(px, py, it);
                }
            for (int py = 0; py < h; ++py)
                for (int px = (py + color) & 1; px < w; px += 2) {
                    const size_t c = size_t(py) * w + px;
                    ++count;
                    const Decision& a = recd[c];
                    const Decision& b = recf[c];
                    int what = 0;
                    if (a.n2 != b.n2)
                        what = 16;
                    else if (a.tsel != b.tsel)
                        what = 1;
                    else if (a.kext != b.kext)
                        what = 2;
                    else if (a.kprop != b.kprop)
                        what = 4;
                    else if (a.kref != b.kref)
                        what = 8;
                    if (!what || rows >= cap) continue;
                    int32_t* o = ids + 5 * size_t(rows);
                    o[0] = it;
                    o[1] = color;
                    o[2] = int32_t(c);
                    o[3] = what;
                    o[4] = a.pm;
                    double* m = margins + 4 * size_t(rows);
                    m[0] = a.m_draw;
                    m[1] = a.m_ext;
                    m[2] = a.m_adopt;
                    m[3] = a.m_refine;
                    if (values) {
                        // FP64 values, FP32 values, FP64 final costs, FP32
                        // final costs, FP64 now, FP32 now, FP32 kext
                        double* v = values + 35 * size_t(rows);
                        for (int i = 0; i < 8; ++i) {
                            v[i] = a.val[i];
                            v[8 + i] = b.val[i];
                            v[16 + i] = a.fc[i];
                            v[24 + i] = b.fc[i];
                        }
                        v[32] = a.now;
                        v[33] = b.now;
                        v[34] = b.kext;
                    }
                    ++rows;
                }
        }
    if (decisions) *decisions = count;
    return rows;
}

// One full PatchMatch pass; precision 64 = oracle, 32 = CPU baseline.
int oracle_pm_run(int precision, const OPMArgs* args) {
    const int rc = check_args(args);
    if (rc) return rc;
#ifdef _OPENMP
    const int saved = omp_get_max_threads();
    if (args->threads > 0) omp_set_num_threads(args->threads);
#endif
    if (precision == 64) {
        Run<double> r(*args);
        r.run();
    } else {
        Run<float> r(*args);
        r.run();
    }
#ifdef _OPENMP
    omp_set_num_threads(saved);
#endif
    return 0;
}

// OpenMP threads of the calls that take no thread count (the fusion twin).
void oracle_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int oracle_max_threads() {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

}  // extern "C"
