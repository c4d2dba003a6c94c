// Device helpers of the PatchMatch kernels (gfx950): counter-based RNG,
// camera-frame geometry, view descriptors and the hoisted bilateral NCC.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gmvs_amd.h"

namespace gmvs {

constexpr float kPiF = 3.14159265358979323846f;
constexpr int kMaxViews = 31;  // sources per reference (32-bit view masks)

// ------------------------------------------------------------------ RNG ---
// Philox4x32-10 (Salmon et al., SC'11): 10 rounds of two 32x32->64 products
// with the published multipliers, key bumped by the Weyl constants between
// rounds. Counter = (pixel, phase, slot / 4, 0), key = (seed, view); a draw
// is word slot % 4 of that block, so every random use has a fixed slot.
__host__ __device__ inline void philox_block(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                             uint32_t k0, uint32_t k1, uint32_t out[4]) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

// Uniform in (0, 1], exact in FP32: ((x >> 8) + 1) * 2^-24.
__host__ __device__ inline float u01(uint32_t x) {
    return (float)((x >> 8) + 1u) * (1.0f / 16777216.0f);
}

// Fixed draw slots (identical in the oracle).
enum : uint32_t {
    RNG_PHASE_INIT = 0x40000000u,
    RNG_INIT_DEPTH = 0,
    RNG_INIT_ANGLE = 1,
    RNG_INIT_NORMAL = 4,
    RNG_VIEW = 0,
    RNG_DEPTH_RAND = 15,
    RNG_DEPTH_PERT = 16,
    RNG_PERT_ANGLE = 17,
    RNG_PRIOR_ANGLE = 20,
    RNG_NORMAL = 24,
    RNG_REJECT_TRIES = 16,
};

struct Rng {
    uint32_t pixel, phase, k0, k1;
    __device__ void block(uint32_t b, uint32_t out[4]) const {
        philox_block(pixel, phase, b, 0u, k0, k1, out);
    }
    __device__ float u(uint32_t slot) const {
        uint32_t o[4];
        block(slot >> 2, o);
        return u01(o[slot & 3u]);
    }
};

// ------------------------------------------------------------- geometry ---
struct RefCam {
    float K[9], R[9], t[3], C[3];
    int h, w;
};

// One source view of one reference view. A = Ks R_rel Kr^-1 and b = Ks t_rel
// are composed in FP64 on the device (pm_setup_views) so that the
// per-hypothesis homography is the FP32 rank-1 update H = A - b m^T with
// m^T = n^T Kr^-1 / d (the factorisation of compute_homography, src 211-296).
struct ViewDesc {
    float A[9];
    float b[3];
    float K[9], R[9], t[3], C[3];  // source camera for the geometric term
    const void* cells;             // packed bilinear cells [H*W] (FP32 or 16-bit)
    const float* depth;            // source depth map [H*W] or nullptr
};

// The same factors in FP64, from which the per-tile local frames are formed.
struct ViewF64 {
    double A[9];
    double b[3];
};

// A view's homography in the local frame of one tile of reference pixels:
// reference positions are taken relative to the tile origin o, source
// positions relative to an integer anchor a, so every FP32 coordinate of a
// tap is a small number (the absolute ones reach thousands of pixels, where
// one FP32 ulp is ~1e-4 px and the NCC's sensitivity to a tap's position
// turns that into cost errors of that size). With m^T = n^T Kr^-1 / d:
//   H'' = T(-a) (A - b m^T) T(o) = A'' - b'' m''^T,
//   A'' = T(-a) A T(o), b'' = T(-a) b, m'' = (m0, m1, m0 ox + m1 oy + m2),
// A'' and b'' composed in FP64 (the column A (o, 1) - a (A (o, 1))_z is
// where the cancellation is). The anchor is the image of the point (cx, cy)
// under the homography of an anchor plane (given by its m in FP64: the
// fronto-parallel plane at the middle inverse depth for a tile, the queried
// plane for a single query), floored, its row moved kAnchorRows up and
// clamped into the image: a tap's cell index relative to the anchor cell is
// then non-negative whenever the tap is at least one row below the anchor
// row, which the clamp-free path checks.
constexpr int kAnchorRows = 64;

struct Frame {
    float A[9];
    float b[3];
    float ax, ay;  // anchor column and row
    int base;      // ay * wp + ax, the anchor's cell index
    int unused;
};

__device__ inline Frame make_frame(const ViewF64& v, int ox, int oy, double cx, double cy,
                                   const double am[3], int w, int h, int wp) {
    const double* A = v.A;
    const double* b = v.b;
    const double s = fma(am[0], cx, fma(am[1], cy, am[2]));
    const double u = fma(A[0], cx, fma(A[1], cy, A[2])) - b[0] * s;
    const double q = fma(A[3], cx, fma(A[4], cy, A[5])) - b[1] * s;
    const double z = fma(A[6], cx, fma(A[7], cy, A[8])) - b[2] * s;
    double X = floor(u / z), Y = floor(q / z) - (double)kAnchorRows;
    if (!(X >= 0.0)) X = 0.0;  // also NaN
    if (!(Y >= 0.0)) Y = 0.0;
    X = fmin(X, (double)(w - 1));
    Y = fmin(Y, (double)(h - 1));
    const double o0 = ox, o1 = oy;
    double Ao[3];
    for (int i = 0; i < 3; ++i) Ao[i] = fma(A[3 * i], o0, fma(A[3 * i + 1], o1, A[3 * i + 2]));
    Frame f;
    for (int j = 0; j < 2; ++j) {
        f.A[j] = (float)(A[j] - X * A[6 + j]);
        f.A[3 + j] = (float)(A[3 + j] - Y * A[6 + j]);
        f.A[6 + j] = (float)A[6 + j];
    }
    f.A[2] = (float)(Ao[0] - X * Ao[2]);
    f.A[5] = (float)(Ao[1] - Y * Ao[2]);
    f.A[8] = (float)Ao[2];
    f.b[0] = (float)(b[0] - X * b[2]);
    f.b[1] = (float)(b[1] - Y * b[2]);
    f.b[2] = (float)b[2];
    f.ax = (float)X;
    f.ay = (float)Y;
    f.base = (int)Y * wp + (int)X;
    f.unused = 0;
    return f;
}

__device__ inline float depth_from_plane(const RefCam& c, float4 pl, float px, float py) {
    return -pl.w * c.K[0] /
           ((px - c.K[2]) * pl.x + (c.K[0] / c.K[4]) * (py - c.K[5]) * pl.y + c.K[0] * pl.z);
}

__device__ inline float dist_to_origin(const RefCam& c, float px, float py, float depth, float4 n) {
    const float X0 = depth * (px - c.K[2]) / c.K[0];
    const float X1 = depth * (py - c.K[5]) / c.K[4];
    return -(n.x * X0 + n.y * X1 + n.z * depth);
}

__device__ inline void view_dir(const RefCam& c, float px, float py, float depth, float v[3]) {
    const float X0 = depth * (px - c.K[2]) / c.K[0];
    const float X1 = depth * (py - c.K[5]) / c.K[4];
    const float X2 = depth;
    const float nrm = sqrtf(X0 * X0 + X1 * X1 + X2 * X2);
    v[0] = X0 / nrm;
    v[1] = X1 / nrm;
    v[2] = X2 / nrm;
}

__device__ inline void normalize3(float4& v) {
    const float s = 1.0f / sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
    v.x *= s;
    v.y *= s;
    v.z *= s;
}

// Random unit normal facing the camera (Marsaglia, A16) from fixed slots.
__device__ inline float4 random_normal(const RefCam& c, float px, float py, const Rng& rng,
                                       uint32_t slot0, float depth) {
    float q1 = 0.f, q2 = 0.f, s = 2.f;
    bool ok = false;
    for (uint32_t a = 0; a < RNG_REJECT_TRIES && !ok; a += 2) {
        // two attempts per Philox block (slots slot0 + 2a .. slot0 + 2a + 3)
        uint32_t o[4];
        rng.block((slot0 + 2 * a) >> 2, o);
        const float a1 = 2.0f * u01(o[0]) - 1.0f, b1 = 2.0f * u01(o[1]) - 1.0f;
        const float s1 = a1 * a1 + b1 * b1;
        if (s1 < 1.0f) {
            q1 = a1;
            q2 = b1;
            s = s1;
            ok = true;
        } else {
            const float a2 = 2.0f * u01(o[2]) - 1.0f, b2 = 2.0f * u01(o[3]) - 1.0f;
            const float s2 = a2 * a2 + b2 * b2;
            if (s2 < 1.0f) {
                q1 = a2;
                q2 = b2;
                s = s2;
                ok = true;
            }
        }
    }
    if (!ok) {
        uint32_t o[4];
        rng.block(slot0 >> 2, o);
        q1 = 2.0f * u01(o[0]) - 1.0f;
        q2 = 2.0f * u01(o[1]) - 1.0f;
        const float r = sqrtf(q1 * q1 + q2 * q2);
        if (r > 0.f) {
            q1 = q1 / r * 0.5f;
            q2 = q2 / r * 0.5f;
        } else {
            q1 = 0.5f;
            q2 = 0.f;
        }
        s = q1 * q1 + q2 * q2;
    }
    const float sq = sqrtf(1.0f - s);
    float4 n = make_float4(2.0f * q1 * sq, 2.0f * q2 * sq, 1.0f - 2.0f * s, 0.f);
    float vd[3];
    view_dir(c, px, py, depth, vd);
    if (n.x * vd[0] + n.y * vd[1] + n.z * vd[2] > 0.f) {
        n.x = -n.x;
        n.y = -n.y;
        n.z = -n.z;
    }
    normalize3(n);
    return n;
}

// e^x through the hardware exp2 (v_exp_f32, ~1 ulp in the exp2 domain).
// Used for the bilateral weights, the view-selection likelihoods and the
// prior scores, where FP32 rounding only moves near-ties.
__device__ inline float fexp(float x) { return __builtin_amdgcn_exp2f(x * 1.44269504088896341f); }

// sin / cos of a small angle (|a| < pi) through v_sin_f32 / v_cos_f32, whose
// argument is in revolutions.
__device__ inline void fsincos(float a, float* s, float* c) {
    const float r = a * 0.159154943091895336f;
    *s = __builtin_amdgcn_sinf(r);
    *c = __builtin_amdgcn_cosf(r);
}

// Euler-angle perturbation of a normal (A17) from three fixed slots.
__device__ inline float4 perturbed_normal(const RefCam& c, float px, float py, float4 n,
                                          float u1, float u2, float u3, float pert) {
    float vd[3];
    view_dir(c, px, py, 1.0f, vd);
    const float a1 = (u1 - 0.5f) * pert, a2 = (u2 - 0.5f) * pert, a3 = (u3 - 0.5f) * pert;
    float s1, c1, s2, c2, s3, c3;
    fsincos(a1, &s1, &c1);
    fsincos(a2, &s2, &c2);
    fsincos(a3, &s3, &c3);
    const float R0 = c2 * c3, R1 = c3 * s1 * s2 - c1 * s3, R2 = s1 * s3 + c1 * c3 * s2;
    const float R3 = c2 * s3, R4 = c1 * c3 + s1 * s2 * s3, R5 = c1 * s2 * s3 - c3 * s1;
    const float R6 = -s2, R7 = c2 * s1, R8 = c1 * c2;
    float4 o = make_float4(R0 * n.x + R1 * n.y + R2 * n.z, R3 * n.x + R4 * n.y + R5 * n.z,
                           R6 * n.x + R7 * n.y + R8 * n.z, n.w);
    if (o.x * vd[0] + o.y * vd[1] + o.z * vd[2] >= 0.f) o = n;
    normalize3(o);
    return o;
}

// m of the fronto-parallel plane at inverse depth inv_d, m^T = n^T Kr^-1 / d
// with n = (0, 0, -1): (0, 0, -inv_d).
__device__ inline void fronto_m(double inv_d, double m[3]) {
    m[0] = 0.0;
    m[1] = 0.0;
    m[2] = -inv_d;
}

// m'' of the local frame with origin (ox, oy): m^T = n^T Kr^-1 / d with its
// constant term moved to the origin,
//   m'' = (m0, m1, m0 (ox - cx) + m1 (oy - cy) + n_z / d) for the plane
//   (n, d) and Kr = [fx 0 cx; 0 fy cy; 0 0 1],
// evaluated in FP64 and rounded once per component. b'' m''^T enters every
// tap position and b'' m''_2 reaches hundreds of pixels; in FP32 the
// divisions and the cancelling constant term leave m'' a few ulps off,
// which moves taps by ~1e-4 px and costs by up to ~1e-4 at 1600x1200.
__device__ inline void plane_m_local(const RefCam& c, float4 pl, float ox, float oy, float m[3]) {
    const double d = pl.w;
    const double a = (double)pl.x / ((double)c.K[0] * d);
    const double b = (double)pl.y / ((double)c.K[4] * d);
    m[0] = (float)a;
    m[1] = (float)b;
    m[2] = (float)fma(a, (double)ox - (double)c.K[2],
                      fma(b, (double)oy - (double)c.K[5], (double)pl.z / d));
}

__device__ inline void homography(const float* __restrict__ A, const float* __restrict__ b,
                                  const float m[3], float H[9]) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) H[3 * i + j] = fmaf(-b[i], m[j], A[3 * i + j]);
}

// Software bilinear of an f32 map with the reference's index rule
// (src 10-33): index = clamp(trunc(c), 0, n-1), fraction = c - index.
__device__ inline float bilinear_map(const float* __restrict__ img, float x, float y, int h,
                                     int w) {
    const float xf = fminf(fmaxf(floorf(x), 0.f), (float)(w - 1));
    const float yf = fminf(fmaxf(floorf(y), 0.f), (float)(h - 1));
    const int wl = (int)xf, hl = (int)yf;
    const float fx = x - xf, fy = y - yf;
    const int wu = (wl == w - 1) ? wl : wl + 1;
    const int hu = (hl == h - 1) ? hl : hl + 1;
    const float v0 = img[hl * w + wl], v1 = img[hl * w + wu];
    const float v2 = img[hu * w + wl], v3 = img[hu * w + wu];
    const float a = fmaf(fx, v1 - v0, v0);
    const float bb = fmaf(fx, v3 - v2, v2);
    return fmaf(fy, bb - a, a);
}

__device__ inline void world_point(float x, float y, float depth, const float* K, const float* R,
                                   const float* C, float X[3]) {
    const float p0 = depth * (x - K[2]) / K[0];
    const float p1 = depth * (y - K[5]) / K[4];
    const float p2 = depth;
    X[0] = R[0] * p0 + R[3] * p1 + R[6] * p2 + C[0];
    X[1] = R[1] * p0 + R[4] * p1 + R[7] * p2 + C[1];
    X[2] = R[2] * p0 + R[5] * p1 + R[8] * p2 + C[2];
}

__device__ inline void project(const float X[3], const float* K, const float* R, const float* t,
                               float& u, float& v) {
    const float t0 = R[0] * X[0] + R[1] * X[1] + R[2] * X[2] + t[0];
    const float t1 = R[3] * X[0] + R[4] * X[1] + R[5] * X[2] + t[1];
    const float t2 = R[6] * X[0] + R[7] * X[1] + R[8] * X[2] + t[2];
    const float d = K[6] * t0 + K[7] * t1 + K[8] * t2;
    u = (K[0] * t0 + K[1] * t1 + K[2] * t2) / d;
    v = (K[3] * t0 + K[4] * t1 + K[5] * t2) / d;
}

// Forward-backward reprojection error against a source depth map, capped at
// 5 px (compute_geom_consistency_cost, src 562-593).
__device__ inline float geom_cost(const RefCam& rc, const ViewDesc& vd, float4 pl, float px,
                                  float py) {
    const float depth = depth_from_plane(rc, pl, px, py);
    float X[3];
    world_point(px, py, depth, rc.K, rc.R, rc.C, X);
    float su, sv;
    project(X, vd.K, vd.R, vd.t, su, sv);
    const float sd = bilinear_map(vd.depth, su, sv, rc.h, rc.w);
    if (sd == 0.0f) return 5.0f;
    float Y[3];
    world_point(su, sv, sd, vd.K, vd.R, vd.C, Y);
    float bu, bv;
    project(Y, rc.K, rc.R, rc.t, bu, bv);
    const float dc = px - bu, dr = py - bv;
    return fminf(5.0f, sqrtf(dc * dc + dr * dr));
}

// ------------------------------------------------------------------ NCC ---
// The bilateral NCC (compute_bilateral_NCC, src 362-445) is evaluated by a
// group of L consecutive lanes per reference pixel (L = 1, 2 or 4). The
// NT x NT taps (i along x, j along y) are split into blocks of NR rows x NC
// columns, one block per lane: L = 1 takes all taps, L = 2 takes every other
// row, L = 4 takes every other row of one half of the columns. Interleaving
// the x offsets keeps the group's four gathers of one load instruction two
// pixels apart, on the same cache lines (a quadrant split put them six
// pixels apart and cost 12 % of the half-sweep time). Each lane hoists the reference side of its
// taps once per pixel -- the bilateral weight w and w * (r - c), c being the
// centre intensity -- and the group reduces its partial sums with DPP quad
// permutes (a fixed, commutative order, so every lane holds the same bits).
// Moments are accumulated about c (shift-invariant, no E[x^2] - E[x]^2
// cancellation).
template <int L>
__device__ inline float group_sum(float v) {
    static_assert(L == 1 || L == 2 || L == 4, "lane group of 1, 2 or 4");
    // update_dpp with old = 0 and bound_ctrl folds into one v_add_f32_dpp
    // (every quad_perm source lane exists, so neither ever applies)
    if constexpr (L >= 2)  // quad_perm [1,0,3,2]
        v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0xB1, 0xF, 0xF, true));
    if constexpr (L >= 4)  // quad_perm [2,3,0,1]
        v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x4E, 0xF, 0xF, true));
    return v;
}

template <int NT, int L>
struct TapBlock {
    static constexpr int NR = (L == 1) ? NT : NT / 2;        // rows (x offsets) per lane
    static constexpr int NC = (L == 4) ? NT / 2 : NT;        // columns (y offsets) per lane
    static_assert(NT % 2 == 0 || L == 1, "split groups need an even tap count");
    // row index of the lane's r-th row, column index of its c-th column
    __device__ static int row(int sub, int r) {
        if constexpr (L == 1) return r;
        else if constexpr (L == 2) return sub + 2 * r;
        else return (sub >> 1) + 2 * r;
    }
    __device__ static int col(int sub, int c) {
        if constexpr (L == 4) return (sub & 1) * NC + c;
        else return c;
    }
    // The extreme tap rows/columns of the patch that this lane owns: for
    // L = 4 one row (row 0 or row NT-1) and one column, so the group's four
    // lanes hold the four corners of the tap rectangle; for L = 2 one row
    // and both columns; for L = 1 all four corners.
    // (patch rows/columns 0 or NT-1, computed, not indexed: a lane-dependent
    // index into the tap arrays would move them to scratch memory)
    static constexpr int NKX = (L == 1) ? 2 : 1;
    static constexpr int NKY = (L == 4) ? 1 : 2;
    __device__ static int kx_row(int sub, int i) {
        if constexpr (L == 1) return i == 0 ? 0 : NT - 1;
        else if constexpr (L == 2) return sub == 0 ? 0 : NT - 1;
        else return (sub >> 1) == 0 ? 0 : NT - 1;
    }
    __device__ static int ky_col(int sub, int i) {
        if constexpr (L == 4) return (sub & 1) == 0 ? 0 : NT - 1;
        else return i == 0 ? 0 : NT - 1;
    }
};

template <int NT, int L>
struct Hoist {
    using B = TapBlock<NT, L>;
    float wt[B::NR * B::NC];
    float wr[B::NR * B::NC];
    // positions relative to the tile origin of the local frame:
    float tx[B::NR];  // clamped tap columns of the lane's rows (src 400-405)
    float ty[B::NC];  // clamped tap rows of the lane's columns
    float kx[B::NKX];  // the lane's corners of the group's tap rectangle
    float ky[B::NKY];
    float inv_w, mean_r, var_r, c;
    float px, py;  // the pixel
};

// The hoisted pixel with its weights read from LDS at each use (pm_refine:
// the tile's weights already sit in LDS, and not holding them in VGPRs
// leaves room for more waves).
template <int NT, int L>
struct HoistRef {
    using B = TapBlock<NT, L>;
    const float* wt;  // LDS, [B::NR * B::NC]
    const float* wr;
    float tx[B::NR];  // relative to the tile origin, as in Hoist
    float ty[B::NC];
    float kx[B::NKX];
    float ky[B::NKY];
    float inv_w, mean_r, var_r, c;
    float px, py;
};

// Fills the tap coordinates of a lane (tx, ty, kx, ky, px, py), relative to
// the tile origin (ox, oy).
template <int NT, int L, class HS>
__device__ inline void hoist_taps(HS& hs, int px, int py, int h, int w, int radius,
                                  int inc, int sub, int ox, int oy) {
    using B = TapBlock<NT, L>;
    hs.px = (float)(px - ox);
    hs.py = (float)(py - oy);
#pragma unroll
    for (int r = 0; r < B::NR; ++r)
        hs.tx[r] = (float)(min(max(px - radius + B::row(sub, r) * inc, 0), w - 1) - ox);
#pragma unroll
    for (int q = 0; q < B::NC; ++q)
        hs.ty[q] = (float)(min(max(py - radius + B::col(sub, q) * inc, 0), h - 1) - oy);
    // (the empty asm keeps the compiler from re-deriving the corners as a
    // lane-dependent pick among tx/ty, which it lowers through scratch memory)
#pragma unroll
    for (int i = 0; i < B::NKX; ++i) {
        int k = B::kx_row(sub, i);
        asm("" : "+v"(k));
        hs.kx[i] = (float)(min(max(px - radius + k * inc, 0), w - 1) - ox);
    }
#pragma unroll
    for (int i = 0; i < B::NKY; ++i) {
        int k = B::ky_col(sub, i);
        asm("" : "+v"(k));
        hs.ky[i] = (float)(min(max(py - radius + k * inc, 0), h - 1) - oy);
    }
}

template <int NT, int L>
__device__ inline void hoist_ref(Hoist<NT, L>& hs, const float* __restrict__ ref, int px, int py,
                                 int h, int w, int radius, int inc, float sigma_spatial,
                                 float sigma_color, int sub, int ox, int oy) {
    using B = TapBlock<NT, L>;
    const float c = ref[py * w + px];
    hs.c = c;
    const float ss = 2.0f * sigma_spatial * sigma_spatial;
    const float sc = 2.0f * sigma_color * sigma_color;
    // the weight's two divisions as products with the reciprocals (weights
    // within ~2 ulp of the quotient form; the NCC is invariant to their
    // common scale and insensitive to ulp-level relative changes)
    const float iss = 1.0f / ss, isc = 1.0f / sc;
    hoist_taps<NT, L, Hoist<NT, L>>(hs, px, py, h, w, radius, inc, sub, ox, oy);
    float W = 0.f, R1 = 0.f, R2 = 0.f;
#pragma unroll
    for (int r = 0; r < B::NR; ++r) {
        const float di = (float)(-radius + B::row(sub, r) * inc);
#pragma unroll
        for (int q = 0; q < B::NC; ++q) {
            const float dj = (float)(-radius + B::col(sub, q) * inc);
            const float rv = ref[((int)hs.ty[q] + oy) * w + (int)hs.tx[r] + ox] - c;
            const float wgt = fexp(-__builtin_amdgcn_sqrtf(di * di + dj * dj) * iss - fabsf(rv) * isc);
            hs.wt[r * B::NC + q] = wgt;
            hs.wr[r * B::NC + q] = wgt * rv;
            W += wgt;
            R1 = fmaf(wgt, rv, R1);
            R2 = fmaf(wgt * rv, rv, R2);
        }
    }
    W = group_sum<L>(W);
    R1 = group_sum<L>(R1);
    R2 = group_sum<L>(R2);
    hs.inv_w = 1.0f / W;
    hs.mean_r = R1 * hs.inv_w;
    hs.var_r = R2 * hs.inv_w - hs.mean_r * hs.mean_r;
}

typedef float __attribute__((ext_vector_type(4))) f32x4;
typedef int __attribute__((ext_vector_type(2))) i32x2;
typedef float __attribute__((ext_vector_type(2))) f32x2;
typedef _Float16 __attribute__((ext_vector_type(2))) f16x2;
typedef __bf16 __attribute__((ext_vector_type(2))) bf16x2;

// The packed cells are read with structured buffer loads: a descriptor whose
// stride is the cell size, over [count] cells, and a per-load record index,
// so a tap's address is one FMA and one conversion (y * w + x is exact in
// FP32 for images of up to 2^24 pixels, which the host checks).
typedef __amdgpu_buffer_rsrc_t rsrc_t;
__device__ f32x4 struct_buffer_load_f32x4(rsrc_t rsrc, int vindex, int voffset, int soffset,
                                          int aux) __asm("llvm.amdgcn.struct.ptr.buffer.load.v4f32");
__device__ i32x2 struct_buffer_load_i32x2(rsrc_t rsrc, int vindex, int voffset, int soffset,
                                          int aux) __asm("llvm.amdgcn.struct.ptr.buffer.load.v2i32");

struct QuadSrc {
    rsrc_t rsrc;  // cells [count]
    int base;     // record index of the view's first cell
};

// gfx9 buffer descriptor word 3: 32-bit data format, no swizzle.
constexpr int kBufWord3 = 0x00020000;

// Storage and arithmetic of the source samples (GMVS_PREC_*):
//   PREC_F32  -- cell (v00, v01-v00, v10-v00, v11-v10-v01+v00) in FP32,
//                16 bytes; bilinear by three FP32 FMAs (the plain path);
//   PREC_F16  -- the same cell in FP16, 8 bytes; the bilinear sample in FP32
//                arithmetic with the FP16 operands widened inside the FMA
//                (v_fma_mix_f32). Integer intensities up to 2048 and their
//                differences are stored exactly, so 8-bit images give the
//                FP32 path's results bit for bit with half its gather bytes;
//   PREC_BF16 -- the cell in BF16, 8 bytes, and the bilinear sample s - c as
//                two BF16 dot products with FP32 accumulation and one FP32
//                FMA: dot((v00, v01-v00), (1, fx))
//                + [fy * dot((v10-v00, d11), (1, fx)) - c]. fx carries BF16
//                rounding (8-bit mantissa); fy stays FP32.
// The NCC moments are accumulated in FP32 about the centre intensity in
// every mode.
enum : int { PREC_F32 = 0, PREC_F16 = 1, PREC_BF16 = 2 };

template <int PR>
struct Cell;

template <>
struct Cell<PREC_F32> {
    static constexpr int kBytes = 16;
    f32x4 q;
    __device__ void load(const QuadSrc& s, int idx) {
        q = struct_buffer_load_f32x4(s.rsrc, s.base + idx, 0, 0, 0);
    }
    __device__ float sample(float fx, float fy, float c) const {
        const float top = fmaf(fx, q.y, q.x);
        const float dyv = fmaf(fx, q.w, q.z);
        return fmaf(fy, dyv, top) - c;
    }
};

template <>
struct Cell<PREC_F16> {
    static constexpr int kBytes = 8;
    i32x2 q;
    __device__ void load(const QuadSrc& s, int idx) {
        q = struct_buffer_load_i32x2(s.rsrc, s.base + idx, 0, 0, 0);
    }
    __device__ float sample(float fx, float fy, float c) const {
        // (elements copied out first: clang's __builtin_bit_cast of a vector
        // element reads the vector's first element)
        const int q0 = q.x, q1 = q.y;
        const f16x2 a = __builtin_bit_cast(f16x2, q0), b = __builtin_bit_cast(f16x2, q1);
        const float top = fmaf(fx, (float)a.y, (float)a.x);
        const float dyv = fmaf(fx, (float)b.y, (float)b.x);
        return fmaf(fy, dyv, top) - c;
    }
};

template <>
struct Cell<PREC_BF16> {
    static constexpr int kBytes = 8;
    i32x2 q;
    __device__ void load(const QuadSrc& s, int idx) {
        q = struct_buffer_load_i32x2(s.rsrc, s.base + idx, 0, 0, 0);
    }
    __device__ float sample(float fx, float fy, float c) const {
        const int q0 = q.x, q1 = q.y;  // see Cell<PREC_F16>::sample
        // both rows of the cell against one weight pair (1, fx): the row
        // values v00 + fx d01 and d10 + fx d11, then the FP32 row blend by fy
        const bf16x2 a = __builtin_convertvector((f32x2){1.0f, fx}, bf16x2);
        // the products are selected as the accumulate-in-place form
        // (v_dot2c): the lower row's blend fy * (d10 + fx d11) - c becomes
        // the accumulator of the upper row's dot product, so only the first
        // product needs its accumulator set (to zero)
        const float dyv = __builtin_amdgcn_fdot2_f32_bf16(__builtin_bit_cast(bf16x2, q1), a, 0.0f, false);
        return __builtin_amdgcn_fdot2_f32_bf16(__builtin_bit_cast(bf16x2, q0), a, fmaf(fy, dyv, -c), false);
    }
};

// The descriptor of a view's cells. cells and count must be uniform across
// the wave (they are: one view of one problem at a time); readfirstlane
// states it, so the descriptor lives in SGPRs even when the compiler cannot
// prove the uniformity (values loaded from the problem table with vector
// loads), instead of wrapping every gather in a waterfall loop.
template <int PR>
__device__ inline QuadSrc quad_src(const void* cells, uint32_t count, int base) {
    const uint64_t a = reinterpret_cast<uint64_t>(cells);
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)a);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(a >> 32));
    void* ua = reinterpret_cast<void*>(((uint64_t)hi << 32) | lo);
    QuadSrc q;
    q.rsrc = __builtin_amdgcn_make_buffer_rsrc(ua, (short)Cell<PR>::kBytes,
                                               (int)__builtin_amdgcn_readfirstlane(count), kBufWord3);
    q.base = base;
    return q;
}

// Source cell grid of one view: the image's h x w cells, padded with kPad
// columns on the right and kPad rows at the bottom, row pitch wp = w + kPad.
// A pad cell is a copy of the nearest edge cell; the edge cells' x (y)
// differences are zero (the reference clamps the upper neighbour index), so
// a tap beyond the right (bottom) edge samples the same value from a pad cell
// at its own floor index as from the edge cell at the clamped index.
constexpr int kPad = 32;

struct SrcGrid {
    int h, w;    // image size (bounds of the centre test and of the clamps)
    int wp;      // cell row pitch, w + kPad
    int fast;    // nonzero: clamp-free taps allowed
};

// The cells of one view as the taps of one tile read them: clamp-free taps
// index relative to the frame's anchor cell, clamped taps absolutely.
struct TapSrc {
    QuadSrc fast;     // records from the anchor cell on
    QuadSrc clamped;  // every record of the view's grid
    float ax, ay;     // the anchor (absolute = local + anchor)
};

template <int PR>
__device__ inline TapSrc tap_src(const void* view_cells, uint32_t vcells, const Frame& fr) {
    TapSrc t;
    const char* c = reinterpret_cast<const char*>(view_cells);
    t.fast = quad_src<PR>(c + (size_t)fr.base * Cell<PR>::kBytes, vcells - (uint32_t)fr.base, 0);
    t.clamped = quad_src<PR>(view_cells, vcells, 0);
    t.ax = fr.ax;
    t.ay = fr.ay;
    return t;
}

// The same over a descriptor of several views' grids (count records), the
// view's grid starting at record view_base: the bases are per lane.
template <int PR>
__device__ inline TapSrc tap_src_at(const void* cells, uint32_t count, int view_base,
                                    const Frame& fr) {
    TapSrc t;
    t.clamped = quad_src<PR>(cells, count, view_base);
    t.fast = t.clamped;
    t.fast.base = view_base + fr.base;
    t.ax = fr.ax;
    t.ay = fr.ay;
    return t;
}

// Floor to int in one instruction (v_cvt_flr_i32_f32).
__device__ inline int cvt_flr(float x) {
    int r;
    asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(r) : "v"(x));
    return r;
}

// Clamp-free test of one lane group's patch: the lanes map the corners of
// the tap rectangle (each lane its own, TapBlock::kx_row/ky_col); when z has
// the centre's sign at every corner, z keeps that sign on the rectangle and
// the projective image of every tap lies in the corners' convex hull. With
// the corners inside [m, w + kPad - m] x [m, h + kPad - m] (m = 1/16 px,
// far above the FP32 rounding of a tap position) every tap's floor cell is
// inside the padded grid and none is left of or above the image, so the
// taps need no clamping: the padded grid gives them the reference's clamped
// values. Returns the mask of the wave's lanes whose corner passes (a NaN
// fails every comparison); inactive lanes read as failing.
template <int NT, int L, class HS>
__device__ inline uint64_t corners_in_grid(const HS& hs, const float H[9], float zc,
                                           const SrcGrid& g, float ax, float ay) {
    using B = TapBlock<NT, L>;
    constexpr float m = 0.0625f;
    // local bounds: the padded grid, and at least one row below the anchor
    // row (a non-negative cell index relative to the anchor cell)
    const float xmin = m - ax, xmax = (float)(g.w + kPad) - m - ax;
    const float ymin = 1.0f + m, ymax = (float)(g.h + kPad) - m - ay;
    // each comparison's lane mask straight from its compare (no per-lane
    // boolean arithmetic)
    uint64_t ok = ~0ull;
#pragma unroll
    for (int i = 0; i < B::NKX; ++i) {
        const float au = fmaf(H[0], hs.kx[i], H[2]);
        const float av = fmaf(H[3], hs.kx[i], H[5]);
        const float az = fmaf(H[6], hs.kx[i], H[8]);
#pragma unroll
        for (int j = 0; j < B::NKY; ++j) {
            const float z = fmaf(H[7], hs.ky[j], az);
            const float rz = __builtin_amdgcn_rcpf(z);
            const float x = fmaf(H[1], hs.ky[j], au) * rz;
            const float y = fmaf(H[4], hs.ky[j], av) * rz;
            ok &= __builtin_amdgcn_ballot_w64(z * zc > 0.0f) & __builtin_amdgcn_ballot_w64(x >= xmin) &
                  __builtin_amdgcn_ballot_w64(x <= xmax) & __builtin_amdgcn_ballot_w64(y >= ymin) &
                  __builtin_amdgcn_ballot_w64(y <= ymax);
        }
    }
    return ok;
}

// One bilateral NCC of the hoisted pixel against one source view. A tap is
// one gather of a packed cell and its bilinear sample. ALLFIRST: each lane
// issues all its gathers before consuming the first (more loads in flight
// per wave, for scattered footprints); else a patch row at a time. All
// lanes of a group return the same cost.
//
// Tap index and fraction, two forms with identical results:
//  - clamped (the reference's rule, src 16-31): index clamp(floor(x), 0,
//    w-1), fraction x - index (negative left of the image: linear
//    extrapolation; beyond the right edge the edge cell's zero differences
//    repeat the edge value);
//  - clamp-free, taken when every lane group of the wave passes
//    corners_in_grid: floor(x) as the index into the padded grid and
//    fract(x) = x - floor(x) as the fraction. Inside the image both forms
//    are the same numbers; on the pad the value only depends on (v00, d10)
//    or (v00, d01) of the edge cell, whose other terms are multiplied by
//    exact zeros.
// The taps of one lane in one of the two index forms, accumulated into the
// moments. NG rows of gathers are in flight together.
template <int NT, int L, int PR, int NG, bool FAST, class HS>
__device__ inline void ncc_taps(const HS& hs, const float H[9], const TapSrc& src,
                                const SrcGrid& g, float& S1, float& S2, float& X) {
    using B = TapBlock<NT, L>;
    // clamp bounds of the local indices, the reference's [0, n - 1]
    const float xlo = -src.ax, xhi = (float)(g.w - 1) - src.ax;
    const float ylo = -src.ay, yhi = (float)(g.h - 1) - src.ay;
    const float wpf = (float)g.wp;
#pragma unroll
    for (int r0 = 0; r0 < B::NR; r0 += NG) {
        Cell<PR> q[NG * B::NC];
        float fx[NG * B::NC], fy[NG * B::NC];
#pragma unroll
        for (int rr = 0; rr < NG; ++rr) {
            const int r = r0 + rr;
            const float au = fmaf(H[0], hs.tx[r], H[2]);
            const float av = fmaf(H[3], hs.tx[r], H[5]);
            const float az = fmaf(H[6], hs.tx[r], H[8]);
#pragma unroll
            for (int j = 0; j < B::NC; ++j) {
                const float rz = __builtin_amdgcn_rcpf(fmaf(H[7], hs.ty[j], az));
                const float x = fmaf(H[1], hs.ty[j], au) * rz;
                const float y = fmaf(H[4], hs.ty[j], av) * rz;
                const int t = rr * B::NC + j;
                if constexpr (FAST) {
                    // y >= 1 and x >= -ax: the index relative to the anchor
                    // cell is non-negative (the unsigned sum wraps exactly)
                    fx[t] = __builtin_amdgcn_fractf(x);
                    fy[t] = __builtin_amdgcn_fractf(y);
                    q[t].load(src.fast, (int)(__umul24((unsigned)cvt_flr(y), (unsigned)g.wp) +
                                              (unsigned)cvt_flr(x)));
                } else {
#pragma clang fp contract(off)
                    // clamp(trunc(c), 0, n-1) of the reference equals
                    // clamp(floor(c), 0, n-1); any NaN/inf lands inside
                    // [0, n-1]. The fraction is the rounded x minus the
                    // index (no contraction into x's product), as fract()
                    // forms it in the clamp-free form.
                    const float xf = __builtin_amdgcn_fmed3f(floorf(x), xlo, xhi);
                    const float yf = __builtin_amdgcn_fmed3f(floorf(y), ylo, yhi);
                    fx[t] = x - xf;
                    fy[t] = y - yf;
                    q[t].load(src.clamped, (int)(uint32_t)fmaf(yf + src.ay, wpf, xf + src.ax));
                }
            }
        }
#pragma unroll
        for (int rr = 0; rr < NG; ++rr)
#pragma unroll
            for (int j = 0; j < B::NC; ++j) {
                const int t = rr * B::NC + j;
                // explicit FMAs: both index forms round identically
                const float s = q[t].sample(fx[t], fy[t], hs.c);
                const float w = hs.wt[(r0 + rr) * B::NC + j];
                S1 = fmaf(w, s, S1);
                S2 = fmaf(w * s, s, S2);
                X = fmaf(hs.wr[(r0 + rr) * B::NC + j], s, X);
            }
        // keep the next rows' gathers behind this batch's use: hoisting them
        // would hold more cells in VGPRs than the occupancy affords
        if constexpr (NG < B::NR) __builtin_amdgcn_sched_barrier(0);
    }
}

// One bilateral NCC of the hoisted pixel against one source view. A tap is
// one gather of a packed cell and its bilinear sample. ALLFIRST: a 4-lane
// group's lane issues all its gathers before consuming the first (more loads
// in flight per wave, for scattered footprints); else (and for narrower
// groups, whose lanes own twice the taps) a patch row at a time. All lanes
// of a group return the same cost.
//
// Tap index and fraction, two forms with identical results:
//  - clamped (the reference's rule, src 16-31): index clamp(floor(x), 0,
//    w-1), fraction x - index (negative left of the image: linear
//    extrapolation; beyond the right edge the edge cell's zero differences
//    repeat the edge value);
//  - clamp-free, taken when every lane group of the wave passes
//    corners_in_grid: floor(x) as the index into the padded grid and
//    fract(x) = x - floor(x) as the fraction. Inside the image both forms
//    are the same numbers; on the pad the value only depends on (v00, d10)
//    or (v00, d01) of the edge cell, whose other terms are multiplied by
//    exact zeros.
template <int NT, int L, int PR, bool ALLFIRST, class HS>
__device__ inline float ncc_eval_t(const HS& hs, const float H[9], const TapSrc& src,
                                   const SrcGrid& g) {
    using B = TapBlock<NT, L>;
    constexpr int NG = (ALLFIRST && L == 4) ? B::NR : 1;
    const float zc = fmaf(H[6], hs.px, fmaf(H[7], hs.py, H[8]));
    const float rcz = __builtin_amdgcn_rcpf(zc);
    const float cx = fmaf(H[0], hs.px, fmaf(H[1], hs.py, H[2])) * rcz;
    const float cy = fmaf(H[3], hs.px, fmaf(H[4], hs.py, H[5])) * rcz;
    // the centre inside [0, w) x [0, h) (src 429-432), in local coordinates
    if (cx >= (float)g.w - src.ax || cx < -src.ax || cy >= (float)g.h - src.ay || cy < -src.ay)
        return 2.0f;
    // wave-uniform choice (the ballot covers the lanes still active here)
    const bool fast =
        g.fast && (__builtin_amdgcn_read_exec() &
                   ~corners_in_grid<NT, L, HS>(hs, H, zc, g, src.ax, src.ay)) == 0;
    float S1 = 0.f, S2 = 0.f, X = 0.f;
    if (fast)
        ncc_taps<NT, L, PR, NG, true, HS>(hs, H, src, g, S1, S2, X);
    else
        ncc_taps<NT, L, PR, NG, false, HS>(hs, H, src, g, S1, S2, X);
    S1 = group_sum<L>(S1);
    S2 = group_sum<L>(S2);
    X = group_sum<L>(X);
    const float ms = S1 * hs.inv_w;
    const float var_s = S2 * hs.inv_w - ms * ms;
    if (hs.var_r < 1e-5f || var_s < 1e-5f) return 2.0f;
    const float cov = X * hs.inv_w - hs.mean_r * ms;
    return fmaxf(0.0f, fminf(2.0f, 1.0f - cov * __builtin_amdgcn_rsqf(hs.var_r * var_s)));
}

template <int NT, int L, int PR, class HS>
__device__ inline float ncc_eval(const HS& hs, const float H[9], const TapSrc& src,
                                 const SrcGrid& g) {
    return ncc_eval_t<NT, L, PR, false, HS>(hs, H, src, g);
}

template <int NT, int L, int PR, class HS>
__device__ inline float ncc_eval_mlp(const HS& hs, const float H[9], const TapSrc& src,
                                     const SrcGrid& g) {
    return ncc_eval_t<NT, L, PR, true, HS>(hs, H, src, g);
}

}  // namespace gmvs
