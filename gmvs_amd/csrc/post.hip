// After fusion (gfx950): the point gather and its text export, and the three
// segment/normal utilities the reference exports beside PlanarInit.
//
//   gmvs_points_gather     the gather of mvs_fusion_segmentaion_cpu.py:158-205
//                          (world points + normals of depth > 0 pixels, the
//                          70-degree view-angle filter, depth_normal_filter)
//   gmvs_format_rows       np.savetxt(fmt="%.4f") of those rows as bytes
//                          (mvs_fusion_segmentaion_cpu.py:207-216)
//   gmvs_sliding_window_normal   sliding_window_normal (gmvs/lib/src/planar_initer.cu:471-535)
//   gmvs_filter_by_var_map       filter_by_var_map (planar_initer.cu:634-684)
//   gmvs_fusion_planar_mask      fusion_planar_mask (planar_initer.cu:537-632)
//
// Every result is bit-identical to a sequential CPU loop over the same
// expressions: this file is compiled with -ffp-contract=off and correctly
// rounded division, and every floating-point sum keeps the reference's order
// (per-pixel window sums row by row; per-segment sums over the segment's
// pixels in row-major order, which a stable radix sort by segment id yields).
// The gather works in FP64, numpy's promotion of the reference's float32
// pixel grid against float64 intrinsics and poses.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>

#include <cmath>
#include <cstring>

#include "abi_common.h"

#pragma clang fp contract(off)

namespace gmvs {
namespace {

constexpr int kBlock = 256;               // threads per block
constexpr int kRounds = 4;                // pixels per thread in the gather
constexpr int kTile = kBlock * kRounds;   // pixels per gather block
constexpr int kScanThreads = 1024;

size_t align_up(size_t x) { return (x + 255) & ~size_t(255); }

// Exclusive prefix over the block of a per-thread value (kBlock threads):
// wave-level inclusive scan by shuffles, then the wave totals through LDS.
// Returns the exclusive prefix; *total receives the block total.
__device__ inline uint64_t block_exclusive_scan(uint64_t v, uint64_t* total) {
    __shared__ uint64_t s_wave[kBlock / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    uint64_t x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint64_t y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) s_wave[wid] = x;
    __syncthreads();
    uint64_t before = 0, all = 0;
#pragma unroll
    for (int i = 0; i < kBlock / 64; ++i) {
        before += (i < wid) ? s_wave[i] : 0;
        all += s_wave[i];
    }
    __syncthreads();  // s_wave is reused by the next call
    *total = all;
    return before + x - v;
}

// Exclusive scan of n (a, b) count pairs in place into 64-bit offsets, by
// one block: each thread scans a contiguous chunk sequentially, the chunk
// totals are scanned across the block. totals[0..1] receive the sums.
__global__ void __launch_bounds__(kScanThreads) scan_pairs(const uint32_t* __restrict__ cnt, int64_t n,
                                                          uint64_t* __restrict__ off,
                                                          uint64_t* __restrict__ totals) {
    __shared__ uint64_t s_a[kScanThreads], s_b[kScanThreads];
    const int t = threadIdx.x;
    const int64_t chunk = (n + kScanThreads - 1) / kScanThreads;
    const int64_t lo = min(n, t * chunk), hi = min(n, lo + chunk);
    uint64_t a = 0, b = 0;
    for (int64_t i = lo; i < hi; ++i) {
        a += cnt[2 * i];
        b += cnt[2 * i + 1];
    }
    s_a[t] = a;
    s_b[t] = b;
    __syncthreads();
    for (int d = 1; d < kScanThreads; d <<= 1) {  // Hillis-Steele, inclusive
        const uint64_t ya = (t >= d) ? s_a[t - d] : 0, yb = (t >= d) ? s_b[t - d] : 0;
        __syncthreads();
        s_a[t] += ya;
        s_b[t] += yb;
        __syncthreads();
    }
    a = s_a[t] - a;
    b = s_b[t] - b;
    for (int64_t i = lo; i < hi; ++i) {
        off[2 * i] = a;
        off[2 * i + 1] = b;
        a += cnt[2 * i];
        b += cnt[2 * i + 1];
    }
    if (t == kScanThreads - 1) {
        totals[0] = s_a[t];
        totals[1] = s_b[t];
    }
}

// ---------------------------------------------------------------- gather --

struct PtsView {
    const float4* dn;  // [H*W] (depth, world normal)
    double R[9];       // camera-to-world rotation, row-major (pose rows 0-2, cols 0-2)
    double C[3];       // camera centre (pose column 3)
};

struct PtsConst {
    int h, w;
    double fx, fy, cx, cy, cos_view;
};

// One pixel of the gather. dir = R (x', y', 1) with x' = (col - cx) / fx and
// y' = (row - cy) / fy (the reference's `dirs_ @ rot_mat.T`), point =
// C + depth * dir, and the filter -dir . n > cos_view, summed in index order
// as numpy's reduction over the last axis does.
struct PtsPixel {
    bool valid, pass;
    double p[3], n[3];
};

__device__ inline PtsPixel pts_pixel(const PtsView& v, const PtsConst& k, int pix, float4 dn) {
    PtsPixel o;
    o.valid = dn.x > 0.0f;
    const int row = pix / k.w, col = pix - row * k.w;
    const double xd = ((double)col - k.cx) / k.fx;
    const double yd = ((double)row - k.cy) / k.fy;
    double dir[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) dir[i] = xd * v.R[3 * i] + yd * v.R[3 * i + 1] + 1.0 * v.R[3 * i + 2];
    const double d = (double)dn.x;
    o.n[0] = (double)dn.y;
    o.n[1] = (double)dn.z;
    o.n[2] = (double)dn.w;
#pragma unroll
    for (int i = 0; i < 3; ++i) o.p[i] = v.C[i] + d * dir[i];
    const double dot = (-dir[0]) * o.n[0] + (-dir[1]) * o.n[1] + (-dir[2]) * o.n[2];
    o.pass = o.valid && (dot > k.cos_view);
    return o;
}

// Pass 1: (valid, filtered) counts of each block of kTile pixels.
__global__ void __launch_bounds__(kBlock) pts_count(PtsView v, PtsConst k, uint32_t* __restrict__ cnt) {
    const int hw = k.h * k.w;
    const int base = blockIdx.x * kTile;
    uint32_t a = 0, b = 0;
#pragma unroll
    for (int r = 0; r < kRounds; ++r) {
        const int pix = base + r * kBlock + threadIdx.x;
        if (pix < hw) {
            const PtsPixel q = pts_pixel(v, k, pix, v.dn[pix]);
            a += q.valid;
            b += q.pass;
        }
    }
    uint64_t ta, tb;
    block_exclusive_scan(a, &ta);
    block_exclusive_scan(b, &tb);
    if (threadIdx.x == 0) {
        cnt[2 * blockIdx.x] = (uint32_t)ta;
        cnt[2 * blockIdx.x + 1] = (uint32_t)tb;
    }
}

// Pass 3: rows in pixel order -- round r of the block holds pixels
// base + r*kBlock + [0, kBlock), so the rank of a pixel is the block offset,
// plus the counts of the earlier rounds, plus the in-round prefix.
__global__ void __launch_bounds__(kBlock) pts_write(PtsView v, PtsConst k, const uint64_t* __restrict__ off,
                                                    double* __restrict__ pts, uint64_t cap,
                                                    double* __restrict__ pts_f, uint64_t cap_f,
                                                    float4* __restrict__ filtered) {
    const int hw = k.h * k.w;
    const int base = blockIdx.x * kTile;
    uint64_t oa = off[2 * blockIdx.x], ob = off[2 * blockIdx.x + 1];
#pragma unroll 1
    for (int r = 0; r < kRounds; ++r) {
        const int pix = base + r * kBlock + threadIdx.x;
        PtsPixel q;
        q.valid = q.pass = false;
        float4 dn = make_float4(0.f, 0.f, 0.f, 0.f);
        if (pix < hw) {
            dn = v.dn[pix];
            q = pts_pixel(v, k, pix, dn);
        }
        uint64_t ta, tb;
        const uint64_t ra = block_exclusive_scan(q.valid, &ta);
        const uint64_t rb = block_exclusive_scan(q.pass, &tb);
        if (q.valid && pts && oa + ra < cap) {
            double* o = pts + 6 * (oa + ra);
            o[0] = q.p[0]; o[1] = q.p[1]; o[2] = q.p[2];
            o[3] = q.n[0]; o[4] = q.n[1]; o[5] = q.n[2];
        }
        if (q.pass && pts_f && ob + rb < cap_f) {
            double* o = pts_f + 6 * (ob + rb);
            o[0] = q.p[0]; o[1] = q.p[1]; o[2] = q.p[2];
            o[3] = q.n[0]; o[4] = q.n[1]; o[5] = q.n[2];
        }
        if (filtered && pix < hw) filtered[pix] = q.pass ? dn : make_float4(0.f, 0.f, 0.f, 0.f);
        oa += ta;
        ob += tb;
    }
}

struct PtsLayout {
    size_t cnt, off, total;
};

PtsLayout pts_layout(int h, int w, int n) {
    const size_t blocks = size_t(n) * ((size_t(h) * w + kTile - 1) / kTile);
    PtsLayout L;
    size_t o = 0;
    L.cnt = o;
    o = align_up(o + 8 * blocks);
    L.off = o;
    o = align_up(o + 16 * blocks);
    L.total = o;
    return L;
}

// ---------------------------------------------------------------- format --

__constant__ uint64_t kPow10[10] = {1ull,      10ull,      100ull,      1000ull,      10000ull,
                                    100000ull, 1000000ull, 10000000ull, 100000000ull, 1000000000ull};

// One value as Python's "%.<dec>f" prints it: the exact binary value rounded
// to dec decimals, ties to even; a sign for every negative value (-0.0 and
// values that round to zero included); nan, inf, -inf. Values whose scaled
// magnitude reaches 2^63 are not supported (*bad is set).
struct Fixed {
    int kind;      // 0 number, 1 nan, 2 inf
    bool neg;
    uint64_t ip;   // integer part
    uint64_t fp;   // fraction digits as an integer < 10^dec
    int ip_digits;
};

__device__ inline Fixed to_fixed(double v, int dec, bool* bad) {
    Fixed f;
    const uint64_t bits = (uint64_t)__double_as_longlong(v);
    f.neg = (bits >> 63) != 0;
    const int ex = (int)((bits >> 52) & 0x7FF);
    uint64_t m = bits & ((1ull << 52) - 1);
    f.ip = f.fp = 0;
    f.ip_digits = 1;
    if (ex == 0x7FF) {
        f.kind = m ? 1 : 2;
        if (m) f.neg = false;  // Python prints nan without a sign
        return f;
    }
    f.kind = 0;
    int e;
    if (ex == 0) {
        e = -1074;
    } else {
        m |= 1ull << 52;
        e = ex - 1075;
    }
    // |v| = m * 2^e; N = round_half_even(|v| * 10^dec)
    const uint64_t p10 = kPow10[dec];
    const unsigned __int128 num = (unsigned __int128)m * p10;  // < 2^83
    unsigned __int128 q;
    if (e >= 0) {
        if (e > 44 || (num >> (63 - e)) != 0) {
            *bad = true;
            return f;
        }
        q = num << e;
    } else {
        const int s = -e;
        if (s > 84) {
            q = 0;  // num < 2^83 < half
        } else {
            q = num >> s;
            const unsigned __int128 r = num - (q << s);
            const unsigned __int128 half = (unsigned __int128)1 << (s - 1);
            if (r > half || (r == half && (q & 1))) q += 1;
        }
    }
    if ((q >> 63) != 0) {
        *bad = true;
        return f;
    }
    const uint64_t N = (uint64_t)q;
    f.ip = N / p10;
    f.fp = N - f.ip * p10;
    int nd = 1;
    for (uint64_t t = f.ip; t >= 10; t /= 10) ++nd;
    f.ip_digits = nd;
    return f;
}

__device__ inline int fixed_len(const Fixed& f, int dec) {
    if (f.kind == 1) return 3;
    if (f.kind == 2) return f.neg ? 4 : 3;
    return (int)f.neg + f.ip_digits + (dec ? 1 + dec : 0);
}

__device__ inline char* fixed_write(const Fixed& f, int dec, char* p) {
    if (f.kind == 1) {
        p[0] = 'n'; p[1] = 'a'; p[2] = 'n';
        return p + 3;
    }
    if (f.neg) *p++ = '-';
    if (f.kind == 2) {
        p[0] = 'i'; p[1] = 'n'; p[2] = 'f';
        return p + 3;
    }
    uint64_t t = f.ip;
    for (int i = f.ip_digits - 1; i >= 0; --i) {
        p[i] = (char)('0' + (int)(t % 10));
        t /= 10;
    }
    p += f.ip_digits;
    if (dec) {
        *p++ = '.';
        uint64_t fr = f.fp;
        for (int i = dec - 1; i >= 0; --i) {
            p[i] = (char)('0' + (int)(fr % 10));
            fr /= 10;
        }
        p += dec;
    }
    return p;
}

// Row length: ncol values, ncol - 1 separating spaces and a newline.
__device__ inline uint32_t row_len(const double* __restrict__ row, int ncol, int dec, bool* bad) {
    uint32_t n = (uint32_t)ncol;
    for (int c = 0; c < ncol; ++c) n += (uint32_t)fixed_len(to_fixed(row[c], dec, bad), dec);
    return n;
}

// Pass 1: bytes of each block of kBlock rows (in the .x slot of a count
// pair; .y stays 0) and the unsupported-value flag.
__global__ void __launch_bounds__(kBlock) fmt_count(const double* __restrict__ rows, uint64_t n, int ncol,
                                                    int dec, uint32_t* __restrict__ cnt,
                                                    uint32_t* __restrict__ bad_flag) {
    const uint64_t r = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    bool bad = false;
    const uint64_t len = (r < n) ? row_len(rows + r * ncol, ncol, dec, &bad) : 0;
    if (bad) bad_flag[0] = 1;
    uint64_t tot;
    block_exclusive_scan(len, &tot);
    if (threadIdx.x == 0) {
        cnt[2 * blockIdx.x] = (uint32_t)tot;
        cnt[2 * blockIdx.x + 1] = 0;
    }
}

// Pass 3: every row written at its byte offset (bytes past cap are dropped).
__global__ void __launch_bounds__(kBlock) fmt_write(const double* __restrict__ rows, uint64_t n, int ncol,
                                                    int dec, const uint64_t* __restrict__ off,
                                                    char* __restrict__ out, uint64_t cap) {
    const uint64_t r = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    bool bad = false;
    const uint64_t len = (r < n) ? row_len(rows + r * ncol, ncol, dec, &bad) : 0;
    uint64_t tot;
    const uint64_t pos = off[2 * blockIdx.x] + block_exclusive_scan(len, &tot);
    if (r >= n || bad) return;
    char buf[24];
    uint64_t o = pos;
    for (int c = 0; c < ncol; ++c) {
        const Fixed f = to_fixed(rows[r * ncol + c], dec, &bad);
        const int l = (int)(fixed_write(f, dec, buf) - buf);
        for (int i = 0; i < l; ++i, ++o)
            if (o < cap) out[o] = buf[i];
        if (o < cap) out[o] = (c + 1 < ncol) ? ' ' : '\n';
        ++o;
    }
}

__global__ void fmt_finish(const uint64_t* __restrict__ totals, const uint32_t* __restrict__ bad_flag,
                           uint64_t* __restrict__ out_len) {
    if (threadIdx.x == 0 && blockIdx.x == 0) out_len[0] = bad_flag[0] ? ~0ull : totals[0];
}

__global__ void fill_u32(uint32_t* __restrict__ p, int64_t n, uint32_t v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

struct FmtLayout {
    size_t cnt, off, totals, bad, total;
};

FmtLayout fmt_layout(uint64_t n_rows) {
    const size_t blocks = (n_rows + kBlock - 1) / kBlock;
    FmtLayout L;
    size_t o = 0;
    L.cnt = o;
    o = align_up(o + 8 * blocks);
    L.off = o;
    o = align_up(o + 16 * blocks);
    L.totals = o;
    o = align_up(o + 16);
    L.bad = o;
    o = align_up(o + 4);
    L.total = o;
    return L;
}

// ------------------------------------------------- sliding window normal --

// sliding_window_normal (planar_initer.cu:471-535): pixels whose planar
// label is 1 get (0, 0, 0); the others the mean normal of the window's
// non-planar pixels, summed row by row as cv::Vec3f additions and divided
// as cv::Vec3f /= int does (multiplication by the double reciprocal of the
// count, rounded to float).
__global__ void __launch_bounds__(kBlock) swn_kernel(const int32_t* __restrict__ planar,
                                                     const float* __restrict__ normal, int h, int w,
                                                     int half, float* __restrict__ out) {
    const int pix = blockIdx.x * kBlock + threadIdx.x;
    if (pix >= h * w) return;
    const int row = pix / w, col = pix - row * w;
    float nx = 0.f, ny = 0.f, nz = 0.f;
    if (planar[pix] != 1) {
        int count = 0;
        const int r0 = max(row - half, 0), r1 = min(row + half, h - 1);
        const int c0 = max(col - half, 0), c1 = min(col + half, w - 1);
        for (int r = r0; r <= r1; ++r) {
            for (int c = c0; c <= c1; ++c) {
                const int q = r * w + c;
                if (planar[q] == 1) continue;
                ++count;
                nx += normal[3 * q];
                ny += normal[3 * q + 1];
                nz += normal[3 * q + 2];
            }
        }
        const double ia = 1.0 / (double)count;  // count >= 1: the centre counts
        nx = (float)((double)nx * ia);
        ny = (float)((double)ny * ia);
        nz = (float)((double)nz * ia);
    }
    out[3 * pix] = nx;
    out[3 * pix + 1] = ny;
    out[3 * pix + 2] = nz;
}

// -------------------------------------------------------------- segments --

// Pixels grouped by segment id, each group in row-major order: a stable
// radix sort of (id, pixel index) pairs, and the [start, end) range of
// every id.
struct SegLayout {
    size_t keys_in, keys_out, idx_in, idx_out, start, end, label, sort_tmp, total;
    size_t sort_bytes;
};

int seg_bits(int num_segments) {
    int b = 1;
    while (b < 31 && (1 << b) < num_segments) ++b;
    return b;
}

size_t sort_tmp_bytes(int hw, int num_segments) {
    size_t bytes = 0;
    const hipError_t e = rocprim::radix_sort_pairs(
        nullptr, bytes, (const uint32_t*)nullptr, (uint32_t*)nullptr, (const uint32_t*)nullptr,
        (uint32_t*)nullptr, (unsigned int)hw, 0u, (unsigned int)seg_bits(num_segments));
    return e == hipSuccess ? bytes : 0;
}

SegLayout seg_layout(int h, int w, int num_segments) {
    const size_t hw = size_t(h) * w, ns = size_t(num_segments);
    SegLayout L;
    size_t o = 0;
    L.keys_in = o;
    o = align_up(o + 4 * hw);
    L.keys_out = o;
    o = align_up(o + 4 * hw);
    L.idx_in = o;
    o = align_up(o + 4 * hw);
    L.idx_out = o;
    o = align_up(o + 4 * hw);
    L.start = o;
    o = align_up(o + 4 * ns);
    L.end = o;
    o = align_up(o + 4 * ns);
    L.label = o;
    o = align_up(o + 4 * ns);
    L.sort_bytes = sort_tmp_bytes((int)hw, num_segments);
    L.sort_tmp = o;
    o = align_up(o + L.sort_bytes + 256);
    L.total = o;
    return L;
}

// Keys: the segment ids (ids outside [0, num_segments) count as 0, which the
// host rules out for the reference's behaviour); values: pixel indices.
__global__ void seg_keys(const int32_t* __restrict__ ids, int hw, int ns, uint32_t* __restrict__ keys,
                         uint32_t* __restrict__ idx) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= hw) return;
    const int s = ids[i];
    keys[i] = (s >= 0 && s < ns) ? (uint32_t)s : 0u;
    idx[i] = (uint32_t)i;
}

__global__ void seg_bounds(const uint32_t* __restrict__ keys, int hw, uint32_t* __restrict__ start,
                           uint32_t* __restrict__ end) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= hw) return;
    const uint32_t k = keys[i];
    if (i == 0 || keys[i - 1] != k) start[k] = (uint32_t)i;
    if (i == hw - 1 || keys[i + 1] != k) end[k] = (uint32_t)(i + 1);
}

// filter_by_var_map, per segment s >= 1: mean = (sequential float sum of
// var over the segment) / (float count); kept if mean <= var_thresh (an empty
// segment gives 0/0 = NaN: not kept).
__global__ void vm_segments(const float* __restrict__ var, const uint32_t* __restrict__ idx,
                            const uint32_t* __restrict__ start, const uint32_t* __restrict__ end,
                            int ns, float var_thresh, int32_t* __restrict__ label) {
    const int s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= ns) return;
    if (s == 0) {
        label[0] = 0;
        return;
    }
    float acc = 0.f, count = 0.f;
    for (uint32_t i = start[s]; i < end[s]; ++i) {
        count += 1.f;
        acc += var[idx[i]];
    }
    const float mean = acc / count;
    label[s] = (mean <= var_thresh) ? s : 0;
}

// fusion_planar_mask, per segment s >= 1 (planar_initer.cu:576-629).
// label: 0 (dropped), s (kept), 300 (object), or -1 (textureless: its id
// 301, 302, ... in segment order is assigned by seg_vps_ids).
__global__ void pm_segments(const int32_t* __restrict__ planar, const float* __restrict__ normals,
                            const uint32_t* __restrict__ idx, const uint32_t* __restrict__ start,
                            const uint32_t* __restrict__ end, int ns, int32_t thresh,
                            float nonplanar_percent, float cos_sim_thresh, float match_ratio_thresh,
                            int32_t* __restrict__ label) {
    const int s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= ns) return;
    if (s == 0) {
        label[0] = 0;
        return;
    }
    int32_t np = 0, pl = 0;
    float sx = 0.f, sy = 0.f, sz = 0.f;
    const uint32_t b = start[s], e = end[s];
    for (uint32_t i = b; i < e; ++i) {
        const uint32_t p = idx[i];
        if (planar[p] == 0) {
            ++np;
            sx += normals[3 * p];
            sy += normals[3 * p + 1];
            sz += normals[3 * p + 2];
        } else {
            ++pl;
        }
    }
    const float nonplanar = (float)np, planar_f = (float)pl;
    const float amount = nonplanar + planar_f;
    int32_t out;
    if (amount <= (float)thresh) {
        out = 0;
    } else if (nonplanar / amount >= nonplanar_percent) {
        // cv::Vec3f /= float: multiplication by the float reciprocal
        const float ia = 1.f / nonplanar;
        const float ax = sx * ia, ay = sy * ia, az = sz * ia;
        int32_t valid = 0;
        for (uint32_t i = b; i < e; ++i) {
            const uint32_t p = idx[i];
            if (planar[p] != 0) continue;
            // cv::Vec3f::dot: float sum from 0 in index order
            float cs = 0.f;
            cs += ax * normals[3 * p];
            cs += ay * normals[3 * p + 1];
            cs += az * normals[3 * p + 2];
            valid += (cs > cos_sim_thresh) ? 1 : 0;
        }
        const float ratio = (float)valid / nonplanar;
        out = (ratio > match_ratio_thresh) ? s : 300;
    } else if ((double)(nonplanar / amount) < 0.25) {
        out = -1;
    } else {
        out = s;
    }
    label[s] = out;
}

// Textureless segments get 301, 302, ... in increasing segment id (the
// reference's vpsarea_seg_id counter), by one block scanning the labels.
__global__ void __launch_bounds__(kScanThreads) seg_vps_ids(int32_t* __restrict__ label, int ns) {
    __shared__ int32_t s_c[kScanThreads];
    const int t = threadIdx.x;
    const int chunk = (ns + kScanThreads - 1) / kScanThreads;
    const int lo = min(ns, t * chunk), hi = min(ns, lo + chunk);
    int32_t c = 0;
    for (int i = lo; i < hi; ++i) c += (label[i] == -1);
    s_c[t] = c;
    __syncthreads();
    for (int d = 1; d < kScanThreads; d <<= 1) {
        const int32_t y = (t >= d) ? s_c[t - d] : 0;
        __syncthreads();
        s_c[t] += y;
        __syncthreads();
    }
    int32_t next = 301 + s_c[t] - c;
    for (int i = lo; i < hi; ++i)
        if (label[i] == -1) label[i] = next++;
}

__global__ void seg_paint(const int32_t* __restrict__ ids, int hw, int ns, const int32_t* __restrict__ label, int32_t* __restrict__ out) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= hw) return;
    const int s = ids[i];
    out[i] = (s >= 0 && s < ns) ? label[s] : 0;
}

// Shared front half of the two segment utilities: keys, stable sort, bounds.
int seg_group(const int32_t* ids, int h, int w, int ns, char* ws, const SegLayout& L, hipStream_t st) {
    const int hw = h * w;
    uint32_t* keys_in = reinterpret_cast<uint32_t*>(ws + L.keys_in);
    uint32_t* keys_out = reinterpret_cast<uint32_t*>(ws + L.keys_out);
    uint32_t* idx_in = reinterpret_cast<uint32_t*>(ws + L.idx_in);
    uint32_t* idx_out = reinterpret_cast<uint32_t*>(ws + L.idx_out);
    uint32_t* start = reinterpret_cast<uint32_t*>(ws + L.start);
    uint32_t* end = reinterpret_cast<uint32_t*>(ws + L.end);
    const dim3 gp((hw + kBlock - 1) / kBlock), gs((ns + kBlock - 1) / kBlock), blk(kBlock);
    GMVS_LAUNCH(K_SEG_SORT, st, seg_keys, gp, blk, 0, st, ids, hw, ns, keys_in, idx_in);
    GMVS_LAUNCH_CHECK();
    {
        GMVS_PROF(st, K_SEG_SORT);
        size_t bytes = L.sort_bytes;
        GMVS_HIP_CHECK(rocprim::radix_sort_pairs(ws + L.sort_tmp, bytes, (const uint32_t*)keys_in, keys_out,
                                                 (const uint32_t*)idx_in, idx_out, (unsigned int)hw, 0u,
                                                 (unsigned int)seg_bits(ns), st));
    }
    GMVS_LAUNCH(K_SEG_SORT, st, fill_u32, gs, blk, 0, st, start, (int64_t)ns, 0u);
    GMVS_LAUNCH_CHECK();
    GMVS_LAUNCH(K_SEG_SORT, st, fill_u32, gs, blk, 0, st, end, (int64_t)ns, 0u);
    GMVS_LAUNCH_CHECK();
    GMVS_LAUNCH(K_SEG_SORT, st, seg_bounds, gp, blk, 0, st, keys_out, hw, start, end);
    GMVS_LAUNCH_CHECK();
    return GMVS_OK;
}

}  // namespace
}  // namespace gmvs

using namespace gmvs;

// ------------------------------------------------------------ C entry points

extern "C" size_t gmvs_points_workspace_bytes(int32_t height, int32_t width, int32_t num_views) {
    if (height <= 0 || width <= 0 || num_views <= 0) return 0;
    return pts_layout(height, width, num_views).total;
}

extern "C" int gmvs_points_gather(const gmvs_points_args* a, void* workspace, size_t workspace_bytes,
                                  void* stream) {
    clear_error();
    GMVS_ARG_CHECK(a, "points_gather: NULL args");
    GMVS_ARG_CHECK(a->height > 0 && a->width > 0 && (int64_t)a->height * a->width <= (1 << 30),
                   "points_gather: bad image size %dx%d", a->width, a->height);
    GMVS_ARG_CHECK(a->num_views >= 1, "points_gather: num_views must be >= 1 (got %d)", a->num_views);
    GMVS_ARG_CHECK(a->depth_normals && a->poses && a->counts, "points_gather: NULL argument");
    GMVS_ARG_CHECK(a->fx != 0.0 && a->fy != 0.0, "points_gather: zero focal length");
    GMVS_ARG_CHECK(!a->capacity || a->points, "points_gather: capacity without points");
    GMVS_ARG_CHECK(!a->capacity_filtered || a->points_filtered,
                   "points_gather: capacity_filtered without points_filtered");
    for (int i = 0; i < a->num_views; ++i)
        GMVS_ARG_CHECK(a->depth_normals[i], "points_gather: depth_normals[%d] is NULL", i);
    const PtsLayout L = pts_layout(a->height, a->width, a->num_views);
    GMVS_ARG_CHECK(workspace, "points_gather: workspace is NULL");
    if (workspace_bytes < L.total) {
        set_error("points_gather: workspace too small: %zu < %zu bytes", workspace_bytes, L.total);
        return GMVS_ERR_NOSPACE;
    }
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    char* ws = reinterpret_cast<char*>(workspace);
    uint32_t* cnt = reinterpret_cast<uint32_t*>(ws + L.cnt);
    uint64_t* off = reinterpret_cast<uint64_t*>(ws + L.off);
    const int hw = a->height * a->width;
    const int nblk = (hw + kTile - 1) / kTile;
    PtsConst k;
    k.h = a->height;
    k.w = a->width;
    k.fx = a->fx;
    k.fy = a->fy;
    k.cx = a->cx;
    k.cy = a->cy;
    k.cos_view = a->cos_view_angle;
    auto view = [&](int i) {
        PtsView v;
        v.dn = reinterpret_cast<const float4*>(a->depth_normals[i]);
        const double* P = a->poses + 16 * i;
        for (int r = 0; r < 3; ++r) {
            for (int c = 0; c < 3; ++c) v.R[3 * r + c] = P[4 * r + c];
            v.C[r] = P[4 * r + 3];
        }
        return v;
    };
    for (int i = 0; i < a->num_views; ++i) {
        GMVS_LAUNCH(K_POINTS, st, pts_count, dim3(nblk), dim3(kBlock), 0, st, view(i), k,
                    cnt + 2 * (size_t)i * nblk);
        GMVS_LAUNCH_CHECK();
    }
    GMVS_LAUNCH(K_POINTS, st, scan_pairs, dim3(1), dim3(kScanThreads), 0, st, cnt,
                (int64_t)a->num_views * nblk, off, a->counts);
    GMVS_LAUNCH_CHECK();
    if (!a->points && !a->points_filtered && !a->filtered) return GMVS_OK;
    for (int i = 0; i < a->num_views; ++i) {
        float4* filt = a->filtered ? reinterpret_cast<float4*>(a->filtered) + (size_t)i * hw : nullptr;
        GMVS_LAUNCH(K_POINTS, st, pts_write, dim3(nblk), dim3(kBlock), 0, st, view(i), k,
                    off + 2 * (size_t)i * nblk, a->points, a->capacity, a->points_filtered,
                    a->capacity_filtered, filt);
        GMVS_LAUNCH_CHECK();
    }
    return GMVS_OK;
}

extern "C" size_t gmvs_format_workspace_bytes(uint64_t n_rows) {
    return fmt_layout(n_rows).total;
}

extern "C" int gmvs_format_rows(const double* rows, uint64_t n_rows, int32_t ncol, int32_t decimals,
                                char* out, uint64_t capacity, uint64_t* out_len, void* workspace,
                                size_t workspace_bytes, void* stream) {
    clear_error();
    GMVS_ARG_CHECK(out_len, "format_rows: out_len is NULL");
    GMVS_ARG_CHECK(ncol >= 1 && ncol <= 64, "format_rows: ncol must be in [1, 64] (got %d)", ncol);
    GMVS_ARG_CHECK(decimals >= 0 && decimals <= 9, "format_rows: decimals must be in [0, 9] (got %d)",
                   decimals);
    GMVS_ARG_CHECK(n_rows == 0 || rows, "format_rows: rows is NULL");
    GMVS_ARG_CHECK(capacity == 0 || out, "format_rows: capacity without out");
    GMVS_ARG_CHECK(n_rows <= (1ull << 40), "format_rows: too many rows");
    const FmtLayout L = fmt_layout(n_rows);
    GMVS_ARG_CHECK(workspace, "format_rows: workspace is NULL");
    if (workspace_bytes < L.total) {
        set_error("format_rows: workspace too small: %zu < %zu bytes", workspace_bytes, L.total);
        return GMVS_ERR_NOSPACE;
    }
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    char* ws = reinterpret_cast<char*>(workspace);
    uint32_t* cnt = reinterpret_cast<uint32_t*>(ws + L.cnt);
    uint64_t* off = reinterpret_cast<uint64_t*>(ws + L.off);
    uint64_t* totals = reinterpret_cast<uint64_t*>(ws + L.totals);
    uint32_t* bad = reinterpret_cast<uint32_t*>(ws + L.bad);
    const uint64_t nblk = (n_rows + kBlock - 1) / kBlock;
    GMVS_LAUNCH(K_FORMAT, st, fill_u32, dim3(1), dim3(64), 0, st, bad, (int64_t)1, 0u);
    GMVS_LAUNCH_CHECK();
    if (nblk) {
        GMVS_LAUNCH(K_FORMAT, st, fmt_count, dim3((unsigned)nblk), dim3(kBlock), 0, st, rows, n_rows, ncol,
                    decimals, cnt, bad);
        GMVS_LAUNCH_CHECK();
    }
    GMVS_LAUNCH(K_FORMAT, st, scan_pairs, dim3(1), dim3(kScanThreads), 0, st, cnt, (int64_t)nblk, off,
                totals);
    GMVS_LAUNCH_CHECK();
    if (nblk && out && capacity) {
        GMVS_LAUNCH(K_FORMAT, st, fmt_write, dim3((unsigned)nblk), dim3(kBlock), 0, st, rows, n_rows, ncol,
                    decimals, off, out, capacity);
        GMVS_LAUNCH_CHECK();
    }
    GMVS_LAUNCH(K_FORMAT, st, fmt_finish, dim3(1), dim3(64), 0, st, totals, bad, out_len);
    GMVS_LAUNCH_CHECK();
    return GMVS_OK;
}

extern "C" int gmvs_sliding_window_normal(const int32_t* planar_region, const float* normal,
                                          int32_t height, int32_t width, int32_t window_size,
                                          float* out, void* stream) {
    clear_error();
    GMVS_ARG_CHECK(planar_region && normal && out, "sliding_window_normal: NULL argument");
    GMVS_ARG_CHECK(height > 0 && width > 0 && (int64_t)height * width <= (1 << 30),
                   "sliding_window_normal: bad size %dx%d", width, height);
    GMVS_ARG_CHECK(window_size >= 1 && window_size <= 4097,
                   "sliding_window_normal: window_size must be in [1, 4097] (got %d)", window_size);
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const int hw = height * width;
    GMVS_LAUNCH(K_SWN, st, swn_kernel, dim3((hw + kBlock - 1) / kBlock), dim3(kBlock), 0, st,
                planar_region, normal, height, width, window_size / 2, out);
    GMVS_LAUNCH_CHECK();
    return GMVS_OK;
}

extern "C" size_t gmvs_segments_workspace_bytes(int32_t height, int32_t width, int32_t num_segments) {
    if (height <= 0 || width <= 0 || num_segments <= 0) return 0;
    return seg_layout(height, width, num_segments).total;
}

static int seg_check(const char* fn, int h, int w, int ns, void* workspace, size_t workspace_bytes,
                     SegLayout* L) {
    GMVS_ARG_CHECK(h > 0 && w > 0 && (int64_t)h * w <= (1 << 30), "%s: bad size %dx%d", fn, w, h);
    GMVS_ARG_CHECK(ns >= 1 && ns <= (1 << 30), "%s: num_segments must be in [1, 2^30] (got %d)", fn, ns);
    *L = seg_layout(h, w, ns);
    GMVS_ARG_CHECK(L->sort_bytes > 0, "%s: radix sort size query failed", fn);
    GMVS_ARG_CHECK(workspace, "%s: workspace is NULL", fn);
    if (workspace_bytes < L->total) {
        set_error("%s: workspace too small: %zu < %zu bytes", fn, workspace_bytes, L->total);
        return GMVS_ERR_NOSPACE;
    }
    return GMVS_OK;
}

extern "C" int gmvs_filter_by_var_map(const float* var_map, const int32_t* segment_ids, int32_t height,
                                      int32_t width, int32_t num_segments, float var_thresh, int32_t* out,
                                      void* workspace, size_t workspace_bytes, void* stream) {
    clear_error();
    GMVS_ARG_CHECK(var_map && segment_ids && out, "filter_by_var_map: NULL argument");
    SegLayout L;
    const int rc = seg_check("filter_by_var_map", height, width, num_segments, workspace, workspace_bytes, &L);
    if (rc != GMVS_OK) return rc;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    char* ws = reinterpret_cast<char*>(workspace);
    const int rc2 = seg_group(segment_ids, height, width, num_segments, ws, L, st);
    if (rc2 != GMVS_OK) return rc2;
    const int hw = height * width;
    int32_t* label = reinterpret_cast<int32_t*>(ws + L.label);
    GMVS_LAUNCH(K_SEG_STATS, st, vm_segments, dim3((num_segments + kBlock - 1) / kBlock), dim3(kBlock), 0, st,
                var_map, reinterpret_cast<const uint32_t*>(ws + L.idx_out),
                reinterpret_cast<const uint32_t*>(ws + L.start), reinterpret_cast<const uint32_t*>(ws + L.end),
                num_segments, var_thresh, label);
    GMVS_LAUNCH_CHECK();
    GMVS_LAUNCH(K_SEG_STATS, st, seg_paint, dim3((hw + kBlock - 1) / kBlock), dim3(kBlock), 0, st,
                segment_ids, hw, num_segments, label, out);
    GMVS_LAUNCH_CHECK();
    return GMVS_OK;
}

extern "C" int gmvs_fusion_planar_mask(const int32_t* seg_ids, const int32_t* planar_mask,
                                       const float* filter_normals, int32_t rows, int32_t cols,
                                       int32_t num_segments, int32_t thresh, float nonplanar_percent,
                                       float cos_sim_thresh, float match_ratio_thresh, int32_t* out,
                                       void* workspace, size_t workspace_bytes, void* stream) {
    clear_error();
    GMVS_ARG_CHECK(seg_ids && planar_mask && filter_normals && out, "fusion_planar_mask: NULL argument");
    SegLayout L;
    const int rc = seg_check("fusion_planar_mask", rows, cols, num_segments, workspace, workspace_bytes, &L);
    if (rc != GMVS_OK) return rc;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    char* ws = reinterpret_cast<char*>(workspace);
    const int rc2 = seg_group(seg_ids, rows, cols, num_segments, ws, L, st);
    if (rc2 != GMVS_OK) return rc2;
    const int hw = rows * cols;
    int32_t* label = reinterpret_cast<int32_t*>(ws + L.label);
    GMVS_LAUNCH(K_SEG_STATS, st, pm_segments, dim3((num_segments + kBlock - 1) / kBlock), dim3(kBlock), 0, st,
                planar_mask, filter_normals, reinterpret_cast<const uint32_t*>(ws + L.idx_out),
                reinterpret_cast<const uint32_t*>(ws + L.start), reinterpret_cast<const uint32_t*>(ws + L.end),
                num_segments, thresh, nonplanar_percent, cos_sim_thresh, match_ratio_thresh, label);
    GMVS_LAUNCH_CHECK();
    GMVS_LAUNCH(K_SEG_STATS, st, seg_vps_ids, dim3(1), dim3(kScanThreads), 0, st, label, num_segments);
    GMVS_LAUNCH_CHECK();
    GMVS_LAUNCH(K_SEG_STATS, st, seg_paint, dim3((hw + kBlock - 1) / kBlock), dim3(kBlock), 0, st,
                seg_ids, hw, num_segments, label, out);
    GMVS_LAUNCH_CHECK();
    return GMVS_OK;
}
