// PatchMatch MVS kernels for gfx950 and their C-ABI drivers.
//
// One half-sweep of the red/black checkerboard (black_red_pixel_update,
// gmvs/src/patch_match.cu:1085-1142; Black/RedPixelUpdate,
// gmvs/lib/src/patch_match_kernel.cu:1579-1653) runs as three kernels:
//   pm_candidates -- adaptive candidate search over the state as it was at
//                    the start of the half-sweep, plus the 8 x V NCC cost
//                    table the view selection needs and the V NCCs of the
//                    pixel's current hypothesis. It writes nothing to the
//                    state, only a per-pixel record (costs, the candidate
//                    planes, the flags) into the workspace;
//   pm_select     -- joint view selection and candidate adoption (plain or
//                    planar-prior scoring, optional geometric term), one
//                    thread per pixel; writes the pixel's adopted state and
//                    its 5 random refinement hypotheses;
//   pm_refine     -- the refinement hypotheses costed over the selected views
//                    only, as load-balanced work units, and adopted.
// Each thread writes only its own pixel, and neighbour planes are read only
// from the candidate records, so every read of the half-sweep sees the
// snapshot (the reference reads same-colour pixels that other threads are
// rewriting, lib 978-1104).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cstring>
#include <mutex>
#include <vector>

#include "abi_common.h"
#include "pm_device.h"

namespace gmvs {

constexpr int kBlock = 256;  // threads per block
constexpr int kXcds = 8;     // MI355X: 8 XCDs, workgroups dispatched round-robin

// Lane groups of L lanes evaluate one pixel's NCCs together: the block of an
// NCC kernel (64 * L threads) holds a tile of 64 pixels, TX x TY with TX = 1 << shift
// (TX counts active columns in the half-sweep kernels and pixel columns in
// pm_init); pm_candidates' blocks hold two horizontally adjacent tiles. The
// shape is chosen per problem (gmvs_pm_args.tile_cols).
struct TileShape {
    int shift;  // log2 of the tile's columns
    __host__ __device__ int tx() const { return 1 << shift; }
    __host__ __device__ int ty() const { return 64 >> shift; }
    __host__ __device__ int tiles(int cols, int rows) const {
        return ((cols + tx() - 1) >> shift) * ((rows + ty() - 1) / ty());
    }
};

// Pixels per NCC tile (one lane group each) and the block of an NCC kernel.
constexpr int kTilePix = 64;
template <int L>
constexpr int ncc_block() { return kTilePix * L; }
// NCC lane-group widths: 2 lanes (18 taps each) or 4 lanes (9 taps each),
// chosen per kernel at run time (gmvs_set_lane_width; defaults below).
// Narrower groups spend fewer instructions per NCC on the per-hypothesis
// set-up, the corner test and the epilogue, which every lane of a group
// repeats; wider groups keep fewer taps per lane and more lanes per pixel,
// which suits the latency-bound refinement kernel (scattered random-depth
// footprints). Measured on the 10-view 1600x1200 ACMP step, all three
// precisions: candidates and init fastest with 2 lanes (-10% / -12% against
// 4), refinement with 4 (-3% to -10% against 2).
inline int default_lanes(int kid, int precision) {
    (void)precision;
    return kid == K_REFINE ? 4 : 2;
}
constexpr int kNT = 6;  // taps per axis: patch 11, increment 2 (the reference defaults)

struct PMConst {
    RefCam rc;
    const float* ref;             // reference intensities [H*W]
    const ViewDesc* views;        // [nv]
    const ViewF64* views64;       // [nv], the homography factors in FP64
    const void* cells;            // packed cells of all views, [nv][(H+kPad)*(W+kPad)]
    const float* init_depth;      // INIT_DEPTH_NORMAL
    const float* init_normal;     // INIT_DEPTH_NORMAL, world frame
    const float4* prior_planes;   // planar prior
    const uint32_t* plane_masks;  // planar prior
    int nv, h, w, wh;             // wh = active columns per row
    int wp, vcells;               // cell row pitch w + kPad; cells per view (h + kPad) * wp
    int fast_taps;                // clamp-free tap path allowed (SrcGrid::fast)
    int radius, inc;
    float sigma_s, sigma_c;
    float dmin, dmax;
    int top_k;
    uint32_t seed, view;
    int init_mode, geom, prior;
    uint32_t compat;              // GMVS_COMPAT_* (0: the reference's quirks)
    TileShape tile;
    __device__ SrcGrid grid() const { return SrcGrid{h, w, wp, fast_taps}; }
};

// Adds per-lane work counts (NCCs, geometric terms) to a kernel's counters:
// one reduction across the wave and one atomic per wave. Every lane of the
// wave must be active.
__device__ inline void count_work(unsigned long long* ctr, uint32_t ncc, uint32_t geom) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        ncc += __shfl_xor(ncc, off);
        geom += __shfl_xor(geom, off);
    }
    if ((threadIdx.x & 63) == 0) {
        if (ncc) atomicAdd(ctr, (unsigned long long)ncc);
        if (geom) atomicAdd(ctr + 1, (unsigned long long)geom);
    }
}

// The local frame (see Frame) of one tile of reference pixels for view v:
// reference positions relative to the tile origin (ox, oy), source positions
// relative to the image of the tile's centre under the fronto-parallel plane
// at the middle inverse depth. tile_frames writes the frames of every view of
// the problem, one thread per view; the block synchronises before the frames
// are read.
__device__ inline void tile_frame(const PMConst& P, int v, int ox, int oy, int cols, int rows,
                                  Frame* dst) {
    double am[3];
    fronto_m(0.5 * (1.0 / (double)P.dmin + 1.0 / (double)P.dmax), am);
    *dst = make_frame(P.views64[v], ox, oy, ox + 0.5 * cols, oy + 0.5 * rows, am, P.w, P.h, P.wp);
}

__device__ inline void tile_frames(const PMConst& P, int ox, int oy, int cols, int rows,
                                   Frame* s_fr) {
    if ((int)threadIdx.x < P.nv)
        tile_frame(P, (int)threadIdx.x, ox, oy, cols, rows, &s_fr[threadIdx.x]);
}

// ------------------------------------------------------------- batching ---
// Every pass runs as a batch of problems (reference views, possibly of
// different scenes and sizes): each kernel of the schedule is launched once
// for the whole batch over a table of work items (tiles or pixel chunks),
// ordered problem by problem.
constexpr int kMaxBatch = 64;

// One problem of a batch: its constants, state and workspace buffers.
struct PMProblem {
    PMConst P;
    float4* planes;         // [H*W] state
    float* costs;           // [H*W] state
    uint32_t* sel;          // [H*W] selected-view masks
    float* ca;              // [9][nv][na] candidate NCC table
    float4* cplane;         // [8][na] candidate planes
    uint8_t* cflag;         // [na] candidate flags
    float4* rtp;            // [6][na] refinement hypotheses, then the adopted
                            // plane (fused selection only)
    uint32_t* rsw;          // [SW_NUM][na] selection record
    float* out_depth;       // [H*W] or nullptr
    float* out_normal;      // [H*W*3] or nullptr
    float* out_cost;        // [H*W] or nullptr
    int na;                 // active pixels of one colour
    int skip_filter;
};

// Work items of one tiling: problem p owns items [start[p], start[p + 1]).
struct WorkTable {
    int n;
    int start[kMaxBatch + 1];
};
enum { WT_FULL = 0,  // 8x8-pixel tiles over the image (pm_init)
       WT_HALF,      // 8x8 tiles over one colour's active pixels (candidates, refine)
       WT_ACTIVE,    // 256-pixel chunks of active pixels (select, filter)
       WT_PIXELS,    // 256-pixel chunks of the image (depth/normal, export)
       WT_HALF_PAIRS,  // pairs of horizontally adjacent WT_HALF tiles (candidates)
       WT_NUM };

__device__ inline int problem_of(const WorkTable& t, int item) {
    int lo = 0, hi = t.n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (t.start[mid] <= item) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Persistent work loop of the heavy kernels: the item space is cut into 8
// contiguous chunks, one queue per XCD (workgroups are dispatched to the
// XCDs round-robin, so block b runs on XCD b % 8); a block drains its own
// XCD's chunk first, so neighbouring tiles (which gather from the same
// source-image regions) share an L2, then takes what is left in the others.
// Returns -1 once every queue is exhausted; every block reaches it. The 8
// counters of a launch start at zero.
__device__ inline int next_item(uint32_t* __restrict__ ctr, int total, int& q, int& tries,
                                int& s_item) {
    __syncthreads();  // the previous item is finished by every thread
    if (threadIdx.x == 0) {
        const int chunk = (total + kXcds - 1) / kXcds;
        int it = -1;
        while (tries < kXcds) {
            const int lo = q * chunk;
            const int hi = min(total, lo + chunk);
            if (lo < hi) {
                const int k = (int)atomicAdd(&ctr[q], 1u);
                if (lo + k < hi) {
                    it = lo + k;
                    break;
                }
            }
            q = (q + 1) % kXcds;
            ++tries;
        }
        s_item = it;
    }
    __syncthreads();
    // block-uniform; readfirstlane tells the compiler so, which keeps the
    // problem record and the buffer descriptors derived from it in SGPRs
    // (a value loaded from LDS is otherwise treated as divergent)
    return __builtin_amdgcn_readfirstlane(s_item);
}

__global__ void pm_zero_words(uint32_t* __restrict__ p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// Writes one problem record and (for problem 0) the work tables; the
// records travel as kernel arguments, so the setup is capturable.
struct WorkTables {
    WorkTable t[WT_NUM];
};
__global__ void pm_setup_problem(PMProblem pr, PMProblem* __restrict__ dst) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *dst = pr;
}
__global__ void pm_setup_tables(WorkTables t, WorkTables* __restrict__ dst) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *dst = t;
}

// ------------------------------------------------------------ setup ---
// FP64 per-view composition of the homography factors:
// A = Ks R_rel Kr^-1, b = Ks t_rel with R_rel = Rs Rr^T,
// t_rel = Rs (C_r - C_s), C = -R^T t (compute_homography, src 211-296; as
// there, Kr^-1 uses fx, fy, cx, cy and Ks contributes K0, K2, K4, K5, K8).
// One single-thread launch per source view; the cameras travel as kernel
// arguments.
__global__ void pm_setup_view(gmvs_camera r, gmvs_camera s, const void* cells,
                              const float* depth, ViewDesc* __restrict__ out,
                              ViewF64* __restrict__ out64) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double Cr[3], Cs[3];
    for (int i = 0; i < 3; ++i) {
        Cr[i] = -((double)r.R[i] * r.t[0] + (double)r.R[3 + i] * r.t[1] + (double)r.R[6 + i] * r.t[2]);
        Cs[i] = -((double)s.R[i] * s.t[0] + (double)s.R[3 + i] * s.t[1] + (double)s.R[6 + i] * s.t[2]);
    }
    double Rrel[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            Rrel[3 * i + j] = (double)s.R[3 * i] * r.R[3 * j] + (double)s.R[3 * i + 1] * r.R[3 * j + 1] +
                              (double)s.R[3 * i + 2] * r.R[3 * j + 2];
    double trel[3];
    for (int i = 0; i < 3; ++i)
        trel[i] = (double)s.R[3 * i] * (Cr[0] - Cs[0]) + (double)s.R[3 * i + 1] * (Cr[1] - Cs[1]) +
                  (double)s.R[3 * i + 2] * (Cr[2] - Cs[2]);
    const double fx = r.K[0], fy = r.K[4], cx = r.K[2], cy = r.K[5];
    const double Ki[9] = {1.0 / fx, 0.0, -cx / fx, 0.0, 1.0 / fy, -cy / fy, 0.0, 0.0, 1.0};
    const double Ks[9] = {s.K[0], 0.0, s.K[2], 0.0, s.K[4], s.K[5], 0.0, 0.0, s.K[8]};
    double M[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            M[3 * i + j] = Rrel[3 * i] * Ki[j] + Rrel[3 * i + 1] * Ki[3 + j] + Rrel[3 * i + 2] * Ki[6 + j];
    ViewDesc d;
    ViewF64 d64;
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) {
            d64.A[3 * i + j] = Ks[3 * i] * M[j] + Ks[3 * i + 1] * M[3 + j] + Ks[3 * i + 2] * M[6 + j];
            d.A[3 * i + j] = (float)d64.A[3 * i + j];
        }
        d64.b[i] = Ks[3 * i] * trel[0] + Ks[3 * i + 1] * trel[1] + Ks[3 * i + 2] * trel[2];
        d.b[i] = (float)d64.b[i];
    }
    *out64 = d64;
    for (int i = 0; i < 9; ++i) {
        d.K[i] = s.K[i];
        d.R[i] = s.R[i];
    }
    for (int i = 0; i < 3; ++i) {
        d.t[i] = s.t[i];
        d.C[i] = -(s.R[i] * s.t[0] + s.R[3 + i] * s.t[1] + s.R[6 + i] * s.t[2]);
    }
    d.cells = cells;
    d.depth = depth;
    *out = d;
}

// Packs an f32 image into bilinear cells (v00, v01-v00, v10-v00,
// v11-v10-v01+v00), with the right/bottom neighbours clamped to the last
// column/row as the reference's software bilinear does (src 20-21). The
// differences are formed in FP32, then stored in the cell precision
// (round to nearest even). The grid is (h + kPad) x (w + kPad), row pitch
// w + kPad; pad cells copy the nearest edge cell (see SrcGrid).
template <int PR>
__global__ void pm_pack_quads(const float* __restrict__ img, void* __restrict__ cells, int h,
                              int w) {
    const int wp = w + kPad;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (h + kPad) * wp) return;
    const int yp = i / wp, xp = i - yp * wp;
    const int y = min(yp, h - 1), x = min(xp, w - 1);
    const int xu = (x == w - 1) ? x : x + 1;
    const int yu = (y == h - 1) ? y : y + 1;
    const float v0 = img[y * w + x], v1 = img[y * w + xu];
    const float v2 = img[yu * w + x], v3 = img[yu * w + xu];
    const float d01 = v1 - v0, d10 = v2 - v0, d11 = (v3 - v2) - (v1 - v0);
    if constexpr (PR == PREC_F32) {
        reinterpret_cast<float4*>(cells)[i] = make_float4(v0, d01, d10, d11);
    } else if constexpr (PR == PREC_F16) {
        const f16x2 lo = {(_Float16)v0, (_Float16)d01}, hi = {(_Float16)d10, (_Float16)d11};
        reinterpret_cast<uint2*>(cells)[i] =
            make_uint2(__builtin_bit_cast(uint32_t, lo), __builtin_bit_cast(uint32_t, hi));
    } else {
        const bf16x2 lo = __builtin_convertvector((f32x2){v0, d01}, bf16x2);
        const bf16x2 hi = __builtin_convertvector((f32x2){d10, d11}, bf16x2);
        reinterpret_cast<uint2*>(cells)[i] =
            make_uint2(__builtin_bit_cast(uint32_t, lo), __builtin_bit_cast(uint32_t, hi));
    }
}

// ---------------------------------------------------------------- init ---
template <int NT, int L, int PR>
__device__ inline void pm_init_body(const PMConst& P, int tile, float4* __restrict__ planes,
                                    float* __restrict__ costs, uint32_t* __restrict__ sel,
                                    float (*s_c)[kTilePix * L], const Frame* s_fr, int ox, int oy,
                                    uint32_t& ncc) {
    const int TX = P.tile.tx(), TY = P.tile.ty();
    const int tx = (P.w + TX - 1) / TX;
    const int slot = threadIdx.x / L, sub = threadIdx.x % L;
    const int px = (tile % tx) * TX + (slot % TX);
    const int py = (tile / tx) * TY + (slot / TX);
    if (px >= P.w || py >= P.h) return;
    const int c = py * P.w + px;
    const RefCam& rc = P.rc;
    const float fpx = (float)px, fpy = (float)py;
    Rng rng{(uint32_t)c, RNG_PHASE_INIT, P.seed, P.view};

    float4 pl;
    if (P.init_mode == GMVS_INIT_DEPTH_NORMAL) {
        const float n0 = P.init_normal[3 * c], n1 = P.init_normal[3 * c + 1], n2 = P.init_normal[3 * c + 2];
        pl.x = rc.R[0] * n0 + rc.R[1] * n1 + rc.R[2] * n2;
        pl.y = rc.R[3] * n0 + rc.R[4] * n1 + rc.R[5] * n2;
        pl.z = rc.R[6] * n0 + rc.R[7] * n1 + rc.R[8] * n2;
        pl.w = dist_to_origin(rc, fpx, fpy, P.init_depth[c], pl);
    } else if (P.init_mode == GMVS_INIT_PRIOR) {
        // Reuse the previous plane (its .w holds a depth), then overwrite it
        // with a perturbed prior plane for masked, costly pixels. Written as
        // straight-line code followed by one branch: the nested short-circuit
        // form (prior && mask && cost) was miscompiled by the ROCm 7.2 gfx950
        // backend (the reuse path read registers clobbered by the mask load).
        const float4 st = planes[c];
        pl = st;
        if (P.compat & GMVS_COMPAT_FIX_PRIOR_NORMALS) {
            // the first pass exported world-frame normals (GetDepthandNormal)
            pl.x = rc.R[0] * st.x + rc.R[1] * st.y + rc.R[2] * st.z;
            pl.y = rc.R[3] * st.x + rc.R[4] * st.y + rc.R[5] * st.z;
            pl.z = rc.R[6] * st.x + rc.R[7] * st.y + rc.R[8] * st.z;
        }
        pl.w = dist_to_origin(rc, fpx, fpy, st.w, pl);
        bool perturb = false;
        if (P.prior) {
            const uint32_t mval = P.plane_masks[c];
            const float cval = costs[c];
            perturb = (mval > 0u) & (cval >= 0.1f);
        }
        if (perturb) {
            const float4 pp = P.prior_planes[c];
            const float pert = 0.02f;
            const float dlo = (1.0f - 3.0f * pert) * pp.w;
            const float dhi = (1.0f + 3.0f * pert) * pp.w;
            uint32_t o[4];
            rng.block(0, o);
            const float dp = u01(o[RNG_INIT_DEPTH]) * (dhi - dlo) + dlo;
            pl = perturbed_normal(rc, fpx, fpy, pp, u01(o[1]), u01(o[2]), u01(o[3]),
                                  3.0f * pert * kPiF);
            pl.w = dp;
        }
    } else {
        const float depth = rng.u(RNG_INIT_DEPTH) * (P.dmax - P.dmin) + P.dmin;
        pl = random_normal(rc, fpx, fpy, rng, RNG_INIT_NORMAL, depth);
        pl.w = dist_to_origin(rc, fpx, fpy, depth, pl);
    }

    Hoist<NT, L> hs;
    hoist_ref<NT, L>(hs, P.ref, px, py, P.h, P.w, P.radius, P.inc, P.sigma_s, P.sigma_c, sub, ox,
                     oy);
    float m[3];
    plane_m_local(rc, pl, (float)ox, (float)oy, m);
    int valid = 0;
    for (int v = 0; v < P.nv; ++v) {
        const ViewDesc& vd = P.views[v];
        const Frame& fr = s_fr[v];
        float H[9];
        homography(fr.A, fr.b, m, H);
        const float cv =
            ncc_eval<NT, L, PR>(hs, H, tap_src<PR>(vd.cells, P.vcells, fr), P.grid());
        s_c[v][threadIdx.x] = cv;
        valid += (cv < 2.0f);
    }
    // Top-k mean of the valid costs and the view mask of costs <= the k-th
    // smallest (compute_multi_view_initial_cost_and_selected_views,
    // src 447-501).
    const int k = min(valid, P.top_k);
    uint32_t bits = 0;
    float cost = 2.0f;
    if (k > 0) {
        uint32_t taken = 0;
        float sum = 0.f, thr = 0.f;
        for (int r = 0; r < k; ++r) {
            float best = 0.f;
            int bi = -1;
            for (int v = 0; v < P.nv; ++v) {
                if ((taken >> v) & 1u) continue;
                const float x = s_c[v][threadIdx.x];
                if (bi < 0 || x < best) {
                    best = x;
                    bi = v;
                }
            }
            taken |= 1u << bi;
            sum += best;
            thr = best;
        }
        for (int v = 0; v < P.nv; ++v)
            if (s_c[v][threadIdx.x] <= thr) bits |= 1u << v;
        cost = sum / (float)k;
    }
    if (sub == 0) {
        planes[c] = pl;
        costs[c] = cost;
        sel[c] = bits;
        ncc += (uint32_t)P.nv;
    }
}

template <int NT, int L, int PR>
__global__ __launch_bounds__(kTilePix * L) void pm_init(const PMProblem* __restrict__ probs,
                                                  const WorkTables* __restrict__ tabs,
                                                  uint32_t* __restrict__ queue,
                                                  unsigned long long* work) {
    __shared__ float s_c[kMaxViews][kTilePix * L];
    __shared__ Frame s_fr[kMaxViews];
    __shared__ int s_item;
    const WorkTable& t = tabs->t[WT_FULL];
    const int total = t.start[t.n];
    uint32_t ncc = 0;
    int q = blockIdx.x % kXcds, tries = 0;
    for (;;) {
        const int item = next_item(queue, total, q, tries, s_item);
        if (item < 0) break;
        const int p = problem_of(t, item);
        const PMProblem& pr = probs[p];
        const int tile = item - t.start[p];
        const int TX = pr.P.tile.tx(), TY = pr.P.tile.ty();
        const int ntx = (pr.P.w + TX - 1) / TX;
        const int ox = (tile % ntx) * TX, oy = (tile / ntx) * TY;
        tile_frames(pr.P, ox, oy, TX, TY, s_fr);  // next_item's barrier retired the readers
        __syncthreads();
        pm_init_body<NT, L, PR>(pr.P, tile, pr.planes, pr.costs, pr.sel, s_c, s_fr, ox, oy, ncc);
    }
    if (work) count_work(work + 2 * K_INIT, ncc, 0);
}

// ------------------------------------------------------- half-sweep, A ---
struct Active {
    int px, py, aidx, sub;
    bool ok;
};

// Pixel `slot` of a tile of one colour's active pixels (slot = lane group in
// the NCC layout; pm_refine and the fused selection epilogue).
__device__ inline Active active_slot(const PMConst& P, int tile, int color, int slot) {
    const int TX = P.tile.tx(), TY = P.tile.ty();
    const int tx = (P.wh + TX - 1) / TX;
    const int ax = (tile % tx) * TX + (slot & (TX - 1));
    const int y = (tile / tx) * TY + (slot >> P.tile.shift);
    Active a;
    a.sub = 0;
    a.py = y;
    a.px = 2 * ax + ((y + color) & 1);
    a.aidx = y * P.wh + ax;
    a.ok = (y < P.h) && (ax < P.wh) && (a.px < P.w);
    return a;
}

// Pixel of thread lt (of the tile's 64 * L threads) in an NCC kernel's tile.
template <int L>
__device__ inline Active active_pixel(const PMConst& P, int tile, int color, int lt) {
    const int TX = P.tile.tx(), TY = P.tile.ty();
    const int tx = (P.wh + TX - 1) / TX;
    const int slot = lt / L;
    const int ax = (tile % tx) * TX + (slot & (TX - 1));
    const int y = (tile / tx) * TY + (slot >> P.tile.shift);
    Active a;
    a.sub = lt % L;
    a.py = y;
    a.px = 2 * ax + ((y + color) & 1);
    a.aidx = y * P.wh + ax;
    a.ok = (y < P.h) && (ax < P.wh) && (a.px < P.w);
    return a;
}

// Adaptive checkerboard candidate search (checkerboard_propagation,
// src 720-913). Order: 0 up_near, 1 up_far, 2 down_near, 3 down_far,
// 4 left_near, 5 left_far, 6 right_near, 7 right_far.
__device__ inline void candidate_search(const float* __restrict__ cs, int px, int py, int h, int w,
                                        bool right_far_min, uint32_t& flags, int pos[8]) {
    const int c = py * w + px;
    flags = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) pos[k] = c;
    if (py > 2) {
        flags |= 1u << 1;
        int best = c - 3 * w;
        float m = cs[best];
        for (int i = 1; i < 11; ++i)
            if (py > 2 + 2 * i) {
                const int q = c - 3 * w - 2 * i * w;
                const float x = cs[q];
                if (x < m) {
                    m = x;
                    best = q;
                }
            }
        pos[1] = best;
    }
    if (py < h - 3) {
        flags |= 1u << 3;
        int best = c + 3 * w;
        float m = cs[best];
        for (int i = 1; i < 11; ++i)
            if (py < h - 3 - 2 * i) {
                const int q = c + 3 * w + 2 * i * w;
                const float x = cs[q];
                if (x < m) {
                    m = x;
                    best = q;
                }
            }
        pos[3] = best;
    }
    if (px > 2) {
        flags |= 1u << 5;
        int best = c - 3;
        float m = cs[best];
        for (int i = 1; i < 11; ++i)
            if (px > 2 + 2 * i) {
                const int q = c - 3 - 2 * i;
                const float x = cs[q];
                if (x < m) {
                    m = x;
                    best = q;
                }
            }
        pos[5] = best;
    }
    if (px < w - 3) {
        flags |= 1u << 7;
        int best = c + 3;
        float m = cs[best];
        for (int i = 1; i < 11; ++i)
            if (px < w - 3 - 2 * i) {
                const int q = c + 3 + 2 * i;
                const float x = cs[q];
                // the reference keeps the maximum here (src 794) unless
                // GMVS_COMPAT_FIX_RIGHT_FAR
                if (right_far_min ? (x < m) : (m < x)) {
                    m = x;
                    best = q;
                }
            }
        pos[7] = best;
    }
    if (py > 0) {
        flags |= 1u << 0;
        const int base = c - w;
        int best = base;
        float m = cs[best];
        for (int i = 0; i < 3; ++i) {
            if (py > 1 + i && px > i) {
                const int q = base - (1 + i) * w - i;
                const float x = cs[q];
                if (x < m) {
                    m = x;
                    best = q;
                }
            }
            if (py > 1 + i && px < w - 1 - i) {
                const int q = base - (1 + i) * w + i;
                const float x = cs[q];
                if (x < m) {
                    m = x;
                    best = q;
                }
            }
        }
        pos[0] = best;
    }
    if (py < h - 1) {
        flags |= 1u << 2;
        const int base = c + w;
        int best = base;
        float m = cs[best];
        for (int i = 0; i < 3; ++i) {
            if (py < h - 2 - i && px > i) {
                const int q = base + (1 + i) * w - i;
                const float x = cs[q];
                if (x < m) {
                    m = x;
                    best = q;
                }
            }
            if (py < h - 2 - i && px < w - 1 - i) {
                const int q = base + (1 + i) * w + i;
                const float x = cs[q];
                if (x < m) {
                    m = x;
                    best = q;
                }
            }
        }
        pos[2] = best;
    }
    if (px > 0) {
        flags |= 1u << 4;
        const int base = c - 1;
        int best = base;
        float m = cs[best];
        for (int i = 0; i < 3; ++i) {
            if (px > 1 + i && py > i) {
                const int q = base - (1 + i) - i * w;
                const float x = cs[q];
                if (x < m) {
                    m = x;
                    best = q;
                }
            }
            if (px > 1 + i && py < h - 1 - i) {
                const int q = base - (1 + i) + i * w;
                const float x = cs[q];
                if (x < m) {
                    m = x;
                    best = q;
                }
            }
        }
        pos[4] = best;
    }
    if (px < w - 1) {
        flags |= 1u << 6;
        const int base = c + 1;
        int best = base;
        float m = cs[best];
        for (int i = 0; i < 3; ++i) {
            if (px < w - 2 - i && py > i) {
                const int q = base + (1 + i) - i * w;
                const float x = cs[q];
                if (x < m) {
                    m = x;
                    best = q;
                }
            }
            if (px < w - 2 - i && py < h - 1 - i) {
                const int q = base + (1 + i) + i * w;
                const float x = cs[q];
                if (x < m) {
                    m = x;
                    best = q;
                }
            }
        }
        pos[6] = best;
    }
}

template <int NT, int L, int PR>
__device__ inline void pm_candidates_body(const PMConst& P, int tile, int color,
                                          const float4* __restrict__ planes,
                                          const float* __restrict__ costs, float* __restrict__ ca,
                                          float4* __restrict__ cplane, uint8_t* __restrict__ cflag,
                                          int NA, float (*s_m)[3][kTilePix * L],
                                          const Frame* s_fr, int ox, int oy, uint32_t& ncc,
                                          int lt) {
    const Active a = active_pixel<L>(P, tile, color, lt);
    if (!a.ok) return;
    const bool lead = a.sub == 0;
    uint32_t flags;
    int pos[9];
    candidate_search(costs, a.px, a.py, P.h, P.w, (P.compat & GMVS_COMPAT_FIX_RIGHT_FAR) != 0u, flags,
                     pos);
    if (lead) {
        cflag[a.aidx] = (uint8_t)flags;
        ncc += (uint32_t)((__popc(flags) + 1) * P.nv);
    }
    Hoist<NT, L> hs;
    hoist_ref<NT, L>(hs, P.ref, a.px, a.py, P.h, P.w, P.radius, P.inc, P.sigma_s, P.sigma_c, a.sub,
                     ox, oy);
    // hypotheses 0-7: the flagged candidates; 8: the pixel's own current
    // plane against every view (the selection weights the selected ones)
    pos[8] = a.py * P.w + a.px;
    const bool cost_default_2 = (P.compat & GMVS_COMPAT_FIX_COST_DEFAULT) != 0u;
    const uint32_t todo = flags | (1u << 8);
    // each lane keeps the rank-1 factors m of its pixel's hypotheses in its
    // own LDS column, so the view loop reads 3 floats per NCC
#pragma unroll
    for (int k = 0; k < 9; ++k)
        if ((todo >> k) & 1u) {
            const float4 pl = planes[pos[k]];
            if (lead && k < 8) cplane[(size_t)k * NA + a.aidx] = pl;
            float m[3];
            plane_m_local(P.rc, pl, (float)ox, (float)oy, m);
            s_m[k][0][lt] = m[0];
            s_m[k][1][lt] = m[1];
            s_m[k][2][lt] = m[2];
        }
    // views outermost: the consecutive NCCs of a view gather from one source
    // image region (the hypotheses are neighbours' planes), which keeps the
    // wave's gathers in L1
    for (int v = 0; v < P.nv; ++v) {
        const ViewDesc& vd = P.views[v];
        const Frame& fr = s_fr[v];  // LDS broadcast reads
        const TapSrc src = tap_src<PR>(vd.cells, P.vcells, fr);
        // a wave-uniform hypothesis index (pixels without hypothesis k sit
        // out that iteration): LDS and cost-table offsets stay scalar
#pragma unroll 1
        for (int k = 0; k < 9; ++k) {
            if (!((todo >> k) & 1u)) {
                // an unflagged candidate keeps the reference's cost_array
                // default (lib 854, src 717): every entry but [0][0] is zero
                // (2 everywhere with GMVS_COMPAT_FIX_COST_DEFAULT)
                if (lead)
                    ca[((size_t)k * P.nv + v) * NA + a.aidx] =
                        (cost_default_2 || (k == 0 && v == 0)) ? 2.0f : 0.0f;
                continue;
            }
            const float m[3] = {s_m[k][0][lt], s_m[k][1][lt], s_m[k][2][lt]};
            float H[9];
            homography(fr.A, fr.b, m, H);
            const float cv = ncc_eval_mlp<NT, L, PR>(hs, H, src, P.grid());
            if (lead) ca[((size_t)k * P.nv + v) * NA + a.aidx] = cv;
        }
    }
}


// ------------------------------------------------------- half-sweep, B ---
// The second half of a half-sweep runs as two kernels:
//   pm_select  -- one thread per active pixel: joint view selection
//                 (src 915-981), candidate adoption with plain or
//                 planar-prior scoring and the optional geometric term
//                 (src 983-1040, lib 1185-1334); writes the adopted state
//                 and the 5 refinement hypotheses (src 595-676,
//                 lib 684-811) with the view weights to the workspace;
//   pm_refine  -- one 8x8 tile of active pixels per block: the refinement
//                 NCCs as work units (pixel, selected view), each costing the
//                 5 hypotheses against one view, dealt round-robin over the
//                 block's 64 lane groups so that trip counts differ by at most
//                 one whatever the per-pixel view counts are; then the
//                 reference's sequential adoption in view order.
constexpr int kK2Pix = kTilePix;  // active pixels per pm_refine block
constexpr int kRefine = 5;               // refinement hypotheses per pixel

__device__ inline int argmin_last8(const float* c) {
    float m = c[0];
    int k = 0;
#pragma unroll
    for (int i = 1; i < 8; ++i)
        if (c[i] <= m) {
            m = c[i];
            k = i;
        }
    return k;
}

__device__ inline int argmax_last8(const float* c) {
    float m = c[0];
    int k = 0;
#pragma unroll
    for (int i = 1; i < 8; ++i)
        if (c[i] >= m) {
            m = c[i];
            k = i;
        }
    return k;
}

// Planar-prior score exp(-c^2/beta) * (gamma + exp(-dd^2/2sd^2) exp(-da^2/2sa^2))
// (lib 1256-1277, 702-707, 784-793).
__device__ inline float prior_score(const PMConst& P, float cost, float depth, float dprior,
                                    float4 pp, float4 pl) {
    const float dsig = (P.dmax - P.dmin) / 64.0f;
    const float two_ds2 = 2.0f * dsig * dsig;
    const float asig = kPiF * (5.0f / 180.0f);
    const float two_as2 = 2.0f * asig * asig;
    const float dd = depth - dprior;
    // unit-normal dot clamped into [-1, 1] (NaN stays NaN), so identical
    // normals give angle 0 rather than acos's NaN just above 1
    float dot = pp.x * pl.x + pp.y * pl.y + pp.z * pl.z;
    dot = dot > 1.0f ? 1.0f : (dot < -1.0f ? -1.0f : dot);
    const float ang = acosf(dot);
    const float pr = 0.5f + fexp(-dd * dd / two_ds2) * fexp(-ang * ang / two_as2);
    return fexp(-cost * cost / 0.18f) * pr;
}

// Cost of candidate k against view v. pm_candidates writes every row: the
// NCC of a flagged candidate, and for an unflagged one the reference's
// cost_array default (cost_array[8][32] = {2.0f} leaves every entry but
// [0][0] at zero, lib 854, src 717), so the read needs no branch and several
// can be in flight.
__device__ inline float ca_entry(const float* __restrict__ ca, int k, int v, int nv, int NA,
                                 int aidx) {
    return ca[((size_t)k * nv + v) * NA + aidx];
}

// Per-pixel record handed from pm_select to pm_refine (workspace, [word][NA]).
enum SelWord {
    SW_SELMASK = 0,  // selected views
    SW_WNIB0,        // 4-bit draw counts, views 0-7 .. 24-31
    SW_WNIB1,
    SW_WNIB2,
    SW_WNIB3,
    SW_RESTRICTED,   // prior score of the adopted hypothesis (float bits)
    SW_DRAND,        // refinement depths (float bits)
    SW_DPERT,
    SW_DCUR,
    SW_COST,         // the adopted cost (float bits), fused selection only
    SW_NUM
};

__device__ inline float nib_weight(const uint32_t wnib[4], int v) {
    const uint32_t wd = (v >> 3) == 0 ? wnib[0] : (v >> 3) == 1 ? wnib[1] : (v >> 3) == 2 ? wnib[2] : wnib[3];
    return (float)((wd >> (4u * (uint32_t)(v & 7))) & 15u);
}

// The active pixel with index aidx (row-major over the half-sweep's columns).
__device__ inline Active active_of_index(const PMConst& P, int color, int aidx) {
    Active a;
    a.sub = 0;
    a.aidx = aidx;
    a.py = aidx / P.wh;
    const int ax = aidx - a.py * P.wh;
    a.px = 2 * ax + ((a.py + color) & 1);
    a.ok = (a.py < P.h) && (a.px < P.w);
    return a;
}

// Joint view selection (src 915-981), candidate adoption with plain or
// planar-prior scoring and the optional geometric term (src 983-1040,
// lib 1185-1334), and the 5 refinement hypotheses (src 595-676,
// lib 684-811) of active pixel aidx (valid, of colour `color`), from the
// candidate table of pm_candidates. cdf: this pixel's column of a
// float[nv][cdf_stride] scratch table; views: the problem's view
// descriptors (read for the geometric term). defer: leave planes/costs
// unchanged and record the adopted plane and cost in the workspace
// (rtp[kRefine], SW_COST) for pm_refine to commit -- required when the
// selection runs inside pm_candidates, whose other tiles still read the
// colour's planes and costs (the candidate search reaches pixels of the
// pixel's own colour, and the half-sweep reads a snapshot). Returns the
// geometric terms evaluated.
__device__ inline uint32_t select_pixel(const PMProblem& pr, int color, int iter, const Active& a,
                                        float* cdf, int cdf_stride, const ViewDesc* views,
                                        bool defer) {
    const PMConst& P = pr.P;
    float4* __restrict__ planes = pr.planes;
    float* __restrict__ costs = pr.costs;
    uint32_t* __restrict__ sel = pr.sel;
    const float* __restrict__ ca = pr.ca;
    const float4* __restrict__ cplane = pr.cplane;
    const uint8_t* __restrict__ cflag = pr.cflag;
    float4* __restrict__ rtp = pr.rtp;
    uint32_t* __restrict__ rsw = pr.rsw;
    const int NA = pr.na;
    const int aidx = a.aidx;
    const int px = a.px, py = a.py;
    const int c = py * P.w + px;
    const float fpx = (float)px, fpy = (float)py;
    const RefCam& rc = P.rc;
    const int nv = P.nv;
    const uint32_t flags = cflag[aidx];
    const Rng rng{(uint32_t)c, (uint32_t)iter, P.seed, P.view};

    // ---- joint view selection (src 915-981)
    uint32_t nbsel[4];
    const int nb[4] = {c - P.w, c + P.w, c - 1, c + 1};
#pragma unroll
    for (int i = 0; i < 4; ++i) nbsel[i] = ((flags >> (2 * i)) & 1u) ? sel[nb[i]] : 0u;
    const float tau = 0.8f * fexp((float)(iter * iter) / -90.0f);
    float psum = 0.f;
    for (int v = 0; v < nv; ++v) {
        float pri = 0.f;
#pragma unroll
        for (int i = 0; i < 4; ++i)
            if ((flags >> (2 * i)) & 1u) pri += ((nbsel[i] >> v) & 1u) ? 0.9f : 0.1f;
        float count = 0.f, tmpw = 0.f;
        int cf = 0;
        float xs[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) xs[k] = ca_entry(ca, k, v, nv, NA, aidx);
        // branch-free (adding +0 leaves the sums' bits unchanged), so the
        // loads above are all issued before the first use
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const float x = xs[k];
            const bool in = x < tau;
            const float e = fexp(x * x / -0.18f);
            tmpw += in ? e : 0.0f;
            count += in ? 1.0f : 0.0f;
            cf += (x > 1.2f);
        }
        float p = 0.f;
        if (count > 2.0f && cf < 3)
            p = tmpw / count;
        else if (cf < 3)
            p = fexp(tau * tau / -0.32f);
        p = p * pri;
        cdf[v * cdf_stride] = p;
        psum += p;
    }
    const float inv = 1.0f / psum;
    float cum = 0.f;
    for (int v = 0; v < nv; ++v) {
        cum += cdf[v * cdf_stride] * inv;
        cdf[v * cdf_stride] = cum;
    }
    float rp[15];
    float u_depth_rand;
    {
        uint32_t o[4];
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            rng.block(b, o);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int s = 4 * b + j;
                if (s < 15)
                    rp[s] = u01(o[j]) - FLT_EPSILON;
                else
                    u_depth_rand = u01(o[j]);
            }
        }
    }
    uint32_t done = 0, selmask = 0;
    uint32_t wnib[4] = {0u, 0u, 0u, 0u};
    float wsum = 0.f;
    for (int v = 0; v < nv; ++v) {
        const float cd = cdf[v * cdf_stride];
        uint32_t hit = 0;
#pragma unroll
        for (int d = 0; d < 15; ++d) hit |= (uint32_t)(cd > rp[d]) << d;
        const uint32_t fresh = hit & ~done;
        done |= fresh;
        const uint32_t cnt = __popc(fresh);
        if (cnt) {
            selmask |= 1u << v;
            const uint32_t add = cnt << (4u * (uint32_t)(v & 7));
            if ((v >> 3) == 0) wnib[0] |= add;
            else if ((v >> 3) == 1) wnib[1] |= add;
            else if ((v >> 3) == 2) wnib[2] |= add;
            else wnib[3] |= add;
            wsum += (float)cnt;
        }
    }

    // ---- candidate final costs (src 983-1009) and the current
    // hypothesis re-costed with the new view weights (its NCCs come from
    // pm_candidates, row 8 of the record). Plain and prior passes walk
    // the selected views once and read the view's 9 costs together; the
    // geometric pass adds a reprojection term per (candidate, view).
    // Every final cost is its weighted sum times one reciprocal of the
    // weight sum: a candidate that is the current plane must compare
    // equal to it (the adoption test is strict), and two divisions by
    // wsum may be lowered to different approximate sequences.
    float fc[8];
    const float rws = 1.0f / wsum;
    float4 cur = planes[c];
    float cnow;
    if (!P.geom) {
        float sk[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) sk[k] = 0.f;
        for (uint32_t mm = selmask; mm; mm &= mm - 1u) {
            const int v = __ffs(mm) - 1;
            const float wv = nib_weight(wnib, v);
            float xv[9];
#pragma unroll
            for (int k = 0; k < 9; ++k) xv[k] = ca_entry(ca, k, v, nv, NA, aidx);
#pragma unroll
            for (int k = 0; k < 9; ++k) sk[k] = fmaf(wv, xv[k], sk[k]);
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) fc[k] = sk[k] * rws;
        cnow = sk[8] * rws;
    } else {
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const bool fk = (flags >> k) & 1u;
            float4 cp = make_float4(0.f, 0.f, 0.f, 1.f);
            if (fk) cp = cplane[(size_t)k * NA + aidx];
            float sk = 0.f;
            for (uint32_t mm = selmask; mm; mm &= mm - 1u) {
                const int v = __ffs(mm) - 1;
                const float wv = nib_weight(wnib, v);
                const float x = ca_entry(ca, k, v, nv, NA, aidx);
                const float g = fk ? x + 0.1f * geom_cost(rc, views[v], cp, fpx, fpy) : x;
                sk = fmaf(wv, g, sk);
            }
            fc[k] = sk * rws;
        }
        float sk = 0.f;
        for (uint32_t mm = selmask; mm; mm &= mm - 1u) {
            const int v = __ffs(mm) - 1;
            const float x = ca_entry(ca, 8, v, nv, NA, aidx);
            const float wv = nib_weight(wnib, v);
            sk = fmaf(wv, x + 0.1f * geom_cost(rc, views[v], cur, fpx, fpy), sk);
        }
        cnow = sk * rws;
    }
    const int kmin = argmin_last8(fc);

    float cost = cnow;
    uint32_t selc = sel[c];
    float depth_now = depth_from_plane(rc, cur, fpx, fpy);
    const float4 cur0 = cur;

    const bool masked = P.prior && P.plane_masks[c] > 0u;
    float4 pp = make_float4(0.f, 0.f, 0.f, 0.f);
    float dprior = 0.f;
    if (P.prior) {
        pp = P.prior_planes[c];
        dprior = depth_from_plane(rc, pp, fpx, fpy);
    }
    float restricted = 0.f;
    if (P.prior && masked) {
        float rs[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            rs[k] = 0.f;
            if ((flags >> k) & 1u) {
                const float4 cp = cplane[(size_t)k * NA + aidx];
                rs[k] = prior_score(P, fc[k], depth_from_plane(rc, cp, fpx, fpy), dprior, pp, cp);
            }
        }
        const int kmax = argmax_last8(rs);
        const float rnow = prior_score(P, cnow, depth_from_plane(rc, cur0, fpx, fpy), dprior, pp, cur0);
        if ((flags >> kmax) & 1u) {
            const float4 cp = cplane[(size_t)kmax * NA + aidx];
            const float db = depth_from_plane(rc, cp, fpx, fpy);
            if (db >= P.dmin && db <= P.dmax && rs[kmax] > rnow) {
                depth_now = db;
                cur = cp;
                cost = fc[kmax];
                restricted = rs[kmax];
                selc = selmask;
            }
        }
    } else if ((flags >> kmin) & 1u) {
        const float4 cp = cplane[(size_t)kmin * NA + aidx];
        const float db = depth_from_plane(rc, cp, fpx, fpy);
        if (db >= P.dmin && db <= P.dmax && fc[kmin] < cnow) {
            depth_now = db;
            cur = cp;
            cost = fc[kmin];
            // the prior path leaves it (lib 1306-1318) unless
            // GMVS_COMPAT_FIX_PRIOR_SELECTION
            if (!P.prior || (P.compat & GMVS_COMPAT_FIX_PRIOR_SELECTION)) selc = selmask;
        }
    }

    // ---- refinement hypotheses (src 595-676, lib 684-811)
    const float pert = 0.02f;
    float drand;
    float4 nrand;
    uint32_t o4[4], o5[4];
    rng.block(4, o4);  // slots 16..19
    if (P.prior && masked) {
        const float dsig = (P.dmax - P.dmin) / 64.0f;
        drand = u_depth_rand * 6.0f * dsig + (dprior - 3.0f * dsig);
        rng.block(5, o5);  // slots 20..23
        nrand = perturbed_normal(rc, fpx, fpy, pp, u01(o5[0]), u01(o5[1]), u01(o5[2]),
                                 kPiF * (5.0f / 180.0f));
    } else {
        drand = u_depth_rand * (P.dmax - P.dmin) + P.dmin;
        nrand = random_normal(rc, fpx, fpy, rng, RNG_NORMAL, depth_now);
    }
    const float dlo = (1.0f - pert) * depth_now;
    const float dhi = (1.0f + pert) * depth_now;
    const float dpert = u01(o4[0]) * (dhi - dlo) + dlo;
    const float4 npert = perturbed_normal(rc, fpx, fpy, cur, u01(o4[1]), u01(o4[2]), u01(o4[3]),
                                          pert * kPiF);
    const float dcur = depth_now;
#pragma unroll
    for (int i = 0; i < kRefine; ++i) {
        const float di = (i == 0 || i == 2) ? drand : (i == 4 ? dpert : dcur);
        float4 tp = (i == 0 || i == 4) ? cur : (i == 3 ? npert : nrand);
        tp.w = dist_to_origin(rc, fpx, fpy, di, tp);
        rtp[(size_t)i * NA + aidx] = tp;
    }
    if (defer) {
        rtp[(size_t)kRefine * NA + aidx] = cur;
        rsw[(size_t)SW_COST * NA + aidx] = __float_as_uint(cost);
    } else {
        planes[c] = cur;
        costs[c] = cost;
    }
    sel[c] = selc;
    rsw[(size_t)SW_SELMASK * NA + aidx] = selmask;
#pragma unroll
    for (int i = 0; i < 4; ++i) rsw[(size_t)(SW_WNIB0 + i) * NA + aidx] = wnib[i];
    rsw[(size_t)SW_RESTRICTED * NA + aidx] = __float_as_uint(restricted);
    rsw[(size_t)SW_DRAND * NA + aidx] = __float_as_uint(drand);
    rsw[(size_t)SW_DPERT * NA + aidx] = __float_as_uint(dpert);
    rsw[(size_t)SW_DCUR * NA + aidx] = __float_as_uint(dcur);
    return P.geom ? ((uint32_t)__popc(flags) + 1u) * (uint32_t)__popc(selmask) : 0u;
}

// G horizontally adjacent tiles per block (G * 64 * L threads, tile g on
// threads [g * 64 * L, (g + 1) * 64 * L)): the tiles' gathers from one source
// region run on one CU and share its L1. Measured on the 10-view 1600x1200
// ACMP step, FP32 cells: 2 tiles per block 5 % faster than 1 (the gathers
// run at 1.00 of the measured rate of random lane pairs in an L1-resident
// window instead of 0.94, a quarter of the L1 accesses instead of 31 %
// going to the L2); 2 x 2 tiles as
// fast as 2 x 1, 4 x 1 and 1 x 2 slower than 2 x 1, 4 x 2 and 2 x 4 (one
// block per CU) slower than 1. Each tile's arithmetic is the same whatever
// G is, so results do not depend on it.
template <int NT, int L, int PR, bool FUSE, int G>
__global__ __launch_bounds__(G * kTilePix * L) __attribute__((amdgpu_waves_per_eu(4))) void pm_candidates(const PMProblem* __restrict__ probs,
                                                        const WorkTables* __restrict__ tabs,
                                                        int color, int iter,
                                                        uint32_t* __restrict__ queue,
                                                        unsigned long long* work) {
    static_assert(G == 1 || !FUSE, "the fused selection epilogue runs on one-tile blocks");
    __shared__ int s_item;
    __shared__ float s_m[G][9][3][kTilePix * L];
    __shared__ Frame s_fr[G][kMaxViews];
    const WorkTable& t = tabs->t[G == 1 ? WT_HALF : WT_HALF_PAIRS];
    const int total = t.start[t.n];
    static_assert(sizeof(s_m) >= sizeof(float) * kMaxViews * kTilePix,
                  "the CDF scratch of the fused selection lives in s_m");
    const int g = threadIdx.x / (kTilePix * L), lt = threadIdx.x % (kTilePix * L);
    uint32_t ncc = 0, geom_count = 0;
    int q = blockIdx.x % kXcds, tries = 0;
    for (;;) {
        const int item = next_item(queue, total, q, tries, s_item);
        if (item < 0) break;
        const int p = problem_of(t, item);
        const PMProblem& pr = probs[p];
        // the item's tiles: tile columns tx0 .. tx0 + G - 1 of tile row ty;
        // a tile spans 2 TX image columns (TX active ones) and TY rows
        const int TX = pr.P.tile.tx(), TY = pr.P.tile.ty();
        const int ntx = (pr.P.wh + TX - 1) / TX;
        const int ngx = (ntx + G - 1) / G;
        const int grp = item - t.start[p];
        const int ty = grp / ngx, tx0 = (grp - ty * ngx) * G;
        // the local frames of the G tiles, one thread per (tile, view)
        // (next_item's barrier retired the readers)
        if ((int)threadIdx.x < G * pr.P.nv) {
            const int gg = (int)threadIdx.x / pr.P.nv, v = (int)threadIdx.x - gg * pr.P.nv;
            tile_frame(pr.P, v, 2 * (tx0 + gg) * TX, ty * TY, 2 * TX, TY, &s_fr[gg][v]);
        }
        __syncthreads();
        const int tcx = tx0 + g;
        const int tile = ty * ntx + tcx;
        const int ox = 2 * tcx * TX, oy = ty * TY;
        if (tcx < ntx)
            pm_candidates_body<NT, L, PR>(pr.P, tile, color, pr.planes, pr.costs, pr.ca, pr.cplane,
                                          pr.cflag, pr.na, s_m[g], s_fr[g], ox, oy, ncc, lt);
        if constexpr (FUSE) {
            // the selection as the tile's epilogue: the block's candidate
            // table and planes are written (the barrier orders the global
            // writes within the workgroup); wave 0 selects for the tile's
            // 64 pixels, the CDF scratch in the m'' staging area. The
            // adopted state goes to the workspace (pm_refine commits it):
            // other tiles of this launch still read the colour's planes
            // and costs.
            __syncthreads();
            if (threadIdx.x < kTilePix) {
                const Active a = active_slot(pr.P, tile, color, threadIdx.x);
                if (a.ok)
                    geom_count += select_pixel(pr, color, iter, a, &s_m[0][0][0][0] + threadIdx.x,
                                               kTilePix, pr.P.views, true);
            }
        }
    }
    if (work) {
        count_work(work + 2 * K_CANDIDATES, ncc, 0);
        if constexpr (FUSE) count_work(work + 2 * K_SELECT, 0u, geom_count);
    }
}

// One block per WT_ACTIVE item (256 active pixels of one problem).
// Dynamic LDS: ViewDesc[nvmax], then the CDF table float[nvmax][kBlock].
__global__ __launch_bounds__(kBlock) void pm_select(const PMProblem* __restrict__ probs,
                                                   const WorkTables* __restrict__ tabs, int color,
                                                   int iter, int nvmax, unsigned long long* work) {
    extern __shared__ __align__(16) unsigned char sel_dyn[];
    ViewDesc* s_views = reinterpret_cast<ViewDesc*>(sel_dyn);
    float* s_cdf = reinterpret_cast<float*>(sel_dyn + sizeof(ViewDesc) * (size_t)nvmax);
    const WorkTable& t = tabs->t[WT_ACTIVE];
    const int item = blockIdx.x;
    const int pidx = problem_of(t, item);
    const PMProblem& pr = probs[pidx];
    const PMConst& P = pr.P;
    if (P.geom) {
        const int n = P.nv * (int)(sizeof(ViewDesc) / 4);
        const uint32_t* src = reinterpret_cast<const uint32_t*>(P.views);
        uint32_t* dst = reinterpret_cast<uint32_t*>(s_views);
        for (int i = threadIdx.x; i < n; i += kBlock) dst[i] = src[i];
        __syncthreads();
    }
    uint32_t geom_count = 0;
    const int aidx = (item - t.start[pidx]) * kBlock + threadIdx.x;
    const Active a = active_of_index(P, color, aidx);
    if (aidx < pr.na && a.ok)
        geom_count = select_pixel(pr, color, iter, a, s_cdf + threadIdx.x, kBlock, s_views, false);
    if (work) count_work(work + 2 * K_SELECT, 0u, geom_count);
}


template <int NT, int L>
struct RefineSmem {
    static constexpr int NW = TapBlock<NT, L>::NR * TapBlock<NT, L>::NC;
    float hw[kK2Pix][L][2 * NW];  // hoisted w and w * (r - c) per (pixel, lane)
    float hp[kK2Pix][4];               // 1 / sum w, mean_r, var_r, centre intensity
    float4 tp[kRefine][kK2Pix];        // refinement hypotheses
    float4 tm[kRefine][kK2Pix];        // their m'' in the tile's frame (w unused)
    float gv[kRefine][kK2Pix];         // geometric terms of a lane group's current unit
    int nunits;
};

// Dynamic LDS of pm_refine: ViewDesc[nv], uint16 units[64 * nv] (slot |
// view << 6), float results[kRefine][64][nv].
__host__ __device__ inline size_t refine_units_off(int nv) { return sizeof(ViewDesc) * (size_t)nv; }
__host__ __device__ inline size_t refine_res_off(int nv) {
    return (refine_units_off(nv) + sizeof(uint16_t) * (size_t)kK2Pix * nv + 15) & ~size_t(15);
}
__host__ __device__ inline size_t refine_dyn_bytes(int nv) {
    return refine_res_off(nv) + sizeof(float) * (size_t)kRefine * kK2Pix * nv;
}

// One tile of pm_refine (every thread of the block calls it: it holds block
// barriers). s_views holds the problem's view descriptors.
template <int NT, int L, int PR>
__device__ inline void pm_refine_body(const PMConst& P, int tile, int color,
                                      float4* __restrict__ planes, float* __restrict__ costs,
                                      const float4* __restrict__ rtp,
                                      const uint32_t* __restrict__ rsw, int NA, RefineSmem<NT, L>& S,
                                      const ViewDesc* s_views, const Frame* s_fr, int ox, int oy,
                                      uint16_t* units, float* res, bool commit, uint32_t& ncc,
                                      uint32_t& geom) {
    constexpr int NW = RefineSmem<NT, L>::NW;
    const RefCam& rc = P.rc;
    const int g = threadIdx.x / L, sub = threadIdx.x % L;
    // the tile's refinement hypotheses
    for (int i = threadIdx.x; i < kK2Pix * kRefine; i += kTilePix * L) {
        const int sl = i % kK2Pix, f = i / kK2Pix;
        const Active a = active_slot(P, tile, color, sl);
        const float4 tp = a.ok ? rtp[(size_t)f * NA + a.aidx] : make_float4(0.f, 0.f, 0.f, 0.f);
        float m[3];
        plane_m_local(rc, tp, (float)ox, (float)oy, m);
        S.tp[f][sl] = tp;
        S.tm[f][sl] = make_float4(m[0], m[1], m[2], 0.f);
    }
    // hoisted reference side of this lane group's pixel
    const Active ag = active_slot(P, tile, color, g);
    if (ag.ok) {
        Hoist<NT, L> hs;
        hoist_ref<NT, L>(hs, P.ref, ag.px, ag.py, P.h, P.w, P.radius, P.inc, P.sigma_s, P.sigma_c, sub,
                         ox, oy);
#pragma unroll
        for (int i = 0; i < NW; ++i) {
            S.hw[g][sub][i] = hs.wt[i];
            S.hw[g][sub][NW + i] = hs.wr[i];
        }
        if (sub == 0) {
            S.hp[g][0] = hs.inv_w;
            S.hp[g][1] = hs.mean_r;
            S.hp[g][2] = hs.var_r;
            S.hp[g][3] = hs.c;
        }
    }
    __syncthreads();
    // work units (slot, view), view-major: the groups working at the same
    // time gather from one source view for neighbouring pixels. Wave 0
    // (one lane per slot) ranks each pixel among the tile's pixels that
    // selected view v with a ballot.
    if (threadIdx.x < kK2Pix) {
        const int sl = threadIdx.x;
        const Active a = active_slot(P, tile, color, sl);
        const uint32_t mask = a.ok ? rsw[(size_t)SW_SELMASK * NA + a.aidx] : 0u;
        const uint64_t below = (sl == 0) ? 0ull : (~0ull >> (64 - sl));
        int base = 0;
        for (int v = 0; v < P.nv; ++v) {
            const bool on = (mask >> v) & 1u;
            const uint64_t b = __ballot(on);
            if (on) units[base + __popcll(b & below)] = (uint16_t)(sl | (v << 6));
            base += __popcll(b);
        }
        if (sl == 0) S.nunits = base;
    }
    __syncthreads();

    const int nunits = S.nunits;
    for (int u = g; u < nunits; u += kK2Pix) {
        const uint32_t code = units[u];
        const int sl = (int)(code & 63u), v = (int)(code >> 6);
        const Active a = active_slot(P, tile, color, sl);
        const float fpx = (float)a.px, fpy = (float)a.py;
        HoistRef<NT, L> hs;
        hs.wt = &S.hw[sl][sub][0];
        hs.wr = &S.hw[sl][sub][NW];
        hoist_taps<NT, L, HoistRef<NT, L>>(hs, a.px, a.py, P.h, P.w, P.radius, P.inc, sub, ox, oy);
        hs.inv_w = S.hp[sl][0];
        hs.mean_r = S.hp[sl][1];
        hs.var_r = S.hp[sl][2];
        hs.c = S.hp[sl][3];
        const ViewDesc& vd = s_views[v];
        const Frame& fr = s_fr[v];
        const TapSrc src =
            tap_src_at<PR>(P.cells, (uint32_t)(P.nv * P.vcells), v * P.vcells, fr);
        if (P.geom) {
            // the unit's geometric terms, hypothesis k on lane k mod L of
            // the group (each lane computing all five repeats the same
            // arithmetic L times); the group's lanes are in one wave, whose
            // LDS accesses complete in order
#pragma unroll 1
            for (int k = sub; k < kRefine; k += L) S.gv[k][g] = geom_cost(rc, vd, S.tp[k][sl], fpx, fpy);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
#pragma unroll 1
        for (int k = 0; k < kRefine; ++k) {
            const float4 tm = S.tm[k][sl];
            const float m[3] = {tm.x, tm.y, tm.z};
            float H[9];
            homography(fr.A, fr.b, m, H);
            float x = ncc_eval_mlp<NT, L, PR>(hs, H, src, P.grid());
            if (P.geom) x = x + 0.1f * S.gv[k][g];
            if (sub == 0) res[(k * kK2Pix + sl) * P.nv + v] = x;
        }
        if (sub == 0) {
            ncc += kRefine;
            if (P.geom) geom += kRefine;
        }
    }
    __syncthreads();

    // sequential adoption over the 5 hypotheses, weighted in view order
    if (threadIdx.x < kK2Pix) {
        const int sl = threadIdx.x;
        const Active a = active_slot(P, tile, color, sl);
        if (a.ok) {
            const int c = a.py * P.w + a.px;
            const float fpx = (float)a.px, fpy = (float)a.py;
            const uint32_t* rec = rsw + a.aidx;
            const uint32_t mask = rec[(size_t)SW_SELMASK * NA];
            uint32_t wnib[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) wnib[i] = rec[(size_t)(SW_WNIB0 + i) * NA];
            float wsum = 0.f;
            for (uint32_t mm = mask; mm; mm &= mm - 1u) wsum += nib_weight(wnib, __ffs(mm) - 1);
            const bool pmask = P.prior && P.plane_masks[c] > 0u;
            float4 pp = make_float4(0.f, 0.f, 0.f, 0.f);
            float dprior = 0.f;
            if (P.prior) {
                pp = P.prior_planes[c];
                dprior = depth_from_plane(rc, pp, fpx, fpy);
            }
            const float drand = __uint_as_float(rec[(size_t)SW_DRAND * NA]);
            const float dpert = __uint_as_float(rec[(size_t)SW_DPERT * NA]);
            const float dcur = __uint_as_float(rec[(size_t)SW_DCUR * NA]);
            float restricted = __uint_as_float(rec[(size_t)SW_RESTRICTED * NA]);
            // with the selection fused into pm_candidates the adopted state
            // is still in the workspace: committed here either way
            float cost = commit ? __uint_as_float(rec[(size_t)SW_COST * NA]) : costs[c];
            int best = -1;
#pragma unroll 1
            for (int i = 0; i < kRefine; ++i) {
                const float4 tp = S.tp[i][sl];
                float s = 0.f;
                for (uint32_t mm = mask; mm; mm &= mm - 1u) {
                    const int v = __ffs(mm) - 1;
                    s += nib_weight(wnib, v) * res[(i * kK2Pix + sl) * P.nv + v];
                }
                const float tc = s / wsum;
                const float di = (i == 0 || i == 2) ? drand : (i == 4 ? dpert : dcur);
                const float db = depth_from_plane(rc, tp, fpx, fpy);
                if (pmask) {
                    const float sc = prior_score(P, tc, di, dprior, pp, tp);
                    if (db >= P.dmin && db <= P.dmax && sc > restricted) {
                        best = i;
                        cost = tc;
                        restricted = sc;
                    }
                } else if (db >= P.dmin && db <= P.dmax && tc < cost) {
                    best = i;
                    cost = tc;
                }
            }
            if (best >= 0) {
                planes[c] = S.tp[best][sl];
                costs[c] = cost;
            } else if (commit) {
                planes[c] = rtp[(size_t)kRefine * NA + a.aidx];
                costs[c] = cost;
            }
        }
    }
}

// Persistent over WT_HALF tiles. Dynamic LDS: ViewDesc[nvmax], uint16
// units[64 * nvmax] (slot | view << 6), float results[kRefine][64][nvmax].
template <int NT, int L, int PR>
__global__ __launch_bounds__(kTilePix * L) __attribute__((amdgpu_waves_per_eu(PR == PREC_F32 ? 3 : 4))) void pm_refine(const PMProblem* __restrict__ probs,
                                                   const WorkTables* __restrict__ tabs, int color,
                                                   int nvmax, int commit,
                                                   uint32_t* __restrict__ queue,
                                                   unsigned long long* work) {
    __shared__ RefineSmem<NT, L> S;
    __shared__ Frame s_fr[kMaxViews];
    __shared__ int s_item;
    extern __shared__ __align__(16) unsigned char ref_dyn[];
    ViewDesc* s_views = reinterpret_cast<ViewDesc*>(ref_dyn);
    uint16_t* units = reinterpret_cast<uint16_t*>(ref_dyn + refine_units_off(nvmax));
    float* res = reinterpret_cast<float*>(ref_dyn + refine_res_off(nvmax));
    const WorkTable& t = tabs->t[WT_HALF];
    const int total = t.start[t.n];
    uint32_t ncc = 0, geom = 0;
    int q = blockIdx.x % kXcds, tries = 0, loaded = -1;
    for (;;) {
        const int item = next_item(queue, total, q, tries, s_item);
        if (item < 0) break;
        const int p = problem_of(t, item);
        const PMProblem& pr = probs[p];
        if (p != loaded) {  // block-uniform; the body's first barrier orders it
            const int n = pr.P.nv * (int)(sizeof(ViewDesc) / 4);
            const uint32_t* src = reinterpret_cast<const uint32_t*>(pr.P.views);
            uint32_t* dst = reinterpret_cast<uint32_t*>(s_views);
            for (int i = threadIdx.x; i < n; i += kTilePix * L) dst[i] = src[i];
            loaded = p;
        }
        const int tile = item - t.start[p];
        const int TX = pr.P.tile.tx(), TY = pr.P.tile.ty();
        const int ntx = (pr.P.wh + TX - 1) / TX;
        const int ox = 2 * (tile % ntx) * TX, oy = (tile / ntx) * TY;
        tile_frames(pr.P, ox, oy, 2 * TX, TY, s_fr);  // the body's first barrier orders it
        pm_refine_body<NT, L, PR>(pr.P, tile, color, pr.planes, pr.costs, pr.rtp, pr.rsw, pr.na, S,
                                  s_views, s_fr, ox, oy, units, res, commit != 0, ncc, geom);
    }
    if (work) count_work(work + 2 * K_REFINE, ncc, geom);
}

// ------------------------------------------------------ finalisation ---
// GetDepthandNormal (lib 1655-1674): .w <- depth, normal camera -> world.
// One block per WT_PIXELS item.
__global__ void pm_depth_normal(const PMProblem* __restrict__ probs,
                                const WorkTables* __restrict__ tabs) {
    const WorkTable& t = tabs->t[WT_PIXELS];
    const int pi = problem_of(t, blockIdx.x);
    const PMConst& P = probs[pi].P;
    float4* __restrict__ planes = probs[pi].planes;
    const int i = (blockIdx.x - t.start[pi]) * blockDim.x + threadIdx.x;
    if (i >= P.h * P.w) return;
    const int py = i / P.w, px = i - py * P.w;
    const RefCam& rc = P.rc;
    float4 p = planes[i];
    p.w = depth_from_plane(rc, p, (float)px, (float)py);
    const float x = rc.R[0] * p.x + rc.R[3] * p.y + rc.R[6] * p.z;
    const float y = rc.R[1] * p.x + rc.R[4] * p.y + rc.R[7] * p.z;
    const float z = rc.R[2] * p.x + rc.R[5] * p.y + rc.R[8] * p.z;
    planes[i] = make_float4(x, y, z, p.w);
}

// CheckerboardFilter (src 1165-1281): median of the depth at the pixel and
// at up to 20 fixed opposite-colour offsets, unless the cost is < 0.001.
// The median is taken by rank counting over the (at most 21) values.
// One block per WT_ACTIVE item; problems with skip_filter set are skipped.
__global__ __launch_bounds__(kBlock) void pm_filter(const PMProblem* __restrict__ probs,
                                                    const WorkTables* __restrict__ tabs,
                                                    int color) {
    const WorkTable& t = tabs->t[WT_ACTIVE];
    const int p = problem_of(t, blockIdx.x);
    if (probs[p].skip_filter) return;
    const PMConst& P = probs[p].P;
    float4* __restrict__ planes = probs[p].planes;
    const float* __restrict__ costs = probs[p].costs;
    const int i = (blockIdx.x - t.start[p]) * blockDim.x + threadIdx.x;
    const int wh = P.wh;
    if (i >= P.h * wh) return;
    const int py = i / wh;
    const int px = 2 * (i - py * wh) + ((py + color) & 1);
    const int w = P.w, h = P.h;
    if (px >= w) return;
    const int c = py * w + px;
    if (costs[c] < 0.001f) return;
    // g[0] is the pixel itself, g[1..20] the reference's offsets in order;
    // gv marks the ones inside the image.
    const int up = c - w, dn = c + w, lf = c - 1, rt = c + 1;
    const int q[20] = {up,         c - 3 * w,  c - 5 * w,  dn,         c + 3 * w,
                       c + 5 * w,  lf,         c - 3,      c - 5,      rt,
                       c + 3,      c + 5,      up + 2,     dn + 2,     up - 2,
                       dn - 2,     lf - 2 * w, rt - 2 * w, lf + 2 * w, rt + 2 * w};
    const bool u[20] = {py > 0,
                        py > 2,
                        py > 4,
                        py < h - 1,
                        py < h - 3,
                        py < h - 5,
                        px > 0,
                        px > 2,
                        px > 4,
                        px < w - 1,
                        px < w - 3,
                        px < w - 5,
                        py > 0 && px < w - 2,
                        py < h - 1 && px < w - 2,
                        py > 0 && px > 1,
                        py < h - 1 && px > 1,
                        px > 0 && py > 2,
                        px < w - 1 && py > 2,
                        px > 0 && py < h - 2,
                        px < w - 1 && py < h - 2};
    float g[21];
    bool gv[21];
    g[0] = planes[c].w;
    gv[0] = true;
#pragma unroll
    for (int k = 0; k < 20; ++k) {
        gv[k + 1] = u[k];
        g[k + 1] = u[k] ? planes[q[k]].w : 0.f;
    }
    int n = 0;
#pragma unroll
    for (int k = 0; k < 21; ++k) n += gv[k];
    // value of rank r: the element e with #(< e) <= r < #(<= e)
    const int m = n / 2;
    float lo = 0.f, hi = 0.f;
#pragma unroll
    for (int k = 0; k < 21; ++k) {
        if (!gv[k]) continue;
        int less = 0, leq = 0;
#pragma unroll
        for (int j = 0; j < 21; ++j) {
            if (!gv[j]) continue;
            less += g[j] < g[k];
            leq += g[j] <= g[k];
        }
        if (less <= m && m < leq) hi = g[k];
        if (less <= m - 1 && m - 1 < leq) lo = g[k];
    }
    planes[c].w = (n % 2 == 0) ? (lo + hi) / 2.0f : hi;
}

// export_plane_hypotheses_cost (src 1284-1313). One block per WT_PIXELS item.
__global__ void pm_export(const PMProblem* __restrict__ probs, const WorkTables* __restrict__ tabs) {
    const WorkTable& t = tabs->t[WT_PIXELS];
    const int p = problem_of(t, blockIdx.x);
    const PMProblem& pr = probs[p];
    const int i = (blockIdx.x - t.start[p]) * blockDim.x + threadIdx.x;
    if (i >= pr.P.h * pr.P.w) return;
    const float4 v = pr.planes[i];
    if (pr.out_depth) pr.out_depth[i] = v.w;
    if (pr.out_normal) {
        pr.out_normal[3 * i + 0] = v.x;
        pr.out_normal[3 * i + 1] = v.y;
        pr.out_normal[3 * i + 2] = v.z;
    }
    if (pr.out_cost) pr.out_cost[i] = pr.costs[i];
}

// Parity hook: V NCCs for a list of (pixel, plane) pairs.
template <int NT, int L, int PR>
__global__ __launch_bounds__(kTilePix * L) void pm_cost_batch(PMConst P, int n,
                                                        const int32_t* __restrict__ xy,
                                                        const float4* __restrict__ planes,
                                                        float* __restrict__ out) {
    const int q = (blockIdx.x * blockDim.x + threadIdx.x) / L;
    const int sub = threadIdx.x % L;
    if (q >= n) return;
    const int px = xy[2 * q], py = xy[2 * q + 1];
    if (px < 0 || px >= P.w || py < 0 || py >= P.h) {
        if (sub == 0)
            for (int v = 0; v < P.nv; ++v) out[(size_t)q * P.nv + v] = __builtin_nanf("");
        return;
    }
    // the local frame of the query: reference positions relative to the
    // pixel, source positions anchored at the pixel's image under the
    // queried plane itself
    Hoist<NT, L> hs;
    hoist_ref<NT, L>(hs, P.ref, px, py, P.h, P.w, P.radius, P.inc, P.sigma_s, P.sigma_c, sub, px,
                     py);
    const float4 pl = planes[q];
    float m[3];
    plane_m_local(P.rc, pl, (float)px, (float)py, m);
    double am[3];
    {
        const RefCam& c = P.rc;
        const double id = 1.0 / (double)pl.w;
        const double a = (double)pl.x / c.K[0], b = (double)pl.y / c.K[4];
        am[0] = a * id;
        am[1] = b * id;
        am[2] = ((double)pl.z - a * c.K[2] - b * c.K[5]) * id;
    }
    for (int v = 0; v < P.nv; ++v) {
        const ViewDesc& vd = P.views[v];
        const Frame fr = make_frame(P.views64[v], px, py, px, py, am, P.w, P.h, P.wp);
        float H[9];
        homography(fr.A, fr.b, m, H);
        const float cv =
            ncc_eval<NT, L, PR>(hs, H, tap_src_at<PR>(vd.cells, P.vcells, 0, fr), P.grid());
        if (sub == 0) out[(size_t)q * P.nv + v] = cv;
    }
}

// ================================================================ host ===
namespace {

size_t align_up(size_t x) { return (x + 255) & ~size_t(255); }

struct PMLayout {
    size_t views, views64, quads, sel, ca, cplane, cflag, rtp, rsw, total;
};

bool valid_precision(int precision) {
    return precision == GMVS_PREC_FP32 || precision == GMVS_PREC_FP16 || precision == GMVS_PREC_BF16;
}

size_t cell_bytes(int precision) { return precision == GMVS_PREC_FP32 ? 16 : 8; }

// cells of one view's padded grid
size_t grid_cells(int h, int w) { return size_t(h + kPad) * size_t(w + kPad); }

PMLayout pm_layout(int h, int w, int num_images, bool sweep, int precision) {
    const int nv = num_images - 1;
    const size_t hw = size_t(h) * w;
    const size_t na = size_t(h) * ((w + 1) / 2);
    PMLayout L;
    size_t o = 0;
    L.views = o;
    o = align_up(o + sizeof(ViewDesc) * nv);
    L.views64 = o;
    o = align_up(o + sizeof(ViewF64) * nv);
    L.quads = o;
    o = align_up(o + cell_bytes(precision) * grid_cells(h, w) * nv);
    L.sel = o;
    if (sweep) {
        o = align_up(o + sizeof(uint32_t) * hw);
        L.ca = o;
        o = align_up(o + sizeof(float) * 9 * nv * na);
        L.cplane = o;
        o = align_up(o + sizeof(float4) * 8 * na);
        L.cflag = o;
        o = align_up(o + na);
        L.rtp = o;
        o = align_up(o + sizeof(float4) * (kRefine + 1) * na);
        L.rsw = o;
        o = align_up(o + sizeof(uint32_t) * SW_NUM * na);
    } else {
        L.ca = L.cplane = L.cflag = L.rtp = L.rsw = o;
    }
    L.total = o;
    return L;
}

int taps_per_axis(const gmvs_params& p, int* radius) {
    *radius = p.patch_size / 2;
    if (p.radius_increment <= 0) return -1;
    int n = 0;
    for (int i = -*radius; i < *radius + 1; i += p.radius_increment) ++n;
    return n;
}

int check_cameras(int num_images, const gmvs_camera* cams) {
    GMVS_ARG_CHECK(cams, "cameras is NULL");
    GMVS_ARG_CHECK(num_images >= 2 && num_images <= kMaxViews + 1,
                   "num_images must be in [2, %d] (got %d)", kMaxViews + 1, num_images);
    const int h = cams[0].height, w = cams[0].width;
    GMVS_ARG_CHECK(h > 0 && w > 0, "reference image size must be positive (got %dx%d)", w, h);
    // tap addresses are formed as y * (w + kPad) + x in FP32 and in 24-bit
    // integer products, exact while the padded grid has at most 2^24 cells
    GMVS_ARG_CHECK((long long)(h + kPad) * (w + kPad) <= (1LL << 24),
                   "image too large (%dx%d; (w + %d) * (h + %d) must be at most 2^24)", w, h, kPad,
                   kPad);
    for (int i = 1; i < num_images; ++i)
        GMVS_ARG_CHECK(cams[i].height == h && cams[i].width == w,
                       "all images of a problem must share the reference size (view %d is %dx%d, "
                       "reference %dx%d)", i, cams[i].width, cams[i].height, w, h);
    return GMVS_OK;
}

// Composes the per-view homography factors (FP64, one tiny launch per view)
// and packs the source images.
template <int PR>
int setup_views(int num_images, const gmvs_camera* cams, const float* const* images,
                const float* const* depths, char* ws, const PMLayout& L, hipStream_t st) {
    const int h = cams[0].height, w = cams[0].width;
    const int nv = num_images - 1;
    char* cells = ws + L.quads;
    const size_t view_bytes = grid_cells(h, w) * Cell<PR>::kBytes;
    ViewDesc* views = reinterpret_cast<ViewDesc*>(ws + L.views);
    ViewF64* views64 = reinterpret_cast<ViewF64*>(ws + L.views64);
    const int blocks = int((grid_cells(h, w) + 255) / 256);
    for (int v = 0; v < nv; ++v) {
        GMVS_LAUNCH(K_SETUP_VIEW, st, pm_setup_view, dim3(1), dim3(64), 0, st, cams[0], cams[v + 1],
                    (const void*)(cells + view_bytes * v), depths ? depths[v + 1] : nullptr, views + v,
                    views64 + v);
        GMVS_LAUNCH_CHECK();
        GMVS_LAUNCH(K_PACK, st, pm_pack_quads<PR>, dim3(blocks), dim3(256), 0, st, images[v + 1],
                    (void*)(cells + view_bytes * v), h, w);
        GMVS_LAUNCH_CHECK();
    }
    return GMVS_OK;
}

// Clamp-free tap indexing allowed (gmvs_set_fast_taps; results do not depend on it).
std::atomic<int> g_fast_taps{1};
// View selection and adoption as pm_candidates' tile epilogue instead of the
// pm_select launch (gmvs_set_select_fusion; results do not depend on it).
std::atomic<int> g_fuse_select{0};

PMConst make_const(int num_images, const gmvs_camera* cams, const float* ref, const ViewDesc* views,
                   const ViewF64* views64, const gmvs_params& prm, int radius) {
    PMConst P;
    std::memset(&P, 0, sizeof(P));
    const gmvs_camera& r = cams[0];
    for (int i = 0; i < 9; ++i) {
        P.rc.K[i] = r.K[i];
        P.rc.R[i] = r.R[i];
    }
    for (int i = 0; i < 3; ++i) {
        P.rc.t[i] = r.t[i];
        P.rc.C[i] = -(r.R[i] * r.t[0] + r.R[3 + i] * r.t[1] + r.R[6 + i] * r.t[2]);
    }
    P.rc.h = r.height;
    P.rc.w = r.width;
    P.ref = ref;
    P.views = views;
    P.views64 = views64;
    P.nv = num_images - 1;
    P.h = r.height;
    P.w = r.width;
    P.wh = (r.width + 1) / 2;
    P.wp = r.width + kPad;
    P.vcells = (int)grid_cells(r.height, r.width);
    P.fast_taps = g_fast_taps.load();
    P.radius = radius;
    P.inc = prm.radius_increment;
    P.sigma_s = prm.sigma_spatial;
    P.sigma_c = prm.sigma_color;
    P.dmin = prm.depth_min;
    P.dmax = prm.depth_max;
    P.top_k = prm.top_k;
    return P;
}

}  // namespace

}  // namespace gmvs

using namespace gmvs;

namespace {

constexpr int kMaxIterations = 1024;
constexpr int kQueueSlots = 1 + 4 * kMaxIterations;  // init + (candidates, refine) x 2 colours

// Batch workspace: problem records, work tables and queue counters, then one
// pm_layout region per problem.
struct BatchLayout {
    size_t probs, tabs, queue, total;
    std::vector<size_t> region;
};

BatchLayout batch_layout(const gmvs_pm_args* args, int n) {
    BatchLayout B;
    size_t o = 0;
    B.probs = o;
    o = align_up(o + sizeof(PMProblem) * n);
    B.tabs = o;
    o = align_up(o + sizeof(WorkTables));
    B.queue = o;
    o = align_up(o + sizeof(uint32_t) * kXcds * kQueueSlots);
    for (int i = 0; i < n; ++i) {
        const gmvs_pm_args& A = args[i];
        B.region.push_back(o);
        o += pm_layout(A.cameras[0].height, A.cameras[0].width, A.num_images, true, A.precision).total;
        o = align_up(o);
    }
    B.total = o;
    return B;
}

// Tile columns of a 64-pixel tile; 0 selects the default 8 x 8.
int tile_shift(int cols) {
    switch (cols) {
        case 4: return 2;
        case 16: return 4;
        case 32: return 5;
        default: return 3;
    }
}

// Argument checks of one problem (the reference's own preconditions made
// explicit).
int check_args(const gmvs_pm_args& A, int* radius) {
    if (int rc = check_cameras(A.num_images, A.cameras)) return rc;
    GMVS_ARG_CHECK(A.images, "images is NULL");
    for (int i = 0; i < A.num_images; ++i) GMVS_ARG_CHECK(A.images[i], "images[%d] is NULL", i);
    GMVS_ARG_CHECK(A.planes && A.costs, "planes/costs state buffers are required");
    GMVS_ARG_CHECK(valid_precision(A.precision), "unsupported precision %d", A.precision);
    GMVS_ARG_CHECK(A.init_mode >= 0 && A.init_mode <= 2, "bad init_mode %d", A.init_mode);
    GMVS_ARG_CHECK(A.params.max_iterations >= 0 && A.params.max_iterations <= kMaxIterations,
                   "max_iterations must be in [0, %d]", kMaxIterations);
    GMVS_ARG_CHECK(A.params.top_k >= 0, "top_k must be >= 0");
    GMVS_ARG_CHECK(A.tile_cols == 0 || A.tile_cols == 4 || A.tile_cols == 8 || A.tile_cols == 16 ||
                       A.tile_cols == 32,
                   "tile_cols must be 0, 4, 8, 16 or 32 (got %d)", A.tile_cols);
    GMVS_ARG_CHECK((A.compat & ~15u) == 0u, "unknown compat bits 0x%x", A.compat & ~15u);
    if (A.init_mode == GMVS_INIT_DEPTH_NORMAL)
        GMVS_ARG_CHECK(A.depths && A.depths[0] && A.ref_normal,
                       "INIT_DEPTH_NORMAL needs depths[0] and ref_normal");
    if (A.geom) {
        GMVS_ARG_CHECK(A.depths, "the geometric term needs source depth maps");
        for (int i = 1; i < A.num_images; ++i) GMVS_ARG_CHECK(A.depths[i], "depths[%d] is NULL", i);
    }
    if (A.params.planar_prior)
        GMVS_ARG_CHECK(A.prior_planes && A.plane_masks, "planar_prior needs prior planes and masks");
    const int nt = taps_per_axis(A.params, radius);
    if (nt != 6) {
        set_error("patch geometry (patch_size=%d, radius_increment=%d) gives %d taps per axis; "
                  "this build is specialised for 6 (the reference defaults 11/2)",
                  A.params.patch_size, A.params.radius_increment, nt);
        return GMVS_ERR_UNSUPPORTED;
    }
    return GMVS_OK;
}

// Blocks per CU of the persistent kernels (0 = as many as fit), set by
// gmvs_set_blocks_per_cu; the autotuner picks them by timing.
std::atomic<int> g_blocks_per_cu[K_NUM];

// Lane-group width per NCC kernel (0 = default_lanes), set by gmvs_set_lane_width.
std::atomic<int> g_lanes[K_NUM];

int lanes_of(int kid, int precision) {
    const int l = g_lanes[kid].load();
    return l ? l : default_lanes(kid, precision);
}

// Blocks of a persistent kernel on the current device: all CUs x the
// occupancy at this block size and dynamic LDS (cached per kernel), or the
// tuned blocks per CU when that is lower.
int persistent_grid(const void* kernel, int block, size_t dyn_lds, int items, int kid) {
    static std::mutex mu;
    static std::vector<std::pair<std::pair<const void*, size_t>, int>> cache;
    int per_cu = 0;
    {
        std::lock_guard<std::mutex> lk(mu);
        for (const auto& e : cache)
            if (e.first.first == kernel && e.first.second == dyn_lds) per_cu = e.second;
    }
    if (per_cu == 0) {
        int bl = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&bl, kernel, block, dyn_lds) != hipSuccess ||
            bl <= 0)
            bl = 1;
        int dev = 0, cus = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            cus <= 0)
            cus = 256;
        per_cu = bl * cus;
        std::lock_guard<std::mutex> lk(mu);
        cache.push_back({{kernel, dyn_lds}, per_cu});
    }
    const int tuned = g_blocks_per_cu[kid].load();
    if (tuned > 0) {
        int dev = 0, cus = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
            cus > 0)
            per_cu = std::min(per_cu, tuned * cus);
    }
    return std::max(1, std::min(per_cu, items));
}

// One batch of passes (init, sweeps, depth/normal, filter, export), every
// kernel launched once for all problems, in cell precision PR.
template <int PR>
int run_batch(const gmvs_pm_args* args, int n, const int* radius, char* ws, const BatchLayout& B,
              hipStream_t st) {
    PMProblem* d_probs = reinterpret_cast<PMProblem*>(ws + B.probs);
    WorkTables* d_tabs = reinterpret_cast<WorkTables*>(ws + B.tabs);
    uint32_t* queue = reinterpret_cast<uint32_t*>(ws + B.queue);
    unsigned long long* work = profile_work_counters();
    const int iters = args[0].params.max_iterations;
    const int nq = kXcds * (1 + 4 * iters);
    GMVS_LAUNCH(K_SETUP_VIEW, st, pm_zero_words, dim3((nq + 255) / 256), dim3(256), 0, st, queue, nq);
    GMVS_LAUNCH_CHECK();

    WorkTables tabs;
    std::memset(&tabs, 0, sizeof(tabs));
    for (int k = 0; k < WT_NUM; ++k) tabs.t[k].n = n;
    int nvmax = 1;
    for (int i = 0; i < n; ++i) {
        const gmvs_pm_args& A = args[i];
        char* rw = ws + B.region[i];
        const int h = A.cameras[0].height, w = A.cameras[0].width;
        const PMLayout L = pm_layout(h, w, A.num_images, true, A.precision);
        if (int rc = setup_views<PR>(A.num_images, A.cameras, A.images, A.depths, rw, L, st)) return rc;
        PMProblem pr;
        std::memset(&pr, 0, sizeof(pr));
        PMConst& P = pr.P;
        P = make_const(A.num_images, A.cameras, A.images[0],
                       reinterpret_cast<const ViewDesc*>(rw + L.views),
                       reinterpret_cast<const ViewF64*>(rw + L.views64), A.params, radius[i]);
        P.cells = rw + L.quads;
        P.init_depth = A.depths ? A.depths[0] : nullptr;
        P.init_normal = A.ref_normal;
        P.prior_planes = reinterpret_cast<const float4*>(A.prior_planes);
        P.plane_masks = A.plane_masks;
        P.seed = A.seed;
        P.view = A.rng_view;
        P.init_mode = A.init_mode;
        P.geom = A.geom ? 1 : 0;
        P.prior = A.params.planar_prior ? 1 : 0;
        P.compat = A.compat;
        P.tile.shift = tile_shift(A.tile_cols);
        pr.planes = reinterpret_cast<float4*>(A.planes);
        pr.costs = A.costs;
        pr.sel = reinterpret_cast<uint32_t*>(rw + L.sel);
        pr.ca = reinterpret_cast<float*>(rw + L.ca);
        pr.cplane = reinterpret_cast<float4*>(rw + L.cplane);
        pr.cflag = reinterpret_cast<uint8_t*>(rw + L.cflag);
        pr.rtp = reinterpret_cast<float4*>(rw + L.rtp);
        pr.rsw = reinterpret_cast<uint32_t*>(rw + L.rsw);
        pr.out_depth = A.out_depth;
        pr.out_normal = A.out_normal;
        pr.out_cost = A.out_cost;
        pr.na = h * P.wh;
        pr.skip_filter = A.skip_filter ? 1 : 0;
        GMVS_LAUNCH(K_SETUP_VIEW, st, pm_setup_problem, dim3(1), dim3(64), 0, st, pr, d_probs + i);
        GMVS_LAUNCH_CHECK();
        nvmax = std::max(nvmax, P.nv);
        const int htx = (P.wh + P.tile.tx() - 1) / P.tile.tx();
        const int hty = (h + P.tile.ty() - 1) / P.tile.ty();
        const int counts[WT_NUM] = {P.tile.tiles(w, h), P.tile.tiles(P.wh, h),
                                    (pr.na + kBlock - 1) / kBlock, (h * w + kBlock - 1) / kBlock,
                                    ((htx + 1) / 2) * hty};
        for (int k = 0; k < WT_NUM; ++k) tabs.t[k].start[i + 1] = tabs.t[k].start[i] + counts[k];
    }
    GMVS_LAUNCH(K_SETUP_VIEW, st, pm_setup_tables, dim3(1), dim3(64), 0, st, tabs, d_tabs);
    GMVS_LAUNCH_CHECK();

    const int n_full = tabs.t[WT_FULL].start[n], n_half = tabs.t[WT_HALF].start[n];
    const int n_pairs = tabs.t[WT_HALF_PAIRS].start[n];
    const int n_act = tabs.t[WT_ACTIVE].start[n], n_pix = tabs.t[WT_PIXELS].start[n];
    const size_t sel_lds = sizeof(ViewDesc) * (size_t)nvmax + sizeof(float) * (size_t)nvmax * kBlock;
    const size_t ref_lds = refine_dyn_bytes(nvmax);
    // lane-group width per kernel: block = tiles x 64 pixels x L lanes
    const int li = lanes_of(K_INIT, PR), lc = lanes_of(K_CANDIDATES, PR), lr = lanes_of(K_REFINE, PR);
    auto k_init = li == 2 ? pm_init<kNT, 2, PR> : pm_init<kNT, 4, PR>;
    const bool fuse = g_fuse_select.load() != 0;
    // the candidate kernel takes tile pairs unless its selection epilogue is on
    const int gc = fuse ? 1 : 2;
    auto k_cand = fuse ? (lc == 2 ? pm_candidates<kNT, 2, PR, true, 1> : pm_candidates<kNT, 4, PR, true, 1>)
                       : (lc == 2 ? pm_candidates<kNT, 2, PR, false, 2> : pm_candidates<kNT, 4, PR, false, 2>);
    auto k_ref = lr == 2 ? pm_refine<kNT, 2, PR> : pm_refine<kNT, 4, PR>;
    const int bi = kTilePix * li, bc = gc * kTilePix * lc, br = kTilePix * lr;
    const dim3 g_init(persistent_grid(reinterpret_cast<const void*>(k_init), bi, 0, n_full, K_INIT));
    const dim3 g_cand(
        persistent_grid(reinterpret_cast<const void*>(k_cand), bc, 0, gc == 1 ? n_half : n_pairs,
                        K_CANDIDATES));
    const dim3 g_ref(
        persistent_grid(reinterpret_cast<const void*>(k_ref), br, ref_lds, n_half, K_REFINE));

    GMVS_LAUNCH(K_INIT, st, k_init, g_init, dim3(bi), 0, st, d_probs, d_tabs, queue, work);
    GMVS_LAUNCH_CHECK();
    for (int it = 0; it < iters; ++it) {
        for (int color = 0; color < 2; ++color) {
            const int slot = 1 + 2 * (2 * it + color);
            GMVS_LAUNCH(K_CANDIDATES, st, k_cand, g_cand, dim3(bc), 0, st, d_probs, d_tabs, color,
                        it, queue + kXcds * slot, work);
            GMVS_LAUNCH_CHECK();
            if (!fuse) {
                GMVS_LAUNCH(K_SELECT, st, pm_select, dim3(n_act), dim3(kBlock), sel_lds, st, d_probs,
                            d_tabs, color, it, nvmax, work);
                GMVS_LAUNCH_CHECK();
            }
            GMVS_LAUNCH(K_REFINE, st, k_ref, g_ref, dim3(br), ref_lds, st, d_probs, d_tabs, color,
                        nvmax, fuse ? 1 : 0, queue + kXcds * (slot + 1), work);
            GMVS_LAUNCH_CHECK();
        }
    }
    GMVS_LAUNCH(K_DEPTH_NORMAL, st, pm_depth_normal, dim3(n_pix), dim3(kBlock), 0, st, d_probs, d_tabs);
    GMVS_LAUNCH_CHECK();
    for (int color = 0; color < 2; ++color) {
        GMVS_LAUNCH(K_FILTER, st, pm_filter, dim3(n_act), dim3(kBlock), 0, st, d_probs, d_tabs, color);
        GMVS_LAUNCH_CHECK();
    }
    GMVS_LAUNCH(K_EXPORT, st, pm_export, dim3(n_pix), dim3(kBlock), 0, st, d_probs, d_tabs);
    GMVS_LAUNCH_CHECK();
    return GMVS_OK;
}

int pm_run_batch(const gmvs_pm_args* args, int n, void* workspace, size_t workspace_bytes,
                 void* stream) {
    GMVS_ARG_CHECK(args, "args is NULL");
    GMVS_ARG_CHECK(n >= 1 && n <= kMaxBatch, "a batch holds 1..%d problems (got %d)", kMaxBatch, n);
    std::vector<int> radius(n, 0);
    int nvmax = 1;
    for (int i = 0; i < n; ++i) {
        if (int rc = check_args(args[i], &radius[i])) return rc;
        GMVS_ARG_CHECK(args[i].precision == args[0].precision,
                       "all problems of a batch share one precision (problem %d)", i);
        GMVS_ARG_CHECK(args[i].params.max_iterations == args[0].params.max_iterations,
                       "all problems of a batch share max_iterations (problem %d)", i);
        nvmax = std::max(nvmax, args[i].num_images - 1);
    }
    {
        // pm_refine keeps per-view descriptors, work units and refinement
        // results in LDS (~1.4 KB per source view on top of the static part)
        int dev = 0, lds = 0;
        GMVS_HIP_CHECK(hipGetDevice(&dev));
        GMVS_HIP_CHECK(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        const size_t need = std::max(sizeof(RefineSmem<kNT, 2>), sizeof(RefineSmem<kNT, 4>)) +
                            refine_dyn_bytes(nvmax);
        if (need > (size_t)lds) {
            set_error("%d source views need %zu bytes of LDS per block; the device allows %d",
                      nvmax, need, lds);
            return GMVS_ERR_UNSUPPORTED;
        }
    }
    const BatchLayout B = batch_layout(args, n);
    GMVS_ARG_CHECK(workspace, "workspace is NULL");
    if (workspace_bytes < B.total) {
        set_error("workspace too small: %zu < %zu bytes", workspace_bytes, B.total);
        return GMVS_ERR_NOSPACE;
    }
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    char* ws = reinterpret_cast<char*>(workspace);
    switch (args[0].precision) {
        case GMVS_PREC_FP16:
            return run_batch<PREC_F16>(args, n, radius.data(), ws, B, st);
        case GMVS_PREC_BF16:
            return run_batch<PREC_BF16>(args, n, radius.data(), ws, B, st);
        default:
            return run_batch<PREC_F32>(args, n, radius.data(), ws, B, st);
    }
}

}  // namespace

template <int PR>
int cost_batch(int num_images, const gmvs_camera* cameras, const float* const* images,
               const gmvs_params& params, int radius, int n, const int32_t* xy, const float* planes,
               float* costs, char* ws, const PMLayout& L, hipStream_t st) {
    if (int rc = setup_views<PR>(num_images, cameras, images, nullptr, ws, L, st)) return rc;
    PMConst P = make_const(num_images, cameras, images[0],
                           reinterpret_cast<const ViewDesc*>(ws + L.views),
                           reinterpret_cast<const ViewF64*>(ws + L.views64), params, radius);
    P.cells = ws + L.quads;
    // the candidate kernel's lane-group width (the costs of a pass's hypotheses)
    const int lg = lanes_of(K_CANDIDATES, PR);
    const int nb = kTilePix * lg;
    auto k = lg == 2 ? pm_cost_batch<kNT, 2, PR> : pm_cost_batch<kNT, 4, PR>;
    GMVS_LAUNCH(K_COST_BATCH, st, k, dim3((n * lg + nb - 1) / nb), dim3(nb), 0, st, P, n, xy,
                reinterpret_cast<const float4*>(planes), costs);
    GMVS_LAUNCH_CHECK();
    return GMVS_OK;
}

extern "C" size_t gmvs_pm_workspace_bytes(int32_t height, int32_t width, int32_t num_images,
                                          int32_t precision) {
    if (height <= 0 || width <= 0 || num_images < 2 || !valid_precision(precision)) return 0;
    gmvs_camera cam;
    std::memset(&cam, 0, sizeof(cam));
    cam.height = height;
    cam.width = width;
    gmvs_pm_args a;
    std::memset(&a, 0, sizeof(a));
    a.cameras = &cam;
    a.num_images = num_images;
    a.precision = precision;
    return batch_layout(&a, 1).total;
}

extern "C" size_t gmvs_pm_batch_workspace_bytes(const gmvs_pm_args* args, int32_t n) {
    if (!args || n < 1 || n > kMaxBatch) return 0;
    for (int i = 0; i < n; ++i)
        if (!args[i].cameras || args[i].num_images < 2 || args[i].cameras[0].height <= 0 ||
            args[i].cameras[0].width <= 0 || !valid_precision(args[i].precision))
            return 0;
    return batch_layout(args, n).total;
}

extern "C" size_t gmvs_cost_workspace_bytes(int32_t height, int32_t width, int32_t num_images,
                                            int32_t precision) {
    if (height <= 0 || width <= 0 || num_images < 2 || !valid_precision(precision)) return 0;
    return pm_layout(height, width, num_images, false, precision).total;
}

extern "C" int gmvs_set_blocks_per_cu(int32_t kernel, int32_t blocks) {
    clear_error();
    GMVS_ARG_CHECK(kernel == GMVS_TUNE_INIT || kernel == GMVS_TUNE_CANDIDATES ||
                       kernel == GMVS_TUNE_REFINE,
                   "unknown tunable kernel %d", kernel);
    GMVS_ARG_CHECK(blocks >= 0 && blocks <= 64, "blocks per CU must be in [0, 64] (got %d)", blocks);
    const int kid = kernel == GMVS_TUNE_INIT ? K_INIT
                    : kernel == GMVS_TUNE_CANDIDATES ? K_CANDIDATES
                                                     : K_REFINE;
    g_blocks_per_cu[kid].store(blocks);
    return GMVS_OK;
}

extern "C" int gmvs_set_lane_width(int32_t kernel, int32_t lanes) {
    clear_error();
    GMVS_ARG_CHECK(kernel == GMVS_TUNE_INIT || kernel == GMVS_TUNE_CANDIDATES ||
                       kernel == GMVS_TUNE_REFINE,
                   "unknown tunable kernel %d", kernel);
    GMVS_ARG_CHECK(lanes == 0 || lanes == 2 || lanes == 4, "lanes must be 0, 2 or 4 (got %d)", lanes);
    const int kid = kernel == GMVS_TUNE_INIT ? K_INIT
                    : kernel == GMVS_TUNE_CANDIDATES ? K_CANDIDATES
                                                     : K_REFINE;
    g_lanes[kid].store(lanes);
    return GMVS_OK;
}

extern "C" int gmvs_set_fast_taps(int32_t enable) {
    clear_error();
    GMVS_ARG_CHECK(enable == 0 || enable == 1, "enable must be 0 or 1 (got %d)", enable);
    g_fast_taps.store(enable);
    return GMVS_OK;
}

extern "C" int gmvs_set_select_fusion(int32_t enable) {
    clear_error();
    GMVS_ARG_CHECK(enable == 0 || enable == 1, "enable must be 0 or 1 (got %d)", enable);
    g_fuse_select.store(enable);
    return GMVS_OK;
}

extern "C" int gmvs_pm_run(const gmvs_pm_args* args, void* workspace, size_t workspace_bytes,
                           void* stream) {
    clear_error();
    return pm_run_batch(args, 1, workspace, workspace_bytes, stream);
}

extern "C" int gmvs_pm_run_batch(const gmvs_pm_args* args, int32_t n, void* workspace,
                                 size_t workspace_bytes, void* stream) {
    clear_error();
    return pm_run_batch(args, n, workspace, workspace_bytes, stream);
}

extern "C" int gmvs_cost_batch(int32_t num_images, const gmvs_camera* cameras,
                               const float* const* images, const gmvs_params* params,
                               int32_t precision, int32_t n, const int32_t* xy,
                               const float* planes, float* costs, void* workspace,
                               size_t workspace_bytes, void* stream) {
    clear_error();
    if (int rc = check_cameras(num_images, cameras)) return rc;
    GMVS_ARG_CHECK(images && params, "images/params is NULL");
    for (int i = 0; i < num_images; ++i) GMVS_ARG_CHECK(images[i], "images[%d] is NULL", i);
    GMVS_ARG_CHECK(valid_precision(precision), "unsupported precision %d", precision);
    GMVS_ARG_CHECK(n >= 0, "n must be >= 0");
    if (n == 0) return GMVS_OK;
    GMVS_ARG_CHECK(xy && planes && costs, "xy/planes/costs is NULL");
    int radius = 0;
    const int nt = taps_per_axis(*params, &radius);
    if (nt != 6) {
        set_error("patch geometry gives %d taps per axis; this build is specialised for 6", nt);
        return GMVS_ERR_UNSUPPORTED;
    }
    const int h = cameras[0].height, w = cameras[0].width;
    const PMLayout L = pm_layout(h, w, num_images, false, precision);
    GMVS_ARG_CHECK(workspace, "workspace is NULL");
    if (workspace_bytes < L.total) {
        set_error("workspace too small: %zu < %zu bytes", workspace_bytes, L.total);
        return GMVS_ERR_NOSPACE;
    }
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    char* ws = reinterpret_cast<char*>(workspace);
    switch (precision) {
        case GMVS_PREC_FP16:
            return cost_batch<PREC_F16>(num_images, cameras, images, *params, radius, n, xy, planes,
                                        costs, ws, L, st);
        case GMVS_PREC_BF16:
            return cost_batch<PREC_BF16>(num_images, cameras, images, *params, radius, n, xy, planes,
                                         costs, ws, L, st);
        default:
            return cost_batch<PREC_F32>(num_images, cameras, images, *params, radius, n, xy, planes,
                                        costs, ws, L, st);
    }
}
