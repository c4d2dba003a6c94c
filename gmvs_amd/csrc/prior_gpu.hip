// Device construction of the ACMP planar prior between the two first-stage
// passes -- the GPU form of gmvs_prior_build (prior.cpp): get_support_points,
// the superpixel filter of triangulation, process_mask, get_prior_plane_params
// and cuda_planar_prior_init (gmvs/lib/src/planar_initer.cu:195-420,686-745).
//
//   prior_support    one thread per 5x5 cell: the cell's lowest-cost pixel
//                    (cost < 2, first minimum in the reference's column-major
//                    scan), kept when its cost is < 0.1;
//   dt_points        the kept cells numbered in the reference's cell order
//                    (a one-block scan, prior_rank), cell -> point table;
//   dt_star          one thread per point: its Delaunay star (delaunay.h),
//                    the triangles whose lowest-index vertex it is into its
//                    slots; dt_emit, after a scan of the counts, lists them
//                    in the host construction's order;
//   prior_triangles  one thread per triangle: the superpixel-filter walk and
//                    the 3-point plane fit (FP64);
//   prior_rank       one block: retained triangles numbered in order;
//   prior_raster     one thread per retained triangle: the barycentric walk
//                    writing its number + 1 into the pixels it visits (the
//                    highest number wins, as the sequential loop's last
//                    writer does);
//   prior_output     one thread per pixel: the prior plane and mask, dropping
//                    pixels whose prior depth leaves [depth_min, depth_max].
//
// No host round trip: triangle counts stay on the device, and the
// per-triangle kernels run over the largest possible count, reading the
// actual one. Results equal gmvs_prior_build's bit for bit: the file is compiled with
// -ffp-contract=off and correctly rounded division and square root, and each
// expression keeps the host code's types and order (the walk mixes float and
// double as the reference's loop does).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <utility>
#include <vector>

#include "abi_common.h"
#include "delaunay.h"

#pragma clang fp contract(off)

namespace gmvs {

namespace {

constexpr int kBlock = 256;
constexpr int kCell = 5;           // support-point cell (planar_initer.cu:228)
constexpr int kRankThreads = 1024;

size_t align_up(size_t x) { return (x + 255) & ~size_t(255); }

// Largest star walk (triangles around one point) before a point is reported
// as failing; a valid star has far fewer.
constexpr int kMaxStarSteps = 4096;
// Triangles one point may own (be the lowest-index vertex of); the support
// points of a 1600x1200 view own at most ~35.
constexpr int kMaxOwned = 64;

struct PriorLayout {
    size_t cells, tris, keep, rank, tri_of, planes;
    size_t pidx, px, py, pkeep, prank, cnt, off, tidx, slots, bcnt, err, total;
    int ncells, nblocks, max_tri;
};

PriorLayout prior_layout(int h, int w) {
    PriorLayout L;
    const int ncx = (w + kCell - 1) / kCell, ncy = (h + kCell - 1) / kCell;
    L.ncells = ncx * ncy;
    L.nblocks = ((ncx + dt::kBlk - 1) / dt::kBlk) * ((ncy + dt::kBlk - 1) / dt::kBlk);
    // a planar triangulation of n points has at most 2n - 5 triangles
    L.max_tri = 2 * L.ncells + 8;
    size_t o = 0;
    L.cells = o;
    o = align_up(o + sizeof(int2) * (size_t)L.ncells);
    L.tris = o;
    o = align_up(o + sizeof(int) * 6 * (size_t)L.max_tri);
    L.keep = o;
    o = align_up(o + sizeof(int) * (size_t)L.max_tri);
    L.rank = o;
    o = align_up(o + sizeof(int) * ((size_t)L.max_tri + 1));
    L.tri_of = o;
    o = align_up(o + sizeof(int) * (size_t)L.max_tri);
    L.planes = o;
    o = align_up(o + sizeof(float4) * (size_t)L.max_tri);
    const size_t nc = (size_t)L.ncells;
    L.pidx = o;
    o = align_up(o + sizeof(int) * nc);
    L.px = o;
    o = align_up(o + sizeof(int) * nc);
    L.py = o;
    o = align_up(o + sizeof(int) * nc);
    L.pkeep = o;
    o = align_up(o + sizeof(int) * nc);
    L.prank = o;
    o = align_up(o + sizeof(int) * (nc + 1));
    L.cnt = o;
    o = align_up(o + sizeof(int) * nc);
    L.off = o;
    o = align_up(o + sizeof(int) * (nc + 1));
    L.tidx = o;
    o = align_up(o + sizeof(int) * 3 * (size_t)L.max_tri);
    L.slots = o;
    o = align_up(o + sizeof(int2) * kMaxOwned * nc);
    L.bcnt = o;
    o = align_up(o + sizeof(int) * (size_t)L.nblocks);
    L.err = o;  // [0] failure flags, [1] triangle count
    o = align_up(o + sizeof(int) * 4);
    L.total = o;
    return L;
}



// get_support_points (planar_initer.cu:225-256): cells in the reference's
// order, columns outer (cell = cx * ncy + cy).
__global__ void prior_support(const float* __restrict__ cost, int h, int w, int ncy, int ncells,
                              int2* __restrict__ cells) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ncells) return;
    const int col = (i / ncy) * kCell, row = (i % ncy) * kCell;
    const int cb = min(w, col + kCell), rb = min(h, row + kCell);
    float mc = 2.0f;
    int bx = 0, by = 0;
    for (int c = col; c < cb; ++c)
        for (int r = row; r < rb; ++r) {
            const float v = cost[r * w + c];
            if (v < 2.0f && mc > v) {
                bx = c;
                by = r;
                mc = v;
            }
        }
    cells[i] = (mc < 0.1f) ? make_int2(bx, by) : make_int2(-1, -1);
}

// Points per block of dt::kBlk x dt::kBlk cells (dt::Grid::bcnt), one
// thread per block.
__global__ void dt_blocks(const int* __restrict__ pidx, int ncx, int ncy, int* __restrict__ bcnt) {
    const int nby = (ncy + dt::kBlk - 1) / dt::kBlk;
    const int nb = ((ncx + dt::kBlk - 1) / dt::kBlk) * nby;
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    const int bx = b / nby, by = b - bx * nby;
    const int x1 = min(ncx, (bx + 1) * dt::kBlk), y1 = min(ncy, (by + 1) * dt::kBlk);
    int n = 0;
    for (int cx = bx * dt::kBlk; cx < x1; ++cx)
        for (int cy = by * dt::kBlk; cy < y1; ++cy) n += pidx[cx * ncy + cy] >= 0;
    bcnt[b] = n;
}

// Kept-cell flags of the support cells (cell_xy x < 0: none).
__global__ void dt_flags(const int2* __restrict__ cells, int ncells, int* __restrict__ keep) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ncells) keep[i] = cells[i].x >= 0 ? 1 : 0;
}

// Points in cell order: pidx[cell] = number of the kept cell (rank) or -1.
__global__ void dt_points(const int2* __restrict__ cells, const int* __restrict__ rank, int ncells,
                          int* __restrict__ pidx, int* __restrict__ px, int* __restrict__ py) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ncells) return;
    const int2 c = cells[i];
    if (c.x < 0) {
        pidx[i] = -1;
        return;
    }
    const int p = rank[i];
    pidx[i] = p;
    px[p] = c.x;
    py[p] = c.y;
}

// Lanes per point of the star walk: its scans (a 7x7-cell square per step,
// and the rare scans of a large circle or half-plane, thousands of cells)
// are shared by the group.
constexpr int kStarLanes = 16;

// One group of kStarLanes lanes per point (npts = rank[ncells], read here):
// the star walk. The triangles the point owns go to its kMaxOwned slots,
// sorted by their second vertex, and their number to cnt[p] (lane 0).
__global__ void dt_star(dt::Grid g, const int* __restrict__ npts_dev, int2* __restrict__ slots,
                        int* __restrict__ cnt, int* __restrict__ err) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int p = t / kStarLanes;
    if (p >= *npts_dev) return;  // the whole group
    const dt::Lanes<kStarLanes> ln{t % kStarLanes};
    int q[kMaxOwned], r[kMaxOwned];
    int n = 0;
    const bool ok = dt::star(g, p, kMaxStarSteps, ln, [&](int a, int b) {
        if (ln.lane == 0 && n < kMaxOwned) {
            // insertion by the second vertex
            int k = n;
            while (k > 0 && q[k - 1] > a) {
                q[k] = q[k - 1];
                r[k] = r[k - 1];
                --k;
            }
            q[k] = a;
            r[k] = b;
        }
        ++n;
    });
    if (ln.lane != 0) return;
    if (!ok) atomicOr(err, 1);
    if (n > kMaxOwned) {
        atomicOr(err, 2);
        n = kMaxOwned;
    }
    int2* out = slots + (size_t)p * kMaxOwned;
    for (int k = 0; k < n; ++k) out[k] = make_int2(q[k], r[k]);
    cnt[p] = n;
}

// One thread per point: its triangles from the slots to their places in the
// list (off = exclusive scan of cnt), as index triples and coordinates.
__global__ void dt_emit(dt::Grid g, const int* __restrict__ npts_dev, const int2* __restrict__ slots,
                        const int* __restrict__ cnt, const int* __restrict__ off, int max_tri,
                        int* __restrict__ tidx, int* __restrict__ tris, int* __restrict__ err) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= *npts_dev) return;
    const int n = cnt[p], o = off[p];
    if (o + n > max_tri) {
        atomicOr(err, 4);
        return;
    }
    const int2* in = slots + (size_t)p * kMaxOwned;
    for (int k = 0; k < n; ++k) {
        const int t = o + k, qk = in[k].x, rk = in[k].y;
        tidx[3 * t] = p;
        tidx[3 * t + 1] = qk;
        tidx[3 * t + 2] = rk;
        tris[6 * t] = g.px[p];
        tris[6 * t + 1] = g.py[p];
        tris[6 * t + 2] = g.px[qk];
        tris[6 * t + 3] = g.py[qk];
        tris[6 * t + 4] = g.px[rk];
        tris[6 * t + 5] = g.py[rk];
    }
}

__device__ inline float edge_len(int dx, int dy) {
    return (float)sqrt((double)dx * dx + (double)dy * dy);
}

// The reference's barycentric walk over a triangle (planar_initer.cu:313-324
// and 372-380): float steps of 1 / longest edge, positions mixing float and
// double as the host loop's expression does. visit(x, y) returns false to
// stop.
template <class F>
__device__ inline void walk_triangle(int ax, int ay, int bx, int by, int cx, int cy, F&& visit) {
    const float L01 = edge_len(ax - bx, ay - by);
    const float L02 = edge_len(ax - cx, ay - cy);
    const float L12 = edge_len(bx - cx, by - cy);
    const float m12 = (L02 < L12) ? L12 : L02;
    const float mx = (L01 < m12) ? m12 : L01;
    const float step = (float)(1.0 / (double)mx);
    for (float p = 0; (double)p < 1.0; p += step) {
        for (float q = 0; (double)q < 1.0 - (double)p; q += step) {
            const double rest = 1.0 - (double)p - (double)q;
            const int x = (int)((double)(p * (float)ax + q * (float)bx) + rest * (double)cx);
            const int y = (int)((double)(p * (float)ay + q * (float)by) + rest * (double)cy);
            if (!visit(x, y)) return;
        }
    }
}

// Superpixel filter (planar_initer.cu:362-410, int32 labels) and the plane of
// each triangle (get_prior_plane_params, :700-745).
__global__ void prior_triangles(const int* __restrict__ tris, const int* __restrict__ ntri,
                                int max_tri, const int32_t* __restrict__ labels, int filter,
                                const float* __restrict__ depth, int h, int w, gmvs_camera cam,
                                int* __restrict__ keep, float4* __restrict__ planes) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= min(*ntri, max_tri)) return;
    const int* v = tris + 6 * t;
    bool retain = true;
    if (filter) {
        float count = 0.f, zero = 0.f;
        int32_t seg = 0;
        walk_triangle(v[0], v[1], v[2], v[3], v[4], v[5], [&](int x, int y) {
            // the walk stays inside the triangle's bounding box; the clamp
            // only guards the read
            const int32_t s = labels[min(max(y, 0), h - 1) * w + min(max(x, 0), w - 1)];
            count += 1.0f;
            if (s == 0) zero += 1.0f;
            if (seg == 0)
                seg = s;
            else if (s != 0 && s != seg)
                retain = false;
            return retain;
        });
        if (count > 0.f && zero / count > 0.1f) retain = false;
    }
    keep[t] = retain ? 1 : 0;
    const float* K = cam.K;
    double X[3][3];
    for (int k = 0; k < 3; ++k) {
        const int px = v[2 * k], py = v[2 * k + 1];
        const double d = depth[py * w + px];
        X[k][0] = d * (px - double(K[2])) / double(K[0]);
        X[k][1] = d * (py - double(K[5])) / double(K[4]);
        X[k][2] = d;
    }
    const double e1[3] = {X[1][0] - X[0][0], X[1][1] - X[0][1], X[1][2] - X[0][2]};
    const double e2[3] = {X[2][0] - X[0][0], X[2][1] - X[0][1], X[2][2] - X[0][2]};
    const double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                         e1[0] * e2[1] - e1[1] * e2[0]};
    double nn = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    const double dd = -(n[0] * X[0][0] + n[1] * X[0][1] + n[2] * X[0][2]);
    if (dd < 0) nn = -nn;
    planes[t] = make_float4(float(n[0] / nn), float(n[1] / nn), float(n[2] / nn), float(dd / nn));
}

// Exclusive prefix sum of keep[0..n) into rank[0..n], rank[n] = total (also
// at *total when given); one block. With n_dev, n = min(*n_dev, n_arg)
// (n_arg bounds the arrays). The block walks the array in chunks of
// kRankThreads * kRankPer elements, each thread kRankPer consecutive ones
// (neighbouring threads read neighbouring words): the threads' sums are
// scanned within each wave by shuffles, the wave totals by the first wave,
// and a running carry joins the chunks.
constexpr int kRankPer = 4;
constexpr int kRankWaves = kRankThreads / 64;

__global__ __launch_bounds__(kRankThreads) void prior_rank(const int* __restrict__ keep, int n_arg,
                                                           const int* __restrict__ n_dev,
                                                           int* __restrict__ rank,
                                                           int* __restrict__ total) {
    __shared__ int s_wave[kRankWaves];
    const int n = n_dev ? min(*n_dev, n_arg) : n_arg;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int carry = 0;
    for (int base = 0; base < n; base += kRankThreads * kRankPer) {
        const int lo = base + (int)threadIdx.x * kRankPer;
        int v[kRankPer];
        int sum = 0;
#pragma unroll
        for (int j = 0; j < kRankPer; ++j) {
            v[j] = lo + j < n ? keep[lo + j] : 0;
            sum += v[j];
        }
        int inc = sum;  // inclusive scan over the wave
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int o = __shfl_up(inc, off);
            if (lane >= off) inc += o;
        }
        if (lane == 63) s_wave[wave] = inc;
        __syncthreads();
        if (wave == 0) {
            int w = lane < kRankWaves ? s_wave[lane] : 0;  // inclusive over the waves
#pragma unroll
            for (int off = 1; off < kRankWaves; off <<= 1) {
                const int o = __shfl_up(w, off);
                if (lane >= off) w += o;
            }
            if (lane < kRankWaves) s_wave[lane] = w;
        }
        __syncthreads();
        int run = carry + (wave > 0 ? s_wave[wave - 1] : 0) + inc - sum;
#pragma unroll
        for (int j = 0; j < kRankPer; ++j) {
            if (lo + j < n) rank[lo + j] = run;
            run += v[j];
        }
        carry += s_wave[kRankWaves - 1];
        __syncthreads();  // s_wave is rewritten by the next chunk
    }
    if (threadIdx.x == 0) {
        rank[n] = carry;
        if (total) *total = carry;
    }
}

__global__ void prior_zero(uint32_t* __restrict__ p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// process_mask (planar_initer.cu:290-324): retained triangle number + 1 into
// the pixels of its walk; later triangles overwrite earlier ones there.
__global__ void prior_raster(const int* __restrict__ tris, const int* __restrict__ ntri,
                             int max_tri, const int* __restrict__ keep, const int* __restrict__ rank,
                             int h, int w, uint32_t* __restrict__ mask) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= min(*ntri, max_tri) || !keep[t]) return;
    const int* v = tris + 6 * t;
    const uint32_t id = (uint32_t)rank[t] + 1u;
    walk_triangle(v[0], v[1], v[2], v[3], v[4], v[5], [&](int x, int y) {
        if (x >= 0 && x < w && y >= 0 && y < h) atomicMax(&mask[(size_t)y * w + x], id);
        return true;
    });
}

// cuda_planar_prior_init and the prior depth check (planar_initer.cu:195-223,
// 330-344): planes of the retained triangles by number.
// mask and plane_masks may be one buffer: each thread reads its own entry
// before writing it.
__global__ void prior_output(const uint32_t* mask, const float4* __restrict__ planes,
                             const int* __restrict__ tri_of, int h, int w, gmvs_camera cam,
                             float dmin, float dmax, float4* __restrict__ prior_planes,
                             uint32_t* plane_masks) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= h * w) return;
    const int y = i / w, x = i - y * w;
    const uint32_t m = mask[i];
    float4 out = make_float4(0.f, 0.f, 0.f, 0.f);
    uint32_t mo = 0u;
    if (m != 0u) {
        const float4 pl = planes[tri_of[m - 1]];
        const float* K = cam.K;
        const float d = -pl.w / ((x - K[2]) * pl.x / K[0] + (y - K[5]) * pl.y / K[4] + pl.z);
        if (d <= dmax && d >= dmin) {
            mo = m;
            out = pl;
        }
    }
    prior_planes[i] = out;
    plane_masks[i] = mo;
}

// tri_of[rank[t]] = t for retained triangles (number -> triangle index).
__global__ void prior_index(const int* __restrict__ keep, const int* __restrict__ rank,
                            const int* __restrict__ ntri, int max_tri, int* __restrict__ tri_of) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < min(*ntri, max_tri) && keep[t]) tri_of[rank[t]] = t;
}

__global__ void dt_clear(int* __restrict__ err) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *err = 0;
}

}  // namespace
}  // namespace gmvs

using namespace gmvs;

extern "C" size_t gmvs_prior_device_workspace_bytes(int32_t height, int32_t width) {
    if (height <= 0 || width <= 0) return 0;
    return prior_layout(height, width).total;
}

// The Delaunay triangles of the support cells on the device (no host
// synchronisation): index triples at tidx and coordinates at tris; at L.err
// the failure flags ([0]: 1 a star walk too long, 2 a point owning more than
// kMaxOwned triangles, 4 more than max_tri triangles) and the count ([1]).
static int run_delaunay(const PriorLayout& L, char* ws, const int2* cells, int ncx, int ncy,
                        hipStream_t st) {
    int* pkeep = reinterpret_cast<int*>(ws + L.pkeep);
    int* prank = reinterpret_cast<int*>(ws + L.prank);
    int* pidx = reinterpret_cast<int*>(ws + L.pidx);
    int* px = reinterpret_cast<int*>(ws + L.px);
    int* py = reinterpret_cast<int*>(ws + L.py);
    int* cnt = reinterpret_cast<int*>(ws + L.cnt);
    int* off = reinterpret_cast<int*>(ws + L.off);
    int* tidx = reinterpret_cast<int*>(ws + L.tidx);
    int* tris = reinterpret_cast<int*>(ws + L.tris);
    int* err = reinterpret_cast<int*>(ws + L.err);
    const int nb = (L.ncells + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(dt_clear, dim3(1), dim3(64), 0, st, err);
    GMVS_LAUNCH_CHECK();
    hipLaunchKernelGGL(dt_flags, dim3(nb), dim3(kBlock), 0, st, cells, L.ncells, pkeep);
    GMVS_LAUNCH_CHECK();
    hipLaunchKernelGGL(prior_rank, dim3(1), dim3(kRankThreads), 0, st, pkeep, L.ncells,
                       (const int*)nullptr, prank, (int*)nullptr);
    GMVS_LAUNCH_CHECK();
    hipLaunchKernelGGL(dt_points, dim3(nb), dim3(kBlock), 0, st, cells, prank, L.ncells, pidx, px, py);
    GMVS_LAUNCH_CHECK();
    int* bcnt = reinterpret_cast<int*>(ws + L.bcnt);
    hipLaunchKernelGGL(dt_blocks, dim3((L.nblocks + kBlock - 1) / kBlock), dim3(kBlock), 0, st, pidx,
                       ncx, ncy, bcnt);
    GMVS_LAUNCH_CHECK();
    dt::Grid g;
    g.pidx = pidx;
    g.px = px;
    g.py = py;
    g.bcnt = bcnt;
    g.ncx = ncx;
    g.ncy = ncy;
    g.cell = kCell;
    g.nby = (ncy + dt::kBlk - 1) / dt::kBlk;
    const int* npts = prank + L.ncells;
    int2* slots = reinterpret_cast<int2*>(ws + L.slots);
    const int sb = (int)(((size_t)L.ncells * kStarLanes + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(dt_star, dim3(sb), dim3(kBlock), 0, st, g, npts, slots, cnt, err);
    GMVS_LAUNCH_CHECK();
    hipLaunchKernelGGL(prior_rank, dim3(1), dim3(kRankThreads), 0, st, cnt, L.ncells, npts, off,
                       err + 1);
    GMVS_LAUNCH_CHECK();
    hipLaunchKernelGGL(dt_emit, dim3(nb), dim3(kBlock), 0, st, g, npts, slots, cnt, off, L.max_tri,
                       tidx, tris, err);
    GMVS_LAUNCH_CHECK();
    return GMVS_OK;
}

extern "C" int gmvs_prior_build_device(const gmvs_camera* ref_camera, const float* depth,
                                       const float* cost, const int32_t* labels,
                                       const gmvs_params* params, float* prior_planes,
                                       uint32_t* plane_masks, int32_t* out_num_triangles,
                                       void* workspace, size_t workspace_bytes, void* stream) {
    clear_error();
    GMVS_ARG_CHECK(ref_camera && depth && cost && params && prior_planes && plane_masks,
                   "prior_build_device: NULL argument");
    const int h = ref_camera->height, w = ref_camera->width;
    GMVS_ARG_CHECK(h > 0 && w > 0 && w < 65536 && h < 65536, "prior_build_device: bad size %dx%d",
                   w, h);
    GMVS_ARG_CHECK(!params->superpixel_filter || labels,
                   "prior_build_device: superpixel_filter needs labels");
    const PriorLayout L = prior_layout(h, w);
    GMVS_ARG_CHECK(workspace, "workspace is NULL");
    if (workspace_bytes < L.total) {
        set_error("workspace too small: %zu < %zu bytes", workspace_bytes, L.total);
        return GMVS_ERR_NOSPACE;
    }
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    char* ws = static_cast<char*>(workspace);
    int2* cells = reinterpret_cast<int2*>(ws + L.cells);
    int* tris = reinterpret_cast<int*>(ws + L.tris);
    int* keep = reinterpret_cast<int*>(ws + L.keep);
    int* rank = reinterpret_cast<int*>(ws + L.rank);
    float4* planes = reinterpret_cast<float4*>(ws + L.planes);
    const int ncy = (h + kCell - 1) / kCell;

    const int ncx = (w + kCell - 1) / kCell;
    int* err = reinterpret_cast<int*>(ws + L.err);
    const int* ntri = err + 1;

    // support points, their Delaunay triangles (on the device; the count
    // stays there, and the per-triangle kernels run over L.max_tri threads)
    hipLaunchKernelGGL(prior_support, dim3((L.ncells + kBlock - 1) / kBlock), dim3(kBlock), 0, st,
                       cost, h, w, ncy, L.ncells, cells);
    GMVS_LAUNCH_CHECK();
    const int rc = run_delaunay(L, ws, cells, ncx, ncy, st);
    if (rc != GMVS_OK) return rc;

    // filter + planes, numbering, rasterisation, per-pixel output; the mask
    // is assembled in plane_masks, then rewritten in place by prior_output
    // (each pixel reads only its own entry)
    const int tb = (L.max_tri + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(prior_triangles, dim3(tb), dim3(kBlock), 0, st, tris, ntri, L.max_tri, labels,
                       params->superpixel_filter ? 1 : 0, depth, h, w, *ref_camera, keep, planes);
    GMVS_LAUNCH_CHECK();
    hipLaunchKernelGGL(prior_rank, dim3(1), dim3(kRankThreads), 0, st, keep, L.max_tri, ntri, rank,
                       err + 2);
    GMVS_LAUNCH_CHECK();
    const int pb = (h * w + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(prior_zero, dim3(pb), dim3(kBlock), 0, st, plane_masks, h * w);
    GMVS_LAUNCH_CHECK();
    int* tri_of = reinterpret_cast<int*>(ws + L.tri_of);
    hipLaunchKernelGGL(prior_raster, dim3(tb), dim3(kBlock), 0, st, tris, ntri, L.max_tri, keep,
                       rank, h, w, plane_masks);
    GMVS_LAUNCH_CHECK();
    hipLaunchKernelGGL(prior_index, dim3(tb), dim3(kBlock), 0, st, keep, rank, ntri, L.max_tri,
                       tri_of);
    GMVS_LAUNCH_CHECK();
    hipLaunchKernelGGL(prior_output, dim3(pb), dim3(kBlock), 0, st, plane_masks, planes, tri_of, h,
                       w, *ref_camera, params->depth_min, params->depth_max,
                       reinterpret_cast<float4*>(prior_planes), plane_masks);
    GMVS_LAUNCH_CHECK();
    // the number of retained triangles and the failure flags: the only
    // synchronisation, and only when the caller asks for the count
    if (!out_num_triangles) return GMVS_OK;
    int hv[4] = {0, 0, 0, 0};
    GMVS_HIP_CHECK(hipMemcpyAsync(hv, err, sizeof(hv), hipMemcpyDeviceToHost, st));
    GMVS_HIP_CHECK(hipStreamSynchronize(st));
    if (hv[0] != 0) {
        set_error("prior_build_device: Delaunay construction failed (flags %d)", hv[0]);
        return GMVS_ERR_INTERNAL;
    }
    const int nkeep = hv[2];
    if (out_num_triangles) *out_num_triangles = nkeep;
    return GMVS_OK;
}

namespace gmvs {
namespace {

// cell_xy points inside their cells (err[0] |= 8 otherwise).
__global__ void dt_check_cells(const int2* __restrict__ cells, int ncy, int ncells,
                               int* __restrict__ err) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ncells) return;
    const int2 c = cells[i];
    if (c.x < 0 && c.y < 0) return;
    if (c.x < 0 || c.y < 0 || c.x / kCell != i / ncy || c.y / kCell != i % ncy) atomicOr(err, 8);
}

}  // namespace
}  // namespace gmvs

extern "C" int gmvs_delaunay_cells_device(const int32_t* cell_xy, int32_t ncx, int32_t ncy,
                                          int32_t* tri_out, int32_t max_tri, int32_t* num_tri,
                                          void* workspace, size_t workspace_bytes, void* stream) {
    clear_error();
    GMVS_ARG_CHECK(ncx > 0 && ncy > 0 && (int64_t)ncx * kCell < 65536 && (int64_t)ncy * kCell < 65536,
                   "delaunay_cells_device: bad grid %dx%d", ncx, ncy);
    GMVS_ARG_CHECK(cell_xy && num_tri && (tri_out || max_tri == 0) && max_tri >= 0,
                   "delaunay_cells_device: bad arguments");
    const PriorLayout L = prior_layout(ncy * kCell, ncx * kCell);
    GMVS_ARG_CHECK(workspace, "workspace is NULL");
    if (workspace_bytes < L.total) {
        set_error("workspace too small: %zu < %zu bytes", workspace_bytes, L.total);
        return GMVS_ERR_NOSPACE;
    }
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    char* ws = static_cast<char*>(workspace);
    const int2* cells = reinterpret_cast<const int2*>(cell_xy);
    int* err = reinterpret_cast<int*>(ws + L.err);
    const int rc = run_delaunay(L, ws, cells, ncx, ncy, st);
    if (rc != GMVS_OK) return rc;
    hipLaunchKernelGGL(dt_check_cells, dim3((L.ncells + kBlock - 1) / kBlock), dim3(kBlock), 0, st,
                       cells, ncy, L.ncells, err);
    GMVS_LAUNCH_CHECK();
    int hv[4] = {0, 0, 0, 0};
    GMVS_HIP_CHECK(hipMemcpyAsync(hv, err, sizeof(hv), hipMemcpyDeviceToHost, st));
    GMVS_HIP_CHECK(hipStreamSynchronize(st));
    GMVS_ARG_CHECK((hv[0] & 8) == 0, "delaunay_cells_device: a point lies outside its cell");
    if (hv[0] != 0) {
        set_error("delaunay_cells_device: construction failed (flags %d)", hv[0]);
        return GMVS_ERR_INTERNAL;
    }
    *num_tri = hv[1];
    if (hv[1] > max_tri) {
        set_error("delaunay_cells_device: %d triangles > max_tri %d", hv[1], max_tri);
        return GMVS_ERR_NOSPACE;
    }
    if (hv[1] > 0)
        GMVS_HIP_CHECK(hipMemcpyAsync(tri_out, ws + L.tidx, sizeof(int) * 3 * (size_t)hv[1],
                                      hipMemcpyDeviceToDevice, st));
    GMVS_HIP_CHECK(hipStreamSynchronize(st));
    return GMVS_OK;
}

extern "C" int gmvs_delaunay_cells_host(const int32_t* cell_xy, int32_t ncx, int32_t ncy,
                                        int32_t* tri_out, int32_t max_tri, int32_t* num_tri) {
    clear_error();
    GMVS_ARG_CHECK(ncx > 0 && ncy > 0 && (int64_t)ncx * kCell < 65536 && (int64_t)ncy * kCell < 65536,
                   "delaunay_cells_host: bad grid %dx%d", ncx, ncy);
    GMVS_ARG_CHECK(cell_xy && num_tri && (tri_out || max_tri == 0) && max_tri >= 0,
                   "delaunay_cells_host: bad arguments");
    const int ncells = ncx * ncy;
    std::vector<int> pidx(ncells, -1), px, py;
    for (int i = 0; i < ncells; ++i) {
        const int x = cell_xy[2 * i], y = cell_xy[2 * i + 1];
        if (x < 0 && y < 0) continue;
        GMVS_ARG_CHECK(x >= 0 && y >= 0 && x / kCell == i / ncy && y / kCell == i % ncy,
                       "delaunay_cells_host: cell %d holds (%d, %d) outside it", i, x, y);
        pidx[i] = (int)px.size();
        px.push_back(x);
        py.push_back(y);
    }
    const int nby = (ncy + dt::kBlk - 1) / dt::kBlk;
    std::vector<int> bcnt((size_t)((ncx + dt::kBlk - 1) / dt::kBlk) * nby);
    dt::block_counts(pidx.data(), ncx, ncy, bcnt.data());
    dt::Grid g;
    g.pidx = pidx.data();
    g.px = px.data();
    g.py = py.data();
    g.bcnt = bcnt.data();
    g.ncx = ncx;
    g.ncy = ncy;
    g.cell = kCell;
    g.nby = nby;
    std::vector<int> out;
    const int npts = (int)px.size();
    for (int p = 0; p < npts; ++p) {
        std::vector<std::pair<int, int>> own;
        const bool ok = dt::star(g, p, kMaxStarSteps, dt::Lanes<1>{0},
                                 [&](int q, int r) { own.push_back({q, r}); });
        if (!ok) {
            set_error("delaunay_cells_host: star walk of point %d too long", p);
            return GMVS_ERR_INTERNAL;
        }
        std::sort(own.begin(), own.end());
        for (const auto& t : own) {
            out.push_back(p);
            out.push_back(t.first);
            out.push_back(t.second);
        }
    }
    const int nt = (int)(out.size() / 3);
    *num_tri = nt;
    if (nt > max_tri) {
        set_error("delaunay_cells_host: %d triangles > max_tri %d", nt, max_tri);
        return GMVS_ERR_NOSPACE;
    }
    std::copy(out.begin(), out.end(), tri_out);
    return GMVS_OK;
}
