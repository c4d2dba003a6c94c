// Device construction of the ACMP planar prior between the two first-stage
// passes -- the GPU form of gmvs_prior_build (prior.cpp): get_support_points,
// the superpixel filter of triangulation, process_mask, get_prior_plane_params
// and cuda_planar_prior_init (gmvs/lib/src/planar_initer.cu:195-420,686-745).
//
//   prior_support    one thread per 5x5 cell: the cell's lowest-cost pixel
//                    (cost < 2, first minimum in the reference's column-major
//                    scan), kept when its cost is < 0.1;
//   (host)           the kept cells in the reference's cell order, Delaunay
//                    (the host triangulation of prior.cpp: exact predicates,
//                    a sequential algorithm);
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
// Results equal gmvs_prior_build's bit for bit: the file is compiled with
// -ffp-contract=off and correctly rounded division and square root, and each
// expression keeps the host code's types and order (the walk mixes float and
// double as the reference's loop does).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "abi_common.h"

#pragma clang fp contract(off)

namespace gmvs {

std::vector<int> host_delaunay(const int32_t* xy, int n);  // prior.cpp

namespace {

constexpr int kBlock = 256;
constexpr int kCell = 5;           // support-point cell (planar_initer.cu:228)
constexpr int kRankThreads = 1024;

size_t align_up(size_t x) { return (x + 255) & ~size_t(255); }

struct PriorLayout {
    size_t cells, tris, keep, rank, tri_of, planes, total;
    int ncells, max_tri;
};

PriorLayout prior_layout(int h, int w) {
    PriorLayout L;
    const int ncx = (w + kCell - 1) / kCell, ncy = (h + kCell - 1) / kCell;
    L.ncells = ncx * ncy;
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
__global__ void prior_triangles(const int* __restrict__ tris, int ntri, const int32_t* __restrict__ labels,
                                int filter, const float* __restrict__ depth, int h, int w,
                                gmvs_camera cam, int* __restrict__ keep, float4* __restrict__ planes) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntri) return;
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

// Exclusive prefix sum of keep[0..n) into rank[0..n], rank[n] = total; one
// block, each thread a contiguous chunk.
__global__ void prior_rank(const int* __restrict__ keep, int n, int* __restrict__ rank) {
    __shared__ int s[kRankThreads];
    const int per = (n + kRankThreads - 1) / kRankThreads;
    const int lo = min(n, (int)threadIdx.x * per), hi = min(n, lo + per);
    int sum = 0;
    for (int i = lo; i < hi; ++i) sum += keep[i];
    s[threadIdx.x] = sum;
    __syncthreads();
    for (int off = 1; off < kRankThreads; off <<= 1) {  // inclusive Hillis-Steele
        const int add = (int)threadIdx.x >= off ? s[threadIdx.x - off] : 0;
        __syncthreads();
        s[threadIdx.x] += add;
        __syncthreads();
    }
    int run = s[threadIdx.x] - sum;
    for (int i = lo; i < hi; ++i) {
        rank[i] = run;
        run += keep[i];
    }
    if (threadIdx.x == kRankThreads - 1) rank[n] = s[kRankThreads - 1];
}

__global__ void prior_zero(uint32_t* __restrict__ p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// process_mask (planar_initer.cu:290-324): retained triangle number + 1 into
// the pixels of its walk; later triangles overwrite earlier ones there.
__global__ void prior_raster(const int* __restrict__ tris, int ntri, const int* __restrict__ keep,
                             const int* __restrict__ rank, int h, int w,
                             uint32_t* __restrict__ mask) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntri || !keep[t]) return;
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
__global__ void prior_index(const int* __restrict__ keep, const int* __restrict__ rank, int ntri,
                            int* __restrict__ tri_of) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < ntri && keep[t]) tri_of[rank[t]] = t;
}

}  // namespace
}  // namespace gmvs

using namespace gmvs;

extern "C" size_t gmvs_prior_device_workspace_bytes(int32_t height, int32_t width) {
    if (height <= 0 || width <= 0) return 0;
    return prior_layout(height, width).total;
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

    // support points, then the kept cells in the reference's order (host)
    hipLaunchKernelGGL(prior_support, dim3((L.ncells + kBlock - 1) / kBlock), dim3(kBlock), 0, st,
                       cost, h, w, ncy, L.ncells, cells);
    GMVS_LAUNCH_CHECK();
    std::vector<int2> hcells(L.ncells);
    GMVS_HIP_CHECK(hipMemcpyAsync(hcells.data(), cells, sizeof(int2) * L.ncells,
                                  hipMemcpyDeviceToHost, st));
    GMVS_HIP_CHECK(hipStreamSynchronize(st));
    std::vector<int32_t> xy;
    xy.reserve(2 * L.ncells);
    for (const int2& c : hcells)
        if (c.x >= 0) {
            xy.push_back(c.x);
            xy.push_back(c.y);
        }
    const int npts = int(xy.size() / 2);
    const std::vector<int> tv = host_delaunay(xy.data(), npts);
    const int ntri = int(tv.size() / 3);
    GMVS_ARG_CHECK(ntri <= L.max_tri, "prior_build_device: %d triangles exceed %d", ntri, L.max_tri);
    std::vector<int> htris(6 * (size_t)std::max(ntri, 1));
    for (int t = 0; t < ntri; ++t)
        for (int k = 0; k < 3; ++k) {
            htris[6 * t + 2 * k] = xy[2 * tv[3 * t + k]];
            htris[6 * t + 2 * k + 1] = xy[2 * tv[3 * t + k] + 1];
        }
    if (ntri > 0)
        GMVS_HIP_CHECK(hipMemcpyAsync(tris, htris.data(), sizeof(int) * 6 * ntri,
                                      hipMemcpyHostToDevice, st));

    // filter + planes, numbering, rasterisation, per-pixel output; the mask
    // is assembled in plane_masks, then rewritten in place by prior_output
    // (each pixel reads only its own entry)
    const int tb = (std::max(ntri, 1) + kBlock - 1) / kBlock;
    if (ntri > 0) {
        hipLaunchKernelGGL(prior_triangles, dim3(tb), dim3(kBlock), 0, st, tris, ntri, labels,
                           params->superpixel_filter ? 1 : 0, depth, h, w, *ref_camera, keep, planes);
        GMVS_LAUNCH_CHECK();
        hipLaunchKernelGGL(prior_rank, dim3(1), dim3(kRankThreads), 0, st, keep, ntri, rank);
        GMVS_LAUNCH_CHECK();
    }
    const int pb = (h * w + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(prior_zero, dim3(pb), dim3(kBlock), 0, st, plane_masks, h * w);
    GMVS_LAUNCH_CHECK();
    int* tri_of = reinterpret_cast<int*>(ws + L.tri_of);
    int nkeep = 0;
    if (ntri > 0) {
        hipLaunchKernelGGL(prior_raster, dim3(tb), dim3(kBlock), 0, st, tris, ntri, keep, rank, h,
                           w, plane_masks);
        GMVS_LAUNCH_CHECK();
        hipLaunchKernelGGL(prior_index, dim3(tb), dim3(kBlock), 0, st, keep, rank, ntri, tri_of);
        GMVS_LAUNCH_CHECK();
        GMVS_HIP_CHECK(hipMemcpyAsync(&nkeep, rank + ntri, sizeof(int), hipMemcpyDeviceToHost, st));
    }
    hipLaunchKernelGGL(prior_output, dim3(pb), dim3(kBlock), 0, st, plane_masks, planes, tri_of, h,
                       w, *ref_camera, params->depth_min, params->depth_max,
                       reinterpret_cast<float4*>(prior_planes), plane_masks);
    GMVS_LAUNCH_CHECK();
    GMVS_HIP_CHECK(hipStreamSynchronize(st));
    if (out_num_triangles) *out_num_triangles = nkeep;
    return GMVS_OK;
}
