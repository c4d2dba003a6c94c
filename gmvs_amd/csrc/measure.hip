// Measurement kernels (gfx950): the peak rate of the gathers the NCC kernels
// issue, the reference point of their gather-bandwidth fraction.
//
// gather_peak: every lane issues structured buffer loads of one cell (16 B
// for FP32 cells, 8 B for 16-bit cells, the loads of the NCC taps) inside a
// per-block window of cells, in one of two patterns:
//   * random lane pairs (pattern 0): each lane pair at a pseudo-random cell
//     of a window small enough to stay in the CU's vector L1 cache, the odd
//     lane on the next cell;
//   * NCC waves (pattern 1): the layout of one tap load of the NCC kernels'
//     waves, 32 pixels of one checkerboard colour (8 active columns, 2 cells
//     apart, x 4 rows) whose two lanes read cells 2 apart, at a wave-uniform
//     pseudo-random offset inside a window of rows of 64 cells.
// Indices cost a few VALU instructions: the kernel is bound by the loads,
// not by its own arithmetic. Eight loads per lane are in flight per round;
// the loaded words are summed so nothing is removed.
#include <hip/hip_runtime.h>

#include "abi_common.h"
#include "pm_device.h"

namespace gmvs {
namespace {

constexpr int kGatherBlock = 256;
constexpr int kGatherInFlight = 8;

constexpr uint32_t kWavePitch = 64;     // cells per window row (pattern 1)
constexpr uint32_t kWaveMinCells = 2048;  // 32 rows: offsets reach row 15 + 3, column 31 + 17

template <int PR, int PATTERN>
__global__ __launch_bounds__(kGatherBlock) void gather_peak(const void* __restrict__ cells,
                                                            uint32_t window, int rounds,
                                                            uint32_t* __restrict__ sink) {
    const uint32_t count = window * gridDim.x;
    const QuadSrc src = quad_src<PR>(cells, count, 0);
    const uint32_t base = blockIdx.x * window;
    uint32_t x = (blockIdx.x * kGatherBlock + threadIdx.x) * 2654435761u + 12345u;
    // wave-uniform stream (pattern 1)
    uint32_t xw = (blockIdx.x * (kGatherBlock / 64) + threadIdx.x / 64) * 2246822519u + 777u;
    const uint32_t lane = threadIdx.x & 63u, pix = lane >> 1;
    uint32_t acc = 0;
    for (int r = 0; r < rounds; ++r) {
        Cell<PR> q[kGatherInFlight];
#pragma unroll
        for (int i = 0; i < kGatherInFlight; ++i) {
            uint32_t k;
            if constexpr (PATTERN == 0) {
                // lane pairs: even lane at cell k, odd lane at cell k + 1
                x = x * 1664525u + 1013904223u;
                k = ((x >> 8) & (window - 2u)) + (threadIdx.x & 1u);
            } else {
                // row cy = oy + pix / 8 <= 18, column 2 * (pix % 8) + (cy & 1)
                // + 2 * (lane & 1) + ox <= 48 < kWavePitch: inside 19 rows
                xw = xw * 1664525u + 1013904223u;
                const uint32_t ox = (xw >> 8) & 31u, oy = (xw >> 20) & 15u;
                const uint32_t cy = oy + (pix >> 3);
                const uint32_t cx = ox + 2u * (pix & 7u) + (cy & 1u) + 2u * (lane & 1u);
                k = cy * kWavePitch + cx;
            }
            q[i].load(src, (int)(base + k));
        }
#pragma unroll
        for (int i = 0; i < kGatherInFlight; ++i) {
            if constexpr (PR == PREC_F32)
                acc += __float_as_uint(q[i].q.x) ^ __float_as_uint(q[i].q.w);
            else
                acc += (uint32_t)q[i].q.x ^ (uint32_t)q[i].q.y;
        }
    }
    if (acc == 0x9E3779B9u) sink[blockIdx.x] = acc;  // practically never true
}

}  // namespace
}  // namespace gmvs

using namespace gmvs;

extern "C" size_t gmvs_gather_peak_buffer_bytes(int32_t cell_bytes, int32_t window_cells) {
    if ((cell_bytes != 16 && cell_bytes != 8) || window_cells < 64 ||
        (window_cells & (window_cells - 1)) != 0)
        return 0;
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return 0;
    const size_t blocks = (size_t)cus * 8;
    return blocks * (size_t)window_cells * cell_bytes + blocks * 4;
}

extern "C" int gmvs_gather_peak(int32_t cell_bytes, int32_t pattern, int32_t window_cells,
                                int32_t rounds, void* buffer, size_t buffer_bytes,
                                double* out_gbps, void* stream) {
    clear_error();
    GMVS_ARG_CHECK(cell_bytes == 16 || cell_bytes == 8, "gather_peak: cell_bytes must be 16 or 8");
    GMVS_ARG_CHECK(pattern == 0 || pattern == 1, "gather_peak: pattern must be 0 or 1");
    GMVS_ARG_CHECK(window_cells >= 64 && (window_cells & (window_cells - 1)) == 0,
                   "gather_peak: the window must be a power of two of at least 64 cells");
    GMVS_ARG_CHECK(pattern == 0 || window_cells >= (int32_t)kWaveMinCells,
                   "gather_peak: the NCC-wave pattern needs a window of at least %u cells",
                   kWaveMinCells);
    GMVS_ARG_CHECK(rounds >= 1 && buffer && out_gbps, "gather_peak: bad argument");
    const size_t need = gmvs_gather_peak_buffer_bytes(cell_bytes, window_cells);
    GMVS_ARG_CHECK(need > 0 && buffer_bytes >= need, "gather_peak: buffer of %zu bytes < %zu",
                   buffer_bytes, need);
    int dev = 0, cus = 0;
    GMVS_HIP_CHECK(hipGetDevice(&dev));
    GMVS_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const int blocks = cus * 8;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    char* buf = reinterpret_cast<char*>(buffer);
    uint32_t* sink = reinterpret_cast<uint32_t*>(buf + (size_t)blocks * window_cells * cell_bytes);
    auto kern = pattern == 0 ? (cell_bytes == 16 ? gather_peak<PREC_F32, 0> : gather_peak<PREC_F16, 0>)
                             : (cell_bytes == 16 ? gather_peak<PREC_F32, 1> : gather_peak<PREC_F16, 1>);
    // warm-up launch, then the timed one
    hipLaunchKernelGGL(kern, dim3(blocks), dim3(kGatherBlock), 0, st, buffer, (uint32_t)window_cells,
                       rounds, sink);
    GMVS_LAUNCH_CHECK();
    hipEvent_t a, b;
    GMVS_HIP_CHECK(hipEventCreate(&a));
    GMVS_HIP_CHECK(hipEventCreate(&b));
    GMVS_HIP_CHECK(hipEventRecord(a, st));
    hipLaunchKernelGGL(kern, dim3(blocks), dim3(kGatherBlock), 0, st, buffer, (uint32_t)window_cells,
                       rounds, sink);
    GMVS_LAUNCH_CHECK();
    GMVS_HIP_CHECK(hipEventRecord(b, st));
    GMVS_HIP_CHECK(hipEventSynchronize(b));
    float ms = 0.f;
    GMVS_HIP_CHECK(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    const double bytes =
        (double)blocks * kGatherBlock * (double)rounds * kGatherInFlight * (double)cell_bytes;
    *out_gbps = bytes / ((double)ms * 1e-3) / 1e9;
    return GMVS_OK;
}
