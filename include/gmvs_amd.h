/*
 * gmvs_amd -- MI355X (gfx950) PatchMatch multi-view stereo hot path.
 *
 * C ABI of libgmvs_amd.so: plain structs, plain pointers and sizes, an int
 * status (0 = ok) and a per-thread error string. No torch or HIP types
 * appear in the signatures; `stream` is a hipStream_t passed as void*
 * (NULL = the null stream). Every device pointer must be on the current
 * HIP device.
 *
 * Each entry point names the reference interface it replaces. Citations are
 * relative to the gmvs source tree (Gorilla-Lab-SCUT/gmvs).
 */
#ifndef GMVS_AMD_H
#define GMVS_AMD_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GMVS_MAX_NGB 32 /* MAX_NGB, gmvs/include/data_structure.hpp:10 */

/* Status codes. */
enum {
    GMVS_OK = 0,
    GMVS_ERR_ARG = 1,      /* invalid argument (shape, count, NULL pointer) */
    GMVS_ERR_HIP = 2,      /* a HIP runtime call failed */
    GMVS_ERR_NOSPACE = 3,  /* workspace too small */
    GMVS_ERR_UNSUPPORTED = 4,
    GMVS_ERR_INTERNAL = 5  /* a device-side consistency check failed */
};

/* mvs::Camera (gmvs/include/data_structure.hpp:45-53): K, world->camera R, t
 * (X_c = R X_w + t), image size and depth range. Same 100-byte layout. */
typedef struct gmvs_camera {
    float K[9];
    float R[9];
    float t[3];
    int32_t height;
    int32_t width;
    float depth_min;
    float depth_max;
} gmvs_camera;

/* mvs::Problem (gmvs/include/data_structure.hpp:39-43). */
typedef struct gmvs_problem {
    int32_t num_ngb;
    int32_t ref_image_id;
    int32_t src_image_ids[GMVS_MAX_NGB];
} gmvs_problem;

/* PatchMatchParams (gmvs/lib/include/planar_initer.hpp:13-31, which is
 * mvs::PatchMatchParams of data_structure.hpp:20-37 plus superpixel_filter). */
typedef struct gmvs_params {
    int32_t max_iterations;
    int32_t patch_size;
    int32_t num_images;
    int32_t max_image_size;
    int32_t radius_increment;
    float sigma_spatial;
    float sigma_color;
    int32_t top_k;
    float baseline;
    float depth_min;
    float depth_max;
    float disparity_min;
    float disparity_max;
    bool geom_consistency;
    bool multi_geometry;
    bool planar_prior;
    bool superpixel_filter;
} gmvs_params;

/* Initialisation of the plane hypotheses (A23). */
enum {
    GMVS_INIT_RANDOM = 0,       /* RandomInitialization random branch, lib 1553-1575 */
    GMVS_INIT_DEPTH_NORMAL = 1, /* initialization, src 1046-1082 */
    GMVS_INIT_PRIOR = 2         /* RandomInitialization prior branch, lib 1513-1551 */
};

/* Storage and arithmetic of the source-image samples. In every mode the
 * homography factors are composed in FP64, tap coordinates are FP32, and the
 * NCC moments are accumulated in FP32 about the patch-centre intensity. */
enum {
    GMVS_PREC_FP32 = 0, /* plain FP32 path: 16-byte bilinear cells, FP32 FMAs */
    GMVS_PREC_FP16 = 1, /* 8-byte FP16 cells, bilinear as FP32 FMAs of the widened
                           FP16 operands (exact for 8-bit images), FP32 sums */
    GMVS_PREC_BF16 = 2  /* 8-byte BF16 cells, bilinear as BF16 dot products, FP32 sums */
};

/* Reference quirks (gmvs_pm_args.compat). 0 mirrors the reference in each
 * case; a set bit switches that one quirk to the corrected behaviour, so
 * both can be run and compared. The oracle implements the same bits. */
enum {
    /* right_far keeps the minimum cost; the reference keeps the maximum
       (`costMin < costs[...]`, lib 954, src 794) */
    GMVS_COMPAT_FIX_RIGHT_FAR = 1,
    /* an out-of-image (unflagged) candidate costs 2 against every view; the
       reference's cost_array[8][32] = {2.0f} sets only [0][0] and leaves the
       rest 0 (lib 854, src 717), so such candidates look good to the view
       selection */
    GMVS_COMPAT_FIX_COST_DEFAULT = 2,
    /* GMVS_INIT_PRIOR rotates the reused first-pass normals, which are in
       the world frame, into the camera frame; the reference uses them as
       camera-frame normals (lib 1546-1550 after 1672-1673) */
    GMVS_COMPAT_FIX_PRIOR_NORMALS = 4,
    /* the planar-prior pass updates the selected views when an unmasked
       pixel adopts a candidate, as the plain pass does; the reference leaves
       them (lib 1306-1318 against 1332) */
    GMVS_COMPAT_FIX_PRIOR_SELECTION = 8
};

/* One PatchMatch pass over one reference view. */
typedef struct gmvs_pm_args {
    int32_t num_images;           /* reference + sources, 2..32 */
    const gmvs_camera* cameras;   /* host [num_images], reference first */
    const float* const* images;   /* host array of num_images device ptrs, [H*W] f32 */
    const float* const* depths;   /* host array of num_images device ptrs, [H*W] f32, or NULL:
                                     [0] = input depth (INIT_DEPTH_NORMAL), [1..] = geom maps */
    const float* ref_normal;      /* device [H*W*3] world normals (INIT_DEPTH_NORMAL) or NULL */
    const float* prior_planes;    /* device [H*W*4] (n, d) per pixel, or NULL */
    const uint32_t* plane_masks;  /* device [H*W] triangle id + 1, 0 = no prior, or NULL */
    float* planes;                /* device [H*W*4] state: in for INIT_PRIOR, out always
                                     ((n_world, depth) after the pass) */
    float* costs;                 /* device [H*W] state: in for INIT_PRIOR, out always */
    gmvs_params params;           /* depth_min/depth_max are used as given */
    int32_t init_mode;            /* GMVS_INIT_* */
    int32_t geom;                 /* 1: add 0.1 * geometric-consistency cost (src 657) */
    uint32_t seed;                /* Philox key word 0 (reference seed 3407) */
    uint32_t rng_view;            /* Philox key word 1 (global reference view id) */
    int32_t precision;            /* GMVS_PREC_* */
    int32_t skip_filter;          /* 1: skip CheckerboardFilter */
    float* out_depth;             /* device [H*W] or NULL */
    float* out_normal;            /* device [H*W*3] or NULL */
    float* out_cost;              /* device [H*W] or NULL */
    int32_t tile_cols;            /* shape of the 64-pixel work tile: columns 4, 8, 16 or 32
                                     (rows = 64 / cols); 0 = 8 x 8. Positions are formed
                                     relative to the tile, so the shape changes costs at
                                     rounding level (chosen by gmvs_amd.autotune) */
    uint32_t compat;              /* GMVS_COMPAT_* bits; 0 = the reference's behaviour */
} gmvs_pm_args;

/* Last error message of the calling thread ("" if none). */
const char* gmvs_last_error(void);

/* Library version string. */
const char* gmvs_version(void);

/* mvs::load_camera (gmvs/src/data_structure.cpp:4-29;
 * gmvs/lib/src/planar_initer.cpp:3-28). */
int gmvs_load_camera(const float K[9], const float R[9], const float t[3], int32_t height,
                     int32_t width, float depth_min, float depth_max, gmvs_camera* out);

/* mvs::load_problem (gmvs/src/data_structure.cpp:31-39). More than
 * GMVS_MAX_NGB sources is an error (the reference asserts). */
int gmvs_load_problem(int32_t ref_id, const int32_t* src_ids, int32_t num_src, gmvs_problem* out);

/* Device workspace needed by gmvs_pm_run for one reference view. */
size_t gmvs_pm_workspace_bytes(int32_t height, int32_t width, int32_t num_images,
                               int32_t precision);

/* One PatchMatch pass: init, max_iterations red/black sweeps, depth/normal,
 * median filter and export. Replaces PatchMatcher::run_patch_match
 * (gmvs/src/patch_match.cu:1353-1519) and PlanarInit::run_patch_match
 * (gmvs/lib/src/patch_match_kernel.cu:1711-1810). Asynchronous on `stream`. */
int gmvs_pm_run(const gmvs_pm_args* args, void* workspace, size_t workspace_bytes, void* stream);

/* Batch of passes: n problems (reference views of one or several scenes,
 * sizes may differ) run as one schedule, every kernel launched once for the
 * whole batch over a table of pixel tiles pulled from per-XCD work queues.
 * Problems share precision and max_iterations; 1 <= n <= 64. Results are
 * bit-identical to n separate gmvs_pm_run calls. No counterpart in the
 * reference, which runs one view per call (PatchMatcher::run_patch_match,
 * gmvs/src/patch_match.cu:1353-1519). The launch sequence allocates nothing
 * and can be captured in a HIP graph. */
size_t gmvs_pm_batch_workspace_bytes(const gmvs_pm_args* args, int32_t n);
int gmvs_pm_run_batch(const gmvs_pm_args* args, int32_t n, void* workspace,
                      size_t workspace_bytes, void* stream);

/* Launch tuning (no counterpart in the reference): blocks per CU of the
 * persistent PatchMatch kernels, 0 = as many as fit (the default). Fewer
 * resident waves can win when the gathers contend for the L1; results do
 * not depend on it. Process-wide; picked by gmvs_amd.autotune. */
enum {
    GMVS_TUNE_INIT = 0,       /* pm_init */
    GMVS_TUNE_CANDIDATES = 1, /* pm_candidates */
    GMVS_TUNE_REFINE = 2      /* pm_refine */
};
int gmvs_set_blocks_per_cu(int32_t kernel, int32_t blocks);

/* Launch tuning (no counterpart in the reference): lanes per pixel of the
 * NCC lane groups of a persistent PatchMatch kernel (GMVS_TUNE_*): 2 or 4,
 * 0 = the default for the run's precision. The block holds 64 pixels x lanes
 * threads. Changes the order of the FP32 NCC sums (rounding-level cost
 * differences), not the algorithm. Process-wide; picked by gmvs_amd.autotune. */
int gmvs_set_lane_width(int32_t kernel, int32_t lanes);

/* Tap indexing of the NCC gathers (no counterpart in the reference): 1 (the
 * default) lets a wave whose patches all map inside the padded cell grid
 * index its taps without clamping; 0 clamps every tap as the reference's
 * software bilinear does (gmvs/src/patch_match.cu:16-31). Both forms give
 * identical results; 0 exists to test that. Process-wide, read when a run
 * is set up. */
int gmvs_set_fast_taps(int32_t enable);

/* Kernel fusion (no counterpart in the reference): 1 runs the joint view
 * selection and candidate adoption (gmvs/src/patch_match.cu:915-1040,
 * gmvs/lib/src/patch_match_kernel.cu:1114-1334) as the epilogue of each
 * candidate tile instead of as a separate launch; 0 (the default) launches
 * them separately. Both give identical results. Process-wide, read when a
 * run is set up. */
int gmvs_set_select_fusion(int32_t enable);

/* Parity hook: V bilateral NCCs for n (pixel, plane) pairs -- the batched form
 * of ComputeMultiViewCostVector (gmvs/lib/src/patch_match_kernel.cu:571-588,
 * gmvs/src/patch_match.cu:503-522). xy: device [n*2] int32, planes: device
 * [n*4] f32 (camera-frame (n, d)), costs: device [n*(num_images-1)] f32. */
size_t gmvs_cost_workspace_bytes(int32_t height, int32_t width, int32_t num_images,
                                 int32_t precision);
int gmvs_cost_batch(int32_t num_images, const gmvs_camera* cameras, const float* const* images,
                    const gmvs_params* params, int32_t precision, int32_t n, const int32_t* xy,
                    const float* planes, float* costs, void* workspace, size_t workspace_bytes,
                    void* stream);

/* Fusion of one reference view -- Fuser::run_fusion + batch_fusion_kernel
 * (gmvs/src/fusion.cu:61-182, 211-275). depths/normals: host arrays of
 * num_images device pointers (reference first). masks: device
 * [num_mask_views*H*W] bytes (bool), indexed by ref_id and src_ids (global
 * view ids, each < num_mask_views), updated in place.
 * out_depth [H*W] and out_normal [H*W*3] are written (zeroed first).
 * Thresholds: reference values are 2 px, 0.01 and cos(0.174533 rad). */
size_t gmvs_fuse_workspace_bytes(int32_t height, int32_t width);
int gmvs_fuse_view(int32_t num_images, const gmvs_camera* cameras, const float* const* depths,
                   const float* const* normals, uint8_t* masks, int32_t num_mask_views,
                   int32_t ref_id,
                   const int32_t* src_ids, int32_t geom_consistent, float reproj_px,
                   float rel_depth, float cos_angle, float* out_depth, float* out_normal,
                   void* workspace, size_t workspace_bytes, void* stream);

/* gmvs_fuse_view with the reprojection check in FP64 (not in the reference,
 * whose batch_fusion_kernel is FP32: gmvs/src/fusion.cu:61-182): the world
 * points, both projections, the reprojection error, the relative depth
 * difference and the averaged point in double; inputs, thresholds and
 * outputs as gmvs_fuse_view. For scenes whose world coordinates are large
 * against their depths (geo-referenced or far-from-origin reconstructions),
 * where FP32 world points carry reprojection errors near the 2 px
 * threshold. Same workspace (gmvs_fuse_workspace_bytes). */
int gmvs_fuse_view_fp64(int32_t num_images, const gmvs_camera* cameras,
                        const float* const* depths, const float* const* normals, uint8_t* masks,
                        int32_t num_mask_views, int32_t ref_id, const int32_t* src_ids,
                        int32_t geom_consistent, float reproj_px, float rel_depth,
                        float cos_angle, float* out_depth, float* out_normal, void* workspace,
                        size_t workspace_bytes, void* stream);

/* No counterpart in the reference: gmvs_fuse_view for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). The same arguments without the workspace:
 * the scratch (25 B per pixel) is a file-scope device array of
 * gmvs_fuse_noalloc_max_pixels() = 2048*1536 pixels, one per device, held
 * until the process ends. Larger images are rejected with GMVS_ERR_ARG, as is
 * a call while kernel profiling is on (it creates timing events). Every call
 * on a device shares that device's array: calls that overlap on one device
 * (other streams or host threads) race. Nothing is allocated or freed. */
int64_t gmvs_fuse_noalloc_max_pixels(void);
int gmvs_fuse_view_noalloc(int32_t num_images, const gmvs_camera* cameras,
                           const float* const* depths, const float* const* normals,
                           uint8_t* masks, int32_t num_mask_views, int32_t ref_id,
                           const int32_t* src_ids, int32_t geom_consistent, float reproj_px,
                           float rel_depth, float cos_angle, float* out_depth, float* out_normal,
                           void* stream);

/* Planar prior construction on the host -- PlanarInit::triangulation and
 * cuda_planar_prior_init (gmvs/lib/src/planar_initer.cu:195-420, 686-745):
 * support points from the pass-1 cost, Delaunay triangulation, optional
 * superpixel filter (int32 labels), 3-point plane fit, rasterisation.
 * depth/cost: host [H*W]; labels: host [H*W] int32 or NULL;
 * out prior_planes: host [H*W*4]; out plane_masks: host [H*W] (0 = none).
 * out_num_triangles may be NULL. */
int gmvs_prior_build(const gmvs_camera* ref_camera, const float* depth, const float* cost,
                     const int32_t* labels, const gmvs_params* params, float* prior_planes,
                     uint32_t* plane_masks, int32_t* out_num_triangles);

/* gmvs_prior_build on the device, with the same results bit for bit:
 * support points, their Delaunay triangulation (one star per point, the
 * host construction's predicates and order), the superpixel filter, plane
 * fit, rasterisation and the depth-range check, all as kernels on the
 * stream. With out_num_triangles NULL the call only enqueues work (no
 * synchronisation); with it, the call waits for the stream, reports the
 * number of retained triangles and returns GMVS_ERR_INTERNAL if the device
 * triangulation failed its own checks. Replaces the same reference steps as
 * gmvs_prior_build (gmvs/lib/src/planar_initer.cu:195-420,686-745).
 * depth, cost: device [H*W] f32; labels: device [H*W] int32 or NULL;
 * prior_planes: device [H*W*4]; plane_masks: device [H*W];
 * workspace: device, gmvs_prior_device_workspace_bytes(H, W) bytes. */
size_t gmvs_prior_device_workspace_bytes(int32_t height, int32_t width);
int gmvs_prior_build_device(const gmvs_camera* ref_camera, const float* depth, const float* cost,
                            const int32_t* labels, const gmvs_params* params, float* prior_planes,
                            uint32_t* plane_masks, int32_t* out_num_triangles, void* workspace,
                            size_t workspace_bytes, void* stream);

/* Delaunay triangulation used by gmvs_prior_build, exported for its tests;
 * replaces the cv::Subdiv2D insert / getTriangleList calls of
 * PlanarInit::triangulation (gmvs/lib/src/planar_initer.cu:354-361).
 * xy: host [n*2] int32 in [0, 65536); tri_out: host [max_tri*3] vertex
 * indices (counter-clockwise) or NULL; *num_tri receives the count. */
int gmvs_delaunay(const int32_t* xy, int32_t n, int32_t* tri_out, int32_t max_tri,
                  int32_t* num_tri);

/* The device triangulation of gmvs_prior_build_device on its own, for tests
 * and callers with support points of their own; replaces the same Subdiv2D
 * calls (gmvs/lib/src/planar_initer.cu:354-361). Points on a grid of 5x5
 * pixel cells, at most one per cell: cell_xy is device [ncx*ncy*2] int32,
 * cell (cx, cy) at index cx*ncy + cy holding (x, y) inside the cell
 * (x / 5 == cx, y / 5 == cy) or (-1, -1); points are numbered in cell
 * order. tri_out: device [max_tri*3], the triangles as gmvs_delaunay lists
 * them for the same numbered points (the same triangles in the same
 * order); *num_tri (host) the count. Synchronises the stream.
 * workspace: device, gmvs_prior_device_workspace_bytes(5*ncy, 5*ncx) bytes.
 * Returns GMVS_ERR_ARG for a point outside its cell, GMVS_ERR_NOSPACE when
 * the triangles exceed max_tri. */
int gmvs_delaunay_cells_device(const int32_t* cell_xy, int32_t ncx, int32_t ncy, int32_t* tri_out,
                               int32_t max_tri, int32_t* num_tri, void* workspace,
                               size_t workspace_bytes, void* stream);

/* The same star construction run on the host (the device kernel's code),
 * for tests on machines without a device. cell_xy, tri_out: host. */
int gmvs_delaunay_cells_host(const int32_t* cell_xy, int32_t ncx, int32_t ncy, int32_t* tri_out,
                             int32_t max_tri, int32_t* num_tri);

/* Point gather after fusion -- the "Gather mvs points" loop of
 * gmvs/scripts/mvs_fusion_segmentaion_cpu.py:158-205. For each view, the
 * pixels of its fused [H,W,4] depth_normal map with depth > 0 become rows
 * [x y z nx ny nz] (float64): point = C + depth * R (x', y', 1) with
 * x' = (col - cx) / fx, y' = (row - cy) / fy, (R | C) the camera-to-world pose
 * (pose/NNNN.txt) and the normal widened from float32. A row passes the
 * view-angle filter when -(R (x', y', 1)) . n > cos_view_angle (the ray is not
 * normalised, as in the reference; cos(70 deg) there). Rows come in view
 * order, then row-major pixel order (numpy's boolean-mask order).
 * depth_normals: host array of num_views device pointers [H*W*4];
 * poses: host [num_views*16] row-major 4x4; filtered: device
 * [num_views*H*W*4] or NULL (depth_normal_filter: the map where the filter
 * passes, else 0); points / points_filtered: device [capacity*6] /
 * [capacity_filtered*6] or NULL (rows past the capacity are dropped);
 * counts: device uint64 [2] receives the total and filtered row counts.
 * Calling with no outputs counts only. */
typedef struct gmvs_points_args {
    int32_t height, width, num_views;
    const float* const* depth_normals;
    const double* poses;
    double fx, fy, cx, cy;
    double cos_view_angle;
    float* filtered;
    double* points;
    double* points_filtered;
    uint64_t capacity, capacity_filtered;
    uint64_t* counts;
} gmvs_points_args;
size_t gmvs_points_workspace_bytes(int32_t height, int32_t width, int32_t num_views);
int gmvs_points_gather(const gmvs_points_args* args, void* workspace, size_t workspace_bytes,
                       void* stream);

/* Text export -- np.savetxt(path, rows, fmt="%.<decimals>f") as the
 * reference writes mvs_depthnorm[_filter].xyz
 * (mvs_fusion_segmentaion_cpu.py:207-216): values joined by ' ', rows ended
 * by '\n', each value the exactly rounded decimal (ties to even) as Python's
 * "%f" prints it, "-" on every negative value, nan / inf / -inf.
 * rows: device [n_rows*ncol] float64; out: device [capacity] bytes or NULL
 * (bytes past the capacity are dropped); out_len: device uint64 receiving
 * the byte count, or UINT64_MAX if a value's magnitude times 10^decimals
 * reaches 2^63 (unsupported). ncol <= 64, decimals <= 9. */
size_t gmvs_format_workspace_bytes(uint64_t n_rows);
int gmvs_format_rows(const double* rows, uint64_t n_rows, int32_t ncol, int32_t decimals, char* out,
                     uint64_t capacity, uint64_t* out_len, void* workspace,
                     size_t workspace_bytes, void* stream);

/* sliding_window_normal (gmvs/lib/src/planar_initer.cu:471-535): device
 * planar_region [H*W] int32 (1 = planar), normal [H*W*3] f32, out
 * [H*W*3] f32: 0 on planar pixels, else the mean normal of the non-planar
 * pixels of the window_size x window_size window, summed row by row and
 * divided as cv::Vec3f /= int does. */
int gmvs_sliding_window_normal(const int32_t* planar_region, const float* normal, int32_t height,
                               int32_t width, int32_t window_size, float* out, void* stream);

/* Per-superpixel utilities. segment ids: device [H*W] int32 in
 * [0, num_segments) with num_segments = max id + 1 (ids outside count as
 * segment 0); per-segment sums run over the segment's pixels in row-major
 * order, as the reference's loops do. Not capturable in a HIP graph (the
 * radix sort clears its counters with hipMemsetAsync). */
size_t gmvs_segments_workspace_bytes(int32_t height, int32_t width, int32_t num_segments);
/* filter_by_var_map (gmvs/lib/src/planar_initer.cu:634-684): out[p] = id
 * if the mean of var_map over segment id (>= 1) is <= var_thresh, else 0. */
int gmvs_filter_by_var_map(const float* var_map, const int32_t* segment_ids, int32_t height,
                           int32_t width, int32_t num_segments, float var_thresh, int32_t* out,
                           void* workspace, size_t workspace_bytes, void* stream);
/* fusion_planar_mask (gmvs/lib/src/planar_initer.cu:537-632): planar_mask
 * [H*W] int32 (0 = non-planar), filter_normals [H*W*3] f32; out [H*W] int32:
 * 0 (small or empty segment), the segment id, 300 (object) or 301, 302, ...
 * (textureless segments in id order). */
int gmvs_fusion_planar_mask(const int32_t* seg_ids, const int32_t* planar_mask,
                            const float* filter_normals, int32_t rows, int32_t cols,
                            int32_t num_segments, int32_t thresh, float nonplanar_percent,
                            float cos_sim_thresh, float match_ratio_thresh, int32_t* out,
                            void* workspace, size_t workspace_bytes, void* stream);

/* No counterpart in the reference: the peak rate of the NCC kernels' tap
 * gathers (structured buffer loads of one 16-byte FP32 cell or one 8-byte
 * 16-bit cell) in GB/s over the whole device, within per-block windows of
 * window_cells cells (a power of two); the reference point of the gather
 * fraction reported by the benchmark. pattern 0: lane pairs on neighbouring
 * cells at random places of the window (keep it L1-sized); pattern 1: the
 * layout of one tap load of an NCC wave (32 pixels of one checkerboard
 * colour, 8 active columns x 4 rows, lanes of a pixel 2 cells apart) at a
 * random offset per load, in rows of 64 cells (window_cells >= 2048).
 * buffer: device, gmvs_gather_peak_buffer_bytes() bytes (its contents are
 * read, never used). Synchronises the stream. */
size_t gmvs_gather_peak_buffer_bytes(int32_t cell_bytes, int32_t window_cells);
int gmvs_gather_peak(int32_t cell_bytes, int32_t pattern, int32_t window_cells, int32_t rounds,
                     void* buffer, size_t buffer_bytes, double* out_gbps, void* stream);

/* Kernel timing. No counterpart in the reference (which only prints
 * wall-clock totals, gmvs/scripts/launch.py:125-126,176-177). When enabled,
 * every kernel launch of the library is bracketed by hipEvents on its
 * stream; gmvs_profile_collect waits for the recorded events and returns the
 * accumulated launches and milliseconds per kernel name (out may be NULL to
 * query the count). The PatchMatch kernels also count the bilateral NCCs
 * (one reference pixel, one plane, one source view) and geometric-term
 * evaluations they perform. */
typedef struct gmvs_kernel_stat {
    char name[32];
    int64_t launches;
    double total_ms;
    uint64_t ncc;
    uint64_t geom;
} gmvs_kernel_stat;
int gmvs_profile_enable(int32_t on);
int gmvs_profile_collect(gmvs_kernel_stat* out, int32_t max_out, int32_t* num_out);
int gmvs_profile_reset(void);

/* Philox4x32-10 block (host), exported for known-answer tests. */
void gmvs_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

#ifdef __cplusplus
}
#endif

#endif /* GMVS_AMD_H */
