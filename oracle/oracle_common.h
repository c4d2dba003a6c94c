// Oracle for the gmvs PatchMatch MVS hot path -- TEST INFRASTRUCTURE ONLY.
//
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
// load this code, and only as the checker (or as the timed CPU baseline).
// The product path (gmvs_amd/) never links, loads or calls it.
//
// This header holds the plain-C data types of the oracle and its own
// Philox4x32-10 counter RNG (Salmon, Moraes, Dror, Shaw, "Parallel random
// numbers: as easy as 1, 2, 3", SC'11; round function, multipliers and Weyl
// key schedule as published there). It is written independently of the
// product's copy so that a mistake in one is caught by the other.
#pragma once

#include <stdint.h>

// Same field layout as mvs::Camera (gmvs/include/data_structure.hpp:45-53).
typedef struct OCamera {
    float K[9];
    float R[9];
    float t[3];
    int32_t height;
    int32_t width;
    float depth_min;
    float depth_max;
} OCamera;

// Same field layout as PatchMatchParams (gmvs/lib/include/planar_initer.hpp:13-31).
typedef struct OParams {
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
} OParams;

// One PatchMatch pass over one reference view (host pointers).
typedef struct OPMArgs {
    const OCamera* cameras;       // [num_images], reference first
    int32_t num_images;
    const float* const* images;   // [num_images] -> [H*W] intensities
    const float* const* depths;   // [num_images] -> [H*W]; [0] = init depth, [1..] = geom src maps; may be NULL
    const float* ref_normal;      // [H*W*3] world-frame normals for init_mode 1; may be NULL
    const float* prior_planes;    // [H*W*4]; may be NULL
    const uint32_t* plane_masks;  // [H*W]; may be NULL
    float* planes;                // [H*W*4] state (in for init_mode 2, out always)
    float* costs;                 // [H*W] state (in for init_mode 2, out always)
    uint32_t* selected_views;     // [H*W] out; may be NULL
    OParams params;
    int32_t init_mode;            // 0 random, 1 from depth/normal, 2 planar prior
    int32_t geom;                 // 1: add the 0.1 * geometric-consistency term
    uint32_t seed;                // Philox key word 0
    uint32_t rng_view;            // Philox key word 1
    int32_t threads;              // OpenMP threads (<=0: default)
    int32_t skip_filter;          // 1: stop after GetDepthandNormal
    float* out_depth;             // [H*W] or NULL
    float* out_normal;            // [H*W*3] or NULL
    float* out_cost;              // [H*W] or NULL
    uint32_t compat;              // GMVS_COMPAT_* bits of the product (0: the reference's quirks)
} OPMArgs;

// The product's GMVS_COMPAT_* bits (include/gmvs_amd.h), restated.
enum : uint32_t {
    OCOMPAT_FIX_RIGHT_FAR = 1,
    OCOMPAT_FIX_COST_DEFAULT = 2,
    OCOMPAT_FIX_PRIOR_NORMALS = 4,
    OCOMPAT_FIX_PRIOR_SELECTION = 8,
};

// ---------------------------------------------------------------- Philox ---
static inline void oracle_philox_round(uint32_t c[4], const uint32_t k[2]) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * (uint64_t)c[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * (uint64_t)c[2];
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    const uint32_t n0 = hi1 ^ c[1] ^ k[0];
    const uint32_t n2 = hi0 ^ c[3] ^ k[1];
    c[0] = n0;
    c[1] = lo1;
    c[2] = n2;
    c[3] = lo0;
}

static inline void oracle_philox4x32_10_block(const uint32_t ctr[4], const uint32_t key[2],
                                              uint32_t out[4]) {
    uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
    uint32_t k[2] = {key[0], key[1]};
    for (int r = 0; r < 10; ++r) {
        if (r) {
            k[0] += 0x9E3779B9u;
            k[1] += 0xBB67AE85u;
        }
        oracle_philox_round(c, k);
    }
    out[0] = c[0];
    out[1] = c[1];
    out[2] = c[2];
    out[3] = c[3];
}

// Draw-slot layout shared with the product (fixed slots, so a flipped
// decision never shifts later draws). Counter = (pixel, phase, slot / 4, 0),
// key = (seed, view); the draw is word slot % 4 of that block.
enum {
    ORNG_PHASE_INIT = 0x40000000,
    // init phase
    ORNG_INIT_DEPTH = 0,
    ORNG_INIT_ANGLE = 1,   // 1..3 (prior-perturbed init)
    ORNG_INIT_NORMAL = 4,  // 4..35, 16 rejection pairs
    // sweep phase (phase = iteration index)
    ORNG_VIEW = 0,         // 0..14
    ORNG_DEPTH_RAND = 15,
    ORNG_DEPTH_PERT = 16,
    ORNG_PERT_ANGLE = 17,  // 17..19
    ORNG_PRIOR_ANGLE = 20, // 20..22
    ORNG_NORMAL = 24,      // 24..55, 16 rejection pairs
    ORNG_REJECT_TRIES = 16,
};
