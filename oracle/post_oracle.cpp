// TEST INFRASTRUCTURE ONLY -- the sequential CPU twin of the post-fusion
// utilities (gmvs/lib/src/planar_initer.cu:471-684). Only tests/ load it, as
// the checker of gmvs_sliding_window_normal, gmvs_filter_by_var_map and
// gmvs_fusion_planar_mask; the product never does.
//
// Each function walks the pixels in the reference's order with the
// reference's FP32 expressions (built with -ffp-contract=off), so the GPU
// results must match bit for bit. The cv::Vec3f operations are restated from
// OpenCV's published matx.hpp (version unpinned in the reference):
//   Vec3f += Vec3f      component-wise float add;
//   Vec3f /= int        multiply by the double reciprocal, round to float;
//   Vec3f /= float      multiply by the float reciprocal;
//   Vec3f::dot          float sum from 0 in index order.
#include <cstdint>
#include <vector>

namespace {

struct V3 {
    float x = 0.f, y = 0.f, z = 0.f;
};

}  // namespace

// sliding_window_normal, planar_initer.cu:471-535 (the host loop; the
// commented-out kernel beside it is not used).
extern "C" void oracle_sliding_window_normal(const int32_t* planar, const float* normal, int h, int w,
                                             int window_size, float* out) {
    const int half = window_size / 2;
    for (int row = 0; row < h; ++row) {
        for (int col = 0; col < w; ++col) {
            float* o = out + 3 * (row * w + col);
            if (planar[row * w + col] == 1) {
                o[0] = o[1] = o[2] = 0.f;
                continue;
            }
            int count = 0;
            V3 n;
            for (int r = row - half; r < row + half + 1; ++r) {
                for (int c = col - half; c < col + half + 1; ++c) {
                    if (r < 0 || r >= h || c < 0 || c >= w) continue;
                    if (planar[r * w + c] == 1) continue;
                    ++count;
                    const float* q = normal + 3 * (r * w + c);
                    n.x = n.x + q[0];
                    n.y = n.y + q[1];
                    n.z = n.z + q[2];
                }
            }
            const double ia = 1.0 / count;
            o[0] = (float)(n.x * ia);
            o[1] = (float)(n.y * ia);
            o[2] = (float)(n.z * ia);
        }
    }
}

// filter_by_var_map, planar_initer.cu:634-684.
extern "C" void oracle_filter_by_var_map(const float* var, const int32_t* seg, int h, int w,
                                         float var_thresh, int32_t* out) {
    int32_t max_id = 0;
    for (int i = 0; i < h * w; ++i) max_id = seg[i] > max_id ? seg[i] : max_id;
    std::vector<float> acc(max_id + 1, 0.f), count(max_id + 1, 0.f);
    std::vector<bool> keep(max_id + 1, false);
    for (int i = 0; i < h * w; ++i) {
        const int32_t s = seg[i];
        if (s == 0) continue;
        count[s] = count[s] + 1.f;
        acc[s] = acc[s] + var[i];
    }
    for (int32_t s = 1; s <= max_id; ++s) {
        const float mean = acc[s] / count[s];
        keep[s] = mean <= var_thresh;
    }
    for (int i = 0; i < h * w; ++i) out[i] = keep[seg[i]] ? seg[i] : 0;
}

// fusion_planar_mask, planar_initer.cu:537-632.
extern "C" void oracle_fusion_planar_mask(const int32_t* seg, const int32_t* planar, const float* normals,
                                          int rows, int cols, int32_t thresh, float nonplanar_percent,
                                          float cos_sim_thresh, float match_ratio_thresh, int32_t* out) {
    int32_t max_id = 0;
    for (int i = 0; i < rows * cols; ++i) max_id = seg[i] > max_id ? seg[i] : max_id;
    std::vector<std::vector<int>> members(max_id + 1);
    std::vector<int32_t> n_nonplanar(max_id + 1, 0), n_planar(max_id + 1, 0);
    for (int i = 0; i < rows * cols; ++i) {
        members[seg[i]].push_back(i);
        if (planar[i] == 0)
            ++n_nonplanar[seg[i]];
        else
            ++n_planar[seg[i]];
    }
    for (int i = 0; i < rows * cols; ++i) out[i] = 0;
    int32_t next_vps = 301;
    for (int32_t s = 1; s <= max_id; ++s) {
        const float nonplanar = (float)n_nonplanar[s], planar_f = (float)n_planar[s];
        const float amount = nonplanar + planar_f;
        if (amount <= (float)thresh) continue;
        int32_t label;
        if (nonplanar / amount >= nonplanar_percent) {
            V3 avg;
            for (int p : members[s]) {
                if (planar[p] != 0) continue;
                avg.x = avg.x + normals[3 * p];
                avg.y = avg.y + normals[3 * p + 1];
                avg.z = avg.z + normals[3 * p + 2];
            }
            const float ia = 1.f / nonplanar;
            avg.x = avg.x * ia;
            avg.y = avg.y * ia;
            avg.z = avg.z * ia;
            int32_t valid = 0;
            for (int p : members[s]) {
                if (planar[p] != 0) continue;
                float cs = 0.f;
                cs = cs + avg.x * normals[3 * p];
                cs = cs + avg.y * normals[3 * p + 1];
                cs = cs + avg.z * normals[3 * p + 2];
                valid += cs > cos_sim_thresh ? 1 : 0;
            }
            const float ratio = (float)valid / nonplanar;
            label = ratio > match_ratio_thresh ? s : 300;
        } else if ((double)(nonplanar / amount) < 0.25) {
            label = next_vps++;
        } else {
            label = s;
        }
        for (int p : members[s]) out[p] = label;
    }
}
