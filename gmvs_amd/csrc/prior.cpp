// Host construction of the ACMP planar prior between the two first-stage
// passes -- PlanarInit::get_support_points, triangulation, process_mask,
// get_prior_plane_params and cuda_planar_prior_init
// (gmvs/lib/src/planar_initer.cu:195-420, 686-745) without OpenCV:
//   * support points: per 5x5 cell the lowest-cost pixel with cost < 2,
//     kept when that cost is < 0.1;
//   * Delaunay triangulation: incremental Bowyer-Watson with exact integer
//     orientation / in-circle predicates, points inserted in Hilbert order,
//     the containing triangle found by walking from the last insertion;
//   * superpixel filter over int32 labels (the reference reads uint8 labels,
//     planar_initer.cu:385, which cannot hold 2000 cells);
//   * 3-point plane (n, d) with |n| = 1 and d >= 0 (the null vector of
//     [X 1] that cv::SVD::solveZ returns, normalised as at :732-743);
//   * barycentric rasterisation into per-pixel triangle ids, dropping pixels
//     whose prior depth leaves [depth_min, depth_max].
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "abi_common.h"
#include "delaunay.h"

namespace gmvs {
namespace {

// A vertex (x + sx*M, y + sy*M). Input points have sx = sy = 0; the three
// super-triangle vertices are at infinity (M -> infinity): their predicates
// are evaluated as polynomials in M and take the sign of the leading nonzero
// coefficient. With the bounding triangle infinitely far away, every triangle
// of input points whose circumcircle is empty survives, including the thin
// triangles along the convex hull that a finite bounding triangle loses.
struct Pt {
    int64_t x, y;
    int sx = 0, sy = 0;
    bool real() const { return sx == 0 && sy == 0; }
};

typedef __int128 i128;

// c[0] + c[1]*M + ... ; the sign of the polynomial as M -> infinity.
template <int N>
inline int limit_sign(const i128 (&c)[N]) {
    for (int i = N - 1; i >= 0; --i)
        if (c[i] != 0) return c[i] > 0 ? 1 : -1;
    return 0;
}

struct Lin {  // c0 + c1*M
    i128 c0, c1;
};
inline Lin sub(const Pt& a, const Pt& b) { return Lin{a.x - b.x, i128(a.sx - b.sx)}; }
inline void mul(const Lin& a, const Lin& b, i128 (&o)[3]) {
    o[0] = a.c0 * b.c0;
    o[1] = a.c0 * b.c1 + a.c1 * b.c0;
    o[2] = a.c1 * b.c1;
}
inline Lin suby(const Pt& a, const Pt& b) { return Lin{a.y - b.y, i128(a.sy - b.sy)}; }

// sign of (b - a) x (c - a): > 0 for a counter-clockwise triangle in x-right,
// y-up axes.
inline int orient(const Pt& a, const Pt& b, const Pt& c) {
    if (a.real() && b.real() && c.real()) {
        const int64_t v = (b.x - a.x) * (c.y - a.y) - (b.y - a.y) * (c.x - a.x);
        return (v > 0) - (v < 0);
    }
    i128 p[3], q[3];
    mul(sub(b, a), suby(c, a), p);
    mul(suby(b, a), sub(c, a), q);
    const i128 r[3] = {p[0] - q[0], p[1] - q[1], p[2] - q[2]};
    return limit_sign(r);
}

// true when the input point d lies inside the circumcircle of the
// positively oriented triangle (a, b, c); ia..id are the point indices,
// which order the symbolic perturbation of cocircular input points
// (dt::in_circle, shared with the device construction).
inline bool in_circle(const Pt& a, int ia, const Pt& b, int ib, const Pt& c, int ic, const Pt& d,
                      int id) {
    if (a.real() && b.real() && c.real())
        return dt::in_circle(a.x, a.y, ia, b.x, b.y, ib, c.x, c.y, ic, d.x, d.y, id);
    const Lin X[3] = {sub(a, d), sub(b, d), sub(c, d)};
    const Lin Y[3] = {suby(a, d), suby(b, d), suby(c, d)};
    i128 lift[3][3], cross[3][3];  // lift[k] = |k - d|^2; cross[k] = the 2x2 minor without k
    for (int k = 0; k < 3; ++k) {
        i128 xx[3], yy[3];
        mul(X[k], X[k], xx);
        mul(Y[k], Y[k], yy);
        for (int j = 0; j < 3; ++j) lift[k][j] = xx[j] + yy[j];
        const int u = (k + 1) % 3, v = (k + 2) % 3;
        i128 p[3], q[3];
        mul(X[u], Y[v], p);
        mul(Y[u], X[v], q);
        for (int j = 0; j < 3; ++j) cross[k][j] = p[j] - q[j];
    }
    // det = sum_k lift[k] * cross[k] (cyclic minors carry the cofactor signs)
    i128 det[5] = {0, 0, 0, 0, 0};
    for (int k = 0; k < 3; ++k)
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) det[i + j] += lift[k][i] * cross[k][j];
    return limit_sign(det) > 0;
}

// Hilbert index of (x, y) on a 2^16 grid (the standard xy -> d
// construction: per level, add the quadrant's offset, then rotate/reflect).
uint64_t hilbert_index(uint32_t x, uint32_t y) {
    const uint32_t n = 1u << 16;
    uint64_t d = 0;
    for (uint32_t s = n / 2; s > 0; s /= 2) {
        const uint32_t rx = (x & s) ? 1u : 0u;
        const uint32_t ry = (y & s) ? 1u : 0u;
        d += uint64_t(s) * s * ((3u * rx) ^ ry);
        if (ry == 0) {
            if (rx == 1) {
                x = n - 1 - x;
                y = n - 1 - y;
            }
            std::swap(x, y);
        }
    }
    return d;
}

struct Tri {
    int v[3];    // counter-clockwise
    int nb[3];   // neighbour across the edge opposite v[i], -1 = none
    bool alive;
};

class Delaunay {
public:
    explicit Delaunay(const std::vector<Pt>& pts) : p_(pts) {
        n_ = int(pts.size());
        p_.push_back(Pt{0, 0, -1, -1});  // the bounding triangle, at infinity
        p_.push_back(Pt{0, 0, 1, -1});
        p_.push_back(Pt{0, 0, 0, 1});
        Tri t;
        t.v[0] = n_;
        t.v[1] = n_ + 1;
        t.v[2] = n_ + 2;
        t.nb[0] = t.nb[1] = t.nb[2] = -1;
        t.alive = true;
        tris_.push_back(t);
    }

    void run() {
        std::vector<int> order(n_);
        for (int i = 0; i < n_; ++i) order[i] = i;
        std::vector<uint64_t> key(n_);
        for (int i = 0; i < n_; ++i)
            key[i] = hilbert_index(uint32_t(p_[i].x) & 0xFFFFu, uint32_t(p_[i].y) & 0xFFFFu);
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return key[a] < key[b]; });
        for (int i : order) insert(i);
    }

    // Triangles whose three vertices are input points, counter-clockwise
    // from their lowest index, listed by that index and then by the next
    // vertex's (the device construction's order).
    std::vector<int> triangles() const {
        std::vector<std::array<int, 3>> t3;
        for (const Tri& t : tris_) {
            if (!t.alive) continue;
            if (t.v[0] >= n_ || t.v[1] >= n_ || t.v[2] >= n_) continue;
            int k = 0;
            if (t.v[1] < t.v[k]) k = 1;
            if (t.v[2] < t.v[k]) k = 2;
            t3.push_back({t.v[k], t.v[(k + 1) % 3], t.v[(k + 2) % 3]});
        }
        std::sort(t3.begin(), t3.end());
        std::vector<int> out;
        out.reserve(3 * t3.size());
        for (const auto& t : t3) out.insert(out.end(), t.begin(), t.end());
        return out;
    }

private:
    std::vector<Pt> p_;
    int n_;
    std::vector<Tri> tris_;
    std::vector<int> free_;
    int last_ = 0;

    int locate(const Pt& q) {
        int t = last_;
        if (t < 0 || t >= int(tris_.size()) || !tris_[t].alive) {
            t = -1;
            for (int i = int(tris_.size()) - 1; i >= 0; --i)
                if (tris_[i].alive) {
                    t = i;
                    break;
                }
        }
        for (size_t steps = 0; steps < 4 * tris_.size() + 16; ++steps) {
            const Tri& T = tris_[t];
            int next = -1;
            // rotate the starting edge to avoid cycling on degenerate walks
            for (int e0 = 0; e0 < 3 && next < 0; ++e0) {
                const int e = (e0 + int(steps)) % 3;
                const Pt& a = p_[T.v[(e + 1) % 3]];
                const Pt& b = p_[T.v[(e + 2) % 3]];
                if (orient(a, b, q) < 0) next = T.nb[e];
            }
            if (next < 0) return t;
            t = next;
        }
        // fall back to an exhaustive search
        for (int i = 0; i < int(tris_.size()); ++i) {
            const Tri& T = tris_[i];
            if (!T.alive) continue;
            if (orient(p_[T.v[0]], p_[T.v[1]], q) >= 0 && orient(p_[T.v[1]], p_[T.v[2]], q) >= 0 &&
                orient(p_[T.v[2]], p_[T.v[0]], q) >= 0)
                return i;
        }
        return -1;
    }

    int new_tri(int a, int b, int c) {
        Tri t;
        t.v[0] = a;
        t.v[1] = b;
        t.v[2] = c;
        t.nb[0] = t.nb[1] = t.nb[2] = -1;
        t.alive = true;
        if (!free_.empty()) {
            const int i = free_.back();
            free_.pop_back();
            tris_[i] = t;
            return i;
        }
        tris_.push_back(t);
        return int(tris_.size()) - 1;
    }

    void insert(int pi) {
        const Pt& q = p_[pi];
        const int t0 = locate(q);
        if (t0 < 0) return;
        for (int k = 0; k < 3; ++k)
            if (p_[tris_[t0].v[k]].real() && p_[tris_[t0].v[k]].x == q.x &&
                p_[tris_[t0].v[k]].y == q.y)
                return;  // duplicate
        // cavity: triangles whose circumcircle contains q, grown from t0
        std::vector<int> bad{t0};
        std::vector<int> stack{t0};
        mark_.resize(tris_.size(), 0);
        ++stamp_;
        if (stamp_ == 0) {
            std::fill(mark_.begin(), mark_.end(), 0);
            stamp_ = 1;
        }
        mark_[t0] = stamp_;
        while (!stack.empty()) {
            const int t = stack.back();
            stack.pop_back();
            for (int e = 0; e < 3; ++e) {
                const int n = tris_[t].nb[e];
                if (n < 0 || mark_[n] == stamp_) continue;
                const Tri& N = tris_[n];
                if (in_circle(p_[N.v[0]], N.v[0], p_[N.v[1]], N.v[1], p_[N.v[2]], N.v[2], q, pi)) {
                    mark_[n] = stamp_;
                    bad.push_back(n);
                    stack.push_back(n);
                }
            }
        }
        // boundary edges (a, b) of the cavity, counter-clockwise, with the
        // outside neighbour
        struct Edge {
            int a, b, out;
        };
        std::vector<Edge> edges;
        for (int t : bad)
            for (int e = 0; e < 3; ++e) {
                const int n = tris_[t].nb[e];
                if (n >= 0 && mark_[n] == stamp_) continue;
                edges.push_back(Edge{tris_[t].v[(e + 1) % 3], tris_[t].v[(e + 2) % 3], n});
            }
        for (int t : bad) {
            tris_[t].alive = false;
            free_.push_back(t);
        }
        // fan of new triangles (a, b, q); neighbours: across (a,b) = out,
        // across (b,q) = the fan triangle starting at b, across (q,a) = the
        // one ending at a
        std::vector<int> made(edges.size());
        for (size_t i = 0; i < edges.size(); ++i) {
            const int t = new_tri(edges[i].a, edges[i].b, pi);
            made[i] = t;
            tris_[t].nb[2] = edges[i].out;
            if (edges[i].out >= 0) {
                Tri& O = tris_[edges[i].out];
                for (int e = 0; e < 3; ++e) {
                    const int oa = O.v[(e + 1) % 3], ob = O.v[(e + 2) % 3];
                    if (oa == edges[i].b && ob == edges[i].a) O.nb[e] = t;
                }
            }
        }
        if (mark_.size() < tris_.size()) mark_.resize(tris_.size(), 0);
        // link fan triangles: map start vertex -> fan triangle
        std::vector<std::pair<int, int>> by_a(edges.size()), by_b(edges.size());
        for (size_t i = 0; i < edges.size(); ++i) {
            by_a[i] = {edges[i].a, made[i]};
            by_b[i] = {edges[i].b, made[i]};
        }
        std::sort(by_a.begin(), by_a.end());
        for (size_t i = 0; i < edges.size(); ++i) {
            const int t = made[i];
            // edge opposite v[0]=a is (b, q): neighbour is the fan triangle whose a == b
            auto it = std::lower_bound(by_a.begin(), by_a.end(), std::make_pair(edges[i].b, -1));
            if (it != by_a.end() && it->first == edges[i].b) tris_[t].nb[0] = it->second;
        }
        std::sort(by_b.begin(), by_b.end());
        for (size_t i = 0; i < edges.size(); ++i) {
            const int t = made[i];
            // edge opposite v[1]=b is (q, a): neighbour is the fan triangle whose b == a
            auto it = std::lower_bound(by_b.begin(), by_b.end(), std::make_pair(edges[i].a, -1));
            if (it != by_b.end() && it->first == edges[i].a) tris_[t].nb[1] = it->second;
        }
        last_ = made.empty() ? 0 : made[0];
    }

    std::vector<uint32_t> mark_;
    uint32_t stamp_ = 0;
};

struct IPt {
    int x, y;
};

// get_support_points (planar_initer.cu:225-256).
std::vector<IPt> support_points(const float* cost, int h, int w) {
    std::vector<IPt> pts;
    const int step = 5;
    for (int col = 0; col < w; col += step)
        for (int row = 0; row < h; row += step) {
            float mc = 2.0f;
            IPt best{0, 0};
            const int cb = std::min(w, col + step), rb = std::min(h, row + step);
            for (int c = col; c < cb; ++c)
                for (int r = row; r < rb; ++r) {
                    const float v = cost[r * w + c];
                    if (v < 2.0f && mc > v) {
                        best = IPt{c, r};
                        mc = v;
                    }
                }
            if (mc < 0.1f) pts.push_back(best);
        }
    return pts;
}

// Barycentric walk over a triangle as the reference rasterises it
// (planar_initer.cu:313-324 and 372-380): float steps of 1 / longest edge.
template <class F>
void walk_triangle(const IPt& a, const IPt& b, const IPt& c, F&& visit) {
    auto len = [](int dx, int dy) { return float(std::sqrt(double(dx) * dx + double(dy) * dy)); };
    const float L01 = len(a.x - b.x, a.y - b.y);
    const float L02 = len(a.x - c.x, a.y - c.y);
    const float L12 = len(b.x - c.x, b.y - c.y);
    const float mx = std::max(L01, std::max(L02, L12));
    const float step = float(1.0 / mx);
    for (float p = 0; p < 1.0; p += step) {
        for (float q = 0; q < 1.0 - p; q += step) {
            const int x = int(p * a.x + q * b.x + (1.0 - p - q) * c.x);
            const int y = int(p * a.y + q * b.y + (1.0 - p - q) * c.y);
            if (!visit(x, y)) return;
        }
    }
}

}  // namespace

// Delaunay triangles (counter-clockwise vertex index triples) of distinct
// integer points in [0, 65536)^2, for the device prior construction.
std::vector<int> host_delaunay(const int32_t* xy, int n) {
    std::vector<Pt> pts(n);
    for (int i = 0; i < n; ++i) pts[i] = Pt{xy[2 * i], xy[2 * i + 1]};
    Delaunay d(pts);
    d.run();
    return d.triangles();
}

}  // namespace gmvs

using namespace gmvs;

extern "C" int gmvs_delaunay(const int32_t* xy, int32_t n, int32_t* tri_out, int32_t max_tri,
                             int32_t* num_tri) {
    clear_error();
    GMVS_ARG_CHECK(n >= 0 && (n == 0 || xy) && num_tri, "delaunay: bad arguments");
    std::vector<Pt> pts(n);
    for (int i = 0; i < n; ++i) {
        GMVS_ARG_CHECK(xy[2 * i] >= 0 && xy[2 * i] < 65536 && xy[2 * i + 1] >= 0 &&
                           xy[2 * i + 1] < 65536,
                       "delaunay: coordinates must lie in [0, 65536)");
        pts[i] = Pt{xy[2 * i], xy[2 * i + 1]};
    }
    Delaunay d(pts);
    d.run();
    const std::vector<int> t = d.triangles();
    const int nt = int(t.size() / 3);
    *num_tri = nt;
    if (tri_out)
        for (int i = 0; i < std::min(nt, max_tri) * 3; ++i) tri_out[i] = t[i];
    return GMVS_OK;
}

extern "C" int gmvs_prior_build(const gmvs_camera* ref_camera, const float* depth,
                                const float* cost, const int32_t* labels,
                                const gmvs_params* params, float* prior_planes,
                                uint32_t* plane_masks, int32_t* out_num_triangles) {
    clear_error();
    GMVS_ARG_CHECK(ref_camera && depth && cost && params && prior_planes && plane_masks,
                   "prior_build: NULL argument");
    const int h = ref_camera->height, w = ref_camera->width;
    GMVS_ARG_CHECK(h > 0 && w > 0 && w < 65536 && h < 65536, "prior_build: bad size %dx%d", w, h);
    GMVS_ARG_CHECK(!params->superpixel_filter || labels,
                   "prior_build: superpixel_filter needs labels");
    const std::vector<IPt> sp = support_points(cost, h, w);
    std::vector<Pt> pts(sp.size());
    for (size_t i = 0; i < sp.size(); ++i) pts[i] = Pt{sp[i].x, sp[i].y};
    Delaunay del(pts);
    del.run();
    const std::vector<int> tv = del.triangles();

    // superpixel filter (planar_initer.cu:362-410)
    std::vector<std::array<IPt, 3>> tris;
    for (size_t i = 0; i < tv.size(); i += 3) {
        const IPt a = sp[tv[i]], b = sp[tv[i + 1]], c = sp[tv[i + 2]];
        bool retain = true;
        if (params->superpixel_filter) {
            float count = 0.f, zero = 0.f;
            int32_t seg = 0;
            walk_triangle(a, b, c, [&](int x, int y) {
                const int32_t s = labels[y * w + x];
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
        if (retain) tris.push_back({a, b, c});
    }

    // process_mask + get_prior_plane_params (planar_initer.cu:290-345, 700-745)
    std::vector<uint32_t> mask(size_t(h) * w, 0u);
    std::vector<float> tplanes;
    const float* K = ref_camera->K;
    uint32_t idx = 0;
    for (const auto& t : tris) {
        walk_triangle(t[0], t[1], t[2], [&](int x, int y) {
            if (x >= 0 && x < w && y >= 0 && y < h) mask[size_t(y) * w + x] = idx + 1;
            return true;
        });
        double X[3][3];
        for (int k = 0; k < 3; ++k) {
            const double d = depth[t[k].y * w + t[k].x];
            X[k][0] = d * (t[k].x - double(K[2])) / double(K[0]);
            X[k][1] = d * (t[k].y - double(K[5])) / double(K[4]);
            X[k][2] = d;
        }
        const double e1[3] = {X[1][0] - X[0][0], X[1][1] - X[0][1], X[1][2] - X[0][2]};
        const double e2[3] = {X[2][0] - X[0][0], X[2][1] - X[0][1], X[2][2] - X[0][2]};
        double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                       e1[0] * e2[1] - e1[1] * e2[0]};
        double nn = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        double dd = -(n[0] * X[0][0] + n[1] * X[0][1] + n[2] * X[0][2]);
        if (dd < 0) nn = -nn;
        tplanes.push_back(float(n[0] / nn));
        tplanes.push_back(float(n[1] / nn));
        tplanes.push_back(float(n[2] / nn));
        tplanes.push_back(float(dd / nn));
        ++idx;
    }
    // prior depth range check (planar_initer.cu:330-344)
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            const size_t i = size_t(y) * w + x;
            const uint32_t m = mask[i];
            float* out = prior_planes + 4 * i;
            if (m == 0) {
                out[0] = out[1] = out[2] = out[3] = 0.f;
                plane_masks[i] = 0;
                continue;
            }
            const float* pl = &tplanes[4 * size_t(m - 1)];
            const float d = -pl[3] / ((x - K[2]) * pl[0] / K[0] + (y - K[5]) * pl[1] / K[4] + pl[2]);
            if (d <= params->depth_max && d >= params->depth_min) {
                plane_masks[i] = m;
                out[0] = pl[0];
                out[1] = pl[1];
                out[2] = pl[2];
                out[3] = pl[3];
            } else {
                plane_masks[i] = 0;
                out[0] = out[1] = out[2] = out[3] = 0.f;
            }
        }
    if (out_num_triangles) *out_num_triangles = int32_t(tris.size());
    return GMVS_OK;
}
