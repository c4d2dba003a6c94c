// Delaunay triangulation of the planar-prior support points, shared by the
// host construction (prior.cpp, incremental Bowyer-Watson) and the device
// construction (prior_gpu.hip, one star per point). Both evaluate the same
// exact predicates with the same symbolic perturbation, so the triangulation
// is unique and the two constructions return the same triangles.
//
// Predicates (integer coordinates in [0, 65536)):
//   orient(a, b, c)   sign of (b - a) x (c - a), exact in 64 bits;
//   in_circle(a, b, c, d) for a counter-clockwise triangle (a, b, c): the
//     sign of the lifted determinant, exact in 128 bits. Four cocircular
//     points (determinant 0) are common on an integer grid, and the
//     triangulation of such a set is not unique; each point i is lifted by
//     an extra infinitesimal eps^(i+1) (a lower index weighs more), which
//     adds eps_a orient(d,b,c) + eps_b orient(d,c,a) + eps_c orient(d,a,b)
//     - eps_d orient(a,b,c) to the determinant. The sign is then that of the
//     term of the lowest index among the four (no three of four distinct
//     cocircular points are collinear, so that term is never zero): the
//     Delaunay triangulation of the perturbed points, the same whatever the
//     insertion order or algorithm.
//
// Device construction, per support point p (points on a grid of square
// cells, at most one per cell, numbered in the cell order; the scans of one
// point shared by a group of lanes, the host running the same code with a
// group of one):
//   1. the nearest point q0 (ties: lower index) -- always a Delaunay
//      neighbour (the circle on the diameter p q0 holds no other point);
//   2. counter-clockwise from the edge (p, q): the third vertex r of the
//      triangle left of p -> q is the point left of the edge whose circle
//      through p and q is smallest on the left (one pass over candidates,
//      keeping r while no candidate lies in circle(p, q, r)). Candidates are
//      the points of the cells within 3 (then 12) cells of p; the result is
//      certified by scanning the cells under the bounding box of the
//      circle's part left of the edge (a left point inside it replaces r;
//      the new circle's left part lies in the old one's). With no left
//      candidate within 12 cells, the cells of the whole left half-plane are
//      scanned: no point there means a hull edge;
//   3. repeated from (p, r) until the star closes at q0, else clockwise from
//      q0 to the other hull edge.
// Large scans (the half-plane of a hull edge, the circles over empty
// regions) walk blocks of kBlk x kBlk cells and skip the blocks an
// occupancy count marks empty; every scan keeps a minimum under a total
// order, so the cells' visiting order does not change its result.
// A triangle is written by its lowest-index vertex; triangles are listed by
// that vertex, then by the index of the next counter-clockwise vertex.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gmvs {
namespace dt {

__host__ __device__ inline int orient(int64_t ax, int64_t ay, int64_t bx, int64_t by, int64_t cx,
                                      int64_t cy) {
    const int64_t v = (bx - ax) * (cy - ay) - (by - ay) * (cx - ax);
    return (v > 0) - (v < 0);
}

// true when d lies inside the circumcircle of the counter-clockwise
// triangle (a, b, c) of the perturbed points (ia..id: point indices).
__host__ __device__ inline bool in_circle(int64_t ax, int64_t ay, int ia, int64_t bx, int64_t by,
                                          int ib, int64_t cx, int64_t cy, int ic, int64_t dx,
                                          int64_t dy, int id) {
    typedef __int128 i128;
    const i128 adx = ax - dx, ady = ay - dy;
    const i128 bdx = bx - dx, bdy = by - dy;
    const i128 cdx = cx - dx, cdy = cy - dy;
    const i128 ad = adx * adx + ady * ady;
    const i128 bd = bdx * bdx + bdy * bdy;
    const i128 cd = cdx * cdx + cdy * cdy;
    const i128 det = ad * (bdx * cdy - bdy * cdx) - bd * (adx * cdy - ady * cdx) +
                     cd * (adx * bdy - ady * bdx);
    if (det != 0) return det > 0;
    // cocircular: the lowest index among the four decides
    int lo = ia, s;
    s = orient(dx, dy, bx, by, cx, cy);  // coefficient of eps_a
    if (ib < lo) {
        lo = ib;
        s = orient(dx, dy, cx, cy, ax, ay);
    }
    if (ic < lo) {
        lo = ic;
        s = orient(dx, dy, ax, ay, bx, by);
    }
    if (id < lo) s = -orient(ax, ay, bx, by, cx, cy);
    return s > 0;
}

// The support points as a cell grid: cell (cx, cy) at cell index
// cx * ncy + cy holds point pidx[cell] (or -1) at (px[p], py[p]). bcnt
// counts the points of each block of kBlk x kBlk cells (block (bx, by) at
// bx * nby + by, nby = ceil(ncy / kBlk)); nullptr scans every cell.
constexpr int kBlk = 8;
// Ranges of at least this many cells are scanned block by block.
constexpr int kBlockScanCells = 512;

struct Grid {
    const int* pidx;  // [ncx * ncy]
    const int* px;    // [npts]
    const int* py;
    const int* bcnt;  // [ceil(ncx / kBlk) * nby] or nullptr
    int ncx, ncy, cell, nby;
};

// Points per block of kBlk x kBlk cells, for Grid::bcnt (host).
inline void block_counts(const int* pidx, int ncx, int ncy, int* bcnt) {
    const int nbx = (ncx + kBlk - 1) / kBlk, nby = (ncy + kBlk - 1) / kBlk;
    for (int i = 0; i < nbx * nby; ++i) bcnt[i] = 0;
    for (int cx = 0; cx < ncx; ++cx)
        for (int cy = 0; cy < ncy; ++cy)
            if (pidx[cx * ncy + cy] >= 0) ++bcnt[(cx / kBlk) * nby + cy / kBlk];
}

// A point's scans are shared by a group of L lanes (L = 1 on the host): the
// lanes take the cells of a range in turn along its longer side, and the
// group combines its lanes' results; L is a power of two, the group a
// contiguous part of a wave.
template <int L>
struct Lanes {
    int lane;  // 0 .. L-1
};

// The lower of two candidates in the pencil of circles through p and q (r
// and o left of p -> q, -1 = none).
__host__ __device__ inline int pencil_min(const Grid& g, int p, int q, int r, int o) {
    if (o < 0 || o == r) return r;
    if (r < 0) return o;
    return in_circle(g.px[p], g.py[p], p, g.px[q], g.py[q], q, g.px[r], g.py[r], r, g.px[o],
                     g.py[o], o)
               ? o
               : r;
}

// The group's lowest candidate, on every lane.
template <int L>
__host__ __device__ inline int group_pencil_min(const Grid& g, int p, int q, int r, Lanes<L> ln) {
    (void)ln;
    if constexpr (L > 1) {
        for (int off = L / 2; off > 0; off >>= 1) r = pencil_min(g, p, q, r, __shfl_xor(r, off, L));
    }
    return r;
}

// Calls visit(cx, cy) on this lane's share of the cells [x0, x1] x
// [y0, y1] (the group's lanes cover each cell once). Small ranges: the lanes
// take the cells in turn along the range's longer side. Large ranges with
// block counts: block by block, skipping empty blocks, the lanes taking the
// cells of a block in turn.
template <int L, class V>
__host__ __device__ inline void for_cells(const Grid& g, int x0, int x1, int y0, int y1,
                                          Lanes<L> ln, V&& visit) {
    if (g.bcnt && (x1 - x0 + 1) * (y1 - y0 + 1) >= kBlockScanCells) {
        for (int bx = x0 / kBlk; bx <= x1 / kBlk; ++bx)
            for (int by = y0 / kBlk; by <= y1 / kBlk; ++by) {
                if (g.bcnt[bx * g.nby + by] == 0) continue;
                for (int k = ln.lane; k < kBlk * kBlk; k += L) {
                    const int cx = bx * kBlk + k % kBlk, cy = by * kBlk + k / kBlk;
                    if (cx >= x0 && cx <= x1 && cy >= y0 && cy <= y1) visit(cx, cy);
                }
            }
    } else if (x1 - x0 > y1 - y0) {
        for (int cy = y0; cy <= y1; ++cy)
            for (int cx = x0 + ln.lane; cx <= x1; cx += L) visit(cx, cy);
    } else {
        for (int cx = x0; cx <= x1; ++cx)
            for (int cy = y0 + ln.lane; cy <= y1; cy += L) visit(cx, cy);
    }
}

// Grows r over the points of the cells [x0, x1] x [y0, y1]: the point left
// of p -> q whose circle through p and q is smallest on the left (r = -1:
// none yet). Every lane of the group returns the same r.
template <int L, bool REDUCE = true>
__host__ __device__ inline int left_min(const Grid& g, int p, int q, int x0, int x1, int y0, int y1,
                                        int r, Lanes<L> ln) {
    const int64_t pxv = g.px[p], pyv = g.py[p], qxv = g.px[q], qyv = g.py[q];
    for_cells(g, x0, x1, y0, y1, ln, [&](int cx, int cy) {
        const int s = g.pidx[cx * g.ncy + cy];
        if (s < 0 || s == p || s == q || s == r) return;
        const int64_t sx = g.px[s], sy = g.py[s];
        if (orient(pxv, pyv, qxv, qyv, sx, sy) <= 0) return;
        if (r < 0 || in_circle(pxv, pyv, p, qxv, qyv, q, g.px[r], g.py[r], r, sx, sy, s)) r = s;
    });
    if constexpr (L > 1 && REDUCE) r = group_pencil_min(g, p, q, r, ln);
    return r;
}

// Cell range of the square of R cells around cell (cx, cy), clipped.
__host__ __device__ inline void square(const Grid& g, int cx, int cy, int R, int& x0, int& x1,
                                       int& y0, int& y1) {
    x0 = cx - R < 0 ? 0 : cx - R;
    y0 = cy - R < 0 ? 0 : cy - R;
    x1 = cx + R >= g.ncx ? g.ncx - 1 : cx + R;
    y1 = cy + R >= g.ncy ? g.ncy - 1 : cy + R;
}

// Cells [c0, c1] (clipped to [0, n - 1]) under the pixel interval [lo, hi].
__host__ __device__ inline void cell_span(double lo, double hi, int cell, int n, int& c0, int& c1) {
    const double a = lo / cell, b = hi / cell;
    c0 = !(a > 0.0) ? 0 : (a > n - 1 ? n - 1 : (int)a);
    c1 = !(b > 0.0) ? 0 : (b > n - 1 ? n - 1 : (int)b);
}

// Grows r over every point left of p -> q: the cells of each band of rows
// (one cell row, or kBlk rows when the grid has block counts) that can hold
// a point of the open half-plane (an interval from the line's crossing of
// the band's pixel rows, widened by one cell).
template <int L>
__host__ __device__ inline int left_min_halfplane(const Grid& g, int p, int q, int r, Lanes<L> ln) {
    const double px = g.px[p], py = g.py[p], qx = g.px[q], qy = g.py[q];
    // s is left of p -> q when (qx - px)(sy - py) - (qy - py)(sx - px) > 0
    const double a = -(qy - py);  // coefficient of sx
    const int band = g.bcnt ? kBlk : 1;
    for (int cy0 = 0; cy0 < g.ncy; cy0 += band) {
        const int cy1 = cy0 + band - 1 < g.ncy ? cy0 + band - 1 : g.ncy - 1;
        const double ylo = (double)cy0 * g.cell, yhi = (double)cy1 * g.cell + (g.cell - 1);
        int x0 = 0, x1 = g.ncx - 1;
        if (a == 0.0) {
            // horizontal edge: the whole band or none of it
            const double f0 = (qx - px) * (ylo - py), f1 = (qx - px) * (yhi - py);
            if (!(f0 > 0.0) && !(f1 > 0.0)) continue;
        } else {
            // a sx + (qx - px)(sy - py) + (qy - py) px > 0: sx beyond t(sy)
            const double t0 = -((qx - px) * (ylo - py) + (qy - py) * px) / a;
            const double t1 = -((qx - px) * (yhi - py) + (qy - py) * px) / a;
            if (a > 0.0) {  // sx > t
                const double t = (t0 < t1 ? t0 : t1) / g.cell - 1.0;
                x0 = !(t > 0.0) ? 0 : (t > g.ncx - 1 ? g.ncx : (int)t);
            } else {  // sx < t
                const double t = (t0 > t1 ? t0 : t1) / g.cell + 1.0;
                x1 = !(t > 0.0) ? -1 : (t > g.ncx - 1 ? g.ncx - 1 : (int)t);
            }
            if (x0 > x1) continue;
        }
        r = left_min<L, false>(g, p, q, x0, x1, cy0, cy1, r, ln);  // lane-local
    }
    return group_pencil_min(g, p, q, r, ln);
}

// The third vertex of the Delaunay triangle left of the Delaunay edge
// p -> q, or -1 on a hull edge.
template <int L>
__host__ __device__ inline int left_vertex(const Grid& g, int p, int q, Lanes<L> ln) {
    const int pcx = g.px[p] / g.cell, pcy = g.py[p] / g.cell;
    int r = -1;
    int x0, x1, y0, y1;
    for (int R = 3; R <= 12; R *= 4) {
        square(g, pcx, pcy, R, x0, x1, y0, y1);
        r = left_min(g, p, q, x0, x1, y0, y1, r, ln);
        if (r >= 0) break;
    }
    if (r < 0) {
        // nothing within 12 cells: the minimum over the whole half-plane
        // (a hull edge when it is empty), certified by construction
        return left_min_halfplane(g, p, q, r, ln);
    }
    // certification: every left point inside circle(p, q, r) lies under the
    // bounding box of the circle's left part. When that box leaves the
    // scanned square, one scan of its cells leaves r minimal over all
    // points: a point replacing r lies inside the old circle, and the left
    // part of each new circle lies inside the old one's
    const double ax = g.px[p], ay = g.py[p], bx = g.px[q], by = g.py[q];
    const double cxv = g.px[r], cyv = g.py[r];
    const double d = 2.0 * (ax * (by - cyv) + bx * (cyv - ay) + cxv * (ay - by));
    const double a2 = ax * ax + ay * ay, b2 = bx * bx + by * by, c2 = cxv * cxv + cyv * cyv;
    const double ux = (a2 * (by - cyv) + b2 * (cyv - ay) + c2 * (ay - by)) / d;
    const double uy = (a2 * (cxv - bx) + b2 * (ax - cxv) + c2 * (bx - ax)) / d;
    const double rad = sqrt((ax - ux) * (ax - ux) + (ay - uy) * (ay - uy));
    // the centre's side of p -> q; right of it, the left part is the cap
    // cut off by the chord pq, within its height of the chord
    const double ex = bx - ax, ey = by - ay;
    const double side = ex * (uy - ay) - ey * (ux - ax);
    double lx0, lx1, ly0, ly1;
    if (side < 0.0) {
        const double hcap = rad - (-side) / sqrt(ex * ex + ey * ey) + 1.0;  // + FP64 slack
        lx0 = (ax < bx ? ax : bx) - hcap;
        lx1 = (ax > bx ? ax : bx) + hcap;
        ly0 = (ay < by ? ay : by) - hcap;
        ly1 = (ay > by ? ay : by) + hcap;
    } else {
        const double rr = rad + 1.0;  // + FP64 slack
        lx0 = ux - rr;
        lx1 = ux + rr;
        ly0 = uy - rr;
        ly1 = uy + rr;
    }
    int bx0, bx1, by0, by1;
    cell_span(lx0, lx1, g.cell, g.ncx, bx0, bx1);
    cell_span(ly0, ly1, g.cell, g.ncy, by0, by1);
    if (bx0 >= x0 && bx1 <= x1 && by0 >= y0 && by1 <= y1) return r;
    return left_min(g, p, q, bx0, bx1, by0, by1, r, ln);
}

// The nearest other point of p (ties: lower index), or -1; the same on
// every lane of the group.
template <int L>
__host__ __device__ inline int nearest(const Grid& g, int p, Lanes<L> ln) {
    const int pcx = g.px[p] / g.cell, pcy = g.py[p] / g.cell;
    const int64_t pxv = g.px[p], pyv = g.py[p];
    int best = -1;
    int64_t bd = 0;
    int x0, x1, y0, y1;
    auto better = [](int64_t d2, int s, int64_t bd2, int b) {
        return s >= 0 && (b < 0 || d2 < bd2 || (d2 == bd2 && s < b));
    };
    for (int R = 2;; R *= 4) {
        square(g, pcx, pcy, R, x0, x1, y0, y1);
        for_cells(g, x0, x1, y0, y1, ln, [&](int cx, int cy) {
            const int s = g.pidx[cx * g.ncy + cy];
            if (s < 0 || s == p) return;
            const int64_t dxv = g.px[s] - pxv, dyv = g.py[s] - pyv;
            const int64_t d2 = dxv * dxv + dyv * dyv;
            if (better(d2, s, bd, best)) {
                best = s;
                bd = d2;
            }
        });
        if constexpr (L > 1) {
            for (int off = L / 2; off > 0; off >>= 1) {
                const int ob = __shfl_xor(best, off, L);
                const int64_t od = __shfl_xor(bd, off, L);
                if (better(od, ob, bd, best)) {
                    best = ob;
                    bd = od;
                }
            }
        }
        const bool whole = x0 == 0 && y0 == 0 && x1 == g.ncx - 1 && y1 == g.ncy - 1;
        if (whole) return best;
        // certified once every cell within sqrt(bd) of p has been scanned
        if (best >= 0) {
            const double rad = sqrt((double)bd) + 1.0;  // + FP64 slack
            int cx0, cx1, cy0, cy1;
            cell_span((double)pxv - rad, (double)pxv + rad, g.cell, g.ncx, cx0, cx1);
            cell_span((double)pyv - rad, (double)pyv + rad, g.cell, g.ncy, cy0, cy1);
            if (cx0 >= x0 && cx1 <= x1 && cy0 >= y0 && cy1 <= y1) return best;
        }
    }
}

// Walks the star of p and calls emit(q, r) for each triangle (p, q, r)
// (counter-clockwise) in which p is the lowest index. Returns false when the
// walk exceeds max_steps (a broken invariant; the caller reports it).
template <int L, class F>
__host__ __device__ inline bool star(const Grid& g, int p, int max_steps, Lanes<L> ln, F&& emit) {
    const int q0 = nearest(g, p, ln);
    if (q0 < 0) return true;
    int q = q0, steps = 0;
    bool closed = false;
    for (;;) {  // counter-clockwise
        if (++steps > max_steps) return false;
        const int r = left_vertex(g, p, q, ln);
        if (r < 0) break;
        if (p < q && p < r) emit(q, r);
        if (r == q0) {
            closed = true;
            break;
        }
        q = r;
    }
    if (closed) return true;
    q = q0;
    for (;;) {  // clockwise: the triangle (p, r, q) right of p -> q
        if (++steps > max_steps) return false;
        const int r = left_vertex(g, q, p, ln);  // left of q -> p: triangle (q, p, r) = (p, r, q)
        if (r < 0) break;
        if (p < q && p < r) emit(r, q);
        q = r;
    }
    return true;
}

}  // namespace dt
}  // namespace gmvs
