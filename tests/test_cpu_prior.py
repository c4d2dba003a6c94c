"""CPU tests of the host planar-prior construction (support points, Delaunay,
superpixel filter, plane fit, rasterisation; gmvs/lib/src/planar_initer.cu:
195-420,686-745).

The Delaunay triangulation is pinned by its defining property (no input point
strictly inside any triangle's circumcircle, checked with exact integer
arithmetic), by Euler's count T = 2n - 2 - h, and against scipy's Qhull on
point sets in general position, where the triangulation is unique. The plane
fit is pinned by depth maps of exact planes, whose fitted planes must equal
the generating plane.
"""
import numpy as np
import pytest

from gmvs_amd import _native as N
from gmvs_amd.lib.build.planar_initer import build_prior, delaunay, delaunay_cells


def _random_points(n, seed, w=1600, h=1200):
    rng = np.random.default_rng(seed)
    pts = np.unique(np.stack([rng.integers(0, w, 4 * n), rng.integers(0, h, 4 * n)], 1), axis=0)
    rng.shuffle(pts)
    return pts[:n].astype(np.int64)


def _incircle_violations(pts, tri):
    """Number of (triangle, point) pairs with the point strictly inside the
    triangle's circumcircle (exact int64 determinant, orientation-aware)."""
    bad = 0
    p = pts.astype(np.int64)
    for a, b, c in tri:
        A, B, C = p[a], p[b], p[c]
        orient = (B[0] - A[0]) * (C[1] - A[1]) - (B[1] - A[1]) * (C[0] - A[0])
        ad, bd, cd = A - p, B - p, C - p
        det = ((ad[:, 0] ** 2 + ad[:, 1] ** 2) * (bd[:, 0] * cd[:, 1] - bd[:, 1] * cd[:, 0])
               - (bd[:, 0] ** 2 + bd[:, 1] ** 2) * (ad[:, 0] * cd[:, 1] - ad[:, 1] * cd[:, 0])
               + (cd[:, 0] ** 2 + cd[:, 1] ** 2) * (ad[:, 0] * bd[:, 1] - ad[:, 1] * bd[:, 0]))
        bad += int(np.sum(det * np.sign(orient) > 0))
    return bad


def _hull_count(pts):
    """Points on the convex hull boundary, collinear ones included
    (monotone chain keeping collinear points)."""
    P = sorted(map(tuple, pts.tolist()))
    if len(P) < 3:
        return len(P)

    def cross(o, a, b):
        return (a[0] - o[0]) * (b[1] - o[1]) - (a[1] - o[1]) * (b[0] - o[0])

    lower, upper = [], []
    for q in P:
        while len(lower) >= 2 and cross(lower[-2], lower[-1], q) < 0:
            lower.pop()
        lower.append(q)
    for q in reversed(P):
        while len(upper) >= 2 and cross(upper[-2], upper[-1], q) < 0:
            upper.pop()
        upper.append(q)
    return len(set(lower[:-1] + upper[:-1]))


@pytest.mark.parametrize("n,seed", [(3, 0), (4, 1), (10, 2), (200, 3), (1500, 4)])
def test_delaunay_empty_circumcircle_and_euler_count(n, seed):
    pts = _random_points(n, seed)
    tri = delaunay(pts)
    a, b, c = pts[tri[:, 0]], pts[tri[:, 1]], pts[tri[:, 2]]
    orient = (b[:, 0] - a[:, 0]) * (c[:, 1] - a[:, 1]) - (b[:, 1] - a[:, 1]) * (c[:, 0] - a[:, 0])
    assert np.all(orient > 0)
    assert _incircle_violations(pts, tri) == 0
    assert len(tri) == 2 * n - 2 - _hull_count(pts)
    assert set(np.unique(tri).tolist()) == set(range(n))


@pytest.mark.parametrize("seed", [5, 6])
def test_delaunay_matches_qhull_in_general_position(seed):
    spatial = pytest.importorskip("scipy.spatial")
    pts = _random_points(400, seed)
    ours = {tuple(sorted(t)) for t in delaunay(pts).tolist()}
    ref = {tuple(sorted(t)) for t in spatial.Delaunay(pts.astype(np.float64)).simplices.tolist()}
    assert ours == ref


def test_delaunay_degenerate_inputs():
    # collinear points: no triangle
    line = np.stack([np.arange(0, 50, 5), np.arange(0, 100, 10)], 1)
    assert len(delaunay(line)) == 0
    assert len(delaunay(np.zeros((0, 2), np.int32))) == 0
    # a square grid (every cell co-circular): any valid Delaunay triangulation
    gx, gy = np.meshgrid(np.arange(0, 50, 5), np.arange(0, 40, 5))
    grid = np.stack([gx.ravel(), gy.ravel()], 1).astype(np.int64)
    tri = delaunay(grid)
    assert _incircle_violations(grid, tri) == 0
    assert len(tri) == 2 * len(grid) - 2 - _hull_count(grid)
    with pytest.raises(ValueError):
        delaunay(np.array([[-1, 3], [4, 5], [6, 0]]))


def cell_points(ncx, ncy, seed, keep=0.8, mode="random", holes=()):
    """Points on a grid of 5x5 cells, at most one per cell, as the support
    points are: [ncx, ncy, 2] with (-1, -1) for an empty cell. mode "random":
    any pixel of the cell; "corner": the cell's corner (a square lattice,
    every cell's four points cocircular); "two": one of the cell's two
    diagonal corners. holes: (cx, cy, r) discs of empty cells."""
    rng = np.random.default_rng(seed)
    cx, cy = np.meshgrid(np.arange(ncx), np.arange(ncy), indexing="ij")
    if mode == "random":
        ox, oy = rng.integers(0, 5, (ncx, ncy)), rng.integers(0, 5, (ncx, ncy))
    elif mode == "corner":
        ox = oy = np.zeros((ncx, ncy), np.int64)
    else:
        ox = oy = 4 * rng.integers(0, 2, (ncx, ncy))
    on = rng.random((ncx, ncy)) < keep
    for hx, hy, r in holes:
        on &= (cx - hx) ** 2 + (cy - hy) ** 2 >= r * r
    c = np.full((ncx, ncy, 2), -1, np.int32)
    c[..., 0] = np.where(on, 5 * cx + ox, -1)
    c[..., 1] = np.where(on, 5 * cy + oy, -1)
    return c


@pytest.mark.parametrize("ncx,ncy,seed,keep,mode,holes", [
    (1, 1, 0, 1.0, "random", ()),
    (1, 9, 1, 1.0, "random", ()),
    (2, 2, 2, 1.0, "corner", ()),
    (12, 9, 3, 1.0, "corner", ()),
    (12, 9, 4, 0.6, "corner", ()),
    (20, 15, 5, 0.9, "two", ()),
    (20, 15, 6, 0.15, "random", ()),
    (40, 30, 7, 0.8, "random", ((20, 15, 8),)),
    (40, 30, 8, 0.5, "two", ((5, 5, 6), (30, 25, 9))),
    (320, 240, 9, 0.7, "random", ((100, 80, 40), (250, 200, 25), (30, 200, 60))),
    # grid sides not multiples of the 8-cell occupancy blocks, large scans
    (83, 61, 10, 0.3, "two", ((40, 30, 20), (0, 60, 15))),
])
def test_star_construction_equals_incremental(ncx, ncy, seed, keep, mode, holes):
    """The device construction's code (each point's star, run on the host)
    against the incremental construction: both triangulate the perturbed
    points, so the triangles and their order agree exactly, on lattices where
    most quadruples are cocircular, on sparse sets and around large empty
    regions (circumcircles far larger than the search square)."""
    c = cell_points(ncx, ncy, seed, keep, mode, holes)
    xy = c.reshape(-1, 2)
    pts = xy[xy[:, 0] >= 0].astype(np.int64)
    star = delaunay_cells(c)
    inc = delaunay(pts) if len(pts) else np.zeros((0, 3), np.int32)
    assert star.shape == inc.shape and np.array_equal(star, inc)
    if len(star):
        # listed by the lowest-index vertex, then by the next vertex
        assert np.all(star[:, 0] < star[:, 1]) and np.all(star[:, 0] < star[:, 2])
        key = star[:, 0].astype(np.int64) * len(pts) + star[:, 1]
        assert np.all(np.diff(key) > 0)
    if 3 <= len(pts) <= 4000:
        assert _incircle_violations(pts, star) == 0
        hull = _hull_count(pts)
        if len(star):
            assert len(star) == 2 * len(pts) - 2 - hull


def test_star_construction_rejects_points_outside_their_cells():
    c = cell_points(4, 4, 0)
    c[1, 2] = (0, 0)  # belongs to cell (0, 0)
    with pytest.raises(ValueError):
        delaunay_cells(c)


def _camera(h, w, f=200.0):
    c = N.CCamera()
    K = [f, 0, w / 2.0, 0, f, h / 2.0, 0, 0, 1]
    R = [1, 0, 0, 0, 1, 0, 0, 0, 1]
    for i in range(9):
        c.K[i] = K[i]
        c.R[i] = R[i]
    c.height, c.width, c.depth_min, c.depth_max = h, w, 1.0, 10.0
    return c


def _plane_depth(h, w, f, n, d):
    """Depth of the plane n.X + d = 0 seen by a pinhole at the origin."""
    ys, xs = np.mgrid[0:h, 0:w].astype(np.float64)
    ray = np.stack([(xs - w / 2.0) / f, (ys - h / 2.0) / f, np.ones_like(xs)], -1)
    return (-d / (ray @ n)).astype(np.float32)


@pytest.mark.parametrize("normal", [(0.0, 0.0, -1.0), (0.3, -0.2, -0.93)])
def test_prior_planes_of_an_exact_plane(normal):
    h, w, f = 120, 160, 200.0
    n = np.array(normal) / np.linalg.norm(normal)
    d = 4.0
    depth = _plane_depth(h, w, f, n, d)
    rng = np.random.default_rng(7)
    cost = (0.05 + 0.01 * rng.random((h, w))).astype(np.float32)
    planes, masks, ntri = build_prior(_camera(h, w, f), depth, cost, None, 0.5, 20.0)
    assert ntri > 100
    m = masks > 0
    assert m.mean() > 0.8  # the support hull covers most of the image
    assert masks.max() <= ntri
    # every masked pixel carries the generating plane (sign fixed by d >= 0)
    assert np.allclose(planes[m][:, :3], n[None, :], atol=2e-4)
    assert np.allclose(planes[m][:, 3], d, rtol=2e-4)
    assert np.all(planes[~m] == 0)


def test_prior_support_points_need_low_cost():
    h, w, f = 60, 80, 100.0
    depth = _plane_depth(h, w, f, np.array([0.0, 0.0, -1.0]), 3.0)
    cost = np.full((h, w), 0.5, np.float32)  # every cell's best cost >= 0.1
    planes, masks, ntri = build_prior(_camera(h, w, f), depth, cost, None, 0.5, 20.0)
    assert ntri == 0 and not masks.any() and not planes.any()


def test_prior_depth_range_masks_out_of_range_pixels():
    h, w, f = 120, 160, 200.0
    n = np.array([0.5, 0.0, -0.866])
    depth = _plane_depth(h, w, f, n, 4.0)
    cost = np.full((h, w), 0.05, np.float32)
    lo, hi = np.percentile(depth, 30), np.percentile(depth, 70)
    _, masks, _ = build_prior(_camera(h, w, f), depth, cost, None, float(lo), float(hi))
    m = masks > 0
    assert m.any()
    assert depth[m].min() >= lo * 0.999 and depth[m].max() <= hi * 1.001


def test_superpixel_filter_drops_triangles_across_labels():
    h, w, f = 120, 160, 200.0
    depth = _plane_depth(h, w, f, np.array([0.0, 0.0, -1.0]), 4.0)
    cost = np.full((h, w), 0.05, np.float32)
    labels = np.where(np.arange(w)[None, :] < 77, 1, 2).astype(np.int32).repeat(h, 0)
    _, m_off, _ = build_prior(_camera(h, w, f), depth, cost, labels, 0.5, 20.0, False)
    _, m_on, _ = build_prior(_camera(h, w, f), depth, cost, labels, 0.5, 20.0, True)

    def straddling(masks):
        ids = np.unique(masks[masks > 0])
        return sum(len(np.unique(labels[masks == i])) > 1 for i in ids)

    assert straddling(m_off) > 0  # without the filter some triangles cross
    assert straddling(m_on) == 0
    assert (m_on > 0).mean() > 0.5


def test_superpixel_filter_drops_unlabelled_triangles():
    h, w, f = 120, 160, 200.0
    depth = _plane_depth(h, w, f, np.array([0.0, 0.0, -1.0]), 4.0)
    cost = np.full((h, w), 0.05, np.float32)
    labels = np.ones((h, w), np.int32)
    labels[:, : w // 2] = 0  # left half unlabelled
    _, masks, _ = build_prior(_camera(h, w, f), depth, cost, labels, 0.5, 20.0, True)
    ids = np.unique(masks[masks > 0])
    assert len(ids) > 0
    for i in ids:  # no retained triangle lies wholly in the unlabelled half
        assert labels[masks == i].max() == 1
    # labels above 255 are kept apart (labels are int32, not uint8)
    labels2 = np.where(np.arange(w)[None, :] < 77, 1, 257).astype(np.int32).repeat(h, 0)
    _, m2, _ = build_prior(_camera(h, w, f), depth, cost, labels2, 0.5, 20.0, True)
    for i in np.unique(m2[m2 > 0]):
        assert len(np.unique(labels2[m2 == i])) == 1
