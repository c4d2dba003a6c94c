"""The post-fusion oracles pinned on the CPU: the sequential twins of
sliding_window_normal / filter_by_var_map / fusion_planar_mask
(gmvs/lib/src/planar_initer.cu:471-684) against FP64 numpy computations of
the same definitions and hand-built known answers, and the numpy restatement
of the point gather (mvs_fusion_segmentaion_cpu.py:158-205) against closed
forms."""
import numpy as np
import pytest

from oracle import oracle as O


def _swn_fp64(planar, normal, ws):
    h, w = planar.shape
    half = ws // 2
    out = np.zeros((h, w, 3))
    for r in range(h):
        for c in range(w):
            if planar[r, c] == 1:
                continue
            r0, r1, c0, c1 = max(r - half, 0), min(r + half + 1, h), max(c - half, 0), min(c + half + 1, w)
            keep = planar[r0:r1, c0:c1] != 1
            out[r, c] = normal[r0:r1, c0:c1][keep].astype(np.float64).mean(0)
    return out


@pytest.mark.parametrize("ws", [1, 3, 4, 7])
def test_sliding_window_normal_twin(ws):
    g = np.random.default_rng(ws)
    planar = (g.random((23, 31)) < 0.3).astype(np.int32)
    planar[g.random((23, 31)) < 0.05] = 2  # labels other than 1 are not planar
    normal = g.standard_normal((23, 31, 3)).astype(np.float32)
    got = O.sliding_window_normal(planar, normal, ws)
    ref = _swn_fp64(planar, normal, ws)
    assert np.all(got[planar == 1] == 0)
    np.testing.assert_allclose(got, ref, rtol=0, atol=2e-6)
    if ws == 1:
        np.testing.assert_array_equal(got[planar != 1], normal[planar != 1])


def test_sliding_window_normal_constant_field():
    planar = np.zeros((9, 9), np.int32)
    planar[4, :] = 1
    normal = np.broadcast_to(np.float32([0.6, -0.8, 0.0]), (9, 9, 3)).copy()
    got = O.sliding_window_normal(planar, normal, 5)
    np.testing.assert_allclose(got[planar == 0], normal[planar == 0], atol=1e-7)


def _segments(h, w, n, seed):
    g = np.random.default_rng(seed)
    cy, cx = g.uniform(0, h, n), g.uniform(0, w, n)
    yy, xx = np.mgrid[0:h, 0:w]
    d = (yy[..., None] - cy) ** 2 + (xx[..., None] - cx) ** 2
    return d.argmin(-1).astype(np.int32)


def test_filter_by_var_map_known_answer():
    seg = np.array([[0, 1, 1, 2], [3, 3, 2, 2], [5, 5, 5, 0]], np.int32)  # id 4 absent
    var = np.array([[9, 0.1, 0.3, 0.5], [0.2, 0.2, 0.5, 0.5], [1.0, 2.0, 3.0, 9]], np.float32)
    got = O.filter_by_var_map(var, seg, 0.25)
    # means: 1 -> 0.2 kept, 2 -> 0.5 dropped, 3 -> 0.2 kept, 5 -> 2.0 dropped; 0 never kept
    np.testing.assert_array_equal(got, np.array([[0, 1, 1, 0], [3, 3, 0, 0], [0, 0, 0, 0]]))


def test_filter_by_var_map_twin_vs_fp64():
    seg = _segments(40, 50, 30, 1)
    seg[seg == 7] = 0
    g = np.random.default_rng(2)
    var = g.random((40, 50)).astype(np.float32)
    thresh = 0.5
    got = O.filter_by_var_map(var, seg, thresh)
    n = seg.max() + 1
    mean = np.bincount(seg.ravel(), var.ravel().astype(np.float64), n) / np.maximum(
        np.bincount(seg.ravel(), minlength=n), 1)
    clear = np.abs(mean - thresh) > 1e-5
    ref = np.where((mean[seg] <= thresh) & (seg > 0), seg, 0)
    sel = clear[seg]
    np.testing.assert_array_equal(got[sel], ref[sel])


def _fpm_fp64(seg, planar, normals, thresh, npp, cst, mrt):
    out = np.zeros_like(seg)
    vps = 301
    for s in range(1, seg.max() + 1):
        m = seg == s
        nonp = m & (planar == 0)
        n_np, amount = nonp.sum(), m.sum()
        if amount <= thresh:
            continue
        if n_np / amount >= npp:
            nv = normals[nonp].astype(np.float64)
            avg = nv.mean(0)
            ratio = ((nv @ avg) > cst).sum() / n_np
            out[m] = s if ratio > mrt else 300
        elif n_np / amount < 0.25:
            out[m] = vps
            vps += 1
        else:
            out[m] = s
    return out


def test_fusion_planar_mask_known_answer():
    # one segment per branch in an 18x10 map (ids 7 absent)
    seg = np.zeros((18, 10), np.int32)
    rows = {1: (0, 2), 2: (2, 5), 3: (5, 8), 4: (8, 11), 5: (11, 12), 6: (12, 15), 8: (15, 18)}
    for s, (r0, r1) in rows.items():
        seg[r0:r1] = s
    planar = np.zeros((18, 10), np.int32)
    planar[8:11] = 1            # segment 4: all planar -> textureless
    planar[15:18] = 1           # segment 8: all planar -> textureless
    planar[12:15, :5] = 1       # segment 6: half planar -> kept
    normals = np.zeros((18, 10, 3), np.float32)
    normals[..., 2] = 1.0
    g = np.random.default_rng(0)
    sc = g.standard_normal((3, 10, 3)).astype(np.float32)
    normals[5:8] = sc / np.linalg.norm(sc, axis=-1, keepdims=True)  # segment 3: scattered -> object
    exp = np.zeros((18, 10), np.int32)
    exp[2:5], exp[5:8], exp[8:11], exp[12:15], exp[15:18] = 2, 300, 301, 6, 302
    np.testing.assert_array_equal(O.fusion_planar_mask(seg, planar, normals, 25, 0.75, 0.8, 0.9), exp)
    exp[0:2] = 1  # 20 pixels > 19: kept
    np.testing.assert_array_equal(O.fusion_planar_mask(seg, planar, normals, 19, 0.75, 0.8, 0.9), exp)


def test_fusion_planar_mask_twin_vs_fp64():
    seg = _segments(48, 64, 40, 3)
    g = np.random.default_rng(4)
    # per-segment planar fraction and normal spread, so every branch occurs
    frac = g.choice([0.0, 0.1, 0.5, 0.9], seg.max() + 1)
    planar = (g.random(seg.shape) < frac[seg]).astype(np.int32)
    base = g.standard_normal((seg.max() + 1, 3))
    base /= np.linalg.norm(base, axis=1, keepdims=True)
    spread = g.choice([0.05, 2.0], seg.max() + 1)
    nrm = base[seg] + spread[seg][..., None] * g.standard_normal(seg.shape + (3,))
    nrm = (nrm / np.linalg.norm(nrm, axis=-1, keepdims=True)).astype(np.float32)
    got = O.fusion_planar_mask(seg, planar, nrm, 30, 0.75, 0.8, 0.9)
    ref = _fpm_fp64(seg, planar, nrm, 30, 0.75, 0.8, 0.9)
    np.testing.assert_array_equal(got, ref)
    assert {0, 300}.issubset(set(np.unique(got))) and (got > 300).any()


def _pose(yaw_deg, C):
    a = np.radians(yaw_deg)
    R = np.array([[np.cos(a), 0, np.sin(a)], [0, 1, 0], [-np.sin(a), 0, np.cos(a)]])
    E = np.eye(4)
    E[:3, :3], E[:3, 3] = R, C
    return E


def test_gather_points_closed_forms():
    h, w = 6, 8
    K = np.array([[100.0, 0, 3.5], [0, 90.0, 2.5], [0, 0, 1]])
    E = _pose(20, [0.3, -0.1, 2.0])
    R = E[:3, :3]
    g = np.random.default_rng(5)
    dn = np.zeros((h, w, 4), np.float32)
    dn[..., 0] = g.uniform(1, 3, (h, w))
    dn[g.random((h, w)) < 0.3, 0] = 0
    # rows 0-2: the world normal -R e_z (facing back along the optical axis):
    #   -dir . n = d_c . e_z = 1 > cos 70 deg, every point passes;
    # rows 3-5: R e_x (perpendicular to it): -dir . n = -(col - cx) / fx lies
    #   in [-0.035, 0.035], below cos 70 deg = 0.342, so no point passes
    dn[:3, :, 1:] = (-R[:, 2]).astype(np.float32)
    dn[3:, :, 1:] = R[:, 0].astype(np.float32)
    pts, ptsf, maps = O.gather_points([dn], [E], K, h, w)
    rows, cols = np.nonzero(dn[..., 0] > 0)
    assert len(pts) == len(rows)
    # back into the camera: X_c = R^T (X - C) = depth * ((c - cx)/fx, (r - cy)/fy, 1)
    Xc = (pts[:, :3] - E[:3, 3]) @ R
    d = dn[rows, cols, 0].astype(np.float64)
    np.testing.assert_allclose(Xc[:, 2], d, rtol=1e-12)
    np.testing.assert_allclose(Xc[:, 0], d * (cols - 3.5) / 100.0, atol=1e-12)
    np.testing.assert_allclose(Xc[:, 1], d * (rows - 2.5) / 90.0, atol=1e-12)
    np.testing.assert_array_equal(pts[:, 3:], dn[rows, cols, 1:])
    keep = rows < 3
    assert len(ptsf) == keep.sum()
    np.testing.assert_array_equal(ptsf, pts[keep])
    np.testing.assert_array_equal(maps[0][rows[keep], cols[keep]], dn[rows[keep], cols[keep]])
    assert np.all(maps[0][3:] == 0) and np.all(maps[0][dn[..., 0] <= 0] == 0)
