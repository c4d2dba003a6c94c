"""Post-fusion kernels against their oracles, through the C ABI:
sliding_window_normal / filter_by_var_map / fusion_planar_mask bit-exact
against the sequential twins (gmvs/lib/src/planar_initer.cu:471-684); the
point gather against the numpy restatement of
mvs_fusion_segmentaion_cpu.py:158-205 (FP64, to 1e-12: the reference's
`dirs_ @ rot_mat.T` goes through BLAS, whose summation order is not pinned);
the .xyz text byte-for-byte against np.savetxt, the reference's own call."""
import io

import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _post():
    from gmvs_amd import post

    return post


def _segments(h, w, n, seed):
    g = np.random.default_rng(seed)
    cy, cx = g.uniform(0, h, n), g.uniform(0, w, n)
    yy, xx = np.mgrid[0:h, 0:w]
    best = np.zeros((h, w), np.int32)
    bd = np.full((h, w), np.inf)
    for i in range(n):  # Voronoi cells, without an [h, w, n] temporary
        d = (yy - cy[i]) ** 2 + (xx - cx[i]) ** 2
        m = d < bd
        best[m], bd[m] = i, d[m]
    return best


@pytest.mark.parametrize("shape,ws", [((23, 31), 1), ((97, 131), 5), ((97, 131), 4), ((64, 48), 31)])
def test_sliding_window_normal_bit_exact(shape, ws):
    P = _post()
    g = np.random.default_rng(ws)
    planar = (g.random(shape) < 0.3).astype(np.int32)
    planar[g.random(shape) < 0.05] = 2
    normal = g.standard_normal(shape + (3,)).astype(np.float32)
    got = P.sliding_window_normal(planar, normal, ws)
    ref = O.sliding_window_normal(planar, normal, ws)
    assert got.dtype == np.float32 and got.shape == ref.shape
    np.testing.assert_array_equal(got.view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("shape,n", [((40, 50), 30), ((480, 640), 900)])
def test_filter_by_var_map_bit_exact(shape, n):
    P = _post()
    seg = _segments(*shape, n, 1)
    seg[seg == 7] = 0
    seg[seg == 11] = n + 5  # a gap in the ids
    g = np.random.default_rng(2)
    var = g.random(shape).astype(np.float32)
    for thresh in (0.3, 0.5):
        got = P.filter_by_var_map(var, seg, thresh)
        ref = O.filter_by_var_map(var, seg, thresh)
        np.testing.assert_array_equal(got, ref)
    # device tensors in, device tensors out
    got_t = P.filter_by_var_map(torch.from_numpy(var).cuda(), torch.from_numpy(seg).cuda(), 0.5)
    assert got_t.is_cuda
    np.testing.assert_array_equal(got_t.cpu().numpy(), O.filter_by_var_map(var, seg, 0.5))


@pytest.mark.parametrize("shape,n,seed", [((48, 64), 40, 3), ((480, 640), 1200, 5)])
def test_fusion_planar_mask_bit_exact(shape, n, seed):
    P = _post()
    seg = _segments(*shape, n, seed)
    g = np.random.default_rng(seed + 1)
    ns = seg.max() + 1
    frac = g.choice([0.0, 0.1, 0.5, 0.9], ns)
    planar = (g.random(shape) < frac[seg]).astype(np.int32)
    base = g.standard_normal((ns, 3))
    base /= np.linalg.norm(base, axis=1, keepdims=True)
    spread = g.choice([0.05, 2.0], ns)
    nrm = base[seg] + spread[seg][..., None] * g.standard_normal(shape + (3,))
    nrm = (nrm / np.linalg.norm(nrm, axis=-1, keepdims=True)).astype(np.float32)
    thresh = 30 if n < 100 else 100
    got = P.fusion_planar_mask(seg, planar, nrm, shape[0], shape[1], thresh, 0.75, 0.8, 0.9)
    ref = O.fusion_planar_mask(seg, planar, nrm, thresh, 0.75, 0.8, 0.9)
    np.testing.assert_array_equal(got, ref)
    assert (got == 300).any() and (got > 300).any() and ((got > 0) & (got < 300)).any()


def test_segment_utilities_reject_negative_ids():
    P = _post()
    seg = np.zeros((4, 4), np.int32)
    seg[0, 0] = -1
    with pytest.raises(ValueError):
        P.filter_by_var_map(np.zeros((4, 4), np.float32), seg, 0.1)


def _pose(yaw_deg, C):
    a = np.radians(yaw_deg)
    R = np.array([[np.cos(a), 0, np.sin(a)], [0, 1, 0], [-np.sin(a), 0, np.cos(a)]])
    E = np.eye(4)
    E[:3, :3], E[:3, 3] = R, C
    return E


def _maps(nv, h, w, seed):
    g = np.random.default_rng(seed)
    dn = np.zeros((nv, h, w, 4), np.float32)
    dn[..., 0] = g.uniform(1.0, 5.0, (nv, h, w))
    dn[g.random((nv, h, w)) < 0.35, 0] = 0
    n = g.standard_normal((nv, h, w, 3))
    n[..., 2] = -np.abs(n[..., 2]) - 0.3  # mostly facing the cameras
    dn[..., 1:] = n / np.linalg.norm(n, axis=-1, keepdims=True)
    return dn


def test_gather_points_matches_restatement():
    P = _post()
    nv, h, w = 3, 120, 160
    K = np.array([[150.0, 0, 79.5], [0, 151.0, 59.5], [0, 0, 1]])
    dn = _maps(nv, h, w, 7)
    poses = [_pose(-15 + 15 * i, [0.2 * i, -0.1, 3.0 - 0.1 * i]) for i in range(nv)]
    pts, ptsf, maps = P.gather_points(list(dn), poses, K)
    rpts, rptsf, rmaps = O.gather_points(list(dn), poses, K, h, w)
    assert pts.shape == rpts.shape and ptsf.shape == rptsf.shape
    np.testing.assert_allclose(pts.cpu().numpy(), rpts, rtol=0, atol=1e-12)
    np.testing.assert_allclose(ptsf.cpu().numpy(), rptsf, rtol=0, atol=1e-12)
    np.testing.assert_array_equal(maps.cpu().numpy(), rmaps)
    assert 0 < len(rptsf) < len(rpts)


def test_gather_points_empty_and_full():
    P = _post()
    K = np.array([[50.0, 0, 10.0], [0, 50.0, 8.0], [0, 0, 1]])
    dn = np.zeros((2, 17, 21, 4), np.float32)
    pts, ptsf, maps = P.gather_points(list(dn), [np.eye(4)] * 2, K)
    assert pts.shape == (0, 6) and ptsf.shape == (0, 6) and not maps.cpu().numpy().any()
    dn[..., 0] = 2.0
    dn[..., 3] = -1.0
    pts, ptsf, maps = P.gather_points(list(dn), [np.eye(4)] * 2, K)
    assert pts.shape == (2 * 17 * 21, 6) and ptsf.shape == pts.shape
    np.testing.assert_array_equal(maps.cpu().numpy(), dn)


def _savetxt(rows, dec):
    b = io.BytesIO()
    np.savetxt(b, rows, fmt=f"%.{dec}f")
    return b.getvalue()


def test_format_rows_matches_savetxt():
    P = _post()
    g = np.random.default_rng(11)
    rows = g.standard_normal((5000, 6)) * np.array([1, 10, 100, 1e4, 1e-3, 1e8])
    special = np.array([-0.0, 0.0, 0.03125, 0.00005, -0.00005, 2.5, 0.5, 1.5, -2.5, 5e-324,
                        1e-310, 123456789.00005, 9.99995, -9.99995, 0.99995, 1e14, -7.25e13,
                        float("nan"), float("inf"), -float("inf")], np.float64)
    rows[: len(special), 0] = special
    rows[: len(special), 5] = special[::-1]
    for dec in (4, 0, 1, 9):
        r = rows if dec < 9 else rows[:, 1:5] / 1e4  # keep |v| * 10^9 < 2^63
        got = P.format_rows(r, dec)
        assert got == _savetxt(r, dec), dec


def test_format_rows_rejects_huge_values(tmp_path):
    P = _post()
    with pytest.raises(ValueError):
        P.format_rows(np.array([[1e300, 1.0]]), 4)
    p = tmp_path / "pts.xyz"
    n = P.savetxt(str(p), np.array([[1.0, -2.0, 3.25]]))
    assert p.read_bytes() == b"1.0000 -2.0000 3.2500\n" and n == 22


def test_gather_script_end_to_end(tmp_path):
    """gmvs_amd.scripts.gather_points on the reference's layout: pose/*.txt,
    intrinsic.txt, depth_normal/*.npy in; depth_normal_filter/*.npy and the
    two .xyz files out."""
    from gmvs_amd.scripts import gather_points

    nv, h, w = 3, 48, 64
    K = np.array([[60.0, 0, 31.5, 0], [0, 60.0, 23.5, 0], [0, 0, 1, 0], [0, 0, 0, 1]])
    data, dnd, save = tmp_path / "data", tmp_path / "dn", tmp_path / "out"
    (data / "pose").mkdir(parents=True)
    dnd.mkdir()
    dn = _maps(nv, h, w, 13)
    poses = [_pose(10 * i, [0.1 * i, 0.0, 2.5]) for i in range(nv)]
    np.savetxt(data / "intrinsic.txt", K)
    for i in range(nv):
        np.savetxt(data / "pose" / f"{i * 10}.txt", poses[i])  # numeric, not zero-padded, names
        np.save(dnd / f"{i * 10}.npy", dn[i])
    rc = gather_points.main(["--depth_normal_dir", str(dnd), "--data_dir", str(data), "--save_dir",
                             str(save), "--height", str(h), "--width", str(w)])
    assert rc == 0
    poses_rt = [np.loadtxt(data / "pose" / f"{i * 10}.txt") for i in range(nv)]
    rpts, rptsf, rmaps = O.gather_points(list(dn), poses_rt, np.loadtxt(data / "intrinsic.txt"), h, w)
    for i in range(nv):
        np.testing.assert_array_equal(np.load(save / "depth_normal_filter" / f"{i * 10}.npy"), rmaps[i])
    for name, ref in (("mvs_depthnorm.xyz", rpts), ("mvs_depthnorm_filter.xyz", rptsf)):
        got = np.loadtxt(save / name)
        assert got.shape == ref.shape
        np.testing.assert_allclose(got, ref, rtol=0, atol=5.1e-5)  # %.4f rounding
