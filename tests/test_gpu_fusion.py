"""GPU fusion vs the sequential twin: bit-exact masks, outputs, counts (the
reference's FP32 check, and the FP64 check of gmvs_fuse_view_fp64 against the
twin's FP64 instantiation)."""
import math

import numpy as np
import pytest
import torch

from oracle import oracle as O
from tests import helpers as T

pytestmark = pytest.mark.gpu


def noisy_maps(sc, rng, noise=0.01, outliers=0.05):
    n = sc.depth.shape[0]
    ds, ns = [], []
    for i in range(n):
        d = sc.depth[i].numpy().copy()
        d *= (1 + noise * rng.standard_normal(d.shape)).astype(np.float32)
        out = rng.random(d.shape) < outliers
        d[out] = rng.uniform(1.5, 7.0, out.sum()).astype(np.float32)
        nn = sc.normal[i].numpy().copy()
        ds.append(d.astype(np.float32))
        ns.append(nn.astype(np.float32))
    return ds, ns


@pytest.mark.parametrize("precision", [32, 64])
@pytest.mark.parametrize("thresholds", [(2.0, 0.01, 0.174533, 3), (1.0, 0.01, math.radians(30), 2)])
def test_fusion_bit_exact_against_twin(thresholds, precision):
    from gmvs_amd import src as S

    reproj, rel, ang, gc = thresholds
    sc = T.scene()
    nviews = sc.depth.shape[0]
    rng = np.random.default_rng(7)
    ds, ns = noisy_maps(sc, rng)
    cams_c = [S.load_camera(c.K.reshape(-1), c.R.reshape(-1), c.t, c.height, c.width,
                            c.depth_min, c.depth_max)._c for c in sc.cameras]
    ocams = T.oracle_cams(sc, range(nviews))
    dev = torch.device("cuda")
    dd = [torch.from_numpy(d).to(dev) for d in ds]
    nd = [torch.from_numpy(n).to(dev) for n in ns]
    h, w = ds[0].shape
    masks_g = torch.zeros(nviews, h, w, dtype=torch.bool, device=dev)
    masks_o = np.zeros((nviews, h, w), np.uint8)
    cos_t = float(np.float32(math.cos(ang)))
    total_g = total_o = 0
    for ref in range(nviews):
        ids = [ref] + sc.pairs[ref]
        od = torch.empty(h, w, device=dev)
        on = torch.empty(h, w, 3, device=dev)
        S.fuse_view(ids, [cams_c[i] for i in ids], [dd[i] for i in ids], [nd[i] for i in ids],
                    masks_g, gc, od, on, reproj, rel, ang, precision=precision)
        torch.cuda.synchronize()
        od_o, on_o, _ = O.fuse_view([ocams[i] for i in ids], [ds[i] for i in ids],
                                    [ns[i] for i in ids], masks_o, ref, ids[1:], gc, reproj, rel,
                                    cos_t, precision=precision)
        gd, gn = od.cpu().numpy(), on.cpu().numpy()
        assert np.array_equal(gd.view(np.uint32), od_o.view(np.uint32)), f"depth of view {ref}"
        assert np.array_equal(gn.view(np.uint32), on_o.view(np.uint32)), f"normal of view {ref}"
        assert np.array_equal(masks_g.cpu().numpy().astype(np.uint8), masks_o), f"masks after {ref}"
        total_g += int((gd > 0).sum())
        total_o += int((od_o > 0).sum())
    print(f"fused points: GPU {total_g}  twin {total_o}")
    assert total_g == total_o and total_g > 0


def test_fusion_bit_exact_full_size():
    """The fusion configuration's view size and thresholds (1600x1200, >= 3
    views, 1 px, 0.01, 30 deg), bit-exact against the twin on every view; 8
    views with 7 sources each keep the sequential CPU twin within seconds."""
    from gmvs_amd import src as S
    from gmvs_amd.synthetic import make_scene

    dev = torch.device("cuda")
    sc = make_scene(1600, 1200, 8, 64, seed=3407 + 5, fx=1500.0, arc_deg=30.0, device=dev)
    depth, normal = sc.depth.cpu(), sc.normal.cpu()
    rng = np.random.default_rng(11)
    ds, ns = [], []
    for i in range(8):
        d = depth[i].numpy().copy()
        d *= (1 + 0.01 * rng.standard_normal(d.shape)).astype(np.float32)
        out = rng.random(d.shape) < 0.05
        d[out] = rng.uniform(1.5, 7.0, out.sum()).astype(np.float32)
        ds.append(d.astype(np.float32))
        ns.append(normal[i].numpy().astype(np.float32).copy())
    cams_c = [S.load_camera(c.K.reshape(-1), c.R.reshape(-1), c.t, c.height, c.width,
                            c.depth_min, c.depth_max)._c for c in sc.cameras]
    ocams = T.oracle_cams(sc, range(8))
    dd = [torch.from_numpy(d).to(dev) for d in ds]
    nd = [torch.from_numpy(n).to(dev) for n in ns]
    h, w = ds[0].shape
    masks_g = torch.zeros(8, h, w, dtype=torch.bool, device=dev)
    masks_o = np.zeros((8, h, w), np.uint8)
    ang = math.radians(30.0)
    cos_t = float(np.float32(math.cos(ang)))
    total = 0
    for ref in range(8):
        ids = [ref] + list(sc.pairs[ref][:7])
        od = torch.empty(h, w, device=dev)
        on = torch.empty(h, w, 3, device=dev)
        S.fuse_view(ids, [cams_c[i] for i in ids], [dd[i] for i in ids], [nd[i] for i in ids],
                    masks_g, 3, od, on, 1.0, 0.01, ang)
        torch.cuda.synchronize()
        od_o, on_o, _ = O.fuse_view([ocams[i] for i in ids], [ds[i] for i in ids],
                                    [ns[i] for i in ids], masks_o, ref, ids[1:], 3, 1.0, 0.01,
                                    cos_t)
        gd, gn = od.cpu().numpy(), on.cpu().numpy()
        assert np.array_equal(gd.view(np.uint32), od_o.view(np.uint32)), f"depth of view {ref}"
        assert np.array_equal(gn.view(np.uint32), on_o.view(np.uint32)), f"normal of view {ref}"
        assert np.array_equal(masks_g.cpu().numpy().astype(np.uint8), masks_o), f"masks after {ref}"
        total += int((gd > 0).sum())
    print(f"fused points (8 views 1600x1200): {total}")
    assert total > 0


def test_fusion_bit_exact_fusion_configuration():
    """The fusion workload the benchmark measures -- 49 views of 1600x1200,
    20 sources per view, 1 px / 0.01 / 30 deg, >= 3 consistent views, 1 %
    depth noise and 5 % outliers -- fused view by view on the GPU and by the
    twin: every view's fused depth and normal maps and the masks after it are
    bit-identical, and so is the fused point count."""
    import bench
    from gmvs_amd import src as S

    dev = torch.device("cuda")
    sc, d, n, pairs = bench.fusion_workload(dev)
    nv = len(sc.cameras)
    ds = [d[i].cpu().numpy() for i in range(nv)]
    ns = [n[i].cpu().numpy() for i in range(nv)]
    cams_c = [S.load_camera(c.K.reshape(-1), c.R.reshape(-1), c.t, c.height, c.width,
                            c.depth_min, c.depth_max)._c for c in sc.cameras]
    ocams = T.oracle_cams(sc, range(nv))
    h, w = ds[0].shape
    masks_g = torch.zeros(nv, h, w, dtype=torch.bool, device=dev)
    masks_o = np.zeros((nv, h, w), np.uint8)
    ang = math.radians(30.0)
    cos_t = float(np.float32(math.cos(ang)))
    total_g = total_o = 0
    for ref in range(nv):
        ids = [ref] + list(pairs[ref])
        od = torch.zeros(h, w, device=dev)
        on = torch.zeros(h, w, 3, device=dev)
        S.fuse_view(ids, [cams_c[i] for i in ids], [d[i] for i in ids], [n[i] for i in ids],
                    masks_g, 3, od, on, 1.0, 0.01, ang)
        od_o, on_o, _ = O.fuse_view([ocams[i] for i in ids], [ds[i] for i in ids],
                                    [ns[i] for i in ids], masks_o, ref, ids[1:], 3, 1.0, 0.01,
                                    cos_t)
        gd, gn = od.cpu().numpy(), on.cpu().numpy()
        assert np.array_equal(gd.view(np.uint32), od_o.view(np.uint32)), f"depth of view {ref}"
        assert np.array_equal(gn.view(np.uint32), on_o.view(np.uint32)), f"normal of view {ref}"
        total_g += int((gd > 0).sum())
        total_o += int((od_o > 0).sum())
    assert np.array_equal(masks_g.cpu().numpy().astype(np.uint8), masks_o)
    print(f"fused points (49 views 1600x1200, 20 sources): GPU {total_g}, twin {total_o}")
    assert total_g == total_o > 0


def _rig_maps(h, w, f, xs, n, d):
    """Ground-truth depth maps of the plane n.X + d = 0 seen by cameras with
    R = I at (x, 0, 0); world normals n (unit, facing the cameras)."""
    ys, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    rx, ry = (xx - w / 2.0) / f, (ys - h / 2.0) / f
    ds = []
    for x in xs:
        s = -(n[0] * x + d) / (n[0] * rx + n[1] * ry + n[2])
        ds.append(s.astype(np.float32))
    nm = np.broadcast_to(np.asarray(n, np.float32), (h, w, 3)).copy()
    return ds, [nm.copy() for _ in xs]


def test_fusion_fp64_check_far_from_the_origin():
    """The far-depth case of the FP64 check: exact maps of one slanted plane
    seen by 5 cameras (R = I, 0.25 apart), the whole rig and scene at the
    world origin and moved by T = (2^20, 2^19, 0) (camera translations
    -(C + T), exact in FP32; depths and normals unchanged). Every pixel that
    projects into enough sources is consistent in exact arithmetic. At the
    origin FP32 and FP64 fuse the same points (to 0.1 %: pixels at the
    rounding boundaries); far from it the FP64 check still fuses them, while
    FP32 world points, rounded to 2^-4 units at 2^20, miss the 1-pixel bar
    (measured: 20,689 / 20,681 at the origin, 6,768 / 20,680 moved, FP32 /
    FP64). GPU and twin agree bit for bit throughout."""
    from gmvs_amd import src as S

    h, w, f = 120, 160, 200.0
    xs = [0.25 * i for i in range(5)]
    n = np.array([0.2, -0.1, -1.0])
    n /= np.linalg.norm(n)
    d = -n[2] * 4.0  # through (0, 0, 4)
    ds, ns = _rig_maps(h, w, f, xs, n, d)
    dev = torch.device("cuda")
    dd = [torch.from_numpy(x).to(dev) for x in ds]
    nd = [torch.from_numpy(x).to(dev) for x in ns]
    K = [f, 0, w / 2.0, 0, f, h / 2.0, 0, 0, 1]
    R = [1.0, 0, 0, 0, 1, 0, 0, 0, 1]
    counts = {}
    for shift in (0.0, 1.0):
        T = np.array([2.0 ** 20, 2.0 ** 19, 0.0]) * shift
        ts = [-(np.array([x, 0.0, 0.0]) + T) for x in xs]
        assert all(np.array_equal(t.astype(np.float32).astype(np.float64), t) for t in ts)
        cams_c = [S.load_camera(K, R, t, h, w, 1.0, 10.0)._c for t in ts]
        ocams = [O.make_camera(np.array(K).reshape(3, 3), np.eye(3), t, h, w, 1.0, 10.0) for t in ts]
        for precision in (32, 64):
            masks_g = torch.zeros(5, h, w, dtype=torch.bool, device=dev)
            masks_o = np.zeros((5, h, w), np.uint8)
            total = 0
            for ref in range(5):
                ids = [ref] + [i for i in range(5) if i != ref]
                od = torch.empty(h, w, device=dev)
                on = torch.empty(h, w, 3, device=dev)
                S.fuse_view(ids, [cams_c[i] for i in ids], [dd[i] for i in ids],
                            [nd[i] for i in ids], masks_g, 2, od, on, 1.0, 0.01,
                            math.radians(30.0), precision=precision)
                torch.cuda.synchronize()
                od_o, _, _ = O.fuse_view([ocams[i] for i in ids], [ds[i] for i in ids],
                                         [ns[i] for i in ids], masks_o, ref, ids[1:], 2, 1.0, 0.01,
                                         float(np.float32(math.cos(math.radians(30.0)))),
                                         precision=precision)
                gd = od.cpu().numpy()
                assert np.array_equal(gd.view(np.uint32), od_o.view(np.uint32))
                total += int((gd > 0).sum())
            counts[(shift, precision)] = total
    print(f"fused points (shift, precision): {counts}")
    # a fused pixel masks the source pixels it used, so the total is about
    # one view's worth of surface points
    base = counts[(0.0, 64)]
    assert base > 0.9 * h * w
    assert abs(counts[(0.0, 32)] - base) <= 0.001 * base
    assert counts[(1.0, 64)] >= 0.999 * base
    assert counts[(1.0, 32)] <= 0.5 * base
