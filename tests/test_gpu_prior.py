"""Device planar-prior construction (gmvs_prior_build_device) against the
host construction (gmvs_prior_build, pinned by test_cpu_prior.py): prior
planes and triangle masks bit for bit, and the retained-triangle count, with
and without the superpixel filter, on support-point sets from 0 to tens of
thousands of points, labels with unlabelled (0) regions, and a depth range
that cuts some triangles' pixels. The device Delaunay step on its own
(gmvs_delaunay_cells_device) against the host's incremental construction on
cocircular lattices, sparse sets and sets with large empty regions."""
import numpy as np
import pytest
import torch

from gmvs_amd.lib.build.planar_initer import build_prior, build_prior_device, delaunay, delaunay_cells
from gmvs_amd.synthetic import make_scene, voronoi_labels
from tests.test_cpu_prior import cell_points

pytestmark = pytest.mark.gpu


def _case(w, h, seed, good_frac, cells, zero_frac):
    from gmvs_amd import src as S

    sc = make_scene(w, h, 2, 8, seed=seed)
    c0 = sc.cameras[0]
    cam = S.load_camera(c0.K.reshape(-1), c0.R.reshape(-1), c0.t, h, w, c0.depth_min, c0.depth_max)._c
    rng = np.random.default_rng(seed)
    depth = sc.depth[0].numpy().astype(np.float32)
    cost = rng.uniform(0.1, 2.5, (h, w)).astype(np.float32)
    good = rng.random((h, w)) < good_frac
    cost[good] = rng.uniform(0.0, 0.1, good.sum()).astype(np.float32)
    labels = voronoi_labels(w, h, cells, seed=seed + 1).astype(np.int32)
    labels[rng.random((h, w)) < zero_frac] = 0
    return cam, depth, cost, labels


@pytest.mark.parametrize("w,h,seed,good,cells,zero,filt", [
    (320, 240, 1, 0.30, 60, 0.0, True),
    (320, 240, 2, 0.30, 60, 0.08, True),
    (320, 240, 3, 0.30, 60, 0.0, False),
    (333, 251, 4, 0.05, 40, 0.02, True),
    (1600, 1200, 5, 0.40, 2000, 0.01, True),
])
def test_prior_device_equals_host(w, h, seed, good, cells, zero, filt):
    cam, depth, cost, labels = _case(w, h, seed, good, cells, zero)
    dmin, dmax = 2.2, 5.5  # narrower than the scene's depths: the range check drops pixels
    hp, hm, hn = build_prior(cam, depth, cost, labels, dmin, dmax, superpixel_filter=filt)
    dev = torch.device("cuda")
    gp, gm, gn = build_prior_device(cam, torch.from_numpy(depth).to(dev),
                                    torch.from_numpy(cost).to(dev),
                                    torch.from_numpy(labels).to(dev), dmin, dmax,
                                    superpixel_filter=filt)
    gp, gm = gp.cpu().numpy(), gm.cpu().numpy()
    print(f"{w}x{h}: triangles host {hn} device {gn}; masked pixels {int((hm > 0).sum())}")
    assert gn == hn and hn > 0
    assert np.array_equal(gm, hm)
    assert np.array_equal(gp.view(np.uint32), hp.view(np.uint32))


def test_prior_device_without_support_points():
    cam, depth, cost, labels = _case(160, 120, 9, 0.0, 10, 0.0, )
    dev = torch.device("cuda")
    gp, gm, gn = build_prior_device(cam, torch.from_numpy(depth).to(dev),
                                    torch.from_numpy(cost).to(dev), None, 1.0, 8.0)
    assert gn == 0
    assert int(gm.abs().sum()) == 0 and float(gp.abs().sum()) == 0.0


@pytest.mark.parametrize("ncx,ncy,seed,keep,mode,holes", [
    (1, 1, 0, 1.0, "random", ()),
    (2, 2, 2, 1.0, "corner", ()),
    (12, 9, 3, 1.0, "corner", ()),
    (20, 15, 5, 0.9, "two", ()),
    (20, 15, 6, 0.15, "random", ()),
    (40, 30, 8, 0.5, "two", ((5, 5, 6), (30, 25, 9))),
    (320, 240, 9, 0.7, "random", ((100, 80, 40), (250, 200, 25), (30, 200, 60))),
    (640, 480, 10, 0.6, "corner", ((300, 200, 120),)),
])
def test_device_delaunay_equals_host(ncx, ncy, seed, keep, mode, holes):
    c = cell_points(ncx, ncy, seed, keep, mode, holes)
    xy = c.reshape(-1, 2)
    pts = xy[xy[:, 0] >= 0]
    host = delaunay(pts) if len(pts) else np.zeros((0, 3), np.int32)
    dev = delaunay_cells(c, device=torch.device("cuda"))
    print(f"{ncx}x{ncy} cells, {len(pts)} points: {len(host)} triangles host, {len(dev)} device")
    assert dev.shape == host.shape and np.array_equal(dev, host)


def test_prior_device_without_count_equals_counted():
    cam, depth, cost, labels = _case(640, 480, 11, 0.3, 200, 0.01)
    dev = torch.device("cuda")
    args = (cam, torch.from_numpy(depth).to(dev), torch.from_numpy(cost).to(dev),
            torch.from_numpy(labels).to(dev), 2.2, 5.5)
    p1, m1, n1 = build_prior_device(*args, superpixel_filter=True)
    p2, m2, n2 = build_prior_device(*args, superpixel_filter=True, count=False)
    torch.cuda.synchronize()
    assert n1 > 0 and n2 is None
    assert torch.equal(m1, m2)
    assert torch.equal(p1.view(torch.int32), p2.view(torch.int32))
