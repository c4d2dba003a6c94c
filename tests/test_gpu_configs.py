"""GPU parity at the benchmark's own sizes, against the FP64 oracle under the
same Philox stream.

* The 10-view ACMP workload (bench.make_workload: 1600x1200, 24 planes +
  10 % Gaussian bumps, 2000-cell Voronoi superpixels, 9 sources): view 0
  through the plain pass, the host prior with the superpixel filter, the
  prior pass and the geometric pass, plus a 4000-pair cost batch. The passes
  run 2 iterations instead of the benchmark's 8 so that the FP64 oracle
  finishes in ~25 s per pass on the host.
* The 4-scene batch (bench.SCENES4: 12 views each at 1024x768, 1600x1200,
  2048x1536 and 3200x2400, 11 sources): all 48 views in one batch captured in
  one HIP graph and replayed (bit-identical to the eager batch), one view of
  each size run on its own (bit-identical to its batched result) and checked
  against FP64 at 2 iterations.
"""
import functools
from types import SimpleNamespace

import numpy as np
import pytest
import torch

from oracle import oracle as O
from tests import helpers as T
from tests.test_gpu_pm import _ccams, _exempt, _pm_gpu, gpu_cost_batch

pytestmark = pytest.mark.gpu

ITERS = 2


# ------------------------------------------------------------ 10-view ACMP ---

@functools.lru_cache(maxsize=1)
def _acmp():
    import bench

    sc, labels = bench.make_workload("cuda")
    ids = [0] + list(sc.pairs[0][:9])
    cams = T.oracle_cams(sc, ids)
    imgs = [sc.images[i].cpu().numpy() for i in ids]
    dmin, dmax = T.depth_range(sc, 0)
    kw = dict(max_iterations=ITERS, depth_min=dmin, depth_max=dmax)
    return sc, labels, ids, cams, imgs, kw


@functools.lru_cache(maxsize=1)
def _acmp_plain_oracle():
    sc, labels, ids, cams, imgs, kw = _acmp()
    return O.pm_run(cams, imgs, O.make_params(**kw), precision=64, rng_view=0)


def test_acmp_cost_batch_matches_fp64():
    sc, labels, ids, cams, imgs, kw = _acmp()
    rng = np.random.default_rng(21)
    xy, planes = T.random_planes(sc, 0, 2000, rng, kw["depth_min"], kw["depth_max"])
    gxy = np.stack([rng.integers(0, 1600, 2000), rng.integers(0, 1200, 2000)], 1).astype(np.int32)
    xy = np.concatenate([xy, gxy])
    host = SimpleNamespace(cameras=sc.cameras, depth=sc.depth.cpu(), normal=sc.normal.cpu())
    planes = np.concatenate([planes, T.gt_planes(host, 0, gxy)])
    gpu = gpu_cost_batch(sc, ids, xy, planes)
    ref, edge, minvar = O.cost_batch(cams, imgs, O.make_params(num_images=len(ids)), xy, planes)
    err = np.abs(gpu - ref)
    bad = (err > 1e-4) & ~_exempt(edge, minvar)
    print(f"1600x1200, 9 sources: pairs={err.size} max|err|={err.max():.3g} "
          f"p99.9={np.quantile(err, 0.999):.3g} beyond_1e-4_non_exempt={bad.sum()}")
    assert bad.sum() == 0


def test_acmp_plain_pass_matches_fp64():
    sc, labels, ids, cams, imgs, kw = _acmp()
    g = _pm_gpu(sc, ids, kw)
    o = _acmp_plain_oracle()
    agree = T.depth_normal_agreement(g["depth"], g["normal"], o["depth"], o["normal"])
    print(f"1600x1200 plain pass ({ITERS} iterations): agreement vs FP64 = {agree:.4f}")
    assert agree >= 0.99


def test_acmp_prior_pass_matches_fp64():
    from gmvs_amd.lib.build import planar_initer as PI

    sc, labels, ids, cams, imgs, kw = _acmp()
    o1 = _acmp_plain_oracle()
    pp, masks, ntri = PI.build_prior(_ccams(sc, [0])[0], o1["depth"], o1["cost"], labels[0],
                                     kw["depth_min"], kw["depth_max"], superpixel_filter=True)
    assert ntri > 1000
    kw2 = dict(kw, planar_prior=True, superpixel_filter=True)
    g = _pm_gpu(sc, ids, kw2, init_mode=2, prior=(pp, masks),
                state=(o1["planes"].copy(), o1["cost"].copy()))
    args = dict(init_mode=2, prior_planes=pp, plane_masks=masks.astype(np.uint32),
                planes=o1["planes"], costs=o1["cost"], rng_view=0)
    o = O.pm_run(cams, imgs, O.make_params(**kw2), precision=64, **args)
    o32 = O.pm_run(cams, imgs, O.make_params(**kw2), precision=32, **args)
    agree = T.depth_normal_agreement(g["depth"], g["normal"], o["depth"], o["normal"])
    agree32 = T.depth_normal_agreement(o32["depth"], o32["normal"], o["depth"], o["normal"])
    rel = np.abs(g["depth"] - o["depth"]) / np.abs(o["depth"])
    print(f"1600x1200 prior pass: {ntri} triangles, masked {np.mean(masks > 0):.3f}; agreement "
          f"vs FP64 = {agree:.4f} (oracle FP32 {agree32:.4f}); depth within 0.5% "
          f"{np.mean(rel <= 0.005):.4f}")
    # the flat top of the restricted score makes masked pixels' normals
    # rounding-dependent (tests/test_cpu_prior_ties.py); depth is not
    assert np.mean(rel <= 0.005) >= 0.99
    assert agree >= agree32 - 0.01


def test_acmp_device_prior_equals_host():
    """The device planar-prior construction (Delaunay included) on the
    support points of a real pass: view 0 of the 10-view workload after the
    GPU plain pass (8 iterations), tens of thousands of support points with
    the gaps a pass leaves (occlusions, slanted planes): equal to the host
    construction bit for bit, and its time."""
    import time

    from gmvs_amd.lib.build import planar_initer as PI

    sc, labels, ids, cams, imgs, kw = _acmp()
    g = _pm_gpu(sc, ids, dict(kw, max_iterations=8))
    cam = _ccams(sc, [0])[0]
    hp, hm, hn = PI.build_prior(cam, g["depth"], g["cost"], labels[0], kw["depth_min"],
                                kw["depth_max"], superpixel_filter=True)
    dev = torch.device("cuda")
    d, c = torch.from_numpy(g["depth"]).to(dev), torch.from_numpy(g["cost"]).to(dev)
    lab = torch.from_numpy(np.ascontiguousarray(labels[0], np.int32)).to(dev)
    args = (cam, d, c, lab, kw["depth_min"], kw["depth_max"])
    PI.build_prior_device(*args, superpixel_filter=True)  # warm
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    gp, gm, gn = PI.build_prior_device(*args, superpixel_filter=True)
    ms = (time.perf_counter() - t0) * 1e3
    print(f"1600x1200 prior from a GPU plain pass: {hn} retained triangles (host), {gn} (device); "
          f"device build {ms:.1f} ms")
    assert gn == hn and hn > 1000
    assert np.array_equal(gm.cpu().numpy(), hm)
    assert np.array_equal(gp.cpu().numpy().view(np.uint32), hp.view(np.uint32))


def test_acmp_geom_pass_matches_fp64():
    sc, labels, ids, cams, imgs, kw = _acmp()
    rng = np.random.default_rng(23)
    depths = [(sc.depth[i].cpu().numpy()
               * (1 + 0.01 * rng.standard_normal(sc.depth[i].shape))).astype(np.float32)
              for i in ids]
    normal = sc.normal[0].cpu().numpy()
    kw3 = dict(kw, geom_consistency=True)
    g = _pm_gpu(sc, ids, kw3, init_mode=1, geom=True, depths=depths, ref_normal=normal)
    args = dict(init_mode=1, geom=True, depths=depths, ref_normal=normal, rng_view=0)
    o = O.pm_run(cams, imgs, O.make_params(**kw3), precision=64, **args)
    o32 = O.pm_run(cams, imgs, O.make_params(**kw3), precision=32, **args)
    agree = T.depth_normal_agreement(g["depth"], g["normal"], o["depth"], o["normal"])
    agree32 = T.depth_normal_agreement(o32["depth"], o32["normal"], o["depth"], o["normal"])
    rel = np.abs(g["depth"] - o["depth"]) / np.abs(o["depth"])
    print(f"1600x1200 geometric pass: agreement vs FP64 = {agree:.4f} (oracle FP32 {agree32:.4f}); "
          f"depth within 0.5% {np.mean(rel <= 0.005):.4f}")
    # The source depth maps hold holes (depth 0 on ~10 % of pixels): the
    # geometric term's bilinear depth sample next to a hole is
    # ill-conditioned (FP32 and FP64 reprojection errors differ by up to
    # ~0.15 px there), and normals drift across near-equal costs, so the
    # depth + 1-degree-normal agreement is bounded by FP32 itself: the GPU
    # must track FP64 at least as well as the oracle's FP32 instantiation.
    assert np.mean(rel <= 0.005) >= 0.99
    assert agree >= agree32 - 0.01


# ------------------------------------------------- 4 scenes, uneven sizes ---

@functools.lru_cache(maxsize=1)
def _scenes4():
    """All 48 plain-pass problems of the 4-scene batch (2 iterations), run as
    one batch captured in one HIP graph; returns the scenes, the problems and
    the eager batch's planes/costs of the first view of each scene (CPU)."""
    import bench
    from gmvs_amd import _native as N
    from gmvs_amd import src as S
    from gmvs_amd.pipeline import cameras_from_synthetic
    from gmvs_amd.synthetic import make_scene

    dev = "cuda"
    scenes, problems = [], []
    for i, (w, h, npl, dr) in enumerate(bench.SCENES4):
        sc = make_scene(w, h, 12, npl, seed=3407 + 40 + i, depth_range=dr,
                        cam_depth_range=(dr[0] + 0.25, dr[1]), device=dev)
        cams = cameras_from_synthetic(sc.cameras)
        imgs = sc.images.to(dev, torch.float32).contiguous()
        for ref in range(12):
            ids = [ref] + list(sc.pairs[ref][:11])
            prm = S.pm_params_for_problem(S.PatchMatchParams(), cams[ref], len(ids))
            prm.max_iterations = ITERS
            problems.append(dict(cams=[cams[j]._c for j in ids], images=[imgs[j] for j in ids],
                                 params=prm, init_mode=N.INIT_RANDOM, geom=False, rng_view=ref,
                                 planes=torch.zeros(h, w, 4, device=dev),
                                 costs=torch.zeros(h, w, device=dev)))
        scenes.append(sc)
    torch.cuda.synchronize()  # the scenes and images were made on the default stream
    ws = S._Workspace()
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        S.run_pm_batch(problems, ws)  # eager (also sizes the workspace)
    torch.cuda.synchronize()
    eager = [(p["planes"].clone(), p["costs"].clone()) for p in problems]
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=stream):
        S.run_pm_batch(problems, ws)
    torch.cuda.synchronize()
    replays_equal = []
    for rep in range(2):
        for p in problems:  # eager work between replays
            p["planes"].fill_(-1.0)
            p["costs"].fill_(-1.0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        # bitwise (a NaN equals itself bit for bit)
        replays_equal.append(all(
            torch.equal(p["planes"].view(torch.int32), e[0].view(torch.int32))
            and torch.equal(p["costs"].view(torch.int32), e[1].view(torch.int32))
            for p, e in zip(problems, eager)))
    del g, ws
    firsts = [12 * k for k in range(4)]
    batched = {k: (eager[k][0].cpu().numpy(), eager[k][1].cpu().numpy()) for k in firsts}
    del eager
    torch.cuda.empty_cache()
    return scenes, problems, batched, replays_equal


def test_scenes4_graph_replay_and_single_views():
    from gmvs_amd import src as S

    scenes, problems, batched, replays_equal = _scenes4()
    print(f"48 views, 162,385,920 pixels: replays equal to eager = {replays_equal}")
    assert all(replays_equal)
    for k, (planes_b, costs_b) in batched.items():
        p = problems[k]
        h, w = p["planes"].shape[:2]
        p1 = dict(p, planes=torch.zeros(h, w, 4, device="cuda"), costs=torch.zeros(h, w, device="cuda"))
        kw = {x: v for x, v in p1.items() if x not in ("cams", "images", "params")}
        S.run_pm(p1["cams"], p1["images"], p1["params"], **kw)
        torch.cuda.synchronize()
        assert np.array_equal(p1["planes"].cpu().numpy().view(np.uint32), planes_b.view(np.uint32)), k
        assert np.array_equal(p1["costs"].cpu().numpy().view(np.uint32), costs_b.view(np.uint32)), k
        print(f"view {k} ({w}x{h}): one-view run equals its batched result")


@pytest.mark.parametrize("scene", [0, 1, 2, 3])
def test_scenes4_view_matches_fp64(scene):
    import bench

    scenes, problems, batched, _ = _scenes4()
    sc = scenes[scene]
    k = 12 * scene
    ids = [0] + list(sc.pairs[0][:11])
    p = problems[k]
    prm = p["params"]
    kw = dict(max_iterations=ITERS, depth_min=prm.depth_min, depth_max=prm.depth_max)
    o = O.pm_run(T.oracle_cams(sc, ids), [sc.images[i].cpu().numpy() for i in ids],
                 O.make_params(**kw), precision=64, rng_view=0)
    planes_b = batched[k][0]
    d, n = planes_b[..., 3], planes_b[..., :3]
    agree = T.depth_normal_agreement(d, n, o["depth"], o["normal"])
    w, h = bench.SCENES4[scene][:2]
    print(f"scene {scene} ({w}x{h}, 11 sources, {ITERS} iterations): agreement vs FP64 = {agree:.4f}")
    assert agree >= 0.99
