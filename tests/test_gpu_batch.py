"""Batched passes (gmvs_pm_run_batch): one schedule for many reference views of
different scenes and sizes must give exactly the results of running the views
one by one, and the schedule must replay correctly from a HIP graph."""
import numpy as np
import pytest
import torch

from gmvs_amd.synthetic import make_scene, voronoi_labels
from tests import helpers as T

pytestmark = pytest.mark.gpu


def _native():
    from gmvs_amd import _native as N
    from gmvs_amd import src as S

    return N, S


def _problems(scenes, refs, iters, precision, dev):
    """Plain-pass problems (ref view + its pair sources) with fresh state
    and output tensors."""
    N, S = _native()
    probs, outs = [], []
    for sc, ref in zip(scenes, refs):
        ids = [ref] + [i for i in range(len(sc.cameras)) if i != ref][:4]
        cams = [S.load_camera(sc.cameras[i].K.reshape(-1), sc.cameras[i].R.reshape(-1),
                              sc.cameras[i].t, sc.cameras[i].height, sc.cameras[i].width,
                              sc.cameras[i].depth_min, sc.cameras[i].depth_max)._c for i in ids]
        imgs = [sc.images[i].to(dev).contiguous() for i in ids]
        h, w = imgs[0].shape
        p = S.PatchMatchParams()
        p.max_iterations = iters
        dmin, dmax = T.depth_range(sc, ref)
        p.depth_min, p.depth_max = dmin, dmax
        cp = p.to_c()
        cp.num_images = len(ids)
        o = dict(planes=torch.zeros(h, w, 4, device=dev), costs=torch.zeros(h, w, device=dev),
                 out_depth=torch.zeros(h, w, device=dev), out_normal=torch.zeros(h, w, 3, device=dev),
                 out_cost=torch.zeros(h, w, device=dev))
        probs.append(dict(cams=cams, images=imgs, params=cp, init_mode=N.INIT_RANDOM, geom=False,
                          rng_view=ref, precision=precision, **o))
        outs.append(o)
    return probs, outs


def _snapshot(outs):
    return [{k: v.cpu().numpy().copy() for k, v in o.items()} for o in outs]


def _assert_same(a, b):
    for x, y in zip(a, b):
        for k in x:
            assert np.array_equal(x[k].view(np.uint32), y[k].view(np.uint32)), k


@pytest.mark.parametrize("precision", [0, 2])
def test_batch_equals_single_runs(precision):
    N, S = _native()
    dev = torch.device("cuda")
    big = T.scene()                                   # 320x240
    small = make_scene(200, 136, 5, 4, seed=3407 + 7)  # another scene and size
    scenes = [big, small, big]
    refs = [0, 0, 2]
    probs, outs = _problems(scenes, refs, 2, precision, dev)
    S.run_pm_batch(probs)
    torch.cuda.synchronize()
    batched = _snapshot(outs)
    for p in probs:
        kw = {k: v for k, v in p.items() if k not in ("cams", "images", "params")}
        S.run_pm(p["cams"], p["images"], p["params"], **kw)
    torch.cuda.synchronize()
    single = _snapshot(outs)
    _assert_same(batched, single)
    assert all(np.isfinite(o["out_depth"]).all() for o in batched)


def test_batch_graph_replay():
    N, S = _native()
    dev = torch.device("cuda")
    sc = T.scene()
    probs, outs = _problems([sc, sc], [0, 1], 2, 0, dev)
    ws = S._Workspace()
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        S.run_pm_batch(probs, ws)  # eager reference (also sizes the workspace)
    torch.cuda.synchronize()
    eager = _snapshot(outs)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=stream):
        S.run_pm_batch(probs, ws)
    torch.cuda.synchronize()
    for rep in range(3):
        for o in outs:  # eager work between replays
            for v in o.values():
                v.fill_(-1.0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        _assert_same(_snapshot(outs), eager)


def test_pipeline_batched_equals_per_view():
    N, S = _native()
    from gmvs_amd.pipeline import MultiViewPipeline, cameras_from_synthetic

    sc = make_scene(240, 180, 5, 5, seed=3407 + 9)
    cams = cameras_from_synthetic(sc.cameras)
    pairs = [p[:4] for p in sc.pairs]
    labels = [voronoi_labels(240, 180, 40, seed=3407 + 20 + i) for i in range(5)]
    res = []
    for batched in (True, False):
        pipe = MultiViewPipeline(sc.images, cams, pairs, labels=labels, prior_threads=4,
                                 batched=batched)
        s1 = pipe.first_stage(iterations=2, planar_prior=True, superpixel_filter=True)
        s2 = pipe.geom_pass(s1, iterations=1)
        torch.cuda.synchronize()
        res.append([t.cpu().numpy() for t in (s1.depth, s1.normal, s2.depth, s2.normal, s2.cost)])
    for a, b in zip(*res):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_pipeline_device_prior_equals_host_prior():
    """The first stage with the planar priors built on the device in stream
    order (no host copies or synchronisation) returns the host-prior
    pipeline's bits."""
    N, S = _native()
    from gmvs_amd.pipeline import MultiViewPipeline, cameras_from_synthetic

    sc = make_scene(240, 180, 5, 5, seed=3407 + 13)
    cams = cameras_from_synthetic(sc.cameras)
    pairs = [p[:4] for p in sc.pairs]
    labels = [voronoi_labels(240, 180, 40, seed=3407 + 40 + i) for i in range(5)]
    res = []
    for device_prior in (True, False):
        pipe = MultiViewPipeline(sc.images, cams, pairs, labels=labels, prior_threads=4,
                                 device_prior=device_prior)
        s1 = pipe.first_stage(iterations=2, planar_prior=True, superpixel_filter=True)
        torch.cuda.synchronize()
        res.append([t.cpu().numpy() for t in (s1.depth, s1.normal, s1.cost)])
    for a, b in zip(*res):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_pipeline_auto_precision_equals_fp32():
    """precision="auto" picks FP16 cells for 8-bit imagery, which hold it
    exactly: the whole pipeline (prior, geometric pass) returns the FP32
    path's bits."""
    N, S = _native()
    from gmvs_amd.pipeline import MultiViewPipeline, cameras_from_synthetic

    sc = make_scene(240, 180, 5, 5, seed=3407 + 11)
    assert S.fp16_cells_exact(sc.images)
    cams = cameras_from_synthetic(sc.cameras)
    pairs = [p[:4] for p in sc.pairs]
    labels = [voronoi_labels(240, 180, 40, seed=3407 + 30 + i) for i in range(5)]
    res = []
    for prec in ("auto", N.PREC_FP32):
        pipe = MultiViewPipeline(sc.images, cams, pairs, labels=labels, prior_threads=4,
                                 precision=prec)
        assert pipe.precision == (N.PREC_FP16 if prec == "auto" else N.PREC_FP32)
        s1 = pipe.first_stage(iterations=2, planar_prior=True, superpixel_filter=True)
        s2 = pipe.geom_pass(s1, iterations=1)
        torch.cuda.synchronize()
        res.append([t.cpu().numpy() for t in (s1.depth, s1.normal, s2.depth, s2.normal, s2.cost)])
    for a, b in zip(*res):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_autotune_picks_from_its_candidates_and_leaves_state_alone():
    """The launch-shape tuner on a two-problem batch: the choice comes from
    the candidate sets, the callers' state tensors are untouched (trials run
    on scratch state started from a short warm pass), and reset() restores
    the library defaults, after which a run equals an untuned one."""
    from gmvs_amd import autotune as AT

    N, S = _native()
    dev = torch.device("cuda")
    scenes = [T.scene(), make_scene(200, 136, 5, 4, seed=3407 + 7)]
    probs, outs = _problems(scenes, [0, 0], 2, 0, dev)
    before = _snapshot(outs)
    AT.reset()
    try:
        choice = AT.autotune(probs, tile_cols=(8, 16), waves=(0, 2))
        assert choice["tile_cols"] in (8, 16)
        assert choice["blocks_candidates"] in (0, AT._blocks(2, 4))
        assert choice["blocks_refine"] in (0, AT._blocks(2, 4))
        torch.cuda.synchronize()
        _assert_same(before, _snapshot(outs))
    finally:
        AT.reset()
    S.run_pm_batch(probs)
    torch.cuda.synchronize()
    tuned_off = _snapshot(outs)
    probs2, outs2 = _problems(scenes, [0, 0], 2, 0, dev)
    S.run_pm_batch(probs2)
    torch.cuda.synchronize()
    _assert_same(tuned_off, _snapshot(outs2))
