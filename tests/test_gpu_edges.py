"""Edge cases of the PatchMatch path against the FP64 oracle: images smaller
than two patch windows (every tap near a border), the largest view count
(31 sources: every bit of the 32-bit view masks below bit 31) and a single
source view."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import helpers as T
from tests.test_gpu_pm import _ccams, _pm_gpu, gpu_cost_batch

pytestmark = pytest.mark.gpu


def _costs_vs_fp64(sc, ids, n=600, seed=3):
    rng = np.random.default_rng(seed)
    dmin, dmax = T.depth_range(sc, 0)
    xy, planes = T.random_planes(sc, 0, n, rng, dmin, dmax)
    gpu = gpu_cost_batch(sc, ids, xy, planes)
    ref, edge, minvar = O.cost_batch(T.oracle_cams(sc, ids), [sc.images[i].numpy() for i in ids],
                                     O.make_params(num_images=len(ids)), xy, planes)
    exempt = (edge < 1e-3) | ((minvar > 0.5e-5) & (minvar < 2e-5))
    err = np.abs(gpu - ref)
    return err, exempt


def _pass_agreement(sc, ids, iters):
    dmin, dmax = T.depth_range(sc, 0)
    kw = dict(max_iterations=iters, depth_min=dmin, depth_max=dmax)
    g = _pm_gpu(sc, ids, kw)
    cams = T.oracle_cams(sc, ids)
    imgs = [sc.images[i].numpy() for i in ids]
    o64 = O.pm_run(cams, imgs, O.make_params(**kw), precision=64, rng_view=0)
    o32 = O.pm_run(cams, imgs, O.make_params(**kw), precision=32, rng_view=0)
    agree = T.depth_normal_agreement(g["depth"], g["normal"], o64["depth"], o64["normal"])
    agree32 = T.depth_normal_agreement(o32["depth"], o32["normal"], o64["depth"], o64["normal"])
    return agree, agree32


@pytest.mark.parametrize("w,h", [(24, 18), (13, 29)])
def test_tiny_images(w, h):
    sc = T.scene(width=w, height=h, views=5, planes=3)
    ids = [0, 1, 2, 3, 4]
    err, exempt = _costs_vs_fp64(sc, ids)
    bad = (err > 1e-4) & ~exempt
    print(f"tiny {w}x{h}: cost max|err| {err.max():.2e}, non-exempt beyond 1e-4: {bad.sum()}")
    assert bad.sum() == 0
    init, _ = _pass_agreement(sc, ids, 0)
    agree, agree32 = _pass_agreement(sc, ids, 3)
    print(f"tiny {w}x{h}: init agreement {init:.4f}; 3-iteration pass {agree:.4f} "
          f"(oracle FP32 {agree32:.4f})")
    assert init >= 0.999
    assert agree >= agree32 - 0.02


def test_thirty_one_sources():
    sc = T.scene(width=48, height=36, views=32, planes=3)
    ids = list(range(32))
    err, exempt = _costs_vs_fp64(sc, ids, n=200)
    bad = (err > 1e-4) & ~exempt
    print(f"31 sources: cost max|err| {err.max():.2e}, non-exempt beyond 1e-4: {bad.sum()}")
    assert bad.sum() == 0
    init, _ = _pass_agreement(sc, ids, 0)
    agree, agree32 = _pass_agreement(sc, ids, 2)
    print(f"31 sources: init agreement {init:.4f}; 2-iteration pass {agree:.4f} "
          f"(oracle FP32 {agree32:.4f})")
    assert init >= 0.999
    assert agree >= agree32 - 0.02


def test_single_source():
    sc = T.scene(width=64, height=48, views=2, planes=3)
    ids = [0, 1]
    err, exempt = _costs_vs_fp64(sc, ids)
    bad = (err > 1e-4) & ~exempt
    assert bad.sum() == 0
    init, _ = _pass_agreement(sc, ids, 0)
    agree, agree32 = _pass_agreement(sc, ids, 3)
    print(f"1 source: init agreement {init:.4f}; 3-iteration pass {agree:.4f} "
          f"(oracle FP32 {agree32:.4f})")
    assert init >= 0.999
    assert agree >= agree32 - 0.02
