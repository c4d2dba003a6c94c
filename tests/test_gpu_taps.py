"""Tap indexing forms and the FP16 cell format (C-ABI calls).

A wave whose patches all map inside the padded cell grid indexes its taps as
floor / fract on the padded grid; any other wave clamps every tap as the
reference's software bilinear does (gmvs/src/patch_match.cu:16-31). The two
forms must give the same bits: gmvs_set_fast_taps(0) forces the clamped form
everywhere. FP16 cells hold 8-bit intensities and their differences exactly
and the bilinear sample is FP32 arithmetic on the widened values, so FP16
costs equal the FP32 path's up to the order of the FP32 sums (FP32 runs
4-lane groups, 16-bit cells 2-lane groups).
"""
import numpy as np
import pytest

from tests import helpers as T
from tests.test_gpu_pm import _params, _pm_gpu, gpu_cost_batch

pytestmark = pytest.mark.gpu


def _with_fast(enable, fn):
    from gmvs_amd import _native as N

    N.set_fast_taps(enable)
    try:
        return fn()
    finally:
        N.set_fast_taps(True)


def _queries(sc, n, seed):
    rng = np.random.default_rng(seed)
    dmin, dmax = T.depth_range(sc, 0)
    xy, planes = T.random_planes(sc, 0, n, rng, dmin, dmax)
    # ground-truth planes at border pixels: patches that cross the image edges
    h, w = sc.cameras[0].height, sc.cameras[0].width
    bx = np.concatenate([np.arange(0, 6), np.arange(w - 6, w), rng.integers(0, w, 24)])
    by = np.concatenate([rng.integers(0, h, 12), np.arange(0, 6), np.arange(h - 6, h),
                         rng.integers(0, h, 12)])
    bxy = np.stack([np.concatenate([bx, rng.integers(0, w, 24)]),
                    np.concatenate([by[:36], rng.integers(0, h, 24)])], 1).astype(np.int32)
    return np.concatenate([xy, bxy]), np.concatenate([planes, T.gt_planes(sc, 0, bxy)])


@pytest.mark.parametrize("precision", [0, 1, 2])
def test_cost_batch_fast_taps_equal_clamped(precision):
    sc = T.scene()
    ids = [0, 1, 2, 3, 4]
    xy, planes = _queries(sc, 3000, 21)
    fast = _with_fast(True, lambda: gpu_cost_batch(sc, ids, xy, planes, precision=precision))
    slow = _with_fast(False, lambda: gpu_cost_batch(sc, ids, xy, planes, precision=precision))
    assert np.array_equal(fast.astype(np.float32).view(np.uint32),
                          slow.astype(np.float32).view(np.uint32))


@pytest.mark.parametrize("precision", [0, 2])
def test_pm_pass_fast_taps_equal_clamped(precision):
    sc = T.scene()
    ids = [0, 1, 2, 3, 4]
    kw = _params(sc, iters=2)
    a = _with_fast(True, lambda: _pm_gpu(sc, ids, kw, precision=precision))
    b = _with_fast(False, lambda: _pm_gpu(sc, ids, kw, precision=precision))
    for k in ("depth", "normal", "cost"):
        assert np.array_equal(a[k].view(np.uint32), b[k].view(np.uint32)), k


def test_fp16_cells_match_fp32_on_8bit_images():
    sc = T.scene()
    ids = [0, 1, 2, 3, 4]
    img = sc.images[0].numpy()
    assert np.array_equal(img, np.round(img)) and img.min() >= 0 and img.max() <= 255
    xy, planes = _queries(sc, 2000, 5)
    c32 = gpu_cost_batch(sc, ids, xy, planes, precision=0)
    c16 = gpu_cost_batch(sc, ids, xy, planes, precision=1)
    d = np.abs(c32 - c16)
    print(f"|FP16 - FP32| max {d.max():.3g}, p99.9 {np.quantile(d, 0.999):.3g}")
    # summation order only (measured: p99.9 1.0e-5, max 3.3e-5; weights rounded
    # to 16 bits, as in the BF16 path, give ~1e-3): within the 1e-4 bar of the
    # FP32 path against the FP64 oracle
    assert np.quantile(d, 0.999) <= 3e-5
    assert d.max() <= 1e-4
