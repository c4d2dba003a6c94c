"""16-bit sample precisions (GMVS_PREC_FP16 / GMVS_PREC_BF16) against the FP64
oracle: per-(pixel, hypothesis) cost error and final depth/normal agreement
under the same Philox stream, on the 5-view 320x240 scene (4 iterations).

The cells hold uint8 intensities and their differences exactly in both
formats. FP16 samples are FP32 arithmetic on the widened cell values
(v_fma_mix_f32), so FP16 carries the FP32 path's error (and its bits, see
test_gpu_taps.py). BF16 samples are BF16 dot products of both cell rows
with the weight pair (1, fx), fx rounded to 8 significant bits, blended by
an FP32 fy; the NCC sums stay FP32. Small cost errors shift the Monte-Carlo
view selection and the adoption order, so the joint depth + 1-degree-normal
agreement drops faster than the depth error.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests import helpers as T
from tests.test_gpu_pm import _exempt, _params, _pm_gpu, gpu_cost_batch

pytestmark = pytest.mark.gpu

# precision code, bar on the 99.9th percentile of |cost error| over the
# non-exempt pairs, floor of the depth (0.5%) + normal (1 deg) agreement,
# floor of the fraction of depths within 0.5%. FP16 is held to the accuracy
# target (costs to 1e-4, >= 99 % of depths within 0.5 %). BF16 does not
# meet it (8 significant bits in the tap weights: costs to ~5e-4, ~99.0 %
# of depths within 0.5 %); its floors here only guard against regressions.
MODES = {"fp16": (1, 1e-4, 0.99, 0.99), "bf16": (2, 2e-3, 0.85, 0.98)}


@pytest.mark.parametrize("mode", sorted(MODES))
def test_cost_error_vs_fp64(mode):
    prec, p999_bar, _, _ = MODES[mode]
    sc = T.scene()
    ids = [0, 1, 2, 3, 4]
    rng = np.random.default_rng(11)
    dmin, dmax = T.depth_range(sc, 0)
    xy, planes = T.random_planes(sc, 0, 4000, rng, dmin, dmax)
    gt_xy = xy.copy()
    gt = T.gt_planes(sc, 0, gt_xy)
    xy = np.concatenate([xy, gt_xy])
    planes = np.concatenate([planes, gt])
    gpu = gpu_cost_batch(sc, ids, xy, planes, precision=prec)
    ref, edge, minvar = O.cost_batch(T.oracle_cams(sc, ids), [sc.images[i].numpy() for i in ids],
                                     O.make_params(num_images=5), xy, planes)
    err = np.abs(gpu - ref)
    keep = ~_exempt(edge, minvar)
    e = err[keep]
    p50, p99, p999 = np.quantile(e, [0.5, 0.99, 0.999])
    print(f"[{mode}] pairs={e.size} median|err|={p50:.2e} p99={p99:.2e} p99.9={p999:.2e} "
          f"max={e.max():.2e}")
    assert p999 <= p999_bar


@pytest.mark.parametrize("mode", sorted(MODES))
def test_pm_agreement_vs_fp64(mode):
    prec, _, agree_bar, depth_bar = MODES[mode]
    sc = T.scene()
    ids = [0, 1, 2, 3, 4]
    kw = _params(sc)
    g = _pm_gpu(sc, ids, kw, precision=prec)
    o = O.pm_run(T.oracle_cams(sc, ids), [sc.images[i].numpy() for i in ids], O.make_params(**kw),
                 precision=64, rng_view=0)
    agree = T.depth_normal_agreement(g["depth"], g["normal"], o["depth"], o["normal"])
    rel = np.abs(g["depth"] - o["depth"]) / np.maximum(o["depth"], 1e-6)
    within = float((rel <= 0.005).mean())
    print(f"[{mode}] agreement vs FP64 oracle (0.5%, 1 deg) = {agree:.4f}; "
          f"depth within 0.5% = {within:.4f}")
    assert agree >= agree_bar
    assert within >= depth_bar


def test_precisions_are_deterministic():
    sc = T.scene()
    ids = [0, 1, 2, 3, 4]
    kw = _params(sc, iters=2)
    for prec in (1, 2):
        a = _pm_gpu(sc, ids, kw, precision=prec)
        b = _pm_gpu(sc, ids, kw, precision=prec)
        for k in ("depth", "normal", "cost"):
            assert np.array_equal(a[k].view(np.uint32), b[k].view(np.uint32)), (prec, k)
