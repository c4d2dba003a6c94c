"""View selection and candidate adoption as their own launch (pm_select) or
as the epilogue of each pm_candidates tile (gmvs_set_select_fusion): the two
schedules run the same per-pixel code and must give the same bits for the
plain, planar-prior and geometric passes, in FP32 and BF16 cells."""
import numpy as np
import pytest

from oracle import oracle as O
from tests import helpers as T
from tests.test_gpu_pm import _ccams, _params, _pm_gpu

pytestmark = pytest.mark.gpu


def _both(fn):
    from gmvs_amd import _native as N

    try:
        N.set_select_fusion(False)
        a = fn()
        N.set_select_fusion(True)
        b = fn()
    finally:
        N.set_select_fusion(False)
    return a, b


def _same(a, b):
    for k in ("depth", "normal", "cost"):
        assert np.array_equal(a[k].view(np.uint32), b[k].view(np.uint32)), k


@pytest.mark.parametrize("precision", [0, 2])
def test_plain_pass_fused_select_equals_separate(precision):
    sc = T.scene()
    ids = [0, 1, 2, 3, 4]
    kw = _params(sc, iters=3)
    _same(*_both(lambda: _pm_gpu(sc, ids, kw, precision=precision)))


def test_prior_pass_fused_select_equals_separate():
    from gmvs_amd.lib.build import planar_initer as PI

    sc = T.scene()
    ids = [0, 1, 2, 3, 4]
    kw = _params(sc, iters=2)
    o1 = O.pm_run(T.oracle_cams(sc, ids), [sc.images[i].numpy() for i in ids], O.make_params(**kw),
                  precision=32, rng_view=0)
    pp, masks, ntri = PI.build_prior(_ccams(sc, [0])[0], o1["depth"], o1["cost"], None,
                                     kw["depth_min"], kw["depth_max"])
    assert ntri > 10
    kw2 = dict(kw, planar_prior=True)
    state = (o1["planes"].copy(), o1["cost"].copy())
    _same(*_both(lambda: _pm_gpu(sc, ids, kw2, init_mode=2, prior=(pp, masks), state=state)))


def test_geom_pass_fused_select_equals_separate():
    sc = T.scene()
    ids = [0, 1, 2, 3, 4]
    kw = _params(sc, iters=2, geom_consistency=True)
    rng = np.random.default_rng(7)
    depths = [(sc.depth[i].numpy() * (1 + 0.01 * rng.standard_normal(sc.depth[i].shape)))
              .astype(np.float32) for i in ids]
    normal = sc.normal[0].numpy()
    _same(*_both(lambda: _pm_gpu(sc, ids, kw, init_mode=1, geom=True, depths=depths,
                                 ref_normal=normal)))
