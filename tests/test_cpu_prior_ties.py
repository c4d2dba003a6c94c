"""Why the planar-prior pass cannot reach 99 % depth+normal agreement with the
FP64 oracle in FP32, and a check that every FP32 decision that parts from
FP64 is a tie below FP32's resolution.

The oracle's FP32 and FP64 instantiations run in lockstep: before every
half-sweep the FP64 state is copied into the FP32 run and both update every
pixel of the colour from that same state, so every recorded divergence is a
root (no earlier difference feeds it). Under the prior, a masked pixel keeps
the candidate with the highest restricted score
exp(-c^2 / 0.18) * (0.5 + exp(-dd^2 / 2 sd^2) exp(-a^2 / 2 sa^2))
(lib 1268-1305) and compares it with the current plane's score; with final
costs c around 1e-3 the score sits on its flat top near 1.5, where
candidates differ by ~1e-10: far below one FP32 ulp (1.2e-7) of the score,
and below the FP32 error of the costs themselves (~1e-7). What decides
these updates in FP32 is rounding, in the reference's FP32 kernels as in
ours; the recorded margins show that no divergence is anything else.
"""
import numpy as np

from oracle import oracle as O
from tests import helpers as T

BETA = 0.18


def _prior_setup():
    from gmvs_amd import src as S
    from gmvs_amd.lib.build import planar_initer as PI

    sc = T.scene()
    ids = [0, 1, 2, 3, 4]
    cams = T.oracle_cams(sc, ids)
    imgs = [sc.images[i].numpy() for i in ids]
    dmin, dmax = T.depth_range(sc, 0)
    kw = dict(max_iterations=3, depth_min=dmin, depth_max=dmax)
    o1 = O.pm_run(cams, imgs, O.make_params(**kw), precision=64, rng_view=0)
    c = sc.cameras[0]
    ccam = S.load_camera(c.K.reshape(-1), c.R.reshape(-1), c.t, c.height, c.width, c.depth_min,
                         c.depth_max)._c
    pp, masks, _ = PI.build_prior(ccam, o1["depth"], o1["cost"], None, dmin, dmax)
    args = dict(init_mode=2, prior_planes=pp, plane_masks=masks.astype(np.uint32),
                planes=o1["planes"], costs=o1["cost"], rng_view=0)
    return sc, cams, imgs, O.make_params(**dict(kw, planar_prior=True)), args


def _argext(v, masked):
    # FindMaxCostIndex / FindMinCostIndex: the last index of the extremum
    r = v[::-1]
    return len(v) - 1 - int(np.argmax(r) if masked else np.argmin(r))


def _budget(val, cost):
    """FP32 error budget of a restricted score `val` of final cost `cost`:
    4 ulps of the score, the score's change for a cost error of 1e-5 (the
    FP32 NCC error; the cost tolerance is 1e-4) and a relative 2e-5 for the
    angle term (acos of a dot product within ~1e-7 of 1 is good to ~3e-4
    rad in FP32, which moves exp(-a^2 / 2 sa^2) by up to ~1e-5)."""
    return (4 * np.spacing(np.float32(abs(val))) + abs(val) * 2 * abs(cost) / BETA * 1e-5
            + abs(val) * 2e-5)


def test_prior_pass_divergences_are_fp32_ties():
    sc, cams, imgs, params, args = _prior_setup()
    L = O.pm_lockstep(cams, imgs, params, **args)
    ids, mg, V = L["ids"], L["margins"], L["values"]
    n = L["decisions"]
    what = ids[:, 3]
    masked = ids[:, 4] == 1
    counts = {k: int((what == b).sum()) for k, b in (("draw", 1), ("argext", 2), ("adopt", 4),
                                                       ("refine", 8), ("ncc_jump", 16))}
    print(f"pixel updates {n}; divergent {len(ids)} ({len(ids) / n:.2%}); {counts}; "
          f"masked {int(masked.sum())}")
    # drawn views differ only where a draw lies within FP32 rounding of a CDF
    # step or a table cost within FP32 rounding of tau or 1.2 (margin column
    # 0), which happens a handful of times in 230,400 updates
    assert counts["draw"] <= 3 and counts["ncc_jump"] <= 3
    if counts["draw"]:
        assert mg[what == 1, 0].max() < 1e-5, mg[what == 1, 0]
    # divergences are a property of the prior's masked branch
    assert (masked | (what == 8)).mean() > 0.995
    ulps = []
    for r in np.flatnonzero((what == 2) | (what == 4)):
        v64, f64 = V[r, 0:8], V[r, 16:24]
        k64 = _argext(v64, masked[r])
        if what[r] == 2:
            other = v64[int(V[r, 34])]
        else:
            other = V[r, 32]  # the current plane's value
        m = abs(v64[k64] - other)
        ulps.append(m / np.spacing(np.float32(abs(v64[k64]))))
        if masked[r]:
            assert m <= _budget(v64[k64], f64[k64]), (ids[r], m, v64, f64)
        else:
            assert m <= 1e-5, (ids[r], m)  # final costs: below the FP32 NCC error
    ulps = np.array(ulps)
    print(f"argmax / adoption margins in FP32 ulps of the score: median {np.median(ulps):.3f}, "
          f"90% {np.quantile(ulps, 0.9):.3f}, 99% {np.quantile(ulps, 0.99):.2f}, "
          f"max {ulps.max():.1f}")
    assert np.quantile(ulps, 0.9) < 1.0  # most are below one ulp: unresolvable in FP32
    ref = mg[what == 8, 3]
    assert ref.max() < 1e-4, ref.max()  # refinement: cost or score within the cost tolerance


def test_plain_pass_divergences_are_rare_and_tiny():
    sc = T.scene()
    ids = [0, 1, 2, 3, 4]
    dmin, dmax = T.depth_range(sc, 0)
    L = O.pm_lockstep(T.oracle_cams(sc, ids), [sc.images[i].numpy() for i in ids],
                      O.make_params(max_iterations=4, depth_min=dmin, depth_max=dmax), rng_view=0)
    ids_, mg = L["ids"], L["margins"]
    frac = len(ids_) / L["decisions"]
    print(f"plain: divergent updates {len(ids_)} of {L['decisions']} ({frac:.2e})")
    assert frac < 1e-3
    what = ids_[:, 3]
    col = {2: 1, 4: 2, 8: 3}
    for b, c in col.items():
        if (what == b).any():
            assert mg[what == b, c].max() < 1e-5
    # a drawn-view flip needs a draw within FP32 rounding of a CDF step or a
    # table cost within FP32 rounding of tau or 1.2; an NCC that lands on its
    # cost-2 branch in one precision only (what = 16) is the discontinuity
    # the cost tests exempt
    if (what == 1).any():
        assert mg[what == 1, 0].max() < 1e-5, mg[what == 1, 0]
    assert (what == 16).sum() <= 3


def test_prior_pass_fp32_depth_within_half_percent():
    """The FP32 port's prior pass, free-running: depth within 0.5 % of FP64
    on >= 99.9 % of pixels; what fails the 1-degree normal test are normals
    that drift across the flat top of the score (median ~1.5 degrees)."""
    sc, cams, imgs, params, args = _prior_setup()
    o = O.pm_run(cams, imgs, params, precision=64, **args)
    o32 = O.pm_run(cams, imgs, params, precision=32, **args)
    rel = np.abs(o32["depth"] - o["depth"]) / np.abs(o["depth"])
    ang = T.angle_deg(o32["normal"], o["normal"])
    agree = T.depth_normal_agreement(o32["depth"], o32["normal"], o["depth"], o["normal"])
    print(f"FP32 port vs FP64: depth within 0.5% {np.mean(rel <= 0.005):.4f}; normal within 1 deg "
          f"{np.mean(ang <= 1):.4f}; both {agree:.4f}")
    assert np.mean(rel <= 0.005) >= 0.999
    assert agree >= 0.97
