"""Autotuner of the PatchMatch launch shape.

Three knobs change how the work is laid out on the GPU:

* the tile shape of the 64-pixel work tiles, per problem (`tile_cols` of
  gmvs_pm_args: 4, 8, 16 or 32 columns);
* the blocks per CU of each persistent kernel (gmvs_set_blocks_per_cu: 0 =
  as many as fit);
* the lanes per pixel of the NCC lane groups of the candidate and
  refinement kernels (gmvs_set_lane_width: 2 or 4; 0 = the library's
  default).

Blocks per CU never change results. The tile shape and the lane width do at
rounding level: tap and source positions are formed relative to the tile's
origin and anchor (the local frame), and the lane width sets the order of the
FP32 NCC sums. The lane width is not
searched by default (the library's defaults, 2 lanes for the candidate
kernel and 4 for refinement, were measured on full passes); pass
`lanes=(0, 2, 4)` to search it anyway. The blocks-per-CU
candidates are given as waves per SIMD and converted with the kernel's
block size (tiles x 64 pixels x lanes; the candidate kernel holds two tiles
per block).

`autotune(problems)` runs one short plain pass of the batch from random
planes (3 iterations), then times, for each setting, plain passes that start
from its depth and normal maps (INIT_DEPTH_NORMAL, `iterations` iterations)
on scratch state, and keeps the fastest: iterations from a coherent state
gather like the converged iterations that dominate a full pass, which
iterations from random planes do not. Results are cached per batch shape for the process, so a
pipeline tunes once.
"""
from __future__ import annotations

from typing import Dict, Optional, Sequence, Tuple

import torch

from . import _native as N
from . import src as S

TILE_COLS = (4, 8, 16, 32)
WAVES_PER_SIMD = (0, 3, 2)  # 0 = as many as fit
LANES = (0,)
# the library's default lane widths (gmvs_set_lane_width(k, 0))
DEFAULT_LANES = {N.TUNE_CANDIDATES: 2, N.TUNE_REFINE: 4}

_cache: Dict[Tuple, dict] = {}


def _key(problems: Sequence[dict]) -> Tuple:
    return tuple((tuple(p["images"][0].shape), len(p["images"]), p.get("precision", N.PREC_FP32))
                 for p in problems)


def _plain(p: dict, iterations: int, tile_cols: int, **kw) -> dict:
    """A plain pass (no prior, no geometric term) of problem p on scratch state."""
    cp = N.CParams.from_buffer_copy(p["params"])
    cp.max_iterations = iterations
    cp.planar_prior = False
    cp.geom_consistency = False
    d = dict(cams=p["cams"], images=p["images"], params=cp, init_mode=N.INIT_RANDOM, geom=False,
             rng_view=p.get("rng_view", 0), planes=torch.empty_like(p["planes"]),
             costs=torch.empty_like(p["costs"]), precision=p.get("precision", N.PREC_FP32),
             tile_cols=tile_cols)
    d.update(kw)
    return d


def _warm_state(problems: Sequence[dict], iterations: int = 3) -> list:
    """(depth, normal) of every problem after a short plain pass from random
    planes. The trials start from these maps, so they time iterations whose
    hypotheses gather from coherent footprints, as the converged iterations
    of a full pass do, not the scattered ones of a random start."""
    out, batch = [], []
    for p in problems:
        h, w = p["images"][0].shape
        dev = p["images"][0].device
        d = torch.empty(h, w, device=dev)
        n = torch.empty(h, w, 3, device=dev)
        batch.append(_plain(p, iterations, 0, out_depth=d, out_normal=n))
        out.append((d, n))
    S.run_pm_batch(batch, S._Workspace())
    return out


def _trial(problems: Sequence[dict], tile_cols: int, iterations: int, state: list) -> list:
    return [_plain(p, iterations, tile_cols, init_mode=N.INIT_DEPTH_NORMAL,
                   depths=[d] + [None] * (len(p["images"]) - 1), ref_normal=n)
            for p, (d, n) in zip(problems, state)]


def _time(problems: Sequence[dict], ws: S._Workspace, kernel: Optional[str] = None,
          repeats: int = 2) -> float:
    """Best of `repeats` timed runs after a warm run: the whole batch (HIP
    events) or one kernel's accumulated time (the library's profiler)."""
    S.run_pm_batch(problems, ws)  # warm (and sizes the workspace)
    torch.cuda.synchronize()
    best = float("inf")
    for _ in range(repeats):
        if kernel is None:
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            S.run_pm_batch(problems, ws)
            b.record()
            torch.cuda.synchronize()
            best = min(best, a.elapsed_time(b))
        else:
            N.profile_reset()
            N.profile_enable(True)
            S.run_pm_batch(problems, ws)
            torch.cuda.synchronize()
            N.profile_enable(False)
            best = min(best, N.profile_collect()[kernel]["ms"])
    return best


def _pick(times: Dict[int, float], default: int, margin: float) -> int:
    """The fastest setting, unless it beats the default by less than
    `margin` (relative): short timings are noisy, and the default is the one
    measured on full workloads."""
    best = min(times, key=times.get)
    return best if times[best] < times[default] * (1.0 - margin) else default


def _tiles_per_block(tune_id: int) -> int:
    """Tiles per block: the candidate kernel takes pairs of tiles unless the
    view selection runs as its tile epilogue."""
    return 2 if tune_id == N.TUNE_CANDIDATES and not N.select_fusion_enabled() else 1


def _blocks(waves_per_simd: int, waves_per_block: int) -> int:
    """Blocks per CU for a waves-per-SIMD target (4 SIMDs per CU, blocks of
    tiles x 64 pixels x lanes threads = tiles x lanes waves); 0 stays 0."""
    return 0 if waves_per_simd == 0 else max(1, 4 * waves_per_simd // waves_per_block)


def autotune(problems: Sequence[dict], iterations: int = 2, tile_cols: Sequence[int] = TILE_COLS,
             waves: Sequence[int] = WAVES_PER_SIMD, lanes: Sequence[int] = LANES,
             apply: bool = True, margin: float = 0.03) -> dict:
    """Best {'tile_cols', 'blocks_candidates', 'blocks_refine',
    'lanes_candidates', 'lanes_refine'} for a batch of make_pm_args keyword
    dicts, keeping the defaults (8 x 8 tiles, full occupancy, the library's
    lane widths) unless another setting is faster by more than `margin`; with
    apply=True the blocks per CU and lane widths are set process-wide. The
    caller's state buffers are not touched."""
    key = _key(problems)
    if key not in _cache:
        ws = S._Workspace()
        state = _warm_state(problems)
        tt = {tc: _time(_trial(problems, tc, iterations, state), ws) for tc in tile_cols}
        best_tc = _pick(tt, 8 if 8 in tt else tile_cols[0], margin)
        trial = _trial(problems, best_tc, iterations, state)
        choice = {"tile_cols": best_tc}
        for name, tune_id, kname in (("blocks_candidates", N.TUNE_CANDIDATES, "pm_candidates"),
                                     ("blocks_refine", N.TUNE_REFINE, "pm_refine")):
            times = {}
            for wv in waves:
                b = _blocks(wv, DEFAULT_LANES[tune_id] * _tiles_per_block(tune_id))
                N.check("gmvs_set_blocks_per_cu", N.lib.gmvs_set_blocks_per_cu(tune_id, b))
                times[b] = _time(trial, ws, kname)
            choice[name] = _pick(times, 0 if 0 in times else min(times), margin)
            N.check("gmvs_set_blocks_per_cu", N.lib.gmvs_set_blocks_per_cu(tune_id, 0))
        for name, tune_id, kname in (("lanes_candidates", N.TUNE_CANDIDATES, "pm_candidates"),
                                     ("lanes_refine", N.TUNE_REFINE, "pm_refine")):
            times = {}
            for ln in lanes:
                N.set_lane_width(tune_id, ln)
                times[ln] = _time(trial, ws, kname) if len(lanes) > 1 else 0.0
            choice[name] = _pick(times, 0 if 0 in times else lanes[0], margin)
            N.set_lane_width(tune_id, 0)
        _cache[key] = choice
    choice = _cache[key]
    if apply:
        N.check("gmvs_set_blocks_per_cu",
                N.lib.gmvs_set_blocks_per_cu(N.TUNE_CANDIDATES, choice["blocks_candidates"]))
        N.check("gmvs_set_blocks_per_cu",
                N.lib.gmvs_set_blocks_per_cu(N.TUNE_REFINE, choice["blocks_refine"]))
        N.set_lane_width(N.TUNE_CANDIDATES, choice.get("lanes_candidates", 0))
        N.set_lane_width(N.TUNE_REFINE, choice.get("lanes_refine", 0))
    return dict(choice)


def reset() -> None:
    """Back to the library defaults for every process-wide knob autotune sets."""
    for tune_id in (N.TUNE_CANDIDATES, N.TUNE_REFINE):
        N.check("gmvs_set_blocks_per_cu", N.lib.gmvs_set_blocks_per_cu(tune_id, 0))
        N.set_lane_width(tune_id, 0)
