"""In-memory multi-view pipeline on one GPU: the stages of
gmvs/scripts/launch.py:122-177 without the per-stage file round trips.

    first_stage()  per reference view: plain PatchMatch pass, host planar
                   prior, prior PatchMatch pass (planar_init.py:113-190)
    geom_pass()    geometric-consistency PatchMatch over all views
                   (consistency.py:27-138)
    fuse()         sequential multi-view fusion (consistency.py:141-211)

All views stay resident in HBM. GPU work is issued on the current torch
stream. The planar prior of a view (support points, Delaunay, plane fit) is
built either on the host, on a thread pool as soon as that view's first pass
has finished on the device, overlapping GPU work of other views; or, with
device_prior=True, on the device (gmvs_prior_build_device: no copies to the
host and no host synchronisation; the same priors bit for bit), on a second
stream that overlaps it with the other group's PatchMatch kernels. Launch
submission is serialised on the calling thread, so the workspaces are reused
in stream order.

With batched=True (the default) the views of a stage run as batches
(gmvs_pm_run_batch: one launch per kernel for the whole batch): the plain
passes in `groups` groups, each group's prior pass as soon as its priors are
built (so the host prior construction of one group overlaps the GPU passes
of the other), and the geometric pass as one batch. Results are identical to
batched=False, which runs one view per call as the reference does.
"""
from __future__ import annotations

import math
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass
from typing import List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _native as N
from . import src as S
from .lib.build.planar_initer import build_prior, build_prior_device


@dataclass
class StageOutput:
    depth: torch.Tensor   # [N,H,W]
    normal: torch.Tensor  # [N,H,W,3] world frame
    cost: torch.Tensor    # [N,H,W]


class MultiViewPipeline:
    def __init__(self, images: torch.Tensor, cameras: Sequence[S.Camera],
                 pairs: Sequence[Sequence[int]], labels: Optional[Sequence[np.ndarray]] = None,
                 params: Optional[S.PatchMatchParams] = None, prior_threads: int = 8,
                 precision: int = N.PREC_FP32, batched: bool = True, groups: int = 2,
                 tile_cols: int = 0, device_prior: bool = False):
        self.dev = S._device()
        self.images = images.to(self.dev, torch.float32).contiguous()
        self.n, self.h, self.w = (int(x) for x in self.images.shape)
        self.cameras = list(cameras)
        self.pairs = [list(p) for p in pairs]
        self.labels = labels
        self.params = params if params is not None else S.PatchMatchParams()
        self.prior_threads = prior_threads
        self.device_prior = bool(device_prior)
        self.precision = S.resolve_precision(precision, self.images)
        self.batched = batched
        self.tile_cols = tile_cols
        self.groups = max(1, min(groups, self.n))
        self._ws = S._Workspace()
        self._fuse_ws = S._Workspace()
        for i, p in enumerate(self.pairs):
            if len(p) > 31:
                raise ValueError(f"view {i} has {len(p)} sources (at most 31)")

    def _ids(self, ref: int) -> List[int]:
        return [ref] + self.pairs[ref]

    def _cparams(self, ref: int, iters: int, prior: bool) -> N.CParams:
        p = S.pm_params_for_problem(self.params, self.cameras[ref], len(self._ids(ref)))
        p.max_iterations = iters
        p.planar_prior = prior
        return p

    def _problem(self, ref: int, iters: int, prior: bool, **kw) -> dict:
        ids = self._ids(ref)
        d = dict(cams=[self.cameras[i]._c for i in ids], images=[self.images[i] for i in ids],
                 params=self._cparams(ref, iters, prior), rng_view=ref, precision=self.precision,
                 tile_cols=self.tile_cols)
        d.update(kw)
        return d

    def _build_prior(self, ref: int, iterations: int, superpixel_filter: bool, host_d, host_c):
        lab = self.labels[ref] if (self.labels is not None and superpixel_filter) else None
        cp = self._cparams(ref, iterations, True)
        pl, mk, ntri = build_prior(self.cameras[ref]._c, host_d[ref].numpy(), host_c[ref].numpy(),
                                   lab, cp.depth_min, cp.depth_max, lab is not None)
        return ref, torch.from_numpy(pl).pin_memory(), torch.from_numpy(mk).pin_memory(), ntri

    def _first_stage_batched(self, iterations: int, planar_prior: bool,
                             superpixel_filter: bool) -> StageOutput:
        n, h, w, dev = self.n, self.h, self.w, self.dev
        planes = torch.zeros(n, h, w, 4, device=dev)
        costs = torch.zeros(n, h, w, device=dev)
        out_d = torch.zeros(n, h, w, device=dev)
        out_n = torch.zeros(n, h, w, 3, device=dev)
        out_c = torch.zeros(n, h, w, device=dev)
        host_d = torch.empty(n, h, w, pin_memory=True)
        host_c = torch.empty(n, h, w, pin_memory=True)
        stream = torch.cuda.current_stream()
        groups = [list(g) for g in np.array_split(np.arange(n), self.groups) if len(g)]
        if not hasattr(self, "_ws_groups") or len(self._ws_groups) != len(groups):
            self._ws_groups = [S._Workspace() for _ in groups]
        if planar_prior and self.device_prior:
            return self._first_stage_device_prior(iterations, superpixel_filter, groups, planes,
                                                  costs, out_d, out_n, out_c)
        done = []
        for gi, g in enumerate(groups):
            S.run_pm_batch([self._problem(r, iterations, False, init_mode=N.INIT_RANDOM, geom=False,
                                          planes=planes[r], costs=costs[r], out_depth=out_d[r],
                                          out_normal=out_n[r], out_cost=out_c[r])
                            for r in g], self._ws_groups[gi], dev)
            if planar_prior:
                for r in g:
                    host_d[r].copy_(out_d[r], non_blocking=True)
                    host_c[r].copy_(out_c[r], non_blocking=True)
                ev = torch.cuda.Event()
                ev.record(stream)
                done.append(ev)
        if not planar_prior:
            return StageOutput(out_d, out_n, out_c)

        prior_planes = torch.empty(n, h, w, 4, device=dev)
        plane_masks = torch.empty(n, h, w, dtype=torch.int32, device=dev)
        self.num_triangles = [0] * n

        def build(gi: int, r: int):
            done[gi].synchronize()
            return self._build_prior(r, iterations, superpixel_filter, host_d, host_c)

        with ThreadPoolExecutor(max_workers=max(1, self.prior_threads)) as ex:
            futures = [[ex.submit(build, gi, r) for r in g] for gi, g in enumerate(groups)]
            keep = []
            for gi, g in enumerate(groups):  # each group's prior pass once its priors exist
                for fut in futures[gi]:
                    r, pl, mk, ntri = fut.result()
                    self.num_triangles[r] = ntri
                    prior_planes[r].copy_(pl, non_blocking=True)
                    plane_masks[r].copy_(mk, non_blocking=True)
                    keep.append((pl, mk))
                S.run_pm_batch([self._problem(r, iterations, True, init_mode=N.INIT_PRIOR, geom=False,
                                              planes=planes[r], costs=costs[r],
                                              prior_planes=prior_planes[r],
                                              plane_masks=plane_masks[r], out_depth=out_d[r],
                                              out_normal=out_n[r], out_cost=out_c[r])
                                for r in g], self._ws_groups[gi], dev)
            torch.cuda.current_stream().synchronize()  # pinned buffers in `keep` outlive the copies
        return StageOutput(out_d, out_n, out_c)

    def _first_stage_device_prior(self, iterations: int, superpixel_filter: bool, groups,
                                  planes, costs, out_d, out_n, out_c) -> StageOutput:
        """The batched first stage with the planar priors built on the device:
        each group's priors on a second stream (default priority) as soon as
        its plain pass is done, overlapping the next group's plain pass and
        the previous group's prior pass, which waits for them by event."""
        n, h, w, dev = self.n, self.h, self.w, self.dev
        prior_planes = torch.empty(n, h, w, 4, device=dev)
        plane_masks = torch.empty(n, h, w, dtype=torch.int32, device=dev)
        labels = None
        if self.labels is not None and superpixel_filter:
            labels = [torch.from_numpy(np.ascontiguousarray(self.labels[r], np.int32)).to(dev)
                      for r in range(n)]
        self.num_triangles = [None] * n
        main = torch.cuda.current_stream()
        if getattr(self, "_side", None) is None:
            self._side = torch.cuda.Stream(device=dev)
        side = self._side
        built = []
        for gi, g in enumerate(groups):
            S.run_pm_batch([self._problem(r, iterations, False, init_mode=N.INIT_RANDOM, geom=False,
                                          planes=planes[r], costs=costs[r], out_depth=out_d[r],
                                          out_normal=out_n[r], out_cost=out_c[r])
                            for r in g], self._ws_groups[gi], dev)
            side.wait_stream(main)
            with torch.cuda.stream(side):
                for r in g:
                    cp = self._cparams(r, iterations, True)
                    pl, mk, _ = build_prior_device(self.cameras[r]._c, out_d[r], out_c[r],
                                                   None if labels is None else labels[r],
                                                   cp.depth_min, cp.depth_max, labels is not None,
                                                   count=False)
                    prior_planes[r].copy_(pl)
                    plane_masks[r].copy_(mk)
                ev = torch.cuda.Event()
                ev.record(side)
                built.append(ev)
        for gi, g in enumerate(groups):
            main.wait_event(built[gi])
            S.run_pm_batch([self._problem(r, iterations, True, init_mode=N.INIT_PRIOR, geom=False,
                                          planes=planes[r], costs=costs[r],
                                          prior_planes=prior_planes[r], plane_masks=plane_masks[r],
                                          out_depth=out_d[r], out_normal=out_n[r], out_cost=out_c[r])
                            for r in g], self._ws_groups[gi], dev)
        # the buffers the side stream used are reused by later work on the main stream
        main.wait_stream(side)
        for t in (out_d, out_c, prior_planes, plane_masks):
            t.record_stream(side)
        if labels is not None:
            for t in labels:
                t.record_stream(side)
        return StageOutput(out_d, out_n, out_c)

    def _geom_pass_batched(self, prev: StageOutput, iterations: int) -> StageOutput:
        n, h, w, dev = self.n, self.h, self.w, self.dev
        out_d = torch.zeros(n, h, w, device=dev)
        out_n = torch.zeros(n, h, w, 3, device=dev)
        out_c = torch.zeros(n, h, w, device=dev)
        planes = torch.empty(n, h, w, 4, device=dev)
        costs = torch.empty(n, h, w, device=dev)
        probs = []
        for r in range(n):
            ids = self._ids(r)
            pb = self._problem(r, iterations, False, init_mode=N.INIT_DEPTH_NORMAL, geom=True,
                               planes=planes[r], costs=costs[r],
                               depths=[prev.depth[i] for i in ids], ref_normal=prev.normal[r],
                               out_depth=out_d[r], out_normal=out_n[r], out_cost=out_c[r])
            pb["params"].geom_consistency = True
            probs.append(pb)
        S.run_pm_batch(probs, self._ws, dev)
        return StageOutput(out_d, out_n, out_c)

    def first_stage(self, iterations: int = 8, planar_prior: bool = True,
                    superpixel_filter: bool = True) -> StageOutput:
        if self.batched:
            return self._first_stage_batched(iterations, planar_prior, superpixel_filter)
        n, h, w, dev = self.n, self.h, self.w, self.dev
        planes = torch.zeros(n, h, w, 4, device=dev)
        costs = torch.zeros(n, h, w, device=dev)
        out_d = torch.zeros(n, h, w, device=dev)
        out_n = torch.zeros(n, h, w, 3, device=dev)
        out_c = torch.zeros(n, h, w, device=dev)
        host_d = torch.empty(n, h, w, pin_memory=True)
        host_c = torch.empty(n, h, w, pin_memory=True)
        done = []
        stream = torch.cuda.current_stream()
        for ref in range(n):
            ids = self._ids(ref)
            S.run_pm([self.cameras[i]._c for i in ids], [self.images[i] for i in ids],
                     self._cparams(ref, iterations, False), init_mode=N.INIT_RANDOM, geom=False,
                     rng_view=ref, planes=planes[ref], costs=costs[ref], out_depth=out_d[ref],
                     out_normal=out_n[ref], out_cost=out_c[ref], workspace=self._ws,
                     precision=self.precision, tile_cols=self.tile_cols)
            if planar_prior:
                host_d[ref].copy_(out_d[ref], non_blocking=True)
                host_c[ref].copy_(out_c[ref], non_blocking=True)
                ev = torch.cuda.Event()
                ev.record(stream)
                done.append(ev)
        if not planar_prior:
            return StageOutput(out_d, out_n, out_c)

        prior_planes = torch.empty(n, h, w, 4, device=dev)
        plane_masks = torch.empty(n, h, w, dtype=torch.int32, device=dev)

        def build(ref: int):
            done[ref].synchronize()
            lab = self.labels[ref] if (self.labels is not None and superpixel_filter) else None
            cp = self._cparams(ref, iterations, True)
            pl, mk, ntri = build_prior(self.cameras[ref]._c, host_d[ref].numpy(),
                                       host_c[ref].numpy(), lab, cp.depth_min, cp.depth_max,
                                       lab is not None)
            return ref, torch.from_numpy(pl).pin_memory(), torch.from_numpy(mk).pin_memory(), ntri

        self.num_triangles = [0] * n
        with ThreadPoolExecutor(max_workers=max(1, self.prior_threads)) as ex:
            futures = [ex.submit(build, ref) for ref in range(n)]
            keep = []
            for fut in futures:  # submit second passes in view order
                ref, pl, mk, ntri = fut.result()
                self.num_triangles[ref] = ntri
                prior_planes[ref].copy_(pl, non_blocking=True)
                plane_masks[ref].copy_(mk, non_blocking=True)
                keep.append((pl, mk))
                ids = self._ids(ref)
                S.run_pm([self.cameras[i]._c for i in ids], [self.images[i] for i in ids],
                         self._cparams(ref, iterations, True), init_mode=N.INIT_PRIOR, geom=False,
                         rng_view=ref, planes=planes[ref], costs=costs[ref],
                         prior_planes=prior_planes[ref], plane_masks=plane_masks[ref],
                         out_depth=out_d[ref], out_normal=out_n[ref], out_cost=out_c[ref],
                         workspace=self._ws, precision=self.precision, tile_cols=self.tile_cols)
            torch.cuda.current_stream().synchronize()  # pinned buffers in `keep` outlive the copies
        return StageOutput(out_d, out_n, out_c)

    def geom_pass(self, prev: StageOutput, iterations: int = 2) -> StageOutput:
        """One geometric-consistency stage: every view re-initialised from the
        previous stage and costed with the forward-backward reprojection
        term against its sources' previous depth maps."""
        if self.batched:
            return self._geom_pass_batched(prev, iterations)
        n, h, w, dev = self.n, self.h, self.w, self.dev
        out_d = torch.zeros(n, h, w, device=dev)
        out_n = torch.zeros(n, h, w, 3, device=dev)
        out_c = torch.zeros(n, h, w, device=dev)
        planes = torch.empty(h, w, 4, device=dev)
        costs = torch.empty(h, w, device=dev)
        for ref in range(n):
            ids = self._ids(ref)
            p = self._cparams(ref, iterations, False)
            p.geom_consistency = True
            S.run_pm([self.cameras[i]._c for i in ids], [self.images[i] for i in ids], p,
                     init_mode=N.INIT_DEPTH_NORMAL, geom=True, rng_view=ref, planes=planes,
                     costs=costs, depths=[prev.depth[i] for i in ids], ref_normal=prev.normal[ref],
                     out_depth=out_d[ref], out_normal=out_n[ref], out_cost=out_c[ref],
                     workspace=self._ws, precision=self.precision, tile_cols=self.tile_cols)
        return StageOutput(out_d, out_n, out_c)

    def fuse(self, stage: StageOutput, geom_consistent: int = 3,
             reproj_px: float = S.FUSE_REPROJ_PX, rel_depth: float = S.FUSE_REL_DEPTH,
             angle_rad: float = S.FUSE_ANGLE_RAD, pairs: Optional[Sequence[Sequence[int]]] = None
             ) -> Tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
        """Sequential fusion of every view; returns (depth [N,H,W],
        normal [N,H,W,3], masks [N,H,W] bool)."""
        n, h, w, dev = self.n, self.h, self.w, self.dev
        pairs = self.pairs if pairs is None else pairs
        masks = torch.zeros(n, h, w, dtype=torch.bool, device=dev)
        out_d = torch.empty(n, h, w, device=dev)
        out_n = torch.empty(n, h, w, 3, device=dev)
        for ref in range(n):
            ids = [ref] + list(pairs[ref])
            S.fuse_view(ids, [self.cameras[i]._c for i in ids], [stage.depth[i] for i in ids],
                        [stage.normal[i] for i in ids], masks, geom_consistent, out_d[ref],
                        out_n[ref], reproj_px, rel_depth, angle_rad, self._fuse_ws)
        return out_d, out_n, masks


def fuse_maps(depths: torch.Tensor, normals: torch.Tensor, cameras: Sequence[S.Camera],
              pairs: Sequence[Sequence[int]], geom_consistent: int = 3,
              reproj_px: float = S.FUSE_REPROJ_PX, rel_depth: float = S.FUSE_REL_DEPTH,
              angle_rad: float = S.FUSE_ANGLE_RAD, workspace: Optional[S._Workspace] = None,
              precision: int = 32) -> Tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
    """Fusion of device-resident depth/normal maps of all views (in view order);
    precision 64 runs the reprojection check in FP64 (gmvs_fuse_view_fp64)."""
    n, h, w = (int(x) for x in depths.shape)
    dev = depths.device
    ws = workspace or S._Workspace()
    masks = torch.zeros(n, h, w, dtype=torch.bool, device=dev)
    out_d = torch.empty(n, h, w, device=dev)
    out_n = torch.empty(n, h, w, 3, device=dev)
    for ref in range(n):
        ids = [ref] + list(pairs[ref])
        S.fuse_view(ids, [cameras[i]._c for i in ids], [depths[i] for i in ids],
                    [normals[i] for i in ids], masks, geom_consistent, out_d[ref], out_n[ref],
                    reproj_px, rel_depth, angle_rad, ws, precision=precision)
    return out_d, out_n, masks


def cameras_from_synthetic(cams) -> List[S.Camera]:
    return [S.load_camera(c.K.reshape(-1), c.R.reshape(-1), c.t, c.height, c.width, c.depth_min,
                          c.depth_max) for c in cams]


def pixel_hypotheses(height: int, width: int, iterations: int) -> int:
    """Pixel-hypothesis count of one pass: 1 at init plus 14 per iteration
    (8 propagated candidates, the current hypothesis, 5 refinements)."""
    return height * width * (1 + 14 * iterations)
