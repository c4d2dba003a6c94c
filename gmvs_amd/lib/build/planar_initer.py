"""Drop-in for the reference's ``gmvs.lib.build.planar_initer`` module
(gmvs/lib/src/bindings.cpp:12-67): ``PlanarInit``, ``PatchMatchParams``,
``Problem``, ``Camera``, ``load_camera``, ``load_problem``.

PlanarInit runs the first stage on one reference view: a plain PatchMatch
pass, then (after ``triangulation`` + ``cuda_planar_prior_init``) the
planar-prior pass. Images live in plain device buffers (no textures); the
hypothesis state stays on the device between the two passes.
``fusion_planar_mask``, ``filter_by_var_map`` and ``sliding_window_normal``
run on the GPU (gmvs_amd.post).
"""
from __future__ import annotations

import ctypes
from typing import List, Optional, Sequence, Tuple

import numpy as np
import torch

from ... import _native as N
from ... import io as gio
from ... import post as _post
from ...src import (Camera, PatchMatchParams, Problem, _Workspace, _device, load_camera,
                    load_problem, resolve_precision, run_pm)

__all__ = ["PlanarInit", "PatchMatchParams", "Problem", "Camera", "load_camera", "load_problem",
           "fusion_planar_mask", "filter_by_var_map", "sliding_window_normal", "build_prior",
           "build_prior_device"]


def build_prior(ref_cam: N.CCamera, depth: np.ndarray, cost: np.ndarray,
                labels: Optional[np.ndarray], depth_min: float, depth_max: float,
                superpixel_filter: bool = False) -> Tuple[np.ndarray, np.ndarray, int]:
    """Host planar-prior construction (gmvs_prior_build). Returns
    (prior_planes [H,W,4] f32, plane_masks [H,W] int32, #triangles)."""
    h, w = ref_cam.height, ref_cam.width
    depth = np.ascontiguousarray(depth, np.float32).reshape(h, w)
    cost = np.ascontiguousarray(cost, np.float32).reshape(h, w)
    lab = None
    if labels is not None:
        lab = np.ascontiguousarray(labels, np.int32).reshape(h, w)
    p = PatchMatchParams()
    p.depth_min = depth_min
    p.depth_max = depth_max
    p.superpixel_filter = superpixel_filter
    cp = p.to_c()
    planes = np.zeros((h, w, 4), np.float32)
    masks = np.zeros((h, w), np.uint32)
    ntri = ctypes.c_int32(0)
    N.check("gmvs_prior_build", N.lib.gmvs_prior_build(
        ctypes.byref(ref_cam), depth.ctypes.data, cost.ctypes.data,
        lab.ctypes.data if lab is not None else None, ctypes.byref(cp), planes.ctypes.data,
        masks.ctypes.data, ctypes.byref(ntri)))
    return planes, masks.astype(np.int32), int(ntri.value)


def build_prior_device(ref_cam: N.CCamera, depth, cost, labels, depth_min: float,
                       depth_max: float, superpixel_filter: bool = False, count: bool = True):
    """Device planar-prior construction (gmvs_prior_build_device; equal to
    build_prior bit for bit). depth, cost: [H,W] float32 device tensors;
    labels: [H,W] int32 device tensor or None. Returns (prior_planes [H,W,4]
    f32, plane_masks [H,W] int32, #triangles), device tensors. Every step,
    the Delaunay triangulation included, runs on the current stream; with
    count=True the call waits for it to report the number of retained
    triangles, with count=False it only enqueues the work (#triangles is
    None)."""
    import torch

    h, w = ref_cam.height, ref_cam.width
    dev = depth.device
    depth = depth.to(torch.float32).contiguous().reshape(h, w)
    cost = cost.to(dev, torch.float32).contiguous().reshape(h, w)
    lab = None if labels is None else labels.to(dev, torch.int32).contiguous().reshape(h, w)
    p = PatchMatchParams()
    p.depth_min = depth_min
    p.depth_max = depth_max
    p.superpixel_filter = superpixel_filter
    cp = p.to_c()
    planes = torch.empty(h, w, 4, dtype=torch.float32, device=dev)
    masks = torch.empty(h, w, dtype=torch.int32, device=dev)
    nbytes = N.lib.gmvs_prior_device_workspace_bytes(h, w)
    ws = torch.empty(nbytes, dtype=torch.uint8, device=dev)
    ntri = ctypes.c_int32(0)
    stream = torch.cuda.current_stream(dev).cuda_stream
    N.check("gmvs_prior_build_device", N.lib.gmvs_prior_build_device(
        ctypes.byref(ref_cam), depth.data_ptr(), cost.data_ptr(),
        lab.data_ptr() if lab is not None else None, ctypes.byref(cp), planes.data_ptr(),
        masks.data_ptr(), ctypes.byref(ntri) if count else None, ws.data_ptr(), nbytes, stream))
    return planes, masks, (int(ntri.value) if count else None)


def delaunay(xy: np.ndarray) -> np.ndarray:
    """Delaunay triangulation of integer points (gmvs_delaunay; the
    triangulation step of planar_initer.cu:347-361). xy: [n,2] ints in
    [0, 65536), distinct. Returns [T,3] int32 vertex indices, each triangle
    with (b - a) x (c - a) > 0 in pixel coordinates (x right, y down)."""
    pts = np.ascontiguousarray(xy, np.int32).reshape(-1, 2)
    ntri = ctypes.c_int32(0)
    N.check("gmvs_delaunay", N.lib.gmvs_delaunay(pts.ctypes.data, len(pts), None, 0,
                                                 ctypes.byref(ntri)))
    tri = np.zeros((max(1, ntri.value), 3), np.int32)
    N.check("gmvs_delaunay", N.lib.gmvs_delaunay(pts.ctypes.data, len(pts), tri.ctypes.data,
                                                 ntri.value, ctypes.byref(ntri)))
    return tri[:ntri.value]


def delaunay_cells(cell_xy: np.ndarray, device=None) -> np.ndarray:
    """The triangulation of points on a grid of 5x5 pixel cells, at most one
    per cell (gmvs_delaunay_cells_device, or _host with device=None: the
    same construction on the host). cell_xy: [ncx, ncy, 2] int32, (x, y)
    inside the cell or (-1, -1). Returns [T,3] int32 indices of the points
    numbered in cell order (column-major), listed as delaunay() lists the
    same points."""
    c = np.ascontiguousarray(cell_xy, np.int32)
    ncx, ncy = int(c.shape[0]), int(c.shape[1])
    ntri = ctypes.c_int32(0)
    max_tri = 2 * ncx * ncy + 8
    if device is None:
        tri = np.zeros((max_tri, 3), np.int32)
        N.check("gmvs_delaunay_cells_host", N.lib.gmvs_delaunay_cells_host(
            c.ctypes.data, ncx, ncy, tri.ctypes.data, max_tri, ctypes.byref(ntri)))
        return tri[:ntri.value]
    import torch

    dc = torch.from_numpy(c).to(device)
    out = torch.zeros(max_tri, 3, dtype=torch.int32, device=device)
    nbytes = N.lib.gmvs_prior_device_workspace_bytes(5 * ncy, 5 * ncx)
    ws = torch.empty(nbytes, dtype=torch.uint8, device=device)
    stream = torch.cuda.current_stream(device).cuda_stream
    N.check("gmvs_delaunay_cells_device", N.lib.gmvs_delaunay_cells_device(
        dc.data_ptr(), ncx, ncy, out.data_ptr(), max_tri, ctypes.byref(ntri), ws.data_ptr(), nbytes,
        stream))
    return out[:ntri.value].cpu().numpy()


class PlanarInit:
    """PlanarInit (gmvs/lib/include/planar_initer.cuh:65-133)."""

    def __init__(self, verbose: bool = False):
        self.verbose = bool(verbose)
        self.params = PatchMatchParams()
        # GMVS_PREC_* or a name ("fp32", "fp16", "bf16", "auto"); not in the
        # reference, which computes in FP32 (the default)
        self.precision = N.PREC_FP32
        self._images: List[np.ndarray] = []
        self._cameras: List[Camera] = []
        self._superpixel: Optional[np.ndarray] = None
        self._dev_images: Optional[List[torch.Tensor]] = None
        self._planes: Optional[torch.Tensor] = None
        self._costs: Optional[torch.Tensor] = None
        self._prior: Optional[Tuple[torch.Tensor, torch.Tensor]] = None
        self._built_prior: Optional[Tuple[torch.Tensor, torch.Tensor]] = None  # built prior (device)
        self._ws = _Workspace()
        # Philox key word 1 (the global reference view id; set by the driver)
        self.rng_view = 0

    # ---- inputs (planar_initer.cu:64-102)
    def add_image(self, image_path: str) -> None:
        self._images.append(gio.read_gray(image_path))

    def add_image_array(self, image: np.ndarray) -> None:
        """Adds an already decoded grey image (float32 [H,W])."""
        self._images.append(np.ascontiguousarray(image, np.float32))

    def add_superpixel(self, superpixel: np.ndarray) -> None:
        self._superpixel = np.ascontiguousarray(superpixel).astype(np.int32)

    def get_image(self, index: int) -> np.ndarray:
        return self._images[index]

    def add_camera(self, K, R, t, height: int, width: int, depth_min: float,
                   depth_max: float) -> None:
        self._cameras.append(load_camera(K, R, t, height, width, depth_min, depth_max))

    def get_camera(self, index: int) -> Camera:
        return self._cameras[index]

    def get_num_images(self) -> int:
        return len(self._images)

    # ---- device state (planar_initer.cu:119-223)
    def cuda_space_init(self) -> None:
        if not self._images or len(self._cameras) < len(self._images):
            raise RuntimeError("add images and cameras before cuda_space_init")
        dev = _device()
        self._dev_images = [torch.from_numpy(im).to(dev).contiguous() for im in self._images]
        h, w = self._images[0].shape
        self._planes = torch.zeros(h, w, 4, dtype=torch.float32, device=dev)
        self._costs = torch.zeros(h, w, dtype=torch.float32, device=dev)

    def cuda_planar_prior_init(self) -> None:
        if self._built_prior is None:
            raise RuntimeError("triangulation must run before cuda_planar_prior_init")
        # the prior was built on the device by triangulation (already resident)
        self._prior = self._built_prior

    # ---- PatchMatch (lib 1711-1810)
    def run_patch_match(self) -> None:
        if self._dev_images is None:
            raise RuntimeError("cuda_space_init must run before run_patch_match")
        n = len(self._dev_images)
        cp = self.params.to_c()
        cp.num_images = n
        prior = bool(self.params.planar_prior)
        if prior and self._prior is None:
            raise RuntimeError("planar_prior is set but cuda_planar_prior_init has not run")
        if self.params.geom_consistency:
            raise RuntimeError("the planar initialiser has no depth inputs for the geometric "
                               "term; use gmvs_amd.src.PatchMatcher")
        run_pm([c._c for c in self._cameras[:n]], self._dev_images, cp,
               init_mode=N.INIT_PRIOR if prior else N.INIT_RANDOM, geom=False, rng_view=self.rng_view,
               planes=self._planes, costs=self._costs,
               prior_planes=self._prior[0] if prior else None,
               plane_masks=self._prior[1] if prior else None, workspace=self._ws,
               precision=resolve_precision(self.precision, self._dev_images))
        torch.cuda.synchronize()
        if self.verbose:
            for it in range(self.params.max_iterations):
                print(f"iteration: {it}")

    def jump_patch_match(self) -> None:
        torch.cuda.synchronize()

    # ---- outputs
    def export_plane(self, depth: np.ndarray, normal: np.ndarray, cost: np.ndarray) -> None:
        """Writes into the caller's arrays in place (planar_initer.cu:104-117)."""
        planes = self._planes.cpu().numpy()
        costs = self._costs.cpu().numpy()
        depth[...] = planes[..., 3]
        normal[...] = planes[..., :3]
        cost[...] = costs

    def get_cost(self, index: int) -> float:
        return float(self._costs.reshape(-1)[index].item())

    def triangulation(self, depth: np.ndarray, triangulation_path: str = " ") -> None:
        """planar_initer.cu:347-420: support points from the last pass's
        costs, Delaunay, superpixel filter, plane fit and rasterisation."""
        cam = self._cameras[0]
        dev = self._costs.device
        h, w = cam.height, cam.width
        d = torch.from_numpy(np.ascontiguousarray(depth, np.float32).reshape(h, w)).to(dev)
        lab = None
        if self._superpixel is not None:
            lab = torch.from_numpy(np.ascontiguousarray(self._superpixel, np.int32)).to(dev)
        # on the device (gmvs_prior_build_device, equal to the host build bit
        # for bit): the cost map stays resident, the Delaunay step included
        planes, masks, ntri = build_prior_device(cam._c, d, self._costs, lab,
                                                 self.params.depth_min, self.params.depth_max,
                                                 bool(self.params.superpixel_filter)
                                                 and self._superpixel is not None)
        self._built_prior = (planes, masks)
        self.num_triangles = ntri
        if triangulation_path.endswith(".png"):
            if self.verbose:
                print(f"save triangulation as {triangulation_path}")
            self._save_prior_png(triangulation_path, masks.cpu().numpy())

    def _save_prior_png(self, path: str, masks: np.ndarray) -> None:
        img = self._images[0]
        edge = np.zeros_like(masks, dtype=bool)
        edge[:, 1:] |= masks[:, 1:] != masks[:, :-1]
        edge[1:, :] |= masks[1:, :] != masks[:-1, :]
        rgb = np.repeat(np.clip(img, 0, 255).astype(np.uint8)[..., None], 3, axis=-1)
        rgb[edge] = (255, 0, 0)
        from PIL import Image

        Image.fromarray(rgb).save(path)


# The segment / normal utilities of the module (planar_initer.cu:471-684),
# on the GPU; numpy in, numpy out, as the reference's cv::Mat bindings.
fusion_planar_mask = _post.fusion_planar_mask
filter_by_var_map = _post.filter_by_var_map
sliding_window_normal = _post.sliding_window_normal
