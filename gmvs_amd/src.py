"""Drop-in for the reference's ``gmvs.src`` extension module
(gmvs/src/bindings.cpp:8-55): ``load_camera``, ``load_problem``,
``PatchMatcher``, ``Fuser``, ``PatchMatchParams``, ``Problem``, ``Camera``.

Same names, argument meaning and error behaviour (``RuntimeError`` for a
tensor on the wrong device, as TORCH_CHECK raises); the arithmetic runs in
the HIP kernels of libgmvs_amd.so. Inputs handed over as CPU tensors are
uploaded once and stay resident in HBM across calls.
"""
from __future__ import annotations

import ctypes
import math
from typing import List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _native as N

RANDOM_SEED = 3407  # gmvs/src/patch_match.cu:8


class PatchMatchParams:
    """mvs::PatchMatchParams (gmvs/include/data_structure.hpp:20-37) with the
    planar initialiser's extra ``superpixel_filter`` flag
    (gmvs/lib/include/planar_initer.hpp:13-31)."""

    _FIELDS = (
        ("max_iterations", int, 3), ("patch_size", int, 11), ("num_images", int, 5),
        ("max_image_size", int, 3200), ("radius_increment", int, 2),
        ("sigma_spatial", float, 5.0), ("sigma_color", float, 3.0), ("top_k", int, 4),
        ("baseline", float, 0.54), ("depth_min", float, 0.0), ("depth_max", float, 1.0),
        ("disparity_min", float, 0.0), ("disparity_max", float, 1.0),
        ("geom_consistency", bool, False), ("multi_geometry", bool, False),
        ("planar_prior", bool, False), ("superpixel_filter", bool, False),
    )

    def __init__(self):
        for name, typ, default in self._FIELDS:
            object.__setattr__(self, name, default)

    def __setattr__(self, name, value):
        for fname, typ, _ in self._FIELDS:
            if fname == name:
                if typ is float:
                    value = float(np.float32(value))  # stored as C float
                else:
                    value = typ(value)
                object.__setattr__(self, name, value)
                return
        raise AttributeError(f"PatchMatchParams has no field {name!r}")

    def to_c(self) -> N.CParams:
        p = N.CParams()
        for name, _, _ in self._FIELDS:
            setattr(p, name, getattr(self, name))
        return p

    def __repr__(self) -> str:
        return "PatchMatchParams(" + ", ".join(f"{n}={getattr(self, n)!r}" for n, _, _ in self._FIELDS) + ")"


class Camera:
    """mvs::Camera (gmvs/include/data_structure.hpp:45-53)."""

    def __init__(self, c: Optional[N.CCamera] = None):
        self._c = c if c is not None else N.CCamera()

    @property
    def K(self) -> List[float]:
        return list(self._c.K)

    @property
    def R(self) -> List[float]:
        return list(self._c.R)

    @property
    def t(self) -> List[float]:
        return list(self._c.t)

    height = property(lambda s: s._c.height, lambda s, v: setattr(s._c, "height", int(v)))
    width = property(lambda s: s._c.width, lambda s, v: setattr(s._c, "width", int(v)))
    depth_min = property(lambda s: s._c.depth_min, lambda s, v: setattr(s._c, "depth_min", float(v)))
    depth_max = property(lambda s: s._c.depth_max, lambda s, v: setattr(s._c, "depth_max", float(v)))

    def __repr__(self) -> str:
        return f"Camera({self.width}x{self.height}, depth=[{self.depth_min}, {self.depth_max}])"


class Problem:
    """mvs::Problem (gmvs/include/data_structure.hpp:39-43)."""

    def __init__(self, c: Optional[N.CProblem] = None):
        self._c = c if c is not None else N.CProblem()

    ref_image_id = property(lambda s: s._c.ref_image_id,
                            lambda s, v: setattr(s._c, "ref_image_id", int(v)))

    @property
    def num_ngb(self) -> int:
        return self._c.num_ngb

    @property
    def src_image_ids(self) -> List[int]:
        return list(self._c.src_image_ids[: self._c.num_ngb])


def load_camera(K: Sequence[float], R: Sequence[float], t: Sequence[float], height: int,
                width: int, depth_min: float, depth_max: float) -> Camera:
    """mvs::load_camera (gmvs/src/data_structure.cpp:4-29)."""
    K = list(K)
    R = list(R)
    t = list(t)
    if len(K) != 9 or len(R) != 9 or len(t) != 3:
        raise ValueError("load_camera expects K[9], R[9], t[3]")
    c = N.CCamera()
    Kc = (ctypes.c_float * 9)(*K)
    Rc = (ctypes.c_float * 9)(*R)
    tc = (ctypes.c_float * 3)(*t)
    N.check("gmvs_load_camera", N.lib.gmvs_load_camera(Kc, Rc, tc, int(height), int(width),
                                                         float(depth_min), float(depth_max),
                                                         ctypes.byref(c)))
    return Camera(c)


def load_problem(ref_id: int, src_ids: Sequence[int]) -> Problem:
    """mvs::load_problem (gmvs/src/data_structure.cpp:31-39); more than 32
    sources raise ValueError (the reference asserts)."""
    src = list(int(s) for s in src_ids)
    arr = (ctypes.c_int32 * max(1, len(src)))(*src) if src else None
    p = N.CProblem()
    N.check("gmvs_load_problem", N.lib.gmvs_load_problem(int(ref_id), arr, len(src), ctypes.byref(p)))
    return Problem(p)


# ---------------------------------------------------------------- helpers ---
def _require_cpu(name: str, t: torch.Tensor) -> None:
    # CHECK_CPU_INPUT (gmvs/include/cuda_utils.cuh:5-11)
    if t.is_cuda:
        raise RuntimeError(f"{name} must be a CPU tensor")
    if not t.is_contiguous():
        raise RuntimeError(f"{name} must be contiguous")


def _require_cuda(name: str, t: torch.Tensor) -> None:
    if not t.is_cuda:
        raise RuntimeError(f"{name} must be a CUDA tensor")
    if not t.is_contiguous():
        raise RuntimeError(f"{name} must be contiguous")


def _stream() -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class _Workspace:
    """A reusable device byte buffer (grown on demand)."""

    def __init__(self):
        self.buf: Optional[torch.Tensor] = None

    def get(self, nbytes: int, device) -> torch.Tensor:
        if self.buf is None or self.buf.numel() < nbytes or self.buf.device != device:
            self.buf = torch.empty(max(nbytes, 256), dtype=torch.uint8, device=device)
        return self.buf


def _device() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("gmvs_amd needs a HIP device (torch.cuda.is_available() is False)")
    return torch.device("cuda", torch.cuda.current_device())


def fp16_cells_exact(images) -> bool:
    """True when FP16 cells hold every bilinear cell of these images exactly:
    integer intensities in [0, 1024] (8-bit imagery is), so the cell's value,
    its x and y differences and its cross difference (|.| <= 2048) are
    integers FP16 represents exactly, and GMVS_PREC_FP16 -- whose bilinear
    sample widens the FP16 operands into FP32 FMAs -- returns the FP32
    path's results bit for bit with half the gather bytes."""
    oks = []
    for im in images:
        t = im if isinstance(im, torch.Tensor) else torch.as_tensor(np.asarray(im))
        t = t.float()
        if t.numel() == 0:
            continue
        # NaN fails every comparison; one host read for all images
        oks.append(((t >= 0.0) & (t <= 1024.0) & (t == torch.round(t))).all())
    return bool(torch.stack(oks).all()) if oks else True


def resolve_precision(precision, images) -> int:
    """GMVS_PREC_* for a precision given as an int or a name ("fp32", "fp16",
    "bf16", "auto"); "auto" (N.PREC_AUTO) is FP16 cells where
    fp16_cells_exact(images), else FP32 cells: the FP32 path's results
    either way."""
    if isinstance(precision, str):
        if precision not in N.PRECISIONS:
            raise ValueError(f"unknown precision {precision!r}")
        precision = N.PRECISIONS[precision]
    if precision == N.PREC_AUTO:
        return N.PREC_FP16 if fp16_cells_exact(images) else N.PREC_FP32
    if precision not in (N.PREC_FP32, N.PREC_FP16, N.PREC_BF16):
        raise ValueError(f"unknown precision {precision!r}")
    return int(precision)


def pm_params_for_problem(params: PatchMatchParams, ref_cam: Camera, num_images: int) -> N.CParams:
    """The per-call parameter overrides of run_patch_match (src 1377-1380):
    num_images = 1 + sources, depth range = camera range x 0.6 / x 1.2."""
    p = params.to_c()
    p.num_images = num_images
    p.depth_min = float(np.float32(float(np.float32(ref_cam.depth_min)) * 0.6))
    p.depth_max = float(np.float32(float(np.float32(ref_cam.depth_max)) * 1.2))
    return p


def make_pm_args(cams: Sequence[N.CCamera], images: Sequence[torch.Tensor], params: N.CParams, *,
                 init_mode: int, geom: bool, rng_view: int, planes: torch.Tensor,
                 costs: torch.Tensor, depths: Optional[Sequence[torch.Tensor]] = None,
                 ref_normal: Optional[torch.Tensor] = None,
                 prior_planes: Optional[torch.Tensor] = None,
                 plane_masks: Optional[torch.Tensor] = None,
                 out_depth: Optional[torch.Tensor] = None,
                 out_normal: Optional[torch.Tensor] = None,
                 out_cost: Optional[torch.Tensor] = None, seed: int = RANDOM_SEED,
                 skip_filter: bool = False, precision: int = N.PREC_FP32, tile_cols: int = 0,
                 compat: int = 0):
    """One gmvs_pm_args record over device tensors. Returns (args, keep): keep
    holds the host arrays the record points to and must outlive the call.
    compat: N.COMPAT_* bits switching reference quirks to their corrected
    behaviour (0: the reference's)."""
    n = len(images)
    cam_arr = N.camera_array(cams)
    img_arr = N.ptr_array([im.data_ptr() for im in images])
    dep_arr = N.ptr_array([d.data_ptr() if d is not None else None for d in depths]) if depths else None
    a = N.CPMArgs()
    a.num_images = n
    a.cameras = ctypes.cast(cam_arr, ctypes.POINTER(N.CCamera))
    a.images = ctypes.cast(img_arr, ctypes.POINTER(ctypes.c_void_p))
    a.depths = ctypes.cast(dep_arr, ctypes.POINTER(ctypes.c_void_p)) if dep_arr is not None else None
    a.ref_normal = ref_normal.data_ptr() if ref_normal is not None else None
    a.prior_planes = prior_planes.data_ptr() if prior_planes is not None else None
    a.plane_masks = plane_masks.data_ptr() if plane_masks is not None else None
    a.planes = planes.data_ptr()
    a.costs = costs.data_ptr()
    a.params = params
    a.init_mode = init_mode
    a.geom = 1 if geom else 0
    a.seed = seed
    a.rng_view = rng_view
    a.precision = precision
    a.skip_filter = 1 if skip_filter else 0
    a.out_depth = out_depth.data_ptr() if out_depth is not None else None
    a.out_normal = out_normal.data_ptr() if out_normal is not None else None
    a.out_cost = out_cost.data_ptr() if out_cost is not None else None
    a.tile_cols = tile_cols
    a.compat = compat
    return a, (cam_arr, img_arr, dep_arr)


def run_pm_batch(problems: Sequence[dict], workspace: Optional[_Workspace] = None,
                 device=None) -> None:
    """gmvs_pm_run_batch over problems given as make_pm_args keyword dicts: one
    schedule, each kernel launched once for all of them (results identical to
    running them one by one). Asynchronous on the current torch stream."""
    made = [make_pm_args(**p) for p in problems]
    arr = (N.CPMArgs * len(made))()
    for i, (a, _) in enumerate(made):
        arr[i] = a
    dev = device if device is not None else problems[0]["images"][0].device
    nbytes = N.lib.gmvs_pm_batch_workspace_bytes(arr, len(made))
    if nbytes == 0:
        raise ValueError("gmvs_pm_batch_workspace_bytes: invalid batch")
    ws = (workspace or _Workspace()).get(nbytes, dev)
    N.check("gmvs_pm_run_batch", N.lib.gmvs_pm_run_batch(arr, len(made), ctypes.c_void_p(ws.data_ptr()),
                                                         ws.numel(), _stream()))


def run_pm(cams: Sequence[N.CCamera], images: Sequence[torch.Tensor], params: N.CParams, *,
           workspace: Optional[_Workspace] = None, **kw) -> None:
    """Thin call of gmvs_pm_run on device tensors (the current torch stream);
    keywords as make_pm_args. precision: N.PREC_FP32 (the plain path),
    N.PREC_FP16 or N.PREC_BF16 (16-bit source samples, FP32 NCC sums)."""
    n = len(images)
    h, w = images[0].shape
    prec = kw.get("precision", N.PREC_FP32)
    ws = (workspace or _Workspace()).get(N.lib.gmvs_pm_workspace_bytes(h, w, n, prec),
                                         images[0].device)
    a, keep = make_pm_args(cams, images, params, **kw)
    N.check("gmvs_pm_run", N.lib.gmvs_pm_run(ctypes.byref(a), ctypes.c_void_p(ws.data_ptr()),
                                             ws.numel(), _stream()))


class PatchMatcher:
    """mvs::PatchMatcher (gmvs/include/patch_match.cuh:22-54): the
    geometric-consistency PatchMatch pass over all views."""

    def __init__(self):
        self.params = PatchMatchParams()
        # GMVS_PREC_* or a name ("fp32", "fp16", "bf16", "auto"); not in the
        # reference, which computes in FP32 (the default)
        self.precision = N.PREC_FP32
        self._problems: List[Problem] = []
        self._cameras: List[Camera] = []
        self.images_host = self.depths_host = self.normals_host = self.costs_host = None
        self._dev = None
        self._prec = None  # (self.precision, resolved GMVS_PREC_*) for the current images
        self._ws = _Workspace()

    def add_samples(self, problems: Sequence[Problem], cameras: Sequence[Camera],
                    images: torch.Tensor, depths: torch.Tensor, normals: torch.Tensor,
                    costs: torch.Tensor) -> None:
        """src 1322-1349: keeps the CPU tensors (no copy)."""
        for name, t in (("images", images), ("depths", depths), ("normals", normals),
                        ("costs", costs)):
            _require_cpu(name, t)
        self._problems = list(problems)
        self._cameras = list(cameras)
        self.images_host, self.depths_host = images, depths
        self.normals_host, self.costs_host = normals, costs
        self._dev = None
        self._prec = None

    def _resolved_precision(self, images) -> int:
        """resolve_precision once per set of images and precision setting
        ("auto" inspects every image of the scene)."""
        if self._prec is None or self._prec[0] != self.precision:
            self._prec = (self.precision, resolve_precision(self.precision, images))
        return self._prec[1]

    def _upload(self):
        if self._dev is None:
            dev = _device()
            self._dev = (self.images_host.to(dev, torch.float32).contiguous(),
                         self.depths_host.to(dev, torch.float32).contiguous(),
                         self.normals_host.to(dev, torch.float32).contiguous())
        return self._dev

    def run_patch_match(self, problem_idx: int, verbose: bool = False
                        ) -> Tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
        """src 1353-1519: init from (depth, world normal), max_iterations
        sweeps with the geometric term, depth/normal, filter, export."""
        if self.images_host is None:
            raise RuntimeError("add_samples must be called before run_patch_match")
        prob = self._problems[problem_idx]
        ref = prob.ref_image_id
        ids = [ref] + prob.src_image_ids
        images, depths, normals = self._upload()
        h, w = int(images.shape[1]), int(images.shape[2])
        dev = images.device
        cams = [self._cameras[i]._c for i in ids]
        params = pm_params_for_problem(self.params, self._cameras[ref], len(ids))
        self.params.num_images = params.num_images
        self.params.depth_min = params.depth_min
        self.params.depth_max = params.depth_max
        planes = torch.empty(h, w, 4, dtype=torch.float32, device=dev)
        costs = torch.empty(h, w, dtype=torch.float32, device=dev)
        out_d = torch.zeros(h, w, dtype=torch.float32, device=dev)
        out_n = torch.zeros(h, w, 3, dtype=torch.float32, device=dev)
        out_c = torch.zeros(h, w, dtype=torch.float32, device=dev)
        run_pm(cams, [images[i] for i in ids], params, init_mode=N.INIT_DEPTH_NORMAL, geom=True,
               rng_view=ref, planes=planes, costs=costs, depths=[depths[i] for i in ids],
               ref_normal=normals[ref], out_depth=out_d, out_normal=out_n, out_cost=out_c,
               workspace=self._ws, precision=self._resolved_precision(images))
        if verbose:
            for it in range(self.params.max_iterations):
                print(f"iteration: {it}")
        return out_d, out_n, out_c


FUSE_REPROJ_PX = 2.0         # gmvs/src/fusion.cu:133
FUSE_REL_DEPTH = 0.01
FUSE_ANGLE_RAD = 0.174533


class Fuser:
    """mvs::Fuser (gmvs/include/fusion.cuh:17-41)."""

    def __init__(self):
        self.depths_host = None
        self.normals_host = None
        self.masks_cuda = None
        self._problems: List[Problem] = []
        self._cameras: List[Camera] = []
        self._dev = None
        self._ws = _Workspace()
        # thresholds of batch_fusion_kernel, exposed as attributes
        self.reproj_px = FUSE_REPROJ_PX
        self.rel_depth = FUSE_REL_DEPTH
        self.angle_rad = FUSE_ANGLE_RAD
        # 32: the reference's FP32 reprojection check; 64: the check in FP64
        # (gmvs_fuse_view_fp64), for world coordinates large against depths
        self.precision = 32

    def load_samples(self, problems: Sequence[Problem], cameras: Sequence[Camera],
                     depths: torch.Tensor, normals: torch.Tensor, masks: torch.Tensor) -> None:
        """fusion.cu:186-209: CPU depth/normal tensors, CUDA bool masks."""
        _require_cpu("depths", depths)
        _require_cpu("normals", normals)
        _require_cuda("masks", masks)
        if masks.dtype != torch.bool:
            raise RuntimeError("masks must be a bool tensor")
        self._problems = list(problems)
        self._cameras = list(cameras)
        self.depths_host, self.normals_host, self.masks_cuda = depths, normals, masks
        self._dev = None

    def _upload(self):
        if self._dev is None:
            dev = self.masks_cuda.device
            self._dev = (self.depths_host.to(dev, torch.float32).contiguous(),
                         self.normals_host.to(dev, torch.float32).contiguous())
        return self._dev

    def run_fusion(self, problem_idx: int, geom_consistent: int
                   ) -> Tuple[torch.Tensor, torch.Tensor]:
        """fusion.cu:211-275. Mutates masks_cuda in place (source pixels
        used by a fused point are marked)."""
        prob = self._problems[problem_idx]
        ref = prob.ref_image_id
        ids = [ref] + prob.src_image_ids
        depths, normals = self._upload()
        h, w = int(depths.shape[1]), int(depths.shape[2])
        dev = depths.device
        out_d = torch.empty(h, w, dtype=torch.float32, device=dev)
        out_n = torch.empty(h, w, 3, dtype=torch.float32, device=dev)
        fuse_view(ids, [self._cameras[i]._c for i in ids], [depths[i] for i in ids],
                  [normals[i] for i in ids], self.masks_cuda, int(geom_consistent), out_d, out_n,
                  self.reproj_px, self.rel_depth, self.angle_rad, self._ws,
                  precision=self.precision)
        return out_d, out_n


def fuse_view(ids: Sequence[int], cams: Sequence[N.CCamera], depths: Sequence[torch.Tensor],
              normals: Sequence[torch.Tensor], masks: torch.Tensor, geom_consistent: int,
              out_depth: torch.Tensor, out_normal: torch.Tensor,
              reproj_px: float = FUSE_REPROJ_PX, rel_depth: float = FUSE_REL_DEPTH,
              angle_rad: float = FUSE_ANGLE_RAD, workspace: Optional[_Workspace] = None,
              precision: int = 32) -> None:
    """Thin call of gmvs_fuse_view (precision 32) or gmvs_fuse_view_fp64
    (precision 64) on device tensors."""
    if precision not in (32, 64):
        raise ValueError(f"fusion precision must be 32 or 64 (got {precision!r})")
    h, w = out_depth.shape
    ws = (workspace or _Workspace()).get(N.lib.gmvs_fuse_workspace_bytes(h, w), out_depth.device)
    cam_arr = N.camera_array(cams)
    src = (ctypes.c_int32 * max(1, len(ids) - 1))(*ids[1:])
    cos_t = float(np.float32(math.cos(angle_rad)))
    name = "gmvs_fuse_view" if precision == 32 else "gmvs_fuse_view_fp64"
    N.check(name, getattr(N.lib, name)(
        len(ids), ctypes.cast(cam_arr, ctypes.POINTER(N.CCamera)),
        N.ptr_array([d.data_ptr() for d in depths]), N.ptr_array([n.data_ptr() for n in normals]),
        ctypes.c_void_p(masks.data_ptr()), int(masks.shape[0]), int(ids[0]), src,
        int(geom_consistent),
        float(reproj_px), float(rel_depth), cos_t, ctypes.c_void_p(out_depth.data_ptr()),
        ctypes.c_void_p(out_normal.data_ptr()), ctypes.c_void_p(ws.data_ptr()), ws.numel(),
        _stream()))
