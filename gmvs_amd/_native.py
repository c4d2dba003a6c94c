"""ctypes binding of libgmvs_amd.so (the C ABI declared in include/gmvs_amd.h).

The library is built in-tree by ``gmvs_amd._build.build()``; importing this
module loads it and raises if it is missing -- there is no fallback path.
``GMVS_AMD_LIB`` names another build of the same library to load instead
(e.g. one compiled with different launch bounds).
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional, Sequence

_PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("GMVS_AMD_LIB") or os.path.join(_PKG_DIR, "libgmvs_amd.so")

MAX_NGB = 32
OK = 0
ERR_ARG = 1
ERR_HIP = 2
ERR_NOSPACE = 3
ERR_UNSUPPORTED = 4

INIT_RANDOM = 0
INIT_DEPTH_NORMAL = 1
INIT_PRIOR = 2

PREC_FP32 = 0
PREC_FP16 = 1
PREC_BF16 = 2
# host-side choice only (gmvs_amd.src.resolve_precision): FP16 cells when they
# hold the images exactly (results identical to PREC_FP32), else FP32 cells
PREC_AUTO = -1
PRECISIONS = {"fp32": PREC_FP32, "fp16": PREC_FP16, "bf16": PREC_BF16, "auto": PREC_AUTO}

# gmvs_pm_args.compat bits (GMVS_COMPAT_*): 0 reproduces the reference's quirks
COMPAT_FIX_RIGHT_FAR = 1
COMPAT_FIX_COST_DEFAULT = 2
COMPAT_FIX_PRIOR_NORMALS = 4
COMPAT_FIX_PRIOR_SELECTION = 8

TUNE_INIT = 0
TUNE_CANDIDATES = 1
TUNE_REFINE = 2


class CCamera(ctypes.Structure):
    _fields_ = [
        ("K", ctypes.c_float * 9),
        ("R", ctypes.c_float * 9),
        ("t", ctypes.c_float * 3),
        ("height", ctypes.c_int32),
        ("width", ctypes.c_int32),
        ("depth_min", ctypes.c_float),
        ("depth_max", ctypes.c_float),
    ]


class CProblem(ctypes.Structure):
    _fields_ = [
        ("num_ngb", ctypes.c_int32),
        ("ref_image_id", ctypes.c_int32),
        ("src_image_ids", ctypes.c_int32 * MAX_NGB),
    ]


class CParams(ctypes.Structure):
    _fields_ = [
        ("max_iterations", ctypes.c_int32),
        ("patch_size", ctypes.c_int32),
        ("num_images", ctypes.c_int32),
        ("max_image_size", ctypes.c_int32),
        ("radius_increment", ctypes.c_int32),
        ("sigma_spatial", ctypes.c_float),
        ("sigma_color", ctypes.c_float),
        ("top_k", ctypes.c_int32),
        ("baseline", ctypes.c_float),
        ("depth_min", ctypes.c_float),
        ("depth_max", ctypes.c_float),
        ("disparity_min", ctypes.c_float),
        ("disparity_max", ctypes.c_float),
        ("geom_consistency", ctypes.c_bool),
        ("multi_geometry", ctypes.c_bool),
        ("planar_prior", ctypes.c_bool),
        ("superpixel_filter", ctypes.c_bool),
    ]


class CPMArgs(ctypes.Structure):
    _fields_ = [
        ("num_images", ctypes.c_int32),
        ("cameras", ctypes.POINTER(CCamera)),
        ("images", ctypes.POINTER(ctypes.c_void_p)),
        ("depths", ctypes.POINTER(ctypes.c_void_p)),
        ("ref_normal", ctypes.c_void_p),
        ("prior_planes", ctypes.c_void_p),
        ("plane_masks", ctypes.c_void_p),
        ("planes", ctypes.c_void_p),
        ("costs", ctypes.c_void_p),
        ("params", CParams),
        ("init_mode", ctypes.c_int32),
        ("geom", ctypes.c_int32),
        ("seed", ctypes.c_uint32),
        ("rng_view", ctypes.c_uint32),
        ("precision", ctypes.c_int32),
        ("skip_filter", ctypes.c_int32),
        ("out_depth", ctypes.c_void_p),
        ("out_normal", ctypes.c_void_p),
        ("out_cost", ctypes.c_void_p),
        ("tile_cols", ctypes.c_int32),
        ("compat", ctypes.c_uint32),
    ]


# Every symbol include/gmvs_amd.h declares (checked by the CPU test suite).
EXPORTS = (
    "gmvs_last_error",
    "gmvs_version",
    "gmvs_load_camera",
    "gmvs_load_problem",
    "gmvs_pm_workspace_bytes",
    "gmvs_pm_run",
    "gmvs_pm_batch_workspace_bytes",
    "gmvs_pm_run_batch",
    "gmvs_set_blocks_per_cu",
    "gmvs_set_lane_width",
    "gmvs_set_fast_taps",
    "gmvs_set_select_fusion",
    "gmvs_cost_workspace_bytes",
    "gmvs_cost_batch",
    "gmvs_fuse_workspace_bytes",
    "gmvs_fuse_view",
    "gmvs_fuse_view_fp64",
    "gmvs_fuse_noalloc_max_pixels",
    "gmvs_fuse_view_noalloc",
    "gmvs_gather_peak_buffer_bytes",
    "gmvs_gather_peak",
    "gmvs_prior_build",
    "gmvs_prior_device_workspace_bytes",
    "gmvs_prior_build_device",
    "gmvs_delaunay",
    "gmvs_delaunay_cells_device",
    "gmvs_delaunay_cells_host",
    "gmvs_points_workspace_bytes",
    "gmvs_points_gather",
    "gmvs_format_workspace_bytes",
    "gmvs_format_rows",
    "gmvs_sliding_window_normal",
    "gmvs_segments_workspace_bytes",
    "gmvs_filter_by_var_map",
    "gmvs_fusion_planar_mask",
    "gmvs_profile_enable",
    "gmvs_profile_collect",
    "gmvs_profile_reset",
    "gmvs_philox4x32_10",
)


class CPointsArgs(ctypes.Structure):
    _fields_ = [
        ("height", ctypes.c_int32),
        ("width", ctypes.c_int32),
        ("num_views", ctypes.c_int32),
        ("depth_normals", ctypes.POINTER(ctypes.c_void_p)),
        ("poses", ctypes.c_void_p),
        ("fx", ctypes.c_double),
        ("fy", ctypes.c_double),
        ("cx", ctypes.c_double),
        ("cy", ctypes.c_double),
        ("cos_view_angle", ctypes.c_double),
        ("filtered", ctypes.c_void_p),
        ("points", ctypes.c_void_p),
        ("points_filtered", ctypes.c_void_p),
        ("capacity", ctypes.c_uint64),
        ("capacity_filtered", ctypes.c_uint64),
        ("counts", ctypes.c_void_p),
    ]


class CKernelStat(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 32), ("launches", ctypes.c_int64),
                ("total_ms", ctypes.c_double), ("ncc", ctypes.c_uint64),
                ("geom", ctypes.c_uint64)]


class GmvsError(RuntimeError):
    """A C-ABI call returned a non-zero status."""

    def __init__(self, fn: str, code: int, msg: str):
        super().__init__(f"{fn} failed (status {code}): {msg}")
        self.code = code


def _load() -> ctypes.CDLL:
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it first with `python -c \"import __graft_entry__ as g; "
            f"g.build()\"` (or python gmvs_amd/_build.py). gmvs_amd has no CPU fallback."
        )
    lib = ctypes.CDLL(LIB_PATH)
    vp, i32, u32, f32, sz = ctypes.c_void_p, ctypes.c_int32, ctypes.c_uint32, ctypes.c_float, ctypes.c_size_t
    lib.gmvs_last_error.restype = ctypes.c_char_p
    lib.gmvs_version.restype = ctypes.c_char_p
    lib.gmvs_load_camera.argtypes = [vp, vp, vp, i32, i32, f32, f32, ctypes.POINTER(CCamera)]
    lib.gmvs_load_problem.argtypes = [i32, vp, i32, ctypes.POINTER(CProblem)]
    lib.gmvs_pm_workspace_bytes.argtypes = [i32, i32, i32, i32]
    lib.gmvs_pm_workspace_bytes.restype = sz
    lib.gmvs_pm_run.argtypes = [ctypes.POINTER(CPMArgs), vp, sz, vp]
    lib.gmvs_pm_batch_workspace_bytes.argtypes = [ctypes.POINTER(CPMArgs), i32]
    lib.gmvs_pm_batch_workspace_bytes.restype = sz
    lib.gmvs_pm_run_batch.argtypes = [ctypes.POINTER(CPMArgs), i32, vp, sz, vp]
    lib.gmvs_set_blocks_per_cu.argtypes = [i32, i32]
    lib.gmvs_set_lane_width.argtypes = [i32, i32]
    lib.gmvs_set_fast_taps.argtypes = [i32]
    lib.gmvs_set_select_fusion.argtypes = [i32]
    lib.gmvs_cost_workspace_bytes.argtypes = [i32, i32, i32, i32]
    lib.gmvs_cost_workspace_bytes.restype = sz
    lib.gmvs_cost_batch.argtypes = [i32, ctypes.POINTER(CCamera), ctypes.POINTER(vp),
                                    ctypes.POINTER(CParams), i32, i32, vp, vp, vp, vp, sz, vp]
    lib.gmvs_fuse_workspace_bytes.argtypes = [i32, i32]
    lib.gmvs_fuse_workspace_bytes.restype = sz
    lib.gmvs_fuse_view.argtypes = [i32, ctypes.POINTER(CCamera), ctypes.POINTER(vp),
                                   ctypes.POINTER(vp), vp, i32, i32, vp, i32, f32, f32, f32, vp,
                                   vp, vp, sz, vp]
    lib.gmvs_fuse_view_fp64.argtypes = lib.gmvs_fuse_view.argtypes
    lib.gmvs_fuse_noalloc_max_pixels.argtypes = []
    lib.gmvs_fuse_noalloc_max_pixels.restype = ctypes.c_int64
    lib.gmvs_fuse_view_noalloc.argtypes = [i32, ctypes.POINTER(CCamera), ctypes.POINTER(vp),
                                           ctypes.POINTER(vp), vp, i32, i32, vp, i32, f32, f32,
                                           f32, vp, vp, vp]
    lib.gmvs_gather_peak_buffer_bytes.argtypes = [i32, i32]
    lib.gmvs_gather_peak_buffer_bytes.restype = sz
    lib.gmvs_gather_peak.argtypes = [i32, i32, i32, i32, vp, sz, ctypes.POINTER(ctypes.c_double), vp]
    lib.gmvs_prior_device_workspace_bytes.argtypes = [i32, i32]
    lib.gmvs_prior_device_workspace_bytes.restype = sz
    lib.gmvs_prior_build_device.argtypes = [ctypes.POINTER(CCamera), vp, vp, vp,
                                            ctypes.POINTER(CParams), vp, vp,
                                            ctypes.POINTER(ctypes.c_int32), vp, sz, vp]
    lib.gmvs_prior_build.argtypes = [ctypes.POINTER(CCamera), vp, vp, vp,
                                     ctypes.POINTER(CParams), vp, vp, ctypes.POINTER(i32)]
    lib.gmvs_delaunay.argtypes = [vp, i32, vp, i32, ctypes.POINTER(i32)]
    lib.gmvs_delaunay_cells_device.argtypes = [vp, i32, i32, vp, i32, ctypes.POINTER(i32), vp, sz,
                                               vp]
    lib.gmvs_delaunay_cells_host.argtypes = [vp, i32, i32, vp, i32, ctypes.POINTER(i32)]
    u64 = ctypes.c_uint64
    lib.gmvs_points_workspace_bytes.argtypes = [i32, i32, i32]
    lib.gmvs_points_workspace_bytes.restype = sz
    lib.gmvs_points_gather.argtypes = [ctypes.POINTER(CPointsArgs), vp, sz, vp]
    lib.gmvs_format_workspace_bytes.argtypes = [u64]
    lib.gmvs_format_workspace_bytes.restype = sz
    lib.gmvs_format_rows.argtypes = [vp, u64, i32, i32, vp, u64, vp, vp, sz, vp]
    lib.gmvs_sliding_window_normal.argtypes = [vp, vp, i32, i32, i32, vp, vp]
    lib.gmvs_segments_workspace_bytes.argtypes = [i32, i32, i32]
    lib.gmvs_segments_workspace_bytes.restype = sz
    lib.gmvs_filter_by_var_map.argtypes = [vp, vp, i32, i32, i32, f32, vp, vp, sz, vp]
    lib.gmvs_fusion_planar_mask.argtypes = [vp, vp, vp, i32, i32, i32, i32, f32, f32, f32, vp, vp,
                                            sz, vp]
    lib.gmvs_philox4x32_10.argtypes = [vp, vp, vp]
    lib.gmvs_philox4x32_10.restype = None
    lib.gmvs_profile_enable.argtypes = [i32]
    lib.gmvs_profile_collect.argtypes = [ctypes.POINTER(CKernelStat), i32, ctypes.POINTER(i32)]
    for name in EXPORTS:
        getattr(lib, name)
    return lib


lib = _load()


def check(fn: str, code: int) -> None:
    if code != OK:
        msg = lib.gmvs_last_error().decode(errors="replace")
        if code == ERR_ARG:
            raise ValueError(f"{fn}: {msg}")
        raise GmvsError(fn, code, msg)


def profile_enable(on: bool = True) -> None:
    check("gmvs_profile_enable", lib.gmvs_profile_enable(1 if on else 0))


def profile_reset() -> None:
    check("gmvs_profile_reset", lib.gmvs_profile_reset())


def profile_collect() -> dict:
    """{kernel name: dict(launches, ms, ncc, geom)} accumulated since the last
    reset (ncc / geom: bilateral NCCs and geometric terms evaluated)."""
    n = ctypes.c_int32(0)
    check("gmvs_profile_collect", lib.gmvs_profile_collect(None, 0, ctypes.byref(n)))
    arr = (CKernelStat * max(1, n.value))()
    check("gmvs_profile_collect", lib.gmvs_profile_collect(arr, n.value, ctypes.byref(n)))
    return {arr[i].name.decode(): dict(launches=int(arr[i].launches), ms=float(arr[i].total_ms),
                                       ncc=int(arr[i].ncc), geom=int(arr[i].geom))
            for i in range(n.value)}


def ptr_array(ptrs: Sequence[Optional[int]]):
    arr = (ctypes.c_void_p * len(ptrs))()
    for i, p in enumerate(ptrs):
        arr[i] = p if p else None
    return arr


def camera_array(cams: Sequence[CCamera]):
    arr = (CCamera * len(cams))()
    for i, c in enumerate(cams):
        arr[i] = c
    return arr


def set_fast_taps(enable: bool) -> None:
    """Clamp-free NCC tap indexing on (default) or off; results are identical."""
    check("gmvs_set_fast_taps", lib.gmvs_set_fast_taps(1 if enable else 0))


_select_fusion = False


def set_select_fusion(enable: bool) -> None:
    """View selection and adoption as the candidate kernel's tile epilogue
    (True) or as their own launch (False, the default); identical results."""
    global _select_fusion
    check("gmvs_set_select_fusion", lib.gmvs_set_select_fusion(1 if enable else 0))
    _select_fusion = bool(enable)


def select_fusion_enabled() -> bool:
    """Whether set_select_fusion(True) is in effect (the candidate kernel then
    holds one tile per block instead of two)."""
    return _select_fusion


def set_lane_width(kernel: int, lanes: int) -> None:
    """Lanes per pixel (2 or 4; 0 = default) of a PatchMatch NCC kernel (TUNE_*)."""
    check("gmvs_set_lane_width", lib.gmvs_set_lane_width(kernel, lanes))


def gather_peak(cell_bytes: int, pattern: int = 1, window_cells: int = 0,
                rounds: int = 2000) -> float:
    """Peak gather rate (GB/s) of the NCC tap loads, measured on the current
    device (gmvs_gather_peak). pattern 1 (default): the NCC waves' tap-load
    layout, window 2048 cells per block; pattern 0: random lane pairs, window
    2 KB per block."""
    import torch

    win = window_cells or (2048 if pattern == 1 else 2048 // cell_bytes)
    nbytes = lib.gmvs_gather_peak_buffer_bytes(cell_bytes, win)
    if nbytes == 0:
        raise ValueError("gather_peak: bad cell size or window")
    buf = torch.zeros(nbytes, dtype=torch.uint8, device="cuda")
    out = ctypes.c_double(0.0)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    check("gmvs_gather_peak", lib.gmvs_gather_peak(cell_bytes, pattern, win, rounds,
                                                     ctypes.c_void_p(buf.data_ptr()), nbytes,
                                                     ctypes.byref(out), stream))
    return out.value
