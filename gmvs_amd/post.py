"""After fusion, on the device: the point gather with its view-angle filter,
the ``.xyz`` text export, and the segment / normal utilities that the
reference's ``planar_initer`` module exports.

* ``gather_points`` -- the "Gather mvs points" loop of
  gmvs/scripts/mvs_fusion_segmentaion_cpu.py:158-205;
* ``format_rows`` / ``savetxt`` -- its ``np.savetxt(..., fmt="%.4f")``
  (mvs_fusion_segmentaion_cpu.py:207-216), formatted on the GPU;
* ``sliding_window_normal``, ``filter_by_var_map``, ``fusion_planar_mask`` --
  gmvs/lib/src/planar_initer.cu:471-535, 634-684, 537-632.

Every function runs the library's HIP kernels (there is no CPU path) and
accepts numpy arrays (returning numpy, as the reference's cv::Mat bindings
do) or device tensors (returning device tensors).
"""
from __future__ import annotations

import ctypes
import math
from typing import List, Optional, Sequence, Tuple, Union

import numpy as np
import torch

from . import _native as N
from .src import _device

ArrayLike = Union[np.ndarray, torch.Tensor]

COS_70 = math.cos(70 / 180 * math.pi)  # mvs_fusion_segmentaion_cpu.py:190-192


def _stream(dev) -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)


def _to_dev(a: ArrayLike, dtype: torch.dtype, dev) -> torch.Tensor:
    t = a if isinstance(a, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(a))
    return t.to(device=dev, dtype=dtype).contiguous()


def _ws(nbytes: int, dev) -> torch.Tensor:
    return torch.empty(max(int(nbytes), 256), dtype=torch.uint8, device=dev)


def gather_points(depth_normals: Sequence[ArrayLike], poses: Sequence[np.ndarray],
                  intrinsic: np.ndarray, cos_view: float = COS_70
                  ) -> Tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
    """World points of every view's fused depth_normal map [H, W, 4].

    poses: camera-to-world 4x4 per view (pose/NNNN.txt); intrinsic: 3x3 or
    4x4 (fx, fy, cx, cy at [0,0], [1,1], [0,2], [1,2]). Returns device tensors
    (points [M, 6] float64, filtered points [Mf, 6] float64, filtered maps
    [V, H, W, 4] float32): rows [x y z nx ny nz] in view order, then pixel
    order; a point is kept by the filter when -dir . n > cos_view with dir the
    unnormalised ray R (x', y', 1)."""
    dev = _device()
    if len(depth_normals) == 0 or len(depth_normals) != len(poses):
        raise ValueError("gather_points: one pose per depth_normal map is needed")
    maps = [_to_dev(d, torch.float32, dev) for d in depth_normals]
    h, w = int(maps[0].shape[0]), int(maps[0].shape[1])
    for m in maps:
        if tuple(m.shape) != (h, w, 4):
            raise ValueError(f"gather_points: depth_normal maps must all be [{h}, {w}, 4], "
                             f"got {list(m.shape)}")
    P = np.ascontiguousarray(np.stack([np.asarray(p, dtype=np.float64).reshape(4, 4)
                                       for p in poses]))
    K = np.asarray(intrinsic, dtype=np.float64)
    nv = len(maps)
    counts = torch.zeros(2, dtype=torch.int64, device=dev)
    filtered = torch.empty(nv, h, w, 4, dtype=torch.float32, device=dev)
    a = N.CPointsArgs()
    a.height, a.width, a.num_views = h, w, nv
    ptrs = N.ptr_array([m.data_ptr() for m in maps])
    a.depth_normals = ptrs
    a.poses = P.ctypes.data
    a.fx, a.fy, a.cx, a.cy = float(K[0, 0]), float(K[1, 1]), float(K[0, 2]), float(K[1, 2])
    a.cos_view_angle = float(cos_view)
    a.counts = counts.data_ptr()
    ws = _ws(N.lib.gmvs_points_workspace_bytes(h, w, nv), dev)
    st = _stream(dev)
    N.check("gmvs_points_gather", N.lib.gmvs_points_gather(ctypes.byref(a), ws.data_ptr(), ws.numel(), st))
    n_all, n_f = (int(x) for x in counts.cpu())
    pts = torch.empty(n_all, 6, dtype=torch.float64, device=dev)
    pts_f = torch.empty(n_f, 6, dtype=torch.float64, device=dev)
    a.filtered = filtered.data_ptr()
    a.points = pts.data_ptr() if n_all else None
    a.points_filtered = pts_f.data_ptr() if n_f else None
    a.capacity, a.capacity_filtered = n_all, n_f
    N.check("gmvs_points_gather", N.lib.gmvs_points_gather(ctypes.byref(a), ws.data_ptr(), ws.numel(), st))
    return pts, pts_f, filtered


def format_rows_device(rows: ArrayLike, decimals: int = 4) -> torch.Tensor:
    """The text np.savetxt(f, rows, fmt=f"%.{decimals}f") writes for a 1-D or
    2-D float64 array, formatted on the GPU: a device uint8 tensor."""
    dev = _device()
    r = _to_dev(rows, torch.float64, dev)
    if r.dim() == 1:
        r = r.reshape(-1, 1)  # np.savetxt writes a 1-D array one value per line
    if r.dim() != 2:
        raise ValueError("format_rows: rows must be 1-D or 2-D")
    n, ncol = int(r.shape[0]), int(r.shape[1])
    if n == 0:
        return torch.empty(0, dtype=torch.uint8, device=dev)
    out_len = torch.zeros(1, dtype=torch.int64, device=dev)
    ws = _ws(N.lib.gmvs_format_workspace_bytes(n), dev)
    st = _stream(dev)
    N.check("gmvs_format_rows", N.lib.gmvs_format_rows(r.data_ptr(), n, ncol, decimals, None, 0,
                                                       out_len.data_ptr(), ws.data_ptr(), ws.numel(), st))
    total = int(out_len.cpu()[0])  # UINT64_MAX reads as -1
    if total < 0:
        raise ValueError("format_rows: a value is too large for fixed-point formatting "
                         f"(|v| * 10^{decimals} >= 2^63)")
    out = torch.empty(total, dtype=torch.uint8, device=dev)
    N.check("gmvs_format_rows", N.lib.gmvs_format_rows(r.data_ptr(), n, ncol, decimals, out.data_ptr(),
                                                       total, out_len.data_ptr(), ws.data_ptr(),
                                                       ws.numel(), st))
    return out


def format_rows(rows: ArrayLike, decimals: int = 4) -> bytes:
    """format_rows_device, copied to the host as bytes."""
    return format_rows_device(rows, decimals).cpu().numpy().tobytes()


def savetxt(path: str, rows: ArrayLike, decimals: int = 4) -> int:
    """np.savetxt(path, rows, fmt=f"%.{decimals}f") with the formatting on the
    GPU; returns the bytes written."""
    text = format_rows_device(rows, decimals).cpu().numpy()
    with open(path, "wb") as f:
        f.write(memoryview(text))
    return int(text.size)


def _like_input(x: torch.Tensor, ref: ArrayLike):
    return x if isinstance(ref, torch.Tensor) else x.cpu().numpy()


def sliding_window_normal(planar_region: ArrayLike, normal: ArrayLike, window_size: int):
    """planar_initer.cu:471-535: [H, W, 3] float32, 0 on planar pixels (label
    1), else the mean normal of the window's non-planar pixels."""
    dev = _device()
    pr = _to_dev(planar_region, torch.int32, dev)
    nm = _to_dev(normal, torch.float32, dev)
    if pr.dim() != 2 or tuple(nm.shape) != (pr.shape[0], pr.shape[1], 3):
        raise ValueError("sliding_window_normal: planar_region [H, W] and normal [H, W, 3] expected")
    h, w = int(pr.shape[0]), int(pr.shape[1])
    out = torch.empty(h, w, 3, dtype=torch.float32, device=dev)
    N.check("gmvs_sliding_window_normal",
            N.lib.gmvs_sliding_window_normal(pr.data_ptr(), nm.data_ptr(), h, w, int(window_size),
                                             out.data_ptr(), _stream(dev)))
    return _like_input(out, normal)


def _segments(ids: ArrayLike, dev) -> Tuple[torch.Tensor, int]:
    t = _to_dev(ids, torch.int32, dev)
    if t.dim() != 2:
        raise ValueError("segment ids must be [H, W]")
    if t.numel() and int(t.min()) < 0:
        raise ValueError("segment ids must be >= 0 (the reference indexes per-segment arrays by id)")
    return t, (int(t.max()) + 1 if t.numel() else 1)


def filter_by_var_map(var_map: ArrayLike, segment_ids_map: ArrayLike, var_thresh: float):
    """planar_initer.cu:634-684: [H, W] int32, the segment id where the
    segment's mean variance is <= var_thresh, else 0."""
    dev = _device()
    seg, ns = _segments(segment_ids_map, dev)
    var = _to_dev(var_map, torch.float32, dev)
    if tuple(var.shape) != tuple(seg.shape):
        raise ValueError("filter_by_var_map: var_map and segment_ids_map must have one shape")
    h, w = int(seg.shape[0]), int(seg.shape[1])
    out = torch.empty(h, w, dtype=torch.int32, device=dev)
    ws = _ws(N.lib.gmvs_segments_workspace_bytes(h, w, ns), dev)
    N.check("gmvs_filter_by_var_map",
            N.lib.gmvs_filter_by_var_map(var.data_ptr(), seg.data_ptr(), h, w, ns, float(var_thresh),
                                         out.data_ptr(), ws.data_ptr(), ws.numel(), _stream(dev)))
    return _like_input(out, segment_ids_map)


def fusion_planar_mask(seg_ids: ArrayLike, planar_mask: ArrayLike, filter_normals: ArrayLike,
                       rows: int, cols: int, thresh: int = 100, nonplanar_percent: float = 0.75,
                       cos_sim_thresh: float = 0.8, match_ratio_thresh: float = 0.9):
    """planar_initer.cu:537-632: [rows, cols] int32 labels -- 0, the segment
    id, 300 (object) or 301, 302, ... (textureless segments in id order)."""
    dev = _device()
    seg, ns = _segments(seg_ids, dev)
    pm = _to_dev(planar_mask, torch.int32, dev)
    fn = _to_dev(filter_normals, torch.float32, dev)
    if tuple(seg.shape) != (rows, cols) or tuple(pm.shape) != (rows, cols) or \
            tuple(fn.shape) != (rows, cols, 3):
        raise ValueError(f"fusion_planar_mask: seg_ids / planar_mask [{rows}, {cols}] and "
                         f"filter_normals [{rows}, {cols}, 3] expected")
    out = torch.empty(rows, cols, dtype=torch.int32, device=dev)
    ws = _ws(N.lib.gmvs_segments_workspace_bytes(rows, cols, ns), dev)
    N.check("gmvs_fusion_planar_mask",
            N.lib.gmvs_fusion_planar_mask(seg.data_ptr(), pm.data_ptr(), fn.data_ptr(), rows, cols, ns,
                                          int(thresh), float(nonplanar_percent), float(cos_sim_thresh),
                                          float(match_ratio_thresh), out.data_ptr(), ws.data_ptr(),
                                          ws.numel(), _stream(dev)))
    return _like_input(out, seg_ids)
