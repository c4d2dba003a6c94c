"""ctypes wrapper of liboracle.so -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker (or the timed CPU baseline). The product
(gmvs_amd/) never imports it.

    pm_run(...)        one PatchMatch pass (precision 64 = FP64 oracle,
                       32 = the OpenMP FP32 CPU baseline)
    cost_batch(...)    FP64 bilateral NCC of (pixel, plane) queries
    geom_batch(...)    FP64 geometric-consistency cost of (pixel, plane) queries
    fuse_view(...)     sequential FP32 fusion twin (bit-exact reference)
    sliding_window_normal / filter_by_var_map / fusion_planar_mask(...)
                       sequential twins of the post-fusion utilities
    gather_points(...) numpy restatement of the post-fusion point gather
    philox(...)        the oracle's own Philox4x32-10 block
    t_*(...)           the decision steps of one half-sweep pixel update on
                       hand-built inputs (candidate search, view selection,
                       final costs, adoption, refinement), FP64
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from typing import List, Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


def build() -> str:
    res = subprocess.run(["make", "-C", HERE, "-j4"], capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + res.stdout + res.stderr)
    return LIB_PATH


class OCamera(ctypes.Structure):
    _fields_ = [("K", ctypes.c_float * 9), ("R", ctypes.c_float * 9), ("t", ctypes.c_float * 3),
                ("height", ctypes.c_int32), ("width", ctypes.c_int32),
                ("depth_min", ctypes.c_float), ("depth_max", ctypes.c_float)]


class OParams(ctypes.Structure):
    _fields_ = [("max_iterations", ctypes.c_int32), ("patch_size", ctypes.c_int32),
                ("num_images", ctypes.c_int32), ("max_image_size", ctypes.c_int32),
                ("radius_increment", ctypes.c_int32), ("sigma_spatial", ctypes.c_float),
                ("sigma_color", ctypes.c_float), ("top_k", ctypes.c_int32),
                ("baseline", ctypes.c_float), ("depth_min", ctypes.c_float),
                ("depth_max", ctypes.c_float), ("disparity_min", ctypes.c_float),
                ("disparity_max", ctypes.c_float), ("geom_consistency", ctypes.c_bool),
                ("multi_geometry", ctypes.c_bool), ("planar_prior", ctypes.c_bool),
                ("superpixel_filter", ctypes.c_bool)]


class OPMArgs(ctypes.Structure):
    _fields_ = [("cameras", ctypes.POINTER(OCamera)), ("num_images", ctypes.c_int32),
                ("images", ctypes.POINTER(ctypes.c_void_p)),
                ("depths", ctypes.POINTER(ctypes.c_void_p)),
                ("ref_normal", ctypes.c_void_p), ("prior_planes", ctypes.c_void_p),
                ("plane_masks", ctypes.c_void_p), ("planes", ctypes.c_void_p),
                ("costs", ctypes.c_void_p), ("selected_views", ctypes.c_void_p),
                ("params", OParams), ("init_mode", ctypes.c_int32), ("geom", ctypes.c_int32),
                ("seed", ctypes.c_uint32), ("rng_view", ctypes.c_uint32),
                ("threads", ctypes.c_int32), ("skip_filter", ctypes.c_int32),
                ("out_depth", ctypes.c_void_p), ("out_normal", ctypes.c_void_p),
                ("out_cost", ctypes.c_void_p), ("compat", ctypes.c_uint32)]


def _load():
    if not os.path.exists(LIB_PATH):
        build()
    lib = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    lib.oracle_philox4x32_10.argtypes = [vp, vp, vp]
    lib.oracle_philox4x32_10.restype = None
    lib.oracle_cost_batch.argtypes = [ctypes.POINTER(OCamera), ctypes.c_int,
                                      ctypes.POINTER(vp), ctypes.POINTER(OParams), ctypes.c_int,
                                      vp, vp, vp, vp, vp]
    lib.oracle_geom_batch.argtypes = [ctypes.POINTER(OCamera), ctypes.c_int, ctypes.POINTER(vp),
                                      ctypes.c_int, vp, vp, vp]
    lib.oracle_geom_batch32.argtypes = lib.oracle_geom_batch.argtypes
    lib.oracle_pm_run.argtypes = [ctypes.c_int, ctypes.POINTER(OPMArgs)]
    lib.oracle_pm_lockstep.argtypes = [ctypes.POINTER(OPMArgs), ctypes.c_int32, vp, vp, vp, vp]
    lib.oracle_pm_lockstep.restype = ctypes.c_int32
    lib.oracle_max_threads.restype = ctypes.c_int
    lib.oracle_set_threads.argtypes = [ctypes.c_int]
    lib.oracle_set_threads.restype = None
    lib.oracle_fuse_view.argtypes = [ctypes.POINTER(OCamera), ctypes.c_int, ctypes.POINTER(vp),
                                     ctypes.POINTER(vp), vp, ctypes.c_int, vp, ctypes.c_int,
                                     ctypes.c_float, ctypes.c_float, ctypes.c_float, vp, vp]
    lib.oracle_fuse_view.restype = ctypes.c_int64
    lib.oracle_fuse_view64.argtypes = lib.oracle_fuse_view.argtypes
    lib.oracle_fuse_view64.restype = ctypes.c_int64
    i32, f32 = ctypes.c_int, ctypes.c_float
    lib.oracle_sliding_window_normal.argtypes = [vp, vp, i32, i32, i32, vp]
    lib.oracle_sliding_window_normal.restype = None
    lib.oracle_filter_by_var_map.argtypes = [vp, vp, i32, i32, f32, vp]
    lib.oracle_filter_by_var_map.restype = None
    lib.oracle_fusion_planar_mask.argtypes = [vp, vp, vp, i32, i32, i32, f32, f32, f32, vp]
    lib.oracle_fusion_planar_mask.restype = None
    d = ctypes.c_double
    lib.oracle_t_candidates.argtypes = [vp, i32, i32, i32, i32, ctypes.c_uint32, vp, vp]
    lib.oracle_t_select_views.argtypes = [vp, i32, vp, vp, i32, vp, vp, vp, vp, vp]
    lib.oracle_t_final_costs.argtypes = [vp, i32, vp, vp, d, vp, vp]
    lib.oracle_t_adopt.argtypes = [i32, vp, vp, vp, d, d, d]
    lib.oracle_t_refine_hypotheses.argtypes = [d, vp, d, vp, d, vp, vp, vp]
    lib.oracle_t_refine_hypotheses.restype = None
    lib.oracle_t_refine_adopt.argtypes = [i32, vp, vp, vp, d, d, vp, vp]
    lib.oracle_t_prior_score.argtypes = [d, d, d, vp, vp, d, d]
    lib.oracle_t_prior_score.restype = d
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32))


def make_camera(K, R, t, height, width, depth_min, depth_max) -> OCamera:
    c = OCamera()
    c.K[:] = [float(x) for x in np.asarray(K, np.float64).reshape(9)]
    c.R[:] = [float(x) for x in np.asarray(R, np.float64).reshape(9)]
    c.t[:] = [float(x) for x in np.asarray(t, np.float64).reshape(3)]
    c.height, c.width = int(height), int(width)
    c.depth_min, c.depth_max = float(depth_min), float(depth_max)
    return c


def make_params(**kw) -> OParams:
    p = OParams()
    defaults = dict(max_iterations=3, patch_size=11, num_images=5, max_image_size=3200,
                    radius_increment=2, sigma_spatial=5.0, sigma_color=3.0, top_k=4,
                    baseline=0.54, depth_min=0.0, depth_max=1.0, disparity_min=0.0,
                    disparity_max=1.0, geom_consistency=False, multi_geometry=False,
                    planar_prior=False, superpixel_filter=False)
    defaults.update(kw)
    for k, v in defaults.items():
        setattr(p, k, v)
    return p


def philox(ctr: Sequence[int], key: Sequence[int]) -> List[int]:
    c = np.asarray(ctr, np.uint32)
    k = np.asarray(key, np.uint32)
    o = np.zeros(4, np.uint32)
    lib().oracle_philox4x32_10(_ptr(c), _ptr(k), _ptr(o))
    return [int(x) for x in o]


def _cams(cams: Sequence[OCamera]):
    arr = (OCamera * len(cams))()
    for i, c in enumerate(cams):
        arr[i] = c
    return arr


def _ptrs(arrs: Sequence[Optional[np.ndarray]]):
    out = (ctypes.c_void_p * len(arrs))()
    for i, a in enumerate(arrs):
        out[i] = None if a is None else a.ctypes.data
    return out


def cost_batch(cams: Sequence[OCamera], images: Sequence[np.ndarray], params: OParams,
               xy: np.ndarray, planes: np.ndarray):
    """Returns (costs, edge_px, min_var), each [n, V] float64."""
    imgs = [_f32(a) for a in images]
    xy = np.ascontiguousarray(xy, np.int32)
    planes = _f32(planes)
    n, nv = xy.shape[0], len(imgs) - 1
    costs = np.zeros((n, nv), np.float64)
    edge = np.zeros((n, nv), np.float64)
    minvar = np.zeros((n, nv), np.float64)
    rc = lib().oracle_cost_batch(_cams(cams), len(imgs), _ptrs(imgs), ctypes.byref(params), n,
                                 _ptr(xy), _ptr(planes), _ptr(costs), _ptr(edge), _ptr(minvar))
    assert rc == 0, rc
    return costs, edge, minvar


def geom_batch(cams: Sequence[OCamera], depths: Sequence[np.ndarray], xy: np.ndarray,
               planes: np.ndarray, precision: int = 64) -> np.ndarray:
    ds = [_f32(a) for a in depths]
    xy = np.ascontiguousarray(xy, np.int32)
    planes = _f32(planes)
    out = np.zeros((xy.shape[0], len(ds) - 1), np.float64)
    fn = lib().oracle_geom_batch if precision == 64 else lib().oracle_geom_batch32
    rc = fn(_cams(cams), len(ds), _ptrs(ds), xy.shape[0], _ptr(xy), _ptr(planes), _ptr(out))
    assert rc == 0, rc
    return out


def pm_run(cams: Sequence[OCamera], images: Sequence[np.ndarray], params: OParams, *,
           precision: int = 64, init_mode: int = 0, geom: bool = False, seed: int = 3407,
           rng_view: int = 0, depths: Optional[Sequence[np.ndarray]] = None,
           ref_normal: Optional[np.ndarray] = None, prior_planes: Optional[np.ndarray] = None,
           plane_masks: Optional[np.ndarray] = None, planes: Optional[np.ndarray] = None,
           costs: Optional[np.ndarray] = None, threads: int = 0, skip_filter: bool = False,
           compat: int = 0, lockstep: int = 0):
    """One PatchMatch pass. Returns dict(depth, normal, cost, planes, selected_views)."""
    imgs = [_f32(a) for a in images]
    h, w = imgs[0].shape
    hw = h * w
    st_planes = np.zeros((hw, 4), np.float32) if planes is None else _f32(planes).reshape(hw, 4).copy()
    st_costs = np.zeros(hw, np.float32) if costs is None else _f32(costs).reshape(hw).copy()
    sel = np.zeros(hw, np.uint32)
    od = np.zeros((h, w), np.float32)
    on = np.zeros((h, w, 3), np.float32)
    oc = np.zeros((h, w), np.float32)
    ds = None if depths is None else [_f32(a) for a in depths]
    rn = None if ref_normal is None else _f32(ref_normal)
    pp = None if prior_planes is None else _f32(prior_planes)
    pmk = None if plane_masks is None else np.ascontiguousarray(plane_masks, np.uint32)
    a = OPMArgs()
    cam_arr = _cams(cams)
    img_ptrs = _ptrs(imgs)
    dep_ptrs = _ptrs(ds) if ds is not None else None
    a.cameras = ctypes.cast(cam_arr, ctypes.POINTER(OCamera))
    a.num_images = len(imgs)
    a.images = ctypes.cast(img_ptrs, ctypes.POINTER(ctypes.c_void_p))
    a.depths = ctypes.cast(dep_ptrs, ctypes.POINTER(ctypes.c_void_p)) if dep_ptrs is not None else None
    a.ref_normal = _ptr(rn)
    a.prior_planes = _ptr(pp)
    a.plane_masks = _ptr(pmk)
    a.planes = _ptr(st_planes)
    a.costs = _ptr(st_costs)
    a.selected_views = _ptr(sel)
    a.params = params
    a.params.num_images = len(imgs)
    a.init_mode = init_mode
    a.geom = 1 if geom else 0
    a.seed = seed
    a.rng_view = rng_view
    a.threads = threads
    a.skip_filter = 1 if skip_filter else 0
    a.compat = compat
    a.out_depth = _ptr(od)
    a.out_normal = _ptr(on)
    a.out_cost = _ptr(oc)
    keep = (cam_arr, img_ptrs, dep_ptrs, imgs, ds, rn, pp, pmk)  # alive until the call returns
    if lockstep:
        cap = lockstep
        ids = np.zeros((cap, 5), np.int32)
        mg = np.zeros((cap, 4), np.float64)
        cnt = np.zeros(1, np.int64)
        vals = np.zeros((cap, 35), np.float64)
        rows = lib().oracle_pm_lockstep(ctypes.byref(a), cap, _ptr(ids), _ptr(mg), _ptr(cnt),
                                        _ptr(vals))
        assert rows >= 0, f"oracle_pm_lockstep returned {rows}"
        del keep
        return dict(ids=ids[:rows], margins=mg[:rows], decisions=int(cnt[0]), values=vals[:rows])
    rc = lib().oracle_pm_run(precision, ctypes.byref(a))
    del keep
    assert rc == 0, f"oracle_pm_run returned {rc}"
    return dict(depth=od, normal=on, cost=oc, planes=st_planes.reshape(h, w, 4),
                selected_views=sel.reshape(h, w))


def pm_lockstep(cams, images, params, cap: int = 1 << 20, **kw):
    """FP32 against FP64 from the same FP64 state at every half-sweep: the
    pixel updates whose decisions differ, as dict(ids [n, 5] = (iteration,
    colour, pixel, what, masked), margins [n, 4] = FP64 (draw, argext, adopt,
    refine) margins, decisions = pixel updates compared). what: 1 drawn views,
    2 argmin/argmax, 4 propagation adoption, 8 refinement adoption, 16 an NCC
    at one of its two discontinuities in one precision only."""
    return pm_run(cams, images, params, lockstep=cap, **kw)


def fuse_view(cams: Sequence[OCamera], depths: Sequence[np.ndarray],
              normals: Sequence[np.ndarray], masks: np.ndarray, ref_id: int,
              src_ids: Sequence[int], geom_consistent: int, reproj_px: float,
              rel_depth: float, cos_angle: float, precision: int = 32):
    """Sequential fusion of one view, the geometry in FP32 (the reference) or
    FP64 (precision 64); masks ([N,H,W] uint8) is updated in place."""
    ds = [_f32(a) for a in depths]
    ns = [_f32(a) for a in normals]
    h, w = ds[0].shape
    od = np.zeros((h, w), np.float32)
    on = np.zeros((h, w, 3), np.float32)
    sids = np.ascontiguousarray(src_ids, np.int32)
    assert masks.dtype == np.uint8 and masks.flags.c_contiguous
    fn = lib().oracle_fuse_view if precision == 32 else lib().oracle_fuse_view64
    n = fn(_cams(cams), len(ds), _ptrs(ds), _ptrs(ns), _ptr(masks), ref_id, _ptr(sids),
           geom_consistent, reproj_px, rel_depth, cos_angle, _ptr(od), _ptr(on))
    return od, on, int(n)


def sliding_window_normal(planar: np.ndarray, normal: np.ndarray, window_size: int) -> np.ndarray:
    """planar_initer.cu:471-535, sequential."""
    pl = np.ascontiguousarray(planar, np.int32)
    nm = _f32(normal)
    h, w = pl.shape
    out = np.zeros((h, w, 3), np.float32)
    lib().oracle_sliding_window_normal(_ptr(pl), _ptr(nm), h, w, int(window_size), _ptr(out))
    return out


def filter_by_var_map(var_map: np.ndarray, seg: np.ndarray, var_thresh: float) -> np.ndarray:
    """planar_initer.cu:634-684, sequential."""
    v = _f32(var_map)
    s = np.ascontiguousarray(seg, np.int32)
    h, w = s.shape
    out = np.zeros((h, w), np.int32)
    lib().oracle_filter_by_var_map(_ptr(v), _ptr(s), h, w, float(var_thresh), _ptr(out))
    return out


def fusion_planar_mask(seg: np.ndarray, planar: np.ndarray, normals: np.ndarray, thresh: int,
                       nonplanar_percent: float, cos_sim_thresh: float,
                       match_ratio_thresh: float) -> np.ndarray:
    """planar_initer.cu:537-632, sequential."""
    s = np.ascontiguousarray(seg, np.int32)
    pl = np.ascontiguousarray(planar, np.int32)
    nm = _f32(normals)
    rows, cols = s.shape
    out = np.zeros((rows, cols), np.int32)
    lib().oracle_fusion_planar_mask(_ptr(s), _ptr(pl), _ptr(nm), rows, cols, int(thresh),
                                    float(nonplanar_percent), float(cos_sim_thresh),
                                    float(match_ratio_thresh), _ptr(out))
    return out


def gather_points(depth_normals, poses, intrinsic, height: int, width: int):
    """The gather loop of mvs_fusion_segmentaion_cpu.py:158-205 restated in
    numpy (numpy >= 2 promotion: the float32 pixel grid against float64
    intrinsics gives float64 rays). Returns (points [M,6], filtered points
    [Mf,6], filtered maps [V,H,W,4] float32)."""
    K = np.asarray(intrinsic, np.float64)
    fx, fy, cx, cy = K[0, 0], K[1, 1], K[0, 2], K[1, 2]
    yy, xx = np.meshgrid(np.arange(height, dtype=np.float32), np.arange(width, dtype=np.float32),
                         indexing="ij")
    xx = (xx - cx) / fx
    yy = (yy - cy) / fy
    dirs_ = np.stack((xx, yy, np.ones_like(xx)), axis=-1).reshape(-1, 3)
    pts_all, pts_f, maps = [], [], []
    for pose, dn in zip(poses, depth_normals):
        E = np.asarray(pose, np.float64).reshape(4, 4)
        origins = np.tile(E[:3, 3], (height * width, 1))
        dirs = dirs_ @ E[:3, :3].T
        depth = dn[..., :1].reshape(-1, 1)
        normal = dn[..., 1:].reshape(-1, 3)
        m = (depth > 0).squeeze()
        points = origins[m] + depth[m] * dirs[m]
        vm = ((-dirs[m]) * normal[m]).sum(-1) > np.cos(70 / 180 * np.pi)
        keep = np.zeros(height * width, bool)
        keep[np.flatnonzero(m)[vm]] = True
        maps.append(np.where(keep.reshape(height, width, 1), dn, 0).astype(np.float32))
        pn = np.concatenate((points, normal[m]), axis=1)
        pts_all.append(pn)
        pts_f.append(pn[vm])
    return np.concatenate(pts_all), np.concatenate(pts_f), np.stack(maps)


# ---- decision steps of the half-sweep (FP64), for known-answer tests ----

def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32))


def t_candidates(costs: np.ndarray, px: int, py: int, compat: int = 0):
    """Candidate search over a [h, w] cost grid: (flag[8] bool, pos[8] flat index or -1)."""
    c = _f64(costs)
    h, w = c.shape
    flag = np.zeros(8, np.int32)
    pos = np.zeros(8, np.int32)
    rc = lib().oracle_t_candidates(_ptr(c), h, w, px, py, compat, _ptr(flag), _ptr(pos))
    assert rc == 0, rc
    return flag.astype(bool), pos


def t_select_views(ca: np.ndarray, flag, nbsel, it: int, draws):
    """View selection from an [8, nv] cost table: (cdf[nv], counts[nv], tsel, wn)."""
    ca = _f64(ca)
    nv = ca.shape[1]
    fl = _i32(flag)
    nb = np.ascontiguousarray(nbsel, np.uint32)
    dr = _f64(draws)
    assert dr.shape == (15,)
    cdf = np.zeros(nv)
    wv = np.zeros(nv)
    tsel = np.zeros(1, np.uint32)
    wn = np.zeros(1)
    rc = lib().oracle_t_select_views(_ptr(ca), nv, _ptr(fl), _ptr(nb), it, _ptr(dr), _ptr(cdf),
                                     _ptr(wv), _ptr(tsel), _ptr(wn))
    assert rc == 0, rc
    return cdf, wv, int(tsel[0]), float(wn[0])


def t_final_costs(ca: np.ndarray, flag, wv, wn: float, geom: Optional[np.ndarray] = None):
    ca = _f64(ca)
    nv = ca.shape[1]
    fl = _i32(flag)
    w_ = _f64(wv)
    g = None if geom is None else _f64(geom)
    fc = np.zeros(8)
    rc = lib().oracle_t_final_costs(_ptr(ca), nv, _ptr(fl), _ptr(w_), float(wn), _ptr(g), _ptr(fc))
    assert rc == 0, rc
    return fc


def t_adopt(prior: bool, score, flag, dk, now: float, dmin: float, dmax: float) -> int:
    s, fl, d = _f64(score), _i32(flag), _f64(dk)
    return int(lib().oracle_t_adopt(1 if prior else 0, _ptr(s), _ptr(fl), _ptr(d), float(now),
                                    float(dmin), float(dmax)))


def t_refine_hypotheses(drand, nrand, dcur, ncur, dpert, npert):
    nr, nc, npt = _f64(nrand), _f64(ncur), _f64(npert)
    depths = np.zeros(5)
    normals = np.zeros((5, 3))
    lib().oracle_t_refine_hypotheses(float(drand), _ptr(nr), float(dcur), _ptr(nc), float(dpert),
                                     _ptr(npt), _ptr(depths), _ptr(normals))
    return depths, normals


def t_refine_adopt(pm: bool, tc, db, sc, dmin: float, dmax: float, cost: float,
                   restricted: float):
    """Returns (last adopted hypothesis or -1, final cost, final restricted score)."""
    t, d, s = _f64(tc), _f64(db), _f64(sc)
    c = np.array([cost], np.float64)
    r = np.array([restricted], np.float64)
    k = lib().oracle_t_refine_adopt(1 if pm else 0, _ptr(t), _ptr(d), _ptr(s), float(dmin),
                                    float(dmax), _ptr(c), _ptr(r))
    return int(k), float(c[0]), float(r[0])


def t_prior_score(cost, depth, depth_prior, n_prior, n, dmin, dmax) -> float:
    a, b = _f64(n_prior), _f64(n)
    return float(lib().oracle_t_prior_score(float(cost), float(depth), float(depth_prior), _ptr(a),
                                            _ptr(b), float(dmin), float(dmax)))


def set_threads(n: int) -> None:
    """OpenMP threads of the fusion twin (and of any call without a thread count)."""
    lib().oracle_set_threads(int(n))
