"""gmvs_fuse_view_noalloc (static device scratch, no workspace) against
gmvs_fuse_view (caller workspace), both called through the C ABI: outputs
and mask updates bit for bit, on two calls in a row, the second on a smaller
view (the static array keeps the first call's contents)."""
import ctypes
import math

import numpy as np
import pytest
import torch

from tests import helpers as T
from tests.test_gpu_fusion import noisy_maps

pytestmark = pytest.mark.gpu


def _call(N, noalloc, cams, ids, dd, nd, masks, gc, od, on, reproj, rel, cos_t, ws=None):
    cam_arr = N.camera_array([cams[i] for i in ids])
    src = (ctypes.c_int32 * max(1, len(ids) - 1))(*ids[1:])
    common = (len(ids), ctypes.cast(cam_arr, ctypes.POINTER(N.CCamera)),
              N.ptr_array([dd[i].data_ptr() for i in ids]),
              N.ptr_array([nd[i].data_ptr() for i in ids]), ctypes.c_void_p(masks.data_ptr()),
              int(masks.shape[0]), int(ids[0]), src, int(gc), float(reproj), float(rel), cos_t,
              ctypes.c_void_p(od.data_ptr()), ctypes.c_void_p(on.data_ptr()))
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    if noalloc:
        rc = N.lib.gmvs_fuse_view_noalloc(*common, stream)
    else:
        rc = N.lib.gmvs_fuse_view(*common, ctypes.c_void_p(ws.data_ptr()), ws.numel(), stream)
    N.check("gmvs_fuse_view_noalloc" if noalloc else "gmvs_fuse_view", rc)


def _scene_maps(w, h, seed):
    from gmvs_amd import src as S

    sc = T.scene(w, h, 5, 6, seed=3407 + seed)
    ds, ns = noisy_maps(sc, np.random.default_rng(seed))
    cams = [S.load_camera(c.K.reshape(-1), c.R.reshape(-1), c.t, c.height, c.width,
                          c.depth_min, c.depth_max)._c for c in sc.cameras]
    dev = torch.device("cuda")
    dd = [torch.from_numpy(d).to(dev) for d in ds]
    nd = [torch.from_numpy(n).to(dev) for n in ns]
    return sc, cams, dd, nd


def test_noalloc_fusion_equals_workspace_fusion():
    from gmvs_amd import _native as N

    reproj, rel, gc = 1.0, 0.01, 2
    cos_t = float(np.float32(math.cos(math.radians(30))))
    dev = torch.device("cuda")
    # two calls in a row: 320x240, then a smaller 200x136 view
    for (w, h, seed) in ((320, 240, 1), (200, 136, 2)):
        sc, cams, dd, nd = _scene_maps(w, h, seed)
        nviews = len(cams)
        masks_a = torch.zeros(nviews, h, w, dtype=torch.bool, device=dev)
        masks_b = torch.zeros_like(masks_a)
        ws = torch.empty(N.lib.gmvs_fuse_workspace_bytes(h, w), dtype=torch.uint8, device=dev)
        total = 0
        for ref in range(nviews):
            ids = [ref] + sc.pairs[ref]
            oa = (torch.full((h, w), 7.0, device=dev), torch.full((h, w, 3), 7.0, device=dev))
            ob = (torch.full((h, w), -3.0, device=dev), torch.full((h, w, 3), -3.0, device=dev))
            _call(N, False, cams, ids, dd, nd, masks_a, gc, *oa, reproj, rel, cos_t, ws=ws)
            _call(N, True, cams, ids, dd, nd, masks_b, gc, *ob, reproj, rel, cos_t)
            torch.cuda.synchronize()
            for x, y, name in ((oa[0], ob[0], "depth"), (oa[1], ob[1], "normal")):
                xa, ya = x.cpu().numpy(), y.cpu().numpy()
                assert np.array_equal(xa.view(np.uint32), ya.view(np.uint32)), (w, h, ref, name)
            assert torch.equal(masks_a, masks_b), (w, h, ref)
            total += int((oa[0] > 0).sum())
        print(f"{w}x{h}: {nviews} views fused, {total} points, equal")
        assert total > 0


def test_noalloc_rejects_oversize_views():
    from gmvs_amd import _native as N

    maxpx = N.lib.gmvs_fuse_noalloc_max_pixels()
    assert maxpx == 2048 * 1536
    K = np.array([[100.0, 0, 50], [0, 100.0, 50], [0, 0, 1]], np.float32)
    from gmvs_amd import src as S

    # one pixel row more than the limit allows; rejected before any launch
    cam = S.load_camera(K.reshape(-1), np.eye(3, dtype=np.float32).reshape(-1),
                        np.zeros(3, np.float32), 1537, 2048, 1.0, 10.0)._c
    cam_arr = N.camera_array([cam])
    dummy = torch.zeros(1, device="cuda")
    ptrs = N.ptr_array([dummy.data_ptr()])
    rc = N.lib.gmvs_fuse_view_noalloc(1, ctypes.cast(cam_arr, ctypes.POINTER(N.CCamera)), ptrs,
                                      ptrs, ctypes.c_void_p(dummy.data_ptr()), 1, 0, None, 1,
                                      1.0, 0.01, 0.5, ctypes.c_void_p(dummy.data_ptr()),
                                      ctypes.c_void_p(dummy.data_ptr()), None)
    assert rc == N.ERR_ARG
    assert "static scratch limit" in N.lib.gmvs_last_error().decode()
