"""End to end through the reference's CLI (gmvs/scripts/launch.py:53-177) on a
dataset written in the reference's on-disk layout: images/NNNN.png,
cams/NNNN_cam.txt, pair.txt and superpixel NNNN.npy in; per-view
mvs/NNNN/{depth,normal,cost}[_geom].npy and the fused
mvs/depth_normal/NNNN.npy ([H, W, 4] float32: depth, world normal) out."""
import os

import numpy as np
import pytest

from gmvs_amd import io as gio
from gmvs_amd.synthetic import make_scene, voronoi_labels

pytestmark = pytest.mark.gpu

W, H, V = 160, 120, 5


def _write_dataset(root, sp_dir):
    sc = make_scene(W, H, V, 4, seed=3407 + 31)
    os.makedirs(os.path.join(root, "images"))
    os.makedirs(os.path.join(root, "cams"))
    os.makedirs(sp_dir)
    for i, cam in enumerate(sc.cameras):
        gio.write_gray(os.path.join(root, "images", f"{i:04}.png"), sc.images[i].numpy())
        gio.write_camera(os.path.join(root, "cams", f"{i:04}_cam.txt"), cam.K, cam.R, cam.t,
                         cam.depth_min, cam.depth_max)
        np.save(os.path.join(sp_dir, f"{i:04}.npy"), voronoi_labels(W, H, 30, seed=3407 + 40 + i))
    gio.write_pair(os.path.join(root, "pair.txt"),
                   [(i, [(s, 100.0 - k) for k, s in enumerate(sc.pairs[i][:4])]) for i in range(V)])
    return sc


def test_launch_end_to_end(tmp_path):
    from gmvs_amd.scripts import launch

    root, sp = str(tmp_path / "scene"), str(tmp_path / "superpixels")
    sc = _write_dataset(root, sp)
    rc = launch.main(["-rd", root, "-sd", sp, "--max_iterations", "2", "-gi", "1", "-gc", "2"])
    assert rc == 0
    fused_any = 0
    for i in range(V):
        d = os.path.join(root, "mvs", f"{i:04}")
        for name in ("depth", "normal", "cost", "depth_geom", "normal_geom", "cost_geom"):
            assert os.path.exists(os.path.join(d, name + ".npy")), (i, name)
        depth = np.load(os.path.join(d, "depth_geom.npy"))
        assert depth.shape == (H, W) and np.isfinite(depth).all()
        gt = sc.depth[i].numpy()
        fg = gt > 0
        rel = np.abs(depth[fg] - gt[fg]) / gt[fg]
        assert np.mean(rel < 0.02) > 0.5, i   # the geometry is recovered on most pixels
        dn = np.load(os.path.join(root, "mvs", "depth_normal", f"{i:04}.npy"))
        assert dn.shape == (H, W, 4) and dn.dtype == np.float32
        kept = dn[..., 0] > 0
        fused_any += int(kept.sum())
        if kept.any():
            assert np.allclose(np.linalg.norm(dn[kept][:, 1:], axis=1), 1.0, atol=1e-3)
            ok = kept & fg
            rel = np.abs(dn[..., 0][ok] - gt[ok]) / gt[ok]
            assert np.mean(rel < 0.02) > 0.8, i
    assert fused_any > 0.3 * H * W


def test_launch_auto_precision_writes_the_fp32_files(tmp_path):
    """--precision auto on 8-bit PNGs (FP16 cells, which hold them exactly)
    writes the same stage and fused files, bit for bit, as the default FP32
    run."""
    from gmvs_amd.scripts import launch

    outs = {}
    for prec in ("fp32", "auto"):
        root, sp = str(tmp_path / prec / "scene"), str(tmp_path / prec / "superpixels")
        _write_dataset(root, sp)
        rc = launch.main(["-rd", root, "-sd", sp, "--max_iterations", "2", "-gi", "1", "-gc", "2",
                          "--precision", prec])
        assert rc == 0
        files = {}
        for i in range(V):
            d = os.path.join(root, "mvs", f"{i:04}")
            for name in ("depth", "normal", "cost", "depth_geom", "normal_geom", "cost_geom"):
                files[(i, name)] = np.load(os.path.join(d, name + ".npy"))
            files[(i, "fused")] = np.load(os.path.join(root, "mvs", "depth_normal", f"{i:04}.npy"))
        outs[prec] = files
    for key, a in outs["fp32"].items():
        b = outs["auto"][key]
        assert a.dtype == b.dtype and np.array_equal(a.view(np.uint8), b.view(np.uint8)), key
