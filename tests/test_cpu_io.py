"""On-disk formats of the launch path, pinned by hand-written fixtures in the
layout colmap2mvs.py:531-558 writes (tests/golden/cam_0003.txt,
pair_small.txt): camera tokens 1-16 extrinsic, 18-26 K, 27 depth_min,
30 depth_max (consistency.py:13-24); pair entries with a negative score are
dropped (launch.py:22-50)."""
import os

import numpy as np
import pytest

from gmvs_amd import io as gio

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_read_camera_fixture():
    K, R, t, (dmin, dmax) = gio.read_camera(os.path.join(GOLDEN, "cam_0003.txt"))
    assert K == [300.0, 0.0, 160.0, 0.0, 301.5, 120.25, 0.0, 0.0, 1.0]
    assert np.allclose(np.array(R).reshape(3, 3) @ np.array(R).reshape(3, 3).T, np.eye(3))
    assert R[2] == pytest.approx(-0.08715574274765817)
    assert t == pytest.approx([0.3486229709906327, -0.05, 4.0152345])
    assert (dmin, dmax) == pytest.approx((2.5, 5.833333))


def test_camera_round_trip(tmp_path):
    K = np.array([[1500.0, 0, 800.5], [0, 1499.0, 600.25], [0, 0, 1]])
    a = np.radians(17.0)
    R = np.array([[np.cos(a), 0, np.sin(a)], [0, 1, 0], [-np.sin(a), 0, np.cos(a)]])
    t = np.array([0.1, -0.2, 4.5])
    p = tmp_path / "0007_cam.txt"
    gio.write_camera(str(p), K, R, t, 1.5, 7.25)
    K2, R2, t2, rng = gio.read_camera(str(p))
    assert np.allclose(np.array(K2).reshape(3, 3), K)
    assert np.allclose(np.array(R2).reshape(3, 3), R)
    assert np.allclose(t2, t)
    assert rng == pytest.approx([1.5, 7.25], abs=1e-6)


def test_read_pair_fixture_drops_negative_scores():
    pairs = gio.read_pair(os.path.join(GOLDEN, "pair_small.txt"))
    assert pairs == [(0, [1]), (1, [0, 2]), (2, [1, 0])]


def test_pair_round_trip(tmp_path):
    p = tmp_path / "pair.txt"
    gio.write_pair(str(p), [(0, [(2, 900), (1, 50)]), (1, [(0, 7)]), (2, [(0, 900)])])
    assert gio.read_pair(str(p)) == [(0, [2, 1]), (1, [0]), (2, [0])]


def test_gray_round_trip(tmp_path):
    img = np.random.default_rng(0).integers(0, 256, (37, 53)).astype(np.float32)
    p = tmp_path / "0000.png"
    gio.write_gray(str(p), img)
    assert np.array_equal(gio.read_gray(str(p)), img)


def test_install_as_gmvs_maps_the_reference_imports():
    """Reference code imports this implementation unchanged after
    gmvs_amd.install_as_gmvs() (run in a child interpreter: sys.modules)."""
    import subprocess
    import sys

    code = (
        "import gmvs_amd\n"
        "gmvs_amd.install_as_gmvs()\n"
        "import gmvs.src as _C\n"
        "from gmvs.lib.build import planar_initer\n"
        "from gmvs import run_patch_match, run_fusion, initialization, Problem, PatchMatchParams\n"
        "import gmvs.scripts.launch\n"
        "import gmvs_amd.src, gmvs_amd.lib.build.planar_initer, gmvs_amd.consistency\n"
        "assert _C is gmvs_amd.src\n"
        "assert planar_initer is gmvs_amd.lib.build.planar_initer\n"
        "assert run_patch_match is gmvs_amd.consistency.run_patch_match\n"
        "assert hasattr(_C, 'PatchMatcher') and hasattr(_C, 'Fuser')\n"
        "print('ok')\n")
    root = __import__("os").path.dirname(__import__("os").path.dirname(__import__("os").path.abspath(__file__)))
    out = subprocess.run([sys.executable, "-c", code], cwd=root, capture_output=True, text=True,
                         timeout=300)
    assert out.returncode == 0 and out.stdout.strip() == "ok", out.stderr


def test_fp16_cells_exact_and_auto_precision():
    import numpy as np
    import torch

    from gmvs_amd import _native as N
    from gmvs_amd import src as S

    rng = np.random.default_rng(0)
    u8 = [torch.from_numpy(rng.integers(0, 256, (12, 16)).astype(np.float32)) for _ in range(3)]
    assert S.fp16_cells_exact(u8)
    assert S.resolve_precision("auto", u8) == N.PREC_FP16
    frac = [u8[0] + 0.5]
    assert not S.fp16_cells_exact(frac)
    assert S.resolve_precision(N.PREC_AUTO, frac) == N.PREC_FP32
    assert not S.fp16_cells_exact([u8[0] * 8])      # beyond 1024
    assert not S.fp16_cells_exact([u8[0] - 1.0])    # negative
    nan = u8[1].clone()
    nan[3, 4] = float("nan")
    assert not S.fp16_cells_exact([u8[0], nan])
    assert S.fp16_cells_exact([torch.zeros(0), u8[2]])  # empty images are skipped
    assert S.resolve_precision("bf16", u8) == N.PREC_BF16
    with pytest.raises(ValueError):
        S.resolve_precision("fp8", u8)


def test_patch_matcher_resolves_precision_once_per_images(monkeypatch):
    import numpy as np
    import torch

    from gmvs_amd import _native as N
    from gmvs_amd import src as S

    calls = []
    real = S.resolve_precision
    monkeypatch.setattr(S, "resolve_precision", lambda p, im: calls.append(p) or real(p, im))
    u8 = torch.from_numpy(np.random.default_rng(1).integers(0, 256, (2, 12, 16)).astype(np.float32))
    pm = S.PatchMatcher()
    pm.precision = "auto"
    assert pm._resolved_precision(u8) == N.PREC_FP16
    assert pm._resolved_precision(u8) == N.PREC_FP16
    assert len(calls) == 1
    pm.precision = "fp32"                      # a new setting is resolved again
    assert pm._resolved_precision(u8) == N.PREC_FP32
    assert len(calls) == 2
    pm.precision = "auto"
    z = torch.zeros(2, 12, 16)
    pm.add_samples([], [], u8 + 0.5, z, torch.zeros(2, 12, 16, 3), z)  # new images
    assert pm._resolved_precision(pm.images_host) == N.PREC_FP32
    assert len(calls) == 3
