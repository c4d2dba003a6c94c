"""CPU tests: Philox known answers (oracle and product), C-ABI exports and
the boundary structs."""
import ctypes
import json
import os
import re

import numpy as np
import pytest

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)

# Published Random123 known-answer vectors for philox4x32_10 (kat_vectors of
# the Random123 distribution): zero counter/key, all-ones, and the pi digits.
RANDOM123_KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


def _kat_cases():
    with open(os.path.join(HERE, "golden", "philox_kat.json")) as f:
        return json.load(f)["cases"]


def test_rocrand_fixture_agrees_with_published_random123():
    cases = {(tuple(c["ctr"]), tuple(c["key"])): c["out"] for c in _kat_cases()}
    for ctr, key, out in RANDOM123_KAT:
        assert cases[(tuple(ctr), tuple(key))] == out


@pytest.mark.parametrize("case", _kat_cases())
def test_oracle_philox_kat(case):
    assert O.philox(case["ctr"], case["key"]) == case["out"]


def _gmvs_philox(ctr, key):
    from gmvs_amd import _native as N

    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    N.lib.gmvs_philox4x32_10(c, k, o)
    return list(o)


@pytest.mark.parametrize("case", _kat_cases())
def test_product_philox_kat(case):
    assert _gmvs_philox(case["ctr"], case["key"]) == case["out"]


def test_uniform_mapping_range():
    # ((x >> 8) + 1) * 2^-24 spans (0, 1] exactly in FP32
    for x, want in ((0, 2.0 ** -24), (0xFFFFFFFF, 1.0), (0x80000000, 0.5 + 2.0 ** -24)):
        u = np.float32(((x >> 8) + 1) * (1.0 / 16777216.0))
        assert float(u) == want


def _header_functions():
    with open(os.path.join(REPO, "include", "gmvs_amd.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gmvs_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from gmvs_amd import _native as N

    declared = _header_functions()
    assert len(declared) >= 12
    for name in declared:
        assert hasattr(N.lib, name), name
    assert set(declared) == set(N.EXPORTS)


def test_struct_layouts_match_reference_pods():
    from gmvs_amd import _native as N

    # mvs::Camera is 100 bytes, mvs::Problem 136 (data_structure.hpp:39-53)
    assert ctypes.sizeof(N.CCamera) == 100
    assert ctypes.sizeof(N.CProblem) == 136
    assert ctypes.sizeof(N.CParams) == 13 * 4 + 4


def test_load_camera_and_problem():
    from gmvs_amd import src as S

    K = [300.0, 0, 160, 0, 300, 120, 0, 0, 1]
    R = [1.0, 0, 0, 0, 1, 0, 0, 0, 1]
    c = S.load_camera(K, R, [0.0, 0.0, 4.0], 240, 320, 2.5, 5.8)
    assert (c.height, c.width) == (240, 320)
    assert c.K == pytest.approx(K)
    assert c.depth_max == pytest.approx(5.8)
    p = S.load_problem(3, [1, 2, 5])
    assert p.ref_image_id == 3 and p.num_ngb == 3 and p.src_image_ids == [1, 2, 5]
    p.ref_image_id = 7
    assert p.ref_image_id == 7
    with pytest.raises(ValueError):
        S.load_problem(0, list(range(33)))
    with pytest.raises(ValueError):
        S.load_camera(K[:8], R, [0, 0, 0], 1, 1, 0, 1)


def test_params_defaults_match_reference():
    from gmvs_amd import src as S

    p = S.PatchMatchParams()
    # data_structure.hpp:20-37
    assert (p.max_iterations, p.patch_size, p.num_images, p.max_image_size,
            p.radius_increment) == (3, 11, 5, 3200, 2)
    assert (p.sigma_spatial, p.sigma_color, p.top_k) == (5.0, 3.0, 4)
    assert (p.geom_consistency, p.planar_prior, p.superpixel_filter) == (False, False, False)
    with pytest.raises(AttributeError):
        p.no_such_field = 1


def test_ctypes_mirrors_match_the_c_header(tmp_path):
    """Every struct the Python binding mirrors has the C layout of
    include/gmvs_amd.h: size and the offset of every field, read from a C
    program compiled against the header."""
    import os
    import shutil
    import subprocess

    from gmvs_amd import _native as N

    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    mirrors = {"gmvs_camera": N.CCamera, "gmvs_problem": N.CProblem, "gmvs_params": N.CParams,
               "gmvs_pm_args": N.CPMArgs, "gmvs_points_args": N.CPointsArgs,
               "gmvs_kernel_stat": N.CKernelStat}
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "gmvs_amd.h"', "int main(void) {"]
    for cname, cls in mirrors.items():
        lines.append(f'    printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            lines.append(f'    printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines += ["    return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    inc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    subprocess.run([cc, "-std=c99", f"-I{inc}", str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {}
    for ln in out:
        if ln:
            cname, fname, val = ln.split()
            got[(cname, fname)] = int(val)
    for cname, cls in mirrors.items():
        assert got[(cname, "size")] == ctypes.sizeof(cls), cname
        for fname, _ in cls._fields_:
            assert got[(cname, fname)] == getattr(cls, fname).offset, (cname, fname)


def test_gather_peak_rejects_bad_arguments_before_any_device_work():
    from gmvs_amd import _native as N

    out = ctypes.c_double(0.0)
    buf = ctypes.c_void_p(1)  # never dereferenced: the checks come first
    # an unknown pattern; the NCC-wave pattern in a window under 2048 cells
    assert N.lib.gmvs_gather_peak(16, 2, 4096, 10, buf, 1 << 20, ctypes.byref(out), None) == 1
    assert b"pattern" in N.lib.gmvs_last_error()
    assert N.lib.gmvs_gather_peak(16, 1, 1024, 10, buf, 1 << 20, ctypes.byref(out), None) == 1
    assert b"2048" in N.lib.gmvs_last_error()
    assert N.lib.gmvs_gather_peak(12, 0, 1024, 10, buf, 1 << 20, ctypes.byref(out), None) == 1
