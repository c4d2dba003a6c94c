"""Image, camera and pair-file I/O in the reference's on-disk layout.

* ``images/NNNN.jpg`` (any format PIL decodes), read as 8-bit grey and
  returned as float32 -- cv2.imread(path, IMREAD_GRAYSCALE) + convertTo(CV_32F)
  in the reference (gmvs/lib/src/planar_initer.cu:64-70, consistency.py:69-72).
  OpenCV is not used: colour images are converted with ITU-R 601-2 luma
  weights (PIL "L"), which can differ from OpenCV's fixed-point rounding by
  one grey level; lossless grey images read identically.
* ``cams/NNNN_cam.txt`` as written by colmap2mvs.py:531-550, parsed as
  read_camera does (consistency.py:13-24).
* ``pair.txt`` as written by colmap2mvs.py:551-558, parsed as
  generate_sample_list does (launch.py:22-50).
"""
from __future__ import annotations

import os
from typing import List, Sequence, Tuple

import numpy as np


def read_gray(path: str) -> np.ndarray:
    if not os.path.exists(path):
        raise FileNotFoundError(path)
    from PIL import Image

    with Image.open(path) as im:
        if im.mode != "L":
            im = im.convert("L")
        return np.asarray(im, dtype=np.uint8).astype(np.float32)


def write_gray(path: str, img: np.ndarray) -> None:
    from PIL import Image

    Image.fromarray(np.clip(np.rint(img), 0, 255).astype(np.uint8), mode="L").save(path)


def read_camera(cam_path: str) -> Tuple[List[float], List[float], List[float], List[float]]:
    """consistency.py:13-24: tokens 1..16 extrinsic 4x4, 18..26 K,
    27 depth_min, 30 depth_max."""
    with open(cam_path, "r") as f:
        token = f.read().split()
    extrins = np.array([float(t) for t in token[1:17]]).reshape(4, 4)
    R = extrins[:3, :3].reshape(9).tolist()
    t = extrins[:3, 3].tolist()
    K = [float(t_) for t_ in token[18:27]]
    return K, R, t, [float(token[27]), float(token[30])]


def write_camera(cam_path: str, K: np.ndarray, R: np.ndarray, t: np.ndarray, depth_min: float,
                 depth_max: float, num_depth: int = 192) -> None:
    """colmap2mvs.py:531-550 layout (depth interval and count in between)."""
    E = np.eye(4)
    E[:3, :3] = R
    E[:3, 3] = t
    with open(cam_path, "w") as f:
        f.write("extrinsic\n")
        for j in range(4):
            f.write(" ".join(repr(float(E[j, k])) for k in range(4)) + " \n")
        f.write("\nintrinsic\n")
        for j in range(3):
            f.write(" ".join(repr(float(K[j, k])) for k in range(3)) + " \n")
        interval = (depth_max - depth_min) / max(1, num_depth - 1)
        f.write("\n%f %f %f %f\n" % (depth_min, interval, num_depth, depth_max))


def read_pair(pair_path: str) -> List[Tuple[int, List[int]]]:
    """launch.py:22-50: returns (ref_id, src_ids) with negative scores dropped."""
    assert os.path.exists(pair_path), f"{pair_path} does not exist"
    with open(pair_path, "r") as f:
        lines = f.read().split("\n")
    out = []
    n = int(lines[0])
    for i in range(n):
        ref_id = int(lines[i * 2 + 1])
        src_line = lines[i * 2 + 2].split(" ")
        num_src = int(src_line[0])
        src = []
        for j in range(num_src):
            sid = int(src_line[j * 2 + 1])
            score = float(src_line[j * 2 + 2])
            if score < 0.0:
                continue
            src.append(sid)
        out.append((ref_id, src))
    return out


def write_pair(pair_path: str, pairs: Sequence[Tuple[int, Sequence[Tuple[int, float]]]]) -> None:
    """colmap2mvs.py:551-558 layout."""
    with open(pair_path, "w") as f:
        f.write("%d\n" % len(pairs))
        for ref, srcs in pairs:
            f.write("%d\n%d " % (ref, len(srcs)))
            for sid, s in srcs:
                f.write("%d %d " % (sid, s))
            f.write("\n")
