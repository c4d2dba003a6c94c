"""``python -m gmvs_amd.scripts.gather_points --depth_normal_dir D --data_dir
A --save_dir S`` -- the point-gather stage of
gmvs/scripts/mvs_fusion_segmentaion_cpu.py:113-216 on the GPU.

Inputs as the reference reads them: ``{A}/pose/*.txt`` (camera-to-world 4x4,
sorted by number), ``{A}/intrinsic.txt`` and ``{D}/*.npy`` fused [H, W, 4]
depth_normal maps. Outputs as it writes them:
``{S}/depth_normal_filter/{name}.npy`` (the maps where the 70-degree
view-angle filter passes), ``{S}/mvs_depthnorm.xyz`` and
``{S}/mvs_depthnorm_filter.xyz`` (``%.4f`` text, one point + normal a line).
The Poisson meshing, ray casting and mask rendering that follow in the
reference (open3d, prim3d) are not part of this package.
"""
from __future__ import annotations

import argparse
import os
import time

import numpy as np


def get_args(argv=None) -> argparse.Namespace:
    parser = argparse.ArgumentParser()
    parser.add_argument("--depth_normal_dir", type=str, required=True,
                        help="directory of the fused NNNN.npy [H, W, 4] maps")
    parser.add_argument("--data_dir", type=str, required=True,
                        help="directory holding pose/ and intrinsic.txt")
    parser.add_argument("--save_dir", "-sd", type=str, required=True,
                        help="directory for depth_normal_filter/ and the .xyz files")
    parser.add_argument("--height", type=int, default=480, help="the height of input images")
    parser.add_argument("--width", type=int, default=640, help="the width of input images")
    return parser.parse_args(argv)


def main(argv=None) -> int:
    from gmvs_amd import post

    args = get_args(argv)
    tic = time.time()
    os.makedirs(os.path.join(args.save_dir, "depth_normal_filter"), exist_ok=True)
    pose_files = sorted(os.listdir(os.path.join(args.data_dir, "pose")),
                        key=lambda x: int(x.split(".")[0]))
    dn_files = sorted(os.listdir(args.depth_normal_dir), key=lambda x: int(x.split(".")[0]))
    if len(pose_files) != len(dn_files):
        raise ValueError(f"{len(pose_files)} poses but {len(dn_files)} depth_normal maps")
    intrinsic = np.loadtxt(os.path.join(args.data_dir, "intrinsic.txt"))
    poses = [np.loadtxt(os.path.join(args.data_dir, "pose", f)) for f in pose_files]
    maps = [np.load(os.path.join(args.depth_normal_dir, f)) for f in dn_files]
    for m in maps:
        if m.shape != (args.height, args.width, 4):
            raise ValueError(f"depth_normal map of shape {m.shape}, expected "
                             f"({args.height}, {args.width}, 4)")
    pts, pts_f, filtered = post.gather_points(maps, poses, intrinsic)
    filtered = filtered.cpu().numpy()
    for f, m in zip(pose_files, filtered):
        np.save(os.path.join(args.save_dir, "depth_normal_filter", f"{f[:-4]}.npy"), m)
    post.savetxt(os.path.join(args.save_dir, "mvs_depthnorm.xyz"), pts)
    post.savetxt(os.path.join(args.save_dir, "mvs_depthnorm_filter.xyz"), pts_f)
    print(f"{len(pts)} points, {len(pts_f)} after the view-angle filter "
          f"({time.time() - tic:.2f} s)")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
