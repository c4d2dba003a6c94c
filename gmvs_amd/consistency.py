"""Geometric-consistency PatchMatch and fusion drivers
(gmvs/consistency.py:13-211), on the HIP boundary of ``gmvs_amd.src``.

File layout, stage order and outputs are the reference's:
``{output_dir}/{id:04}/{depth,normal,cost}[_geom].npy`` in, ``*_geom.npy``
out; fusion writes ``{output_dir}/{id:04}.npy`` as [H, W, 4] float32
(depth, unit normal) and prints the fused point count.
"""
import os
from typing import List, Optional, Tuple

import numpy as np
import torch
from tqdm import tqdm

from . import io as gio
from .primitives import Problem

read_camera = gio.read_camera


def _image_path(result_dir: str, ref_image_id: int) -> str:
    base = os.path.join(result_dir, "images", f"{ref_image_id:04}")
    for ext in (".jpg", ".png"):
        if os.path.exists(base + ext):
            return base + ext
    raise FileNotFoundError(base + ".jpg")


def run_patch_match(
    result_dir: str,
    output_dir: str,
    problems: List[Problem],
    verbose: bool = False,
    first_stage: bool = True,
    precision="fp32",
) -> Tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
    """consistency.py:27-138. ``precision`` (not in the reference): "fp32"
    (the reference's arithmetic), "fp16", "bf16" or "auto" (FP16 cells where
    they hold the images exactly: the FP32 results)."""
    from . import src as _C

    num_imgs = len(problems)
    pbar = tqdm(range(num_imgs), total=num_imgs, disable=not verbose)
    cpp_problems, cameras, gray_images, depths, normals, costs = [], [], [], [], [], []
    for i in pbar:
        problem = problems[i]
        ref_image_id = problem.ref_image_id
        camera_path = os.path.join(result_dir, "cams", f"{ref_image_id:04}_cam.txt")
        save_dir = os.path.join(output_dir, f"{ref_image_id:04}")
        suffix = "" if first_stage else "_geom"
        depth = torch.from_numpy(np.load(os.path.join(save_dir, f"depth{suffix}.npy")))
        normal = torch.from_numpy(np.load(os.path.join(save_dir, f"normal{suffix}.npy")))
        cost = torch.from_numpy(np.load(os.path.join(save_dir, f"cost{suffix}.npy")))
        gray = torch.from_numpy(gio.read_gray(_image_path(result_dir, ref_image_id)))
        height, width = depth.shape
        K, R, t, (depth_min, depth_max) = read_camera(camera_path)
        cameras.append(_C.load_camera(K, R, t, height, width, depth_min, depth_max))
        cpp_problems.append(problem._to_cpp())
        gray_images.append(gray)
        depths.append(depth)
        normals.append(normal)
        costs.append(cost)

    matcher = _C.PatchMatcher()
    matcher.precision = precision
    matcher.params.geom_consistency = True
    matcher.params.max_iterations = 2

    gray_images = torch.stack(gray_images)
    depths = torch.stack(depths)
    normals = torch.stack(normals)
    costs = torch.stack(costs)
    matcher.add_samples(cpp_problems, cameras, gray_images, depths, normals, costs)

    output_depths = torch.zeros_like(depths)
    output_normals = torch.zeros_like(normals)
    output_costs = torch.zeros_like(costs)
    process_sequence = np.arange(len(problems))
    np.random.shuffle(process_sequence)
    for i in tqdm(process_sequence, total=len(problems), disable=not verbose):
        problem = problems[i]
        ref_image_id = problem.ref_image_id
        depth, normal, cost = matcher.run_patch_match(int(i), verbose)
        save_dir = os.path.join(output_dir, f"{ref_image_id:04}")
        np.save(os.path.join(save_dir, "depth_geom.npy"), depth.cpu().numpy())
        np.save(os.path.join(save_dir, "normal_geom.npy"), normal.cpu().numpy())
        np.save(os.path.join(save_dir, "cost_geom.npy"), cost.cpu().numpy())
        output_depths[i] = depth
        output_normals[i] = normal
        output_costs[i] = cost
    return output_depths, output_normals, output_costs


def run_fusion(
    depths: torch.Tensor,
    normals: torch.Tensor,
    result_dir: str,
    problems: List[Problem],
    geom_consistent: int = 2,
    output_dir: Optional[str] = None,
    gi_iters=None,
) -> int:
    """consistency.py:141-211. Returns the fused point count (it also prints
    it, as the reference does)."""
    from . import src as _C

    num_imgs, height, width = depths.shape
    cpp_problems, cameras = [], []
    print("Begin funsion")
    for i in range(num_imgs):
        ref_image_id = problems[i].ref_image_id
        camera_path = os.path.join(result_dir, "cams", f"{ref_image_id:04}_cam.txt")
        K, R, t, (depth_min, depth_max) = read_camera(camera_path)
        cameras.append(_C.load_camera(K, R, t, height, width, depth_min, depth_max))
        cpp_problems.append(problems[i]._to_cpp())

    masks = torch.zeros([num_imgs, height, width]).bool().cuda()
    fuser = _C.Fuser()
    fuser.load_samples(cpp_problems, cameras, depths.cpu().contiguous(),
                       normals.cpu().contiguous(), masks)
    out_dir = f"{output_dir}_gi{gi_iters}" if gi_iters is not None else output_dir
    if out_dir is not None:
        os.makedirs(out_dir, exist_ok=True)

    valid = 0
    for i in range(num_imgs):
        proj_depth, proj_normal = fuser.run_fusion(i, geom_consistent)
        depth = proj_depth.cpu().numpy()
        normal = proj_normal.cpu().numpy()
        normal_norm = np.linalg.norm(normal, axis=-1, ord=2, keepdims=True)
        normal_norm = np.where(normal_norm > 0.5, normal_norm, 1.0)
        normal /= normal_norm
        valid += int((depth > 0).sum())
        depth_normal = np.concatenate([depth[..., None], normal], axis=-1).astype(np.float32)
        if out_dir is not None:
            np.save(os.path.join(out_dir, f"{int(problems[i].ref_image_id):04d}.npy"),
                    depth_normal)
    print(f"valid again: {valid}")
    return valid
