"""First-stage driver (gmvs/planar_init.py:26-190): copy an input
depth/normal, run plain PatchMatch only (``no_init``), or plain PatchMatch
followed by the planar-prior pass (default)."""
import os
import shutil
from typing import Optional

import numpy as np

from . import io as gio
from .consistency import _image_path
from .primitives import Problem

read_camera = gio.read_camera


def _load_views(initer, result_dir: str, problem: Problem):
    cam_dir = os.path.join(result_dir, "cams")
    ref = problem.ref_image_id
    initer.add_image(_image_path(result_dir, ref))
    height, width = initer.get_image(0).shape
    K, R, t, (dmin, dmax) = read_camera(os.path.join(cam_dir, f"{ref:04}_cam.txt"))
    initer.add_camera(K, R, t, height, width, dmin, dmax)
    for src in problem.src_image_ids:
        initer.add_image(_image_path(result_dir, src))
        K, R, t, (dmin, dmax) = read_camera(os.path.join(cam_dir, f"{src:04}_cam.txt"))
        initer.add_camera(K, R, t, height, width, dmin, dmax)
    return height, width


def initialization(
    result_dir: str,
    output_dir: str,
    problem: Problem,
    verbose: bool = False,
    init_plane: bool = False,
    no_init: bool = False,
    superpixel_dir: Optional[str] = None,
    input_depth_normal_dir: Optional[str] = None,
    max_iterations: Optional[int] = None,
    precision="fp32",
) -> None:
    """``max_iterations`` (not in the reference) overrides the default of
    PatchMatchParams (3) for both passes; ``precision`` (not in the
    reference) as for consistency.run_patch_match."""
    ref_image_id = problem.ref_image_id
    save_dir = os.path.join(output_dir, f"{ref_image_id:04}")
    os.makedirs(save_dir, exist_ok=True)

    if init_plane:
        src_dir = os.path.join(input_depth_normal_dir, f"{ref_image_id:04}")
        for name in ("depth.npy", "normal.npy", "cost.npy"):
            shutil.copy2(os.path.join(src_dir, name), os.path.join(save_dir, name))
        return

    from .lib.build import planar_initer as _C

    initer = _C.PlanarInit(verbose)
    initer.precision = precision
    initer.rng_view = ref_image_id
    height, width = _load_views(initer, result_dir, problem)
    if not no_init and superpixel_dir:
        initer.add_superpixel(np.load(os.path.join(superpixel_dir, f"{ref_image_id:04d}.npy")))
    initer.params.depth_min = initer.get_camera(0).depth_min * 0.6
    initer.params.depth_max = initer.get_camera(0).depth_max * 1.2
    if max_iterations is not None:
        initer.params.max_iterations = max_iterations
    if verbose:
        print(f"depth range: {initer.params.depth_min} {initer.params.depth_max}")
    initer.params.num_images = initer.get_num_images()
    if verbose:
        print(f"num images: {initer.params.num_images}")
    initer.cuda_space_init()
    initer.run_patch_match()

    depth = np.zeros([height, width], dtype=np.float32)
    normal = np.zeros([height, width, 3], dtype=np.float32)
    cost = np.zeros([height, width], dtype=np.float32)
    initer.export_plane(depth, normal, cost)

    if not no_init:
        if verbose:
            print("Run Planar Prior Assisted PatchMatch MVS ...")
        initer.params.planar_prior = True
        initer.params.superpixel_filter = superpixel_dir is not None
        initer.triangulation(depth, os.path.join(save_dir, "triangulation.png"))
        initer.cuda_planar_prior_init()
        initer.run_patch_match()
        initer.export_plane(depth, normal, cost)

    np.save(os.path.join(save_dir, "depth.npy"), depth)
    np.save(os.path.join(save_dir, "normal.npy"), normal)
    np.save(os.path.join(save_dir, "cost.npy"), cost)
