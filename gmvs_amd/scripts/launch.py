"""``python -m gmvs_amd.scripts.launch -rd R [...]`` -- the reference CLI
(gmvs/scripts/launch.py:22-177): parse pair.txt, run the first stage per
view, ``geom_iters`` geometric-consistency passes, then fusion into
``{R}/{mvs_suffix}/depth_normal/NNNN.npy``."""
import argparse
import os
import random
import time
from typing import List

import numpy as np
from tqdm import tqdm

seed = 3407
random.seed(seed)
np.random.seed(seed)
os.environ["PYTHONHASHSEED"] = str(seed)


def generate_sample_list(result_dir: str) -> List["Problem"]:
    """launch.py:22-50."""
    from gmvs_amd import io as gio
    from gmvs_amd.primitives import Problem

    pairs = gio.read_pair(os.path.join(result_dir, "pair.txt"))
    return [Problem(ref, src) for ref, src in pairs]


def get_args(argv=None) -> argparse.Namespace:
    parser = argparse.ArgumentParser()
    parser.add_argument("--res_dir", "-rd", type=str, required=True,
                        help="root directory holding images/, cams/ and pair.txt")
    parser.add_argument("--mvs_suffix", "-ms", type=str, default="mvs",
                        help="sub-directory under the root for the mvs results")
    parser.add_argument("--input_depth_normal_dir", "-idnd", type=str, default=None,
                        help="directory with */depth.npy, */normal.npy, */cost.npy inputs")
    parser.add_argument("--dn_input", "-di", action="store_true",
                        help="initialise from the depth/normal inputs")
    parser.add_argument("--superpixel_dir", "-sd", type=str, default=None,
                        help="directory of superpixel label arrays NNNN.npy")
    parser.add_argument("--geom_iters", "-gi", type=int, default=2,
                        help="number of geometric-consistency passes")
    parser.add_argument("--geom_cons", "-gc", type=int, default=3,
                        help="consistent views required by fusion")
    parser.add_argument("--no_init", "-ni", action="store_true", help="no planar-prior pass")
    parser.add_argument("--verbose", "-v", action="store_true", help="verbose description")
    parser.add_argument("--max_iterations", type=int, default=None,
                        help="first-stage iterations (default: PatchMatchParams' 3)")
    parser.add_argument("--precision", default="fp32", choices=("fp32", "fp16", "bf16", "auto"),
                        help="source samples: fp32 (the reference's arithmetic), fp16/bf16 "
                             "cells, or auto (fp16 cells where they hold the images exactly)")
    return parser.parse_args(argv)


def main(argv=None) -> int:
    from gmvs_amd.consistency import run_fusion, run_patch_match
    from gmvs_amd.planar_init import initialization

    args = get_args(argv)
    time_tic = time.time()
    result_root = args.res_dir
    problems = generate_sample_list(result_root)
    num_imgs = len(problems)
    print(f"there are {num_imgs} problems to be processed!")

    output_dir = os.path.join(result_root, args.mvs_suffix)
    os.makedirs(output_dir, exist_ok=True)
    output_depth_normal_dir = os.path.join(output_dir, "depth_normal")
    os.makedirs(output_depth_normal_dir, exist_ok=True)

    for i in tqdm(range(num_imgs), total=num_imgs, disable=not args.verbose):
        initialization(
            result_dir=result_root,
            output_dir=output_dir,
            problem=problems[i],
            verbose=args.verbose,
            init_plane=args.dn_input,
            no_init=args.no_init,
            superpixel_dir=args.superpixel_dir,
            input_depth_normal_dir=args.input_depth_normal_dir,
            max_iterations=args.max_iterations,
            precision=args.precision,
        )

    valid = None
    for i in range(args.geom_iters):
        depths, normals, costs = run_patch_match(
            result_dir=result_root,
            output_dir=output_dir,
            problems=problems,
            verbose=args.verbose,
            first_stage=(i == 0),
            precision=args.precision,
        )
    if args.geom_iters > 0:
        valid = run_fusion(depths, normals, result_root, problems,
                           geom_consistent=int(args.geom_cons),
                           output_dir=output_depth_normal_dir)
    print("Time Used for MVS: ", time.time() - time_tic)
    return 0 if valid is None or valid >= 0 else 1


if __name__ == "__main__":
    raise SystemExit(main())
