"""``Problem`` and ``PatchMatchParams`` dataclasses (gmvs/primitives.py:7-85)."""
import json
from dataclasses import dataclass
from typing import List, Type


@dataclass
class Problem:
    ref_image_id: int
    src_image_ids: List[int]

    def _to_cpp(self, planar_init: bool = False) -> Type:
        """The boundary object (gmvs_amd.src / planar_initer load_problem)."""
        if planar_init:
            from gmvs_amd.lib.build import planar_initer

            return planar_initer.load_problem(self.ref_image_id, self.src_image_ids)
        from gmvs_amd import src as _C

        return _C.load_problem(self.ref_image_id, self.src_image_ids)

    def __repr__(self) -> str:
        return json.dumps(self.__dict__, indent=4, ensure_ascii=False)


@dataclass
class PatchMatchParams:
    max_iterations: int = 3
    patch_size: int = 11
    num_images: int = 5
    max_image_size: int = 3200
    radius_increment: int = 2
    sigma_spatial: float = 5.0
    sigma_color: float = 3.0
    top_k: int = 4
    baseline: float = 0.54
    depth_min: float = 0.0
    depth_max: float = 1.0
    disparity_min: float = 0.0
    disparity_max: float = 1.0
    geom_consistency: bool = False
    multi_geometry: bool = False
    planar_prior: bool = False

    def _to_cpp(self) -> Type:
        from gmvs_amd.lib.build import planar_initer

        opt = planar_initer.PatchMatchParams()
        for k, v in self.__dict__.items():
            setattr(opt, k, v)
        return opt

    def __repr__(self) -> str:
        return json.dumps(self.__dict__, indent=4, ensure_ascii=False)
