"""Shared test helpers: synthetic scenes in oracle/product form."""
from __future__ import annotations

import functools
import math

import numpy as np

from gmvs_amd.synthetic import make_scene
from oracle import oracle as O

SEED = 3407


@functools.lru_cache(maxsize=8)
def scene(width=320, height=240, views=5, planes=6, seed=SEED + 1, labels=0, bump=0.0):
    # plane normals uniform on a 60-degree cap around the direction facing
    # camera 0 (the suite's thresholds and known answers are set on it)
    return make_scene(width, height, views, planes, seed=seed, num_labels=labels,
                      bump_fraction=bump)


@functools.lru_cache(maxsize=2)
def hemisphere_scene(seed=SEED + 1):
    """The CPU-runnable configuration as written: 6 planes whose normals are
    uniform on the hemisphere facing camera 0, depths U[2, 6] at their seeds,
    5 views at 320x240 (fx = fy = 300) on a 30-degree arc. A plane close to
    grazing leaves any depth range within its Voronoi cell, so each plane is
    drawn again until its depth over the cell stays inside the search range
    [1.5, 7]: the normals are uniform on the hemisphere given that."""
    return make_scene(320, 240, 5, 6, seed=seed, cap_deg=90.0, depth_band=(1.5, 7.0))


def oracle_cams(sc, ids):
    return [O.make_camera(sc.cameras[i].K, sc.cameras[i].R, sc.cameras[i].t, sc.cameras[i].height,
                          sc.cameras[i].width, sc.cameras[i].depth_min, sc.cameras[i].depth_max)
            for i in ids]


def depth_range(sc, ref):
    c = sc.cameras[ref]
    dmin = float(np.float32(float(np.float32(c.depth_min)) * 0.6))
    dmax = float(np.float32(float(np.float32(c.depth_max)) * 1.2))
    return dmin, dmax


def random_planes(sc, ref, n, rng, dmin, dmax):
    """Random (pixel, camera-frame plane) queries facing the camera."""
    cam = sc.cameras[ref]
    H, W = cam.height, cam.width
    xs = rng.integers(0, W, n)
    ys = rng.integers(0, H, n)
    depth = rng.uniform(dmin, dmax, n)
    # normals within 60 degrees of -view ray
    out = np.zeros((n, 4), np.float32)
    K = cam.K
    for i in range(n):
        ray = np.array([(xs[i] - K[0, 2]) / K[0, 0], (ys[i] - K[1, 2]) / K[1, 1], 1.0])
        ray /= np.linalg.norm(ray)
        u = rng.uniform(math.cos(math.radians(60)), 1.0)
        phi = rng.uniform(0, 2 * math.pi)
        s = math.sqrt(1 - u * u)
        a = np.array([1.0, 0, 0]) if abs(ray[0]) < 0.9 else np.array([0, 1.0, 0])
        e1 = np.cross(ray, a)
        e1 /= np.linalg.norm(e1)
        e2 = np.cross(ray, e1)
        n_ = -u * ray + s * (math.cos(phi) * e1 + math.sin(phi) * e2)
        X = np.array([depth[i] * (xs[i] - K[0, 2]) / K[0, 0], depth[i] * (ys[i] - K[1, 2]) / K[1, 1], depth[i]])
        out[i, :3] = n_
        out[i, 3] = -n_ @ X
    return np.stack([xs, ys], 1).astype(np.int32), out


def gt_planes(sc, ref, xy):
    """Camera-frame planes of the ground truth at pixels xy."""
    cam = sc.cameras[ref]
    K, R = cam.K, cam.R
    out = np.zeros((len(xy), 4), np.float32)
    d = sc.depth[ref].numpy()
    nw = sc.normal[ref].numpy()
    for i, (x, y) in enumerate(xy):
        n = R @ nw[y, x]
        z = d[y, x]
        X = np.array([z * (x - K[0, 2]) / K[0, 0], z * (y - K[1, 2]) / K[1, 1], z])
        out[i, :3] = n
        out[i, 3] = -n @ X
    return out


def angle_deg(n1, n2):
    n1 = n1 / np.maximum(np.linalg.norm(n1, axis=-1, keepdims=True), 1e-12)
    n2 = n2 / np.maximum(np.linalg.norm(n2, axis=-1, keepdims=True), 1e-12)
    c = np.clip((n1 * n2).sum(-1), -1.0, 1.0)
    return np.degrees(np.arccos(c))


def depth_normal_agreement(d_a, n_a, d_b, n_b, rel_tol=0.005, ang_tol=1.0):
    """Fraction of pixels whose depth agrees within rel_tol and normal within ang_tol degrees."""
    valid = np.isfinite(d_b) & (np.abs(d_b) > 0)
    rel = np.abs(d_a - d_b) / np.where(valid, np.abs(d_b), 1.0)
    ang = angle_deg(n_a, n_b)
    ok = valid & (rel <= rel_tol) & (ang <= ang_tol)
    both_nan = ~np.isfinite(d_a) & ~np.isfinite(d_b)
    return float((ok | both_nan).mean())


def world_state_facing_camera(sc, ref, seed=0):
    """A [H, W, 4] (world normal, depth) state as the first pass exports it:
    the ground-truth depth and random unit normals within ~30 degrees of the
    reversed viewing ray (well conditioned for the depth <-> plane round
    trip), rotated into the world frame (n_w = R^T n_c)."""
    cam = sc.cameras[ref]
    K = np.asarray(cam.K, np.float64)
    R = np.asarray(cam.R, np.float64)
    H, W = cam.height, cam.width
    y, x = np.mgrid[0:H, 0:W].astype(np.float64)
    ray = np.stack([(x - K[0, 2]) / K[0, 0], (y - K[1, 2]) / K[1, 1], np.ones_like(x)], -1)
    ray /= np.linalg.norm(ray, axis=-1, keepdims=True)
    rng = np.random.default_rng(seed)
    n_c = -ray + 0.3 * rng.standard_normal((H, W, 3))
    n_c /= np.linalg.norm(n_c, axis=-1, keepdims=True)
    n_w = n_c @ R  # row form of R^T n_c
    depth = sc.depth[ref].numpy().astype(np.float64)
    return np.concatenate([n_w, depth[..., None]], -1).astype(np.float32)
