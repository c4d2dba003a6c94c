"""Seeded synthetic multi-view scenes for tests and benchmarks.

A scene is piecewise planar as seen from view 0: the image plane of view 0 is
partitioned into Voronoi cells, cell i shows plane i, and the plane patch is
the back-projection of the cell onto the plane. Planes have normals drawn
uniformly on a spherical cap facing view 0 and pass through the point at a
uniform random depth on the ray of their cell seed. Every plane carries a
4-octave value-noise texture in plane-local coordinates (rendered to uint8
grey values); optional Gaussian bumps displace the surface along the plane
normal. Other views are rendered by ray casting against all patches (nearest
hit wins) with a textured background plane behind the scene.

Cameras are pinhole (fx = fy, principal point at the image centre) on an arc
(or a grid of yaw/pitch angles) of radius ``radius`` looking at the origin;
view 0 sits at yaw 0, at (0, 0, -radius). ``X_cam = R X_world + t``.

Everything is torch so that large scenes can be generated on the GPU.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np
import torch

_M32 = 0xFFFFFFFF


@dataclass
class Camera:
    K: np.ndarray  # [3,3]
    R: np.ndarray  # [3,3]
    t: np.ndarray  # [3]
    height: int
    width: int
    depth_min: float
    depth_max: float

    @property
    def center(self) -> np.ndarray:
        return -self.R.T @ self.t


@dataclass
class Scene:
    images: torch.Tensor           # [N,H,W] float32 holding uint8 values
    cameras: List[Camera]
    depth: torch.Tensor            # [N,H,W] ground-truth depth (0 = background)
    normal: torch.Tensor           # [N,H,W,3] ground-truth world normals
    labels: Optional[torch.Tensor] = None  # [H,W] int32 superpixels of view 0
    planes: np.ndarray = field(default_factory=lambda: np.zeros((0, 4)))  # world (n, d)
    pairs: List[List[int]] = field(default_factory=list)  # source ids per view


def look_at(center: np.ndarray, target: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
    z = target - center
    z = z / np.linalg.norm(z)
    down = np.array([0.0, 1.0, 0.0])
    x = np.cross(down, z)
    x = x / np.linalg.norm(x)
    y = np.cross(z, x)
    R = np.stack([x, y, z], 0)
    t = -R @ center
    return R, t


def arc_cameras(num_views: int, width: int, height: int, fx: float, arc_deg: float = 30.0,
                radius: float = 4.0, pitch_deg: Sequence[float] = (0.0,),
                depth_min: float = 2.5, depth_max: float = 35.0 / 6.0) -> List[Camera]:
    """Cameras on yaw arcs (one per pitch), view 0 at yaw 0 / first pitch."""
    K = np.array([[fx, 0.0, width / 2.0], [0.0, fx, height / 2.0], [0.0, 0.0, 1.0]])
    per_row = int(math.ceil(num_views / len(pitch_deg)))
    yaws = np.linspace(-arc_deg / 2, arc_deg / 2, per_row) if per_row > 1 else np.zeros(1)
    # view 0 at yaw 0: order yaws by |yaw|
    yaws = sorted(yaws, key=lambda a: (abs(a), a))
    cams = []
    for p in pitch_deg:
        for yv in yaws:
            if len(cams) == num_views:
                break
            a, b = math.radians(yv), math.radians(p)
            c = radius * np.array([math.sin(a) * math.cos(b), math.sin(b), -math.cos(a) * math.cos(b)])
            R, t = look_at(c, np.zeros(3))
            cams.append(Camera(K.copy(), R, t, height, width, depth_min, depth_max))
    return cams


def _hash2(ix: torch.Tensor, iy: torch.Tensor, seed: int) -> torch.Tensor:
    h = (ix * 374761393 + iy * 668265263 + seed * 1442695041) & _M32
    h = ((h ^ (h >> 13)) * 1274126177) & _M32
    h = h ^ (h >> 16)
    return h.to(torch.float32) * (1.0 / 4294967296.0)


def value_noise(a: torch.Tensor, b: torch.Tensor, seed: int, octaves: int = 4) -> torch.Tensor:
    """Sum of octaves of smoothstep-interpolated lattice noise, in [0, 1)."""
    out = torch.zeros_like(a)
    amp, norm = 1.0, 0.0
    for o in range(octaves):
        f = float(2 ** o)
        x, y = a * f, b * f
        x0, y0 = torch.floor(x), torch.floor(y)
        fx, fy = x - x0, y - y0
        sx, sy = fx * fx * (3 - 2 * fx), fy * fy * (3 - 2 * fy)
        ix, iy = x0.to(torch.int64), y0.to(torch.int64)
        s = seed * 131 + o
        v00, v10 = _hash2(ix, iy, s), _hash2(ix + 1, iy, s)
        v01, v11 = _hash2(ix, iy + 1, s), _hash2(ix + 1, iy + 1, s)
        v = (v00 * (1 - sx) + v10 * sx) * (1 - sy) + (v01 * (1 - sx) + v11 * sx) * sy
        out = out + amp * v
        norm += amp
        amp *= 0.5
    return out / norm


def _pixel_rays(cam: Camera, device) -> torch.Tensor:
    """World-frame ray directions [H,W,3] (not normalised, z_cam = 1)."""
    H, W = cam.height, cam.width
    K = torch.tensor(cam.K, dtype=torch.float64, device=device)
    R = torch.tensor(cam.R, dtype=torch.float64, device=device)
    ys, xs = torch.meshgrid(torch.arange(H, device=device, dtype=torch.float64),
                            torch.arange(W, device=device, dtype=torch.float64), indexing="ij")
    dc = torch.stack([(xs - K[0, 2]) / K[0, 0], (ys - K[1, 2]) / K[1, 1], torch.ones_like(xs)], -1)
    return dc @ R  # R^T d


def make_scene(width: int, height: int, num_views: int, num_planes: int, seed: int,
               fx: Optional[float] = None, arc_deg: float = 30.0, radius: float = 4.0,
               pitch_deg: Sequence[float] = (0.0,), depth_range: Tuple[float, float] = (2.0, 6.0),
               bump_fraction: float = 0.0, num_labels: int = 0, cap_deg: float = 60.0,
               feature_px: float = 6.0, device: str = "cpu",
               cam_depth_range: Tuple[float, float] = (2.5, 35.0 / 6.0),
               depth_band: Optional[Tuple[float, float]] = None) -> Scene:
    gen = np.random.default_rng(seed)
    if fx is None:
        fx = 300.0 * width / 320.0
    cams = arc_cameras(num_views, width, height, fx, arc_deg, radius, pitch_deg,
                       cam_depth_range[0], cam_depth_range[1])
    dev = torch.device(device)
    c0 = cams[0]

    # Voronoi seeds in view 0 and planes through them
    seeds = np.stack([gen.uniform(0, width, num_planes), gen.uniform(0, height, num_planes)], 1)
    ys, xs = torch.meshgrid(torch.arange(height, device=dev, dtype=torch.float32),
                            torch.arange(width, device=dev, dtype=torch.float32), indexing="ij")
    sd = torch.tensor(seeds, dtype=torch.float32, device=dev)
    owner0 = torch.zeros(height, width, dtype=torch.int64, device=dev)
    best = torch.full((height, width), float("inf"), device=dev)
    for i in range(num_planes):
        d2 = (xs - sd[i, 0]) ** 2 + (ys - sd[i, 1]) ** 2
        m = d2 < best
        owner0 = torch.where(m, torch.full_like(owner0, i), owner0)
        best = torch.minimum(best, d2)

    planes = np.zeros((num_planes, 4))
    cosmax = math.cos(math.radians(cap_deg))
    C0 = c0.center
    if depth_band is not None:
        # camera-0 rays of every pixel, for the depth of a plane over its cell
        rays0 = np.stack([(xs.cpu().numpy().astype(np.float64) - c0.K[0, 2]) / c0.K[0, 0],
                          (ys.cpu().numpy().astype(np.float64) - c0.K[1, 2]) / c0.K[1, 1],
                          np.ones((height, width))], -1)
        own_np = owner0.cpu().numpy()
    for i in range(num_planes):
        ray = np.array([(seeds[i, 0] - c0.K[0, 2]) / c0.K[0, 0], (seeds[i, 1] - c0.K[1, 2]) / c0.K[1, 1], 1.0])
        for _ in range(10000):
            # uniform on the cap around -z of camera 0 (facing the camera);
            # cap_deg = 90 is the whole facing hemisphere
            u = gen.uniform(cosmax, 1.0)
            phi = gen.uniform(0, 2 * math.pi)
            s = math.sqrt(1 - u * u)
            n_cam = np.array([s * math.cos(phi), s * math.sin(phi), -u])
            depth = gen.uniform(*depth_range)
            Xc = ray * depth
            if depth_band is None:
                break
            # with a depth band, a plane is drawn again until its depth over
            # its Voronoi cell stays inside the band (a plane close to
            # grazing leaves any band, or the half-space in front of the
            # camera, within its cell): the normals are uniform on the
            # hemisphere given that condition
            den = rays0[own_np == i] @ n_cam
            dcell = (n_cam @ Xc) / np.where(den == 0.0, np.nan, den)
            if np.all(np.isfinite(dcell)) and dcell.min() >= depth_band[0] and dcell.max() <= depth_band[1]:
                break
        else:
            raise RuntimeError(f"plane {i}: no normal keeps its cell inside depth_band {depth_band}")
        n_w = c0.R.T @ n_cam
        Xw = c0.R.T @ Xc + C0
        planes[i, :3] = n_w
        planes[i, 3] = -n_w @ Xw
    # background plane behind everything, facing view 0
    zb = depth_range[1] * 1.6
    bg = np.concatenate([c0.R.T @ np.array([0, 0, -1.0]), [0.0]])
    bg[3] = -bg[:3] @ (c0.R.T @ np.array([0, 0, zb]) + C0)

    # Gaussian bumps: world centres on random view-0 surface points
    n_bumps = 0
    bumps = []
    if bump_fraction > 0:
        area = width * height * bump_fraction
        r_px = max(6.0, 0.04 * min(width, height))
        n_bumps = max(1, int(round(area / (math.pi * r_px * r_px))))
        for _ in range(n_bumps):
            px, py = gen.uniform(0, width), gen.uniform(0, height)
            i = int(owner0[int(py), int(px)].item())
            ray = np.array([(px - c0.K[0, 2]) / c0.K[0, 0], (py - c0.K[1, 2]) / c0.K[1, 1], 1.0])
            dw = c0.R.T @ ray
            tt = -(planes[i, :3] @ C0 + planes[i, 3]) / (planes[i, :3] @ dw)
            Xw = C0 + tt * dw
            sigma = r_px / c0.K[0, 0] * tt * 0.6
            amp = gen.uniform(-1, 1) * sigma * 0.8
            bumps.append((Xw, sigma, amp, i))

    # plane-local texture frames
    frames = []
    for i in range(num_planes + 1):
        n = planes[i, :3] if i < num_planes else bg[:3]
        a = np.array([1.0, 0.0, 0.0]) if abs(n[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
        e1 = np.cross(n, a)
        e1 /= np.linalg.norm(e1)
        e2 = np.cross(n, e1)
        frames.append((e1, e2))
    mean_depth = 0.5 * (depth_range[0] + depth_range[1])
    freq = fx / (mean_depth * feature_px * 2.0)

    P = torch.tensor(planes, dtype=torch.float64, device=dev)
    images, depths, normals = [], [], []
    for cam in cams:
        C = torch.tensor(cam.center, dtype=torch.float64, device=dev)
        d = _pixel_rays(cam, dev)  # [H,W,3]
        zcam = torch.tensor(cam.R[2], dtype=torch.float64, device=dev)
        best_t = torch.full((height, width), float("inf"), dtype=torch.float64, device=dev)
        best_i = torch.full((height, width), -1, dtype=torch.int64, device=dev)
        K0 = torch.tensor(c0.K, dtype=torch.float64, device=dev)
        R0 = torch.tensor(c0.R, dtype=torch.float64, device=dev)
        t0 = torch.tensor(c0.t, dtype=torch.float64, device=dev)
        for i in range(num_planes):
            n, dd = P[i, :3], P[i, 3]
            den = d @ n
            tt = -(C @ n + dd) / den
            X = C + tt[..., None] * d
            Xc0 = X @ R0.T + t0
            u = K0[0, 0] * Xc0[..., 0] / Xc0[..., 2] + K0[0, 2]
            v = K0[1, 1] * Xc0[..., 1] / Xc0[..., 2] + K0[1, 2]
            ui = torch.round(u).long()
            vi = torch.round(v).long()
            inb = (ui >= 0) & (ui < width) & (vi >= 0) & (vi < height) & (Xc0[..., 2] > 0) & (tt > 0)
            own = torch.full_like(ui, -2)
            own[inb] = owner0[vi[inb], ui[inb]]
            hit = inb & (own == i) & (tt < best_t)
            best_t = torch.where(hit, tt, best_t)
            best_i = torch.where(hit, torch.full_like(best_i, i), best_i)
        # background
        nb = torch.tensor(bg[:3], dtype=torch.float64, device=dev)
        tb = -(C @ nb + float(bg[3])) / (d @ nb)
        is_bg = best_i < 0
        best_t = torch.where(is_bg, tb, best_t)
        X = C + best_t[..., None] * d
        # bumps: move along the ray by the normal displacement
        if bumps:
            for (Xb, sg, amp, pi) in bumps:
                Xb_t = torch.tensor(Xb, dtype=torch.float64, device=dev)
                r2 = ((X - Xb_t) ** 2).sum(-1)
                hgt = amp * torch.exp(-r2 / (2 * sg * sg))
                sel = best_i == pi
                n = P[pi, :3]
                dt = hgt / (-(d @ n)).clamp(min=1e-3)
                best_t = torch.where(sel, best_t - dt, best_t)
            X = C + best_t[..., None] * d
        tex = torch.zeros(height, width, dtype=torch.float32, device=dev)
        for i in range(num_planes + 1):
            m = (best_i == i) if i < num_planes else is_bg
            if not bool(m.any()):
                continue
            e1 = torch.tensor(frames[i][0], dtype=torch.float64, device=dev)
            e2 = torch.tensor(frames[i][1], dtype=torch.float64, device=dev)
            a = (X[m] @ e1) * freq
            b = (X[m] @ e2) * freq
            tex[m] = value_noise(a.float() + 17.0 * i, b.float() - 5.0 * i, seed + 7 * i)
        img = torch.clamp(torch.round(tex * 255.0), 0, 255)
        depth = best_t * (d @ zcam)  # z in this camera
        depth = torch.where(is_bg, torch.zeros_like(depth), depth)
        nw = torch.zeros(height, width, 3, dtype=torch.float64, device=dev)
        for i in range(num_planes):
            m = best_i == i
            nw[m] = P[i, :3]
        images.append(img)
        depths.append(depth.float())
        normals.append(nw.float())
    labels = None
    if num_labels > 0:
        lx = torch.tensor(gen.uniform(0, width, num_labels), dtype=torch.float32, device=dev)
        ly = torch.tensor(gen.uniform(0, height, num_labels), dtype=torch.float32, device=dev)
        labels = torch.zeros(height, width, dtype=torch.int32, device=dev)
        bestl = torch.full((height, width), float("inf"), device=dev)
        for i in range(num_labels):
            d2 = (xs - lx[i]) ** 2 + (ys - ly[i]) ** 2
            m = d2 < bestl
            labels = torch.where(m, torch.full_like(labels, i + 1), labels)
            bestl = torch.minimum(bestl, d2)
    pairs = pair_sources(cams, min(num_views - 1, 31))
    return Scene(torch.stack(images), cams, torch.stack(depths), torch.stack(normals), labels,
                 planes, pairs)


def pair_sources(cams: List[Camera], max_src: int) -> List[List[int]]:
    """Per view, the other views ordered by the angle between optical axes."""
    out = []
    for i, ci in enumerate(cams):
        zi = ci.R[2]
        order = sorted((j for j in range(len(cams)) if j != i),
                       key=lambda j: (-float(zi @ cams[j].R[2]), j))
        out.append(order[:max_src])
    return out


def voronoi_labels(width: int, height: int, cells: int, seed: int) -> np.ndarray:
    """int32 labels 1..cells of a seeded Voronoi partition of the image."""
    from scipy.spatial import cKDTree

    gen = np.random.default_rng(seed)
    pts = np.stack([gen.uniform(0, width, cells), gen.uniform(0, height, cells)], 1)
    ys, xs = np.mgrid[0:height, 0:width]
    q = np.stack([xs.ravel(), ys.ravel()], 1).astype(np.float64)
    _, idx = cKDTree(pts).query(q, k=1)
    return (idx.astype(np.int32) + 1).reshape(height, width)
