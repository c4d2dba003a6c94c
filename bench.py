"""Benchmark of the PatchMatch MVS hot path on one MI355X.

One step = the full ACMP pipeline over the synthetic 10-view 1600x1200 scene
(24 planes + Gaussian bumps, 2000-cell Voronoi superpixels): for every
reference view a plain PatchMatch pass and a planar-prior pass (8 iterations
each, top-4 initial cost, host prior construction overlapped with the GPU),
then one geometric-consistency pass (2 iterations) over all views, then
fusion of all views (>= 3 consistent views). Inputs are resident in HBM
before the timed region starts.

Prints one JSON line: value = pixel-hypotheses per second (Mpix-hyp/s; a
pixel-hypothesis is one plane costed against all V source views) of the FP32
path with each stage's views batched into one schedule (bit-identical to
running them one by one), plus the same workload view by view in FP32 (the
plain path) and batched with FP16/BF16 samples, the fused Mpts/s of the
49-view fusion workload, the roofline of the dominant kernel from live
HIP-event timing, and the host-CPU baseline.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)

PEAK_FP32_TFLOPS = 157.3   # MI355X published FP32 peak
PEAK_HBM_GBS = 8000.0      # MI355X published HBM3E bandwidth
# Algorithmic FP32 operations of one 36-tap bilateral NCC: per tap ~22 (map,
# divide, index and fraction, bilinear, three weighted moments) plus ~40 for
# the per-NCC homography, centre test and epilogue: 36 * 22 + 40.
OPS_PER_NCC = 36 * 22 + 40
OPS_PER_GEOM = 110.0       # one forward-backward reprojection
GATHER_BYTES_PER_NCC = 36 * 4 * 4  # 36 taps x 4 bilinear texels x 4 B


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_workload(device: str):
    import numpy as np
    import torch

    from gmvs_amd.synthetic import make_scene, voronoi_labels

    sc = make_scene(1600, 1200, 10, 24, seed=3407 + 2, fx=1500.0, arc_deg=40.0,
                    bump_fraction=0.10, device=device)
    labels = [voronoi_labels(1600, 1200, 2000, seed=3407 + 100 + i) for i in range(10)]
    return sc, labels


def candidate_ncc_count(h: int, w: int, nv: int) -> int:
    """NCCs computed by one pm_candidates launch over both colours' sum / 2:
    flagged candidates per active pixel (the flags depend on the pixel
    position only) times the V source views."""
    import numpy as np

    ys, xs = np.mgrid[0:h, 0:w]
    flags = ((ys > 2).astype(np.int64) + (ys < h - 3) + (xs > 2) + (xs < w - 3) + (ys > 0)
             + (ys < h - 1) + (xs > 0) + (xs < w - 1))
    return int(flags.sum()) * nv // 2


def host_cpu_info() -> dict:
    """nproc, the OpenMP thread count used and the lscpu model line."""
    import subprocess

    info = {"nproc": os.cpu_count()}
    try:
        info["affinity"] = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        pass
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for ln in out.splitlines():
            k, _, v = ln.partition(":")
            if k.strip() in ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core",
                             "CPU(s)", "CPU max MHz"):
                info[k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return info


def host_threads() -> int:
    """All host threads this process may use: OMP_NUM_THREADS when the
    environment sets it (the machine's share for this GPU), else the
    affinity mask."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    try:
        return len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        return os.cpu_count() or 1


def cpu_baseline(sc, labels, threads: int, iterations: int = 8, geom_iterations: int = 2):
    """The FP32 OpenMP port of the reference loop on one reference view of
    the workload, each pass of the benchmark's mix timed on its own at one
    iteration: plain pass, host prior construction (superpixel filter),
    prior pass, and a geometric pass against the scene's depth maps with
    1 % noise. Median of 3 runs after a warm-up run. The rate is the
    benchmark step's pixel-hypotheses of one view over the time the port
    takes for them, each pass's time per pixel-hypothesis applied to that
    pass's share of the step (iterations / geom_iterations per pass) plus
    the host prior. Returns (Mpix-hyp/s, per-pass seconds, runs)."""
    import numpy as np

    from gmvs_amd import src as S
    from gmvs_amd.lib.build import planar_initer as PI
    from oracle import oracle as O

    ids = [0] + list(sc.pairs[0][:9])
    cams = [O.make_camera(sc.cameras[i].K, sc.cameras[i].R, sc.cameras[i].t, sc.cameras[i].height,
                          sc.cameras[i].width, sc.cameras[i].depth_min, sc.cameras[i].depth_max)
            for i in ids]
    imgs = [sc.images[i].cpu().numpy() for i in ids]
    c0 = sc.cameras[0]
    dmin = float(np.float32(float(np.float32(c0.depth_min)) * 0.6))
    dmax = float(np.float32(float(np.float32(c0.depth_max)) * 1.2))
    h, w = imgs[0].shape
    rng = np.random.default_rng(3407)
    depths = [(sc.depth[i].cpu().numpy() * (1 + 0.01 * rng.standard_normal((h, w)))).astype(np.float32)
              for i in ids]
    normal = sc.normal[0].cpu().numpy()
    ccam = S.load_camera(c0.K.reshape(-1), c0.R.reshape(-1), c0.t, h, w, c0.depth_min,
                         c0.depth_max)._c
    lab = labels[0].cpu().numpy() if hasattr(labels[0], "cpu") else np.asarray(labels[0])

    def one_view():
        t = {}
        t0 = time.perf_counter()
        p1 = O.make_params(max_iterations=1, depth_min=dmin, depth_max=dmax)
        o1 = O.pm_run(cams, imgs, p1, precision=32, threads=threads, rng_view=0)
        t["plain"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        pp, pm, _ = PI.build_prior(ccam, o1["depth"], o1["cost"], lab, dmin, dmax,
                                   superpixel_filter=True)
        t["prior_build"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        p2 = O.make_params(max_iterations=1, depth_min=dmin, depth_max=dmax, planar_prior=True,
                           superpixel_filter=True)
        O.pm_run(cams, imgs, p2, precision=32, threads=threads, init_mode=2, prior_planes=pp,
                 plane_masks=pm.astype(np.uint32), planes=o1["planes"], costs=o1["cost"],
                 rng_view=0)
        t["prior"] = time.perf_counter() - t0
        t0 = time.perf_counter()
        p3 = O.make_params(max_iterations=1, depth_min=dmin, depth_max=dmax,
                           geom_consistency=True)
        O.pm_run(cams, imgs, p3, precision=32, threads=threads, init_mode=1, geom=True,
                 depths=depths, ref_normal=normal, rng_view=0)
        t["geom"] = time.perf_counter() - t0
        return t

    one_view()  # warm-up
    runs = [one_view() for _ in range(3)]
    med = {k: float(np.median([r[k] for r in runs])) for k in runs[0]}
    ph1 = h * w * (1 + 14 * 1)  # pixel-hypotheses of a one-iteration pass
    share = {"plain": h * w * (1 + 14 * iterations), "prior": h * w * (1 + 14 * iterations),
             "geom": h * w * (1 + 14 * geom_iterations)}
    t_step = med["prior_build"] + sum(med[k] / ph1 * share[k] for k in share)
    return sum(share.values()) / t_step / 1e6, med, runs


def cpu_fusion_baseline(fd_in, fn_in, fcams, fpairs, threads: int):
    """The fusion CPU twin (the sequential restatement of batch_fusion_kernel)
    on the first views of the fusion workload, in view order: one thread on
    2 views, then `threads` OpenMP threads on 8 views. Mpts/s of input points."""
    import numpy as np

    from oracle import oracle as O

    d = fd_in.cpu().numpy()
    n = fn_in.cpu().numpy()
    ocams = [O.make_camera(c.K, c.R, c.t, c.height, c.width, c.depth_min, c.depth_max)
             for c in fcams]
    nv, h, w = d.shape
    cos_t = float(np.float32(math.cos(math.radians(30.0))))
    res = {}
    for label, nthreads, nviews in (("sequential", 1, 2), ("openmp", threads, 8)):
        O.set_threads(nthreads)
        masks = np.zeros((nv, h, w), np.uint8)
        t0 = time.perf_counter()
        for ref in range(nviews):
            ids = [ref] + list(fpairs[ref])
            O.fuse_view([ocams[i] for i in ids], [d[i] for i in ids], [n[i] for i in ids], masks,
                        ref, ids[1:], 3, 1.0, 0.01, cos_t)
        dt = time.perf_counter() - t0
        res[label] = {"mpts_per_s": round(nviews * h * w / dt / 1e6, 2), "threads": nthreads,
                      "views": nviews, "s": round(dt, 2)}
    O.set_threads(threads)
    return res


def fusion_workload(device: str):
    """49 views at 1600x1200 from a 64-plane scene, 1% depth noise, 5% outliers;
    20 source views per reference ranked by viewing angle."""
    import numpy as np
    import torch

    from gmvs_amd.synthetic import make_scene, pair_sources

    sc = make_scene(1600, 1200, 49, 64, seed=3407 + 5, fx=1500.0, arc_deg=30.0,
                    pitch_deg=(-12.0, -8.0, -4.0, 0.0, 4.0, 8.0, 12.0), device=device)
    g = torch.Generator(device=device).manual_seed(3407)
    d = sc.depth.clone()
    d = d * (1 + 0.01 * torch.randn(d.shape, generator=g, device=device))
    out = torch.rand(d.shape, generator=g, device=device) < 0.05
    d = torch.where(out, 1.5 + 5.5 * torch.rand(d.shape, generator=g, device=device), d)
    pairs = pair_sources(sc.cameras, 20)
    return sc, d.contiguous(), sc.normal.contiguous(), pairs


def post_fusion(fsc, od, on, cpu_sample: bool):
    """The stage after fusion (mvs_fusion_segmentaion_cpu.py:158-216) on the
    fused maps: depth_normal maps as consistency.py:193-210 builds them, the
    GPU point gather with its view-angle filter, and the %.4f text of both
    point sets (formatting plus the copy of the text to the host). With
    cpu_sample, the numpy gather and np.savetxt -- the reference's own code
    path -- on a bounded sample: 2 views and the first 200,000 rows."""
    import io as _io

    import numpy as np
    import torch

    from gmvs_amd import post

    nv, h, w = od.shape
    nrm = on.clone()
    nn = torch.linalg.norm(nrm, dim=-1, keepdim=True)
    nrm = nrm / torch.where(nn > 0.5, nn, torch.ones_like(nn))
    maps = torch.cat([od[..., None], nrm], dim=-1).contiguous()
    poses = []
    for c in fsc.cameras:
        E = np.eye(4)
        E[:3, :3] = c.R.T
        E[:3, 3] = -c.R.T @ c.t
        poses.append(E)
    K = fsc.cameras[0].K
    post.gather_points([maps[0]], poses[:1], K)  # warm
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    pts, pts_f, filt = post.gather_points(list(maps), poses, K)
    torch.cuda.synchronize()
    tg = time.perf_counter() - t0
    from gmvs_amd import _native as N

    post.format_rows_device(pts_f[:1000])  # warm
    torch.cuda.synchronize()
    N.profile_reset()
    N.profile_enable(True)
    t0 = time.perf_counter()
    texts = [post.format_rows_device(pts), post.format_rows_device(pts_f)]
    torch.cuda.synchronize()
    tf = time.perf_counter() - t0
    N.profile_enable(False)
    kms = N.profile_collect().get("format_rows", {}).get("ms", float("nan"))
    t0 = time.perf_counter()
    txt = sum(len(memoryview(t.cpu().numpy())) for t in texts)
    th = time.perf_counter() - t0
    del texts
    nrows = int(pts.shape[0] + pts_f.shape[0])
    res = {"input_points": nv * h * w, "points": int(pts.shape[0]), "filtered_points": int(pts_f.shape[0]),
           "gather_ms": round(tg * 1e3, 2), "gather_mpts_per_s": round(nv * h * w / tg / 1e6, 1),
           "text_bytes": txt, "format_ms": round(tf * 1e3, 2),
           "format_mrows_per_s": round(nrows / tf / 1e6, 2), "format_kernels_ms": round(kms, 2),
           "text_to_host_ms": round(th * 1e3, 2)}
    del filt
    if cpu_sample:
        from oracle import oracle as O

        m2 = maps[:2].cpu().numpy()
        t0 = time.perf_counter()
        O.gather_points(list(m2), poses[:2], K, h, w)
        cg = time.perf_counter() - t0
        rows = pts[:200000].cpu().numpy()
        t0 = time.perf_counter()
        np.savetxt(_io.BytesIO(), rows, fmt="%.4f")
        cf = time.perf_counter() - t0
        res["cpu_numpy"] = {"gather_mpts_per_s": round(2 * h * w / cg / 1e6, 2),
                            "savetxt_mrows_per_s": round(len(rows) / cf / 1e6, 3), "cores": 1,
                            "sample": "2 views gathered, 200,000 rows formatted"}
    return res


def _short_kernel(name: str) -> str:
    name = name.strip().strip('"')
    if name.startswith("void "):
        name = name[5:]
    name = name.replace("(anonymous namespace)::", "")
    depth = 0
    for i, ch in enumerate(name):  # cut the argument list, keep template args
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            name = name[:i]
            break
    return name[:90]


def summarize_profile(prof_dir: str) -> str:
    """Compact text summary of rocprofv3 CSV output in prof_dir: the
    kernel-stats table and, for every counter-collection CSV, the mean value
    per dispatch of each counter for each kernel."""
    import csv
    import glob
    from collections import defaultdict

    out = []
    for path in sorted(glob.glob(os.path.join(prof_dir, "**", "*kernel_stats.csv"), recursive=True)):
        out.append(f"== {os.path.basename(path)} (top 20 by total time, then the other gmvs kernels)")
        with open(path, newline="") as f:
            rows = list(csv.DictReader(f))
        rows.sort(key=lambda r: -float(r.get("TotalDurationNs", 0) or 0))
        out.append(f"{'kernel':60s} {'calls':>7s} {'total_ms':>10s} {'avg_us':>10s} {'min_us':>9s} {'max_us':>9s}")
        # the top 20, then this library's kernels below them
        for r in rows[:20] + [r for r in rows[20:] if _short_kernel(r["Name"]).startswith("gmvs::")]:
            out.append(f"{_short_kernel(r['Name'])[:60]:60s} {int(r['Calls']):7d} "
                       f"{float(r['TotalDurationNs']) / 1e6:10.2f} {float(r['AverageNs']) / 1e3:10.1f} "
                       f"{float(r['MinNs']) / 1e3:9.1f} {float(r['MaxNs']) / 1e3:9.1f}")
    for path in sorted(glob.glob(os.path.join(prof_dir, "**", "*counter_collection.csv"),
                                       recursive=True)):
        acc = defaultdict(float)
        disp = defaultdict(set)
        top = []  # the largest single dispatches per counter (finds the copy probe)
        with open(path, newline="") as f:
            for r in csv.DictReader(f):
                k = _short_kernel(r["Kernel_Name"])
                v = float(r["Counter_Value"])
                acc[(k, r["Counter_Name"])] += v
                disp[k].add(r["Dispatch_Id"])
                if "copy" in k.lower():
                    top.append((v, r["Counter_Name"], k, r["Dispatch_Id"], r.get("Grid_Size", "")))
                    top = sorted(top, reverse=True)[:4]
        counters = sorted({c for _, c in acc})
        out.append(f"== {os.path.basename(path)}: mean per dispatch")
        out.append(f"{'kernel':60s} {'disp':>6s} " + " ".join(f"{c:>20s}" for c in counters))
        kern = sorted(disp, key=lambda k: -sum(acc[(k, c)] for c in counters))
        for k in kern[:25]:
            n = len(disp[k])
            out.append(f"{k[:60]:60s} {n:6d} " + " ".join(f"{acc[(k, c)] / n:20.1f}" for c in counters))
        out.append("largest copy-kernel dispatches: value, counter, kernel, dispatch id, grid size")
        for t in top:
            out.append(f"  {t[0]:.1f} {t[1]} {t[2]} {t[3]} {t[4]}")
    return "\n".join(out)


SCENES4 = [  # config 4: (width, height, planes, plane depth range)
    (1024, 768, 16, (2.0, 6.0)),
    (1600, 1200, 27, (2.5, 7.0)),
    (2048, 1536, 38, (3.0, 8.0)),
    (3200, 2400, 48, (2.0, 5.0)),
]


def run_scenes4(args) -> int:
    """Config 4: 4 synthetic scenes of 12 views at 1024x768, 1600x1200,
    2048x1536 and 3200x2400 (48 reference views, 11 sources each, plain
    pass, 8 iterations), every view as reference. Timed: the plain path (FP32,
    one view per call) and the whole 48-view batch replayed from one HIP graph
    in FP32, FP16 and BF16, tile shape and blocks per CU from the autotuner."""
    import numpy as np
    import torch

    from gmvs_amd import _native as N
    from gmvs_amd import src as S
    from gmvs_amd.autotune import autotune
    from gmvs_amd.autotune import reset as autotune_reset
    from gmvs_amd.pipeline import cameras_from_synthetic
    from gmvs_amd.synthetic import make_scene

    dev = "cuda"
    problems, ph = [], 0
    for i, (w, h, npl, dr) in enumerate(SCENES4):
        log(f"scene {i}: {w}x{h}, {npl} planes, 12 views")
        sc = make_scene(w, h, 12, npl, seed=3407 + 40 + i, depth_range=dr,
                        cam_depth_range=(dr[0] + 0.25, dr[1]), device=dev)
        cams = cameras_from_synthetic(sc.cameras)
        imgs = sc.images.to(dev, torch.float32).contiguous()
        for ref in range(12):
            ids = [ref] + list(sc.pairs[ref][:11])
            prm = S.pm_params_for_problem(S.PatchMatchParams(), cams[ref], len(ids))
            prm.max_iterations = 8
            problems.append(dict(cams=[cams[j]._c for j in ids], images=[imgs[j] for j in ids],
                                 params=prm, init_mode=N.INIT_RANDOM, geom=False, rng_view=ref,
                                 planes=torch.zeros(h, w, 4, device=dev),
                                 costs=torch.zeros(h, w, device=dev)))
            ph += h * w * (1 + 14 * 8)
        del sc
    torch.cuda.synchronize()

    results = {}
    log("plain FP32, one view per call")
    ws = S._Workspace()

    def plain():
        for p in problems:
            kw = {k: v for k, v in p.items() if k not in ("cams", "images", "params")}
            S.run_pm(p["cams"], p["images"], p["params"], workspace=ws, **kw)

    plain()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plain()
    torch.cuda.synchronize()
    plain_rate = ph * args.steps / (time.perf_counter() - t0) / 1e6
    results["fp32_per_view"] = {"value": round(plain_rate, 2)}
    log(f"fp32_per_view: {plain_rate:.1f} Mpix-hyp/s")
    del ws

    runs = (("fp32_default", N.PREC_FP32, False), ("fp32", N.PREC_FP32, True),
            ("fp16", N.PREC_FP16, True), ("bf16", N.PREC_BF16, True))
    for name, prec, tune in runs:
        batch = [dict(p, precision=prec) for p in problems]
        tuned = autotune(batch) if tune else {"tile_cols": 8, "blocks_candidates": 0,
                                              "blocks_refine": 0, "lanes_candidates": 0,
                                              "lanes_refine": 0}
        batch = [dict(p, tile_cols=tuned["tile_cols"]) for p in batch]
        ws = S._Workspace()
        stream = torch.cuda.Stream()
        with torch.cuda.stream(stream):
            S.run_pm_batch(batch, ws)  # eager run sizes the workspace
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=stream):
            S.run_pm_batch(batch, ws)
        g.replay()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            g.replay()
        torch.cuda.synchronize()
        rate = ph * args.steps / (time.perf_counter() - t0) / 1e6
        results[f"{name}_batched_graph"] = {"value": round(rate, 2),
                                            "speedup_vs_plain_fp32": round(rate / plain_rate, 3),
                                            "autotune": tuned}
        log(f"{name} batched graph: {rate:.1f} Mpix-hyp/s ({rate / plain_rate:.3f}x), tuned {tuned}")
        del g, ws
        torch.cuda.empty_cache()
        autotune_reset()
    autotune_reset()
    best = results["bf16_batched_graph"]["value"]
    line = {
        "metric": "PatchMatch Mpix-hyp/s + fusion Mpts/s; batched BF16 vs plain FP32 at fixed err",
        "value": best, "unit": "Mpix-hyp/s", "n_gpus": 1, "steps": args.steps, "warmup": 1,
        "ms_per_step": round(ph / best / 1e3, 2), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "bf16", "data": "synthetic",
        "config": {"workload": "scenes4_48v_uneven_plain_8it_graph", "views": len(problems),
                   "sizes": [f"{w}x{h}" for w, h, _, _ in SCENES4], "sources": 11,
                   "iterations": 8, "pixel_hypotheses_per_step": ph},
        "variants": results,
    }
    print(json.dumps(line), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fusion", action="store_true")
    ap.add_argument("--variants", default="plain,fp16,bf16",
                    help="other runs of the same workload: plain (FP32, one view per call), "
                         "fp16 / bf16 (batched, 16-bit samples), fused (batched FP32 with the "
                         "view selection fused into the candidate kernel), devprior (batched "
                         "FP32 with the planar priors built on the device); comma list, empty "
                         "for none")
    ap.add_argument("--fuse-select", action="store_true",
                    help="run the view selection as the candidate kernel's tile epilogue in the "
                         "measured pipeline (gmvs_set_select_fusion; identical results)")
    ap.add_argument("--workload", default="acmp10", choices=("acmp10", "scenes4"),
                    help="acmp10: the 10-view ACMP pipeline (default); scenes4: 4 scenes x 12 "
                         "views of uneven sizes in one batched HIP graph")
    ap.add_argument("--summarize-profile", metavar="DIR", default=None,
                    help="print a summary of rocprofv3 CSV output in DIR and exit")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="before the workload, run one device-to-device copy of this many MiB "
                         "(the known-size reference for hardware byte counters)")
    args = ap.parse_args()
    if args.summarize_profile:
        print(summarize_profile(args.summarize_profile))
        return 0
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; --gpus must be 1 "
                                   "(views would shard across GPUs with one all-gather of "
                                   "depth/normal maps per stage; not built)"}))
        return 2
    if args.workload == "scenes4":
        return run_scenes4(args)

    import numpy as np
    import torch

    from gmvs_amd import _native as N
    from gmvs_amd.pipeline import MultiViewPipeline, cameras_from_synthetic, fuse_maps, pixel_hypotheses

    dev = "cuda"
    N.set_select_fusion(args.fuse_select)
    if args.copy_probe_mib > 0:
        nel = args.copy_probe_mib * (1 << 18)
        src = torch.ones(nel, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy probe: {nel * 4} bytes read, {nel * 4} bytes written")
        del src, dst
    log("generating the 10-view 1600x1200 scene ...")
    sc, labels = make_workload(dev)
    cams = cameras_from_synthetic(sc.cameras)
    pairs = [p[:9] for p in sc.pairs]
    pipe = MultiViewPipeline(sc.images, cams, pairs, labels=labels, prior_threads=8)
    n, h, w = pipe.n, pipe.h, pipe.w
    ph_step = n * (2 * pixel_hypotheses(h, w, 8) + pixel_hypotheses(h, w, 2))

    def step(p=pipe):
        s1 = p.first_stage(iterations=8, planar_prior=True, superpixel_filter=True)
        s2 = p.geom_pass(s1, iterations=2)
        fd, _, _ = p.fuse(s2, geom_consistent=3)
        return fd

    for i in range(args.warmup):
        log(f"warmup step {i + 1}/{args.warmup}")
        step()
    torch.cuda.synchronize()
    N.profile_reset()
    N.profile_enable(True)
    log(f"timing {args.steps} steps")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        fd = step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    N.profile_enable(False)
    stats = N.profile_collect()
    fused = int((fd > 0).sum().item())
    ms_per_step = elapsed / args.steps * 1e3
    value = ph_step * args.steps / elapsed / 1e6
    log(f"{ms_per_step:.1f} ms/step, {value:.1f} Mpix-hyp/s, fused points {fused}")
    for k, s in sorted(stats.items(), key=lambda kv: -kv[1]["ms"]):
        log(f"  {k:18s} launches {s['launches']:6d}  total {s['ms']:9.2f} ms  "
            f"avg {s['ms'] / s['launches']:8.3f} ms  ncc {s['ncc']:>13d}  geom {s['geom']:>12d}")

    # roofline of the dominant kernel (largest accumulated time): its
    # algorithmic FP32 work per launch (counted NCCs and geometric terms x the
    # per-unit flops) over its average launch duration from the HIP events
    dom_name, dom = max(stats.items(), key=lambda kv: kv[1]["ms"])
    avg_ms = dom["ms"] / dom["launches"]
    ncc_per_launch = dom["ncc"] / dom["launches"]
    geom_per_launch = dom["geom"] / dom["launches"]
    ops = ncc_per_launch * OPS_PER_NCC + geom_per_launch * OPS_PER_GEOM
    ach = ops / (avg_ms * 1e-3) / 1e12
    # the tap gathers of that kernel against the measured peak of the same
    # loads in the NCC waves' layout (FP32 cells: 16 B per tap); random lane
    # pairs in an L1-sized window for comparison
    gather_ach = ncc_per_launch * GATHER_BYTES_PER_NCC / (avg_ms * 1e-3) / 1e9
    gather_peak = N.gather_peak(16, pattern=1)
    gather_random = N.gather_peak(16, pattern=0)
    log(f"gather peak (16-byte cells): NCC-wave layout {gather_peak:.0f} GB/s, "
        f"random lane pairs {gather_random:.0f} GB/s")
    # HBM bytes and VALU issue utilisation per launch of that kernel from the
    # committed rocprofv3 passes over this workload (FETCH_SIZE / WRITE_SIZE
    # scaled by a copy of known size; SQ_INSTS_VALU per launch over one wave
    # instruction per CU per clock at 2.4 GHz); counters cannot be read inside
    # a timed run, so each figure names the commit its profile was taken at
    traffic, hbm, valu = None, None, None
    tpath = os.path.join(HERE, "profiles", "pm_traffic_acmp_10v_1600x1200.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tj = json.load(f)
        tk = tj["kernels"].get(dom_name)
        if tk:
            traffic = tk["traffic_bytes_per_launch"]
            gbs = traffic / (avg_ms * 1e-3) / 1e9
            hbm = {"achieved_GBps": round(gbs, 1), "peak_GBps": PEAK_HBM_GBS,
                   "frac": round(gbs / PEAK_HBM_GBS, 4), "profile": os.path.basename(tpath),
                   "profile_commit": tj.get("commit")}
            if "valu_insts_per_launch" in tk:
                util = tk["valu_insts_per_launch"] / (256 * 2.4e9 * tk["avg_ms"] * 1e-3)
                valu = {"issue_util": round(util, 4), "insts_per_launch": tk["valu_insts_per_launch"],
                        "insts_per_ncc": round(tk["valu_insts_per_launch"] / ncc_per_launch, 2),
                        "profile": os.path.basename(tpath), "profile_commit": tj.get("commit")}
    roof = {"bound": "valu", "achieved": round(ach, 3), "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s", "frac": round(ach / PEAK_FP32_TFLOPS, 4), "traffic": traffic,
            "ops_per_ncc": OPS_PER_NCC, "hbm": hbm, "valu": valu,
            "gather": {"achieved_GBps": round(gather_ach, 1), "peak_GBps": round(gather_peak, 1),
                       "frac": round(gather_ach / gather_peak, 4),
                       "random_pairs_GBps": round(gather_random, 1),
                       "bytes_per_ncc": GATHER_BYTES_PER_NCC},
            "kernel": dom_name, "avg_ms": round(avg_ms, 4),
            "ncc_per_launch": int(ncc_per_launch), "geom_per_launch": int(geom_per_launch)}
    if "pm_candidates" in stats:
        s = stats["pm_candidates"]
        log(f"pm_candidates NCCs per launch: counted {s['ncc'] / s['launches']:.0f}; per view "
            f"and half-sweep (analytic) {candidate_ncc_count(h, w, 9)}")
    total_ncc = sum(s["ncc"] for s in stats.values())
    total_kernel_ms = sum(s["ms"] for s in stats.values())
    log(f"NCC evaluations per step {total_ncc // args.steps}, "
        f"{total_ncc / (elapsed) / 1e9:.2f} G NCC/s wall, kernel time {total_kernel_ms / args.steps:.1f} ms/step")

    # The same workload run view by view in FP32 (the reference's structure:
    # one view per call; the plain path the batched variants are measured
    # against) and batched with 16-bit source samples (FP32 NCC sums).
    specs = []
    for v in [v for v in args.variants.split(",") if v]:
        if v == "plain":
            specs.append(("fp32_per_view", N.PREC_FP32, False, False, False))
        elif v == "fused":
            # selection and adoption as the candidate kernel's tile epilogue
            specs.append(("fp32_batched_fused_select", N.PREC_FP32, True, True, False))
        elif v == "devprior":
            # planar priors built on the device in stream order
            specs.append(("fp32_batched_device_prior", N.PREC_FP32, True, False, True))
        else:
            specs.append((f"{v}_batched", N.PRECISIONS[v], True, False, False))
    variants = {}
    for name, prec, batched, fuse_sel, dev_prior in specs:
        vp = MultiViewPipeline(sc.images, cams, pairs, labels=labels, prior_threads=8,
                               precision=prec, batched=batched, device_prior=dev_prior)
        N.set_select_fusion(fuse_sel)
        for i in range(args.warmup):
            step(vp)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(args.steps):
            vfd = step(vp)
        torch.cuda.synchronize()
        vt = time.perf_counter() - t0
        N.set_select_fusion(args.fuse_select)
        vval = ph_step * args.steps / vt / 1e6
        vfused = int((vfd > 0).sum().item())
        # the fused depth map against the FP32 batched run's, bit for bit
        same = bool(torch.equal(vfd.contiguous().view(torch.int32), fd.contiguous().view(torch.int32)))
        variants[name] = {"value": round(vval, 2), "ms_per_step": round(vt / args.steps * 1e3, 2),
                          "fused_points": vfused, "fused_depth_bits_equal_fp32": same}
        log(f"{name}: {variants[name]}")
        del vp
    plain = variants.get("fp32_per_view", {}).get("value")
    if plain:
        for v in variants.values():
            v["speedup_vs_plain_fp32"] = round(v["value"] / plain, 3)
        log(f"fp32_batched: speedup vs plain fp32 {value / plain:.3f}")

    fusion = None
    if not args.no_fusion:
        log("fusion workload: 49 views at 1600x1200, 20 sources each")
        fsc, fd_in, fn_in, fpairs = fusion_workload(dev)
        fcams = cameras_from_synthetic(fsc.cameras)
        fuse_maps(fd_in, fn_in, fcams, fpairs, geom_consistent=3, reproj_px=1.0, rel_depth=0.01,
                  angle_rad=math.radians(30.0))
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        od, on, mk = fuse_maps(fd_in, fn_in, fcams, fpairs, geom_consistent=3, reproj_px=1.0,
                               rel_depth=0.01, angle_rad=math.radians(30.0))
        torch.cuda.synchronize()
        ft = time.perf_counter() - t1
        pts = fd_in.numel()
        fusion = {"mpts_per_s": round(pts / ft / 1e6, 1), "input_points": pts,
                  "fused_points": int((od > 0).sum().item()), "ms": round(ft * 1e3, 2)}
        # the same fusion with the reprojection check in FP64
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        od64, _, _ = fuse_maps(fd_in, fn_in, fcams, fpairs, geom_consistent=3, reproj_px=1.0,
                               rel_depth=0.01, angle_rad=math.radians(30.0), precision=64)
        torch.cuda.synchronize()
        ft64 = time.perf_counter() - t1
        fusion["fp64_check"] = {"mpts_per_s": round(pts / ft64 / 1e6, 1),
                                "fused_points": int((od64 > 0).sum().item()),
                                "ms": round(ft64 * 1e3, 2)}
        del od64
        log(f"fusion: {fusion}")
        fusion["post"] = post_fusion(fsc, od, on, not args.no_cpu_baseline)
        log(f"post-fusion gather and text: {fusion['post']}")
        del od, on, mk
        if not args.no_cpu_baseline:
            log("fusion CPU twin ...")
            fusion["cpu_twin"] = cpu_fusion_baseline(fd_in, fn_in, fcams, fpairs, host_threads())
            log(f"fusion CPU twin: {fusion['cpu_twin']}")

    cpu = None
    if not args.no_cpu_baseline:
        threads = host_threads()
        log(f"CPU baseline (FP32 OpenMP port, {threads} threads) ...")
        rate, med, runs = cpu_baseline(sc, labels, threads)
        cpu = {"value": round(rate, 3), "unit": "Mpix-hyp/s", "cores": threads, "kind": "port",
               "sample": ("view 0 of the workload: plain pass, host prior, prior pass and "
                          "geometric pass at 1 iteration each, median of 3 runs after a warm-up "
                          f"({sum(med.values()):.1f} s per run); the rate weights each pass's "
                          "time per pixel-hypothesis by its share of the step (8, 8 and 2 "
                          "iterations) and adds the host prior"),
               "pass_s": {k: round(v, 3) for k, v in med.items()},
               "runs_s": [round(sum(r.values()), 2) for r in runs], "host": host_cpu_info(),
               "gpu_vs_cpu": round(value / rate, 1)}
        log(f"CPU baseline: {cpu}")

    line = {
        "metric": "PatchMatch Mpix-hyp/s + fusion Mpts/s; batched BF16 vs plain FP32 at fixed err",
        "value": round(value, 2),
        "unit": "Mpix-hyp/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "acmp_10v_1600x1200_prior_geom_fusion", "batched": True,
                   "select_fused": bool(args.fuse_select),
                   "plain_fp32_per_view": plain, "views": n,
                   "height": h, "width": w, "sources": 9, "iterations": 8, "geom_iterations": 2,
                   "superpixels": 2000, "planes": 24, "pixel_hypotheses_per_step": ph_step,
                   "fused_points": fused},
        # the NCCs the kernels evaluated (refinement costs only the selected
        # views, so fewer than pixel-hypotheses x sources), per step and per
        # second of the timed region
        "ncc": {"per_step": int(total_ncc // args.steps),
                "g_per_s": round(total_ncc / elapsed / 1e9, 2)},
        "fusion": fusion,
        "variants": variants,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
