#!/usr/bin/env python3
"""Benchmark of the hot path on one MI355X: fwd+bwd Mpix/s and plain tracking-step latency.

Workload (one "step"): the 1M-Gaussian, 640x480 mapping view — projection, (tile, depth)
binning, blend forward and backward and the 2D->3D gradients (fvgs_mapping_step) — against a
synthetic observation (the GPU's own forward render minus N(0, 1e-2) perturbations), FP32. The
map and the frame are resident in HBM before the timed region.

Printed JSON line: value = pixels / step time (Mpix/s); `roofline` for the dominant kernel (the
slower of the blend forward and backward), measured with the library's HIP events on its own
stream over the timed steps;
`cpu_baseline` = the FP64 oracle (OpenMP) timed on the host cores over the same view; `tracking`
= plain Gauss-Newton tracking steps (500k Gaussians, 640x480, 20 iterations from a 2 deg / 3 cm
perturbation), p50/p99 step latency.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu-baseline] [--no-tracking]
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--width", type=int, default=640)
    ap.add_argument("--height", type=int, default=480)
    ap.add_argument("--track-n", type=int, default=500_000)
    ap.add_argument("--track-iters", type=int, default=20)
    ap.add_argument("--track-solves", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-tracking", action="store_true")
    ap.add_argument("--cpu-runs", type=int, default=3)
    ap.add_argument("--breakdown", action="store_true",
                    help="after the timed steps, run the forward-only blend K times (profiling)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="first run 3 device-to-device copies of this many MiB: the reference "
                         "kernel that calibrates rocprofv3's FETCH_SIZE / WRITE_SIZE counters")
    return ap.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def mapping_frame(ctx, scene, intr, pose, w, h):
    """Observation = GPU forward render at the pose minus seeded N(0, 1e-2) noise."""
    from f_vigs_slam_amd import scenes as S
    ctx.set_frame(np.zeros((h, w, 3), np.float32), np.ones((h, w), np.float32))
    ctx.prepare_rasterization(pose, intr, w, h)
    ctx.forward()
    rgb, depth = ctx.read("RGB"), ctx.read("DEPTH")
    up_c, up_d = S.upstream_grads(w, h, seed=1)
    return (rgb - up_c).astype(np.float32), (depth - up_d).astype(np.float32)


def cpu_baseline(scene, intr, pose, obs_rgb, obs_depth, w, h, runs):
    """The oracle (FP64, OpenMP) on the same view: binning + forward + backward + 2D->3D."""
    import oracle as O
    cp = np.asarray(pose[:3], np.float32)
    cq = np.asarray(pose[3:], np.float32)

    def once():
        t0 = time.perf_counter()
        pf, b, rec = O.pipeline(scene, cp, cq, intr, w, h)
        fwd = O.forward(w, h, b, rec, scene["color"], scene["opacity"], obs_rgb)
        d2 = O.backward(w, h, b, rec, scene["color"], scene["opacity"], fwd, obs_depth)
        O.delta3d(scene["pos"], scene["scale"], scene["quat"], cp, cq, intr, d2)
        return time.perf_counter() - t0

    once()
    ts = [once() for _ in range(runs)]
    t = statistics.median(ts)
    return dict(value=round(w * h / t / 1e6, 4), unit="Mpix/s", cores=O.num_threads(),
                kind="port",
                sample=f"the full {len(scene['pos'])}-Gaussian {w}x{h} view (binning FP32 + "
                       f"forward/backward/2D->3D FP64), median of {runs} runs after 1 warm-up, "
                       f"{t:.3f} s per view")


def copy_probe(mib):
    """3 plain device copies of `mib` MiB (reads and writes exactly that many bytes each)."""
    import torch
    src = torch.ones(mib * (1 << 20) // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    log(f"copy probe: 3 x {mib} MiB")


def tracking(args, log_fn):
    from f_vigs_slam_amd import Context
    from f_vigs_slam_amd import scenes as S
    w, h = 640, 480
    ctx = Context()
    scene = S.make_scene(args.track_n, seed=21)
    intr = S.intrinsics(w, h)
    gt = np.array([0, 0, 0, 0, 0, 0, 1.0])
    ctx.set_scene(scene)
    ctx.render_tracking(gt, intr, w, h)
    obs_rgb, obs_depth = ctx.read("TRACK_RGB"), ctx.read("TRACK_DEPTH")
    ctx.set_frame(obs_rgb, obs_depth)
    ctx.compute_gradients()
    steps, errs = [], []
    for k in range(args.track_solves + 1):
        start = S.perturbed_pose(21 + k, rot_deg=2.0, trans_m=0.03)
        pose, ms = ctx.track(intr, args.track_iters, start)
        if k > 0:   # the first solve warms up
            steps.extend(ms.tolist())
            errs.append(S.pose_error(pose, gt))
        log_fn(f"tracking solve {k}: {np.mean(ms):.3f} ms/step, final error "
               f"{S.pose_error(pose, gt)[0]:.2e} rad {S.pose_error(pose, gt)[1]:.2e} m")
    st = np.array(steps)
    ctx.close()
    return dict(form="plain", n_gaussians=args.track_n, frame=f"{w}x{h}",
                iters=args.track_iters, solves=args.track_solves,
                p50_ms=round(float(np.percentile(st, 50)), 4),
                p99_ms=round(float(np.percentile(st, 99)), 4),
                mean_ms=round(float(st.mean()), 4),
                final_rot_err_rad=float(max(e[0] for e in errs)),
                final_trans_err_m=float(max(e[1] for e in errs)))


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported in this build "
              "(the hot path runs on one MI355X; see the sharding notes in DESIGN.md).",
              file=sys.stderr)
        sys.exit(2)
    from f_vigs_slam_amd import Context
    from f_vigs_slam_amd import scenes as S

    w, h, n = args.width, args.height, args.n
    if args.copy_probe_mib:
        copy_probe(args.copy_probe_mib)
    log(f"scene {n} Gaussians {w}x{h}")
    scene = S.make_scene(n, seed=0, width=w, height=h)
    intr = S.intrinsics(w, h)
    pose = np.array([0, 0, 0, 0, 0, 0, 1.0])
    ctx = Context()
    ctx.set_scene(scene)
    obs_rgb, obs_depth = mapping_frame(ctx, scene, intr, pose, w, h)
    ctx.set_frame(obs_rgb, obs_depth)
    for _ in range(args.warmup):
        ctx.mapping_step(pose, intr)
    ctx.set_timing(True)
    stage = np.zeros(5)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.mapping_step(pose, intr)
        stage += np.array(ctx.last_timing())
    ctx.sync()
    t = (time.perf_counter() - t0) / args.steps
    ctx.set_timing(False)
    stage /= args.steps
    st = ctx.stats()
    M, K, P = st["M"], st["K"], st["pixels"]
    if args.breakdown:
        ctx.prepare_rasterization(pose, intr, w, h)
        for _ in range(args.steps):
            ctx.forward()
    log(f"step {t * 1e3:.3f} ms; stages binning {stage[0]:.3f} clear {stage[1]:.3f} "
        f"forward {stage[2]:.3f} backward {stage[3]:.3f} 3d {stage[4]:.3f} ms; M={M} K={K}")
    # Algorithmic bytes (compulsory traffic, each counted once), K = walked entries, P = pixels:
    #   k_forward:  52 B per walked entry (index 4 + 48 B of record) + 48 B per pixel
    #               (12 B observed colour in; colour 12, depth 4, T 4, n_contrib 4, error 12 out);
    #   k_backward: 52 B per walked entry + 52 B per pixel (colour 12, depth 4, n_contrib 4,
    #               error 12, observed depth 4 in; one 16 B checkpoint per pixel and bucket);
    #   whole step: 108 N + 44 M + 104 K + 36 P.
    # The roofline line is for the slower blend kernel, each timed by its own events.
    fwd_bytes = 52.0 * K + 48.0 * P
    bwd_bytes = 52.0 * K + 52.0 * P
    if stage[3] >= stage[2]:
        kname, kbytes, kms = "k_backward (blend backward, A15)", bwd_bytes, stage[3]
    else:
        kname, kbytes, kms = "k_forward (blend forward, A14)", fwd_bytes, stage[2]
    achieved = kbytes / (kms * 1e-3) / 1e9
    step_bytes = 108.0 * n + 44.0 * M + 104.0 * K + 36.0 * P
    result = {
        "metric": "fwd+bwd Mpix/s @1M Gaussians 640x480; p99 tracking-step ms, served vs plain",
        "value": round(P / t / 1e6, 4),
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded config-0 distributions; observation = own render - N(0,1e-2))",
        "config": {"workload": "plain fwd+bwd view: projection, binning, blend forward, blend "
                               "backward, 2D->3D gradients",
                   "n_gaussians": n, "width": w, "height": h, "instances_M": M,
                   "walked_K": K, "pixels_P": P,
                   "stage_ms": {"binning": round(stage[0], 4), "clear": round(stage[1], 4),
                                "forward": round(stage[2], 4), "backward": round(stage[3], 4),
                                "grad3d": round(stage[4], 4)},
                   "step_algorithmic_GBs": round(step_bytes / t / 1e9, 2)},
        "roofline": {"kernel": kname, "bound": "hbm",
                     "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
                     "algorithmic_bytes": int(kbytes), "avg_ms": round(kms, 4)},
    }
    ctx.close()
    if not args.no_tracking:
        try:
            result["tracking"] = tracking(args, log)
        except Exception as e:  # report, do not hide the main number
            result["tracking"] = {"error": str(e)}
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle) ...")
        result["cpu_baseline"] = cpu_baseline(scene, intr, pose, obs_rgb, obs_depth, w, h,
                                              args.cpu_runs)
        result["cpu_baseline"]["speedup"] = round(result["value"] / result["cpu_baseline"]["value"], 2)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
