"""Shared helpers of the parity tests: seeded workloads, oracle runs, tolerance checks."""
from __future__ import annotations

import numpy as np

import oracle as O
from f_vigs_slam_amd import scenes as S

IDENTITY7 = np.array([0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0])


def workload(n, seed=0, width=640, height=480, pose7=None):
    scene = S.make_scene(n, seed=seed, width=width, height=height)
    intr = S.intrinsics(width, height)
    pose = IDENTITY7 if pose7 is None else np.asarray(pose7, np.float64)
    return scene, intr, pose


def cam(pose7):
    p = np.asarray(pose7, np.float64)
    return p[:3].astype(np.float32), p[3:].astype(np.float32)


def oracle_view(scene, pose7, intr, width, height):
    cp, cq = cam(pose7)
    return O.pipeline(scene, cp, cq, intr, width, height)


def mapping_observations(scene, b, rec, width, height, seed=1):
    """Observed colour/depth = oracle forward render minus N(0, 1e-2) upstream perturbations."""
    zero = np.zeros((height * width, 3), np.float32)
    f0 = O.forward(width, height, b, rec, scene["color"], scene["opacity"], zero)
    up_c, up_d = S.upstream_grads(width, height, seed=seed)
    obs_rgb = (f0["rgb"].reshape(height, width, 3) - up_c).astype(np.float32)
    obs_depth = (f0["depth"].reshape(height, width) - up_d).astype(np.float32)
    return obs_rgb, obs_depth


def close_fraction(got, ref, atol, rtol):
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    err = np.abs(got - ref)
    ok = err <= atol + rtol * np.abs(ref)
    return float(ok.mean()), float(err.max()) if err.size else 0.0


def grad_agreement(got, ref, rel=1e-3):
    """Fraction of entries with |got - ref| <= rel * (|ref| + 1e-3 * max|ref|): relative error
    with a floor at a thousandth of the field's largest magnitude."""
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    scale = np.abs(ref).max() if ref.size else 0.0
    ok = np.abs(got - ref) <= rel * (np.abs(ref) + 1e-3 * scale)
    return float(ok.mean())


def grad_report(got, ref):
    """Error summary of a gradient field: max-norm relative error |g - r|_inf / |r|_inf and
    quantiles of the per-entry error normalised the same way."""
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    scale = max(np.abs(ref).max(), 1e-30)
    e = np.abs(got - ref).max(axis=-1) / scale if got.ndim > 1 else np.abs(got - ref) / scale
    return dict(maxnorm=float(e.max()), q50=float(np.quantile(e, 0.5)),
                q99=float(np.quantile(e, 0.99)), q999=float(np.quantile(e, 0.999)))


def rel_l2(got, ref):
    """||got - ref||_2 / ||ref||_2 over the whole field."""
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    return float(np.linalg.norm(got - ref) / max(np.linalg.norm(ref), 1e-300))


def quantize_reference(scene, chunk=256):
    """The precision form's storage restated in numpy: per chunk of 256 Gaussians an FP32 origin
    at the centre of the positions' bounding box, FP16 (round to nearest even) offsets and FP16
    scale, quaternion, colour and opacity. Returns the dequantised FP32 map as the kernels read
    it (dict like the scene)."""
    f32 = np.float32
    pos = np.asarray(scene["pos"], f32)
    n = len(pos)
    out = {}
    deq = np.empty_like(pos)
    for c0 in range(0, n, chunk):
        blk = pos[c0:c0 + chunk]
        centre = ((blk.min(0) + blk.max(0)) * f32(0.5)).astype(f32)
        off = (blk - centre).astype(np.float16).astype(f32)
        deq[c0:c0 + chunk] = centre + off
    out["pos"] = deq
    for k in ("scale", "quat", "color", "opacity"):
        out[k] = np.asarray(scene[k], f32).astype(np.float16).astype(f32)
    return out
