"""GPU parity of the Adam map update (fvgs_adam_step / fvgs_optimize_keyframe) against the
oracle's FP64 restatement of updateGaussiansParametersAdam_kernel (GSCudaKernels.cu:2722-2882),
through the C ABI.

The oracle is fed the GPU's own 3D gradients, so the test isolates the update (the gradients'
parity has its own tests). Bar: every map entry within 2e-6 * (1 + |oracle|) of the oracle's
(FP32 arithmetic with the reference's fast intrinsics: __powf, __frsqrt_rn, __logf, __expf);
Adam moments within 1e-4 of the oracle's relative to each moment column's largest entry; step
counts exact; Gaussians without gradient (n = 0) bit-identical to their previous values. Two
consecutive steps, the second from a new pose (first moments no longer equal to one gradient).
"""
import numpy as np
import pytest

import oracle as O
from f_vigs_slam_amd import Context, FvgsError
from f_vigs_slam_amd import _lib as L
from f_vigs_slam_amd import scenes as S

pytestmark = pytest.mark.gpu

W, H, N, ETA = 320, 240, 20000, 2e-3
FIELDS = (("pos", 0, 3), ("scale", 3, 6), ("quat", 6, 10), ("color", 10, 13),
          ("opacity", 13, 14))


def _observation(scene, pose, intr):
    with Context() as c:
        c.set_scene(scene)
        c.render_tracking(pose, intr, W, H)
        rgb = np.clip(c.read("TRACK_RGB").reshape(H, W, 3) * 0.9 + 0.03, 0, 1).astype(np.float32)
        depth = (c.read("TRACK_DEPTH").reshape(H, W) * 1.02).astype(np.float32)
    return rgb, depth


def _d16(g):
    d = np.zeros((len(g), 16))
    d[:, 0:3], d[:, 3:6], d[:, 6:9] = g["position"], g["scale"], g["orientation"]
    d[:, 9:12], d[:, 12], d[:, 13] = g["color"], g["alpha"], g["n"]
    return d


def _as_map(m14):
    out = {k: np.ascontiguousarray(m14[:, a:z]) for k, a, z in FIELDS}
    out["opacity"] = np.ascontiguousarray(m14[:, 13])
    return out


def _check_map(tag, got14, want, prev14, d):
    for k, a, z in FIELDS:
        g = got14[:, a:z].astype(np.float64).reshape(len(got14), -1)
        r = np.asarray(want[k], np.float64).reshape(len(got14), -1)
        err = np.abs(g - r) / (1.0 + np.abs(r))
        print(f"[adam] {tag} {k} max err {err.max():.2e}", flush=True)
        assert err.max() <= 2e-6, (tag, k)
    idle = d[:, 13] == 0
    assert idle.any() and (~idle).any()
    np.testing.assert_array_equal(got14[idle], prev14[idle])
    assert np.abs(got14[~idle] - prev14[~idle]).max() > 0


def _check_state(tag, got27, want27):
    np.testing.assert_array_equal(got27[:, 26], want27[:, 26])
    for c in range(26):
        s = max(np.abs(want27[:, c]).max(), 1e-30)
        err = np.abs(got27[:, c] - want27[:, c]).max() / s
        assert err <= 1e-4, (tag, c, err)


def test_adam_two_steps_match_oracle():
    scene = S.make_scene(N, seed=11, width=W, height=H)
    intr = S.intrinsics(W, H)
    pose1 = S.perturbed_pose(1, rot_deg=1.0, trans_m=0.02)
    pose2 = S.perturbed_pose(2, rot_deg=3.0, trans_m=0.05)
    rgb, depth = _observation(scene, np.array([0, 0, 0, 0, 0, 0, 1.0]), intr)
    ctx = Context()
    ctx.set_scene(scene)
    ctx.set_frame(rgb, depth)
    with pytest.raises(FvgsError):
        ctx.adam_step(ETA)                       # no gradients yet
    m0 = ctx.read("MAP").copy()
    state = np.zeros((N, 27))
    # step 1
    ctx.mapping_step(pose1, intr)
    d1 = _d16(ctx.read("DELTA3D"))
    want1 = O.adam(_as_map(m0), state, d1, ETA)
    ctx.adam_step(ETA)
    m1 = ctx.read("MAP").copy()
    _check_map("step 1", m1, want1, m0, d1)
    _check_state("step 1", ctx.read("ADAM_STATE"), state)
    with pytest.raises(FvgsError):
        ctx.adam_step(ETA)                       # gradients are consumed by the update
    # step 2: the oracle continues from the GPU's map after step 1 and its own moments
    ctx.mapping_step(pose2, intr)
    d2 = _d16(ctx.read("DELTA3D"))
    want2 = O.adam(_as_map(m1), state, d2, ETA)
    ctx.adam_step(ETA)
    m2 = ctx.read("MAP").copy()
    _check_map("step 2", m2, want2, m1, d2)
    _check_state("step 2", ctx.read("ADAM_STATE"), state)
    assert state[:, 26].max() == 2
    # a new map resets the moments
    ctx.set_scene(scene)
    ctx.mapping_step(pose1, intr)
    ctx.adam_step(ETA)
    assert ctx.read("ADAM_STATE")[:, 26].max() == 1
    ctx.close()


def test_optimize_keyframe_updates_device_map_in_place():
    """fvgs_optimize_keyframe on a map set with FVGS_MEM_DEVICE updates the caller's arrays in
    place (as the reference updates gaussians_); here the arrays are another context's map
    copies. The result meets the oracle's bar for the gradients of that same step. (Two
    contexts' gradients differ in the order of the backward's float atomics, which Adam's
    normalisation amplifies where a Gaussian's gradient nearly cancels, so runs are not
    compared with each other.)"""
    scene = S.make_scene(N, seed=12, width=W, height=H)
    intr = S.intrinsics(W, H)
    pose = S.perturbed_pose(3, rot_deg=1.0, trans_m=0.02)
    rgb, depth = _observation(scene, np.array([0, 0, 0, 0, 0, 0, 1.0]), intr)
    owner = Context()
    owner.set_scene(scene)
    ctx = Context()
    ctx.set_scene(owner.map_device())
    ctx.set_frame(rgb, depth)
    ctx.optimize_keyframe(pose, intr, ETA)
    d = _d16(ctx.read("DELTA3D"))
    want = O.adam(scene, np.zeros((N, 27)), d, ETA)
    m = owner.read_map()
    got = np.concatenate([m[k].reshape(N, -1) for k, _, _ in FIELDS], axis=1)
    prev = np.concatenate([scene[k].reshape(N, -1) for k, _, _ in FIELDS], axis=1)
    _check_map("device map", got, want, prev, d)
    np.testing.assert_array_equal(ctx.read("MAP"), got)
    ctx.close()
    owner.close()


def test_adam_rejected_in_precision_form():
    scene = S.make_scene(2000, seed=13, width=W, height=H)
    intr = S.intrinsics(W, H)
    rgb, depth = _observation(scene, np.array([0, 0, 0, 0, 0, 0, 1.0]), intr)
    ctx = Context(precision=L.PRECISION_F16)
    ctx.set_scene(scene)
    ctx.set_frame(rgb, depth)
    with pytest.raises(FvgsError) as e:
        ctx.optimize_keyframe(np.array([0, 0, 0, 0, 0, 0, 1.0]), intr, ETA)
    assert e.value.code == L.FVGS_E_STATE
    ctx.close()
