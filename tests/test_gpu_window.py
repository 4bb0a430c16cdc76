"""GPU tests of the batched mapping window (fvgs_window.h) through the C ABI.

The window's summed 3D gradients equal the sum of plain per-view mapping steps (max-norm
relative 1e-5: float atomics of the backward add in a different order; n counts exact), on its
first run (which autotunes and captures the graph) and on a replay with new poses (the captured
kernels read the poses from device memory).
"""
import numpy as np
import pytest

from f_vigs_slam_amd import Context, Window
from f_vigs_slam_amd import scenes as S

pytestmark = pytest.mark.gpu

SIZES = [(320, 240), (320, 240), (424, 240), (640, 360)]


def _views(scene, poses):
    out = []
    ctx = Context()
    ctx.set_scene(scene)
    for p, (w, h) in zip(poses, SIZES):
        intr = S.intrinsics(w, h, f=0.8 * w)
        ctx.render_tracking(p, intr, w, h)
        rgb = (ctx.read("TRACK_RGB").reshape(h, w, 3) * 0.95).astype(np.float32)
        depth = (ctx.read("TRACK_DEPTH").reshape(h, w) + 0.01).astype(np.float32)
        out.append((intr, rgb, depth))
    ctx.close()
    return out


def _plain_sum(scene, poses, views):
    ctx = Context()
    ctx.set_scene(scene)
    acc = None
    for p, (intr, rgb, depth) in zip(poses, views):
        ctx.set_frame(rgb, depth)
        ctx.mapping_step(p, intr)
        g = ctx.read("DELTA3D")
        if acc is None:
            acc = {f: g[f].astype(np.float64) for f in g.dtype.names}
        else:
            for f in g.dtype.names:
                acc[f] = acc[f] + g[f]
    ctx.close()
    return acc


def _compare(got, want, tag):
    np.testing.assert_array_equal(got["n"], want["n"])
    assert want["n"].sum() > 0
    for f in ("position", "scale", "orientation", "color", "alpha"):
        ref = want[f]
        err = np.abs(got[f] - ref).max() / max(np.abs(ref).max(), 1e-30)
        print(f"[window] {tag} {f} rel err {err:.2e}", flush=True)
        assert err <= 1e-5


def test_window_equals_sum_of_plain_steps_and_replays_new_poses():
    scene = S.make_map_cube(100000, seed=3)
    poses = S.arc_poses(len(SIZES), radius=2.0, span_deg=60.0)
    views = _views(scene, poses)
    win = Window(max_views=len(SIZES))
    win.set_scene(scene)
    for k, (intr, rgb, depth) in enumerate(views):
        win.set_view(k, rgb, depth)
    intrs = [v[0] for v in views]
    win.run(poses, intrs)
    st = win.stats()
    print(f"[window] autotuner: {st}", flush=True)
    assert st["candidates"] >= 4 and st["best_ms"] <= st["worst_ms"]
    _compare(win.gradients(), _plain_sum(scene, poses, views), "first run")
    # same shapes, new poses: the captured graph is replayed
    poses2 = [np.concatenate([p[:3] + np.array([0.01, -0.02, 0.015]), p[3:]]) for p in poses]
    win.run(poses2, intrs)
    st2 = win.stats()
    assert st2["candidates"] == st["candidates"]
    _compare(win.gradients(), _plain_sum(scene, poses2, views), "replay")
    print(f"[window] replay {st2['window_ms']:.3f} ms for {len(SIZES)} views", flush=True)
    win.close()
