"""GPU parity of the raw-image frame pyramid (fvgs_set_frame_raw / fvgs_select_level, the device
form of GSSlam::initAndCopyImgs, GSSlam.cu:1913-1969) against the oracle, through the C ABI.

Bars: colour and depth of every level bit-exact (integer 8-bit pyrDown; float pyrDown in the
same operation order, no contraction; single-rounding conversions); Sobel gradients within
2e-6 (the oracle sums in FP64); pose Jacobians on a selected level bit-identical to the same
level's images given through fvgs_set_frame (same kernels on the same data).
"""
import numpy as np
import pytest

import oracle as O
from f_vigs_slam_amd import Context, FvgsError
from f_vigs_slam_amd import scenes as S

pytestmark = pytest.mark.gpu


def _images(h, w, seed, u16=True):
    rng = np.random.default_rng(seed)
    col = rng.integers(0, 256, (h, w, 3), dtype=np.uint8)
    if u16:
        depth = rng.integers(300, 6000, (h, w), dtype=np.uint16)
        depth[rng.random((h, w)) < 0.02] = 0            # missing depth
    else:
        depth = rng.uniform(0.3, 6.0, (h, w)).astype(np.float32)
    return col, depth


@pytest.mark.parametrize("h,w,u16,levels", [(480, 640, True, 3), (479, 641, False, 4),
                                            (37, 5, True, 3)])
def test_frame_pyramid_matches_oracle(h, w, u16, levels):
    col, depth = _images(h, w, seed=h + w, u16=u16)
    want = O.frame_pyramid(col, depth, levels)
    with Context() as ctx:
        ctx.set_frame_raw(col, depth, levels=levels)
        for lv in range(levels):
            lw, lh = ctx.select_level(lv)
            rgb, dep, gx, gy = want[lv]
            assert (lh, lw) == dep.shape
            print(f"[frame] {w}x{h} level {lv}: {lw}x{lh}", flush=True)
            np.testing.assert_array_equal(ctx.read("FRAME_RGB"), rgb)
            np.testing.assert_array_equal(ctx.read("FRAME_DEPTH"), dep)
            np.testing.assert_allclose(ctx.read("FRAME_GX"), gx, rtol=0, atol=2e-6)
            np.testing.assert_allclose(ctx.read("FRAME_GY"), gy, rtol=0, atol=2e-6)
        with pytest.raises(FvgsError):
            ctx.select_level(levels)


def test_pose_jacobians_on_a_selected_level():
    w, h = 640, 480
    scene = S.make_scene(30000, seed=5, width=w, height=h)
    intr = S.intrinsics(w, h)
    pose = np.array([0, 0, 0, 0, 0, 0, 1.0])
    with Context() as c:
        c.set_scene(scene)
        c.render_tracking(pose, intr, w, h)
        col = np.clip(c.read("TRACK_RGB").reshape(h, w, 3) * 255.0 + 0.5, 0, 255).astype(np.uint8)
        dmm = np.clip(c.read("TRACK_DEPTH").reshape(h, w) * 1000.0, 0, 65535).astype(np.uint16)
    level = 1
    s = 1.0 / (1 << level)
    intr_l = np.asarray(intr, np.float32) * np.float32(s)
    with Context() as a, Context() as b:
        a.set_scene(scene)
        a.set_frame_raw(col, dmm, levels=3)
        lw, lh = a.select_level(level)
        jtj_a, jtr_a = a.pose_jacobians(pose, intr_l)
        rgb, dep, gx, gy = (a.read("FRAME_RGB").copy(), a.read("FRAME_DEPTH").copy(),
                            a.read("FRAME_GX").copy(), a.read("FRAME_GY").copy())
        b.set_scene(scene)
        b.set_frame(rgb, dep, gx, gy)
        jtj_b, jtr_b = b.pose_jacobians(pose, intr_l)
    print(f"[frame] level {level} {lw}x{lh}: |JtJ| {np.abs(jtj_a).max():.3e}", flush=True)
    assert np.abs(jtj_a).max() > 0
    np.testing.assert_array_equal(jtj_a, jtj_b)
    np.testing.assert_array_equal(jtr_a, jtr_b)
