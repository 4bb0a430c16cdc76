"""Pinning the oracle's map maintenance (GSCudaKernels.cu:87-311, 1455-2068) on the CPU: closed
forms on hand-built scenes, an independent FP64 restatement with the conic as a 2x2 matrix
(the reference reads the stored (xx, yy, xy) slots as [[xx, yy], [yy, xy]], Q14), and scipy
rotations for the initialisation's camera convention (its float4 read (w, x, y, z))."""
import numpy as np
import pytest
from scipy.spatial.transform import Rotation as Rot

import oracle as O
from f_vigs_slam_amd import scenes as S

W, H = 64, 48
INTR = np.array([60.0, 60.0, 31.5, 23.5], np.float32)
ID_POSE = (np.zeros(3, np.float32), np.array([0, 0, 0, 1], np.float32))


def _scene(pos, scale, opac, color=None):
    n = len(pos)
    return dict(pos=np.asarray(pos, np.float32), scale=np.asarray(scale, np.float32),
                quat=np.tile(np.array([0, 0, 0, 1], np.float32), (n, 1)),
                color=np.asarray(color if color is not None else np.full((n, 3), 0.5),
                                 np.float32),
                opacity=np.asarray(opac, np.float32))


def _view(sc):
    pf = O.project_f32(sc["pos"], sc["scale"], sc["quat"], *ID_POSE, INTR, W, H)
    return pf, O.bin_f32(pf, W, H)


def _pixel_of(p):
    return (int(round(INTR[0] * p[0] / p[2] + INTR[2])), int(round(INTR[1] * p[1] / p[2] + INTR[3])))


# ---- visibility ----------------------------------------------------------------------------
def test_visibility_occlusion_and_opacity():
    # a large near Gaussian in front of a small far one on the same ray
    sc = _scene([[0, 0, 1.0], [0, 0, 2.0]], [[0.2, 0.2, 0.2], [0.02, 0.02, 0.02]], [0.99, 0.9])
    pf, b = _view(sc)
    vis = O.visibility(W, H, pf, b, sc["opacity"])
    assert vis.tolist() == [1, 0]                      # T = 0.01 < 0.5 after the near one
    sc["opacity"][0] = 0.3                             # T = 0.7 stays above 0.5
    assert O.visibility(W, H, pf, b, sc["opacity"]).tolist() == [1, 1]
    sc["opacity"][:] = 0.0                             # alpha < 1/255 never marks (with Q14's
    assert O.visibility(W, H, pf, b, sc["opacity"]).tolist() == [0, 0]   # indefinite q, exp(-q/2)
    # can exceed 1, so only opacity 0 rules a Gaussian out everywhere)


def test_visibility_marks_culled_zero_records_in_tile0():
    """Q1: a culled Gaussian's zero record (tile 0, key 0) has q = 0 and counts at the map's
    opacity, as the reference's walk (no v <= 0 test) does."""
    sc = _scene([[0, 0, -1.0], [0.5, 0.3, 2.0]], [[0.05] * 3, [0.01] * 3], [0.3, 0.5])
    pf, b = _view(sc)
    assert b["ranges"][0, 0] == 0 and b["idx"][0] == 0     # zero segment first in tile 0
    vis = O.visibility(W, H, pf, b, sc["opacity"])
    assert vis[0] == 1
    sc["opacity"][0] = 0.003
    assert O.visibility(W, H, pf, b, sc["opacity"])[0] == 0


def _fp64_visibility(sc, pf, b):
    """Independent FP64 restatement: per pixel, the tile's first 256 entries in sorted order."""
    vis = np.zeros(len(sc["opacity"]), np.uint8)
    ntx = (W + 15) // 16
    for y in range(H):
        for x in range(W):
            t = (y // 16) * ntx + x // 16
            s, e = b["ranges"][t]
            T = 1.0
            for g in b["idx"][s:min(e, s + 256)]:
                xx, yy, xy = (float(c) for c in pf["conic"][g])
                C = np.array([[xx, yy], [yy, xy]])
                d = np.array([pf["pos2d"][g, 0] - x, pf["pos2d"][g, 1] - y], float)
                with np.errstate(over="ignore"):     # Q14's indefinite q
                    a = min(0.99, float(sc["opacity"][g]) * np.exp(-0.5 * d @ C @ d))
                if a < 1 / 255:
                    continue
                vis[g] = 1
                T *= 1 - a
                if T < 0.5:
                    break
    return vis


def test_visibility_matches_fp64_restatement():
    sc = S.make_scene(400, seed=8, width=W, height=H, intr=INTR)
    pf, b = _view(sc)
    got = O.visibility(W, H, pf, b, sc["opacity"])
    want = _fp64_visibility(sc, pf, b)
    assert got.sum() > 20
    assert (got != want).sum() <= 2          # FP32 vs FP64 threshold flips only


def test_visibility_walks_only_the_first_256_entries():
    # 256 small Gaussians at the tile centre, pixel (8, 8) (in front: keys use the depth at the
    # tile centre, Q2) and one strong one behind, centred on pixel (12, 12) of tile 0, where the
    # small ones are faint: beyond the 256-entry cap it is never marked.
    rng = np.random.default_rng(0)
    z_front = 1.0 + rng.uniform(0, 0.1, 256)
    front = np.stack([(8 - INTR[2]) * z_front / INTR[0], (8 - INTR[3]) * z_front / INTR[1],
                      z_front], 1)
    back = [[(12 - INTR[2]) * 2.0 / INTR[0], (12 - INTR[3]) * 2.0 / INTR[1], 2.0]]
    sc = _scene(np.concatenate([front, back]), [[0.004] * 3] * 256 + [[0.02] * 3],
                [0.5] * 256 + [0.9])
    pf, b = _view(sc)
    s, e = b["ranges"][0]
    assert e - s > 256 and b["idx"][s + 256] == 256
    assert O.visibility(W, H, pf, b, sc["opacity"])[256] == 0
    sc2 = {k: v[256:] for k, v in sc.items()}              # alone it is visible
    pf2, b2 = _view(sc2)
    assert O.visibility(W, H, pf2, b2, sc2["opacity"])[0] == 1


# ---- outliers --------------------------------------------------------------------------------
def test_outliers_single_gaussian_closed_form():
    p = [0.1, -0.05, 2.0]
    sc = _scene([p], [[0.03] * 3], [0.7])
    pf, b = _view(sc)
    far = np.full((H, W), 5.0, np.float32)                 # rendered 2 < 0.8 * 5: all outlier
    prob, total = O.outliers(W, H, pf, b, sc["opacity"], far)

    def fp64(obs):
        """Independent FP64 sums over the pixels of the tiles the Gaussian is listed in."""
        ntx = (W + 15) // 16
        wp = wt = 0.0
        xx, yy, xy = (float(c) for c in pf["conic"][0])
        C = np.array([[xx, yy], [yy, xy]])
        for t in range(len(b["ranges"])):
            s, e = b["ranges"][t]
            if 0 not in b["idx"][s:e]:
                continue
            tx, ty = (t % ntx) * 16, (t // ntx) * 16
            for y in range(ty, min(ty + 16, H)):
                for x in range(tx, min(tx + 16, W)):
                    d = np.array([x - pf["pos2d"][0, 0], y - pf["pos2d"][0, 1]], float)
                    a = 0.7 * np.exp(-0.5 * d @ C @ d)
                    dr = pf["depth"][0] + pf["phat"][0] @ d
                    if a >= 1 / 255 and dr > 0:
                        wt += a
                        if obs[y, x] > 0.1 and dr < 0.8 * obs[y, x]:
                            wp += a
        return wp, wt

    wp, wt = fp64(far)
    assert total[0] == pytest.approx(wt, rel=1e-5) and prob[0] == pytest.approx(wp, rel=1e-5)
    assert prob[0] > 0.5 * total[0]
    near = np.full((H, W), 2.2, np.float32)                # rendered ~2 vs 0.8 * 2.2
    prob2, total2 = O.outliers(W, H, pf, b, sc["opacity"], near)
    wp2, wt2 = fp64(near)
    assert total2[0] == total[0] and prob2[0] == pytest.approx(wp2, rel=1e-5, abs=1e-30)
    assert prob2[0] < prob[0]
    keep = O.outlier_keep(prob, total, 0.6)
    assert keep[0] == (0 if (total[0] > 1 and prob[0] / total[0] > 0.6) else 1)


# ---- density mask ----------------------------------------------------------------------------
def test_density_mask_closed_forms():
    p = [0.0, 0.0, 3.0]
    sc = _scene([p], [[0.2] * 3], [0.99])
    pf, b = _view(sc)
    x, y = _pixel_of(p)
    obs = np.full((H, W), 2.0, np.float32)
    m = O.density_mask(W, H, pf, b, sc["opacity"], obs)
    assert m[y, x] == 1.0                    # rendered 3 behind observed 2 by > 0.2 * 2
    assert m[0, 0] == 1.0                    # nothing rendered: prod = 1 > 0.5
    obs[:] = 4.0
    m = O.density_mask(W, H, pf, b, sc["opacity"], obs)
    assert m[y, x] == 0.0                    # rendered in front of the observation
    obs[:] = 0.4
    assert not O.density_mask(W, H, pf, b, sc["opacity"], obs).any()   # no valid depth
    sc["opacity"][0] = 0.3                   # prod = 0.7 > 0.5 at the centre: the value is prod
    obs[:] = 2.0
    m = O.density_mask(W, H, pf, b, sc["opacity"], obs)
    assert 0.5 < m[y, x] < 0.75             # ~0.7 (q at the pixel centre is not exactly 0)


# ---- prune -----------------------------------------------------------------------------------
def test_prune_rules():
    scale = np.array([[0.1, 0.1, 0.1], [0.1, 0.1, 0.1], [0.1, 0.004, 0.002], [0.004, 0.004, 0.004],
                      [0.1, 0.006, 0.1]], np.float32)
    opac = np.array([0.5, 0.04, 0.5, 0.5, 0.5], np.float32)
    # kept; low opacity; middle/max = 0.04 < 0.05; max < 0.005; middle/max = 1
    assert O.prune_keep(scale, opac).tolist() == [1, 0, 0, 0, 1]
    sc = dict(pos=np.arange(15, dtype=np.float32).reshape(5, 3), opacity=opac)
    kept = O.compact(sc, O.prune_keep(scale, opac))
    np.testing.assert_array_equal(kept["pos"], sc["pos"][[0, 4]])


# ---- normals, initialisation, densification -----------------------------------------------
def test_normals_of_planes():
    yy, xx = np.mgrid[0:H, 0:W].astype(np.float32)
    d = (2.0 + 0.01 * xx - 0.02 * yy).astype(np.float32)
    n = O.normals(d)
    want = np.array([-0.01, 0.02, 1.0]) / np.linalg.norm([-0.01, 0.02, 1.0])
    np.testing.assert_allclose(n[5:-5, 5:-5], np.broadcast_to(want, n[5:-5, 5:-5].shape),
                               atol=1e-6)
    np.testing.assert_array_equal(n[0, 0], [0, 0, 1])      # border
    d[10, 10] = 0.0                                         # invalid neighbour
    np.testing.assert_array_equal(O.normals(d)[10, 11], [0, 0, 1])


@pytest.mark.parametrize("rot", [None, [0.0, 0.4, 0.1]])
def test_init_from_rgbd_backprojects_in_the_cameras_frame(rot):
    rng = np.random.default_rng(2)
    bgr = rng.integers(0, 256, (H, W, 3), dtype=np.uint8)
    d = np.full((H, W), 2.0, np.float32)
    d[0, 10] = 0.0                                           # cell (1, 0) skipped
    R = Rot.identity() if rot is None else Rot.from_rotvec(rot)
    xyzw = R.as_quat().astype(np.float32)
    pose7 = np.array([0.1, -0.2, 0.3, xyzw[3], xyzw[0], xyzw[1], xyzw[2]], np.float32)
    out = O.init_from_rgbd(bgr, d, pose7, INTR, sample_px=10, init_scale=0.05,
                           init_opacity=0.8)
    cells = [(u, v) for v in range(0, H, 10) for u in range(0, W, 10) if (u, v) != (10, 0)]
    assert len(out["pos"]) == len(cells)
    uv = np.array(cells, float)
    pc = np.stack([(uv[:, 0] - INTR[2]) * 2.0 / INTR[0], (uv[:, 1] - INTR[3]) * 2.0 / INTR[1],
                   np.full(len(uv), 2.0)], 1)
    np.testing.assert_allclose(out["pos"], R.apply(pc) + pose7[:3], atol=2e-6)
    np.testing.assert_array_equal(out["scale"], np.float32(0.05))
    np.testing.assert_array_equal(out["opacity"], np.float32(0.8))
    px = np.array([bgr[v, u] for u, v in cells])
    np.testing.assert_allclose(out["color"], px[:, ::-1] / 255.0, atol=1e-7)   # B,G,R -> R,G,B
    # flat depth: normal (0, 0, 1) in the camera; the Gaussian's z axis is the world normal
    q = out["quat"]                                           # slots (w, x, y, z)
    z_axis = Rot.from_quat(np.stack([q[:, 1], q[:, 2], q[:, 3], q[:, 0]], 1)).apply([0, 0, 1])
    np.testing.assert_allclose(z_axis, np.broadcast_to(R.apply([0, 0, 1]), z_axis.shape),
                               atol=1e-6)


def test_densify_cell_and_pixel_modes():
    bgr = np.full((H, W, 3), (10, 20, 30), np.uint8)
    d = np.full((H, W), 2.0, np.float32)
    mask = np.zeros((H, W), np.float32)
    mask[0:10, 0:10] = 1.0                       # a full cell
    mask[10:20, 0:4] = 1.0                       # 40 of 100 pixels: below half, no Gaussian
    pose7 = np.array([0, 0, 0, 1, 0, 0, 0], np.float32)     # identity read (w, x, y, z)
    out = O.densify_cells(mask, bgr, d, pose7, INTR, sample_px=10)
    assert len(out["pos"]) == 1
    c = 4.5                                      # mean pixel of the cell
    np.testing.assert_allclose(out["pos"][0], [(c - INTR[2]) * 2 / INTR[0],
                                               (c - INTR[3]) * 2 / INTR[1], 2.0], atol=1e-6)
    s = 0.5 * (2.0 * 10 / INTR[0] + 2.0 * 10 / INTR[1])
    np.testing.assert_allclose(out["scale"][0], [s, s, 0.1 * s], rtol=1e-6)
    np.testing.assert_allclose(out["color"][0], [30 / 255, 20 / 255, 10 / 255], rtol=1e-5)
    np.testing.assert_array_equal(out["quat"][0], pose7[3:])   # camera float4 times identity
    assert out["opacity"][0] == 1.0
    one = O.densify_cells(mask, bgr, d, pose7, INTR, sample_px=1)
    assert len(one["pos"]) == 140
    s1 = 0.5 * (0.8 * 2.0 / INTR[0] + 0.8 * 2.0 / INTR[1])
    np.testing.assert_allclose(one["scale"][0], [s1, s1, 0.1 * s1], rtol=1e-6)
