"""Python handle over fvgs_window (include/fvgs/fvgs_window.h): a batched mapping window.

Several keyframe views of one map run forward + backward + 2D->3D as one captured hipGraph
(autotuned schedule per window shape); the per-Gaussian 3D gradients are summed over the views
(run), or the views are optimised in sequence with Adam on the shared map, the reference's
optimizeGaussians loop, replayed as one captured hipGraph per sequence shape (optimize).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib as L
from .context import FvgsError, default_params, make_intr, make_pose


class Window:
    def __init__(self, max_views=8, **params):
        self._lib = L.load()
        self._h = ctypes.c_void_p()
        p = default_params(**params)
        rc = self._lib.fvgs_window_create(ctypes.byref(p), max_views, ctypes.byref(self._h))
        if rc != L.FVGS_OK:
            raise FvgsError(rc, "fvgs_window_create failed (no GPU?)")
        self.n = 0

    def _check(self, rc):
        if rc != L.FVGS_OK:
            raise FvgsError(rc, self._lib.fvgs_window_last_error(self._h).decode())

    def close(self):
        if self._h:
            self._lib.fvgs_window_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_scene(self, scene):
        a = [np.ascontiguousarray(scene[k], np.float32)
             for k in ("pos", "scale", "quat", "color", "opacity")]
        self._check(self._lib.fvgs_window_set_gaussians(self._h, len(a[4]),
                                                        *(L.ptr(x) for x in a), L.MEM_HOST))
        self.n = len(a[4])

    def set_view(self, slot, rgb, depth):
        rgb = np.ascontiguousarray(rgb, np.float32)
        depth = np.ascontiguousarray(depth, np.float32)
        h, w = depth.shape
        self._check(self._lib.fvgs_window_set_view(self._h, slot, w, h, L.ptr(rgb), L.ptr(depth),
                                                   L.MEM_HOST))

    def run(self, poses, intrs):
        nv = len(poses)
        P = (L.Pose * nv)(*[make_pose(p) for p in poses])
        I = (L.Intrinsics * nv)(*[make_intr(i) for i in intrs])
        self._check(self._lib.fvgs_window_run(self._h, nv, ctypes.cast(P, ctypes.c_void_p),
                                              ctypes.cast(I, ctypes.c_void_p)))

    def optimize(self, slots, poses, intrs, eta):
        """fvgs_window_optimize: optimizeGaussiansKeyframe on slots[s] at poses[s] for each
        step s in order, Adam on the shared map (GSSlam::optimizeGaussians' loop)."""
        ns = len(slots)
        S = (ctypes.c_int * ns)(*[int(v) for v in slots])
        P = (L.Pose * ns)(*[make_pose(p) for p in poses])
        I = (L.Intrinsics * ns)(*[make_intr(i) for i in intrs])
        self._check(self._lib.fvgs_window_optimize(self._h, ns, ctypes.cast(S, ctypes.c_void_p),
                                                   ctypes.cast(P, ctypes.c_void_p),
                                                   ctypes.cast(I, ctypes.c_void_p), float(eta)))

    def read_map(self):
        """The window's map as a scene dict (pos, scale, quat, color, opacity)."""
        out = dict(pos=np.zeros((self.n, 3), np.float32), scale=np.zeros((self.n, 3), np.float32),
                   quat=np.zeros((self.n, 4), np.float32), color=np.zeros((self.n, 3), np.float32),
                   opacity=np.zeros(self.n, np.float32))
        self._check(self._lib.fvgs_window_read_map(
            self._h, *(L.ptr(out[k]) for k in ("pos", "scale", "quat", "color", "opacity"))))
        return out

    def gradients(self):
        out = np.zeros(self.n, L.DELTA3D_DTYPE)
        self._check(self._lib.fvgs_window_read(self._h, L.ptr(out), out.nbytes))
        return out

    def stats(self):
        s = L.WindowStats()
        self._check(self._lib.fvgs_window_get_stats(self._h, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in L.WindowStats._fields_}
