"""Python handle over fvgs_server (include/fvgs/fvgs_server.h): many sessions on one GPU.

Sessions keep their maps in device memory and are admitted against a memory budget; tracking
and mapping requests are queued and run on concurrent lanes (HIP streams). ``submit_track`` /
``submit_mapping`` return tickets, ``wait`` returns the result. No counterpart in the reference.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib as L
from .context import FvgsError, default_params, make_intr, make_pose


def _arr(a, dtype=np.float32):
    return np.ascontiguousarray(a, dtype)


class Server:
    def __init__(self, lanes=4, hbm_budget=0, max_sessions=64, **params):
        self._lib = L.load()
        p = L.ServerParams()
        self._lib.fvgs_server_default_params(ctypes.byref(p))
        p.lanes, p.max_sessions, p.hbm_budget = lanes, max_sessions, hbm_budget
        p.params = default_params(**params)
        self._h = ctypes.c_void_p()
        rc = self._lib.fvgs_server_create(ctypes.byref(p), ctypes.byref(self._h))
        if rc != L.FVGS_OK:
            raise FvgsError(rc, "fvgs_server_create failed (no GPU?)")
        self._keep = {}   # ticket -> host arrays that must outlive the request

    def _check(self, rc):
        if rc != L.FVGS_OK:
            raise FvgsError(rc, self._lib.fvgs_server_last_error(self._h).decode())

    def close(self):
        if self._h:
            self._lib.fvgs_server_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def open_session(self, scene, frame_w, frame_h):
        a = [_arr(scene[k]) for k in ("pos", "scale", "quat", "color", "opacity")]
        sid = ctypes.c_int(0)
        self._check(self._lib.fvgs_server_open_session(
            self._h, len(a[4]), *(L.ptr(x) for x in a), L.MEM_HOST, frame_w, frame_h,
            ctypes.byref(sid)))
        return sid.value

    def close_session(self, sid):
        self._check(self._lib.fvgs_server_close_session(self._h, sid))

    def add_keyframe(self, sid, pose7, rgb, depth):
        rgb, depth = _arr(rgb), _arr(depth)
        h, w = depth.shape
        kf = ctypes.c_int(0)
        self._check(self._lib.fvgs_server_add_keyframe(
            self._h, sid, ctypes.byref(make_pose(pose7)), w, h, L.ptr(rgb), L.ptr(depth),
            L.MEM_HOST, ctypes.byref(kf)))
        return kf.value

    def submit_track(self, sid, rgb, depth, intr, iters, pose7):
        """rgb / depth: host numpy arrays, or torch CUDA tensors (used in place)."""
        if hasattr(rgb, "data_ptr"):
            h, w = depth.shape
            rp, dp, mem = rgb.data_ptr(), depth.data_ptr(), L.MEM_DEVICE
            keep = (rgb, depth)
        else:
            rgb, depth = _arr(rgb), _arr(depth)
            h, w = depth.shape
            rp, dp, mem = rgb.ctypes.data, depth.ctypes.data, L.MEM_HOST
            keep = (rgb, depth)
        pose = _arr(pose7, np.float64)
        t = ctypes.c_uint64(0)
        self._check(self._lib.fvgs_server_submit_track(
            self._h, sid, w, h, ctypes.c_void_p(rp), ctypes.c_void_p(dp), mem,
            ctypes.byref(make_intr(intr)), iters, L.ptr(pose), ctypes.byref(t)))
        self._keep[t.value] = keep
        return t.value

    def submit_mapping(self, sid, intr, keyframes):
        kfs = np.ascontiguousarray(keyframes, np.int32)
        t = ctypes.c_uint64(0)
        self._check(self._lib.fvgs_server_submit_mapping(
            self._h, sid, ctypes.byref(make_intr(intr)), len(kfs), L.ptr(kfs), ctypes.byref(t)))
        return t.value

    def wait(self, ticket):
        """Returns a dict: status, kind, session, iters, pose7, queue_ms, run_ms, device_ms."""
        r = L.RequestResult()
        rc = self._lib.fvgs_server_wait(self._h, ticket, ctypes.byref(r))
        self._keep.pop(ticket, None)
        out = dict(status=r.status, kind=r.kind, session=r.session, iters=r.iters,
                   pose7=np.array(r.pose7[:]), queue_ms=r.queue_ms, run_ms=r.run_ms,
                   device_ms=r.device_ms)
        if rc != L.FVGS_OK and r.status == 0:
            self._check(rc)
        return out

    def gradients(self, sid, n):
        out = np.zeros(n, L.DELTA3D_DTYPE)
        self._check(self._lib.fvgs_server_read_gradients(self._h, sid, L.ptr(out), out.nbytes))
        return out

    def memory(self):
        s, l, b = ctypes.c_size_t(0), ctypes.c_size_t(0), ctypes.c_size_t(0)
        self._check(self._lib.fvgs_server_memory(self._h, ctypes.byref(s), ctypes.byref(l),
                                                 ctypes.byref(b)))
        return dict(sessions=s.value, lanes=l.value, budget=b.value)
