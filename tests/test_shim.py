"""The C++ GSSlam shim (f_vigs_slam_amd/compat/GSSlam.hpp) compiles against the C ABI with a
plain host compiler, and its Ceres-cost conditioning (RgbdPoseCost.cu:57-102) matches a numpy
eigen-decomposition. The GPU half (computeRgbdPoseJacobians and optimizeGaussiansKeyframe
through the shim, driven by tests/shim_driver.cpp) is in test_gpu_shim.py."""
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SHIM = os.path.join(ROOT, "f_vigs_slam_amd", "compat", "GSSlam.hpp")
LIBDIR = os.path.join(ROOT, "f_vigs_slam_amd")

PROG = r'''
#include <cstdio>
#include "GSSLAM_HPP"
// instantiates the keyframe loop (GSSlam::optimizeGaussians) with a caller-side sampler
void keyframe_loop(fvgs_compat::GSSlam& s) {
    s.optimizeGaussians(2, 0.002f, [](int count, int total) {
        std::vector<int> v;
        for (int i = 0; i < count && i < total; i++) v.push_back(i);
        return v;
    });
}
int main() {
    double A[36], b[6], r[6], J[42];
    for (int i = 0; i < 36; i++) if (scanf("%lf", &A[i]) != 1) return 2;
    for (int i = 0; i < 6; i++) if (scanf("%lf", &b[i]) != 1) return 2;
    fvgs_compat::rgbdPoseCost(A, b, r, J);
    for (int i = 0; i < 6; i++) printf("%.17g\n", r[i]);
    for (int i = 0; i < 42; i++) printf("%.17g\n", J[i]);
    return 0;
}
'''


def build(tmp_path):
    src = tmp_path / "cost.cpp"
    src.write_text(PROG.replace("GSSLAM_HPP", SHIM))
    exe = tmp_path / "cost"
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-Werror", str(src), "-o", str(exe),
                    f"-L{LIBDIR}", "-lfvgs", f"-Wl,-rpath,{LIBDIR}"], check=True)
    return exe


def test_rgbd_pose_cost_matches_numpy_eigh(tmp_path):
    exe = build(tmp_path)
    rng = np.random.default_rng(4)
    for rank in (6, 4):
        X = rng.standard_normal((6, rank))
        A = X @ X.T * 10.0
        b = rng.standard_normal(6)
        out = subprocess.run([str(exe)], input=" ".join(repr(float(x)) for x in [*A.ravel(), *b]),
                             capture_output=True, text=True, check=True).stdout.split()
        vals = np.array([float(v) for v in out])
        r, J = vals[:6], vals[6:].reshape(6, 7)
        assert np.all(J[:, 6] == 0)
        Js = J[:, :6]
        # J*^T J* reproduces JtJ (on its numerically non-null part), J*^T r* = projected Jtr
        np.testing.assert_allclose(Js.T @ Js, A, atol=1e-9 * np.abs(A).max())
        w, U = np.linalg.eigh(A)
        keep = w > 1e-4
        P = U[:, keep] @ U[:, keep].T
        np.testing.assert_allclose(Js.T @ r, P @ b, atol=1e-9)
        # the Gauss-Newton step of the cost equals the pseudo-inverse step
        step = -np.linalg.lstsq(Js, r, rcond=None)[0]
        np.testing.assert_allclose(step, -np.linalg.pinv(A, rcond=1e-10) @ b, atol=1e-8)
