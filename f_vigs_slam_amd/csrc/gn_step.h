// FP64 Gauss-Newton step shared by the host (fvgs_gn_step) and the device-resident tracking
// loop (k_gn_step), so that both give the same bits.
//
// Solves JtJ * d = Jtr by Cholesky (JtJ from its 21-entry upper triangle, row-major r <= c),
// then applies PoseLocalParameterization::Plus (PoseLocalParameterization.cpp:22-44) to
// pose7 = (px, py, pz, qx, qy, qz, qw): p += d[0:3], q = normalize(q * [1, d[3:6] / 2]).
// Returns false (pose untouched) when JtJ is not positive definite.
#pragma once
#include <hip/hip_runtime.h>

namespace fvgs {

__host__ __device__ inline bool gn_solve_plus(const double* jtj21, const double* jtr6,
                                              double* pose7, double* delta6) {
#pragma clang fp contract(off)
    double A[6][6];
    int id = 0;
    for (int r = 0; r < 6; r++)
        for (int cc = r; cc < 6; cc++) {
            A[r][cc] = jtj21[id];
            A[cc][r] = jtj21[id];
            id++;
        }
    double L[6][6];
    for (int r = 0; r < 6; r++)
        for (int cc = 0; cc < 6; cc++) L[r][cc] = 0.0;
    for (int j = 0; j < 6; j++) {
        double sacc = A[j][j];
        for (int k = 0; k < j; k++) sacc -= L[j][k] * L[j][k];
        if (!(sacc > 0.0) || !(sacc <= 1.7976931348623157e308)) return false;   // NaN, inf
        L[j][j] = sqrt(sacc);
        for (int i = j + 1; i < 6; i++) {
            double t = A[i][j];
            for (int k = 0; k < j; k++) t -= L[i][k] * L[j][k];
            L[i][j] = t / L[j][j];
        }
    }
    double y[6], d[6];
    for (int i = 0; i < 6; i++) {
        double t = jtr6[i];
        for (int k = 0; k < i; k++) t -= L[i][k] * y[k];
        y[i] = t / L[i][i];
    }
    for (int i = 5; i >= 0; i--) {
        double t = y[i];
        for (int k = i + 1; k < 6; k++) t -= L[k][i] * d[k];
        d[i] = t / L[i][i];
    }
    if (delta6)
        for (int i = 0; i < 6; i++) delta6[i] = d[i];
    pose7[0] += d[0];
    pose7[1] += d[1];
    pose7[2] += d[2];
    const double qx = pose7[3], qy = pose7[4], qz = pose7[5], qw = pose7[6];
    const double bx = 0.5 * d[3], by = 0.5 * d[4], bz = 0.5 * d[5];
    const double rw = qw - qx * bx - qy * by - qz * bz;
    const double rx = qw * bx + qx + qy * bz - qz * by;
    const double ry = qw * by - qx * bz + qy + qz * bx;
    const double rz = qw * bz + qx * by - qy * bx + qz;
    const double nrm = sqrt(rw * rw + rx * rx + ry * ry + rz * rz);
    pose7[3] = rx / nrm;
    pose7[4] = ry / nrm;
    pose7[5] = rz / nrm;
    pose7[6] = rw / nrm;
    return true;
}

}  // namespace fvgs
