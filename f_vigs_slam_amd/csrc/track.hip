// RGB-D pose Jacobians for gfx950.
//
// Replaces getRgbdPoseJacobians (GSCudaKernels.cu:922-1392) and the host-side reduction of
// GSSlam::computeRgbdPoseJacobians (GSSlam.cu:2167-2226).
//
// k_pose: one workgroup per 16x16 tile, one thread per pixel. Pass 1 re-renders the pixel
// (LDS-staged records, wave-uniform early exit), pass 2 walks the same list again for the
// per-Gaussian terms. Each thread accumulates its 21 + 6 products in registers; the tile's
// partial is a wave butterfly reduction plus a fixed-order sum of the four waves, written
// once per tile (no atomics, deterministic). k_pose_reduce sums the tile partials in FP64 in
// tile order. The reference's divergent barriers inside `if (inside)` (Q12) are replaced by a
// block-uniform batch loop in which every pixel sees the whole list.
#include "fvgs_internal.h"
#include "gn_step.h"

namespace fvgs {

namespace {

constexpr float kAlphaMin = 1.0f / 255.0f;

__device__ __forceinline__ void qmat_t(float x, float y, float z, float w, float R[3][3]) {
    R[0][0] = 1.0f - 2.0f * (y * y + z * z);
    R[0][1] = 2.0f * (x * y - w * z);
    R[0][2] = 2.0f * (x * z + w * y);
    R[1][0] = 2.0f * (x * y + w * z);
    R[1][1] = 1.0f - 2.0f * (x * x + z * z);
    R[1][2] = 2.0f * (y * z - w * x);
    R[2][0] = 2.0f * (x * z - w * y);
    R[2][1] = 2.0f * (y * z + w * x);
    R[2][2] = 1.0f - 2.0f * (x * x + y * y);
}

// acc[0..5] += Jpose * b ; acc[6..26] += upper(Jpose * H * Jpose^T) with
// Jpose = [R ; d * [ray]x] (6x3) (GSCudaKernels.cu:1116-1165, 1293-1336). Each term is formed in
// FP32 as in the reference; the running sums are FP64: JtJ sums terms spanning ~10 orders of
// magnitude (pass-2 terms carry 1/(1 - alpha)), and FP32 running sums lose positive
// definiteness.
__device__ __forceinline__ void accumulate_pose(double acc[27], const float H[9], const float b3[3],
                                                const float R[3][3], float d, const float ray[3]) {
    float Jp[18];
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int k = 0; k < 3; ++k) Jp[r * 3 + k] = R[r][k];
    // d * [ray]x = d * [[0, -rz, ry], [rz, 0, -rx], [-ry, rx, 0]]
    Jp[9] = 0.f;           Jp[10] = -d * ray[2];  Jp[11] = d * ray[1];
    Jp[12] = d * ray[2];   Jp[13] = 0.f;          Jp[14] = -d * ray[0];
    Jp[15] = -d * ray[1];  Jp[16] = d * ray[0];   Jp[17] = 0.f;
    float tmp[18];
#pragma unroll
    for (int r = 0; r < 6; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c)
            tmp[r * 3 + c] = Jp[r * 3] * H[c] + Jp[r * 3 + 1] * H[3 + c] + Jp[r * 3 + 2] * H[6 + c];
    int id = 6;
#pragma unroll
    for (int r = 0; r < 6; ++r)
#pragma unroll
        for (int c = r; c < 6; ++c)
            acc[id++] += (double)(tmp[r * 3] * Jp[c * 3] + tmp[r * 3 + 1] * Jp[c * 3 + 1] +
                                  tmp[r * 3 + 2] * Jp[c * 3 + 2]);
#pragma unroll
    for (int r = 0; r < 6; ++r)
        acc[r] += (double)(Jp[r * 3] * b3[0] + Jp[r * 3 + 1] * b3[1] + Jp[r * 3 + 2] * b3[2]);
}

__global__ __launch_bounds__(256) void k_pose(View vh, const View* __restrict__ vd,
                                              const uint32_t* __restrict__ ranges,
                                              const uint32_t* __restrict__ idx,
                                              const Rec* __restrict__ rec,
                                              const float* __restrict__ obs_rgb,
                                              const float* __restrict__ obs_depth,
                                              const float* __restrict__ gxi,
                                              const float* __restrict__ gyi, BlendParams bp,
                                              double* __restrict__ tiles27,
                                              float* __restrict__ out_rgb,
                                              float* __restrict__ out_depth) {
    const View v = vd ? *vd : vh;   // device-resident view (pose updated on the device) or argument
    __shared__ Rec recs[256];
    __shared__ double red[4][27];
    const int tile = blockIdx.x, tid = threadIdx.x;
    const int x = (tile % v.ntx) * kTile + (tid & 15), y = (tile / v.ntx) * kTile + (tid >> 4);
    const bool in_bounds = x < v.W && y < v.H;
    const size_t p = in_bounds ? (size_t)y * v.W + x : 0;
    // Render-only mode (obs_depth == nullptr): every in-bounds pixel renders; no Jacobians.
    const bool render_only = obs_depth == nullptr;
    const float img_depth = in_bounds ? (render_only ? 1.0f : obs_depth[p]) : 0.0f;
    bool inside = in_bounds && img_depth > 0.1f;
    const uint32_t s = ranges[2 * tile], n = ranges[2 * tile + 1] - s;
    const float px = (float)x, py = (float)y;

    double acc[27];
#pragma unroll
    for (int k = 0; k < 27; ++k) acc[k] = 0.0;

    // ---------------- pass 1: render (GSCudaKernels.cu:978-1036) ----------------
    float C0 = 0.f, C1 = 0.f, C2 = 0.f, D = 0.f, final_T = 1.0f;
    uint32_t n_contrib = 0;
    if (n > 0) {
        float T = 1.f;
        uint32_t contributor = 0, last = 0;
        bool done = !inside;
        for (uint32_t base = 0; base < n; base += 256) {
            if (__syncthreads_count(!done) == 0) break;
            if (base + tid < n) recs[tid] = rec[idx[s + base + tid]];
            __syncthreads();
            const int cnt = (int)min(256u, n - base);
            if (!done) {
                for (int i = 0; i < cnt; ++i) {
                    ++contributor;
                    const float4 r0 = recs[i].r0, r1 = recs[i].r1;
                    const float dx = r0.x - px, dy = r0.y - py;
                    const float q = r0.z * dx * dx + 2.0f * r0.w * dx * dy + r1.x * dy * dy;  // Q14
                    const float4 r2 = recs[i].r2;
                    const float a = fminf(0.99f, r2.w * __expf(-0.5f * q));
                    if (a < kAlphaMin || q <= 0.0f) continue;
                    const float tT = T * (1.0f - a);
                    if (tT < 0.0001f) { done = true; break; }
                    C0 += r2.x * a * T; C1 += r2.y * a * T; C2 += r2.z * a * T;
                    const float d = r1.y + dx * r1.z + dy * r1.w;
                    if (T > 0.5f && tT < 0.5f) D = d;
                    T = tT;
                    last = contributor;
                }
            }
        }
        if (inside) {
            final_T = T;
            n_contrib = last;
            C0 += T * bp.bg[0]; C1 += T * bp.bg[1]; C2 += T * bp.bg[2];
        }
    }
    if (out_rgb && in_bounds) {
        out_rgb[3 * p] = C0; out_rgb[3 * p + 1] = C1; out_rgb[3 * p + 2] = C2;
        out_depth[p] = D;
    }
    if (render_only) return;   // uniform over the block
    inside = inside && (final_T < bp.alpha_thresh);

    float R[3][3];
    qmat_t(v.qx, v.qy, v.qz, v.qw, R);
    float e0 = 0.f, e1 = 0.f, e2 = 0.f, de = 0.f, wc = 0.f;
    if (inside) {
        // ---------------- pixel term (GSCudaKernels.cu:1044-1168) ----------------
        e0 = C0 - obs_rgb[3 * p]; e1 = C1 - obs_rgb[3 * p + 1]; e2 = C2 - obs_rgb[3 * p + 2];
        de = (img_depth > 0.1f) ? (D - img_depth) : 0.0f;
        const float cl = sqrtf(e0 * e0 + e1 * e1 + e2 * e2);
        wc = (cl < bp.color_thresh) ? 1.0f : bp.color_thresh / cl;
        const float ray[3] = {(px - v.cx) / v.fx, (py - v.cy) / v.fy, 1.0f};
        const float gX[3] = {gxi[3 * p], gxi[3 * p + 1], gxi[3 * p + 2]};
        const float gY[3] = {gyi[3 * p], gyi[3 * p + 1], gyi[3 * p + 2]};
        const float iid = 1.0f / img_depth;
        const float Jt[6] = {v.fx * iid, 0.f, 0.f, v.fy * iid, -v.fx * ray[0] * iid,
                             -v.fy * ray[1] * iid};
        float A[9];
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int k = 0; k < 3; ++k) A[r * 3 + k] = Jt[2 * r] * gX[k] + Jt[2 * r + 1] * gY[k];
        float H[9], b3[3];
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c)
                H[r * 3 + c] = wc * (A[r * 3] * A[c * 3] + A[r * 3 + 1] * A[c * 3 + 1] +
                                     A[r * 3 + 2] * A[c * 3 + 2]);
#pragma unroll
        for (int r = 0; r < 3; ++r) b3[r] = -wc * (A[r * 3] * e0 + A[r * 3 + 1] * e1 + A[r * 3 + 2] * e2);
        const float ld = fabsf(de);
        float wd = (ld < bp.depth_thresh) ? 1.0f : bp.depth_thresh / ld;
        wd /= img_depth;
        wd *= bp.depth_weight;
#pragma unroll
        for (int r = 0; r < 3; ++r) {
#pragma unroll
            for (int c = 0; c < 3; ++c) H[r * 3 + c] += wd * ray[r] * ray[c];
            b3[r] += wd * de * ray[r];
        }
        accumulate_pose(acc, H, b3, R, img_depth, ray);
    }

    // ---------------- pass 2: per-Gaussian terms (GSCudaKernels.cu:1174-1348) ----------------
    float prod = 1.f, T2 = 1.f, a0 = 0.f, a1 = 0.f, a2 = 0.f;
    bool active = inside && n_contrib > 0;
    for (uint32_t base = 0; base < n; base += 256) {
        const bool want = active && base < n_contrib;
        if (__syncthreads_count(want) == 0) break;
        if (base + tid < n) recs[tid] = rec[idx[s + base + tid]];
        __syncthreads();
        if (!want) continue;
        const int cnt = (int)min(256u, n - base);
        for (int i = 0; i < cnt && base + i < n_contrib; ++i) {
            const float4 r0 = recs[i].r0, r1 = recs[i].r1, r2 = recs[i].r2;
            const float dx = r0.x - px, dy = r0.y - py;
            const float q = r0.z * dx * dx + 2.0f * r0.w * dx * dy + r1.x * dy * dy;
            const float a = fminf(0.99f, r2.w * __expf(-0.5f * q));   // Q8: no q <= 0 test
            // Q15: a culled Gaussian (zero record, depth 0) would divide by zero depth below.
            if (a < kAlphaMin || r1.y == 0.0f) continue;
            const float d = r1.y + dx * r1.z + dy * r1.w;
            float da0 = r2.x * prod, da1 = r2.y * prod, da2 = r2.z * prod;
            a0 += a * da0; a1 += a * da1; a2 += a * da2;
            const float iom = 1.0f / (1.0f - a);
            da0 -= (C0 - a0) * iom;
            da1 -= (C1 - a1) * iom;
            da2 -= (C2 - a2) * iom;
            const float dm0 = r0.z * dx + r0.w * dy, dm1 = r0.w * dx + r1.x * dy;
            const float rg[3] = {(r0.x - v.cx) / v.fx, (r0.y - v.cy) / v.fy, 1.0f};
            const float id_ = 1.0f / d;
            const float Jt0 = v.fx * id_, Jt3 = v.fy * id_, Jt4 = -v.fx * rg[0] * id_,
                        Jt5 = -v.fy * rg[1] * id_;
            const float jd[3] = {Jt0 * dm0, Jt3 * dm1, Jt4 * dm0 + Jt5 * dm1};
            float Jc[9];
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                Jc[r * 3] = jd[r] * a * da0;
                Jc[r * 3 + 1] = jd[r] * a * da1;
                Jc[r * 3 + 2] = jd[r] * a * da2;
            }
            float H[9], b3[3];
#pragma unroll
            for (int r = 0; r < 3; ++r)
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    H[r * 3 + c] = wc * (Jc[r * 3] * Jc[c * 3] + Jc[r * 3 + 1] * Jc[c * 3 + 1] +
                                         Jc[r * 3 + 2] * Jc[c * 3 + 2]);
#pragma unroll
            for (int r = 0; r < 3; ++r)
                b3[r] = -wc * (Jc[r * 3] * e0 + Jc[r * 3 + 1] * e1 + Jc[r * 3 + 2] * e2);
            prod *= (1.0f - a);
            if (T2 > 0.5f && prod <= 0.5f && img_depth > 0.5f) {
                T2 = prod;
                const float ld = fabsf(de);
                float wd = (ld < bp.depth_thresh) ? 1.0f : bp.depth_thresh / ld;
                wd /= img_depth;
                wd *= bp.depth_weight;
#pragma unroll
                for (int r = 0; r < 3; ++r) {
#pragma unroll
                    for (int c = 0; c < 3; ++c) H[r * 3 + c] += wd * rg[r] * rg[c];
                    b3[r] += wd * de * rg[r];
                }
            }
            accumulate_pose(acc, H, b3, R, d, rg);
            if (prod < 0.001f) { active = false; break; }
        }
    }

    // ---------------- tile reduction (fixed order) ----------------
    const int w = tid >> 6, lane = tid & 63;
#pragma unroll
    for (int k = 0; k < 27; ++k) {
        double vk = acc[k];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) vk += __shfl_xor(vk, o);
        if (lane == 0) red[w][k] = vk;
    }
    __syncthreads();
    if (tid < 27) tiles27[(size_t)tile * 27 + tid] = ((red[0][tid] + red[1][tid]) + red[2][tid]) + red[3][tid];
}

// FP64 sum of the per-tile partials: block k sums component k over tiles in a fixed order.
__global__ __launch_bounds__(256) void k_pose_reduce(const double* __restrict__ tiles27, int ntiles,
                                                     double* __restrict__ out27) {
    __shared__ double part[256];
    const int k = blockIdx.x, tid = threadIdx.x;
    double sum = 0.0;
    for (int t = tid; t < ntiles; t += 256) sum += tiles27[(size_t)t * 27 + k];
    part[tid] = sum;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (tid < off) part[tid] += part[tid + off];
        __syncthreads();
    }
    if (tid == 0) out27[k] = part[0];
}

// 3x3 Sobel of an interleaved RGB float image, scale 1, BORDER_REFLECT_101 (the tracking
// gradients the reference builds with cv::Sobel(CV_32F, 1, 0, 3) / (0, 1, 3),
// GSSlam.cu:1959-1962). One thread per pixel, 16x16 blocks.
__device__ __forceinline__ int reflect101(int i, int n) {
    if (n == 1) return 0;
    if (i < 0) i = -i;
    if (i >= n) i = 2 * n - 2 - i;
    return i;
}

__global__ __launch_bounds__(256) void k_sobel(const float* __restrict__ img, int W, int H,
                                               float* __restrict__ gx, float* __restrict__ gy) {
    const int x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= W || y >= H) return;
    const int xm = reflect101(x - 1, W), xp = reflect101(x + 1, W);
    const int ym = reflect101(y - 1, H), yp = reflect101(y + 1, H);
    const size_t p = (size_t)y * W + x;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        auto P = [&](int yy, int xx) { return img[3 * ((size_t)yy * W + xx) + c]; };
        const float sx = (P(ym, xp) - P(ym, xm)) + 2.0f * (P(y, xp) - P(y, xm)) + (P(yp, xp) - P(yp, xm));
        const float sy = (P(yp, xm) - P(ym, xm)) + 2.0f * (P(yp, x) - P(ym, x)) + (P(yp, xp) - P(ym, xp));
        gx[3 * p + c] = sx;
        gy[3 * p + c] = sy;
    }
}

// One Gauss-Newton step of the device-resident tracking loop: the 27 FP64 sums (Jtr, then the
// JtJ upper triangle) -> Cholesky solve + Plus on the FP64 pose, and the float pose of the view
// that the next iteration's kernels read. After a failed step (JtJ not positive definite) the
// pose stays as it was and later steps are skipped (*fail is sticky); *status records 0 (step
// taken), -5 (not positive definite) or 1 (skipped).
__global__ __launch_bounds__(64) void k_gn_step(const double* __restrict__ sum27,
                                                double* __restrict__ pose7, View* __restrict__ vd,
                                                int* __restrict__ fail, int* __restrict__ status) {
    if (threadIdx.x != 0) return;
    if (*fail) { *status = 1; return; }
    double p[7];
    for (int k = 0; k < 7; ++k) p[k] = pose7[k];
    if (!gn_solve_plus(sum27 + 6, sum27, p, nullptr)) {
        *fail = 1;
        *status = -5;
        return;
    }
    for (int k = 0; k < 7; ++k) pose7[k] = p[k];
    vd->px = (float)p[0]; vd->py = (float)p[1]; vd->pz = (float)p[2];
    vd->qx = (float)p[3]; vd->qy = (float)p[4]; vd->qz = (float)p[5]; vd->qw = (float)p[6];
    *status = 0;
}

}  // namespace

hipError_t launch_sobel(const float* img, int W, int H, float* gx, float* gy, hipStream_t s) {
    hipLaunchKernelGGL(k_sobel, dim3((W + 15) / 16, (H + 15) / 16), dim3(256), 0, s, img, W, H,
                       gx, gy);
    return hipGetLastError();
}

hipError_t launch_pose(const View& v, const View* vd, const uint32_t* ranges,
                       const uint32_t* idx, const Rec* rec, const float* obs_rgb,
                       const float* obs_depth, const float* gx, const float* gy,
                       const BlendParams& bp, double* tiles27, float* out_rgb, float* out_depth,
                       hipStream_t s) {
    hipLaunchKernelGGL(k_pose, dim3(v.ntx * v.nty), dim3(256), 0, s, v, vd, ranges, idx, rec,
                       obs_rgb, obs_depth, gx, gy, bp, tiles27, out_rgb, out_depth);
    return hipGetLastError();
}

hipError_t launch_gn_step(const double* sum27, double* pose7, View* vd, int* fail, int* status,
                          hipStream_t s) {
    hipLaunchKernelGGL(k_gn_step, dim3(1), dim3(64), 0, s, sum27, pose7, vd, fail, status);
    return hipGetLastError();
}

hipError_t launch_pose_reduce(const double* tiles27, int ntiles, double* out27, hipStream_t s) {
    hipLaunchKernelGGL(k_pose_reduce, dim3(27), dim3(256), 0, s, tiles27, ntiles, out27);
    return hipGetLastError();
}

}  // namespace fvgs
