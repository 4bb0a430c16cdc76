// Device helpers shared by the blend kernels (raster.hip) and the pose kernel (track.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "fvgs_internal.h"

namespace fvgs {

constexpr float kAlphaMin = 1.0f / 255.0f;   // GSCudaKernels.cu:2189, 974

// Wave-level culling for the blend walks. A wave owns an 8x8 pixel quadrant of the tile; an entry
// can change a pixel's state only where alpha >= 1/255, i.e. q <= 2 ln(255 opacity), and (in the
// forward) q > 0. Interval arithmetic bounds q = c0 dx^2 + 2 c1 dx dy + c2 dy^2 over the quadrant's
// dx, dy ranges; an entry whose bounds exclude every pixel of a quadrant (with a relative slack
// far above FP32 rounding of either evaluation) is skipped by that wave. Skipped entries leave
// every pixel of the wave untouched, so results are those of the full walk. NaN bounds keep the
// entry. Returns bit w set for each quadrant w whose wave must evaluate the entry.
// Pixel block of wave w (of 4) in a 16x16 tile: QW x (64 / QW) pixels, QW = 8 (2x2 quadrants)
// or 16 (four row bands).
template <int QW>
__device__ __forceinline__ int wave_x0(int w) { return (w % (kTile / QW)) * QW; }
template <int QW>
__device__ __forceinline__ int wave_y0(int w) { return (w / (kTile / QW)) * (64 / QW); }

// Whether the wave of pixel block w must evaluate the entry (r0.x, r0.y mean; r0.z, r0.w, r1.x
// conic; r2.w opacity).
template <int QW = 8>
__device__ __forceinline__ bool quadrant_hit(const float4& r0, const float4& r1, const float4& r2,
                                             int tx0, int ty0, int w, bool need_pos_q) {
    constexpr int QH = 64 / QW;
    const float u = r0.x, vv = r0.y, c0 = r0.z, c1x2 = 2.0f * r0.w, c2 = r1.x;
    const float qcut = 2.0f * __logf(255.0f * r2.w) + 1e-3f;
    const int bx = tx0 + wave_x0<QW>(w), by = ty0 + wave_y0<QW>(w);
    const float ax0 = (float)bx - u, ax1 = (float)(bx + QW - 1) - u;
    const float ay0 = (float)by - vv, ay1 = (float)(by + QH - 1) - vv;
    const float xa = ax0 * ax0, xb = ax1 * ax1, ya = ay0 * ay0, yb = ay1 * ay1;
    const float xxlo = (ax0 <= 0.0f && ax1 >= 0.0f) ? 0.0f : fminf(xa, xb), xxhi = fmaxf(xa, xb);
    const float yylo = (ay0 <= 0.0f && ay1 >= 0.0f) ? 0.0f : fminf(ya, yb), yyhi = fmaxf(ya, yb);
    const float p0 = ax0 * ay0, p1 = ax0 * ay1, p2 = ax1 * ay0, p3 = ax1 * ay1;
    const float xylo = fminf(fminf(p0, p1), fminf(p2, p3)), xyhi = fmaxf(fmaxf(p0, p1), fmaxf(p2, p3));
    const float t0lo = c0 >= 0.0f ? c0 * xxlo : c0 * xxhi, t0hi = c0 >= 0.0f ? c0 * xxhi : c0 * xxlo;
    const float t1lo = c1x2 >= 0.0f ? c1x2 * xylo : c1x2 * xyhi,
                t1hi = c1x2 >= 0.0f ? c1x2 * xyhi : c1x2 * xylo;
    const float t2lo = c2 >= 0.0f ? c2 * yylo : c2 * yyhi, t2hi = c2 >= 0.0f ? c2 * yyhi : c2 * yylo;
    const float qlo = t0lo + t1lo + t2lo, qhi = t0hi + t1hi + t2hi;
    const float slack = 1e-5f * (fabsf(t0lo) + fabsf(t0hi) + fabsf(t1lo) + fabsf(t1hi) +
                                 fabsf(t2lo) + fabsf(t2hi)) + 1e-30f;
    // A zero form (the culled Gaussians' zero record) gives q = 0 at every pixel exactly.
    const bool zero_form = c0 == 0.0f && c1x2 == 0.0f && c2 == 0.0f;
    const bool skip = (need_pos_q && (zero_form || qhi + slack <= 0.0f)) || (qlo - slack > qcut);
    return !skip;
}

template <int QW = 8>
__device__ __forceinline__ uint32_t quadrant_mask(const Rec& r, int tx0, int ty0, bool need_pos_q) {
    uint32_t mask = 0;
#pragma unroll 1
    for (int w = 0; w < 4; ++w)
        if (quadrant_hit<QW>(r.r0, r.r1, r.r2, tx0, ty0, w, need_pos_q)) mask |= 1u << w;
    return mask;
}

}  // namespace fvgs
