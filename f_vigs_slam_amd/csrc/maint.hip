// Map maintenance on the device: visibility and covisibility of a view, prune, outlier removal,
// density mask and densification, normals from depth and initialisation from RGB-D
// (GSCudaKernels.cu:87-311, 1455-2068; host sequences GSSlam.cu:271-406, 1351-1904).
//
// The walks read the screen records, tile ranges and sorted indices that the binning built for
// the view, and the opacity from the map as the reference does (so the culled Gaussians' zero
// records in tile 0 count at full opacity in the visibility and density walks, Q1). Arithmetic
// is FP32 in the reference's operation order with contraction off. Appends (densify,
// initialisation) and removals (prune, outliers) are deterministic: new Gaussians in cell order
// and survivors in their previous order, placed by a prefix sum instead of an atomic counter.
#include <hip/hip_runtime.h>

#include "fvgs_internal.h"

// FP32 in the reference's operation order throughout this file (helpers included): no FMA
// contraction, so results match the oracle's -ffp-contract=off restatement bit for bit.
#pragma clang fp contract(off)

namespace fvgs {
namespace {

constexpr float kMinA = 1.0f / 255.0f;   // GSCudaKernels.cu:1520, 1697, 1813

struct Entry {   // one staged list entry: centre, conic as stored, map opacity, Gaussian
    float u, v, c0, c1, c2, o;
    uint32_t g;
    float pad;
};

__device__ __forceinline__ Entry stage(const Rec* rec, const MapRef& map, uint32_t g) {
    const Rec r = rec[g];
    float c[3], o;
    map.appearance(g, c, o);
    Entry e;
    e.u = r.r0.x; e.v = r.r0.y; e.c0 = r.r0.z; e.c1 = r.r0.w; e.c2 = r.r1.x;
    e.o = o; e.g = g; e.pad = 0.f;
    return e;
}

// computeGaussiansVisibility_kernel (GSCudaKernels.cu:1455-1537): the first min(len, 256)
// entries of each tile; per pixel, front to back, alpha = min(0.99, o exp(-q/2)) with
// q = c0 dx dx + 2 c1 dx dy + c2 dy dy, dx = u - x; an entry with alpha >= 1/255 marks its
// Gaussian visible and T *= 1 - alpha; the pixel stops once T < 0.5.
__global__ __launch_bounds__(256) void k_visibility(View v, const uint32_t* __restrict__ ranges,
                                                    const uint32_t* __restrict__ idx,
                                                    const Rec* __restrict__ rec, MapRef map,
                                                    uint8_t* __restrict__ vis) {
    __shared__ Entry es[256];
    const int tile = blockIdx.x, tid = threadIdx.x;
    const int x = (tile % v.ntx) * kTile + (tid & 15), y = (tile / v.ntx) * kTile + (tid >> 4);
    const uint32_t s = ranges[2 * tile];
    const int n = (int)min(ranges[2 * tile + 1] - s, 256u);
    if (tid < n) es[tid] = stage(rec, map, idx[s + tid]);
    __syncthreads();
    if (x >= v.W || y >= v.H) return;
    const float px = (float)x, py = (float)y;
    float T = 1.0f;
    for (int i = 0; i < n; ++i) {
        const Entry e = es[i];
        const float dx = e.u - px, dy = e.v - py;
        const float q = e.c0 * dx * dx + 2.0f * e.c1 * dx * dy + e.c2 * dy * dy;
        const float a = fminf(0.99f, e.o * expf(-0.5f * q));
        if (a < kMinA) continue;
        vis[e.g] = 1;
        T *= (1.0f - a);
        if (T < 0.5f) break;
    }
}

// computeGaussiansCovisibility_kernel (GSCudaKernels.cu:1539-1560): |a AND b| and |a OR b|,
// reduced per wave (one atomic per wave and counter).
__global__ __launch_bounds__(256) void k_covisibility(const uint8_t* __restrict__ a,
                                                     const uint8_t* __restrict__ b, int n,
                                                     uint32_t* __restrict__ counts) {
    const int g = blockIdx.x * 256 + threadIdx.x;
    const bool va = g < n && a[g], vb = g < n && b[g];
    const uint32_t inter = (uint32_t)__popcll(__ballot(va && vb));
    const uint32_t uni = (uint32_t)__popcll(__ballot(va || vb));
    if ((threadIdx.x & 63) == 0) {
        if (inter) atomicAdd(&counts[0], inter);
        if (uni) atomicAdd(&counts[1], uni);
    }
}

// computeOutliers_kernel (GSCudaKernels.cu:1620-1715), every entry of the tile against every
// pixel: dx = x - u, rendered depth z + p.(dx, dy), alpha = o exp(-(c0 dx dx + c2 dy dy +
// 2 c1 dx dy)/2); pairs with alpha >= 1/255 and positive rendered depth add alpha to the
// Gaussian's total, and to its outlier mass where the observed depth exceeds 0.1 and the
// rendered depth is below 0.8 of it. Transposed: one thread per entry sums its 256 pixels (the
// tile's observed depths in LDS), one atomic per (Gaussian, tile) instead of per pixel.
__global__ __launch_bounds__(256) void k_outliers(View v, const uint32_t* __restrict__ ranges,
                                                  const uint32_t* __restrict__ idx,
                                                  const Rec* __restrict__ rec, MapRef map,
                                                  const float* __restrict__ obs_depth,
                                                  float* __restrict__ prob,
                                                  float* __restrict__ total) {
    __shared__ float sd[256];
    __shared__ uint8_t sin[256];
    const int tile = blockIdx.x, tid = threadIdx.x;
    const int tx0 = (tile % v.ntx) * kTile, ty0 = (tile / v.ntx) * kTile;
    {
        const int x = tx0 + (tid & 15), y = ty0 + (tid >> 4);
        const bool inside = x < v.W && y < v.H;
        sin[tid] = inside;
        sd[tid] = inside ? obs_depth[(size_t)y * v.W + x] : 0.f;
    }
    __syncthreads();
    const uint32_t s = ranges[2 * tile], n = ranges[2 * tile + 1] - s;
    for (uint32_t k = tid; k < n; k += 256) {
        const uint32_t g = idx[s + k];
        const Rec r = rec[g];
        float c[3], o;
        map.appearance(g, c, o);
        const float u = r.r0.x, vv = r.r0.y, c0 = r.r0.z, c1 = r.r0.w, c2 = r.r1.x;
        const float z = r.r1.y, phx = r.r1.z, phy = r.r1.w;
        float ta = 0.f, op = 0.f;
        for (int p = 0; p < 256; ++p) {
            if (!sin[p]) continue;
            const float dx = (float)(tx0 + (p & 15)) - u, dy = (float)(ty0 + (p >> 4)) - vv;
            const float dr = z + phx * dx + phy * dy;
            const float gv = expf(-0.5f * (c0 * dx * dx + c2 * dy * dy + 2.0f * c1 * dx * dy));
            const float a = o * gv;
            if (a < kMinA || dr <= 0.0f) continue;
            ta += a;
            const float d = sd[p];
            if (d > 0.1f && dr < 0.8f * d) op += a;
        }
        if (ta != 0.f) atomicAdd(&total[g], ta);
        if (op != 0.f) atomicAdd(&prob[g], op);
    }
}

// computeDensityMask_kernel (GSCudaKernels.cu:1741-1857): per pixel with observed depth >= 0.5,
// front to back over the whole list, alpha = min(0.99, o exp(-q/2)) (q as in the visibility,
// dx = u - x), prod *= 1 - alpha; the rendered depth z + p.(dx, dy) is taken where prod first
// falls below 0.5; stop below 1e-4. Mask: prod if prod > 0.5, else 1 where the rendered surface
// lies behind the observed one by more than 0.2 of it, else 0 (0 also without valid depth).
__global__ __launch_bounds__(256) void k_density_mask(View v, const uint32_t* __restrict__ ranges,
                                                      const uint32_t* __restrict__ idx,
                                                      const Rec* __restrict__ rec, MapRef map,
                                                      const float* __restrict__ obs_depth,
                                                      float* __restrict__ mask) {
    __shared__ Entry es[256];
    __shared__ float4 ez[256];   // (depth, phx, phy, -)
    const int tile = blockIdx.x, tid = threadIdx.x;
    const int x = (tile % v.ntx) * kTile + (tid & 15), y = (tile / v.ntx) * kTile + (tid >> 4);
    const bool inside = x < v.W && y < v.H;
    const float dobs = inside ? obs_depth[(size_t)y * v.W + x] : 0.f;
    const uint32_t s = ranges[2 * tile], n = ranges[2 * tile + 1] - s;
    const float px = (float)x, py = (float)y;
    float prod = 1.0f, T = 1.0f, dr = 0.0f;
    bool done = (dobs < 0.5f) || !inside;
    for (uint32_t base = 0; base < n; base += 256) {
        if (__syncthreads_count(!done) == 0) break;
        if (base + tid < n) {
            const uint32_t g = idx[s + base + tid];
            es[tid] = stage(rec, map, g);
            const float4 r1 = rec[g].r1;
            ez[tid] = make_float4(r1.y, r1.z, r1.w, 0.f);
        }
        __syncthreads();
        const int cnt = (int)min(256u, n - base);
        for (int i = 0; !done && i < cnt; ++i) {
            const Entry e = es[i];
            const float dx = e.u - px, dy = e.v - py;
            const float q = e.c0 * dx * dx + 2.0f * e.c1 * dx * dy + e.c2 * dy * dy;
            const float a = fminf(0.99f, e.o * expf(-0.5f * q));
            if (a < kMinA) continue;
            prod *= (1.0f - a);
            if (T > 0.5f && prod < 0.5f) {
                T = prod;
                const float4 zz = ez[i];
                dr = zz.x + zz.y * dx + zz.z * dy;
            }
            if (prod < 0.0001f) done = true;
        }
    }
    if (!inside) return;
    float val = 0.0f;
    if (dobs > 0.5f) {
        if (prod > 0.5f) {
            val = prod;
        } else {
            const float err = dr - dobs;
            if (dobs < dr && err > 0.2f * dobs) val = 1.0f;
        }
    }
    mask[(size_t)y * v.W + x] = val;
}

// pruneGaussians_kernel (GSCudaKernels.cu:1566-1618): removed when opacity < alpha_thr, or the
// middle scale over the largest is below ratio_thr, or the largest scale is below 0.005.
__global__ __launch_bounds__(256) void k_prune_keep(const float* __restrict__ scale,
                                                    const float* __restrict__ opac, int n,
                                                    float alpha_thr, float ratio_thr,
                                                    uint8_t* __restrict__ keep) {
    const int g = blockIdx.x * 256 + threadIdx.x;
    if (g >= n) return;
    bool rm;
    if (opac[g] < alpha_thr) {
        rm = true;
    } else {
        const float sx = scale[3 * g], sy = scale[3 * g + 1], sz = scale[3 * g + 2];
        const float smax = fmaxf(sx, fmaxf(sy, sz)), smin = fminf(sx, fminf(sy, sz));
        const float smed = sx + sy + sz - smax - smin;
        rm = (smed / smax < ratio_thr) || (smax < 0.005f);
    }
    keep[g] = rm ? 0 : 1;
}

// removeOutliers_kernel (GSCudaKernels.cu:1717-1739): removed when total > 1 and
// prob / total > threshold.
__global__ __launch_bounds__(256) void k_outlier_keep(const float* __restrict__ prob,
                                                      const float* __restrict__ total, int n,
                                                      float thr, uint8_t* __restrict__ keep) {
    const int g = blockIdx.x * 256 + threadIdx.x;
    if (g >= n) return;
    keep[g] = (total[g] > 1.0f && prob[g] / total[g] > thr) ? 0 : 1;
}

// ---- deterministic compaction: counts per 256-item block, one scan, ordered scatter --------
__device__ __forceinline__ uint32_t block_rank(bool flag, uint32_t* wcount, uint32_t* total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint64_t m = __ballot(flag);
    const uint32_t below = (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) wcount[w] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t off = 0;
    for (int k = 0; k < w; ++k) off += wcount[k];
    if (total) *total = wcount[0] + wcount[1] + wcount[2] + wcount[3];
    return off + below;
}

__global__ __launch_bounds__(256) void k_count_flags(const uint8_t* __restrict__ flag, int n,
                                                     uint32_t* __restrict__ bsum) {
    __shared__ uint32_t wc[4];
    const int g = blockIdx.x * 256 + threadIdx.x;
    uint32_t tot;
    block_rank(g < n && flag[g], wc, &tot);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

// Exclusive scan of nb block counts in place; bsum[nb] = total. One workgroup, 1024 per round.
__global__ __launch_bounds__(1024) void k_scan_counts(uint32_t* __restrict__ bsum, int nb) {
    __shared__ uint32_t buf[1024];
    __shared__ uint32_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int base = 0; base < nb; base += 1024) {
        const int i = base + (int)threadIdx.x;
        const uint32_t x = i < nb ? bsum[i] : 0u;
        buf[threadIdx.x] = x;
        __syncthreads();
        for (int d = 1; d < 1024; d <<= 1) {   // Hillis-Steele inclusive scan
            const uint32_t t = threadIdx.x >= (unsigned)d ? buf[threadIdx.x - d] : 0u;
            __syncthreads();
            buf[threadIdx.x] += t;
            __syncthreads();
        }
        if (i < nb) bsum[i] = carry + buf[threadIdx.x] - x;
        __syncthreads();
        if (threadIdx.x == 0) carry += buf[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) bsum[nb] = carry;
}

// Survivors (keep = 1) to dst[rank] in their order: map arrays and, when given, Adam states.
__global__ __launch_bounds__(256) void k_compact_map(const uint8_t* __restrict__ keep, int n,
                                                     const uint32_t* __restrict__ boff,
                                                     MapArrays src, MapArrays dst,
                                                     const AdamState* __restrict__ asrc,
                                                     AdamState* __restrict__ adst) {
    __shared__ uint32_t wc[4];
    const int g = blockIdx.x * 256 + threadIdx.x;
    const bool k = g < n && keep[g];
    const uint32_t r = boff[blockIdx.x] + block_rank(k, wc, nullptr);
    if (!k) return;
    for (int c = 0; c < 3; ++c) {
        dst.pos[3 * (size_t)r + c] = src.pos[3 * (size_t)g + c];
        dst.scale[3 * (size_t)r + c] = src.scale[3 * (size_t)g + c];
        dst.color[3 * (size_t)r + c] = src.color[3 * (size_t)g + c];
    }
    reinterpret_cast<float4*>(dst.quat)[r] = reinterpret_cast<const float4*>(src.quat)[g];
    dst.opac[r] = src.opac[g];
    if (adst) adst[r] = asrc[g];
}

// ---- quaternion helpers of the initialisation kernels: float4 slots read (w, x, y, z) -------
__device__ __forceinline__ float3 cross3(float3 a, float3 b) {
    return make_float3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
// rotateByQuaternion (GSCudaKernels.cu:20-26): v + w t + qv x t, t = 2 (qv x v)
__device__ __forceinline__ float3 rotate_wxyz(float4 q, float3 v) {
    const float3 qv = make_float3(q.y, q.z, q.w);
    const float3 c = cross3(qv, v);
    const float3 t = make_float3(2.0f * c.x, 2.0f * c.y, 2.0f * c.z);
    const float3 c2 = cross3(qv, t);
    return make_float3(v.x + q.x * t.x + c2.x, v.y + q.x * t.y + c2.y, v.z + q.x * t.z + c2.z);
}
__device__ __forceinline__ float3 normalize_or_zero(float3 v) {   // GSCudaKernels.cu:28-37
    const float n = sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
    if (n > 1e-8f) {
        const float inv = 1.0f / n;
        return make_float3(v.x * inv, v.y * inv, v.z * inv);
    }
    return make_float3(0.f, 0.f, 0.f);
}
__device__ __forceinline__ float4 qmul_wxyz(float4 a, float4 b) {   // GSCudaKernels.cu:39-50
    const float aw = a.x, ax = a.y, ay = a.z, az = a.w;
    const float bw = b.x, bx = b.y, by = b.z, bz = b.w;
    return make_float4(aw * bw - ax * bx - ay * by - az * bz, aw * bx + ax * bw + ay * bz - az * by,
                       aw * by - ax * bz + ay * bw + az * bx, aw * bz + ax * by - ay * bx + az * bw);
}
__device__ __forceinline__ float4 q_from_two(float3 from, float3 to) {   // GSCudaKernels.cu:52-83
    const float3 u = normalize_or_zero(from), v = normalize_or_zero(to);
    const float dot = u.x * v.x + u.y * v.y + u.z * v.z;
    if (dot < -0.999999f) {
        const float3 axis = (fabsf(u.x) < 0.1f) ? make_float3(1.f, 0.f, 0.f) : make_float3(0.f, 1.f, 0.f);
        const float3 o = normalize_or_zero(cross3(u, axis));
        return make_float4(0.f, o.x, o.y, o.z);
    }
    const float3 c = cross3(u, v);
    float4 q = make_float4(1.0f + dot, c.x, c.y, c.z);
    const float n = sqrtf(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
    if (n > 1e-8f) {
        const float inv = 1.0f / n;
        q = make_float4(q.x * inv, q.y * inv, q.z * inv, q.w * inv);
    } else {
        q = make_float4(1.f, 0.f, 0.f, 0.f);
    }
    return q;
}

// computeNormalsFromDepth_kernel (GSCudaKernels.cu:87-162)
__global__ __launch_bounds__(256) void k_normals(const float* __restrict__ depth, int W, int H,
                                                 float* __restrict__ nrm) {
    const int x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= W || y >= H) return;
    float3 n = make_float3(0.f, 0.f, 1.f);
    if (x > 0 && x < W - 1 && y > 0 && y < H - 1) {
        const float zc = depth[(size_t)y * W + x], zl = depth[(size_t)y * W + x - 1],
                    zr = depth[(size_t)y * W + x + 1], zu = depth[(size_t)(y - 1) * W + x],
                    zd = depth[(size_t)(y + 1) * W + x];
        if (zc > 0.01f && zl > 0.01f && zr > 0.01f && zu > 0.01f && zd > 0.01f) {
            const float dzdx = (zr - zl) / 2.0f, dzdy = (zd - zu) / 2.0f;
            n = cross3(make_float3(1.f, 0.f, dzdx), make_float3(0.f, 1.f, dzdy));
            const float norm = sqrtf(n.x * n.x + n.y * n.y + n.z * n.z);
            if (norm > 1e-6f) { n.x /= norm; n.y /= norm; n.z /= norm; }
            if (n.z < 0.0f) { n.x = -n.x; n.y = -n.y; n.z = -n.z; }
        }
    }
    nrm[3 * ((size_t)y * W + x)] = n.x;
    nrm[3 * ((size_t)y * W + x) + 1] = n.y;
    nrm[3 * ((size_t)y * W + x) + 2] = n.z;
}

// A candidate Gaussian per sampling cell: pos3 scale3 quat4 (float4 slots as the reference
// writes them) colour3 opacity; keep[cell] = 1 when it is emitted.
constexpr int kCand = 14;

__device__ __forceinline__ void put_cand(float* c, float3 p, float3 s, float4 q, float3 col,
                                         float o) {
    c[0] = p.x; c[1] = p.y; c[2] = p.z; c[3] = s.x; c[4] = s.y; c[5] = s.z;
    c[6] = q.x; c[7] = q.y; c[8] = q.z; c[9] = q.w; c[10] = col.x; c[11] = col.y; c[12] = col.z;
    c[13] = o;
}

// initGaussiansFromRgbd_kernel (GSCudaKernels.cu:164-311): one Gaussian per sampling cell at its
// first pixel with valid depth; orientation aligning z with the world normal; colour from the
// 8-bit (B, G, R) pixel; the camera's float4 read (w, x, y, z) as these kernels do.
__global__ __launch_bounds__(256) void k_init_cells(View v, const uint8_t* __restrict__ bgr,
                                                    const float* __restrict__ depth,
                                                    const float* __restrict__ nrm, int sdx,
                                                    int sdy, float init_scale, float init_opac,
                                                    float* __restrict__ cand,
                                                    uint8_t* __restrict__ keep) {
    const int sw = (v.W + sdx - 1) / sdx, sh = (v.H + sdy - 1) / sdy;
    const int cell = blockIdx.x * 256 + threadIdx.x;
    if (cell >= sw * sh) return;
    keep[cell] = 0;
    const int u = (cell % sw) * sdx, vv = (cell / sw) * sdy;
    if (u >= v.W || vv >= v.H) return;
    const size_t p = (size_t)vv * v.W + u;
    const float z = depth[p];
    if (!(z > 0.0f) || isinf(z) || isnan(z)) return;
    float3 n = make_float3(nrm[3 * p], nrm[3 * p + 1], nrm[3 * p + 2]);
    const float nl = sqrtf(n.x * n.x + n.y * n.y + n.z * n.z);
    if (nl < 1e-6f) n = make_float3(0.f, 0.f, 1.f);
    else { n.x /= nl; n.y /= nl; n.z /= nl; }
    const float cx = ((float)u - v.cx) * z / v.fx, cy = ((float)vv - v.cy) * z / v.fy;
    const float4 cam = make_float4(v.qx, v.qy, v.qz, v.qw);
    const float3 r = rotate_wxyz(cam, make_float3(cx, cy, z));
    const float3 pw = make_float3(v.px + r.x, v.py + r.y, v.pz + r.z);
    float3 nw = rotate_wxyz(cam, n);
    const float nwl = sqrtf(nw.x * nw.x + nw.y * nw.y + nw.z * nw.z);
    if (nwl > 1e-6f) { nw.x /= nwl; nw.y /= nwl; nw.z /= nwl; }
    const float3 f = make_float3(0.f, 0.f, 1.f);
    float3 ax = make_float3(f.y * nw.z - f.z * nw.y, f.z * nw.x - f.x * nw.z, f.x * nw.y - f.y * nw.x);
    const float al = sqrtf(ax.x * ax.x + ax.y * ax.y + ax.z * ax.z);
    const float ac = f.x * nw.x + f.y * nw.y + f.z * nw.z;
    float4 q;
    if (al < 1e-6f) {
        q = ac > 0.0f ? make_float4(1.f, 0.f, 0.f, 0.f) : make_float4(0.f, 1.f, 0.f, 0.f);
    } else {
        ax.x /= al; ax.y /= al; ax.z /= al;
        const float hs = sqrtf(fmaxf(0.0f, (1.0f - ac) / 2.0f));
        const float hc = sqrtf(fmaxf(0.0f, (1.0f + ac) / 2.0f));
        q = make_float4(hc, hs * ax.x, hs * ax.y, hs * ax.z);
    }
    const uint8_t* px = bgr + 3 * p;
    put_cand(cand + (size_t)kCand * cell, pw, make_float3(init_scale, init_scale, init_scale), q,
             make_float3(px[2] / 255.0f, px[1] / 255.0f, px[0] / 255.0f), init_opac);
    keep[cell] = 1;
}

// densifyGaussians_kernel (GSCudaKernels.cu:1859-2068) without normals (the reference passes
// none): cells of sdx x sdy pixels (sdx > 1) emit one Gaussian at the mean of their masked
// pixels with valid depth when at least half the cell qualifies; sdx == 1 emits one per masked
// pixel. Opacity 1; orientation = camera float4 (w, x, y, z) times the quaternion from z to the
// (camera-facing) normal.
__global__ __launch_bounds__(256) void k_densify_cells(View v, const float* __restrict__ mask,
                                                       const uint8_t* __restrict__ bgr,
                                                       const float* __restrict__ depth, int sdx,
                                                       int sdy, float* __restrict__ cand,
                                                       uint8_t* __restrict__ keep) {
    const int sw = (v.W + sdx - 1) / sdx, sh = (v.H + sdy - 1) / sdy;
    const int cell = blockIdx.x * 256 + threadIdx.x;
    if (cell >= sw * sh) return;
    keep[cell] = 0;
    const float4 cam = make_float4(v.qx, v.qy, v.qz, v.qw);
    float3 pos_img, col;
    float sc, depth_z;
    if (sdx > 1) {
        const int u0 = (cell % sw) * sdx, v0 = (cell / sw) * sdy;
        if (u0 >= v.W || v0 >= v.H) return;
        float3 ip = make_float3(0.f, 0.f, 0.f), ra = make_float3(0.f, 0.f, 0.f);
        int cnt = 0;
        for (int u = u0; u < u0 + sdx && u < v.W; u++)
            for (int y = v0; y < v0 + sdy && y < v.H; y++) {
                const size_t p = (size_t)y * v.W + u;
                if (mask[p] > 0.0f) {
                    const float d = depth[p];
                    if (!(d > 0.0f)) continue;
                    const uint8_t* px = bgr + 3 * p;
                    ip.x += (float)u; ip.y += (float)y; ip.z += d;
                    ra.x += px[2] / 255.0f; ra.y += px[1] / 255.0f; ra.z += px[0] / 255.0f;
                    cnt++;
                }
            }
        if (cnt < ((sdx * sdy) >> 1)) return;
        const float inv = 1.0f / (float)cnt;
        ip.x *= inv; ip.y *= inv; ip.z *= inv;
        ra.x *= inv; ra.y *= inv; ra.z *= inv;
        pos_img = ip;
        col = ra;
        depth_z = ip.z;
        const float sx = ip.z * (float)sdx / v.fx, sy = ip.z * (float)sdy / v.fy;
        sc = 0.5f * (sx + sy);
    } else {
        const int x = cell % sw, y = cell / sw;
        if (x >= v.W || y >= v.H) return;
        const size_t p = (size_t)y * v.W + x;
        if (mask[p] <= 0.0f) return;
        const float d = depth[p];
        if (!(d > 0.0f)) return;
        const uint8_t* px = bgr + 3 * p;
        col = make_float3(px[2] / 255.0f, px[1] / 255.0f, px[0] / 255.0f);
        pos_img = make_float3((float)x, (float)y, d);
        depth_z = d;
        const float sx = 0.8f * d * (float)sdx / v.fx, sy = 0.8f * d * (float)sdy / v.fy;
        sc = 0.5f * (sx + sy);
    }
    const float3 pc = make_float3(depth_z * (pos_img.x - v.cx) / v.fx,
                                  depth_z * (pos_img.y - v.cy) / v.fy, depth_z);
    const float3 r = rotate_wxyz(cam, pc);
    const float3 pw = make_float3(v.px + r.x, v.py + r.y, v.pz + r.z);
    const float3 normal = make_float3(0.f, 0.f, 1.f);
    const float4 q = qmul_wxyz(cam, q_from_two(make_float3(0.f, 0.f, 1.f), normal));
    put_cand(cand + (size_t)kCand * cell, pw, make_float3(sc, sc, 0.1f * sc), q, col, 1.0f);
    keep[cell] = 1;
}

// Emitted candidates to map[n0 + rank] in cell order, up to `cap` Gaussians in total; new Adam
// states (when given) are zero.
__global__ __launch_bounds__(256) void k_append_cands(const uint8_t* __restrict__ keep, int ncell,
                                                      const uint32_t* __restrict__ boff,
                                                      const float* __restrict__ cand, uint32_t n0,
                                                      uint32_t cap, MapArrays dst,
                                                      AdamState* __restrict__ adst) {
    __shared__ uint32_t wc[4];
    const int cell = blockIdx.x * 256 + threadIdx.x;
    const bool k = cell < ncell && keep[cell];
    const uint32_t r = boff[blockIdx.x] + block_rank(k, wc, nullptr);
    if (!k || n0 + r >= cap) return;
    const size_t g = n0 + r;
    const float* c = cand + (size_t)kCand * cell;
    for (int i = 0; i < 3; ++i) {
        dst.pos[3 * g + i] = c[i];
        dst.scale[3 * g + i] = c[3 + i];
        dst.color[3 * g + i] = c[10 + i];
    }
    reinterpret_cast<float4*>(dst.quat)[g] = make_float4(c[6], c[7], c[8], c[9]);
    dst.opac[g] = c[13];
    if (adst) adst[g] = AdamState{};
}

}  // namespace

hipError_t launch_visibility(const View& v, const uint32_t* ranges, const uint32_t* idx,
                             const Rec* rec, const MapRef& map, uint8_t* vis, hipStream_t s) {
    hipLaunchKernelGGL(k_visibility, dim3(v.ntx * v.nty), dim3(256), 0, s, v, ranges, idx, rec, map,
                       vis);
    return hipGetLastError();
}

hipError_t launch_covisibility(const uint8_t* a, const uint8_t* b, int n, uint32_t* counts,
                               hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_covisibility, dim3((n + 255) / 256), dim3(256), 0, s, a, b, n, counts);
    return hipGetLastError();
}

hipError_t launch_outliers(const View& v, const uint32_t* ranges, const uint32_t* idx,
                           const Rec* rec, const MapRef& map, const float* obs_depth, float* prob,
                           float* total, hipStream_t s) {
    hipLaunchKernelGGL(k_outliers, dim3(v.ntx * v.nty), dim3(256), 0, s, v, ranges, idx, rec, map,
                       obs_depth, prob, total);
    return hipGetLastError();
}

hipError_t launch_density_mask(const View& v, const uint32_t* ranges, const uint32_t* idx,
                               const Rec* rec, const MapRef& map, const float* obs_depth,
                               float* mask, hipStream_t s) {
    hipLaunchKernelGGL(k_density_mask, dim3(v.ntx * v.nty), dim3(256), 0, s, v, ranges, idx, rec,
                       map, obs_depth, mask);
    return hipGetLastError();
}

hipError_t launch_prune_keep(const float* scale, const float* opac, int n, float alpha_thr,
                             float ratio_thr, uint8_t* keep, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_prune_keep, dim3((n + 255) / 256), dim3(256), 0, s, scale, opac, n,
                       alpha_thr, ratio_thr, keep);
    return hipGetLastError();
}

hipError_t launch_outlier_keep(const float* prob, const float* total, int n, float thr,
                               uint8_t* keep, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_outlier_keep, dim3((n + 255) / 256), dim3(256), 0, s, prob, total, n,
                       thr, keep);
    return hipGetLastError();
}

hipError_t launch_flag_offsets(const uint8_t* flag, int n, uint32_t* boff, hipStream_t s) {
    const int nb = (n + 255) / 256;
    if (nb > 0)
        hipLaunchKernelGGL(k_count_flags, dim3(nb), dim3(256), 0, s, flag, n, boff);
    hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(1024), 0, s, boff, nb);
    return hipGetLastError();
}

hipError_t launch_compact_map(const uint8_t* keep, int n, const uint32_t* boff, MapArrays src,
                              MapArrays dst, const AdamState* asrc, AdamState* adst,
                              hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_compact_map, dim3((n + 255) / 256), dim3(256), 0, s, keep, n, boff, src,
                       dst, asrc, adst);
    return hipGetLastError();
}

hipError_t launch_normals(const float* depth, int W, int H, float* nrm, hipStream_t s) {
    hipLaunchKernelGGL(k_normals, dim3((W + 15) / 16, (H + 15) / 16), dim3(256), 0, s, depth, W, H,
                       nrm);
    return hipGetLastError();
}

hipError_t launch_init_cells(const View& v, const uint8_t* bgr, const float* depth,
                             const float* nrm, int sdx, int sdy, float init_scale,
                             float init_opac, float* cand, uint8_t* keep, hipStream_t s) {
    const int ncell = ((v.W + sdx - 1) / sdx) * ((v.H + sdy - 1) / sdy);
    hipLaunchKernelGGL(k_init_cells, dim3((ncell + 255) / 256), dim3(256), 0, s, v, bgr, depth, nrm,
                       sdx, sdy, init_scale, init_opac, cand, keep);
    return hipGetLastError();
}

hipError_t launch_densify_cells(const View& v, const float* mask, const uint8_t* bgr,
                                const float* depth, int sdx, int sdy, float* cand, uint8_t* keep,
                                hipStream_t s) {
    const int ncell = ((v.W + sdx - 1) / sdx) * ((v.H + sdy - 1) / sdy);
    hipLaunchKernelGGL(k_densify_cells, dim3((ncell + 255) / 256), dim3(256), 0, s, v, mask, bgr,
                       depth, sdx, sdy, cand, keep);
    return hipGetLastError();
}

hipError_t launch_append_cands(const uint8_t* keep, int ncell, const uint32_t* boff,
                               const float* cand, uint32_t n0, uint32_t cap, MapArrays dst,
                               AdamState* adst, hipStream_t s) {
    if (ncell <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_append_cands, dim3((ncell + 255) / 256), dim3(256), 0, s, keep, ncell,
                       boff, cand, n0, cap, dst, adst);
    return hipGetLastError();
}

}  // namespace fvgs
