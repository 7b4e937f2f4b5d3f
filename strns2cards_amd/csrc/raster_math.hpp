// Per-triangle arithmetic of the reference rasterizer (rasterize.cu), shared
// by the drop-in kernels, the fused triangle -> bitset kernel and the
// served card-fitting kernel.
#pragma once

#include "common.hpp"

namespace s2c {

__device__ __forceinline__ float4 ld4(const float* p, long long row) {
    return reinterpret_cast<const float4*>(p)[row];
}

// rasterize.cu:9-59 for one (b, t).  Returns h_len * w_len.
__device__ __forceinline__ int tri_rect(int H, int W, float4 P0, float4 P1, float4 P2, int4& r) {
    const float fH = static_cast<float>(H);
    const float fW = static_cast<float>(W);
    float x0 = fW, y0 = fH, x1 = 0.f, y1 = 0.f;
    const float4 P[3] = {P0, P1, P2};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const float inv_w = 1.f / P[i].w;
        const float x = 0.5f * (1.f + P[i].x * inv_w) * fW;
        const float y = 0.5f * (1.f - P[i].y * inv_w) * fH;
        x0 = isnan(x) ? x0 : fminf(x, x0);
        x1 = isnan(x) ? x1 : fmaxf(x, x1);
        y0 = isnan(y) ? y0 : fminf(y, y0);
        y1 = isnan(y) ? y1 : fmaxf(y, y1);
    }
    const int h0 = dmax(f2i(floorf(y0)), 0);
    const int h1 = dmin(f2i(floorf(y1)), H - 1) + 1;
    const int w0 = dmax(f2i(floorf(x0)), 0);
    const int w1 = dmin(f2i(floorf(x1)), W - 1) + 1;
    const int hl = dmax(h1 - h0, 0);
    const int wl = dmax(w1 - w0, 0);
    r = make_int4(h0, hl, w0, wl);
    return hl * wl;
}

// rasterize.cu:132-163 for pixel (x, y).  Returns true when covered.
__device__ __forceinline__ bool tri_cover(int H, int W, int x, int y, float4 P0, float4 P1,
                                          float4 P2, float& b0, float& b1, float& z) {
    if (x < 0 || x >= W || y < 0 || y >= H) return false;
    const float px = -1.f + 2.f * (static_cast<float>(x) + 0.5f) / static_cast<float>(W);
    const float py = -1.f + 2.f * (static_cast<float>(y) + 0.5f) / static_cast<float>(H);
    const float q0x = P0.x - px * P0.w, q0y = -P0.y - py * P0.w;
    const float q1x = P1.x - px * P1.w, q1y = -P1.y - py * P1.w;
    const float q2x = P2.x - px * P2.w, q2y = -P2.y - py * P2.w;
    const float A0 = q1x * q2y - q1y * q2x;
    const float A1 = q2x * q0y - q2y * q0x;
    const float A2 = q0x * q1y - q0y * q1x;
    const float A = A0 + A1 + A2;
    if (A == 0.f) return false;
    const float invA = 1.f / A;
    b0 = A0 * invA;
    b1 = A1 * invA;
    const float zc = A0 * P0.z + A1 * P1.z + A2 * P2.z;
    const float wc = A0 * P0.w + A1 * P1.w + A2 * P2.w;
    z = zc / wc;
    if (!isfinite(b0) || !isfinite(b1) || !isfinite(z)) return false;
    if (b0 < 0.f || b1 < 0.f || b0 + b1 > 1.f || z < -1.f || z > 1.f) return false;
    return true;
}

}  // namespace s2c
