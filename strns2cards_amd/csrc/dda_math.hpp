// Per-primitive and per-fragment DDA arithmetic, shared by the drop-in
// kernels (dda.hip), the fused DDA->mask kernel and the antialias pipeline.
// Every expression follows the reference's order of operations so that,
// with contraction off, pixels / ids are bit-exact with the CPU oracle.
#pragma once

#include "common.hpp"

namespace s2c {

struct DdaSpan {
    float sx, sy, off0, off1;  // frag_spans row (dda.cu:62-65, 82-85)
    float axis, slope;         // frag_slopes row (dda.cu:48-49, 68-69)
};

// dda.cu:9-87 for one primitive.  Returns the candidate count; `sp` is
// written only for non-degenerate edges (the reference writes nothing for a
// zero-length edge).  `p0`/`p1` receive the two clip-space vertices.
__device__ __forceinline__ int dda_span(int H, int W, float4 p0, float4 p1, DdaSpan& sp,
                                        bool& wrote) {
    // x / 1 == x exactly in IEEE arithmetic, so orthographic vertices
    // (w == 1) skip the four correctly rounded divisions bit-exactly
    float x0n, y0n, x1n, y1n;
    if (p0.w == 1.f && p1.w == 1.f) {
        x0n = p0.x;
        y0n = p0.y;
        x1n = p1.x;
        y1n = p1.y;
    } else {
        x0n = p0.x / p0.w;
        y0n = p0.y / p0.w;
        x1n = p1.x / p1.w;
        y1n = p1.y / p1.w;
    }
    const float fH = static_cast<float>(H);
    const float fW = static_cast<float>(W);
    float x0 = 0.5f * (x0n + 1.f) * fW;
    float y0 = 0.5f * (1.f - y0n) * fH;
    float x1 = 0.5f * (x1n + 1.f) * fW;
    float y1 = 0.5f * (1.f - y1n) * fH;
    const float dx = x1 - x0;
    const float dy = y1 - y0;
    wrote = false;
    if (dx == 0.f && dy == 0.f) return 0;
    wrote = true;
    // dda.cu's two branches (x-major when |dx| >= |dy|, else y-major) as
    // one sequence on the selected major / minor coordinates: the same
    // operations on the same values, without the divergent branch.
    const bool ym = !(fabsf(dx) >= fabsf(dy));  // y-major (also for NaN, as in dda.cu)
    float a0 = ym ? y0 : x0, b0 = ym ? x0 : y0;   // major, minor of vertex 0
    float a1 = ym ? y1 : x1, b1 = ym ? x1 : y1;
    const float s = (ym ? dx : dy) / (ym ? dy : dx);
    sp.axis = ym ? 1.f : 0.f;
    sp.slope = s;
    const bool flip = a0 > a1;  // start from the lower major coordinate
    const float ta = a0;
    a0 = flip ? a1 : a0;
    b0 = flip ? b1 : b0;
    a1 = flip ? ta : a1;
    const float sa = dmax(floorf(a0), 0.f);
    const float sb = b0 + s * (sa - a0);
    const int i0 = f2i(sa);
    const int i1 = dmin(f2i(floorf(a1)), (ym ? H : W) - 1) + 1;
    sp.sx = ym ? sb : sa;
    sp.sy = ym ? sa : sb;
    sp.off0 = a0 - sa;
    sp.off1 = a1 - sa;
    return dmax(i1 - i0, 0);
}

// dda.cu:124-144.
__device__ __forceinline__ float dda_t1(int H, int W, float4 p0, float4 p1, float x, float y,
                                        float axis) {
    float t0, t1;
    if (axis == 0.f) {
        const float qx = -1.f + 2.f * (x / static_cast<float>(W));
        const float a = p0.x - qx * p0.w;
        const float c = p1.x - qx * p1.w;
        t0 = fabsf(c);
        t1 = fabsf(a);
    } else {
        const float qy = -1.f + 2.f * (y / static_cast<float>(H));
        const float a = -p0.y - qy * p0.w;
        const float c = -p1.y - qy * p1.w;
        t0 = fabsf(c);
        t1 = fabsf(a);
    }
    const float denom = t0 + t1;
    if (denom < 1e-6f) return 0.5f;
    return t1 / denom;
}

struct DdaFrag {
    int h, w;
    float t1, zw;
};

// dda.cu:185-216 for candidate k of a primitive.  Returns true when the
// fragment is on screen and its depth lies in [-1, 1].
__device__ __forceinline__ bool dda_frag(int H, int W, const DdaSpan& sp, int k, float4 p0,
                                         float4 p1, DdaFrag& f) {
    const float d = fminf(fmaxf(static_cast<float>(k) + 0.5f, sp.off0), sp.off1);
    const float ddx = (sp.axis == 0.f) ? d : d * sp.slope;
    const float ddy = (sp.axis == 0.f) ? d * sp.slope : d;
    const float x = sp.sx + ddx;
    const float y = sp.sy + ddy;
    const int h = f2i(floorf(y));
    const int w = f2i(floorf(x));
    if (h < 0 || h >= H || w < 0 || w >= W) return false;
    const float t1 = dda_t1(H, W, p0, p1, x, y, sp.axis);
    const float t0 = 1.f - t1;
    const float pz = t0 * p0.z + t1 * p1.z;
    const float pw = t0 * p0.w + t1 * p1.w;
    const float zw = pz / pw;
    if (!isfinite(zw) || zw < -1.f || zw > 1.f) return false;
    f.h = h; f.w = w; f.t1 = t1; f.zw = zw;
    return true;
}

__device__ __forceinline__ float4 load_pos(const float* pos, long long row) {
    return reinterpret_cast<const float4*>(pos)[row];
}

// True when every fragment of the edge passes the depth test of
// dda.cu:210-216 whatever its t1, so a mask pass may skip t1 and z.
// For t1 in [0, 1] and t0 = fl(1 - t1), the exact (t0 z0 + t1 z1) /
// (t0 w0 + t1 w1) is a weighted mean of the endpoints' NDC depths (weights
// t0*w0, t1*w1 >= 0).  The fp32 evaluation of dda.cu:211-214 adds at most
// 2u(Z/w_min + 1) + u with u = 2^-24 and Z = max|z_clip| <= w_max, i.e.
// below 35u ~= 2.1e-6 when w_max <= 16 w_min.  So with both NDC depths in
// [-1 + 1e-5, 1 - 1e-5] and w in [1e-30, 1e30] (no under/overflow in the
// products), zw stays finite and inside [-1, 1].
// The NDC depths are compared without dividing: |z| <= lim2 * w with
// lim2 = 1 - 2e-5 implies the rounded quotient |fl(z / w)| <= 1 - 1e-5
// (relative error of the product and quotient below 2^-23), so this test
// only ever answers "passes" where the divided test above would.
// t1 is in [0, 1] only while dda.cu:124-144's |p.x - q p.w| terms are
// finite: a sample that reaches the depth test is on screen, so q is in
// [-1, 1], and |x|, |y| <= 1e30 with w <= 1e30 keeps them finite (an
// infinite coordinate makes t1 = inf / inf = NaN, which fails the test).
__device__ __forceinline__ bool dda_depth_always_passes(float4 p0, float4 p1) {
    const float lo = fminf(p0.w, p1.w), hi = fmaxf(p0.w, p1.w);
    if (!(lo >= 1e-30f && hi <= 1e30f && hi <= 16.f * lo)) return false;
    constexpr float cmax = 1e30f;  // each comparison is false for NaN
    if (!(fabsf(p0.x) <= cmax && fabsf(p0.y) <= cmax && fabsf(p1.x) <= cmax &&
          fabsf(p1.y) <= cmax))
        return false;
    constexpr float lim2 = 1.f - 2e-5f;
    return fabsf(p0.z) <= lim2 * p0.w && fabsf(p1.z) <= lim2 * p1.w;  // false for NaN
}

// dda.cu:185-197: pixel of candidate k, without t1 / depth.  Returns false
// when it falls off screen.
__device__ __forceinline__ bool dda_pixel(int H, int W, const DdaSpan& sp, int k, int& h, int& w) {
    const float d = fminf(fmaxf(static_cast<float>(k) + 0.5f, sp.off0), sp.off1);
    const float ddx = (sp.axis == 0.f) ? d : d * sp.slope;
    const float ddy = (sp.axis == 0.f) ? d * sp.slope : d;
    const float x = sp.sx + ddx;
    const float y = sp.sy + ddy;
    h = f2i(floorf(y));
    w = f2i(floorf(x));
    // 0 <= h < H and 0 <= w < W, as one unsigned compare each
    return static_cast<unsigned>(h) < static_cast<unsigned>(H) &&
           static_cast<unsigned>(w) < static_cast<unsigned>(W);
}

// Survival of candidate k for a cluster mask: the pixel test, then the
// depth test unless it is known to pass for the whole edge.
__device__ __forceinline__ bool dda_mask_frag(int H, int W, const DdaSpan& sp, int k, float4 p0,
                                              float4 p1, bool depth_ok, int& h, int& w) {
    if (depth_ok) return dda_pixel(H, W, sp, k, h, w);
    DdaFrag f;
    if (!dda_frag(H, W, sp, k, p0, p1, f)) return false;
    h = f.h;
    w = f.w;
    return true;
}

}  // namespace s2c
