// Per-edge and per-sample arithmetic of the antialiased cluster-bitset
// gradient (reference: pystrns2cards/src/antialias.cu), shared by the
// drop-in kernels and the served card-fitting kernel.  Every expression is
// evaluated in the reference's order.
#pragma once

#include "common.hpp"

namespace s2c {

// antialias.cu:18-35: homogeneous vertex -> pixel coordinates.
__device__ __forceinline__ void aa_project(float4 q, float fW, float fH, float& x, float& y) {
    const float nx = q.x / q.w;
    const float ny = q.y / q.w;
    x = fW * 0.5f * (1.f + nx);
    y = fH * 0.5f * (1.f - ny);
}

// antialias.cu:56-93 for one edge with its (up to two) opposite vertices.
// Returns false for a zero-length edge (nothing written).  `silhouette` is
// the reference's dot0 * dot1 >= 0 test; `n` the outward unit normal.
__device__ __forceinline__ bool aa_edge(float fW, float fH, float4 P0, float4 P1, bool has2,
                                        float4 Q2, bool has3, float4 Q3, bool& silhouette,
                                        float2& n) {
    float x0, y0, x1, y1;
    aa_project(P0, fW, fH, x0, y0);
    aa_project(P1, fW, fH, x1, y1);
    const float ex = x1 - x0, ey = y1 - y0;
    const float len = sqrtf(ex * ex + ey * ey);
    if (len == 0.f) return false;
    const float nx = -ey / len;
    const float ny = ex / len;
    float d0 = 0.f, d1 = 0.f;
    if (has2) {
        float x2, y2;
        aa_project(Q2, fW, fH, x2, y2);
        d0 = nx * (x2 - x0) + ny * (y2 - y0);
    }
    if (has3) {
        float x2, y2;
        aa_project(Q3, fW, fH, x2, y2);
        d1 = nx * (x2 - x0) + ny * (y2 - y0);
    }
    silhouette = d0 * d1 >= 0.f;
    const float sign = 2.f * (.5f - static_cast<float>(d0 > 0.f || d1 > 0.f));
    n = make_float2(sign * nx, sign * ny);
    return true;
}

// One DDA sample of a silhouette edge (antialias.cu:198-254).
struct AaSample {
    bool live;              // passes the interior-sample and neighbour tests
    int nh, nw;             // neighbour pixel across the edge
    int ph, pw_;            // pixel whose coverage gates the sample (c01)
    float t0, t1, px, py, pw, fx, fy;
};

__device__ __forceinline__ AaSample aa_sample(int H, int W, int h, int w, float axis, float t1,
                                              float4 P0, float4 P1, float2 n) {
    AaSample s;
    const float fH = static_cast<float>(H), fW = static_cast<float>(W);
    s.t1 = t1;
    s.t0 = 1.f - t1;
    s.px = s.t0 * P0.x + s.t1 * P1.x;
    s.py = s.t0 * P0.y + s.t1 * P1.y;
    s.pw = s.t0 * P0.w + s.t1 * P1.w;
    s.fx = 0.5f * (1.f + s.px / s.pw) * fW;
    s.fy = 0.5f * (1.f - s.py / s.pw) * fH;
    s.nh = h;
    s.nw = w;
    const float fxr = s.fx - floorf(s.fx);
    const float fyr = s.fy - floorf(s.fy);
    constexpr float eps = 1e-4f;
    float dot = 0.f;
    bool ok;
    if (axis == 0.f) {
        ok = isfinite(fxr) && !(fabsf(fxr - 0.5f) >= eps);
        const int d = 2 * static_cast<int>(fyr > 0.5f) - 1;
        s.nh += d;
        dot += static_cast<float>(d) * n.y;
    } else {
        ok = isfinite(fyr) && !(fabsf(fyr - 0.5f) >= eps);
        const int d = 2 * static_cast<int>(fxr > 0.5f) - 1;
        s.nw += d;
        dot += static_cast<float>(d) * n.x;
    }
    s.live = ok && !(s.nw < 0 || s.nw >= W || s.nh < 0 || s.nh >= H || dot == 0.f);
    s.ph = (dot > 0.f) ? s.nh : h;
    s.pw_ = (dot > 0.f) ? s.nw : w;
    return s;
}

// antialias.cu:147-176 (compute_adjoint_sum with gaussian_weight,
// antialias.cu:121-132).  `code(h, w)` returns the pixel's residual bits:
// bit 0 = rendered (c0), bit 1 = target (c1).  Window pixels outside the
// radius (weight 0) or with c0 == c1 (residual 0) add an exact zero in the
// reference and are skipped, so expf runs only where card and target
// differ; the remaining terms are summed in the reference's order.
template <class Code>
__device__ __forceinline__ float aa_adjoint_sum(float x, float y, int H, int W, float R, float rho,
                                                const Code& code) {
    constexpr float pi = 3.14159265358979323846f;
    const float sigma = R / 4.f;
    const float c = 1.f / (sigma * sigma * 2.f * pi);
    const float alpha = 0.5f / (sigma * sigma);
    const float R2 = R * R;
    const int Ri = f2i(floorf(R));
    const int cw = f2i(floorf(x)), ch = f2i(floorf(y));
    const int h0 = dmax(ch - Ri, 0), h1 = dmin(ch + Ri, H - 1) + 1;
    const int w0 = dmax(cw - Ri, 0), w1 = dmin(cw + Ri, W - 1) + 1;
    float dc = 0.f;
    for (int h = h0; h < h1; ++h) {
        const float dy = static_cast<float>(h) + 0.5f - y;
        for (int w = w0; w < w1; ++w) {
            const float dx = static_cast<float>(w) + 0.5f - x;
            const float n2 = dx * dx + dy * dy;
            if (!(n2 < R2)) continue;
            const unsigned cc = code(h, w);
            const unsigned c0 = cc & 1u, c1 = cc >> 1;
            if (c0 == c1) continue;
            const float wt = c * expf(-alpha * n2) * 1.f;
            // wt * (c0 - c1) * (c1 == 0 ? rho : 1) with c0 != c1, exactly:
            // multiplying by +-1 is exact
            dc += c1 ? -wt : wt * rho;
        }
    }
    return dc;
}

// Widest window row aa_adjoint_sum_rows takes (2 * floor(R) + 1 pixels).
constexpr int kAaRowMax = 32;

// Bits c .. c + n - 1 (1 <= n <= 32) of a packed bit row, bit i for c + i.
// Reads the next word only when the range reaches into it.
__device__ __forceinline__ unsigned bits_at(const unsigned* row, int c, int n) {
    const int wi = c >> 5, sh = c & 31;
    unsigned long long v = row[wi];
    if (sh + n > 32) v |= static_cast<unsigned long long>(row[wi + 1]) << 32;
    v >>= sh;
    return n == 32 ? static_cast<unsigned>(v) : static_cast<unsigned>(v) & ((1u << n) - 1u);
}

// aa_adjoint_sum for residuals read a window row at a time:
// rows(h, w0, n, diff, tgt) gives, for pixels w0 .. w0 + n - 1 of row h
// (1 <= n <= kAaRowMax), the bits where the rendered and target bits differ
// and the target bits, bit i for pixel w0 + i.  Rows without a difference
// are skipped whole and only the differing pixels are visited, by
// increasing w within increasing h: the terms aa_adjoint_sum adds (all
// others are exact zeros there), in the same order, with the same
// arithmetic, so the same sum.  The caller ensures 2 * floor(R) + 1 <=
// kAaRowMax.  FAST_EXP evaluates the Gaussian with __expf, as the
// reference kernel does (antialias.cu:131), instead of expf (the CPU
// oracle's function); the two differ by about an ulp.
template <bool FAST_EXP, class Rows>
__device__ __forceinline__ float aa_adjoint_sum_rows(float x, float y, int H, int W, float R,
                                                     float rho, const Rows& rows) {
    constexpr float pi = 3.14159265358979323846f;
    const float sigma = R / 4.f;
    const float c = 1.f / (sigma * sigma * 2.f * pi);
    const float alpha = 0.5f / (sigma * sigma);
    const float R2 = R * R;
    const int Ri = f2i(floorf(R));
    const int cw = f2i(floorf(x)), ch = f2i(floorf(y));
    const int h0 = dmax(ch - Ri, 0), h1 = dmin(ch + Ri, H - 1) + 1;
    const int w0 = dmax(cw - Ri, 0), w1 = dmin(cw + Ri, W - 1) + 1;
    const int n = w1 - w0;
    float dc = 0.f;
    if (n <= 0) return dc;
    for (int h = h0; h < h1; ++h) {
        unsigned diff, tgt;
        rows(h, w0, n, diff, tgt);
        if (diff == 0u) continue;
        const float dy = static_cast<float>(h) + 0.5f - y;
        while (diff != 0u) {
            const int i = __builtin_ctz(diff);
            diff &= diff - 1u;
            const float dx = static_cast<float>(w0 + i) + 0.5f - x;
            const float n2 = dx * dx + dy * dy;
            if (!(n2 < R2)) continue;
            const unsigned c1 = (tgt >> i) & 1u;
            const float wt = c * (FAST_EXP ? __expf(-alpha * n2) : expf(-alpha * n2)) * 1.f;
            dc += c1 ? -wt : wt * rho;  // as in aa_adjoint_sum
        }
    }
    return dc;
}

// antialias.cu:264-269 for a gated sum dc.
__device__ __forceinline__ bool aa_clip_grad(const AaSample& s, int H, int W, float2 n, float dc,
                                             float& dpx, float& dpy, float& dpw) {
    if (dc == 0.f) return false;
    const float fH = static_cast<float>(H), fW = static_cast<float>(W);
    const float dx = dc * 0.5f * fW * n.x;
    const float dy = dc * -0.5f * fH * n.y;
    dpx = dx / s.pw;
    dpy = dy / s.pw;
    dpw = (dx * s.px + dy * s.py) / (s.pw * s.pw);
    return isfinite(dpx) && isfinite(dpy) && isfinite(dpw);
}

// antialias.cu:256-270: from the gated sum to the clip-space gradient of
// the sample point.  Returns false when nothing is to be added.
template <class Code>
__device__ __forceinline__ bool aa_sample_grad(const AaSample& s, int H, int W, float R, float rho,
                                               float2 n, const Code& code, float& dpx, float& dpy,
                                               float& dpw) {
    const float c01 = 1.f - static_cast<float>(code(s.ph, s.pw_) & 1u);
    if (c01 == 0.f) return false;  // dc = 0 * sum (antialias.cu:261-262)
    const float dc = c01 * aa_adjoint_sum(s.fx, s.fy, H, W, R, rho, code);
    return aa_clip_grad(s, H, W, n, dc, dpx, dpy, dpw);
}

// aa_sample_grad with the window read a row at a time (aa_adjoint_sum_rows)
// when it is at most kAaRowMax pixels wide; the same result (with FAST_EXP,
// up to the exponential's last bits).
template <bool FAST_EXP, class Code, class Rows>
__device__ __forceinline__ bool aa_sample_grad_rows(const AaSample& s, int H, int W, float R,
                                                    float rho, float2 n, const Code& code,
                                                    const Rows& rows, float& dpx, float& dpy,
                                                    float& dpw) {
    const float c01 = 1.f - static_cast<float>(code(s.ph, s.pw_) & 1u);
    if (c01 == 0.f) return false;  // dc = 0 * sum (antialias.cu:261-262)
    const float sum = 2 * f2i(floorf(R)) + 1 <= kAaRowMax
                          ? aa_adjoint_sum_rows<FAST_EXP>(s.fx, s.fy, H, W, R, rho, rows)
                          : aa_adjoint_sum(s.fx, s.fy, H, W, R, rho, code);
    return aa_clip_grad(s, H, W, n, c01 * sum, dpx, dpy, dpw);
}
}  // namespace s2c
