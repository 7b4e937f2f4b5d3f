// Area-sampled (box-filtered) triangle coverage and its adjoint.
//
// coverage(pixel) = sum over triangles of area(triangle ∩ pixel box), with
// the pixel box [w, w+1] x [h, h+1] in the screen coordinates of
// dda.cu:33-36 (x = 0.5 (x_ndc + 1) W, y = 0.5 (1 - y_ndc) H).  It is
// continuous in the vertices.  Its derivative follows from the transport
// theorem: moving vertex a of edge a->b (parameter t along the edge, the
// part inside the box t in [t0, t1]) changes the area by
//   n_out * |a b| * integral_{t0}^{t1} (1 - t) dt
// (and b by the integral of t), n_out the triangle's outward unit normal.
#pragma once

#include "common.hpp"

namespace s2c {

struct ScreenTri {
    float2 p[3];  // oriented so that cross(p1 - p0, p2 - p0) > 0
    int map[3];   // original vertex slot of p[i]
    bool valid;
};

// Clip-space vertices -> screen-space triangle with positive orientation.
// Triangles with a vertex at or behind the eye (w <= 0) or with zero screen
// area are invalid and contribute nothing.
__device__ __host__ inline ScreenTri screen_tri(float4 P0, float4 P1, float4 P2, float fW, float fH) {
    ScreenTri t;
    const float4 P[3] = {P0, P1, P2};
    for (int i = 0; i < 3; ++i) {
        t.p[i] = make_float2(0.5f * (1.f + P[i].x / P[i].w) * fW, 0.5f * (1.f - P[i].y / P[i].w) * fH);
        t.map[i] = i;
    }
    const float cr = (t.p[1].x - t.p[0].x) * (t.p[2].y - t.p[0].y) -
                     (t.p[1].y - t.p[0].y) * (t.p[2].x - t.p[0].x);
    t.valid = cr != 0.f && fabsf(cr) <= 3.4028235e38f && P0.w > 0.f && P1.w > 0.f && P2.w > 0.f;
    if (cr < 0.f) {
        const float2 q = t.p[1];
        t.p[1] = t.p[2];
        t.p[2] = q;
        t.map[1] = 2;
        t.map[2] = 1;
    }
    return t;
}

// Area of the (positively oriented) triangle intersected with the box
// [bx, bx+1] x [by, by+1].  By Green's theorem over the triangle cut to the
// strip by <= y <= by+1,
//   area = sum over edges of  integral clamp(x - bx, 0, 1) dy
// taken over the part of each edge inside the strip (the cut lines are
// horizontal, dy = 0 there, and clamp(.) is continuous in x, so no edge
// that lies on a box side needs a tie rule).  clamp(x(t) - bx, 0, 1) is
// linear between the points where x crosses bx and bx+1, so splitting the
// edge there and applying the trapezoid rule is exact.
__device__ __host__ inline float edge_strip_area(float2 a, float2 d) {
    // a, d in box-local coordinates (box = [0,1]^2)
    if (d.y == 0.f) return 0.f;
    const float ty0 = (0.f - a.y) / d.y, ty1 = (1.f - a.y) / d.y;
    const float lo = fmaxf(0.f, fminf(ty0, ty1)), hi = fminf(1.f, fmaxf(ty0, ty1));
    if (!(hi > lo)) return 0.f;
    float s1 = hi, s2 = hi;
    if (d.x != 0.f) {
        const float tx0 = (0.f - a.x) / d.x, tx1 = (1.f - a.x) / d.x;
        s1 = fminf(fmaxf(fminf(tx0, tx1), lo), hi);
        s2 = fminf(fmaxf(fmaxf(tx0, tx1), lo), hi);
    }
    const float f0 = fminf(fmaxf(a.x + lo * d.x, 0.f), 1.f);
    const float f1 = fminf(fmaxf(a.x + s1 * d.x, 0.f), 1.f);
    const float f2 = fminf(fmaxf(a.x + s2 * d.x, 0.f), 1.f);
    const float f3 = fminf(fmaxf(a.x + hi * d.x, 0.f), 1.f);
    const float integral = (f0 + f1) * (s1 - lo) + (f1 + f2) * (s2 - s1) + (f2 + f3) * (hi - s2);
    return 0.5f * integral * d.y;
}

__device__ __host__ inline float clip_area(const ScreenTri& t, float bx, float by) {
    float s = 0.f;
#pragma unroll
    for (int e = 0; e < 3; ++e) {
        const float2 a = make_float2(t.p[e].x - bx, t.p[e].y - by);
        const float2 b = make_float2(t.p[(e + 1) % 3].x - bx, t.p[(e + 1) % 3].y - by);
        s += edge_strip_area(a, make_float2(b.x - a.x, b.y - a.y));
    }
    return s;
}

// One Liang-Barsky side: keeps the part of [t0, t1] where q - t p >= 0.
__device__ __host__ inline void lb_side(float pp, float qq, float& t0, float& t1) {
    if (pp == 0.f) {
        if (qq < 0.f) t1 = -1.f;  // parallel and outside: empty
        return;
    }
    const float r = qq / pp;
    if (pp < 0.f) t0 = fmaxf(t0, r);
    else t1 = fminf(t1, r);
}

// Part [t0, t1] of segment p + t d inside the box (Liang-Barsky).
__device__ __host__ inline bool clip_segment(float2 p, float2 d, float bx, float by, float& t0,
                                             float& t1) {
    float a = 0.f, b = 1.f;
    lb_side(-d.x, p.x - bx, a, b);
    lb_side(d.x, bx + 1.f - p.x, a, b);
    lb_side(-d.y, p.y - by, a, b);
    lb_side(d.y, by + 1.f - p.y, a, b);
    t0 = a;
    t1 = b;
    return b > a;
}

// Adds g * d area(t ∩ box) / d p[i] (screen space) to grad[i].
__device__ __host__ inline void clip_area_grad(const ScreenTri& t, float bx, float by, float g,
                                               float2& g0, float2& g1, float2& g2) {
#pragma unroll
    for (int e = 0; e < 3; ++e) {
        const float2 a = t.p[e], b = t.p[e == 2 ? 0 : e + 1];
        const float2 d = make_float2(b.x - a.x, b.y - a.y);
        float t0, t1;
        if (!clip_segment(a, d, bx, by, t0, t1)) continue;
        const float len = t1 - t0, mid = 0.5f * (t0 + t1);
        const float i1 = len * mid;          // integral of t
        const float i0 = len * (1.f - mid);  // integral of (1 - t)
        // outward normal times |d| for a positively oriented triangle
        const float nx = g * d.y, ny = -g * d.x;
        float2& ga = e == 0 ? g0 : e == 1 ? g1 : g2;
        float2& gb = e == 0 ? g1 : e == 1 ? g2 : g0;
        ga.x += nx * i0;
        ga.y += ny * i0;
        gb.x += nx * i1;
        gb.y += ny * i1;
    }
}

// Screen-space gradient of a vertex -> clip-space (x, y, w) gradient.
__device__ __host__ inline float3 screen_to_clip_grad(float4 P, float2 gs, float fW, float fH) {
    const float gx = gs.x * 0.5f * fW / P.w;
    const float gy = -gs.y * 0.5f * fH / P.w;
    const float gw = -gs.x * 0.5f * fW * P.x / (P.w * P.w) + gs.y * 0.5f * fH * P.y / (P.w * P.w);
    return make_float3(gx, gy, gw);
}

}  // namespace s2c
