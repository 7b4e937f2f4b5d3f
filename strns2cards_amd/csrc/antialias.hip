// Silhouette-edge marking and the antialiased cluster-bitset edge gradient
// (reference: pystrns2cards/src/antialias.cu).
#include "common.hpp"

namespace s2c {

// antialias.cu:8-15.
__device__ __forceinline__ int third_vertex(const int* tri, int t, int v0, int v1) {
    const int i0 = tri[t * 3 + 0], i1 = tri[t * 3 + 1], i2 = tri[t * 3 + 2];
    return (i0 != v0 && i0 != v1) ? i0 : (i1 != v0 && i1 != v1) ? i1
         : (i2 != v0 && i2 != v1) ? i2 : -1;
}

// antialias.cu:18-35.
__device__ __forceinline__ void project(const float* p, int v, float fW, float fH, float& x,
                                        float& y) {
    const float4 q = reinterpret_cast<const float4*>(p)[v];
    const float nx = q.x / q.w;
    const float ny = q.y / q.w;
    x = fW * 0.5f * (1.f + nx);
    y = fH * 0.5f * (1.f - ny);
}

// antialias.cu:38-94.
__global__ void __launch_bounds__(kBlock)
mark_edges_kernel(int H, int W, long long BE, int V, const float* __restrict__ pos,
                  const int* __restrict__ tri, int E, const int* __restrict__ edges,
                  const int* __restrict__ edge2tri, int* __restrict__ prim_ids,
                  float* __restrict__ normals) {
    const long long idx = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
    if (idx >= BE) return;
    const int b = static_cast<int>(idx / E);
    const int e = static_cast<int>(idx - static_cast<long long>(b) * E);
    const float* p = pos + static_cast<long long>(b) * V * 4;
    const float fH = static_cast<float>(H), fW = static_cast<float>(W);
    const int v0 = edges[e * 2 + 0], v1 = edges[e * 2 + 1];
    float x0, y0, x1, y1;
    project(p, v0, fW, fH, x0, y0);
    project(p, v1, fW, fH, x1, y1);
    const float ex = x1 - x0, ey = y1 - y0;
    const float len = sqrtf(ex * ex + ey * ey);
    if (len == 0.f) return;
    const float nx = -ey / len;
    const float ny = ex / len;
    float d0 = 0.f, d1 = 0.f;
    const int t0 = edge2tri[e * 2 + 0], t1 = edge2tri[e * 2 + 1];
    if (t0 >= 0) {
        float x2, y2;
        project(p, third_vertex(tri, t0, v0, v1), fW, fH, x2, y2);
        d0 = nx * (x2 - x0) + ny * (y2 - y0);
    }
    if (t1 >= 0) {
        float x2, y2;
        project(p, third_vertex(tri, t1, v0, v1), fW, fH, x2, y2);
        d1 = nx * (x2 - x0) + ny * (y2 - y0);
    }
    if (d0 * d1 >= 0.f) prim_ids[idx] = static_cast<int>(idx);
    const float sign = 2.f * (.5f - static_cast<float>(d0 > 0.f || d1 > 0.f));
    reinterpret_cast<float2*>(normals)[idx] = make_float2(sign * nx, sign * ny);
}

__device__ __forceinline__ unsigned bit_u(const int* bs, long long pix, int S, int k) {
    return (static_cast<unsigned>(bs[pix * S + (k >> 5)]) >> (k & 31)) & 1u;
}

// Residual bits of a window pixel: bit 0 = rendered (c0), bit 1 = target (c1).
struct ResidualSrc {
    const unsigned char* box;  // staged bytes, or nullptr for global reads
    int h0, w0, width;
    const int* bitset;
    const int* target;
    long long plane;
    int W, S, k;
    __device__ __forceinline__ unsigned at(int h, int w) const {
        if (box) return box[(h - h0) * width + (w - w0)];
        const long long pix = plane + w + static_cast<long long>(h) * W;
        return bit_u(bitset, pix, S, k) | (bit_u(target, pix, S, k) << 1);
    }
};

// antialias.cu:147-176 (compute_adjoint_sum with gaussian_weight,
// antialias.cu:121-132).  Window pixels outside the radius (weight 0) or
// with c0 == c1 (residual 0) add an exact zero in the reference and are
// skipped here, so expf runs only where card and target differ; the
// remaining terms are summed in the reference's order (h, then w).
__device__ __forceinline__ float adjoint_sum(float x, float y, int H, int W, float R, float rho,
                                             const ResidualSrc& src) {
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
            const unsigned code = src.at(h, w);
            const unsigned c0 = code & 1u, c1 = code >> 1;
            if (c0 == c1) continue;
            const float wt = c * expf(-alpha * n2) * 1.f;
            dc += wt * (static_cast<float>(c0) - static_cast<float>(c1)) * (c1 == 0u ? rho : 1.f);
        }
    }
    return dc;
}

// antialias.cu:179-279.  A wave takes 64 consecutive fragments.  When all of
// its live samples share one view and cluster (consecutive silhouette
// fragments of one card) and the union of their windows fits kAaBox bytes,
// the wave stages that box's residual bits in LDS once; otherwise every
// window pixel is read from the bitsets directly.
constexpr int kAaWaves = kBlock / kWave;
constexpr int kAaBox = 8192;

__global__ void __launch_bounds__(kBlock)
antialias_bwd_kernel(int N, const int* __restrict__ frag_pix, const float* __restrict__ fa,
                     int H, int W, int S, const int* __restrict__ bitset,
                     const int* __restrict__ target, int V, const float* __restrict__ pos, int E,
                     const int* __restrict__ edges, const float* __restrict__ normals,
                     const int* __restrict__ edge2cluster, float* __restrict__ d_pos, float R,
                     float rho) {
    __shared__ unsigned char box[kAaWaves][kAaBox];
    const int wv = threadIdx.x / kWave;
    const long long i = (static_cast<long long>(blockIdx.x) * kAaWaves + wv) * kWave + lane_id();
    const float fH = static_cast<float>(H), fW = static_cast<float>(W);

    // ---- per-sample setup, as the reference (antialias.cu:198-256)
    bool live = false;
    int b = 0, h = 0, w = 0, nh = 0, nw = 0, k = 0, v0 = 0, v1 = 0;
    float t0 = 0.f, t1 = 0.f, px = 0.f, py = 0.f, pw = 1.f, fx = 0.f, fy = 0.f, dot = 0.f;
    float2 n = make_float2(0.f, 0.f);
    if (i < N) {
        b = frag_pix[i * 3 + 0];
        h = frag_pix[i * 3 + 1];
        w = frag_pix[i * 3 + 2];
        const float4 a = reinterpret_cast<const float4*>(fa)[i];
        const float axis = a.x;
        t1 = a.y;
        t0 = 1.f - t1;
        const int e = f2i(a.w) - 1;
        if (e >= 0) {
            const float* p = pos + static_cast<long long>(b) * V * 4;
            v0 = edges[e * 2 + 0];
            v1 = edges[e * 2 + 1];
            const float4 P0 = reinterpret_cast<const float4*>(p)[v0];
            const float4 P1 = reinterpret_cast<const float4*>(p)[v1];
            px = t0 * P0.x + t1 * P1.x;
            py = t0 * P0.y + t1 * P1.y;
            pw = t0 * P0.w + t1 * P1.w;
            fx = 0.5f * (1.f + px / pw) * fW;
            fy = 0.5f * (1.f - py / pw) * fH;
            n = reinterpret_cast<const float2*>(normals + static_cast<long long>(b) * E * 2)[e];
            nh = h;
            nw = w;
            const float fxr = fx - floorf(fx);
            const float fyr = fy - floorf(fy);
            constexpr float eps = 1e-4f;
            bool ok;
            if (axis == 0.f) {
                ok = isfinite(fxr) && !(fabsf(fxr - 0.5f) >= eps);
                const int d = 2 * static_cast<int>(fyr > 0.5f) - 1;
                nh += d;
                dot += static_cast<float>(d) * n.y;
            } else {
                ok = isfinite(fyr) && !(fabsf(fyr - 0.5f) >= eps);
                const int d = 2 * static_cast<int>(fxr > 0.5f) - 1;
                nw += d;
                dot += static_cast<float>(d) * n.x;
            }
            // (h, w) outside the image only arises from a malformed fragment
            // list; such samples are dropped instead of read out of bounds.
            live = ok && !(nw < 0 || nw >= W || nh < 0 || nh >= H || dot == 0.f) &&
                   h >= 0 && h < H && w >= 0 && w < W && b >= 0;
            if (live) k = edge2cluster[e];
        }
    }
    const int ch = live ? f2i(floorf(fy)) : 0, cw = live ? f2i(floorf(fx)) : 0;
    const int Ri = f2i(floorf(R));
    const int ph = (dot > 0.f) ? nh : h;  // pixel whose coverage gates the sample (c01)
    const int pwx = (dot > 0.f) ? nw : w;

    // ---- wave-uniform staging decision
    const unsigned long long lm = __ballot(live);
    int bh0 = 0, bw0 = 0, bwid = 0, area = 0;
    bool staged = false;
    if (lm) {
        const int first = __ffsll(static_cast<long long>(lm)) - 1;
        const int b_ref = __shfl(b, first, kWave), k_ref = __shfl(k, first, kWave);
        const bool same = !live || (b == b_ref && k == k_ref);
        if (__ballot(!same) == 0) {
            const int lo_h = live ? dmin(dmax(ch - Ri, 0), ph) : INT_MAX;
            const int hi_h = live ? dmax(dmin(ch + Ri, H - 1), ph) : INT_MIN;
            const int lo_w = live ? dmin(dmax(cw - Ri, 0), pwx) : INT_MAX;
            const int hi_w = live ? dmax(dmin(cw + Ri, W - 1), pwx) : INT_MIN;
            bh0 = wave_min(lo_h);
            bw0 = wave_min(lo_w);
            const long long hh = static_cast<long long>(wave_max(hi_h)) - bh0 + 1;
            const long long ww = static_cast<long long>(wave_max(hi_w)) - bw0 + 1;
            if (hh > 0 && ww > 0 && hh * ww <= kAaBox) {
                staged = true;
                bwid = static_cast<int>(ww);
                area = static_cast<int>(hh * ww);
                const long long plane = static_cast<long long>(H) * W * b_ref;
                for (int q = lane_id(); q < area; q += kWave) {
                    const int y = bh0 + q / bwid, x = bw0 + q % bwid;  // in [0,H) x [0,W)
                    const long long pix = plane + x + static_cast<long long>(y) * W;
                    box[wv][q] = static_cast<unsigned char>(bit_u(bitset, pix, S, k_ref) |
                                                            (bit_u(target, pix, S, k_ref) << 1));
                }
            }
        }
    }
    __syncthreads();
    if (!live) return;

    ResidualSrc src{staged ? box[wv] : nullptr, bh0, bw0, bwid, bitset, target,
                    static_cast<long long>(H) * W * b, W, S, k};
    const float c01 = 1.f - static_cast<float>(src.at(ph, pwx) & 1u);
    if (c01 == 0.f) return;  // dc = 0 * sum: nothing to add (antialias.cu:261-262)
    const float dc = c01 * adjoint_sum(fx, fy, H, W, R, rho, src);
    if (dc == 0.f) return;
    const float dx = dc * 0.5f * fW * n.x;
    const float dy = dc * -0.5f * fH * n.y;
    const float dpx = dx / pw;
    const float dpy = dy / pw;
    const float dpw = (dx * px + dy * py) / (pw * pw);
    if (isfinite(dpx) && isfinite(dpy) && isfinite(dpw)) {
        float* dp = d_pos + static_cast<long long>(b) * V * 4;
        atomicAdd(&dp[v0 * 4 + 0], dpx * t0);
        atomicAdd(&dp[v0 * 4 + 1], dpy * t0);
        atomicAdd(&dp[v0 * 4 + 3], dpw * t0);
        atomicAdd(&dp[v1 * 4 + 0], dpx * t1);
        atomicAdd(&dp[v1 * 4 + 1], dpy * t1);
        atomicAdd(&dp[v1 * 4 + 3], dpw * t1);
    }
}

}  // namespace s2c

using namespace s2c;

extern "C" int s2c_mark_discontinuity_edges(int H, int W, int B, int V, const float* pos,
                                            const int* tri, int E, const int* edges,
                                            const int* edge2tri, int* prim_ids, float* normals,
                                            void* stream) {
    S2C_CHECK(H >= 0 && W >= 0 && B >= 0 && V >= 0 && E >= 0,
              "mark_discontinuity_edges: negative size");
    const long long BE = static_cast<long long>(B) * E;
    S2C_CHECK(BE <= INT_MAX, "mark_discontinuity_edges: B*E exceeds INT32_MAX");
    if (BE == 0) return S2C_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    mark_edges_kernel<<<grid_for(BE, kBlock), kBlock, 0, s>>>(H, W, BE, V, pos, tri, E, edges,
                                                              edge2tri, prim_ids, normals);
    S2C_LAUNCH_CHECK();
    return S2C_OK;
}

extern "C" int s2c_backward_antialiased_cluster_bitset(
    int num_frags, const int* frag_pix, const float* frag_attrs_dda, int H, int W, int num_slots,
    const int* bitset, const int* target_bitset, int V, const float* pos, int E,
    const int* edges, const float* normals, const int* edge2cluster, float* d_pos,
    float kernel_radius, float rho, void* stream) {
    S2C_CHECK(num_frags >= 0 && H >= 0 && W >= 0 && num_slots >= 0 && V >= 0 && E >= 0,
              "backward_antialiased_cluster_bitset: negative size");
    if (num_frags == 0) return S2C_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    antialias_bwd_kernel<<<grid_for(num_frags, kBlock), kBlock, 0, s>>>(
        num_frags, frag_pix, frag_attrs_dda, H, W, num_slots, bitset, target_bitset, V, pos, E,
        edges, normals, edge2cluster, d_pos, kernel_radius, rho);
    S2C_LAUNCH_CHECK();
    return S2C_OK;
}
