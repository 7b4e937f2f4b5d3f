// Batched / served card fitting: one workgroup per fitting job.
//
// The per-iteration sequence is 3_fitting.py:381-407 for one card strip of
// a cluster:
//   pos       = from_differential(M, param)           (3_fitting.py:384)
//   pos_ndc_h = [pos, 1] @ MVP^T                      (3_fitting.py:385-386)
//   bitset    = render_bitset_from_triangles(...)     (3_fitting.py:154-163)
//   loss      = antialiased_cluster_bitset_loss(...)  (antialias.py:66-130)
//             + w_pln * planarity + w_lap * ||L pos||^2 (3_fitting.py:400-401)
//   UniformAdam step on param, then ExponentialLR decay (optimizer.py:33-68)
// with M = I + lam (L_h + w L_w) of the strip (3_fitting.py:347-356) and
// param = M pos at the job's first step (3_fitting.py:365); the gradient
// reaches param through M^{-1} (M is symmetric).  Both solves use the
// strip's pivots, factored once per launch (grid_math.hpp).  lam = 0 turns
// the reparameterization off (pos itself is optimized).
// A job's render and target masks (<= 256 x 256, one bit per pixel) live
// in LDS as packed bits, its vertices and Adam moments in registers, so a
// launch runs many iterations with only workgroup barriers and no HBM
// traffic beyond the job's start and end.  Per-edge gradient sums are
// reduced per vertex in a fixed edge order, so a job's trajectory depends
// only on its own data: the same alone or batched with any other jobs.
//
// Card topology is the two-row strip of wisp_to_grid_mesh
// (texture_utils.py:525-599) with N = Q + 1 columns: vertex (row r, col c)
// is r*N + c; quad c is [c, N+c, N+c+1, c+1], split into triangles
// [c, N+c, N+c+1] and [c, N+c+1, c+1].
#include "antialias_math.hpp"
#include "dda_math.hpp"
#include "grid_math.hpp"
#include "raster_math.hpp"

namespace s2c {

constexpr int kFitMaxRes = 256;
constexpr int kFitMaxWords = kFitMaxRes * kFitMaxRes / 32;
constexpr int kFitMaxQuads = 32;
constexpr int kFitMaxVerts = 2 * (kFitMaxQuads + 1);
constexpr int kFitMaxTris = 2 * kFitMaxQuads;
constexpr int kFitMaxEdges = 4 * kFitMaxQuads + 1;
// Silhouette samples evaluated per workgroup pass: a template parameter
// (LDS per workgroup against passes per iteration), chosen at run time from
// kFitChunks; a job's sums do not depend on it.
constexpr int kFitChunkDefault = 512;

struct FitHyper {
    float lr0, b1, b2, gamma, w_pln, w_lap, rho, R, lam, ww;
};

__device__ __forceinline__ int3 strip_tri(int t, int N) {
    const int c = t >> 1;
    return (t & 1) ? make_int3(c, N + c + 1, c + 1) : make_int3(c, N + c, N + c + 1);
}

// Unique edges in a fixed order (vertex pair as (min, max), as
// build_edges_from_triangles stores them, mesh_utils.cpp:19-22) with their
// adjacent triangles (-1 = boundary):
//   [0, N)         width edges  (c, N+c)       tris 2c (c<Q), 2c-1 (c>0)
//   [N, N+Q)       row-0 edges  (c, c+1)       tri 2c+1
//   [N+Q, N+2Q)    row-1 edges  (N+c, N+c+1)   tri 2c
//   [N+2Q, N+3Q)   diagonals    (c, N+c+1)     tris 2c, 2c+1
struct StripEdge {
    int v0, v1, ta, tb;
};

__device__ __forceinline__ StripEdge strip_edge(int e, int Q) {
    const int N = Q + 1;
    if (e < N) return {e, N + e, e < Q ? 2 * e : -1, e > 0 ? 2 * e - 1 : -1};
    e -= N;
    if (e < Q) return {e, e + 1, 2 * e + 1, -1};
    e -= Q;
    if (e < Q) return {N + e, N + e + 1, 2 * e, -1};
    e -= Q;
    return {e, N + e + 1, 2 * e, 2 * e + 1};
}

__device__ __forceinline__ int third_of(int3 t, int a, int b) {
    return (t.x != a && t.x != b) ? t.x : (t.y != a && t.y != b) ? t.y : t.z;
}

__device__ __forceinline__ unsigned get_bit(const unsigned* bits, int p) {
    return (bits[p >> 5] >> (p & 31)) & 1u;
}

struct FitCode {
    const unsigned* render;
    const unsigned* target;
    int res;
    __device__ __forceinline__ unsigned operator()(int h, int w) const {
        const int p = h * res + w;
        return get_bit(render, p) | (get_bit(target, p) << 1);
    }
};

// Window rows of the job's masks for aa_adjoint_sum_rows.
struct FitRows {
    const unsigned* render;
    const unsigned* target;
    int res;
    __device__ __forceinline__ void operator()(int h, int w0, int n, unsigned& diff,
                                               unsigned& tgt) const {
        const int p0 = h * res + w0;
        tgt = bits_at(target, p0, n);
        diff = bits_at(render, p0, n) ^ tgt;
    }
};

// Rasterize the strip into `render` (cleared by the caller): every covered
// pixel of every triangle (rasterize.cu:90-172) sets its bit.
__device__ __forceinline__ void fit_raster(int res, int Q, const float4* clip, unsigned* render,
                                           int4* rect, int* excl, int* sh) {
    const int N = Q + 1, T = 2 * Q;
    int cnt = 0;
    int4 r = make_int4(0, 0, 0, 0);
    if (threadIdx.x < T) {
        const int3 t = strip_tri(threadIdx.x, N);
        cnt = tri_rect(res, res, clip[t.x], clip[t.y], clip[t.z], r);
        rect[threadIdx.x] = r;
    }
    int total;
    excl[threadIdx.x] = block_excl_scan(cnt, sh, &total);
    __syncthreads();
    for (int j = threadIdx.x; j < total; j += kBlock) {
        int lo = 0, hi = T;  // last triangle with excl <= j
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (excl[mid] <= j) lo = mid;
            else hi = mid;
        }
        const int4 rr = rect[lo];
        const int l = j - excl[lo];
        const int y = rr.x + l / rr.w, x = rr.z + l % rr.w;
        const int3 t = strip_tri(lo, N);
        float b0, b1, z;
        if (tri_cover(res, res, x, y, clip[t.x], clip[t.y], clip[t.z], b0, b1, z)) {
            const int p = y * res + x;
            atomicOr(&render[p >> 5], 1u << (p & 31));
        }
    }
}

__device__ __forceinline__ float block_max_nonneg(float v, float* red) {
    float m = v;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) m = fmaxf(m, __shfl_xor(m, d, kWave));
    if (lane_id() == 0) red[threadIdx.x / kWave] = m;
    __syncthreads();
    float r = red[0];
    for (int i = 1; i < kBlock / kWave; ++i) r = fmaxf(r, red[i]);
    __syncthreads();
    return r;
}

template <int kFitChunk>
__global__ void __launch_bounds__(kBlock)
fit_cards_kernel(const int* __restrict__ slots, int iters, const int* __restrict__ res_arr,
                 const int* __restrict__ quad_arr, const int* __restrict__ vert_off,
                 const int* __restrict__ bits_off, const float* __restrict__ mvp_arr,
                 float* __restrict__ verts, float* __restrict__ params,
                 float* __restrict__ adam_m, float* __restrict__ adam_v,
                 int* __restrict__ steps, const int* __restrict__ limit,
                 const unsigned* __restrict__ target_bits, FitHyper hp,
                 float* __restrict__ loss_out) {
    __shared__ unsigned render[kFitMaxWords];
    __shared__ unsigned target[kFitMaxWords];
    __shared__ float4 clip[kFitMaxVerts];
    __shared__ float P[kFitMaxVerts][3];
    __shared__ float U[kFitMaxVerts][3];     // param (largesteps)
    __shared__ float G[kFitMaxVerts][3];     // gradient handed to the solve
    __shared__ float Ysol[3 * kFitMaxVerts]; // modes of a solve
    __shared__ float IM[kFitMaxVerts];       // pivots of the strip's systems
    __shared__ float Qb[4];                  // 2 x 2 row basis
    __shared__ float lp[kFitMaxVerts][3];
    __shared__ float acc[kFitMaxEdges][6];
    __shared__ DdaSpan espan[kFitMaxEdges];
    __shared__ float2 enorm[kFitMaxEdges];
    __shared__ int eexcl[kFitMaxEdges];
    __shared__ float contrib[kFitChunk][6];
    __shared__ int4 rect[kFitMaxTris];
    __shared__ int excl[kBlock];
    __shared__ int sh[kBlock / kWave + 1];
    __shared__ float red[kBlock / kWave];
    __shared__ float M[16];

    const int job = slots[blockIdx.x];
    if (job < 0) return;  // inactive slot (workgroup-uniform)
    // A job with an iteration limit runs only the iterations it has left;
    // once it has none, an advancing launch leaves it (and its loss) as is.
    int n_it = iters;
    if (limit != nullptr && iters > 0) {
        const int left = limit[job] - steps[job];
        if (left <= 0) return;  // workgroup-uniform
        n_it = dmin(iters, left);
    }
    const int res = res_arr[job], Q = quad_arr[job];
    if (res < 1 || res > kFitMaxRes || Q < 1 || Q > kFitMaxQuads) {
        if (threadIdx.x == 0) loss_out[job] = -1.f;
        return;  // workgroup-uniform
    }
    const int N = Q + 1, V = 2 * N, E = 4 * Q + 1;
    const int words = (res * res + 31) / 32;
    const int tid = threadIdx.x;
    if (tid < 16) M[tid] = mvp_arr[static_cast<long long>(job) * 16 + tid];
    const unsigned* tg = target_bits + bits_off[job];
    for (int i = tid; i < words; i += kBlock) target[i] = tg[i];
    const long long vo = vert_off[job];
    float g1x = 0.f, g1y = 0.f, g1z = 0.f, g2 = 0.f;
    if (tid < V) {
        for (int k = 0; k < 3; ++k) P[tid][k] = verts[(vo + tid) * 3 + k];
        g1x = adam_m[(vo + tid) * 3 + 0];
        g1y = adam_m[(vo + tid) * 3 + 1];
        g1z = adam_m[(vo + tid) * 3 + 2];
        g2 = adam_v[vo + tid];
    }
    int step = steps[job];
    const FitCode code{render, target, res};
    const FitRows rows{render, target, res};
    const bool ls = hp.lam > 0.f;  // workgroup-uniform
    if (ls) {
        if (tid < 4) Qb[tid] = grid_dct_q(tid >> 1, tid & 1, 2);
        if (tid < 2) grid_factor(2, N, tid, hp.lam, hp.ww, IM + tid * N);
        if (step > 0 && tid < V)
            for (int k = 0; k < 3; ++k) U[tid][k] = params[(vo + tid) * 3 + k];
    }
    __syncthreads();
    if (ls && step == 0) {  // param = to_differential(M, pos)
        for (int q = tid; q < 3 * V; q += kBlock)
            (&U[0][0])[q] = grid_apply_at(2, N, hp.lam, hp.ww, &P[0][0], q);
        __syncthreads();
    }
    // in-place M^{-1} over a (V, 3) LDS array (three workgroup barriers; the
    // sweeps over LDS: their register form would cost this kernel occupancy)
    auto solve = [&](float* a) {
        grid_project<kBlock>(2, N, Qb, a, Ysol, tid);
        __syncthreads();
        if (tid < kWave) grid_solve_wave(2, N, hp.lam, IM, Ysol, tid);
        __syncthreads();
        grid_back_project<kBlock>(2, N, Qb, Ysol, [&](int q, float v) { a[q] = v; }, tid);
        __syncthreads();
    };

    for (int it = 0; it <= n_it; ++it) {
        // ---- pos = from_differential(M, param) (3_fitting.py:384)
        if (ls) {
            for (int q = tid; q < 3 * V; q += kBlock) (&P[0][0])[q] = (&U[0][0])[q];
            __syncthreads();
            solve(&P[0][0]);
        }
        // ---- project (3_fitting.py:385-386) and clear
        if (tid < V) {
            float c[4];
#pragma unroll
            for (int j = 0; j < 4; ++j)
                c[j] = M[j * 4 + 0] * P[tid][0] + M[j * 4 + 1] * P[tid][1] +
                       M[j * 4 + 2] * P[tid][2] + M[j * 4 + 3];
            clip[tid] = make_float4(c[0], c[1], c[2], c[3]);
        }
        for (int i = tid; i < words; i += kBlock) render[i] = 0u;
        __syncthreads();
        // ---- raster (render_bitset_from_triangles)
        fit_raster(res, Q, clip, render, rect, excl, sh);
        __syncthreads();
        if (it == n_it) break;  // final render of the updated card

        // ---- silhouette edges and their antialiased gradient.  Each edge's
        // DDA samples are spread over the whole workgroup (kFitChunk at a
        // time); a sample's six contributions go to LDS and the edge's
        // thread then adds them in sample order, so the sums are exactly
        // those of a serial walk of the edge.
        int cnt = 0;
        if (tid < E) {
            const StripEdge se = strip_edge(tid, Q);
            const float4 P0 = clip[se.v0], P1 = clip[se.v1];
            const float4 none = make_float4(0.f, 0.f, 0.f, 1.f);
            const float4 Q2 = se.ta >= 0 ? clip[third_of(strip_tri(se.ta, N), se.v0, se.v1)] : none;
            const float4 Q3 = se.tb >= 0 ? clip[third_of(strip_tri(se.tb, N), se.v0, se.v1)] : none;
            const float fres = static_cast<float>(res);
            bool sil = false;
            float2 n = make_float2(0.f, 0.f);
            if (aa_edge(fres, fres, P0, P1, se.ta >= 0, Q2, se.tb >= 0, Q3, sil, n) && sil) {
                bool wrote;
                cnt = dda_span(res, res, P0, P1, espan[tid], wrote);
            }
            enorm[tid] = n;
        }
        int nsamp;
        const int first = block_excl_scan(cnt, sh, &nsamp);
        if (tid < kFitMaxEdges) eexcl[tid] = first;
        float a[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        __syncthreads();
        for (int base = 0; base < nsamp; base += kFitChunk) {
            const int lim = dmin(nsamp - base, kFitChunk);
            for (int q = tid; q < lim; q += kBlock) {
                const int j = base + q;
                int lo = 0, hi = E;  // last edge with eexcl <= j
                while (hi - lo > 1) {
                    const int mid = (lo + hi) >> 1;
                    if (eexcl[mid] <= j) lo = mid;
                    else hi = mid;
                }
                const StripEdge se = strip_edge(lo, Q);
                const float4 P0 = clip[se.v0], P1 = clip[se.v1];
                const DdaSpan sp = espan[lo];
                const float2 n = enorm[lo];
                float c[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
                DdaFrag fr;
                if (dda_frag(res, res, sp, j - eexcl[lo], P0, P1, fr)) {
                    const AaSample s = aa_sample(res, res, fr.h, fr.w, sp.axis, fr.t1, P0, P1, n);
                    float dpx, dpy, dpw;
                    if (s.live &&
                        aa_sample_grad_rows<true>(s, res, res, hp.R, hp.rho, n, code, rows, dpx,
                                                  dpy, dpw)) {
                        c[0] = dpx * s.t0; c[1] = dpy * s.t0; c[2] = dpw * s.t0;
                        c[3] = dpx * s.t1; c[4] = dpy * s.t1; c[5] = dpw * s.t1;
                    }
                }
#pragma unroll
                for (int k = 0; k < 6; ++k) contrib[q][k] = c[k];
            }
            __syncthreads();
            if (tid < E && cnt > 0) {
                const int s0 = dmax(first - base, 0), s1 = dmin(first + cnt - base, lim);
                for (int q = s0; q < s1; ++q) {
#pragma unroll
                    for (int k = 0; k < 6; ++k) a[k] += contrib[q][k];
                }
            }
            __syncthreads();
        }
        if (tid < E) {
#pragma unroll
            for (int k = 0; k < 6; ++k) acc[tid][k] = a[k];
        }
        // ---- uniform Laplacian over the strip's row and width edges
        if (tid < V) {
            const int r = tid / N, c = tid - r * N;
            float deg = 1.f;
            float s0 = P[tid + (r ? -N : N)][0], s1 = P[tid + (r ? -N : N)][1],
                  s2 = P[tid + (r ? -N : N)][2];
            if (c > 0) { deg += 1.f; s0 += P[tid - 1][0]; s1 += P[tid - 1][1]; s2 += P[tid - 1][2]; }
            if (c < Q) { deg += 1.f; s0 += P[tid + 1][0]; s1 += P[tid + 1][1]; s2 += P[tid + 1][2]; }
            lp[tid][0] = deg * P[tid][0] - s0;
            lp[tid][1] = deg * P[tid][1] - s1;
            lp[tid][2] = deg * P[tid][2] - s2;
        }
        __syncthreads();

        // ---- per-vertex gradient and UniformAdam (optimizer.py:33-68)
        float gx = 0.f, gy = 0.f, gz = 0.f, rms = 0.f;
        if (tid < V) {
            // d loss / d clip (x, y, w): the vertex's incident edges (at most
            // four; strip_edge's numbering) in increasing edge index, each
            // adding its end's sums -- the order of a scan over all edges
            const int r = tid / N, c = tid - r * N;
            float cx = 0.f, cy = 0.f, cw = 0.f;
            // edge e's sums for its end k (0: v0, 3: v1)
            auto add = [&](int e, int k) {
                cx += acc[e][k + 0];
                cy += acc[e][k + 1];
                cw += acc[e][k + 2];
            };
            if (r == 0) {
                add(c, 0);                              // width edge, v0
                if (c > 0) add(N + c - 1, 3);           // row 0 left, v1
                if (c < Q) add(N + c, 0);               // row 0 right, v0
                if (c < Q) add(N + 2 * Q + c, 0);       // diagonal, v0
            } else {
                add(c, 3);                              // width edge, v1
                if (c > 0) add(N + Q + c - 1, 3);       // row 1 left, v1
                if (c < Q) add(N + Q + c, 0);           // row 1 right, v0
                if (c > 0) add(N + 2 * Q + c - 1, 3);   // diagonal, v1
            }
            gx = cx * M[0] + cy * M[4] + cw * M[12];
            gy = cx * M[1] + cy * M[5] + cw * M[13];
            gz = cx * M[2] + cy * M[6] + cw * M[14];
            // d/dp of w_lap * sum |L p|^2 = 2 w_lap * L (L p)
            {
                float deg = 1.f;
                const int o = tid + (r ? -N : N);
                float s0 = lp[o][0], s1 = lp[o][1], s2 = lp[o][2];
                if (c > 0) { deg += 1.f; s0 += lp[tid - 1][0]; s1 += lp[tid - 1][1]; s2 += lp[tid - 1][2]; }
                if (c < Q) { deg += 1.f; s0 += lp[tid + 1][0]; s1 += lp[tid + 1][1]; s2 += lp[tid + 1][2]; }
                const float k2 = 2.f * hp.w_lap;
                gx += k2 * (deg * lp[tid][0] - s0);
                gy += k2 * (deg * lp[tid][1] - s1);
                gz += k2 * (deg * lp[tid][2] - s2);
            }
            // d/dp of w_pln * sum |(q0 + q2) - (q1 + q3)|^2 over quads
            // [c, N+c, N+c+1, c+1]: +2 w diag for q0, q2 and -2 w diag for q1, q3.
            for (int q = c - 1; q <= c; ++q) {
                if (q < 0 || q >= Q) continue;
                const int i0 = q, i1 = N + q, i2 = N + q + 1, i3 = q + 1;
                const float sgn = (tid == i0 || tid == i2) ? 2.f * hp.w_pln : -2.f * hp.w_pln;
                gx += sgn * ((P[i0][0] + P[i2][0]) - (P[i1][0] + P[i3][0]));
                gy += sgn * ((P[i0][1] + P[i2][1]) - (P[i1][1] + P[i3][1]));
                gz += sgn * ((P[i0][2] + P[i2][2]) - (P[i1][2] + P[i3][2]));
            }
        }
        if (ls) {  // d loss / d param = M^{-1} d loss / d pos
            if (tid < V) {
                G[tid][0] = gx;
                G[tid][1] = gy;
                G[tid][2] = gz;
            }
            __syncthreads();
            solve(&G[0][0]);
            if (tid < V) {
                gx = G[tid][0];
                gy = G[tid][1];
                gz = G[tid][2];
            }
        }
        if (tid < V) {
            const float t = static_cast<float>(step + 1);
            g1x = hp.b1 * g1x + (1.f - hp.b1) * gx;
            g1y = hp.b1 * g1y + (1.f - hp.b1) * gy;
            g1z = hp.b1 * g1z + (1.f - hp.b1) * gz;
            g2 = hp.b2 * g2 + (1.f - hp.b2) * (gx * gx + gy * gy + gz * gz);
            const float m2 = g2 / (1.f - powf(hp.b2, t));
            rms = sqrtf(m2);
        }
        const float denom = block_max_nonneg(rms, red) + 1e-8f;
        if (tid < V) {
            const float t = static_cast<float>(step + 1);
            const float c1 = 1.f - powf(hp.b1, t);
            const float lr = hp.lr0 * powf(hp.gamma, static_cast<float>(step));
            float* p = ls ? U[tid] : P[tid];
            p[0] -= lr * ((g1x / c1) / denom);
            p[1] -= lr * ((g1y / c1) / denom);
            p[2] -= lr * ((g1z / c1) / denom);
        }
        ++step;
        __syncthreads();
    }

    // ---- coverage loss of the final render: popcount(render XOR target)
    int mism = 0;
    for (int i = tid; i < words; i += kBlock) mism += __popc(render[i] ^ target[i]);
    int total;
    block_excl_scan(mism, sh, &total);
    if (tid == 0) {
        loss_out[job] = static_cast<float>(total);
        steps[job] = step;
    }
    // iters == 0 only measures the current state: with large steps at step 0
    // the render used M^{-1} (M pos), which differs from pos in the last
    // bits, so nothing but the loss is written back
    if (n_it > 0 && tid < V) {
        for (int k = 0; k < 3; ++k) verts[(vo + tid) * 3 + k] = P[tid][k];
        if (ls)
            for (int k = 0; k < 3; ++k) params[(vo + tid) * 3 + k] = U[tid][k];
        adam_m[(vo + tid) * 3 + 0] = g1x;
        adam_m[(vo + tid) * 3 + 1] = g1y;
        adam_m[(vo + tid) * 3 + 2] = g1z;
        adam_v[vo + tid] = g2;
    }
}

// Admission: one workgroup per admitted job writes its slot's tables from
// one packed record (kFitAdmitInts ints: slot, res, quads, word offset,
// vertex count, iteration limit; 16 + 3 * max_verts floats: the 4 x 4
// clip matrix, then the vertices), zeroes its optimizer state and
// activates the slot.
constexpr int kFitAdmitInts = 8;

__global__ void __launch_bounds__(kBlock)
fit_admit_kernel(const int* __restrict__ rec, const float* __restrict__ frec,
                 int* __restrict__ slots, int* __restrict__ res, int* __restrict__ quads,
                 int* __restrict__ bits_off, int* __restrict__ limit, float* __restrict__ mvp,
                 int* __restrict__ steps, float* __restrict__ verts,
                 float* __restrict__ adam_m, float* __restrict__ adam_v, int capacity,
                 int max_verts) {
    const int* r = rec + static_cast<long long>(blockIdx.x) * kFitAdmitInts;
    const float* f = frec + static_cast<long long>(blockIdx.x) * (16 + 3 * max_verts);
    const int s = r[0], nv = r[4];
    if (s < 0 || s >= capacity || nv < 0 || nv > kFitMaxVerts) return;  // workgroup-uniform
    const long long v0 = static_cast<long long>(s) * max_verts;
    const int tid = threadIdx.x;
    if (tid == 0) {
        res[s] = r[1];
        quads[s] = r[2];
        bits_off[s] = r[3];
        if (limit != nullptr) limit[s] = r[5];
        steps[s] = 0;
        slots[s] = s;
    }
    if (tid < 16) mvp[static_cast<long long>(s) * 16 + tid] = f[tid];
    for (int q = tid; q < 3 * nv; q += kBlock) {
        verts[v0 * 3 + q] = f[16 + q];
        adam_m[v0 * 3 + q] = 0.f;
    }
    for (int v = tid; v < nv; v += kBlock) adam_v[v0 + v] = 0.f;
}

// Target masks: every edge of a job's strands (consecutive vertices of
// each strand) DDA-rasterized (dda.cu:9-225, all edges in cluster 0) into a
// packed res x res bit mask.
__global__ void __launch_bounds__(kBlock)
fit_targets_kernel(const int* __restrict__ res_arr, const int* __restrict__ bits_off,
                   const int* __restrict__ strand_off, const int* __restrict__ num_strands,
                   int nv, const float* __restrict__ strand_clip, unsigned* __restrict__ target_bits) {
    __shared__ unsigned bits[kFitMaxWords];
    const int job = blockIdx.x;
    const int res = res_arr[job];
    if (res < 1 || res > kFitMaxRes || nv < 2) return;  // workgroup-uniform
    const int words = (res * res + 31) / 32;
    for (int i = threadIdx.x; i < words; i += kBlock) bits[i] = 0u;
    __syncthreads();
    const long long v0 = strand_off[job];
    const int ne = num_strands[job] * (nv - 1);
    const float4* p = reinterpret_cast<const float4*>(strand_clip);
    for (int e = threadIdx.x; e < ne; e += kBlock) {
        const int s = e / (nv - 1), i = e - s * (nv - 1);
        const long long a = v0 + static_cast<long long>(s) * nv + i;
        const float4 P0 = p[a], P1 = p[a + 1];
        DdaSpan sp;
        bool wrote;
        const int cnt = dda_span(res, res, P0, P1, sp, wrote);
        const bool depth_ok = dda_depth_always_passes(P0, P1);
        for (int k = 0; k < cnt; ++k) {
            int h, w;
            if (dda_mask_frag(res, res, sp, k, P0, P1, depth_ok, h, w)) {
                const int q = h * res + w;
                atomicOr(&bits[q >> 5], 1u << (q & 31));
            }
        }
    }
    __syncthreads();
    unsigned* out = target_bits + bits_off[job];
    for (int i = threadIdx.x; i < words; i += kBlock) out[i] = bits[i];
}

}  // namespace s2c

using namespace s2c;

extern "C" int s2c_fit_targets(int num_jobs, const int* res, const int* bits_off,
                               const int* strand_off, const int* num_strands,
                               int verts_per_strand, const float* strand_clip,
                               unsigned* target_bits, void* stream) {
    S2C_CHECK(num_jobs >= 0 && verts_per_strand >= 2, "fit_targets: bad sizes");
    if (num_jobs == 0) return S2C_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    fit_targets_kernel<<<num_jobs, kBlock, 0, s>>>(res, bits_off, strand_off, num_strands,
                                                   verts_per_strand, strand_clip, target_bits);
    S2C_LAUNCH_CHECK();
    return S2C_OK;
}

extern "C" int s2c_fit_admit(int num_jobs, const int* records, const float* frecords,
                             int capacity, int max_verts, int* slots, int* res, int* quads,
                             int* bits_off, int* limit, float* mvp, int* steps, float* verts,
                             float* adam_m, float* adam_v, void* stream) {
    S2C_CHECK(num_jobs >= 0 && capacity >= 0, "fit_admit: negative size");
    S2C_CHECK(max_verts >= kFitMaxVerts, "fit_admit: max_verts below the strip's vertex count");
    if (num_jobs == 0) return S2C_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    fit_admit_kernel<<<num_jobs, kBlock, 0, s>>>(records, frecords, slots, res, quads, bits_off,
                                                 limit, mvp, steps, verts, adam_m, adam_v,
                                                 capacity, max_verts);
    S2C_LAUNCH_CHECK();
    return S2C_OK;
}

extern "C" int s2c_fit_cards(int num_slots, const int* slots, int iters, const int* res,
                             const int* quads, const int* vert_off, const int* bits_off,
                             const float* mvp, float* verts, float* params, float* adam_m,
                             float* adam_v,
                             int* steps, const int* limit, const unsigned* target_bits,
                             const float* hyper, float* loss_out, int chunk, void* stream) {
    S2C_CHECK(num_slots >= 0 && iters >= 0, "fit_cards: negative size");
    S2C_CHECK(hyper != nullptr, "fit_cards: null hyper-parameters");
    if (chunk == 0) chunk = kFitChunkDefault;
    S2C_CHECK(chunk == 256 || chunk == 512 || chunk == 1024,
              "fit_cards: chunk must be 0, 256, 512 or 1024");
    if (num_slots == 0) return S2C_OK;
    FitHyper hp{hyper[0], hyper[1], hyper[2], hyper[3], hyper[4],
                hyper[5], hyper[6], hyper[7], hyper[8], hyper[9]};
    S2C_CHECK(hp.R > 0.f && hp.R < 64.f, "fit_cards: kernel radius must be in (0, 64)");
    S2C_CHECK(hp.lam >= 0.f && hp.ww >= 0.f,
              "fit_cards: laplacian lambda and w weight must be >= 0");
    hipStream_t s = static_cast<hipStream_t>(stream);
#define S2C_FIT_LAUNCH(C)                                                                      \
    fit_cards_kernel<C><<<num_slots, kBlock, 0, s>>>(slots, iters, res, quads, vert_off,     \
                                                     bits_off, mvp, verts, params, adam_m,   \
                                                     adam_v, steps, limit,                   \
                                                     target_bits, hp, loss_out)
    if (chunk == 256) S2C_FIT_LAUNCH(256);
    else if (chunk == 512) S2C_FIT_LAUNCH(512);
    else S2C_FIT_LAUNCH(1024);
#undef S2C_FIT_LAUNCH
    S2C_LAUNCH_CHECK();
    return S2C_OK;
}
