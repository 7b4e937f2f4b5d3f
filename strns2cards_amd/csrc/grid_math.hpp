// Exact solve of the large-steps system M = I + lam * (L_h + w L_w) of one
// grid mesh (3_fitting.py:347-356; grid of S rows by N columns, vertex
// (i, j) at i*N + j, h_edges along rows, w_edges across them,
// texture_utils.py:566-584), shared by the standalone solve and the
// fitting kernel.
//
// In (column, row) order M = I + lam * (L_N (x) I_S + w I_N (x) L_S), L_S the
// Laplacian of the S-vertex path, whose orthonormal eigenbasis is the DCT-II
// Q[i][k] = c_k cos(pi (i + 1/2) k / S) (c_0 = sqrt(1/S), else sqrt(2/S)),
// eigenvalues mu_k = 2 - 2 cos(pi k / S).  Projecting every column onto Q
// splits M x = u into S tridiagonal systems along the columns,
//   T_k y = r,  T_k = (1 + lam w mu_k) I + lam L_N,  r_j = (Q^T u_j)_k,
// then x_j = Q y_j.  T_k is strictly diagonally dominant (each row's excess
// is 1 + lam w mu_k >= 1), so Thomas elimination needs no pivoting; its
// pivots depend on (k, N, lam, w) only and are factored once:
//   m_0 = t_0,  m_j = t_j - lam^2 / m_{j-1},  t_j = 1 + lam w mu_k + lam deg_j,
// stored as im_j = 1 / m_j.  A solve is then division-free:
//   forward   d_j = lam im_j d_{j-1} + im_j r_j   (d_{-1} = 0)
//   backward  y_j = lam im_j y_{j+1} + d_j        (y_N = 0).
#pragma once

#include "common.hpp"

namespace s2c {

// v / n and v % n for 0 <= v < 2^22 without an integer divide: the float
// quotient is within 1 of the exact one (two roundings of relative size
// 2^-24 on a value below 2^22), and one correction step fixes it.
struct DivN {
    int n;
    float inv;
    __device__ __forceinline__ explicit DivN(int n_) : n(n_), inv(1.f / static_cast<float>(n_)) {}
    __device__ __forceinline__ int div(int v, int& rem) const {
        int d = static_cast<int>(static_cast<float>(v) * inv);
        int r = v - d * n;
        if (r < 0) { --d; r += n; }
        else if (r >= n) { ++d; r -= n; }
        rem = r;
        return d;
    }
};

__device__ __forceinline__ float grid_dct_q(int i, int k, int S) {
    const float pi = 3.14159265358979323846f;
    const float fs = static_cast<float>(S);
    const float c = k == 0 ? sqrtf(1.f / fs) : sqrtf(2.f / fs);
    return c * cosf(pi * (static_cast<float>(i) + 0.5f) * static_cast<float>(k) / fs);
}

// Pivot reciprocals im[0 .. N) of T_k (one lane, serial in j).
__device__ __forceinline__ void grid_factor(int S, int N, int k, float lam, float ww, float* im) {
    const float pi = 3.14159265358979323846f;
    const float mu = 2.f - 2.f * cosf(pi * static_cast<float>(k) / static_cast<float>(S));
    const float dk = 1.f + lam * ww * mu;
    const float lam2 = lam * lam;
    float prev = 0.f;  // im_{j-1}
    for (int j = 0; j < N; ++j) {
        const float deg = static_cast<float>((j > 0) + (j < N - 1));
        const float m = (dk + lam * deg) - lam2 * prev;
        prev = 1.f / m;
        im[j] = prev;
    }
}

// Chunk length of a lane in the wave solve of a grid with 3S systems of N:
// odd, so that the lanes' strided LDS accesses fall on distinct banks.
__device__ __forceinline__ int grid_chunk(int S, int N) {
    return ((3 * S * N + kWave - 1) / kWave) | 1;
}

// One wave: y <- T^{-1} y for the 3S systems of one grid stored back to
// back in Y[(k*3 + c)*N + j], system k*3 + c using the pivots im[k*N ..].
// Both sweeps are first-order affine recurrences over the whole array
// (a coefficient 0 at each system's first / last column keeps the systems
// apart), so each lane composes the maps of its contiguous chunk, a wave
// scan of the composed maps gives every chunk its incoming value, and the
// lane replays its chunk from it:
//   forward   d_q = (lam im_j) d_{q-1} + im_j y_q      (lam im_j -> 0 at j = 0)
//   backward  y_q = (lam im_j) y_{q+1} + d_q           (-> 0 at j = N - 1)
__device__ __forceinline__ void grid_solve_wave(int S, int N, float lam, const float* im,
                                                float* Y, int lane) {
    const int n = 3 * S * N;
    const int L = grid_chunk(S, N);
    const int lo = min(lane * L, n), hi = min(lo + L, n);
    const DivN dn(N);
    int j0;
    const int kc0 = dn.div(lo, j0);
    // forward sweep: the chunk maps d_in to A d_in + B
    float A = 1.f, B = 0.f;
    {
        int kc = kc0, j = j0;
        for (int q = lo; q < hi; ++q) {
            const float m = im[(kc / 3) * N + j];
            const float a = j == 0 ? 0.f : lam * m;
            A = a * A;
            B = a * B + m * Y[q];
            if (++j == N) { j = 0; ++kc; }
        }
    }
    for (int s = 1; s < kWave; s <<= 1) {  // compose with the earlier lanes
        const float pA = __shfl_up(A, s, kWave), pB = __shfl_up(B, s, kWave);
        if (lane >= s) {
            B = A * pB + B;
            A = A * pA;
        }
    }
    float d = __shfl_up(B, 1, kWave);
    if (lane == 0) d = 0.f;
    {
        int kc = kc0, j = j0;
        for (int q = lo; q < hi; ++q) {
            const float m = im[(kc / 3) * N + j];
            const float a = j == 0 ? 0.f : lam * m;
            d = a * d + m * Y[q];
            Y[q] = d;
            if (++j == N) { j = 0; ++kc; }
        }
    }
    // backward sweep over the same chunk, last element first
    int j1;
    const int kc1 = dn.div(max(hi - 1, 0), j1);
    A = 1.f;
    B = 0.f;
    {
        int kc = kc1, j = j1;
        for (int q = hi - 1; q >= lo; --q) {
            const float c = j == N - 1 ? 0.f : lam * im[(kc / 3) * N + j];
            A = c * A;
            B = c * B + Y[q];
            if (--j < 0) { j = N - 1; --kc; }
        }
    }
    for (int s = 1; s < kWave; s <<= 1) {  // compose with the later lanes
        const float nA = __shfl_down(A, s, kWave), nB = __shfl_down(B, s, kWave);
        if (lane + s < kWave) {
            B = A * nB + B;
            A = A * nA;
        }
    }
    float y = __shfl_down(B, 1, kWave);
    if (lane == kWave - 1) y = 0.f;
    {
        int kc = kc1, j = j1;
        for (int q = hi - 1; q >= lo; --q) {
            const float c = j == N - 1 ? 0.f : lam * im[(kc / 3) * N + j];
            y = c * y + Y[q];
            Y[q] = y;
            if (--j < 0) { j = N - 1; --kc; }
        }
    }
}

// One wave, for grids whose chunk fits kL registers: the projection
// Y = Q^T in (per column) and both elimination sweeps of grid_solve_wave in
// registers, the result written to Y[(k*3 + c)*N + j].  `in` is the grid
// vertex-major in LDS, `im` the pivots (im[k*N + j]).  Same chunks and the
// same operations per element as grid_project + grid_solve_wave.
template <int kL>
__device__ __forceinline__ void grid_solve_wave_regs(int S, int N, float lam, const float* Qb,
                                                     const float* in, const float* im, float* Y,
                                                     int lane) {
    const int n = 3 * S * N;
    const int L = grid_chunk(S, N);
    const int lo = min(lane * L, n), hi = min(lo + L, n);
    const int cnt = hi - lo;
    const DivN dn(N);
    int j;
    int kc = dn.div(lo, j);
    // per element: value r, pivot m and the sweep coefficients (forward a,
    // backward c); elements past the chunk are identity maps (a = c = 1,
    // r = m = 0)
    float r[kL], m[kL], a[kL], c[kL];
#pragma unroll
    for (int e = 0; e < kL; ++e) {
        const bool ok = e < cnt;
        const int k = ok ? kc / 3 : 0, cc = ok ? kc - 3 * (kc / 3) : 0, jj = ok ? j : 0;
        float acc = 0.f;
        for (int i = 0; i < S; ++i) acc += Qb[i * S + k] * in[(i * N + jj) * 3 + cc];
        const float mm = im[k * N + jj];
        r[e] = ok ? acc : 0.f;
        m[e] = ok ? mm : 0.f;
        a[e] = ok ? (jj == 0 ? 0.f : lam * mm) : 1.f;
        c[e] = ok ? (jj == N - 1 ? 0.f : lam * mm) : 1.f;
        if (++j == N) { j = 0; ++kc; }
    }
    float A = 1.f, B = 0.f;
#pragma unroll
    for (int e = 0; e < kL; ++e) {
        A = a[e] * A;
        B = a[e] * B + m[e] * r[e];
    }
    for (int s = 1; s < kWave; s <<= 1) {
        const float pA = __shfl_up(A, s, kWave), pB = __shfl_up(B, s, kWave);
        if (lane >= s) {
            B = A * pB + B;
            A = A * pA;
        }
    }
    float d = __shfl_up(B, 1, kWave);
    if (lane == 0) d = 0.f;
#pragma unroll
    for (int e = 0; e < kL; ++e) {
        d = a[e] * d + m[e] * r[e];
        r[e] = e < cnt ? d : 0.f;  // padding stays an identity map
    }
    A = 1.f;
    B = 0.f;
#pragma unroll
    for (int e = kL - 1; e >= 0; --e) {
        A = c[e] * A;
        B = c[e] * B + r[e];
    }
    for (int s = 1; s < kWave; s <<= 1) {
        const float nA = __shfl_down(A, s, kWave), nB = __shfl_down(B, s, kWave);
        if (lane + s < kWave) {
            B = A * nB + B;
            A = A * nA;
        }
    }
    float y = __shfl_down(B, 1, kWave);
    if (lane == kWave - 1) y = 0.f;
#pragma unroll
    for (int e = kL - 1; e >= 0; --e) {
        y = c[e] * y + r[e];
        if (e < cnt) Y[lo + e] = y;
    }
}

// Y[(k*3 + c)*N + j] = sum_i Q[i][k] * in[(i*N + j)*3 + c]   (Q^T per column)
template <int kThreads>
__device__ __forceinline__ void grid_project(int S, int N, const float* Qb, const float* in,
                                             float* Y, int tid) {
    const int n = 3 * S * N;
    const DivN dn(N);
    for (int q = tid; q < n; q += kThreads) {
        int j;
        const int kc = dn.div(q, j);
        const int k = kc / 3, c = kc - 3 * k;
        float r = 0.f;
        for (int i = 0; i < S; ++i) r += Qb[i * S + k] * in[(i * N + j) * 3 + c];
        Y[q] = r;
    }
}

// out[(i*N + j)*3 + c] = sum_k Q[i][k] * Y[(k*3 + c)*N + j]   (Q per column)
template <int kThreads, class Out>
__device__ __forceinline__ void grid_back_project(int S, int N, const float* Qb, const float* Y,
                                                  Out out, int tid) {
    const int n = 3 * S * N;
    const DivN dn(N);
    for (int q = tid; q < n; q += kThreads) {
        const int v = q / 3, c = q - 3 * v;
        int j;
        const int i = dn.div(v, j);
        float acc = 0.f;
        for (int k = 0; k < S; ++k) acc += Qb[i * S + k] * Y[(k * 3 + c) * N + j];
        out(q, acc);
    }
}

// (L_h + w L_w) x at vertex-major index q = v*3 + c of one grid: the
// degree-weighted differences with the row (h) and column (w) neighbours.
__device__ __forceinline__ float grid_laplacian_at(int S, int N, float ww, const float* x,
                                                   int q) {
    const int v = q / 3;
    const int i = v / N, j = v - i * N;
    const float xv = x[q];
    float h = 0.f, w = 0.f;
    if (j > 0) h += xv - x[q - 3];
    if (j < N - 1) h += xv - x[q + 3];
    if (i > 0) w += xv - x[q - 3 * N];
    if (i < S - 1) w += xv - x[q + 3 * N];
    return h + ww * w;
}

// u = M x at vertex-major index q = v*3 + c of one grid (x vertex-major).
__device__ __forceinline__ float grid_apply_at(int S, int N, float lam, float ww, const float* x,
                                               int q) {
    const int v = q / 3;
    const int i = v / N, j = v - i * N;
    const float xv = x[q];
    float h = 0.f, w = 0.f;
    if (j > 0) h += xv - x[q - 3];
    if (j < N - 1) h += xv - x[q + 3];
    if (i > 0) w += xv - x[q - 3 * N];
    if (i < S - 1) w += xv - x[q + 3 * N];
    return xv + lam * (h + ww * w);
}

}  // namespace s2c
