// The large-steps parameterization of card grids (3_fitting.py:346-384):
// u = M x and x = M^{-1} u with M = I + lam * (L_h + w * L_w) for the
// concatenated grid meshes of texture_utils.py:525-655.  The reference
// factors M once with a sparse Cholesky (largesteps 0.2.2 over cholespy
// 2.1.0) and solves with it every iteration; here the factor is the set of
// Thomas pivots of each grid shape's decoupled tridiagonal systems
// (grid_math.hpp) and a solve is one workgroup (one wave) per grid:
// the grid's 3 S N floats staged in LDS with coalesced loads, projected onto
// the row basis, both elimination sweeps as wave scans (over registers when
// a lane's chunk has at most 16 elements, else over LDS), and the
// back-projection stored coalesced.
#include "grid_math.hpp"

namespace s2c {

constexpr size_t kGridLdsMax = 65536;

// LDS bytes of the solve for grids of at most `max_verts` vertices and
// `max_rows` rows: two arrays of 3 * max_verts floats and the S x S basis.
inline size_t grid_solve_lds(int max_verts, int max_rows) {
    return sizeof(float) * (6 * static_cast<size_t>(max_verts) +
                            static_cast<size_t>(max_rows) * max_rows);
}

// factor[fac_off[g] + k*N + j] = 1 / m_j of T_k (one lane per row mode k).
__global__ void __launch_bounds__(kWave)
grid_factor_kernel(const int* __restrict__ rows, const int* __restrict__ cols,
                   const int* __restrict__ fac_off, float lam, float ww,
                   float* __restrict__ factor) {
    const int g = blockIdx.x;
    const int S = rows[g], N = cols[g];
    float* im = factor + fac_off[g];
    for (int k = threadIdx.x; k < S; k += kWave) grid_factor(S, N, k, lam, ww, im + k * N);
}

// u = M x, one workgroup per grid.
__global__ void __launch_bounds__(kBlock)
grid_apply_kernel(const int* __restrict__ rows, const int* __restrict__ cols,
                  const int* __restrict__ voff, float lam, float ww,
                  const float* __restrict__ x, float* __restrict__ u) {
    const int g = blockIdx.x;
    const int S = rows[g], N = cols[g];
    const long long base = 3ll * voff[g];
    const int n = 3 * S * N;
    for (int q = threadIdx.x; q < n; q += kBlock)
        u[base + q] = grid_apply_at(S, N, lam, ww, x + base, q);
}

// y = (L_h + w L_w) x, one workgroup per grid: the uniform Laplacian of
// the grid's quad edges with the w_edges weighted (w = 1: the
// laplacian_edges of h_edges + w_edges, 3_fitting.py:359).
__global__ void __launch_bounds__(kBlock)
grid_laplacian_kernel(const int* __restrict__ rows, const int* __restrict__ cols,
                      const int* __restrict__ voff, float ww, const float* __restrict__ x,
                      float* __restrict__ y) {
    const int g = blockIdx.x;
    const int S = rows[g], N = cols[g];
    const long long base = 3ll * voff[g];
    const int n = 3 * S * N;
    for (int q = threadIdx.x; q < n; q += kBlock)
        y[base + q] = grid_laplacian_at(S, N, ww, x + base, q);
}

// Quad diagonals of the grids (the planarity term, 3_fitting.py:181-199):
// quad (i, j) of grid g, at row qoff[g] + i (N - 1) + j, has the corners
// [(i,j), (i+1,j), (i+1,j+1), (i,j+1)] of texture_utils.py:557-565 and
// D = (p0 + p2) - (p1 + p3).
__global__ void __launch_bounds__(kBlock)
grid_quad_diag_kernel(const int* __restrict__ rows, const int* __restrict__ cols,
                      const int* __restrict__ voff, const int* __restrict__ qoff,
                      const float* __restrict__ pos, float* __restrict__ D) {
    const int g = blockIdx.x;
    const int S = rows[g], N = cols[g];
    const long long v0 = voff[g], q0 = qoff[g];
    const int nq = (S - 1) * (N - 1);
    for (int q = threadIdx.x; q < 3 * nq; q += kBlock) {
        const int k = q / 3, c = q - 3 * k;
        const int i = k / (N - 1), j = k - i * (N - 1);
        const long long a = v0 + static_cast<long long>(i) * N + j;
        const float p0 = pos[3 * a + c], p1 = pos[3 * (a + N) + c];
        const float p2 = pos[3 * (a + N + 1) + c], p3 = pos[3 * (a + 1) + c];
        D[3 * (q0 + k) + c] = (p0 + p2) - (p1 + p3);
    }
}

// The adjoint: dpos[v] = sum over the quads at v of +dD (corners 0, 2) or
// -dD (corners 1, 3), in the fixed order: v as corner 0 of quad (i, j),
// corner 1 of (i-1, j), corner 2 of (i-1, j-1), corner 3 of (i, j-1).
__global__ void __launch_bounds__(kBlock)
grid_quad_diag_bwd_kernel(const int* __restrict__ rows, const int* __restrict__ cols,
                          const int* __restrict__ voff, const int* __restrict__ qoff,
                          const float* __restrict__ dD, float* __restrict__ dpos) {
    const int g = blockIdx.x;
    const int S = rows[g], N = cols[g];
    const long long v0 = voff[g], q0 = qoff[g];
    for (int q = threadIdx.x; q < 3 * S * N; q += kBlock) {
        const int v = q / 3, c = q - 3 * v;
        const int i = v / N, j = v - i * N;
        auto dq = [&](int qi, int qj) { return dD[3 * (q0 + qi * (N - 1) + qj) + c]; };
        float acc = 0.f;
        if (i < S - 1 && j < N - 1) acc += dq(i, j);
        if (i > 0 && j < N - 1) acc -= dq(i - 1, j);
        if (i > 0 && j > 0) acc += dq(i - 1, j - 1);
        if (i < S - 1 && j > 0) acc -= dq(i, j - 1);
        dpos[3 * (v0 + v) + c] = acc;
    }
}

// x = M^{-1} u, one workgroup (one wave) per grid.
__global__ void __launch_bounds__(kWave)
grid_solve_kernel(const int* __restrict__ rows, const int* __restrict__ cols,
                  const int* __restrict__ voff, const int* __restrict__ fac_off,
                  const float* __restrict__ factor, int max_verts, int max_rows, float lam,
                  const float* __restrict__ u, float* __restrict__ x) {
    extern __shared__ float lds[];
    const int g = blockIdx.x;
    const int S = rows[g], N = cols[g];
    const long long base = 3ll * voff[g];
    const int t = threadIdx.x;
    if (S < 1 || N < 1 || S > max_rows || N > max_verts / S) {
        // a layout the launch was not sized for: the grid reads as NaN
        // (S * N is not formed before N <= max_verts / S holds)
        if (S >= 1 && N >= 1 && N <= INT_MAX / 3 / S)
            for (int q = t; q < 3 * S * N; q += kWave) x[base + q] = __int_as_float(0x7fc00000);
        return;
    }
    const int n = 3 * S * N;
    float* U = lds;                  // staged u [(i N + j) 3 + c]; then the pivots
    float* Y = lds + 3 * max_verts;  // modes [(k 3 + c) N + j]
    float* Qb = Y + 3 * max_verts;   // basis [i S + k]
    for (int q = t; q < n; q += kWave) U[q] = u[base + q];
    for (int q = t; q < S * S; q += kWave) Qb[q] = grid_dct_q(q / S, q % S, S);
    __syncthreads();
    const float* im = factor + fac_off[g];
    const int L = grid_chunk(S, N);
    if (L <= 8) {  // chunks in registers (wave-uniform branch)
        grid_solve_wave_regs<8>(S, N, lam, Qb, U, im, Y, t);
    } else if (L <= 16) {
        grid_solve_wave_regs<16>(S, N, lam, Qb, U, im, Y, t);
    } else {  // long grids: the sweeps over LDS
        grid_project<kWave>(S, N, Qb, U, Y, t);
        __syncthreads();
        for (int q = t; q < S * N; q += kWave) U[q] = im[q];
        __syncthreads();
        grid_solve_wave(S, N, lam, U, Y, t);
    }
    __syncthreads();
    float* xg = x + base;
    grid_back_project<kWave>(S, N, Qb, Y, [&](int q, float v) { xg[q] = v; }, t);
}

}  // namespace s2c

using namespace s2c;

extern "C" int s2c_grid_system_factor(int num_grids, const int* rows, const int* cols,
                                      const int* fac_off, float lam, float w_weight,
                                      float* factor, void* stream) {
    S2C_CHECK(num_grids >= 0, "grid_system_factor: num_grids must be >= 0");
    S2C_CHECK(lam >= 0.f && w_weight >= 0.f, "grid_system_factor: lam and w_weight must be >= 0");
    if (num_grids == 0) return S2C_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    grid_factor_kernel<<<num_grids, kWave, 0, s>>>(rows, cols, fac_off, lam, w_weight, factor);
    S2C_LAUNCH_CHECK();
    return S2C_OK;
}

extern "C" int s2c_grid_system_apply(int num_grids, const int* rows, const int* cols,
                                     const int* voff, float lam, float w_weight, const float* x,
                                     float* u, void* stream) {
    S2C_CHECK(num_grids >= 0, "grid_system_apply: num_grids must be >= 0");
    if (num_grids == 0) return S2C_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    grid_apply_kernel<<<num_grids, kBlock, 0, s>>>(rows, cols, voff, lam, w_weight, x, u);
    S2C_LAUNCH_CHECK();
    return S2C_OK;
}

extern "C" int s2c_grid_laplacian(int num_grids, const int* rows, const int* cols,
                                  const int* voff, float w_weight, const float* x, float* y,
                                  void* stream) {
    S2C_CHECK(num_grids >= 0, "grid_laplacian: num_grids must be >= 0");
    if (num_grids == 0) return S2C_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    grid_laplacian_kernel<<<num_grids, kBlock, 0, s>>>(rows, cols, voff, w_weight, x, y);
    S2C_LAUNCH_CHECK();
    return S2C_OK;
}

extern "C" int s2c_grid_quad_diagonals(int num_grids, const int* rows, const int* cols,
                                       const int* voff, const int* qoff, const float* pos,
                                       float* D, void* stream) {
    S2C_CHECK(num_grids >= 0, "grid_quad_diagonals: num_grids must be >= 0");
    if (num_grids == 0) return S2C_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    grid_quad_diag_kernel<<<num_grids, kBlock, 0, s>>>(rows, cols, voff, qoff, pos, D);
    S2C_LAUNCH_CHECK();
    return S2C_OK;
}

extern "C" int s2c_grid_quad_diagonals_backward(int num_grids, const int* rows, const int* cols,
                                                const int* voff, const int* qoff,
                                                const float* dD, float* dpos, void* stream) {
    S2C_CHECK(num_grids >= 0, "grid_quad_diagonals_backward: num_grids must be >= 0");
    if (num_grids == 0) return S2C_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    grid_quad_diag_bwd_kernel<<<num_grids, kBlock, 0, s>>>(rows, cols, voff, qoff, dD, dpos);
    S2C_LAUNCH_CHECK();
    return S2C_OK;
}

extern "C" int s2c_grid_system_solve(int num_grids, const int* rows, const int* cols,
                                     const int* voff, const int* fac_off, const float* factor,
                                     int max_verts, int max_rows, float lam, const float* u,
                                     float* x, void* stream) {
    S2C_CHECK(num_grids >= 0, "grid_system_solve: num_grids must be >= 0");
    S2C_CHECK(max_verts >= 1 && max_rows >= 1 && max_rows <= max_verts,
              "grid_system_solve: bad max_verts / max_rows");
    S2C_CHECK(lam >= 0.f, "grid_system_solve: lam must be >= 0");
    const size_t lds = grid_solve_lds(max_verts, max_rows);
    S2C_CHECK(lds <= kGridLdsMax,
              "grid_system_solve: a grid of %d vertices (%d rows) needs %zu bytes of LDS, over "
              "the %zu of one workgroup", max_verts, max_rows, lds, kGridLdsMax);
    if (num_grids == 0) return S2C_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    grid_solve_kernel<<<num_grids, kWave, lds, s>>>(rows, cols, voff, fac_off, factor, max_verts,
                                                    max_rows, lam, u, x);
    S2C_LAUNCH_CHECK();
    return S2C_OK;
}
