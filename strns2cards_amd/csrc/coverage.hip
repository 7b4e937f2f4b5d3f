// Area-sampled card coverage (forward) and its hand-written adjoint
// (backward) for the differentiable card-fitting loss.  Geometry is
// recomputed from the clip-space vertices in the backward pass; nothing but
// the vertices and the triangle list is kept between the two passes.
#include "coverage_math.hpp"
#include "raster_math.hpp"

#include <climits>

namespace s2c {

struct CovTri {
    ScreenTri st;
    int h0, rows;  // pixel rows [h0, h0 + rows) of the bounding box, clamped
    int b, t;
};

__device__ __forceinline__ int cov_rows(const ScreenTri& st, int H, int W, int& h0) {
    h0 = 0;
    if (!st.valid) return 0;
    const float xmin = fminf(fminf(st.p[0].x, st.p[1].x), st.p[2].x);
    const float xmax = fmaxf(fmaxf(st.p[0].x, st.p[1].x), st.p[2].x);
    const float ymin = fminf(fminf(st.p[0].y, st.p[1].y), st.p[2].y);
    const float ymax = fmaxf(fmaxf(st.p[0].y, st.p[1].y), st.p[2].y);
    if (!(xmax >= 0.f && xmin < static_cast<float>(W))) return 0;
    h0 = dmax(f2i(floorf(ymin)), 0);
    const int h1 = dmin(f2i(floorf(ymax)), H - 1);
    return dmax(h1 - h0 + 1, 0);
}

// Pixel columns [w0, w1] of row y that the triangle can cover: the x
// extent of the triangle cut to the strip y <= Y <= y+1 (vertices inside
// the strip and the edges' crossings of its two lines), widened by 1e-4 px
// so that rounding never drops a covered pixel; clamped to the image.
__device__ __forceinline__ bool cov_span(const ScreenTri& st, float y, int W, int& w0, int& w1) {
    float lo = 3.4e38f, hi = -3.4e38f;
#pragma unroll
    for (int e = 0; e < 3; ++e) {
        const float2 a = st.p[e], b = st.p[e == 2 ? 0 : e + 1];
        if (a.y >= y && a.y <= y + 1.f) {
            lo = fminf(lo, a.x);
            hi = fmaxf(hi, a.x);
        }
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const float yl = y + static_cast<float>(k);
            if ((a.y - yl) * (b.y - yl) < 0.f) {
                const float x = a.x + (yl - a.y) / (b.y - a.y) * (b.x - a.x);
                lo = fminf(lo, x);
                hi = fmaxf(hi, x);
            }
        }
    }
    if (!(hi >= lo)) return false;
    w0 = dmax(f2i(floorf(lo - 1e-4f)), 0);
    w1 = dmin(f2i(floorf(hi + 1e-4f)), W - 1);
    return w1 >= w0;
}

// Loads the workgroup's triangles into LDS and returns the number of
// (triangle, row) items; excl[] gets each triangle's first item.
__device__ __forceinline__ int cov_load(long long BT, int H, int W, int V,
                                        const float* __restrict__ pos, int T,
                                        const int* __restrict__ tri, CovTri* tris, int* excl,
                                        int* sh) {
    const long long q = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
    int cnt = 0;
    if (q < BT) {
        CovTri c;
        c.b = static_cast<int>(q / T);
        c.t = static_cast<int>(q - static_cast<long long>(c.b) * T);
        const long long vb = static_cast<long long>(c.b) * V;
        c.st = screen_tri(ld4(pos, vb + tri[c.t * 3 + 0]), ld4(pos, vb + tri[c.t * 3 + 1]),
                          ld4(pos, vb + tri[c.t * 3 + 2]), static_cast<float>(W),
                          static_cast<float>(H));
        cnt = cov_rows(c.st, H, W, c.h0);
        c.rows = cnt;
        tris[threadIdx.x] = c;
    }
    int total;
    excl[threadIdx.x] = block_excl_scan(cnt, sh, &total);
    __syncthreads();
    return total;
}

__device__ __forceinline__ int cov_find(const int* excl, int j) {
    int lo = 0, hi = kBlock;  // last slot with excl <= j
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (excl[mid] <= j) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Work decomposition: a workgroup takes kBlock consecutive (view,
// triangle) pairs and spreads their (triangle, pixel row) items over its
// threads.  The backward visits a row's span in one thread (its gradient
// sums stay in registers); the forward spreads the rows' pixels over the
// threads.
__global__ void __launch_bounds__(kBlock)
coverage_fwd_kernel(long long BT, int H, int W, int V, const float* __restrict__ pos, int T,
                    const int* __restrict__ tri, float* __restrict__ cov) {
    __shared__ CovTri tris[kBlock];
    __shared__ int excl[kBlock];
    __shared__ int sh[kBlock / kWave + 1];
    __shared__ int it_s[kBlock], it_y[kBlock], it_w0[kBlock], it_ex[kBlock];
    const int total = cov_load(BT, H, W, V, pos, T, tri, tris, excl, sh);
    // kBlock rows at a time; their pixels are spread over the threads so
    // that neighbouring lanes add into neighbouring pixels (coalesced atomics)
    for (int jb = 0; jb < total; jb += kBlock) {
        const int j = jb + threadIdx.x;
        int cnt = 0, si = 0, y = 0, w0 = 0;
        if (j < total) {
            si = cov_find(excl, j);
            const CovTri& c = tris[si];
            y = c.h0 + (j - excl[si]);
            int w1;
            if (cov_span(c.st, static_cast<float>(y), W, w0, w1)) cnt = w1 - w0 + 1;
        }
        int npx;
        const int ex = block_excl_scan(cnt, sh, &npx);
        it_s[threadIdx.x] = si;
        it_y[threadIdx.x] = y;
        it_w0[threadIdx.x] = w0;
        it_ex[threadIdx.x] = ex;
        __syncthreads();
        for (int p = threadIdx.x; p < npx; p += kBlock) {
            const int k = cov_find(it_ex, p);
            const CovTri& c = tris[it_s[k]];
            const int yy = it_y[k], x = it_w0[k] + (p - it_ex[k]);
            const float a = clip_area(c.st, static_cast<float>(x), static_cast<float>(yy));
            if (a != 0.f) atomicAdd(&cov[(static_cast<long long>(c.b) * H + yy) * W + x], a);
        }
        __syncthreads();
    }
}

__global__ void __launch_bounds__(kBlock)
coverage_bwd_kernel(long long BT, int H, int W, int V, const float* __restrict__ pos, int T,
                    const int* __restrict__ tri, const float* __restrict__ d_cov,
                    float* __restrict__ d_pos) {
    __shared__ CovTri tris[kBlock];
    __shared__ int excl[kBlock];
    __shared__ int sh[kBlock / kWave + 1];
    __shared__ float gacc[kBlock][6];
    for (int k = 0; k < 6; ++k) gacc[threadIdx.x][k] = 0.f;
    const int total = cov_load(BT, H, W, V, pos, T, tri, tris, excl, sh);
    for (int j = threadIdx.x; j < total; j += kBlock) {
        const int s = cov_find(excl, j);
        const CovTri& c = tris[s];
        const int y = c.h0 + (j - excl[s]);
        int w0, w1;
        if (!cov_span(c.st, static_cast<float>(y), W, w0, w1)) continue;
        const float* row = d_cov + (static_cast<long long>(c.b) * H + y) * W;
        float2 gr[3] = {make_float2(0.f, 0.f), make_float2(0.f, 0.f), make_float2(0.f, 0.f)};
        for (int x = w0; x <= w1; ++x) {
            const float g = row[x];
            if (g == 0.f) continue;
            clip_area_grad(c.st, static_cast<float>(x), static_cast<float>(y), g, gr[0], gr[1],
                           gr[2]);
        }
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            if (gr[i].x != 0.f) atomicAdd(&gacc[s][2 * i + 0], gr[i].x);
            if (gr[i].y != 0.f) atomicAdd(&gacc[s][2 * i + 1], gr[i].y);
        }
    }
    __syncthreads();
    const long long q = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
    if (q >= BT) return;
    const CovTri& c = tris[threadIdx.x];
    if (!c.st.valid) return;
    const float fW = static_cast<float>(W), fH = static_cast<float>(H);
    for (int i = 0; i < 3; ++i) {
        const float2 gs = make_float2(gacc[threadIdx.x][2 * i], gacc[threadIdx.x][2 * i + 1]);
        if (gs.x == 0.f && gs.y == 0.f) continue;
        const int v = tri[c.t * 3 + c.st.map[i]];
        const long long row = static_cast<long long>(c.b) * V + v;
        const float3 g3 = screen_to_clip_grad(ld4(pos, row), gs, fW, fH);
        atomicAdd(&d_pos[row * 4 + 0], g3.x);
        atomicAdd(&d_pos[row * 4 + 1], g3.y);
        atomicAdd(&d_pos[row * 4 + 3], g3.z);
    }
}

}  // namespace s2c

using namespace s2c;

extern "C" int s2c_card_coverage_forward(int B, int H, int W, int V, const float* pos, int T,
                                         const int* tri, float* coverage, void* stream) {
    S2C_CHECK(B >= 0 && H >= 0 && W >= 0 && V >= 0 && T >= 0, "card_coverage_forward: negative size");
    S2C_CHECK(static_cast<long long>(H) * kBlock <= INT_MAX,
              "card_coverage_forward: image taller than 8,388,607 rows");
    const long long BT = static_cast<long long>(B) * T;
    if (BT == 0 || H == 0 || W == 0) return S2C_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    coverage_fwd_kernel<<<grid_for(BT, kBlock), kBlock, 0, s>>>(BT, H, W, V, pos, T, tri, coverage);
    S2C_LAUNCH_CHECK();
    return S2C_OK;
}

extern "C" int s2c_card_coverage_backward(int B, int H, int W, int V, const float* pos, int T,
                                          const int* tri, const float* d_coverage, float* d_pos,
                                          void* stream) {
    S2C_CHECK(B >= 0 && H >= 0 && W >= 0 && V >= 0 && T >= 0,
              "card_coverage_backward: negative size");
    S2C_CHECK(static_cast<long long>(H) * kBlock <= INT_MAX,
              "card_coverage_backward: image taller than 8,388,607 rows");
    const long long BT = static_cast<long long>(B) * T;
    if (BT == 0 || H == 0 || W == 0) return S2C_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    coverage_bwd_kernel<<<grid_for(BT, kBlock), kBlock, 0, s>>>(BT, H, W, V, pos, T, tri,
                                                                d_coverage, d_pos);
    S2C_LAUNCH_CHECK();
    return S2C_OK;
}
