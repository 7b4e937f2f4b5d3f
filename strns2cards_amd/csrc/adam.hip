// UniformAdam step for one parameter of n points x dim channels
// (reference: pystrns2cards/optimizer.py:33-68), in two launches and no host
// synchronization: the moments with per-workgroup maxima of the
// bias-corrected RMS, then the shared denominator and the update.
#include <cmath>

#include "common.hpp"

namespace s2c {

// At most this many workgroups stride over the points, so the partial
// maxima stay few enough for every workgroup of the second launch to reduce
// them all.
constexpr int kAdamBlocks = 1024;

// The optimizer's scalars as float32, each rounded once from the double the
// reference's Python arithmetic produces (b1 ** step etc.).
struct AdamScalars {
    float b1, a1, b2, a2, c1, c2, lr;
};

// optimizer.py:55-61.  g1 <- g1*b1 + (1-b1)*g; g2 <- g2*b2 + (1-b2)*sum_c g_c^2;
// partial[block] = the largest sqrt(g2 / c2) of the block's points, as float
// bits: the values are >= 0 or NaN (sign cleared), and those bit patterns
// order as int32 with every NaN above +inf, so a NaN wins the maximum as it
// does in torch's max().
__global__ void __launch_bounds__(kBlock)
uniform_adam_moments_kernel(long long n, int dim, AdamScalars sc, const float* __restrict__ grad,
                            float* __restrict__ g1, float* __restrict__ g2,
                            int* __restrict__ partial) {
    __shared__ int sh[kBlock / kWave];
    int best = 0;
    const long long stride = static_cast<long long>(gridDim.x) * kBlock;
    for (long long i = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x; i < n;
         i += stride) {
        float ss = 0.f;
        for (int c = 0; c < dim; ++c) {
            const long long k = i * dim + c;
            const float g = grad[k];
            g1[k] = g1[k] * sc.b1 + sc.a1 * g;
            ss += g * g;
        }
        const float v = g2[i] * sc.b2 + sc.a2 * ss;
        g2[i] = v;
        best = dmax(best, static_cast<int>(__float_as_uint(fabsf(sqrtf(v / sc.c2)))));
    }
    best = wave_max(best);
    if (lane_id() == 0) sh[threadIdx.x / kWave] = best;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kBlock / kWave; ++w) best = dmax(best, sh[w]);
        partial[blockIdx.x] = best;
    }
}

// optimizer.py:63-68.  denom = max sqrt(m2) + 1e-8; p <- p - lr * (m1 / denom)
// with m1 = g1 / c1.
__global__ void __launch_bounds__(kBlock)
uniform_adam_apply_kernel(long long total, int nparts, AdamScalars sc,
                          const int* __restrict__ partial, const float* __restrict__ g1,
                          float* __restrict__ param) {
    __shared__ float denom_sh;
    if (threadIdx.x < kWave) {  // the whole first wave
        int m = 0;
        for (int j = threadIdx.x; j < nparts; j += kWave) m = dmax(m, partial[j]);
        m = wave_max(m);
        if (threadIdx.x == 0) denom_sh = __int_as_float(m) + 1e-8f;
    }
    __syncthreads();
    const float denom = denom_sh;
    const long long stride = static_cast<long long>(gridDim.x) * kBlock;
    for (long long k = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x; k < total;
         k += stride)
        param[k] = param[k] - sc.lr * ((g1[k] / sc.c1) / denom);
}

}  // namespace s2c

using namespace s2c;

extern "C" int s2c_uniform_adam_step(int n, int dim, double lr, double b1, double b2, int step,
                                     float* param, const float* grad, float* g1, float* g2,
                                     void* stream) {
    S2C_CHECK(n >= 0 && dim >= 1, "uniform_adam_step: n must be >= 0 and dim >= 1");
    S2C_CHECK(step >= 1, "uniform_adam_step: step counts from 1 (after the increment)");
    if (n == 0) return S2C_OK;
    AdamScalars sc;
    sc.b1 = static_cast<float>(b1);
    sc.a1 = static_cast<float>(1.0 - b1);
    sc.b2 = static_cast<float>(b2);
    sc.a2 = static_cast<float>(1.0 - b2);
    sc.c1 = static_cast<float>(1.0 - std::pow(b1, static_cast<double>(step)));
    sc.c2 = static_cast<float>(1.0 - std::pow(b2, static_cast<double>(step)));
    sc.lr = static_cast<float>(lr);
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int nb = static_cast<int>(dmin<long long>(grid_for(n, kBlock), kAdamBlocks));
    Temp partial;
    if (int r = partial.alloc(sizeof(int) * static_cast<size_t>(nb), s)) return r;
    uniform_adam_moments_kernel<<<nb, kBlock, 0, s>>>(n, dim, sc, grad, g1, g2,
                                                      partial.as<int>());
    S2C_LAUNCH_CHECK();
    const long long total = static_cast<long long>(n) * dim;
    const int na = static_cast<int>(dmin<long long>(grid_for(total, kBlock), kAdamBlocks));
    uniform_adam_apply_kernel<<<na, kBlock, 0, s>>>(total, nb, sc, partial.as<int>(), g1, param);
    S2C_LAUNCH_CHECK();
    return S2C_OK;
}
