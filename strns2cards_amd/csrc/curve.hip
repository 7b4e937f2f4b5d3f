// Gaussian arclength smoothing and rotation skinning of strands
// (reference: pystrns2cards/src/curve_utils.cu).
//
// The reference runs one thread per strand and loops over all n_v^2 pairs
// (curve_utils.cu:24-131), which leaves a 256-CU part mostly idle for tens of
// thousands of strands.  Here one thread owns one output vertex (strand,
// i); the strand's arclength row is staged in LDS by the workgroup, so the
// inner j-loop reads LDS and the per-vertex sum keeps the reference's
// summation order (j ascending), which keeps the result within fp32
// round-off of the oracle.
#include "common.hpp"

namespace s2c {

constexpr int kCurveMaxLds = 4096;  // arclength floats staged per workgroup

// Each workgroup covers kBlock consecutive (strand, vertex) outputs; the
// strands they touch are contiguous, so their arclength rows form one
// contiguous range that is copied to LDS when it fits.
__global__ void __launch_bounds__(kBlock)
smooth_kernel(int N, int P, const float* __restrict__ x, const float* __restrict__ s,
              float* __restrict__ out, float beta) {
    __shared__ float sl[kCurveMaxLds];
    const long long total = static_cast<long long>(N) * P;
    const long long g0 = static_cast<long long>(blockIdx.x) * kBlock;
    const long long n_first = g0 / P;
    const long long g_last = dmin<long long>(g0 + kBlock, total) - 1;
    const long long n_last = g_last / P;
    const long long span = (n_last - n_first + 1) * P;
    const bool staged = span <= kCurveMaxLds;
    if (staged) {
        for (long long k = threadIdx.x; k < span; k += kBlock) sl[k] = s[n_first * P + k];
    }
    __syncthreads();
    const long long g = g0 + threadIdx.x;
    if (g >= total) return;
    const long long n = g / P;
    const int i = static_cast<int>(g - n * P);
    const float* row = staged ? sl + (n - n_first) * P : s + n * P;
    const float* xs = x + n * P * 3;
    const float inv_b2 = 1.0f / (beta * beta);
    const float cut = 3.0f * beta;
    const float si = row[i];
    float a0 = 0.f, a1 = 0.f, a2 = 0.f, sw = 0.f;
    for (int j = 0; j < P; ++j) {
        const float dist = fabsf(si - row[j]);
        if (dist > cut) continue;
        const float w = expf(-(dist * dist) * inv_b2);
        a0 += w * xs[j * 3 + 0];
        a1 += w * xs[j * 3 + 1];
        a2 += w * xs[j * 3 + 2];
        sw += w;
    }
    if (sw > 1e-6f) {
        float* o = out + g * 3;
        o[0] = a0 / sw;
        o[1] = a1 / sw;
        o[2] = a2 / sw;
    }
}

__global__ void __launch_bounds__(kBlock)
skinning_kernel(int N, int P, const float* __restrict__ x, const float* __restrict__ s,
                const float* __restrict__ hnd, const float* __restrict__ can,
                const float* __restrict__ R, float* __restrict__ out, float beta) {
    __shared__ float sl[kCurveMaxLds];
    const long long total = static_cast<long long>(N) * P;
    const long long g0 = static_cast<long long>(blockIdx.x) * kBlock;
    const long long n_first = g0 / P;
    const long long g_last = dmin<long long>(g0 + kBlock, total) - 1;
    const long long n_last = g_last / P;
    const long long span = (n_last - n_first + 1) * P;
    const bool staged = span <= kCurveMaxLds;
    if (staged) {
        for (long long k = threadIdx.x; k < span; k += kBlock) sl[k] = s[n_first * P + k];
    }
    __syncthreads();
    const long long g = g0 + threadIdx.x;
    if (g >= total) return;
    const long long n = g / P;
    const int i = static_cast<int>(g - n * P);
    const float* row = staged ? sl + (n - n_first) * P : s + n * P;
    const long long o3 = n * P * 3;
    const float inv_b2 = 1.0f / (beta * beta);
    const float cut = 3.0f * beta;
    const float si = row[i];
    const float xi0 = x[o3 + i * 3 + 0], xi1 = x[o3 + i * 3 + 1], xi2 = x[o3 + i * 3 + 2];
    float a0 = 0.f, a1 = 0.f, a2 = 0.f, sw = 0.f;
    for (int j = 0; j < P; ++j) {
        const float dist = fabsf(si - row[j]);
        if (dist > cut) continue;
        const float w = expf(-(dist * dist) * inv_b2);
        const float* hj = hnd + o3 + j * 3;
        const float* cj = can + o3 + j * 3;
        const float* Rj = R + (n * P + j) * 9;
        const float l0 = xi0 - hj[0], l1 = xi1 - hj[1], l2 = xi2 - hj[2];
        const float r0 = Rj[0] * l0 + Rj[1] * l1 + Rj[2] * l2;
        const float r1 = Rj[3] * l0 + Rj[4] * l1 + Rj[5] * l2;
        const float r2 = Rj[6] * l0 + Rj[7] * l1 + Rj[8] * l2;
        a0 += w * (r0 + cj[0]);
        a1 += w * (r1 + cj[1]);
        a2 += w * (r2 + cj[2]);
        sw += w;
    }
    float* o = out + g * 3;
    if (sw > 1e-8f) {
        o[0] = a0 / sw;
        o[1] = a1 / sw;
        o[2] = a2 / sw;
    } else {
        o[0] = xi0;
        o[1] = xi1;
        o[2] = xi2;
    }
}

// ------------------------------------------------------------ Bishop frame
//
// Rotation-minimizing frames by double reflection (curve_utils.py:104-165):
// tangents are the normalized sums of the adjacent segments, the root frame
// comes from tangent_frame (curve_utils.py:76-102), and each next normal is
// the previous one reflected across the bisector planes of the segment and
// of the tangent change.  The recurrence is sequential along a strand, so
// one thread owns one strand; lanes read their strands' points in lock-step
// and each 128-byte line is reused for the next points from L1.

__device__ __forceinline__ float3 f3(const float* p) { return make_float3(p[0], p[1], p[2]); }
__device__ __forceinline__ float3 sub3(float3 a, float3 b) {
    return make_float3(a.x - b.x, a.y - b.y, a.z - b.z);
}
__device__ __forceinline__ float dot3(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float3 axpy3(float a, float3 x, float3 y) {  // y - a x
    return make_float3(y.x - a * x.x, y.y - a * x.y, y.z - a * x.z);
}
__device__ __forceinline__ float3 unit3(float3 v) {
    const float n = fmaxf(sqrtf(dot3(v, v)), 1e-8f);
    return make_float3(v.x / n, v.y / n, v.z / n);
}
__device__ __forceinline__ float3 cross3(float3 a, float3 b) {
    return make_float3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ void put3(float* p, float3 v) {
    p[0] = v.x;
    p[1] = v.y;
    p[2] = v.z;
}

__global__ void __launch_bounds__(kBlock)
bishop_kernel(int N, int S, const float* __restrict__ x, float* __restrict__ T,
              float* __restrict__ Nr, float* __restrict__ Bn) {
    const int n = blockIdx.x * kBlock + threadIdx.x;
    if (n >= N) return;
    const long long o = static_cast<long long>(n) * S * 3;
    const float* p = x + o;
    // tangents
    float3 prev_seg = make_float3(0.f, 0.f, 0.f);
    for (int i = 0; i < S; ++i) {
        const float3 seg = i + 1 < S ? sub3(f3(p + 3 * (i + 1)), f3(p + 3 * i))
                                     : make_float3(0.f, 0.f, 0.f);
        const float3 t = make_float3(seg.x + prev_seg.x, seg.y + prev_seg.y, seg.z + prev_seg.z);
        put3(T + o + 3 * i, unit3(t));
        prev_seg = seg;
    }
    if (S == 0) return;
    // root frame (tangent_frame)
    const float3 t0 = unit3(f3(T + o));
    const float sg = t0.z < 0.f ? -1.f : 1.f;
    const float a = -1.f / (sg + t0.z);
    const float b = t0.x * t0.y * a;
    float3 nprev = make_float3(1.f + sg * t0.x * t0.x * a, sg * b, -sg * t0.x);
    put3(Nr + o, nprev);
    put3(Bn + o, make_float3(b, sg + t0.y * t0.y * a, -t0.y));
    float3 tprev = f3(T + o);
    for (int i = 1; i < S; ++i) {
        const float3 v1 = sub3(f3(p + 3 * i), f3(p + 3 * (i - 1)));
        const float c1 = fmaxf(dot3(v1, v1), 1e-8f);
        const float3 nL = axpy3(2.f * (dot3(v1, nprev) / c1), v1, nprev);
        const float3 tL = axpy3(2.f * (dot3(v1, tprev) / c1), v1, tprev);
        const float3 tcur = f3(T + o + 3 * i);
        const float3 v2 = sub3(tcur, tL);
        const float c2 = fmaxf(dot3(v2, v2), 1e-8f);
        const float3 n1 = unit3(axpy3(2.f * (dot3(v2, nL) / c2), v2, nL));
        put3(Nr + o + 3 * i, n1);
        put3(Bn + o + 3 * i, unit3(cross3(tcur, n1)));
        nprev = n1;
        tprev = tcur;
    }
}

// ------------------------------------------------------ arclength resampling
//
// resample_strands_by_arclength (hair_loader.cpp:261-307): per strand the
// cumulative segment lengths normalized by the total; output sample j of T
// sits at t = j / (T - 1) and interpolates the two points around the first
// normalized length greater than t (an upper bound, clamped to [1, N-1]);
// the last sample copies the last point.  Pass 1 (one thread per strand)
// writes the normalized lengths in the reference's sequential order; pass 2
// (one thread per output sample) searches and interpolates.

__global__ void __launch_bounds__(kBlock)
resample_lengths_kernel(int num_strands, const int* __restrict__ num_samples,
                        const long long* __restrict__ offsets, const float* __restrict__ points,
                        float* __restrict__ lens) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= num_strands) return;
    const int N = num_samples[i];
    const long long o = offsets[i];
    const float* p = points + o * 3;
    float* l = lens + o;
    float acc = 0.f;
    l[0] = 0.f;
    for (int j = 1; j < N; ++j) {
        const float dx = p[j * 3 + 0] - p[(j - 1) * 3 + 0];
        const float dy = p[j * 3 + 1] - p[(j - 1) * 3 + 1];
        const float dz = p[j * 3 + 2] - p[(j - 1) * 3 + 2];
        acc = acc + sqrtf(dx * dx + dy * dy + dz * dz);
        l[j] = acc;
    }
    for (int j = 0; j < N; ++j) l[j] = l[j] / acc;
}

__global__ void __launch_bounds__(kBlock)
resample_points_kernel(int num_strands, int T, const int* __restrict__ num_samples,
                       const long long* __restrict__ offsets, const float* __restrict__ points,
                       const float* __restrict__ lens, float* __restrict__ out) {
    const long long q = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
    if (q >= static_cast<long long>(num_strands) * T) return;
    const int i = static_cast<int>(q / T), j = static_cast<int>(q - static_cast<long long>(i) * T);
    const int N = num_samples[i];
    const float* p = points + offsets[i] * 3;
    float* o = out + q * 3;
    if (j == T - 1) {
        o[0] = p[(N - 1) * 3 + 0];
        o[1] = p[(N - 1) * 3 + 1];
        o[2] = p[(N - 1) * 3 + 2];
        return;
    }
    const float* l = lens + offsets[i];
    const float t = static_cast<float>(j) / static_cast<float>(T - 1);
    int first = 0, len = N;  // first index with t < l[idx] (upper bound)
    while (len > 0) {
        const int half = len >> 1;
        if (t < l[first + half]) {
            len = half;
        } else {
            first += half + 1;
            len -= half + 1;
        }
    }
    const int idx = dmin(dmax(first, 1), N - 1);
    const float den = l[idx] - l[idx - 1];
    const float t0 = (l[idx] - t) / den;
    const float t1 = (t - l[idx - 1]) / den;
    for (int k = 0; k < 3; ++k) o[k] = t0 * p[(idx - 1) * 3 + k] + t1 * p[idx * 3 + k];
}

}  // namespace s2c

using namespace s2c;

extern "C" int s2c_resample_strands_by_arclength(int num_strands, const int* num_samples,
                                                 const long long* offsets, const float* points,
                                                 int target_num_samples, float* lengths,
                                                 float* out, void* stream) {
    S2C_CHECK(num_strands >= 0 && target_num_samples >= 1,
              "resample_strands_by_arclength: bad size");
    if (num_strands == 0) return S2C_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    resample_lengths_kernel<<<grid_for(num_strands, kBlock), kBlock, 0, s>>>(
        num_strands, num_samples, offsets, points, lengths);
    S2C_LAUNCH_CHECK();
    const long long n = static_cast<long long>(num_strands) * target_num_samples;
    resample_points_kernel<<<grid_for(n, kBlock), kBlock, 0, s>>>(
        num_strands, target_num_samples, num_samples, offsets, points, lengths, out);
    S2C_LAUNCH_CHECK();
    return S2C_OK;
}

extern "C" int s2c_bishop_frame(int num_strands, int num_points, const float* strands,
                                float* tangents, float* normals, float* binormals,
                                void* stream) {
    S2C_CHECK(num_strands >= 0 && num_points >= 0, "bishop_frame: negative size");
    if (num_strands == 0 || num_points == 0) return S2C_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    bishop_kernel<<<grid_for(num_strands, kBlock), kBlock, 0, s>>>(num_strands, num_points, strands,
                                                                  tangents, normals, binormals);
    S2C_LAUNCH_CHECK();
    return S2C_OK;
}

extern "C" int s2c_smooth_strands(int num_strands, int num_points, const float* strands,
                                  const float* arclength, float* smoothed_strands, float beta,
                                  void* stream) {
    S2C_CHECK(num_strands >= 0 && num_points >= 0, "smooth_strands: negative size");
    const long long n = static_cast<long long>(num_strands) * num_points;
    if (n == 0) return S2C_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    smooth_kernel<<<grid_for(n, kBlock), kBlock, 0, s>>>(num_strands, num_points, strands,
                                                         arclength, smoothed_strands, beta);
    S2C_LAUNCH_CHECK();
    return S2C_OK;
}

extern "C" int s2c_skinning(int num_strands, int num_points, const float* strands,
                            const float* arclength, const float* handle_strands,
                            const float* canonical_handles, const float* R,
                            float* skinned_strands, float beta, void* stream) {
    S2C_CHECK(num_strands >= 0 && num_points >= 0, "skinning: negative size");
    const long long n = static_cast<long long>(num_strands) * num_points;
    if (n == 0) return S2C_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    skinning_kernel<<<grid_for(n, kBlock), kBlock, 0, s>>>(num_strands, num_points, strands,
                                                           arclength, handle_strands,
                                                           canonical_handles, R,
                                                           skinned_strands, beta);
    S2C_LAUNCH_CHECK();
    return S2C_OK;
}
