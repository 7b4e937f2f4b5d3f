// DDA strand-line rasterization (reference: pystrns2cards/src/dda.cu).
#include "dda_math.hpp"

namespace s2c {

// ------------------------------------------------------------ span + scan

// dda.cu:9-87.  Counts go to `prefix` (scanned in place afterwards).
__global__ void __launch_bounds__(kBlock)
dda_span_kernel(int P, const int* __restrict__ edge_ids, int H, int W, int V,
                const float* __restrict__ pos, int E, const int* __restrict__ edges,
                int* __restrict__ counts, float* __restrict__ slopes,
                float* __restrict__ spans) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= P) return;
    const int id = edge_ids[i];
    const int b = id / E;
    const int e = id % E;
    const int2 ev = reinterpret_cast<const int2*>(edges)[e];
    const long long vb = static_cast<long long>(b) * V;
    const float4 p0 = load_pos(pos, vb + ev.x);
    const float4 p1 = load_pos(pos, vb + ev.y);
    DdaSpan sp;
    bool wrote;
    const int c = dda_span(H, W, p0, p1, sp, wrote);
    counts[i] = c;
    if (wrote) {
        reinterpret_cast<float2*>(slopes)[i] = make_float2(sp.axis, sp.slope);
        reinterpret_cast<float4*>(spans)[i] = make_float4(sp.sx, sp.sy, sp.off0, sp.off1);
    }
}

// ------------------------------------------------------- fragment expansion
// Each workgroup owns kExpand consecutive fragment slots.  Thread 0 finds
// the primitive range of the workgroup once (two searches over the global
// prefix); every slot then searches only inside that range, which stays in
// L1/L2.  This replaces the per-fragment search over the whole prefix array
// (dda.cu:164-172).

constexpr int kExpandItems = 4;
constexpr int kExpand = kBlock * kExpandItems;

// First index i in [lo, hi) with prefix[i] > key (prefix is inclusive and
// non-decreasing); hi if none.
__device__ __forceinline__ int upper_bound(const int* prefix, int lo, int hi, int key) {
    while (lo < hi) {
        const int mid = lo + ((hi - lo) >> 1);
        if (prefix[mid] <= key) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

__global__ void __launch_bounds__(kBlock)
dda_fragments_kernel(int P, int N, const int* __restrict__ prefix,
                     const int* __restrict__ edge_ids, const float* __restrict__ slopes,
                     const float* __restrict__ spans, int H, int W, int V,
                     const float* __restrict__ pos, int E, const int* __restrict__ edges,
                     int* __restrict__ frag_pix, float* __restrict__ frag_attrs) {
    __shared__ int range[2];
    const long long f0 = static_cast<long long>(blockIdx.x) * kExpand;
    if (threadIdx.x == 0) {
        const long long f1 = dmin<long long>(f0 + kExpand, N) - 1;
        range[0] = upper_bound(prefix, 0, P, static_cast<int>(f0));
        range[1] = dmin(upper_bound(prefix, range[0], P, static_cast<int>(f1)) + 1, P);
    }
    __syncthreads();
    const int lo = range[0], hi = range[1];
#pragma unroll 1
    for (int j = 0; j < kExpandItems; ++j) {
        const long long f = f0 + j * kBlock + threadIdx.x;
        if (f >= N) break;
        const int fi = static_cast<int>(f);
        const int p = upper_bound(prefix, lo, hi, fi);
        if (p >= P) continue;
        const int base = (p == 0) ? 0 : prefix[p - 1];
        const int id = edge_ids[p];
        const int b = id / E;
        const int e = id % E;
        const float2 sl = reinterpret_cast<const float2*>(slopes)[p];
        const float4 sn = reinterpret_cast<const float4*>(spans)[p];
        DdaSpan sp{sn.x, sn.y, sn.z, sn.w, sl.x, sl.y};
        const int2 ev = reinterpret_cast<const int2*>(edges)[e];
        const long long vb = static_cast<long long>(b) * V;
        const float4 p0 = load_pos(pos, vb + ev.x);
        const float4 p1 = load_pos(pos, vb + ev.y);
        DdaFrag fr;
        if (!dda_frag(H, W, sp, fi - base, p0, p1, fr)) continue;
        frag_pix[f * 3 + 0] = b;
        frag_pix[f * 3 + 1] = fr.h;
        frag_pix[f * 3 + 2] = fr.w;
        reinterpret_cast<float4*>(frag_attrs)[f] =
            make_float4(sp.axis, fr.t1, fr.zw, static_cast<float>(e + 1));
    }
}

// ---------------------------------------------------- attribute interp

// dda.cu:258-284.
__global__ void __launch_bounds__(kBlock)
dda_interp_kernel(long long npix, const float* __restrict__ rast, int C,
                  const float* __restrict__ attr, const int* __restrict__ edges,
                  float* __restrict__ image) {
    const long long i = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= npix) return;
    const float4 r = reinterpret_cast<const float4*>(rast)[i];
    const float t1 = r.y;
    const float t0 = 1.f - t1;
    const int e = f2i(r.w) - 1;
    if (e < 0) return;
    const float* a0 = attr + static_cast<long long>(edges[e * 2 + 0]) * C;
    const float* a1 = attr + static_cast<long long>(edges[e * 2 + 1]) * C;
    for (int c = 0; c < C; ++c) image[i * C + c] = t0 * a0[c] + t1 * a1[c];
}

// dda.cu:300-327.
__global__ void __launch_bounds__(kBlock)
dda_interp_bwd_kernel(long long npix, const float* __restrict__ rast, int C,
                      const float* __restrict__ d_image, const int* __restrict__ edges,
                      float* __restrict__ d_attr) {
    const long long i = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= npix) return;
    const float4 r = reinterpret_cast<const float4*>(rast)[i];
    const float t1 = r.y;
    const float t0 = 1.f - t1;
    const int e = f2i(r.w) - 1;
    if (e < 0) return;
    float* d0 = d_attr + static_cast<long long>(edges[e * 2 + 0]) * C;
    float* d1 = d_attr + static_cast<long long>(edges[e * 2 + 1]) * C;
    for (int c = 0; c < C; ++c) {
        const float g = d_image[i * C + c];
        atomicAdd(&d0[c], g * t0);
        atomicAdd(&d1[c], g * t1);
    }
}

}  // namespace s2c

using namespace s2c;

extern "C" int s2c_dda_compute_span(int num_prims, const int* edge_ids, int H, int W, int V,
                                    const float* pos, int E, const int* edges,
                                    int* frag_prefix_sum, float* frag_slopes,
                                    float* frag_spans, int* num_frags_out, void* stream) {
    S2C_CHECK(num_frags_out != nullptr, "dda_compute_span: null count pointer");
    *num_frags_out = 0;
    S2C_CHECK(num_prims >= 0 && H >= 0 && W >= 0 && V >= 0 && E >= 0,
              "dda_compute_span: negative size");
    if (num_prims == 0) return S2C_OK;
    S2C_CHECK(E > 0, "dda_compute_span: E must be positive when primitives are given");
    hipStream_t s = static_cast<hipStream_t>(stream);
    dda_span_kernel<<<grid_for(num_prims, kBlock), kBlock, 0, s>>>(
        num_prims, edge_ids, H, W, V, pos, E, edges, frag_prefix_sum, frag_slopes, frag_spans);
    S2C_LAUNCH_CHECK();
    Temp scratch;
    if (int r = scratch.alloc(sizeof(int) * scan_scratch_ints(num_prims), s)) return r;
    if (int r = inclusive_scan_inplace(frag_prefix_sum, num_prims, scratch.as<int>(), s)) return r;
    return read_scan_total(scratch.as<int>(), num_prims, s, num_frags_out, "dda_compute_span");
}

extern "C" int s2c_dda_compute_fragments(int num_prims, int num_frags,
                                         const int* frag_prefix_sum, const int* edge_ids,
                                         const float* frag_slopes, const float* frag_spans,
                                         int H, int W, int V, const float* pos, int E,
                                         const int* edges, int* frag_pix, float* frag_attrs,
                                         void* stream) {
    S2C_CHECK(num_prims >= 0 && num_frags >= 0, "dda_compute_fragments: negative size");
    if (num_frags == 0 || num_prims == 0) return S2C_OK;
    S2C_CHECK(E > 0, "dda_compute_fragments: E must be positive");
    hipStream_t s = static_cast<hipStream_t>(stream);
    dda_fragments_kernel<<<grid_for(num_frags, kExpand), kBlock, 0, s>>>(
        num_prims, num_frags, frag_prefix_sum, edge_ids, frag_slopes, frag_spans, H, W, V, pos,
        E, edges, frag_pix, frag_attrs);
    S2C_LAUNCH_CHECK();
    return S2C_OK;
}

extern "C" int s2c_dda_interpolate_attributes(int B, int H, int W, int C, const float* rast_dda,
                                              const float* attr, const int* edges,
                                              float* image, void* stream) {
    S2C_CHECK(B >= 0 && H >= 0 && W >= 0 && C >= 0, "dda_interpolate_attributes: negative size");
    const long long n = static_cast<long long>(B) * H * W;
    if (n == 0 || C == 0) return S2C_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    dda_interp_kernel<<<grid_for(n, kBlock), kBlock, 0, s>>>(n, rast_dda, C, attr, edges, image);
    S2C_LAUNCH_CHECK();
    return S2C_OK;
}

extern "C" int s2c_backward_dda_interpolate_attributes(int B, int H, int W,
                                                       const float* rast_dda, int C,
                                                       const float* d_image, const int* edges,
                                                       float* d_attr, void* stream) {
    S2C_CHECK(B >= 0 && H >= 0 && W >= 0 && C >= 0,
              "backward_dda_interpolate_attributes: negative size");
    const long long n = static_cast<long long>(B) * H * W;
    if (n == 0 || C == 0) return S2C_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    dda_interp_bwd_kernel<<<grid_for(n, kBlock), kBlock, 0, s>>>(n, rast_dda, C, d_image, edges,
                                                                 d_attr);
    S2C_LAUNCH_CHECK();
    return S2C_OK;
}
