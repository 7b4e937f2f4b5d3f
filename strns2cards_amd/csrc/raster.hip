// Triangle rasterization, depth test, barycentric interpolation and cluster
// masks (reference: pystrns2cards/src/rasterize.cu), plus the fused
// triangle -> cluster-bitset kernel used by card fitting.
#include <mutex>

#include "raster_math.hpp"

namespace s2c {

__global__ void __launch_bounds__(kBlock)
tri_rects_kernel(int H, int W, int B, int V, const float* __restrict__ pos, int T,
                 const int* __restrict__ tri, int* __restrict__ rects, int* __restrict__ counts) {
    const long long i = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= static_cast<long long>(B) * T) return;
    const int b = static_cast<int>(i / T);
    const int t = static_cast<int>(i - static_cast<long long>(b) * T);
    const long long vb = static_cast<long long>(b) * V;
    const float4 P0 = ld4(pos, vb + tri[t * 3 + 0]);
    const float4 P1 = ld4(pos, vb + tri[t * 3 + 1]);
    const float4 P2 = ld4(pos, vb + tri[t * 3 + 2]);
    int4 r;
    counts[i] = tri_rect(H, W, P0, P1, P2, r);
    reinterpret_cast<int4*>(rects)[i] = r;
}

constexpr int kExpandItems = 4;
constexpr int kExpand = kBlock * kExpandItems;

__device__ __forceinline__ int upper_bound_i(const int* prefix, int lo, int hi, int key) {
    while (lo < hi) {
        const int mid = lo + ((hi - lo) >> 1);
        if (prefix[mid] <= key) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// rasterize.cu:90-172, with the per-workgroup primitive range found once
// (see dda_fragments_kernel).
__global__ void __launch_bounds__(kBlock)
tri_fragments_kernel(int H, int W, int V, const float* __restrict__ pos, int T,
                     const int* __restrict__ tri, int num_tris, int N,
                     const int* __restrict__ prefix, const int* __restrict__ rects,
                     int* __restrict__ frag_pix, float* __restrict__ frag_attrs) {
    __shared__ int range[2];
    const long long f0 = static_cast<long long>(blockIdx.x) * kExpand;
    if (threadIdx.x == 0) {
        const long long f1 = dmin<long long>(f0 + kExpand, N) - 1;
        range[0] = upper_bound_i(prefix, 0, num_tris, static_cast<int>(f0));
        range[1] = dmin(upper_bound_i(prefix, range[0], num_tris, static_cast<int>(f1)) + 1,
                        num_tris);
    }
    __syncthreads();
    const int lo = range[0], hi = range[1];
#pragma unroll 1
    for (int j = 0; j < kExpandItems; ++j) {
        const long long f = f0 + j * kBlock + threadIdx.x;
        if (f >= N) break;
        const int fi = static_cast<int>(f);
        const int ti = upper_bound_i(prefix, lo, hi, fi);
        if (ti >= num_tris) continue;
        const int4 r = reinterpret_cast<const int4*>(rects)[ti];
        if (r.w <= 0 || r.y <= 0) continue;
        const int b = ti / T;
        const int t = ti % T;
        const int l = fi - (ti == 0 ? 0 : prefix[ti - 1]);
        const int y = r.x + l / r.w;
        const int x = r.z + l % r.w;
        const long long vb = static_cast<long long>(b) * V;
        const float4 P0 = ld4(pos, vb + tri[t * 3 + 0]);
        const float4 P1 = ld4(pos, vb + tri[t * 3 + 1]);
        const float4 P2 = ld4(pos, vb + tri[t * 3 + 2]);
        float b0, b1, z;
        if (!tri_cover(H, W, x, y, P0, P1, P2, b0, b1, z)) continue;
        frag_pix[f * 3 + 0] = b;
        frag_pix[f * 3 + 1] = y;
        frag_pix[f * 3 + 2] = x;
        reinterpret_cast<float4*>(frag_attrs)[f] = make_float4(b0, b1, z, static_cast<float>(t + 1));
    }
}

// ------------------------------------------------------------ depth test
// rasterize.cu:197-251.

__global__ void __launch_bounds__(kBlock)
depth_key_kernel(int N, const int* __restrict__ frag_pix, const float* __restrict__ frag_attrs,
                 int H, int W, long long* __restrict__ keys) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= N) return;
    const int b = frag_pix[i * 3 + 0];
    const int h = frag_pix[i * 3 + 1];
    const int w = frag_pix[i * 3 + 2];
    const float zz = frag_attrs[static_cast<long long>(i) * 4 + 2] + 2.f;
    const unsigned long long k =
        (static_cast<unsigned long long>(__float_as_uint(zz)) << 32) | static_cast<unsigned>(i);
    atomicMin(&keys[(static_cast<long long>(b) * H + h) * W + w], static_cast<long long>(k));
}

__global__ void __launch_bounds__(kBlock)
depth_gather_kernel(long long npix, const long long* __restrict__ keys,
                    const float* __restrict__ frag_attrs, float* __restrict__ rast) {
    const long long i = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= npix) return;
    const int f = static_cast<int>(static_cast<unsigned>(keys[i] & 0xFFFFFFFFll));
    if (f < 0) return;
    reinterpret_cast<float4*>(rast)[i] = reinterpret_cast<const float4*>(frag_attrs)[f];
}

// --------------------------------------------------- barycentric interp

// rasterize.cu:367-400.
__global__ void __launch_bounds__(kBlock)
tri_interp_kernel(long long npix, const float* __restrict__ rast, int C,
                  const float* __restrict__ attr, const int* __restrict__ tri,
                  float* __restrict__ image) {
    const long long i = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= npix) return;
    const float4 r = reinterpret_cast<const float4*>(rast)[i];
    const float b0 = r.x, b1 = r.y;
    const float b2 = 1.f - b0 - b1;
    const int t = f2i(r.w) - 1;
    if (t < 0) return;
    const float* a0 = attr + static_cast<long long>(tri[t * 3 + 0]) * C;
    const float* a1 = attr + static_cast<long long>(tri[t * 3 + 1]) * C;
    const float* a2 = attr + static_cast<long long>(tri[t * 3 + 2]) * C;
    for (int c = 0; c < C; ++c) image[i * C + c] = b0 * a0[c] + b1 * a1[c] + b2 * a2[c];
}

// rasterize.cu:402-438.
__global__ void __launch_bounds__(kBlock)
tri_interp_bwd_kernel(long long npix, const float* __restrict__ rast, int C,
                      const float* __restrict__ d_image, const int* __restrict__ tri,
                      float* __restrict__ d_attr) {
    const long long i = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= npix) return;
    const float4 r = reinterpret_cast<const float4*>(rast)[i];
    const float b0 = r.x, b1 = r.y;
    const float b2 = 1.f - b0 - b1;
    const int t = f2i(r.w) - 1;
    if (t < 0) return;
    float* g0 = d_attr + static_cast<long long>(tri[t * 3 + 0]) * C;
    float* g1 = d_attr + static_cast<long long>(tri[t * 3 + 1]) * C;
    float* g2 = d_attr + static_cast<long long>(tri[t * 3 + 2]) * C;
    for (int c = 0; c < C; ++c) {
        const float g = d_image[i * C + c];
        atomicAdd(&g0[c], g * b0);
        atomicAdd(&g1[c], g * b1);
        atomicAdd(&g2[c], g * b2);
    }
}

// ------------------------------------------------- cluster mask (list)

// rasterize.cu:474-498.
__global__ void __launch_bounds__(kBlock)
cluster_mask_kernel(int N, const int* __restrict__ frag_pix, const float* __restrict__ frag_attrs,
                    const int* __restrict__ prim2cluster, int H, int W, int S,
                    int* __restrict__ bitset) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= N) return;
    const int prim = f2i(frag_attrs[static_cast<long long>(i) * 4 + 3]) - 1;
    if (prim < 0) return;
    const int k = prim2cluster[prim];
    const long long idx =
        ((static_cast<long long>(frag_pix[i * 3 + 0]) * H + frag_pix[i * 3 + 1]) * W +
         frag_pix[i * 3 + 2]) * S + k / 32;
    atomicOr(&bitset[idx], static_cast<int>(1u << (k % 32)));
}

// --------------------------------------------- fused triangle -> bitset
// Workgroup = kBlock consecutive (b, t) pairs; their bounding rectangles are
// block-scanned in LDS and every thread walks the block's candidate pixels.

struct TriPrim {
    float4 P0, P1, P2;
    int4 r;
    int b, cluster;
};

// kMaskTris triangles per workgroup of kBlock threads: a wave sets them up
// and scans their candidate counts, then all threads share the candidates.
// Few triangles per workgroup keep many workgroups on each CU and bound the
// tail of the largest triangles.
constexpr int kMaskTris = 32;

__global__ void __launch_bounds__(kBlock)
tri_mask_kernel(long long BT, int H, int W, int V, const float* __restrict__ pos, int T,
                const int* __restrict__ tri, const int* __restrict__ tri2cluster, int S,
                int* __restrict__ bitset) {
    __shared__ TriPrim prim[kMaskTris];
    __shared__ int excl[kMaskTris];
    __shared__ int sum;
    const long long q = static_cast<long long>(blockIdx.x) * kMaskTris + threadIdx.x;
    int cnt = 0;
    if (threadIdx.x < kMaskTris && q < BT) {
        const int b = static_cast<int>(q / T);
        const int t = static_cast<int>(q - static_cast<long long>(b) * T);
        const long long vb = static_cast<long long>(b) * V;
        TriPrim m;
        m.P0 = ld4(pos, vb + tri[t * 3 + 0]);
        m.P1 = ld4(pos, vb + tri[t * 3 + 1]);
        m.P2 = ld4(pos, vb + tri[t * 3 + 2]);
        cnt = tri_rect(H, W, m.P0, m.P1, m.P2, m.r);
        m.b = b;
        m.cluster = cnt > 0 ? tri2cluster[t] : 0;
        prim[threadIdx.x] = m;
    }
    if (threadIdx.x < kWave) {  // wave 0 (lanes past kMaskTris count 0)
        const int incl = wave_incl_scan(cnt);
        if (threadIdx.x < kMaskTris) excl[threadIdx.x] = incl - cnt;
        if (threadIdx.x == kWave - 1) sum = incl;
    }
    __syncthreads();
    const int total = sum;
#pragma unroll 1
    for (int j = threadIdx.x; j < total; j += kBlock) {
        int lo = 0, hi = kMaskTris;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (excl[mid] <= j) lo = mid;
            else hi = mid;
        }
        const TriPrim& m = prim[lo];
        const int l = j - excl[lo];
        const int y = m.r.x + l / m.r.w;
        const int x = m.r.z + l % m.r.w;
        float b0, b1, z;
        if (!tri_cover(H, W, x, y, m.P0, m.P1, m.P2, b0, b1, z)) continue;
        const long long idx =
            ((static_cast<long long>(m.b) * H + y) * W + x) * S + (m.cluster >> 5);
        atomicOr(&bitset[idx], static_cast<int>(1u << (m.cluster & 31)));
    }
}

}  // namespace s2c

using namespace s2c;

extern "C" int s2c_compute_triangle_rects(int H, int W, int B, int V, const float* pos, int T,
                                          const int* tri, int* triangle_rects,
                                          int* frag_prefix_sum, int* num_frags_out,
                                          void* stream) {
    S2C_CHECK(num_frags_out != nullptr, "compute_triangle_rects: null count pointer");
    *num_frags_out = 0;
    S2C_CHECK(H >= 0 && W >= 0 && B >= 0 && V >= 0 && T >= 0,
              "compute_triangle_rects: negative size");
    const long long n = static_cast<long long>(B) * T;
    S2C_CHECK(n <= INT_MAX, "compute_triangle_rects: B*T exceeds INT32_MAX");
    if (n == 0) return S2C_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    tri_rects_kernel<<<grid_for(n, kBlock), kBlock, 0, s>>>(H, W, B, V, pos, T, tri,
                                                            triangle_rects, frag_prefix_sum);
    S2C_LAUNCH_CHECK();
    Temp sums;
    if (int r = sums.alloc(sizeof(int) * scan_scratch_ints(n), s)) return r;
    if (int r = inclusive_scan_inplace(frag_prefix_sum, n, sums.as<int>(), s)) return r;
    return read_scan_total(sums.as<int>(), n, s, num_frags_out, "compute_triangle_rects");
}

extern "C" int s2c_compute_fragments(int H, int W, int V, const float* pos, int T,
                                     const int* tri, int num_tris, int num_frags,
                                     const int* frag_prefix_sum, const int* triangle_rects,
                                     int* frag_pix, float* frag_attrs, void* stream) {
    S2C_CHECK(num_tris >= 0 && num_frags >= 0 && T >= 0, "compute_fragments: negative size");
    if (num_frags == 0 || num_tris == 0) return S2C_OK;
    S2C_CHECK(T > 0, "compute_fragments: T must be positive");
    hipStream_t s = static_cast<hipStream_t>(stream);
    tri_fragments_kernel<<<grid_for(num_frags, kExpand), kBlock, 0, s>>>(
        H, W, V, pos, T, tri, num_tris, num_frags, frag_prefix_sum, triangle_rects, frag_pix,
        frag_attrs);
    S2C_LAUNCH_CHECK();
    return S2C_OK;
}

namespace s2c {

// The depth test's three launches over a key buffer of npix int64: a fill
// kernel (no memset node, so the sequence can be captured in a graph), the
// per-fragment atomicMin of the keys, and the gather of the winners.
static int launch_depth_test(long long npix, int H, int W, int num_fragments,
                             const int* frag_pix, const float* frag_attrs, long long* keys,
                             float* rast_out, hipStream_t s) {
    fill_i64_kernel<<<grid_for(dmin<long long>(npix, 1ll << 22), kBlock), kBlock, 0, s>>>(
        keys, LLONG_MAX, npix);
    S2C_LAUNCH_CHECK();
    if (num_fragments > 0) {
        depth_key_kernel<<<grid_for(num_fragments, kBlock), kBlock, 0, s>>>(
            num_fragments, frag_pix, frag_attrs, H, W, keys);
        S2C_LAUNCH_CHECK();
    }
    depth_gather_kernel<<<grid_for(npix, kBlock), kBlock, 0, s>>>(npix, keys, frag_attrs,
                                                                   rast_out);
    S2C_LAUNCH_CHECK();
    return S2C_OK;
}

// Key buffer of s2c_depth_test_noalloc: one array per GPU, loaded with the
// code object and held until the process ends (32 MiB: 4,194,304 pixels,
// e.g. 4 views at 1024 x 1024 or 16 at 512 x 512).  Host code never uses
// this name as an address; it uses the device address that
// hipGetSymbolAddress returns for the current device.
constexpr long long kDepthScratchPixels = 1ll << 22;
static __device__ __align__(256) long long scratch[kDepthScratchPixels];

constexpr int kMaxDevices = 64;

// Device address of `scratch` on the current device, looked up on the first
// call on that device (with it current) and kept per device.
static int depth_scratch(long long** out) {
    static std::mutex mu;
    static long long* ptr[kMaxDevices] = {};
    int dev = 0;
    S2C_HIP(hipGetDevice(&dev));
    S2C_CHECK(dev >= 0 && dev < kMaxDevices, "depth_test_noalloc: device %d beyond %d", dev,
              kMaxDevices);
    std::lock_guard<std::mutex> held(mu);
    if (ptr[dev] == nullptr) {
        void* p = nullptr;
        S2C_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        S2C_CHECK(p != nullptr, "depth_test_noalloc: no device address for the key buffer");
        ptr[dev] = static_cast<long long*>(p);
    }
    *out = ptr[dev];
    return S2C_OK;
}

}  // namespace s2c

extern "C" int s2c_depth_test(int B, int H, int W, int num_fragments, const int* frag_pix,
                              const float* frag_attrs, float* rast_out, void* stream) {
    S2C_CHECK(B >= 0 && H >= 0 && W >= 0 && num_fragments >= 0, "depth_test: negative size");
    const long long npix = static_cast<long long>(B) * H * W;
    if (npix == 0) return S2C_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    Temp keys;
    if (int r = keys.alloc(sizeof(long long) * npix, s)) return r;
    return launch_depth_test(npix, H, W, num_fragments, frag_pix, frag_attrs,
                             keys.as<long long>(), rast_out, s);
}

extern "C" int s2c_depth_test_noalloc(int B, int H, int W, int num_fragments,
                                      const int* frag_pix, const float* frag_attrs,
                                      float* rast_out, void* stream) {
    S2C_CHECK(B >= 0 && H >= 0 && W >= 0 && num_fragments >= 0,
              "depth_test_noalloc: negative size");
    const long long npix = static_cast<long long>(B) * H * W;
    S2C_CHECK(npix <= kDepthScratchPixels,
              "depth_test_noalloc: B*H*W = %lld pixels exceeds the key buffer's %lld", npix,
              kDepthScratchPixels);
    if (npix == 0) return S2C_OK;
    long long* keys = nullptr;
    if (int r = depth_scratch(&keys)) return r;
    return launch_depth_test(npix, H, W, num_fragments, frag_pix, frag_attrs, keys, rast_out,
                             static_cast<hipStream_t>(stream));
}

extern "C" int s2c_interpolate_triangle_attributes(int B, int H, int W, int attr_dim,
                                                   const float* rast, const float* attr,
                                                   const int* tri, float* image, void* stream) {
    S2C_CHECK(B >= 0 && H >= 0 && W >= 0 && attr_dim >= 0,
              "interpolate_triangle_attributes: negative size");
    const long long n = static_cast<long long>(B) * H * W;
    if (n == 0 || attr_dim == 0) return S2C_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    tri_interp_kernel<<<grid_for(n, kBlock), kBlock, 0, s>>>(n, rast, attr_dim, attr, tri, image);
    S2C_LAUNCH_CHECK();
    return S2C_OK;
}

extern "C" int s2c_backward_interpolate_triangle_attributes(int B, int H, int W, int attr_dim,
                                                            const float* rast,
                                                            const float* d_image,
                                                            const int* tri, float* d_attr,
                                                            void* stream) {
    S2C_CHECK(B >= 0 && H >= 0 && W >= 0 && attr_dim >= 0,
              "backward_interpolate_triangle_attributes: negative size");
    const long long n = static_cast<long long>(B) * H * W;
    if (n == 0 || attr_dim == 0) return S2C_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    tri_interp_bwd_kernel<<<grid_for(n, kBlock), kBlock, 0, s>>>(n, rast, attr_dim, d_image, tri,
                                                                 d_attr);
    S2C_LAUNCH_CHECK();
    return S2C_OK;
}

// Largest view index frag_pix[i][0] (rasterize.py:197, B = max + 1): four
// rows (three 16-byte loads) per thread and step, a workgroup maximum and
// one atomicMax per workgroup (one per wave, all on one address, took 0.26
// ms for 16K waves).
constexpr int kMaxViewGrid = 1024;

__global__ void __launch_bounds__(kBlock)
max_frag_view_kernel(long long n, const int* __restrict__ frag_pix, int* __restrict__ out) {
    __shared__ int wm[kBlock / kWave];
    int m = INT_MIN;
    const long long quads = n / 4;
    const long long stride = static_cast<long long>(gridDim.x) * kBlock;
    const long long t0 = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
    const int4* p4 = reinterpret_cast<const int4*>(frag_pix);
    for (long long q = t0; q < quads; q += stride) {
        const int4 a = p4[q * 3], b = p4[q * 3 + 1], c = p4[q * 3 + 2];
        m = dmax(dmax(m, a.x), dmax(a.w, dmax(b.z, c.y)));
    }
    for (long long i = quads * 4 + t0; i < n; i += stride) m = dmax(m, frag_pix[i * 3]);
    m = wave_max(m);
    if (lane_id() == 0) wm[threadIdx.x / kWave] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < kBlock / kWave; ++k) m = dmax(m, wm[k]);
        atomicMax(out, m);
    }
}

extern "C" int s2c_cluster_mask_from_fragments(int num_frags, const int* frag_pix,
                                               const float* frag_attrs,
                                               const int* prim2cluster, int H, int W,
                                               int num_slots, int* bitset, void* stream) {
    S2C_CHECK(num_frags >= 0 && H >= 0 && W >= 0 && num_slots >= 0,
              "cluster_mask_from_fragments: negative size");
    if (num_frags == 0) return S2C_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    cluster_mask_kernel<<<grid_for(num_frags, kBlock), kBlock, 0, s>>>(
        num_frags, frag_pix, frag_attrs, prim2cluster, H, W, num_slots, bitset);
    S2C_LAUNCH_CHECK();
    return S2C_OK;
}

extern "C" int s2c_triangle_cluster_mask(int B, int H, int W, int V, const float* pos, int T,
                                         const int* tri, const int* tri2cluster, int num_slots,
                                         int* bitset, void* stream) {
    S2C_CHECK(B >= 0 && H >= 0 && W >= 0 && V >= 0 && T >= 0 && num_slots >= 0,
              "triangle_cluster_mask: negative size");
    hipStream_t s = static_cast<hipStream_t>(stream);
    const long long words = static_cast<long long>(B) * H * W * num_slots;
    if (int r = fill_u32(bitset, 0u, words, s)) return r;
    const long long BT = static_cast<long long>(B) * T;
    if (BT == 0 || words == 0) return S2C_OK;
    tri_mask_kernel<<<grid_for(BT, kMaskTris), kBlock, 0, s>>>(BT, H, W, V, pos, T, tri,
                                                               tri2cluster, num_slots, bitset);
    S2C_LAUNCH_CHECK();
    return S2C_OK;
}

extern "C" int s2c_max_frag_view(int num_frags, const int* frag_pix, int* max_out,
                                 void* stream) {
    S2C_CHECK(max_out != nullptr, "max_frag_view: null output pointer");
    S2C_CHECK(num_frags > 0, "max_frag_view: no fragments");
    S2C_CHECK((reinterpret_cast<uintptr_t>(frag_pix) & 15) == 0,
              "max_frag_view: frag_pix not 16-byte aligned");
    hipStream_t s = static_cast<hipStream_t>(stream);
    Temp t;
    if (int r = t.alloc(sizeof(int), s)) return r;
    if (int r = fill_u32(t.ptr, 0x80000000u, 1, s)) return r;
    const long long quads = (static_cast<long long>(num_frags) + 3) / 4;
    const long long grid = dmin((quads + kBlock - 1) / kBlock, static_cast<long long>(kMaxViewGrid));
    max_frag_view_kernel<<<static_cast<unsigned>(grid), kBlock, 0, s>>>(num_frags, frag_pix,
                                                                          t.as<int>());
    S2C_LAUNCH_CHECK();
    return read_device_int(t.as<int>(), s, max_out);
}
