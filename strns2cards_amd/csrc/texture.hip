// Software-MSAA texture patches (render_texture, texture_utils.py:106-288).
//
// The reference renders each wisp's tube mesh 64 times, once per jitter of
// an 8 x 8 sample grid, each a separate rasterize + interpolate call, and
// accumulates on the host side with torch ops.  Here the 64 jittered copies
// of the vertices are made in one kernel and rasterized as 64 views of one
// batched rasterize (the drop-in raster kernels; per-view results are those
// of separate calls), then one kernel resolves every pixel over its samples
// in the reference's order:
//   acc_rgb += rgb * valid, count += valid, alpha += w_s * valid   (s = 0..63)
//   rgb = acc_rgb / max(count, 1), (0, 1, 0) where count = 0, then 0.5 (rgb + 1)
// The alpha is blurred with a normalized 3-tap Gaussian (reflect borders),
// horizontally then vertically, and the patch is flipped vertically,
// clamped, quantized to bytes and pasted into the atlas.
#include "dda_math.hpp"
#include "raster_math.hpp"

namespace s2c {

// out[s, v] = (x + dx[s], y + dy[s], z, 1) for the (V, 3) vertices.
__global__ void __launch_bounds__(kBlock)
msaa_jitter_kernel(int V, const float* __restrict__ pos, int S, const float* __restrict__ dx,
                   const float* __restrict__ dy, float* __restrict__ out) {
    const long long i = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= static_cast<long long>(S) * V) return;
    const int s = static_cast<int>(i / V), v = static_cast<int>(i - static_cast<long long>(s) * V);
    const float* p = pos + 3ll * v;
    reinterpret_cast<float4*>(out)[i] = make_float4(p[0] + dx[s], p[1] + dy[s], p[2], 1.f);
}

// One thread per pixel: the samples' interpolated attribute (as
// rasterize.cu:367-400 interpolates) averaged over the covered samples and
// the weighted coverage.
__global__ void __launch_bounds__(kBlock)
msaa_resolve_kernel(int S, long long npix, const float* __restrict__ rast,
                    const float* __restrict__ attr, const int* __restrict__ tri,
                    const float* __restrict__ weights, float* __restrict__ rgb,
                    float* __restrict__ alpha) {
    const long long i = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= npix) return;
    float a0 = 0.f, a1 = 0.f, a2 = 0.f, cnt = 0.f, al = 0.f;
    for (int s = 0; s < S; ++s) {
        const float4 r = reinterpret_cast<const float4*>(rast)[static_cast<long long>(s) * npix + i];
        const bool valid = r.w > 0.f;
        float c0 = 0.f, c1 = 0.f, c2 = 0.f;
        const int t = f2i(r.w) - 1;
        if (t >= 0) {
            const float b0 = r.x, b1 = r.y, b2 = 1.f - b0 - b1;
            const float* p0 = attr + 3ll * tri[t * 3 + 0];
            const float* p1 = attr + 3ll * tri[t * 3 + 1];
            const float* p2 = attr + 3ll * tri[t * 3 + 2];
            c0 = b0 * p0[0] + b1 * p1[0] + b2 * p2[0];
            c1 = b0 * p0[1] + b1 * p1[1] + b2 * p2[1];
            c2 = b0 * p0[2] + b1 * p1[2] + b2 * p2[2];
        }
        const float v = valid ? 1.f : 0.f;
        a0 += c0 * v;
        a1 += c1 * v;
        a2 += c2 * v;
        cnt += v;
        al += weights[s] * v;
    }
    float r0 = 0.f, r1 = 1.f, r2 = 0.f;
    if (cnt > 0.f) {
        r0 = a0 / cnt;
        r1 = a1 / cnt;
        r2 = a2 / cnt;
    }
    rgb[3 * i + 0] = 0.5f * (r0 + 1.f);
    rgb[3 * i + 1] = 0.5f * (r1 + 1.f);
    rgb[3 * i + 2] = 0.5f * (r2 + 1.f);
    alpha[i] = al;
}

// ---------------------------------------------------------------- fused form
// The same patch without fragment lists.  At one pixel a triangle yields at
// most one fragment, and the drop-in depth test orders fragments of one
// view by their stable index, i.e. by triangle index at a given pixel: so
// the minimum of (bits(z + 2) << 32 | t) over the covering triangles
// (signed, as rasterize.cu:198-228) picks the same fragment, and its
// barycentrics are recomputed by the same coverage test.

struct MsaaPrim {
    float4 P0, P1, P2;
    int4 r;
    int b, t;
};

// Depth keys (S, H, W) of the S jittered views (pos (S, V, 4)), filled with
// INT64_MAX by the caller: one lane per (view, triangle) counts its
// candidates, a block scan spreads them over the workgroup, and each
// covered candidate takes an atomic minimum.
__global__ void __launch_bounds__(kBlock)
msaa_depth_kernel(long long BT, int H, int W, int V, const float* __restrict__ pos, int T,
                  const int* __restrict__ tri, long long* __restrict__ keys) {
    __shared__ MsaaPrim prim[kBlock];
    __shared__ int excl[kBlock];
    __shared__ int sh[kBlock / kWave + 1];
    const long long q = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
    int cnt = 0;
    if (q < BT) {
        const int b = static_cast<int>(q / T);
        const int t = static_cast<int>(q - static_cast<long long>(b) * T);
        const long long vb = static_cast<long long>(b) * V;
        MsaaPrim m;
        m.P0 = ld4(pos, vb + tri[t * 3 + 0]);
        m.P1 = ld4(pos, vb + tri[t * 3 + 1]);
        m.P2 = ld4(pos, vb + tri[t * 3 + 2]);
        cnt = tri_rect(H, W, m.P0, m.P1, m.P2, m.r);
        m.b = b;
        m.t = t;
        prim[threadIdx.x] = m;
    }
    int total;
    excl[threadIdx.x] = block_excl_scan(cnt, sh, &total);
    __syncthreads();
#pragma unroll 1
    for (int j = threadIdx.x; j < total; j += kBlock) {
        int lo = 0, hi = kBlock;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (excl[mid] <= j) lo = mid;
            else hi = mid;
        }
        const MsaaPrim& m = prim[lo];
        const int l = j - excl[lo];
        const int y = m.r.x + l / m.r.w;
        const int x = m.r.z + l % m.r.w;
        float b0, b1, z;
        if (!tri_cover(H, W, x, y, m.P0, m.P1, m.P2, b0, b1, z)) continue;
        const unsigned long long k =
            (static_cast<unsigned long long>(__float_as_uint(z + 2.f)) << 32) |
            static_cast<unsigned>(m.t);
        atomicMin(&keys[(static_cast<long long>(m.b) * H + y) * W + x], static_cast<long long>(k));
    }
}

// The resolve of msaa_resolve_kernel from depth keys: the winning
// triangle's barycentrics recomputed at the pixel in its view.
__global__ void __launch_bounds__(kBlock)
msaa_resolve_keys_kernel(int S, int H, int W, int V, const float* __restrict__ pos,
                         const long long* __restrict__ keys, const float* __restrict__ attr,
                         const int* __restrict__ tri, const float* __restrict__ weights,
                         float* __restrict__ rgb, float* __restrict__ alpha) {
    const long long npix = static_cast<long long>(H) * W;
    const long long i = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= npix) return;
    const int y = static_cast<int>(i / W), x = static_cast<int>(i - static_cast<long long>(y) * W);
    float a0 = 0.f, a1 = 0.f, a2 = 0.f, cnt = 0.f, al = 0.f;
    for (int s = 0; s < S; ++s) {
        const int t = static_cast<int>(static_cast<unsigned>(keys[s * npix + i] & 0xFFFFFFFFll));
        float c0 = 0.f, c1 = 0.f, c2 = 0.f, v = 0.f;
        if (t >= 0) {
            const long long vb = static_cast<long long>(s) * V;
            const int i0 = tri[t * 3 + 0], i1 = tri[t * 3 + 1], i2 = tri[t * 3 + 2];
            float b0, b1, z;
            tri_cover(H, W, x, y, ld4(pos, vb + i0), ld4(pos, vb + i1), ld4(pos, vb + i2), b0, b1,
                      z);
            const float b2 = 1.f - b0 - b1;
            const float* p0 = attr + 3ll * i0;
            const float* p1 = attr + 3ll * i1;
            const float* p2 = attr + 3ll * i2;
            c0 = b0 * p0[0] + b1 * p1[0] + b2 * p2[0];
            c1 = b0 * p0[1] + b1 * p1[1] + b2 * p2[1];
            c2 = b0 * p0[2] + b1 * p1[2] + b2 * p2[2];
            v = 1.f;
        }
        a0 += c0 * v;
        a1 += c1 * v;
        a2 += c2 * v;
        cnt += v;
        al += weights[s] * v;
    }
    float r0 = 0.f, r1 = 1.f, r2 = 0.f;
    if (cnt > 0.f) {
        r0 = a0 / cnt;
        r1 = a1 / cnt;
        r2 = a2 / cnt;
    }
    rgb[3 * i + 0] = 0.5f * (r0 + 1.f);
    rgb[3 * i + 1] = 0.5f * (r1 + 1.f);
    rgb[3 * i + 2] = 0.5f * (r2 + 1.f);
    alpha[i] = al;
}

// --------------------------------------------------------------- patch masks
// render_patches_reduced (texture_utils.py:290-411): per wisp, the DDA
// fragments of its strands' polylines, the depth test and the
// interpolation of a constant 1 (1 - t1 + t1, which may round below 1),
// flipped and quantized.  All wisps in one launch, without fragment lists:
// a DDA walk visits distinct major-axis columns, so an edge has at most one
// fragment per pixel, the drop-in depth test's tie order at a pixel is the
// edge order, and the minimum of (bits(zw + 2) << 32 | edge) keeps its
// fragment; the resolve recomputes that edge's fragment at the pixel.
// Wisp g: rows[g] strands of cols[g] points from vertex voff[g] (clip
// float4, w = 1 as the reference builds them); edge e of the wisp joins
// point (e / (cols - 1), e % (cols - 1)) and the next one
// (build_edge_indices_from_wisp, __init__.py:165-175).

__device__ __forceinline__ void wisp_edge(int e, int N, int& v0) {
    const int n = e / (N - 1);
    v0 = n * N + (e - n * (N - 1));
}

__global__ void __launch_bounds__(kBlock)
patch_mask_depth_kernel(int H, int W, const int* __restrict__ rows, const int* __restrict__ cols,
                        const long long* __restrict__ voff, const float* __restrict__ pos,
                        long long* __restrict__ keys) {
    const int g = blockIdx.y;
    const int N = cols[g];
    if (N < 2) return;
    const int E = rows[g] * (N - 1);
    const int e = blockIdx.x * kBlock + threadIdx.x;
    if (e >= E) return;
    int v0;
    wisp_edge(e, N, v0);
    const float4 P0 = load_pos(pos, voff[g] + v0), P1 = load_pos(pos, voff[g] + v0 + 1);
    DdaSpan sp;
    bool wrote;
    const int cnt = dda_span(H, W, P0, P1, sp, wrote);
    long long* kg = keys + static_cast<long long>(g) * H * W;
    for (int k = 0; k < cnt; ++k) {
        DdaFrag f;
        if (!dda_frag(H, W, sp, k, P0, P1, f)) continue;
        const unsigned long long key =
            (static_cast<unsigned long long>(__float_as_uint(f.zw + 2.f)) << 32) |
            static_cast<unsigned>(e);
        atomicMin(&kg[static_cast<long long>(f.h) * W + f.w], static_cast<long long>(key));
    }
}

// out[g, H - 1 - h, w] = byte(clamp(t0 * 1 + t1 * 1)) of the winner, 0
// where no fragment.
__global__ void __launch_bounds__(kBlock)
patch_mask_resolve_kernel(int H, int W, const int* __restrict__ cols,
                          const long long* __restrict__ voff, const float* __restrict__ pos,
                          const long long* __restrict__ keys, unsigned char* __restrict__ out) {
    const int g = blockIdx.y;
    const long long npix = static_cast<long long>(H) * W;
    const long long i = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= npix) return;
    const int h = static_cast<int>(i / W), w = static_cast<int>(i - static_cast<long long>(h) * W);
    const int e = static_cast<int>(static_cast<unsigned>(keys[g * npix + i] & 0xFFFFFFFFll));
    float m = 0.f;
    if (e >= 0) {
        int v0;
        wisp_edge(e, cols[g], v0);
        const float4 P0 = load_pos(pos, voff[g] + v0), P1 = load_pos(pos, voff[g] + v0 + 1);
        DdaSpan sp;
        bool wrote;
        dda_span(H, W, P0, P1, sp, wrote);
        // the candidate on this pixel's major-axis column (k - 1 and k + 1
        // as well, should a clamped end of the walk round across a column)
        // (were two of them on the pixel, the depth test keeps the smaller
        // key, then the earlier candidate: so does this search)
        const int k = (sp.axis == 0.f ? w : h) - f2i(sp.axis == 0.f ? sp.sx : sp.sy);
        unsigned best = 0xFFFFFFFFu;
        for (int kk = k - 1; kk <= k + 1; ++kk) {
            DdaFrag f;
            if (kk >= 0 && dda_frag(H, W, sp, kk, P0, P1, f) && f.h == h && f.w == w) {
                const unsigned zb = __float_as_uint(f.zw + 2.f);
                if (best == 0xFFFFFFFFu || static_cast<int>(zb) < static_cast<int>(best)) {
                    best = zb;
                    const float t1 = f.t1, t0 = 1.f - t1;
                    m = t0 * 1.f + t1 * 1.f;
                }
            }
        }
    }
    const float c = fminf(fmaxf(m, 0.f), 1.f);
    out[g * npix + static_cast<long long>(H - 1 - h) * W + w] =
        static_cast<unsigned char>(static_cast<int>(c * 255.f));
}

__device__ __forceinline__ int reflect_index(int i, int n) {
    return i < 0 ? -i : (i >= n ? 2 * n - 2 - i : i);
}

// One 3-tap pass with reflect borders: along x (dir 0) or y (dir 1),
// out = (k0 in[-1] + k1 in[0]) + k2 in[+1].
__global__ void __launch_bounds__(kBlock)
blur3_kernel(int H, int W, int dir, float k0, float k1, float k2, const float* __restrict__ in,
             float* __restrict__ out) {
    const long long i = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= static_cast<long long>(H) * W) return;
    const int y = static_cast<int>(i / W), x = static_cast<int>(i - static_cast<long long>(y) * W);
    float a, b = in[i], c;
    if (dir == 0) {
        a = in[static_cast<long long>(y) * W + reflect_index(x - 1, W)];
        c = in[static_cast<long long>(y) * W + reflect_index(x + 1, W)];
    } else {
        a = in[static_cast<long long>(reflect_index(y - 1, H)) * W + x];
        c = in[static_cast<long long>(reflect_index(y + 1, H)) * W + x];
    }
    out[i] = (k0 * a + k1 * b) + k2 * c;
}

__device__ __forceinline__ unsigned char to_byte(float v) {
    const float c = fminf(fmaxf(v, 0.f), 1.f);  // NaN -> 0 (fmaxf returns the number)
    return static_cast<unsigned char>(static_cast<int>(c * 255.f));
}

// atlas[ay0 + r, ax0 + c] = flip(patch)[cy0 + r, cx0 + c] for the
// (ay1 - ay0) x (ax1 - ax0) atlas rectangle, rgb and alpha as bytes.
__global__ void __launch_bounds__(kBlock)
atlas_paste_kernel(int H, int W, const float* __restrict__ rgb, const float* __restrict__ alpha,
                   int cx0, int cy0, int ax0, int ay0, int aw, int ah, int R,
                   unsigned char* __restrict__ atlas_rgb, unsigned char* __restrict__ atlas_alpha) {
    const long long i = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= static_cast<long long>(aw) * ah) return;
    const int r = static_cast<int>(i / aw), c = static_cast<int>(i - static_cast<long long>(r) * aw);
    const int py = H - 1 - (cy0 + r), px = cx0 + c;  // flipped patch row
    const long long p = static_cast<long long>(py) * W + px;
    const long long a = static_cast<long long>(ay0 + r) * R + (ax0 + c);
    atlas_rgb[3 * a + 0] = to_byte(rgb[3 * p + 0]);
    atlas_rgb[3 * a + 1] = to_byte(rgb[3 * p + 1]);
    atlas_rgb[3 * a + 2] = to_byte(rgb[3 * p + 2]);
    atlas_alpha[a] = to_byte(alpha[p]);
}

}  // namespace s2c

using namespace s2c;

extern "C" int s2c_msaa_jitter(int V, const float* pos, int num_samples, const float* dx,
                               const float* dy, float* out, void* stream) {
    S2C_CHECK(V >= 0 && num_samples >= 0, "msaa_jitter: negative size");
    if (V == 0 || num_samples == 0) return S2C_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const long long n = static_cast<long long>(V) * num_samples;
    msaa_jitter_kernel<<<grid_for(n, kBlock), kBlock, 0, s>>>(V, pos, num_samples, dx, dy, out);
    S2C_LAUNCH_CHECK();
    return S2C_OK;
}

extern "C" int s2c_msaa_resolve(int num_samples, int H, int W, const float* rast,
                                const float* attr, const int* tri, const float* weights,
                                float* rgb, float* alpha, void* stream) {
    S2C_CHECK(num_samples >= 0 && H >= 0 && W >= 0, "msaa_resolve: negative size");
    const long long npix = static_cast<long long>(H) * W;
    if (npix == 0) return S2C_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    msaa_resolve_kernel<<<grid_for(npix, kBlock), kBlock, 0, s>>>(num_samples, npix, rast, attr,
                                                                  tri, weights, rgb, alpha);
    S2C_LAUNCH_CHECK();
    return S2C_OK;
}

extern "C" int s2c_msaa_patch(int V, int T, const float* jittered, const int* tri,
                              const float* attr, int num_samples, int H, int W,
                              const float* weights, long long* keys, float* rgb, float* alpha,
                              void* stream) {
    S2C_CHECK(V >= 0 && T >= 0 && num_samples >= 0 && H >= 0 && W >= 0,
              "msaa_patch: negative size");
    const long long npix = static_cast<long long>(H) * W;
    if (npix == 0) return S2C_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const long long nkeys = npix * num_samples;
    if (nkeys > 0) {
        fill_i64_kernel<<<grid_for(nkeys, kBlock), kBlock, 0, s>>>(keys, LLONG_MAX, nkeys);
        S2C_LAUNCH_CHECK();
    }
    const long long BT = static_cast<long long>(num_samples) * T;
    if (BT > 0) {
        msaa_depth_kernel<<<grid_for(BT, kBlock), kBlock, 0, s>>>(BT, H, W, V, jittered, T, tri,
                                                                  keys);
        S2C_LAUNCH_CHECK();
    }
    msaa_resolve_keys_kernel<<<grid_for(npix, kBlock), kBlock, 0, s>>>(
        num_samples, H, W, V, jittered, keys, attr, tri, weights, rgb, alpha);
    S2C_LAUNCH_CHECK();
    return S2C_OK;
}

extern "C" int s2c_patch_masks(int num_wisps, int H, int W, const int* rows, const int* cols,
                               const long long* voff, int max_edges, const float* pos,
                               long long* keys, unsigned char* out, void* stream) {
    S2C_CHECK(num_wisps >= 0 && H >= 0 && W >= 0 && max_edges >= 0, "patch_masks: negative size");
    S2C_CHECK(num_wisps <= 65535, "patch_masks: at most 65535 wisps per call");
    const long long npix = static_cast<long long>(H) * W;
    if (num_wisps == 0 || npix == 0) return S2C_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const long long nkeys = npix * num_wisps;
    fill_i64_kernel<<<grid_for(nkeys, kBlock), kBlock, 0, s>>>(keys, LLONG_MAX, nkeys);
    S2C_LAUNCH_CHECK();
    if (max_edges > 0) {
        patch_mask_depth_kernel<<<dim3(grid_for(max_edges, kBlock), num_wisps), kBlock, 0, s>>>(
            H, W, rows, cols, voff, pos, keys);
        S2C_LAUNCH_CHECK();
    }
    patch_mask_resolve_kernel<<<dim3(grid_for(npix, kBlock), num_wisps), kBlock, 0, s>>>(
        H, W, cols, voff, pos, keys, out);
    S2C_LAUNCH_CHECK();
    return S2C_OK;
}

extern "C" int s2c_blur3x3_reflect(int H, int W, const float* kernel3, const float* in,
                                   float* tmp, float* out, void* stream) {
    S2C_CHECK(H >= 2 && W >= 2, "blur3x3_reflect: reflect borders need H, W >= 2");
    S2C_CHECK(kernel3 != nullptr, "blur3x3_reflect: null kernel");
    hipStream_t s = static_cast<hipStream_t>(stream);
    const long long n = static_cast<long long>(H) * W;
    blur3_kernel<<<grid_for(n, kBlock), kBlock, 0, s>>>(H, W, 0, kernel3[0], kernel3[1],
                                                        kernel3[2], in, tmp);
    S2C_LAUNCH_CHECK();
    blur3_kernel<<<grid_for(n, kBlock), kBlock, 0, s>>>(H, W, 1, kernel3[0], kernel3[1],
                                                        kernel3[2], tmp, out);
    S2C_LAUNCH_CHECK();
    return S2C_OK;
}

extern "C" int s2c_atlas_paste(int H, int W, const float* rgb, const float* alpha, int crop_x0,
                               int crop_y0, int atlas_x0, int atlas_y0, int atlas_x1,
                               int atlas_y1, int R, unsigned char* atlas_rgb,
                               unsigned char* atlas_alpha, void* stream) {
    const int aw = atlas_x1 - atlas_x0, ah = atlas_y1 - atlas_y0;
    S2C_CHECK(aw >= 0 && ah >= 0, "atlas_paste: empty or inverted atlas rectangle");
    S2C_CHECK(atlas_x0 >= 0 && atlas_y0 >= 0 && atlas_x1 <= R && atlas_y1 <= R,
              "atlas_paste: rectangle outside the %d x %d atlas", R, R);
    S2C_CHECK(crop_x0 >= 0 && crop_y0 >= 0 && crop_x0 + aw <= W && crop_y0 + ah <= H,
              "atlas_paste: crop outside the %d x %d patch", H, W);
    if (aw == 0 || ah == 0) return S2C_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const long long n = static_cast<long long>(aw) * ah;
    atlas_paste_kernel<<<grid_for(n, kBlock), kBlock, 0, s>>>(
        H, W, rgb, alpha, crop_x0, crop_y0, atlas_x0, atlas_y0, aw, ah, R, atlas_rgb,
        atlas_alpha);
    S2C_LAUNCH_CHECK();
    return S2C_OK;
}
