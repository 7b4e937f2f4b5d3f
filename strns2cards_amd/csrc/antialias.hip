// Silhouette-edge marking and the antialiased cluster-bitset edge gradient
// (reference: pystrns2cards/src/antialias.cu).
#include <cmath>

#include "antialias_math.hpp"
#include "dda_math.hpp"

namespace s2c {

// antialias.cu:8-15.
__device__ __forceinline__ int third_vertex(const int* tri, int t, int v0, int v1) {
    const int i0 = tri[t * 3 + 0], i1 = tri[t * 3 + 1], i2 = tri[t * 3 + 2];
    return (i0 != v0 && i0 != v1) ? i0 : (i1 != v0 && i1 != v1) ? i1
         : (i2 != v0 && i2 != v1) ? i2 : -1;
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
    const float4* p = reinterpret_cast<const float4*>(pos + static_cast<long long>(b) * V * 4);
    const float fH = static_cast<float>(H), fW = static_cast<float>(W);
    const int v0 = edges[e * 2 + 0], v1 = edges[e * 2 + 1];
    const int t0 = edge2tri[e * 2 + 0], t1 = edge2tri[e * 2 + 1];
    // A triangle that does not contain the edge has no third vertex; the
    // reference would read pos[-1] there, here it counts as absent.
    const int a2 = t0 >= 0 ? third_vertex(tri, t0, v0, v1) : -1;
    const int a3 = t1 >= 0 ? third_vertex(tri, t1, v0, v1) : -1;
    const float4 none = make_float4(0.f, 0.f, 0.f, 1.f);
    bool sil;
    float2 n;
    if (!aa_edge(fW, fH, p[v0], p[v1], a2 >= 0, a2 >= 0 ? p[a2] : none, a3 >= 0,
                 a3 >= 0 ? p[a3] : none, sil, n))
        return;
    if (sil) prim_ids[idx] = static_cast<int>(idx);
    reinterpret_cast<float2*>(normals)[idx] = n;
}

__device__ __forceinline__ unsigned bit_u(const int* bs, long long pix, int S, int k) {
    return (static_cast<unsigned>(bs[pix * S + (k >> 5)]) >> (k & 31)) & 1u;
}

// Word planes of the two bitsets: planes[(b*S + j)*H*W + h*W + w] =
// (bitset[b,h,w,j], target[b,h,w,j]).  With S words per pixel a window
// read from the (B,H,W,S) layout takes one word of each S*4-byte pixel
// record, so a window row touches a line per pixel; in word planes the
// row is contiguous.  Only the pixels the samples read are copied (below).
// Below 16 words per pixel the records win: the transpose costs more than the
// planes save (measured on 13 views at 200^2 and 512^2 with a transpose of
// every pixel: records 4-10% faster at 7 and 10 words, planes 3-8% faster
// at 16 and 20, 25% at 32).
constexpr int kPlaneMinSlots = 16;

// Only the planes' pixels that some sample reads are needed: the window of
// a sample (floor(R) pixels around its point on the edge) and the pixel
// that gates it (the fragment's or its neighbour's), all within
// floor(R) + 2 pixels of the fragment.  The silhouette fragments mark the
// kPlaneTile x kPlaneTile tiles within that reach (plain stores of 1, so
// no atomics), and the transpose skips the pixels of unmarked tiles (a
// wave's staged box may cover such pixels; no sample reads them).
// Measured and not kept: marking the words of each tile as well (one bit
// per word, atomicOr): the marks cost 0.13-0.85 ms in atomics and the
// transpose, still reading whole records, gained 14 us.
constexpr int kPlaneTileLog2 = 4;
constexpr int kPlaneTile = 1 << kPlaneTileLog2;

struct PlaneTiles {
    int tx, ty;       // tiles per row / column of a view
    unsigned* mark;   // mark[(b * ty + y) * tx + x] != 0: transpose the tile
};

__device__ __forceinline__ void mark_plane_box(const PlaneTiles& t, int b, int h0, int h1, int w0,
                                               int w1, int H, int W) {
    h0 = dmax(h0, 0);
    w0 = dmax(w0, 0);
    h1 = dmin(h1, H - 1);
    w1 = dmin(w1, W - 1);
    if (h0 > h1 || w0 > w1) return;
    for (int y = h0 >> kPlaneTileLog2; y <= (h1 >> kPlaneTileLog2); ++y)
        for (int x = w0 >> kPlaneTileLog2; x <= (w1 >> kPlaneTileLog2); ++x)
            t.mark[(static_cast<long long>(b) * t.ty + y) * t.tx + x] = 1u;
}

// Marks from a fragment list (frag_pix rows (b, h, w)); fragments of views
// >= B are dropped by the backward and mark nothing.
__global__ void __launch_bounds__(kBlock)
plane_tiles_frags_kernel(int N, const int* __restrict__ frag_pix, int H, int W, int B, int reach,
                         PlaneTiles t) {
    const long long i = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= N) return;
    const int b = frag_pix[i * 3 + 0], h = frag_pix[i * 3 + 1], w = frag_pix[i * 3 + 2];
    if (b < 0 || b >= B) return;
    mark_plane_box(t, b, h - reach, h + reach, w - reach, w + reach, H, W);
}

// Screen coordinate 0.5 (1 + c / cw) * n of the sample formula (aa_sample),
// clamped into [-1, n] so that it converts to int safely.
__device__ __forceinline__ int screen_clamped(float c, float cw, float n, bool flip) {
    const float v = 0.5f * (1.f + (flip ? -c : c) / cw) * n;
    return f2i(floorf(fminf(fmaxf(v, -1.f), n)));
}

// Marks from the marked-edge array (b*E + e or -1): every sample of a
// marked edge lies on its screen segment, so the segment's bounding box
// grown by the reach covers its fragments' reads.  With w <= 0 at an end
// (or a non-finite one) the projected samples are not bounded by the ends'
// projections, and the whole view is marked.
__global__ void __launch_bounds__(kBlock)
plane_tiles_edges_kernel(long long BE, const int* __restrict__ prim_ids, int E,
                         const int* __restrict__ edges, int V, const float* __restrict__ pos,
                         int H, int W, int reach, PlaneTiles t) {
    const long long q = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
    if (q >= BE || prim_ids[q] < 0) return;
    const int b = static_cast<int>(q / E);
    const int e = static_cast<int>(q - static_cast<long long>(b) * E);
    const float4* p = reinterpret_cast<const float4*>(pos + static_cast<long long>(b) * V * 4);
    const float4 p0 = p[edges[e * 2 + 0]], p1 = p[edges[e * 2 + 1]];
    const bool finite = isfinite(p0.x) && isfinite(p0.y) && isfinite(p1.x) && isfinite(p1.y);
    if (!(p0.w > 0.f && p1.w > 0.f && finite)) {
        mark_plane_box(t, b, 0, H - 1, 0, W - 1, H, W);
        return;
    }
    const float fW = static_cast<float>(W), fH = static_cast<float>(H);
    const int xa = screen_clamped(p0.x, p0.w, fW, false), xb = screen_clamped(p1.x, p1.w, fW, false);
    const int ya = screen_clamped(p0.y, p0.w, fH, true), yb = screen_clamped(p1.y, p1.w, fH, true);
    mark_plane_box(t, b, dmin(ya, yb) - reach, dmax(ya, yb) + reach, dmin(xa, xb) - reach,
                   dmax(xa, xb) + reach, H, W);
}

// One workgroup per kPlanePix consecutive pixels: coalesced reads of the
// records of the pixels in marked tiles, coalesced plane-row writes.
constexpr int kPlanePix = 64;

__global__ void __launch_bounds__(kBlock)
residual_planes_kernel(long long NP, long long HW, int W, int S, const int* __restrict__ bitset,
                       const int* __restrict__ target, int2* __restrict__ planes,
                       PlaneTiles t) {
    __shared__ int2 tile[kPlanePix][33];
    __shared__ unsigned char take[kPlanePix];
    __shared__ int any;
    const long long p0 = static_cast<long long>(blockIdx.x) * kPlanePix;
    const int np = static_cast<int>(dmin(static_cast<long long>(kPlanePix), NP - p0));
    if (threadIdx.x == 0) any = 0;
    __syncthreads();
    if (threadIdx.x < np) {
        const long long p = p0 + threadIdx.x;
        const long long b = p / HW, hw = p - b * HW;
        const int h = static_cast<int>(hw / W), w = static_cast<int>(hw - static_cast<long long>(h) * W);
        const bool m = t.mark[(b * t.ty + (h >> kPlaneTileLog2)) * t.tx + (w >> kPlaneTileLog2)] != 0u;
        take[threadIdx.x] = m;
        if (m) any = 1;
    }
    __syncthreads();
    if (!any) return;  // workgroup-uniform
    for (int j0 = 0; j0 < S; j0 += 32) {
        const int nj = dmin(32, S - j0);
        for (int i = threadIdx.x; i < np * nj; i += kBlock) {
            const int pl = i / nj, j = i - pl * nj;
            if (!take[pl]) continue;
            const long long src = (p0 + pl) * S + j0 + j;
            tile[pl][j] = make_int2(bitset[src], target[src]);
        }
        __syncthreads();
        for (int i = threadIdx.x; i < np * nj; i += kBlock) {
            const int j = i / np, pl = i - j * np;
            if (!take[pl]) continue;
            const long long p = p0 + pl;
            const long long b = p / HW, hw = p - b * HW;
            planes[(b * S + j0 + j) * HW + hw] = tile[pl][j];
        }
        __syncthreads();
    }
}

// The marks' reach, floor(R) + 2 pixels, from the host: the whole image for
// a radius beyond it or NaN, 2 pixels for a negative one.
inline int plane_reach(float R, int H, int W) {
    const int cap = H + W + 2;
    if (!(R < static_cast<float>(cap))) return cap;  // also NaN
    return (R > 0.f ? static_cast<int>(std::floor(R)) : 0) + 2;
}

// Word planes and tile marks of one call in one temporary: the planes, then
// the marks at a 256-byte boundary.
struct PlaneBuffers {
    Temp mem;
    int2* planes = nullptr;
    PlaneTiles tiles{0, 0, nullptr};
    int alloc(int B, int H, int W, int S, hipStream_t s) {
        const size_t pb = sizeof(int2) * static_cast<size_t>(B) * H * W * S;
        const size_t po = (pb + 255) / 256 * 256;
        tiles.tx = (W + kPlaneTile - 1) >> kPlaneTileLog2;
        tiles.ty = (H + kPlaneTile - 1) >> kPlaneTileLog2;
        const long long nm = static_cast<long long>(B) * tiles.tx * tiles.ty;
        if (int r = mem.alloc(po + 4 * static_cast<size_t>(nm), s)) return r;
        planes = mem.as<int2>();
        tiles.mark = reinterpret_cast<unsigned*>(mem.as<char>() + po);
        return fill_u32(tiles.mark, 0u, nm, s);
    }
};

// Residual bits (bit 0 = rendered, bit 1 = target) of cluster k at a pixel
// of the plane starting at `base` (plane of word k / 32).
__device__ __forceinline__ unsigned plane_code(const int2* planes, long long base, int W, int h,
                                               int w, int k) {
    const int2 v = planes[base + w + static_cast<long long>(h) * W];
    const int kb = k & 31;
    return ((static_cast<unsigned>(v.x) >> kb) & 1u) |
           (((static_cast<unsigned>(v.y) >> kb) & 1u) << 1);
}

// Residual bits from the (B,H,W,S) bitsets or their word planes.
struct ResidualSrc {
    const int* bitset;
    const int* target;
    const int2* planes;  // word planes, or nullptr to read the records
    long long plane;     // first pixel of the view / first entry of the word plane
    int W, S, k;
    __device__ __forceinline__ unsigned operator()(int h, int w) const {
        if (planes) return plane_code(planes, plane, W, h, w, k);
        const long long pix = plane + w + static_cast<long long>(h) * W;
        return bit_u(bitset, pix, S, k) | (bit_u(target, pix, S, k) << 1);
    }
};

// A wave's staged box of residual bits: rows h0 .. h0 + rows - 1 and
// columns w0 .. w0 + wpr * 32 - 1, one bit per pixel in two planes
// (rendered, target) of `wpr` words per row.
struct BoxBits {
    const unsigned* render;
    const unsigned* target;
    int h0, w0, wpr;
    __device__ __forceinline__ unsigned operator()(int h, int w) const {  // as ResidualSrc
        const int i = (h - h0) * wpr * 32 + (w - w0);
        return ((render[i >> 5] >> (i & 31)) & 1u) | (((target[i >> 5] >> (i & 31)) & 1u) << 1);
    }
    // the window-row form (aa_adjoint_sum_rows)
    __device__ __forceinline__ void operator()(int h, int w, int n, unsigned& diff,
                                               unsigned& tgt) const {
        const int r = (h - h0) * wpr;
        tgt = bits_at(target + r, w - w0, n);
        diff = bits_at(render + r, w - w0, n) ^ tgt;
    }
};

// Adds (x, y, w) to components 0, 1, 3 of d_pos[key] for every live lane
// of the wave, with one atomic per run of lanes that share a key: the
// wave's consecutive samples lie on a few edges, so most lanes would
// otherwise add into the same two vertices.  A run's values are summed by
// a segmented wave scan; the run's last lane issues the atomics.  Every
// lane of the wave must call it.
__device__ __forceinline__ void wave_add_runs(float* d_pos, bool live, long long key, float x,
                                              float y, float w) {
    const int lane = lane_id();
    if (!live) key = -1 - lane;  // a run of its own, never added
    const long long prev = __shfl_up(key, 1, kWave);
    int start = (lane == 0 || prev != key) ? lane : 0;  // first lane of the run
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const int t = __shfl_up(start, d, kWave);
        if (lane >= d) start = dmax(start, t);
    }
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const float tx = __shfl_up(x, d, kWave), ty = __shfl_up(y, d, kWave),
                    tw = __shfl_up(w, d, kWave);
        if (lane - d >= start) {
            x += tx;
            y += ty;
            w += tw;
        }
    }
    const long long next = __shfl_down(key, 1, kWave);
    if (live && (lane == kWave - 1 || next != key)) {
        float* dp = d_pos + key * 4;
        atomicAdd(&dp[0], x);
        atomicAdd(&dp[1], y);
        atomicAdd(&dp[3], w);
    }
}

// antialias.cu:179-279.  A wave takes 64 consecutive fragments.  When all of
// its live samples share one view and cluster (consecutive silhouette
// fragments of one card) and the union of their windows fits the wave's
// kAaBoxWords, the wave stages that box's residual bits in LDS once, as
// bit rows (ballots of 64 pixels), and each sample reads its window a row
// at a time (aa_adjoint_sum_rows: the same terms in the same order);
// otherwise every window pixel is read from the bitsets directly.  PLANES:
// the bits come from the word planes of B views (residual_planes_kernel);
// fragments of views >= B are dropped.
constexpr int kAaWaves = kBlock / kWave;
// Words per wave for the staged box's two bit planes (4 KB): up to 512
// words of rows, e.g. 15 rows of 34 words (1,088 pixels).  The box was a
// byte per pixel (8 KB for 8,192 pixels), read a pixel at a time.
constexpr int kAaBoxWords = 1024;

// Orders a wave's own LDS accesses: the LDS executes one wave's operations
// in issue order, so a buffer private to the wave needs only a compiler
// fence at wavefront scope between its writes and reads.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One sample per lane, every lane of the wave calling.  `box` is the wave's
// own LDS box, so only the wave synchronizes (waves of a workgroup never
// wait for each other).  `has`: the lane holds a fragment (b, h, w) of edge
// e with the DDA axis and t1 of frag_attrs_dda (dda.cu:218-224).
template <bool PLANES>
__device__ __forceinline__ void aa_wave_fragment(
    bool has, int b, int h, int w, float axis, float t1, int e, int H, int W, int S,
    const int* __restrict__ bitset, const int* __restrict__ target, int V,
    const float* __restrict__ pos, int E, const int* __restrict__ edges,
    const float* __restrict__ normals, const int* __restrict__ edge2cluster,
    float* __restrict__ d_pos, float R, float rho, int B, const int2* __restrict__ planes,
    unsigned* box) {
    // ---- per-sample setup, as the reference (antialias.cu:198-256)
    bool live = false;
    int k = 0, v0 = 0, v1 = 0;
    float2 n = make_float2(0.f, 0.f);
    AaSample smp{};
    if (has && e >= 0) {
        const float4* p = reinterpret_cast<const float4*>(pos + static_cast<long long>(b) * V * 4);
        v0 = edges[e * 2 + 0];
        v1 = edges[e * 2 + 1];
        n = reinterpret_cast<const float2*>(normals + static_cast<long long>(b) * E * 2)[e];
        smp = aa_sample(H, W, h, w, axis, t1, p[v0], p[v1], n);
        // (h, w) outside the image only arises from a malformed fragment
        // list; such samples are dropped instead of read out of bounds.
        live = smp.live && h >= 0 && h < H && w >= 0 && w < W && b >= 0 && (!PLANES || b < B);
        if (live) k = edge2cluster[e];
    }
    const int ch = live ? f2i(floorf(smp.fy)) : 0, cw = live ? f2i(floorf(smp.fx)) : 0;
    const int Ri = f2i(floorf(R));

    // ---- wave-uniform staging decision
    const unsigned long long lm = __ballot(live);
    int bh0 = 0, bw0 = 0, wpr = 0;
    bool staged = false;
    if (lm) {
        const int first = __ffsll(static_cast<long long>(lm)) - 1;
        const int b_ref = __shfl(b, first, kWave), k_ref = __shfl(k, first, kWave);
        const bool same = !live || (b == b_ref && k == k_ref);
        if (__ballot(!same) == 0) {
            const int lo_h = live ? dmin(dmax(ch - Ri, 0), smp.ph) : INT_MAX;
            const int hi_h = live ? dmax(dmin(ch + Ri, H - 1), smp.ph) : INT_MIN;
            const int lo_w = live ? dmin(dmax(cw - Ri, 0), smp.pw_) : INT_MAX;
            const int hi_w = live ? dmax(dmin(cw + Ri, W - 1), smp.pw_) : INT_MIN;
            bh0 = wave_min(lo_h);
            bw0 = wave_min(lo_w);
            const long long hh = static_cast<long long>(wave_max(hi_h)) - bh0 + 1;
            const long long ww = static_cast<long long>(wave_max(hi_w)) - bw0 + 1;
            const long long words = hh * ((ww + 31) / 32);
            if (hh > 0 && ww > 0 && 2 * words <= kAaBoxWords) {
                staged = true;
                wpr = static_cast<int>((ww + 31) / 32);
                const int lane = lane_id();
                const long long plane = static_cast<long long>(H) * W * b_ref;
                const long long pbase = (static_cast<long long>(b_ref) * S + (k_ref >> 5)) *
                                        (static_cast<long long>(H) * W);
                unsigned* rbits = box;
                unsigned* tbits = box + kAaBoxWords / 2;
                // 64 pixels of a row per step: ballots of the two bits,
                // lanes 0 and 1 store the step's two words of each plane
                for (int y = 0; y < static_cast<int>(hh); ++y) {
                    for (int c0 = 0; c0 < wpr * 32; c0 += kWave) {
                        const int x = bw0 + c0 + lane;
                        unsigned code = 0u;
                        if (c0 + lane < ww) {  // (bh0 + y, x) in [0,H) x [0,W)
                            if (PLANES) {
                                code = plane_code(planes, pbase, W, bh0 + y, x, k_ref);
                            } else {
                                const long long pix =
                                    plane + x + static_cast<long long>(bh0 + y) * W;
                                code = bit_u(bitset, pix, S, k_ref) |
                                       (bit_u(target, pix, S, k_ref) << 1);
                            }
                        }
                        const unsigned long long rb = __ballot(code & 1u);
                        const unsigned long long tb = __ballot(code & 2u);
                        const int wi = (c0 >> 5) + lane;
                        if (lane < 2 && wi < wpr) {
                            rbits[y * wpr + wi] = static_cast<unsigned>(lane ? rb >> 32 : rb);
                            tbits[y * wpr + wi] = static_cast<unsigned>(lane ? tb >> 32 : tb);
                        }
                    }
                }
            }
        }
    }
    wave_lds_sync();

    // the sample's contributions to its two vertices (antialias.cu:272-277)
    bool adds = false;
    float dpx = 0.f, dpy = 0.f, dpw = 0.f;
    if (live) {
        if (staged) {
            const BoxBits bx{box, box + kAaBoxWords / 2, bh0, bw0, wpr};
            adds = aa_sample_grad_rows<false>(smp, H, W, R, rho, n, bx, bx, dpx, dpy, dpw);
        } else {
            const long long hw = static_cast<long long>(H) * W;
            const ResidualSrc src{bitset, target, PLANES ? planes : nullptr,
                                  PLANES ? (static_cast<long long>(b) * S + (k >> 5)) * hw : hw * b,
                                  W, S, k};
            adds = aa_sample_grad(smp, H, W, R, rho, n, src, dpx, dpy, dpw);
        }
    }
    const long long vb = static_cast<long long>(b) * V;
    wave_add_runs(d_pos, adds, vb + v0, dpx * smp.t0, dpy * smp.t0, dpw * smp.t0);
    wave_add_runs(d_pos, adds, vb + v1, dpx * smp.t1, dpy * smp.t1, dpw * smp.t1);
    wave_lds_sync();  // the box is reused by the wave's next sample
}

template <bool PLANES>
__global__ void __launch_bounds__(kBlock)
antialias_bwd_kernel(int N, const int* __restrict__ frag_pix, const float* __restrict__ fa,
                     int H, int W, int S, const int* __restrict__ bitset,
                     const int* __restrict__ target, int V, const float* __restrict__ pos, int E,
                     const int* __restrict__ edges, const float* __restrict__ normals,
                     const int* __restrict__ edge2cluster, float* __restrict__ d_pos, float R,
                     float rho, int B, const int2* __restrict__ planes) {
    __shared__ unsigned box[kAaWaves][kAaBoxWords];
    const int wv = threadIdx.x / kWave;
    const long long i = (static_cast<long long>(blockIdx.x) * kAaWaves + wv) * kWave + lane_id();
    int b = 0, h = 0, w = 0, e = -1;
    float axis = 0.f, t1 = 0.f;
    if (i < N) {
        b = frag_pix[i * 3 + 0];
        h = frag_pix[i * 3 + 1];
        w = frag_pix[i * 3 + 2];
        const float4 a = reinterpret_cast<const float4*>(fa)[i];
        e = f2i(a.w) - 1;
        axis = a.x;
        t1 = a.y;
    }
    aa_wave_fragment<PLANES>(i < N, b, h, w, axis, t1, e, H, W, S, bitset, target, V, pos,
                                    E, edges, normals, edge2cluster, d_pos, R, rho, B, planes,
                                    box[wv]);
}

// The same samples without a fragment list.  Each wave owns 64*M
// consecutive (view, edge) items of the marked-edge array (mark_edges_kernel:
// b*E + e, or -1), M consecutive ones per lane, and DDA-walks the marked ones
// itself: a wave scan of their candidate counts (dda_span, dda.cu:9-87)
// spreads the candidates over the lanes in (edge, candidate) order, as the
// stable fragment list orders them, and every candidate that
// dda_compute_fragments would keep (dda.cu:185-216) becomes a lane's sample.
// A lane finds its item by a binary search over the wave's exclusive prefix
// and a scan of the owning lane's M partial sums (shuffles), and recomputes
// that item's span.  No count leaves the device and no wave waits for
// another.  M > 1 fills the wave's last pass better where few edges are
// marked or their spans are short.
constexpr int kAaEdgeItems = 2;

template <bool PLANES, int M>
__global__ void __launch_bounds__(kBlock)
antialias_edges_bwd_kernel(long long BE, int H, int W, int S, const int* __restrict__ bitset,
                           const int* __restrict__ target, int V, const float* __restrict__ pos,
                           int E, const int* __restrict__ edges,
                           const int* __restrict__ prim_ids, const float* __restrict__ normals,
                           const int* __restrict__ edge2cluster, float* __restrict__ d_pos,
                           float R, float rho, int B, const int2* __restrict__ planes) {
    __shared__ unsigned box[kAaWaves][kAaBoxWords];
    const int wv = threadIdx.x / kWave;
    const int lane = lane_id();
    const long long q0 = (static_cast<long long>(blockIdx.x) * kAaWaves + wv) * kWave * M;
    int sub[M];  // sub[m]: candidates of the lane's items before item m
    int lsum = 0;
#pragma unroll
    for (int m = 0; m < M; ++m) {
        sub[m] = lsum;
        const long long q = q0 + static_cast<long long>(lane) * M + m;
        if (q < BE && prim_ids[q] >= 0) {
            const int b = static_cast<int>(q / E);
            const int e = static_cast<int>(q - static_cast<long long>(b) * E);
            const float4* p =
                reinterpret_cast<const float4*>(pos + static_cast<long long>(b) * V * 4);
            DdaSpan sp;
            bool wrote;
            lsum += dda_span(H, W, p[edges[e * 2 + 0]], p[edges[e * 2 + 1]], sp, wrote);
        }
    }
    const int incl = wave_incl_scan(lsum);
    const int excl = incl - lsum;
    const int total = __shfl(incl, kWave - 1, kWave);
#pragma unroll 1
    for (int base = 0; base < total; base += kWave) {  // wave-uniform trip count
        const int j = base + lane;
        // last lane with excl <= j; excl is non-decreasing and excl[0] = 0
        int lo = 0;
#pragma unroll
        for (int step = kWave / 2; step > 0; step >>= 1) {
            const int v = __shfl(excl, lo + step, kWave);
            if (v <= j) lo += step;
        }
        int r = j - __shfl(excl, lo, kWave);
        // last item of lane lo with sub <= r (it has candidates left)
        int m_sel = 0, off = 0;
#pragma unroll
        for (int m = 1; m < M; ++m) {
            const int v = __shfl(sub[m], lo, kWave);
            if (v <= r) {
                m_sel = m;
                off = v;
            }
        }
        const int k = r - off;
        bool has = false;
        int b = 0, h = 0, w = 0, e = -1;
        float axis = 0.f, t1 = 0.f;
        if (j < total) {
            // the item has more than k >= 0 candidates: a marked edge in [0, BE)
            const long long qi = q0 + static_cast<long long>(lo) * M + m_sel;
            const int bi = static_cast<int>(qi / E);
            const int ei = static_cast<int>(qi - static_cast<long long>(bi) * E);
            const float4* p =
                reinterpret_cast<const float4*>(pos + static_cast<long long>(bi) * V * 4);
            const float4 p0 = p[edges[ei * 2 + 0]], p1 = p[edges[ei * 2 + 1]];
            DdaSpan sp;
            bool wrote;
            dda_span(H, W, p0, p1, sp, wrote);
            DdaFrag f;
            if (dda_frag(H, W, sp, k, p0, p1, f)) {
                has = true;
                b = bi;
                h = f.h;
                w = f.w;
                e = ei;
                axis = sp.axis;
                t1 = f.t1;
            }
        }
        aa_wave_fragment<PLANES>(has, b, h, w, axis, t1, e, H, W, S, bitset, target, V, pos,
                                       E, edges, normals, edge2cluster, d_pos, R, rho, B, planes,
                                       box[wv]);
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
    return s2c_backward_antialiased_cluster_bitset_views(
        0, num_frags, frag_pix, frag_attrs_dda, H, W, num_slots, bitset, target_bitset, V, pos,
        E, edges, normals, edge2cluster, d_pos, kernel_radius, rho, stream);
}

extern "C" int s2c_backward_antialiased_cluster_bitset_edges(
    int num_views, int H, int W, int num_slots, const int* bitset, const int* target_bitset,
    int V, const float* pos, int E, const int* edges, const int* prim_ids, const float* normals,
    const int* edge2cluster, float* d_pos, float kernel_radius, float rho, void* stream) {
    S2C_CHECK(num_views >= 0 && H >= 0 && W >= 0 && num_slots >= 0 && V >= 0 && E >= 0,
              "backward_antialiased_cluster_bitset_edges: negative size");
    const long long BE = static_cast<long long>(num_views) * E;
    S2C_CHECK(BE <= INT_MAX, "backward_antialiased_cluster_bitset_edges: B*E exceeds INT32_MAX");
    const long long np = static_cast<long long>(num_views) * H * W;
    if (BE == 0 || np == 0) return S2C_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (num_slots < kPlaneMinSlots) {
        antialias_edges_bwd_kernel<false, kAaEdgeItems>
            <<<grid_for(BE, kBlock * kAaEdgeItems), kBlock, 0, s>>>(
            BE, H, W, num_slots, bitset, target_bitset, V, pos, E, edges, prim_ids, normals,
            edge2cluster, d_pos, kernel_radius, rho, num_views, nullptr);
        S2C_LAUNCH_CHECK();
        return S2C_OK;
    }
    PlaneBuffers pl;
    if (int r = pl.alloc(num_views, H, W, num_slots, s)) return r;
    const int reach = plane_reach(kernel_radius, H, W);
    plane_tiles_edges_kernel<<<grid_for(BE, kBlock), kBlock, 0, s>>>(
        BE, prim_ids, E, edges, V, pos, H, W, reach, pl.tiles);
    S2C_LAUNCH_CHECK();
    residual_planes_kernel<<<grid_for(np, kPlanePix), kBlock, 0, s>>>(
        np, static_cast<long long>(H) * W, W, num_slots, bitset, target_bitset, pl.planes,
        pl.tiles);
    S2C_LAUNCH_CHECK();
    antialias_edges_bwd_kernel<true, kAaEdgeItems>
        <<<grid_for(BE, kBlock * kAaEdgeItems), kBlock, 0, s>>>(
        BE, H, W, num_slots, bitset, target_bitset, V, pos, E, edges, prim_ids, normals,
        edge2cluster, d_pos, kernel_radius, rho, num_views, pl.planes);
    S2C_LAUNCH_CHECK();
    return S2C_OK;
}

extern "C" int s2c_backward_antialiased_cluster_bitset_views(
    int num_views, int num_frags, const int* frag_pix, const float* frag_attrs_dda, int H, int W,
    int num_slots, const int* bitset, const int* target_bitset, int V, const float* pos, int E,
    const int* edges, const float* normals, const int* edge2cluster, float* d_pos,
    float kernel_radius, float rho, void* stream) {
    S2C_CHECK(num_views >= 0 && num_frags >= 0 && H >= 0 && W >= 0 && num_slots >= 0 && V >= 0 &&
                  E >= 0,
              "backward_antialiased_cluster_bitset: negative size");
    if (num_frags == 0) return S2C_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const long long np = static_cast<long long>(num_views) * H * W;
    if (num_views == 0 || num_slots < kPlaneMinSlots || np == 0) {
        antialias_bwd_kernel<false><<<grid_for(num_frags, kBlock), kBlock, 0, s>>>(
            num_frags, frag_pix, frag_attrs_dda, H, W, num_slots, bitset, target_bitset, V, pos,
            E, edges, normals, edge2cluster, d_pos, kernel_radius, rho, 0, nullptr);
        S2C_LAUNCH_CHECK();
        return S2C_OK;
    }
    PlaneBuffers pl;
    if (int r = pl.alloc(num_views, H, W, num_slots, s)) return r;
    const int reach = plane_reach(kernel_radius, H, W);
    plane_tiles_frags_kernel<<<grid_for(num_frags, kBlock), kBlock, 0, s>>>(
        num_frags, frag_pix, H, W, num_views, reach, pl.tiles);
    S2C_LAUNCH_CHECK();
    residual_planes_kernel<<<grid_for(np, kPlanePix), kBlock, 0, s>>>(
        np, static_cast<long long>(H) * W, W, num_slots, bitset, target_bitset, pl.planes,
        pl.tiles);
    S2C_LAUNCH_CHECK();
    antialias_bwd_kernel<true><<<grid_for(num_frags, kBlock), kBlock, 0, s>>>(
        num_frags, frag_pix, frag_attrs_dda, H, W, num_slots, bitset, target_bitset, V, pos, E,
        edges, normals, edge2cluster, d_pos, kernel_radius, rho, num_views, pl.planes);
    S2C_LAUNCH_CHECK();
    return S2C_OK;
}
