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
// Each workgroup owns kExpand consecutive candidate slots.  The primitive of
// every slot comes from an LDS table that the workgroup fills from the
// primitives overlapping its slots (their prefix ranges, read coalesced),
// so no slot searches the prefix array (dda.cu:164-172 searches it per
// fragment).  The first primitive of each workgroup comes from
// dda_block_first_kernel, one thread per primitive.

constexpr int kExpandItems = 4;
constexpr int kExpand = kBlock * kExpandItems;

// first[b] = the primitive holding item b * rows of an inclusive prefix,
// for every b <= last with b * rows < prefix[P - 1].
__global__ void __launch_bounds__(kBlock)
dda_block_first_kernel(int P, const int* __restrict__ prefix, int rows, int last,
                       int* __restrict__ first) {
    const int p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= P) return;
    const long long beg = p == 0 ? 0 : prefix[p - 1];
    const long long end = prefix[p];
    for (long long b = dmax((beg + rows - 1) / rows, 0ll); b <= last && b * rows < end; ++b)
        first[b] = p;
}

// Per candidate slot of a workgroup: its primitive (-1 past the
// candidates), candidate index within the primitive, view, edge and vertex
// indices.  Filled by one thread per primitive (coalesced prefix and id
// loads, one edge gather), so a slot's own loads are one level deep.
struct SlotTable {
    int p[kExpand], k[kExpand], b[kExpand], e[kExpand], v0[kExpand], v1[kExpand];
};

// Ends with a barrier.
__device__ __forceinline__ void fill_slots(int blk, int P, int total,
                                           const int* __restrict__ prefix,
                                           const int* __restrict__ first,
                                           const int* __restrict__ edge_ids, int E,
                                           const int* __restrict__ edges, SlotTable& t) {
    for (int s = threadIdx.x; s < kExpand; s += kBlock) t.p[s] = -1;
    __syncthreads();
    const long long f0 = static_cast<long long>(blk) * kExpand;
    if (f0 < total) {
        const int lo = dmin(dmax(first[blk], 0), P - 1);
        const int hi = f0 + kExpand < total ? dmin(dmax(first[blk + 1], lo), P - 1) : P - 1;
        for (int p = lo + static_cast<int>(threadIdx.x); p <= hi; p += kBlock) {
            const long long beg = p == 0 ? 0 : prefix[p - 1];
            const int s0 = static_cast<int>(dmax(beg - f0, 0ll));
            const int s1 = static_cast<int>(dmin(static_cast<long long>(prefix[p]) - f0,
                                                 static_cast<long long>(kExpand)));
            if (s0 >= s1) continue;
            const int id = edge_ids[p];
            if (id < 0) continue;
            const int b = id / E;
            const int e = id - b * E;
            const int2 ev = reinterpret_cast<const int2*>(edges)[e];
            for (int q = s0; q < s1; ++q) {
                t.p[q] = p;
                t.k[q] = static_cast<int>(f0 + q - beg);
                t.b[q] = b;
                t.e[q] = e;
                t.v0[q] = ev.x;
                t.v1[q] = ev.y;
            }
        }
    }
    __syncthreads();
}

struct SlotFrag {
    int b, e;
    float axis;
    DdaFrag fr;
};

// dda.cu:174-225 for candidate slot s of workgroup blk: false when the slot
// is past the candidates or its fragment fails the screen / depth test.
__device__ __forceinline__ bool slot_fragment(int blk, int s, int N, const SlotTable& t,
                                              const float* __restrict__ slopes,
                                              const float* __restrict__ spans, int H, int W,
                                              int V, const float* __restrict__ pos,
                                              SlotFrag& o) {
    const long long f = static_cast<long long>(blk) * kExpand + s;
    const int p = t.p[s];
    if (f >= N || p < 0) return false;
    o.b = t.b[s];
    o.e = t.e[s];
    const float2 sl = reinterpret_cast<const float2*>(slopes)[p];
    const float4 sn = reinterpret_cast<const float4*>(spans)[p];
    const DdaSpan sp{sn.x, sn.y, sn.z, sn.w, sl.x, sl.y};
    o.axis = sl.x;
    const long long vb = static_cast<long long>(o.b) * V;
    const float4 p0 = load_pos(pos, vb + t.v0[s]);
    const float4 p1 = load_pos(pos, vb + t.v1[s]);
    return dda_frag(H, W, sp, t.k[s], p0, p1, o.fr);
}

__device__ __forceinline__ void store_fragment(long long o, const SlotFrag& q,
                                               int* __restrict__ frag_pix,
                                               float* __restrict__ frag_attrs) {
    frag_pix[o * 3 + 0] = q.b;
    frag_pix[o * 3 + 1] = q.fr.h;
    frag_pix[o * 3 + 2] = q.fr.w;
    reinterpret_cast<float4*>(frag_attrs)[o] =
        make_float4(q.axis, q.fr.t1, q.fr.zw, static_cast<float>(q.e + 1));
}

// dda.cu:146-225: every candidate slot; failing ones are not written.
__global__ void __launch_bounds__(kBlock)
dda_fragments_kernel(int P, int N, const int* __restrict__ prefix,
                     const int* __restrict__ first, const int* __restrict__ edge_ids,
                     const float* __restrict__ slopes, const float* __restrict__ spans, int H,
                     int W, int V, const float* __restrict__ pos, int E,
                     const int* __restrict__ edges, int* __restrict__ frag_pix,
                     float* __restrict__ frag_attrs) {
    __shared__ SlotTable tab;
    const int blk = blockIdx.x;
    fill_slots(blk, P, prefix[P - 1], prefix, first, edge_ids, E, edges, tab);
#pragma unroll
    for (int j = 0; j < kExpandItems; ++j) {
        const int s = j * kBlock + threadIdx.x;
        SlotFrag q;
        if (slot_fragment(blk, s, N, tab, slopes, spans, H, W, V, pos, q))
            store_fragment(static_cast<long long>(blk) * kExpand + s, q, frag_pix, frag_attrs);
    }
}

// Expansion and stable compaction fused (dda.py:60-76: dda_compute_fragments
// into a -1-filled candidate list, then filter_valid_fragments), in the
// space of surviving fragments instead of candidates.  A counting pass
// walks each primitive's candidates and counts the survivors (the
// division-free pixel test where the whole edge passes the depth test, the
// full test otherwise); an inclusive scan of the counts gives each
// primitive's place in the output.  The emitting pass gives each workgroup
// kExpand consecutive output rows: one thread per overlapping primitive
// walks its candidates again, places the survivors that fall in the
// workgroup's rows into LDS, and the workgroup writes its rows with
// coalesced stores.  No candidate list, slopes or spans are stored.
// edge_ids may be null: primitive p is then edge p.

__device__ __forceinline__ int prim_id(const int* __restrict__ edge_ids, int p) {
    return edge_ids ? edge_ids[p] : p;
}

__global__ void __launch_bounds__(kBlock)
dda_survivors_kernel(int P, const int* __restrict__ edge_ids, int H, int W, int V,
                     const float* __restrict__ pos, int E, const int* __restrict__ edges,
                     int* __restrict__ counts) {
    const int p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= P) return;
    const int id = prim_id(edge_ids, p);
    int n = 0;
    if (id >= 0) {
        const int b = id / E;
        const int e = id - b * E;
        const int2 ev = reinterpret_cast<const int2*>(edges)[e];
        const long long vb = static_cast<long long>(b) * V;
        const float4 p0 = load_pos(pos, vb + ev.x);
        const float4 p1 = load_pos(pos, vb + ev.y);
        DdaSpan sp;
        bool wrote;
        const int cnt = dda_span(H, W, p0, p1, sp, wrote);
        const bool depth_ok = dda_depth_always_passes(p0, p1);
        for (int k = 0; k < cnt; ++k) {
            int h, w;
            n += dda_mask_frag(H, W, sp, k, p0, p1, depth_ok, h, w);
        }
    }
    counts[p] = n;
}

constexpr int kEmitRows = 512;  // output rows per workgroup (1,024: 0.66 ms, 512: 0.56 ms)
constexpr int kEmitThreads = 256;

__global__ void __launch_bounds__(kEmitThreads)
dda_emit_kernel(int P, int N, const int* __restrict__ prefix, const int* __restrict__ first,
                const int* __restrict__ edge_ids, int H, int W, int V,
                const float* __restrict__ pos, int E, const int* __restrict__ edges,
                int* __restrict__ frag_pix, float* __restrict__ frag_attrs) {
    __shared__ float4 stage_attr[kEmitRows];
    __shared__ __align__(16) int stage_pix[kEmitRows * 3];
    const int blk = blockIdx.x;
    const long long f0 = static_cast<long long>(blk) * kEmitRows;
    const int total = prefix[P - 1];
    const int rows = static_cast<int>(dmin(static_cast<long long>(dmin(total, N)) - f0,
                                           static_cast<long long>(kEmitRows)));
    if (rows <= 0) return;  // workgroup-uniform
    const int lo = dmin(dmax(first[blk], 0), P - 1);
    const int hi = f0 + kEmitRows < total ? dmin(dmax(first[blk + 1], lo), P - 1) : P - 1;
    for (int p = lo + static_cast<int>(threadIdx.x); p <= hi; p += kEmitThreads) {
        const long long beg = p == 0 ? 0 : prefix[p - 1];
        if (prefix[p] <= beg || beg >= f0 + kEmitRows || prefix[p] <= f0) continue;
        const int id = prim_id(edge_ids, p);
        if (id < 0) continue;
        const int b = id / E;
        const int e = id - b * E;
        const int2 ev = reinterpret_cast<const int2*>(edges)[e];
        const long long vb = static_cast<long long>(b) * V;
        const float4 p0 = load_pos(pos, vb + ev.x);
        const float4 p1 = load_pos(pos, vb + ev.y);
        DdaSpan sp;
        bool wrote;
        const int cnt = dda_span(H, W, p0, p1, sp, wrote);
        long long q = beg;  // output row of the next survivor
        for (int k = 0; k < cnt && q < f0 + kEmitRows; ++k) {
            DdaFrag fr;
            if (!dda_frag(H, W, sp, k, p0, p1, fr)) continue;
            if (q >= f0) {
                const int r = static_cast<int>(q - f0);
                stage_pix[r * 3 + 0] = b;
                stage_pix[r * 3 + 1] = fr.h;
                stage_pix[r * 3 + 2] = fr.w;
                stage_attr[r] = make_float4(sp.axis, fr.t1, fr.zw, static_cast<float>(e + 1));
            }
            ++q;
        }
    }
    __syncthreads();
    // rows start at f0 * 12 bytes, a multiple of 16: 16-byte stores, then
    // the tail of the last workgroup's rows
    const int quads = rows * 3 / 4;
    int4* pout = reinterpret_cast<int4*>(frag_pix + f0 * 3);
    const int4* pst = reinterpret_cast<const int4*>(stage_pix);
    for (int i = threadIdx.x; i < quads; i += kEmitThreads) pout[i] = pst[i];
    for (int i = quads * 4 + threadIdx.x; i < rows * 3; i += kEmitThreads)
        frag_pix[f0 * 3 + i] = stage_pix[i];
    float4* out = reinterpret_cast<float4*>(frag_attrs) + f0;
    for (int i = threadIdx.x; i < rows; i += kEmitThreads) out[i] = stage_attr[i];
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

namespace {
// first[] of the workgroups of `rows` items each in `t`, followed by
// `extra` ints.
int block_first(int P, int N, int rows, const int* prefix, hipStream_t s, Temp& t, int extra,
                int** first) {
    const int blocks = static_cast<int>((static_cast<long long>(N) + rows - 1) / rows);
    if (int r = t.alloc(sizeof(int) * (static_cast<size_t>(blocks) + 1 + extra), s)) return r;
    *first = t.as<int>();
    dda_block_first_kernel<<<grid_for(P, kBlock), kBlock, 0, s>>>(P, prefix, rows, blocks,
                                                                  *first);
    S2C_LAUNCH_CHECK();
    return S2C_OK;
}
}  // namespace

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
    Temp t;
    int* first = nullptr;
    if (int r = block_first(num_prims, num_frags, kExpand, frag_prefix_sum, s, t, 0, &first))
        return r;
    dda_fragments_kernel<<<grid_for(num_frags, kExpand), kBlock, 0, s>>>(
        num_prims, num_frags, frag_prefix_sum, first, edge_ids, frag_slopes, frag_spans, H, W,
        V, pos, E, edges, frag_pix, frag_attrs);
    S2C_LAUNCH_CHECK();
    return S2C_OK;
}

extern "C" int s2c_dda_count_fragments(int num_prims, const int* edge_ids, int H, int W,
                                       int V, const float* pos, int E, const int* edges,
                                       int* valid_prefix, int* num_frags_out, void* stream) {
    S2C_CHECK(num_frags_out != nullptr, "dda_count_fragments: null count pointer");
    *num_frags_out = 0;
    S2C_CHECK(num_prims >= 0 && H >= 0 && W >= 0 && V >= 0 && E >= 0,
              "dda_count_fragments: negative size");
    if (num_prims == 0) return S2C_OK;
    S2C_CHECK(E > 0, "dda_count_fragments: E must be positive when primitives are given");
    hipStream_t s = static_cast<hipStream_t>(stream);
    dda_survivors_kernel<<<grid_for(num_prims, kBlock), kBlock, 0, s>>>(
        num_prims, edge_ids, H, W, V, pos, E, edges, valid_prefix);
    S2C_LAUNCH_CHECK();
    Temp scratch;
    if (int r = scratch.alloc(sizeof(int) * scan_scratch_ints(num_prims), s)) return r;
    if (int r = inclusive_scan_inplace(valid_prefix, num_prims, scratch.as<int>(), s)) return r;
    return read_scan_total(scratch.as<int>(), num_prims, s, num_frags_out, "dda_count_fragments");
}

extern "C" int s2c_dda_emit_fragments(int num_prims, int num_frags, const int* valid_prefix,
                                      const int* edge_ids, int H, int W, int V,
                                      const float* pos, int E, const int* edges, int* frag_pix,
                                      float* frag_attrs, void* stream) {
    S2C_CHECK(num_prims >= 0 && num_frags >= 0, "dda_emit_fragments: negative size");
    if (num_frags == 0 || num_prims == 0) return S2C_OK;
    S2C_CHECK(E > 0, "dda_emit_fragments: E must be positive");
    hipStream_t s = static_cast<hipStream_t>(stream);
    Temp t;
    int* first = nullptr;
    if (int r = block_first(num_prims, num_frags, kEmitRows, valid_prefix, s, t, 0, &first))
        return r;
    dda_emit_kernel<<<grid_for(num_frags, kEmitRows), kEmitThreads, 0, s>>>(
        num_prims, num_frags, valid_prefix, first, edge_ids, H, W, V, pos, E, edges, frag_pix,
        frag_attrs);
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
