// Fused DDA -> cluster bitset (the target-mask sequence of
// 3_fitting.py:254-261: dda.py:5-76 followed by rasterize.py:170-214).
//
// Work decomposition: a workgroup takes kPrimsPerBlock consecutive edges
// (b*E + e order), kPrimsPerThread consecutive edges per thread, so each
// thread walks along a strand and its consecutive fragments are
// neighbouring pixels.  A thread walks the first kSerialFrags candidates of
// each of its edges itself; the rest of a longer edge is deferred and walked
// by the whole workgroup, so one long edge cannot stall a wave.
#include "binning.hpp"
#include "dda_math.hpp"

namespace s2c {

constexpr int kPrimsPerThread = 4;
constexpr int kPrimsPerBlock = kBlock * kPrimsPerThread;
constexpr int kSerialFrags = 64;
constexpr int kDeferred = 256;  // long edges per workgroup walked by all threads

struct EdgeView {
    DdaSpan sp;
    float4 p0, p1;
    int b, cluster, cnt;
    bool depth_ok;
};

__device__ __forceinline__ EdgeView load_edge(long long q, int H, int W, int V,
                                              const float* __restrict__ pos, int E,
                                              const int* __restrict__ edges,
                                              const int* __restrict__ edge2cluster) {
    EdgeView v;
    v.b = static_cast<int>(q / E);
    const int e = static_cast<int>(q - static_cast<long long>(v.b) * E);
    const int2 ev = reinterpret_cast<const int2*>(edges)[e];
    const long long vb = static_cast<long long>(v.b) * V;
    v.p0 = load_pos(pos, vb + ev.x);
    v.p1 = load_pos(pos, vb + ev.y);
    bool wrote;
    v.cnt = dda_span(H, W, v.p0, v.p1, v.sp, wrote);
    v.cluster = v.cnt > 0 ? edge2cluster[e] : 0;
    v.depth_ok = dda_depth_always_passes(v.p0, v.p1);
    return v;
}

// emit(b, h, w, cluster) for every surviving fragment of the workgroup's
// edges.  `deferred` / `ndeferred` are LDS scratch (kDeferred ints + 1,
// ndeferred zeroed and synchronized by the caller).  Contains
// __syncthreads(): every thread of the workgroup must call it.
template <class Emit>
__device__ __forceinline__ void walk_edges(long long P, int H, int W, int V,
                                           const float* __restrict__ pos, int E,
                                           const int* __restrict__ edges,
                                           const int* __restrict__ edge2cluster, int* deferred,
                                           int* ndeferred, Emit&& emit,
                                           long long blk = blockIdx.x) {
    const long long q0 = blk * kPrimsPerBlock +
                         static_cast<long long>(threadIdx.x) * kPrimsPerThread;
    // The vertex and cluster loads of all kPrimsPerThread edges are issued
    // before the first walk, so the walk of an edge overlaps the later
    // edges' loads.  (Each edge load waits inside its own branch; issuing
    // them unconditionally, all in flight at once, raised the VGPRs past
    // 72 and measured no faster.)
    int b0 = 0, e0 = 0;
    if (q0 < P) {
        b0 = static_cast<int>(q0 / E);
        e0 = static_cast<int>(q0 - static_cast<long long>(b0) * E);
    }
    int bv[kPrimsPerThread], ev[kPrimsPerThread];
    bool ok[kPrimsPerThread];
    int2 vv[kPrimsPerThread];
#pragma unroll
    for (int i = 0; i < kPrimsPerThread; ++i) {
        int e = e0 + i, b = b0;
        while (e >= E) {  // views shorter than kPrimsPerThread edges wrap more than once
            e -= E;
            ++b;
        }
        ok[i] = q0 + i < P;
        bv[i] = b;
        ev[i] = e;
        vv[i] = ok[i] ? reinterpret_cast<const int2*>(edges)[e] : make_int2(0, 0);
    }
    float4 pa[kPrimsPerThread], pb[kPrimsPerThread];
    int cl[kPrimsPerThread];
#pragma unroll
    for (int i = 0; i < kPrimsPerThread; ++i) {
        const long long vb = static_cast<long long>(bv[i]) * V;
        pa[i] = ok[i] ? load_pos(pos, vb + vv[i].x) : make_float4(0.f, 0.f, 0.f, 1.f);
        pb[i] = ok[i] ? load_pos(pos, vb + vv[i].y) : make_float4(0.f, 0.f, 0.f, 1.f);
        cl[i] = ok[i] ? edge2cluster[ev[i]] : 0;
    }
#pragma unroll
    for (int i = 0; i < kPrimsPerThread; ++i) {
        if (!ok[i]) break;
        DdaSpan sp;
        bool wrote;
        const int cnt = dda_span(H, W, pa[i], pb[i], sp, wrote);
        const bool depth_ok = dda_depth_always_passes(pa[i], pb[i]);
        const int n = dmin(cnt, kSerialFrags);
        for (int k = 0; k < n; ++k) {
            int h, w;
            if (dda_mask_frag(H, W, sp, k, pa[i], pb[i], depth_ok, h, w)) emit(bv[i], h, w, cl[i]);
        }
        if (cnt > kSerialFrags) {
            const int d = atomicAdd(ndeferred, 1);
            if (d < kDeferred) {
                deferred[d] = static_cast<int>(q0 + i - blk * kPrimsPerBlock);
            } else {  // list full: the thread walks the rest of its edge itself
                for (int k = kSerialFrags; k < cnt; ++k) {
                    int h, w;
                    if (dda_mask_frag(H, W, sp, k, pa[i], pb[i], depth_ok, h, w))
                        emit(bv[i], h, w, cl[i]);
                }
            }
        }
    }
    __syncthreads();
    const int nd = dmin(*ndeferred, kDeferred);
    const long long qb = blk * kPrimsPerBlock;
#pragma unroll 1
    for (int d = 0; d < nd; ++d) {
        const EdgeView v = load_edge(qb + deferred[d], H, W, V, pos, E, edges, edge2cluster);
        for (int k = kSerialFrags + threadIdx.x; k < v.cnt; k += kBlock) {
            int h, w;
            if (dda_mask_frag(H, W, v.sp, k, v.p0, v.p1, v.depth_ok, h, w)) emit(v.b, h, w, v.cluster);
        }
    }
}

// Direct form: one global atomicOr per fragment into a zeroed bitset, for
// the edges of block index `blk`.
__device__ __forceinline__ void dda_mask_direct(long long P, int H, int W, int V,
                                                const float* __restrict__ pos, int E,
                                                const int* __restrict__ edges,
                                                const int* __restrict__ edge2cluster, int S,
                                                int* __restrict__ bitset, long long blk) {
    __shared__ int deferred[kDeferred];
    __shared__ int ndeferred;
    if (threadIdx.x == 0) ndeferred = 0;
    __syncthreads();
    walk_edges(
        P, H, W, V, pos, E, edges, edge2cluster, deferred, &ndeferred,
        [&](int b, int h, int w, int c) {
            // ids beyond the bitset set no bit, as in the binned form
            if (static_cast<unsigned>(c) >= static_cast<unsigned>(S) * 32u) return;
            const long long idx = ((static_cast<long long>(b) * H + h) * W + w) * S + (c >> 5);
            atomicOr(&bitset[idx], static_cast<int>(1u << (c & 31)));
        },
        blk);
    __syncthreads();  // `deferred` is reused by the caller's next block
}

// For tile bitsets that do not fit LDS.
__global__ void __launch_bounds__(kBlock)
dda_mask_direct_kernel(long long P, int H, int W, int V, const float* __restrict__ pos, int E,
                       const int* __restrict__ edges, const int* __restrict__ edge2cluster, int S,
                       int* __restrict__ bitset) {
    dda_mask_direct(P, H, W, V, pos, E, edges, edge2cluster, S, bitset, blockIdx.x);
}

// Behind the binned pass: OR every fragment of the stage workgroups whose
// key table overflowed (record count -1) directly into the resolved
// bitset.  Workgroup i checks stage workgroups [i * kBlock, (i + 1) * kBlock)
// with one load per thread (grid-striding), so when none overflowed the
// kernel costs a few microseconds.
constexpr int kFallbackGrid = 256;

// Calls f(b) for every overflowed stage workgroup b of this workgroup's
// windows (workgroup-uniform).
template <class F>
__device__ __forceinline__ void for_overflowed(long long blocks, const int* __restrict__ block_recs,
                                               F&& f) {
    __shared__ int list[kBlock];
    __shared__ int nlist;
    for (long long w0 = static_cast<long long>(blockIdx.x) * kBlock; w0 < blocks;
         w0 += static_cast<long long>(gridDim.x) * kBlock) {
        if (threadIdx.x == 0) nlist = 0;
        __syncthreads();
        const long long b = w0 + threadIdx.x;
        if (b < blocks && block_recs[b] < 0) list[atomicAdd(&nlist, 1)] = static_cast<int>(b);
        __syncthreads();
        const int n = nlist;
        for (int i = 0; i < n; ++i) f(static_cast<long long>(list[i]));
        __syncthreads();  // list and nlist are rewritten by the next window
    }
}

__global__ void __launch_bounds__(kBlock)
dda_mask_fallback_kernel(long long P, int H, int W, int V, const float* __restrict__ pos, int E,
                         const int* __restrict__ edges, const int* __restrict__ edge2cluster,
                         int S, int* __restrict__ bitset, long long blocks,
                         const int* __restrict__ block_recs) {
    for_overflowed(blocks, block_recs, [&](long long blk) {
        dda_mask_direct(P, H, W, V, pos, E, edges, edge2cluster, S, bitset, blk);
    });
}

// One global OR for fragment i of a list (rasterize.cu:474-498), with
// ListSource's checks.
__device__ __forceinline__ void list_or_one(long long i, int B, int H, int W,
                                            const int* __restrict__ frag_pix,
                                            const float* __restrict__ frag_attrs,
                                            const int* __restrict__ prim2cluster, unsigned kmax,
                                            int S, int* __restrict__ bitset) {
    const int prim = f2i(frag_attrs[i * 4 + 3]) - 1;
    if (prim < 0) return;
    const int c = prim2cluster[prim];
    const int b = frag_pix[i * 3 + 0], h = frag_pix[i * 3 + 1], w = frag_pix[i * 3 + 2];
    if (static_cast<unsigned>(c) >= kmax || static_cast<unsigned>(b) >= static_cast<unsigned>(B) ||
        static_cast<unsigned>(h) >= static_cast<unsigned>(H) ||
        static_cast<unsigned>(w) >= static_cast<unsigned>(W))
        return;
    const long long idx = ((static_cast<long long>(b) * H + h) * W + w) * S + (c >> 5);
    atomicOr(&bitset[idx], static_cast<int>(1u << (c & 31)));
}

// Into a zeroed bitset, grid-striding over the list.
__global__ void __launch_bounds__(kBlock)
list_mask_direct_kernel(long long N, int B, int H, int W, const int* __restrict__ frag_pix,
                        const float* __restrict__ frag_attrs,
                        const int* __restrict__ prim2cluster, unsigned kmax, int S,
                        int* __restrict__ bitset) {
    const long long stride = static_cast<long long>(gridDim.x) * kBlock;
    for (long long i = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x; i < N;
         i += stride)
        list_or_one(i, B, H, W, frag_pix, frag_attrs, prim2cluster, kmax, S, bitset);
}

// Fragments per workgroup of the list's stage pass.
constexpr int kListFrags = 4096;

// The list's counterpart of dda_mask_fallback_kernel: the kListFrags
// fragments of each overflowed stage workgroup.
__global__ void __launch_bounds__(kBlock)
list_mask_fallback_kernel(long long N, int B, int H, int W, const int* __restrict__ frag_pix,
                          const float* __restrict__ frag_attrs,
                          const int* __restrict__ prim2cluster, unsigned kmax, int S,
                          int* __restrict__ bitset, long long blocks,
                          const int* __restrict__ block_recs) {
    for_overflowed(blocks, block_recs, [&](long long blk) {
        const long long f0 = blk * kListFrags;
        const long long f1 = dmin(f0 + kListFrags, N);
        for (long long f = f0 + threadIdx.x; f < f1; f += kBlock)
            list_or_one(f, B, H, W, frag_pix, frag_attrs, prim2cluster, kmax, S, bitset);
    });
}

// Binned form: one walk.  Each fragment sets its pixel's bit in the LDS
// mask of its (cell, cluster) key (cell_key), the key's slot found in a
// kHash-slot LDS table once per change of key along a thread's fragments;
// a fragment repeating a pixel and cluster of the workgroup sets no new
// bit.  After the walk a block scan of the used slots gives each key its
// record in the workgroup's own region: (tile, cluster | cell << 16) and
// the key's mask.  A workgroup whose table overflows writes no record and
// marks its count -1 for the fallback kernel.  Cluster ids outside
// [0, kmax) set no bit.

// The fragment sources of the binned pass.  walk(emit) calls
// emit(b, h, w, cluster) for every fragment of the workgroup's share of the
// work (every thread calls it: it may hold barriers); thread 0 calls
// reset() before a second walk, with a barrier after it.

// The DDA walk of kPrimsPerBlock consecutive edges (dda.cu:124-225).
struct EdgeSource {
    long long P;
    int H, W, V;
    const float* __restrict__ pos;
    int E;
    const int* __restrict__ edges;
    const int* __restrict__ edge2cluster;
    int* deferred;  // LDS, kDeferred ints
    int* ndeferred;
    template <class Emit>
    __device__ __forceinline__ void walk(Emit&& emit) const {
        walk_edges(P, H, W, V, pos, E, edges, edge2cluster, deferred, ndeferred, emit);
    }
    __device__ __forceinline__ void reset() const { *ndeferred = 0; }
};

// kListFrags consecutive fragments of a list (rasterize.cu:474-498) in
// kListRounds rounds; in each, thread t takes 4 consecutive fragments, read
// with 16-byte loads (a wave's loads cover 3 KB of frag_pix and 4 KB of
// frag_attrs contiguously), so a thread's fragments of a round mostly share
// a tile.  A fragment with prim id <= 0, a pixel outside the B x H x W
// image or a cluster beyond the bitset sets no bit.
constexpr int kListRounds = kListFrags / (4 * kBlock);

struct ListSource {
    long long N;
    int B, H, W;
    const int* __restrict__ frag_pix;
    const float* __restrict__ frag_attrs;
    const int* __restrict__ prim2cluster;
    unsigned kmax;
    template <class Emit>
    __device__ __forceinline__ void walk(Emit&& emit) const {
#pragma unroll 1
        for (int r = 0; r < kListRounds; ++r) {
            const long long i0 = static_cast<long long>(blockIdx.x) * kListFrags +
                                 static_cast<long long>(r) * (kBlock * 4) +
                                 static_cast<long long>(threadIdx.x) * 4;
            int pix[12], prim[4], c[4];
            if (i0 + 4 <= N) {  // 48 and 64 bytes from 16-byte aligned rows
                const int4* pp = reinterpret_cast<const int4*>(frag_pix + i0 * 3);
                const int4 a = pp[0], b = pp[1], d = pp[2];
                pix[0] = a.x; pix[1] = a.y; pix[2] = a.z; pix[3] = a.w;
                pix[4] = b.x; pix[5] = b.y; pix[6] = b.z; pix[7] = b.w;
                pix[8] = d.x; pix[9] = d.y; pix[10] = d.z; pix[11] = d.w;
                const float4* pa = reinterpret_cast<const float4*>(frag_attrs) + i0;
#pragma unroll
                for (int u = 0; u < 4; ++u) prim[u] = f2i(pa[u].w) - 1;
            } else {
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const long long i = i0 + u;
                    prim[u] = -1;
                    pix[u * 3] = pix[u * 3 + 1] = pix[u * 3 + 2] = 0;
                    if (i < N) {
                        pix[u * 3 + 0] = frag_pix[i * 3 + 0];
                        pix[u * 3 + 1] = frag_pix[i * 3 + 1];
                        pix[u * 3 + 2] = frag_pix[i * 3 + 2];
                        prim[u] = f2i(frag_attrs[i * 4 + 3]) - 1;
                    }
                }
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) c[u] = prim[u] >= 0 ? prim2cluster[prim[u]] : -1;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (prim[u] < 0 || static_cast<unsigned>(c[u]) >= kmax) continue;
                const int b = pix[u * 3], h = pix[u * 3 + 1], w = pix[u * 3 + 2];
                if (static_cast<unsigned>(b) >= static_cast<unsigned>(B) ||
                    static_cast<unsigned>(h) >= static_cast<unsigned>(H) ||
                    static_cast<unsigned>(w) >= static_cast<unsigned>(W))
                    continue;
                emit(b, h, w, c[u]);
            }
        }
    }
    __device__ __forceinline__ void reset() const {}
};

// The stage pass over one workgroup's fragments from `src`.  `keys` and
// `mask` are LDS (kHash ints, kHash << mw_log2 words), `sh` LDS scratch of
// block_excl_scan.
template <class Source>
__device__ __forceinline__ void cell_stage(const Source& src, const TileGrid g, unsigned kmax,
                                           const BinOut out, int* keys, unsigned* mask,
                                           int& full, int* sh) {
    const int mwl = g.mw_log2;
    for (int s = threadIdx.x; s < kHash; s += kBlock) keys[s] = -1;
    for (int i = threadIdx.x; i < (kHash << mwl) / 4; i += kBlock)
        reinterpret_cast<uint4*>(mask)[i] = make_uint4(0u, 0u, 0u, 0u);
    if (threadIdx.x == 0) {
        src.reset();
        full = 0;
    }
    __syncthreads();
    // current key and the byte offset of its mask in `mask` (-1: none)
    int cur = -1, row = -1;
    src.walk([&](int b, int h, int w, int c) {
        if (static_cast<unsigned>(c) >= kmax) return;
        const int key = cell_key(g, b, h, w, c);
        if (key != cur) {
            cur = key;
            const int slot = full ? -1 : hash_insert(keys, key);
            if (slot < 0) full = 1;
            row = slot < 0 ? -1 : slot << (mwl + 2);
        }
        if (row >= 0) {
            int word;
            unsigned bit;
            cell_bit(g, h, w, word, bit);
            atomicOr(reinterpret_cast<unsigned*>(reinterpret_cast<char*>(mask) + row) + word,
                     1u << bit);
        }
    });
    __syncthreads();
    if (full) {  // workgroup-uniform
        if (threadIdx.x == 0) out.block_recs[blockIdx.x] = -1;
        return;
    }
    constexpr int per_thread = kHash / kBlock;
    int key[per_thread], used = 0;
#pragma unroll
    for (int j = 0; j < per_thread; ++j) {
        key[j] = keys[threadIdx.x * per_thread + j];
        used += key[j] >= 0;
    }
    int nrec;
    int rec = block_excl_scan(used, sh, &nrec);  // this thread's first record
    if (threadIdx.x == 0) out.block_recs[blockIdx.x] = nrec;
    const long long first = static_cast<long long>(blockIdx.x) * kBinRecs;
    const int cmask = (1 << g.kb) - 1;
#pragma unroll
    for (int j = 0; j < per_thread; ++j) {
        if (key[j] < 0) continue;
        int tile, cell;
        cell_tile(g, key[j], tile, cell);
        const long long r = first + rec++;
        const unsigned* m = mask + ((threadIdx.x * per_thread + j) << mwl);
        uint4 m0, m1;
        if (mwl == 3) {  // 16 x 16 cells: all 8 words
            m0 = reinterpret_cast<const uint4*>(m)[0];
            m1 = reinterpret_cast<const uint4*>(m)[1];
        } else {
            unsigned v[kMaskWordsMax] = {};
            for (int i = 0; i < (1 << mwl); ++i) v[i] = m[i];
            m0 = make_uint4(v[0], v[1], v[2], v[3]);
            m1 = make_uint4(v[4], v[5], v[6], v[7]);
        }
        out.rec_ti[r] = make_int2(tile, (key[j] & cmask) | (cell << 16));
        out.rec_mask[2 * r] = m0;
        out.rec_mask[2 * r + 1] = m1;
    }
}

// LDS of a stage workgroup (18 KB, with the walk's deferred list 19 KB).
#define S2C_STAGE_LDS                                                                  \
    __shared__ int keys[kHash];     /* (cell, cluster) key of each slot */              \
    __shared__ __align__(16) unsigned mask[kHash * kMaskWordsMax]; /* slot masks */     \
    __shared__ int full;                                                               \
    __shared__ int sh[kBlock / kWave + 1];

__global__ void __launch_bounds__(kBlock)
dda_cell_stage_kernel(long long P, int H, int W, int V, const float* __restrict__ pos, int E,
                      const int* __restrict__ edges, const int* __restrict__ edge2cluster,
                      unsigned kmax, TileGrid g, BinOut out) {
    S2C_STAGE_LDS
    __shared__ int deferred[kDeferred];
    __shared__ int ndeferred;
    const EdgeSource src{P, H, W, V, pos, E, edges, edge2cluster, deferred, &ndeferred};
    cell_stage(src, g, kmax, out, keys, mask, full, sh);
}

// The stage pass over a fragment list: kListFrags fragments per workgroup.
__global__ void __launch_bounds__(kBlock)
list_cell_stage_kernel(long long N, int B, int H, int W, const int* __restrict__ frag_pix,
                       const float* __restrict__ frag_attrs, const int* __restrict__ prim2cluster,
                       unsigned kmax, TileGrid g, BinOut out) {
    S2C_STAGE_LDS
    const ListSource src{N, B, H, W, frag_pix, frag_attrs, prim2cluster, kmax};
    cell_stage(src, g, kmax, out, keys, mask, full, sh);
}
#undef S2C_STAGE_LDS

}  // namespace s2c

using namespace s2c;

extern "C" int s2c_dda_cluster_mask(int B, int H, int W, int V, const float* pos, int E,
                                    const int* edges, const int* edge2cluster, int num_slots,
                                    int* bitset, void* stream) {
    return s2c_dda_cluster_mask_tiled(B, H, W, V, pos, E, edges, edge2cluster, num_slots, 0,
                                      bitset, stream);
}

namespace {

// The binned pass after its tile grid is chosen: workspace, stage
// (stage(plan) launches the stage kernel over plan.blocks
// workgroups), resolve, and the fallback behind it (fallback(plan) launches
// the kernel that ORs the fragments of the workgroups whose key table
// overflowed).  Nothing waits for the device.
template <class Stage, class Fallback>
int binned_mask(BinPlan& plan, int H, int W, int num_slots, int* bitset, hipStream_t s,
                Stage&& stage, Fallback&& fallback) {
    std::lock_guard<std::mutex> held(BinWorkspace::lock());
    void *work = nullptr, *mbuf = nullptr, *obuf = nullptr;
    if (int r = BinWorkspace::get(s, BinWorkspace::kPlanInts,
                                  sizeof(int) * bin_plan_ints(plan.g, plan.blocks), &work))
        return r;
    bin_carve(plan, static_cast<int*>(work));
    const size_t recs = static_cast<size_t>(plan.records());
    if (int r = BinWorkspace::get(s, BinWorkspace::kMasks, 2 * sizeof(uint4) * recs, &mbuf))
        return r;
    if (int r = BinWorkspace::get(s, BinWorkspace::kRuns, sizeof(int2) * recs, &obuf))
        return r;
    plan.rec_mask = static_cast<uint4*>(mbuf);
    plan.runs = static_cast<int2*>(obuf);
    stage(plan);
    S2C_LAUNCH_CHECK();
    if (int r = bin_resolve(plan, H, W, num_slots, bitset, s)) return r;
    fallback(plan);
    S2C_LAUNCH_CHECK();
    return S2C_OK;
}

unsigned fallback_grid(long long blocks) {
    return static_cast<unsigned>(
        dmin((blocks + kBlock - 1) / kBlock, static_cast<long long>(kFallbackGrid)));
}

}  // namespace

extern "C" int s2c_dda_cluster_mask_tiled(int B, int H, int W, int V, const float* pos, int E,
                                          const int* edges, const int* edge2cluster,
                                          int num_slots, int tile_log2, int* bitset,
                                          void* stream) {
    S2C_CHECK(B >= 0 && H >= 0 && W >= 0 && V >= 0 && E >= 0 && num_slots >= 0,
              "dda_cluster_mask: negative size");
    S2C_CHECK(tile_log2 == 0 || (tile_log2 >= 2 && tile_log2 <= 6),
              "dda_cluster_mask: tile_log2 must be 0 or in [2, 6]");
    hipStream_t s = static_cast<hipStream_t>(stream);
    const long long words = static_cast<long long>(B) * H * W * num_slots;
    const long long P = static_cast<long long>(B) * E;
    if (words == 0) return S2C_OK;
    BinPlan plan;
    plan.blocks = (P + kPrimsPerBlock - 1) / kPrimsPerBlock;
    const bool tiled = P > 0 && choose_tiles(B, H, W, num_slots, 4096, plan.g, tile_log2);
    S2C_CHECK(tiled || tile_log2 == 0 || P == 0,
              "dda_cluster_mask: the requested tile side does not fit this shape (its tile "
              "bitset exceeds LDS, or the (cell, cluster) keys exceed 31 bits)");
    if (!tiled) {
        // Empty input, a tile bitset too large for LDS, or (cell, cluster)
        // keys beyond 31 bits: clear + direct ORs.
        if (int r = fill_u32(bitset, 0u, words, s)) return r;
        if (P == 0) return S2C_OK;
        dda_mask_direct_kernel<<<static_cast<unsigned>(plan.blocks), kBlock, 0, s>>>(
            P, H, W, V, pos, E, edges, edge2cluster, num_slots, bitset);
        S2C_LAUNCH_CHECK();
        return S2C_OK;
    }
    S2C_CHECK(plan.records() < INT_MAX / 2 && plan.sort_blocks() * plan.g.nt < INT_MAX / 2,
              "dda_cluster_mask: too many primitives for 32-bit record indices");
    const unsigned kmax = static_cast<unsigned>(num_slots) * 32u;
    return binned_mask(
        plan, H, W, num_slots, bitset, s,
        [&](const BinPlan& p) {
            dda_cell_stage_kernel<<<static_cast<unsigned>(p.blocks), kBlock, 0, s>>>(
                P, H, W, V, pos, E, edges, edge2cluster, kmax, p.g, p.out());
        },
        [&](const BinPlan& p) {
            dda_mask_fallback_kernel<<<fallback_grid(p.blocks), kBlock, 0, s>>>(
                P, H, W, V, pos, E, edges, edge2cluster, num_slots, bitset, p.blocks,
                p.block_recs);
        });
}

extern "C" int s2c_fragment_cluster_mask(int B, int H, int W, int num_frags, const int* frag_pix,
                                         const float* frag_attrs, const int* prim2cluster,
                                         int num_slots, int* bitset, void* stream) {
    S2C_CHECK(B >= 0 && H >= 0 && W >= 0 && num_frags >= 0 && num_slots >= 0,
              "fragment_cluster_mask: negative size");
    hipStream_t s = static_cast<hipStream_t>(stream);
    const long long words = static_cast<long long>(B) * H * W * num_slots;
    if (words == 0) return S2C_OK;
    const long long N = num_frags;
    const unsigned kmax = static_cast<unsigned>(dmin(static_cast<long long>(num_slots) * 32,
                                                     static_cast<long long>(UINT_MAX)));
    BinPlan plan;
    plan.blocks = (N + kListFrags - 1) / kListFrags;
    const bool tiled = N > 0 && choose_tiles(B, H, W, num_slots, 4096, plan.g);
    if (!tiled) {
        if (int r = fill_u32(bitset, 0u, words, s)) return r;
        if (N == 0) return S2C_OK;
        list_mask_direct_kernel<<<grid_for(N, kBlock), kBlock, 0, s>>>(
            N, B, H, W, frag_pix, frag_attrs, prim2cluster, kmax, num_slots, bitset);
        S2C_LAUNCH_CHECK();
        return S2C_OK;
    }
    S2C_CHECK(plan.records() < INT_MAX / 2 && plan.sort_blocks() * plan.g.nt < INT_MAX / 2,
              "fragment_cluster_mask: too many fragments for 32-bit record indices");
    return binned_mask(
        plan, H, W, num_slots, bitset, s,
        [&](const BinPlan& p) {
            list_cell_stage_kernel<<<static_cast<unsigned>(p.blocks), kBlock, 0, s>>>(
                N, B, H, W, frag_pix, frag_attrs, prim2cluster, kmax, p.g, p.out());
        },
        [&](const BinPlan& p) {
            list_mask_fallback_kernel<<<fallback_grid(p.blocks), kBlock, 0, s>>>(
                N, B, H, W, frag_pix, frag_attrs, prim2cluster, kmax, num_slots, bitset,
                p.blocks, p.block_recs);
        });
}

extern "C" int s2c_dda_cluster_mask_release(void* stream) {
    std::lock_guard<std::mutex> held(BinWorkspace::lock());
    return BinWorkspace::release(static_cast<hipStream_t>(stream));
}
