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
    // All loads of the thread's edges are issued before any is used (one
    // memory round trip for kPrimsPerThread edges instead of one per edge).
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

// Behind the binned pass: when its overflow areas were too small (the
// resolve then wrote every tile as zeros), OR every fragment directly, the
// grid striding over the walk's blocks; otherwise every workgroup exits at
// once (a small grid, so the common case costs a few microseconds).
constexpr int kFallbackGrid = 1024;

__global__ void __launch_bounds__(kBlock)
dda_mask_fallback_kernel(long long P, int H, int W, int V, const float* __restrict__ pos, int E,
                         const int* __restrict__ edges, const int* __restrict__ edge2cluster,
                         int S, int* __restrict__ bitset, const int* __restrict__ counters,
                         int ovf_entry_cap, int ovf_cap, long long blocks) {
    const int ce = counters[kEntriesOvf], cr = counters[kRunsOvf];
    if (ce >= 0 && ce <= ovf_entry_cap && cr >= 0 && cr <= ovf_cap) return;  // workgroup-uniform
    for (long long blk = blockIdx.x; blk < blocks; blk += gridDim.x)
        dda_mask_direct(P, H, W, V, pos, E, edges, edge2cluster, S, bitset, blk);
}

// One global OR per listed fragment (rasterize.cu:474-498) into a zeroed
// bitset, with ListSource's checks; grid-striding.
__device__ __forceinline__ void list_mask_or(long long N, int B, int H, int W,
                                             const int* __restrict__ frag_pix,
                                             const float* __restrict__ frag_attrs,
                                             const int* __restrict__ prim2cluster, unsigned kmax,
                                             int S, int* __restrict__ bitset) {
    const long long stride = static_cast<long long>(gridDim.x) * kBlock;
    for (long long i = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x; i < N;
         i += stride) {
        const int prim = f2i(frag_attrs[i * 4 + 3]) - 1;
        if (prim < 0) continue;
        const int c = prim2cluster[prim];
        const int b = frag_pix[i * 3 + 0], h = frag_pix[i * 3 + 1], w = frag_pix[i * 3 + 2];
        if (static_cast<unsigned>(c) >= kmax || static_cast<unsigned>(b) >= static_cast<unsigned>(B) ||
            static_cast<unsigned>(h) >= static_cast<unsigned>(H) ||
            static_cast<unsigned>(w) >= static_cast<unsigned>(W))
            continue;
        const long long idx = ((static_cast<long long>(b) * H + h) * W + w) * S + (c >> 5);
        atomicOr(&bitset[idx], static_cast<int>(1u << (c & 31)));
    }
}

__global__ void __launch_bounds__(kBlock)
list_mask_direct_kernel(long long N, int B, int H, int W, const int* __restrict__ frag_pix,
                        const float* __restrict__ frag_attrs,
                        const int* __restrict__ prim2cluster, unsigned kmax, int S,
                        int* __restrict__ bitset) {
    list_mask_or(N, B, H, W, frag_pix, frag_attrs, prim2cluster, kmax, S, bitset);
}

__global__ void __launch_bounds__(kBlock)
list_mask_fallback_kernel(long long N, int B, int H, int W, const int* __restrict__ frag_pix,
                          const float* __restrict__ frag_attrs,
                          const int* __restrict__ prim2cluster, unsigned kmax, int S,
                          int* __restrict__ bitset, const int* __restrict__ counters,
                          int ovf_entry_cap, int ovf_cap) {
    const int ce = counters[kEntriesOvf], cr = counters[kRunsOvf];
    if (ce >= 0 && ce <= ovf_entry_cap && cr >= 0 && cr <= ovf_cap) return;  // workgroup-uniform
    list_mask_or(N, B, H, W, frag_pix, frag_attrs, prim2cluster, kmax, S, bitset);
}

// Binned form: one walk.  Each thread stages its fragments in its own
// kStage / kBlock words of LDS as slot << 22 | entry (the tile's LDS hash
// slot and the packed entry), and the workgroup counts fragments per slot
// with one LDS atomic per run of same-tile fragments.  After the walk the
// slots' counts are scanned into offsets within the workgroup's fixed
// entry region, a run record per (workgroup, tile) pair goes to its fixed
// record region at an index from a block scan, each thread puts its column
// back into the stage in slot order (one LDS atomic reserving each run's
// span), and the workgroup copies the stage to its region with coalesced
// stores.  A workgroup whose table overflows (> kHash distinct tiles) makes
// one run per fragment in the overflow areas; one whose thread has more
// fragments than its stage share takes its region from the overflow area
// and walks again to place them.  Nothing is written past the overflow
// capacities; their use is counted regardless, so the host can size them
// and launch again.  (choose_tiles keeps entries below 2^22.)
constexpr int kPerThread = kStage / kBlock;
constexpr unsigned kEntryMask = (1u << 22) - 1u;

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

// kStage consecutive fragments of a list (rasterize.cu:474-498) in
// kListRounds rounds; in each, thread t takes 4 consecutive fragments, read
// with 16-byte loads (a wave's loads cover 3 KB of frag_pix and 4 KB of
// frag_attrs contiguously), so a thread's fragments of a round mostly share
// a tile.  A fragment with prim id <= 0, a pixel outside the B x H x W
// image or a cluster beyond the bitset sets no bit.
constexpr int kListRounds = kPerThread / 4;

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
            const long long i0 = static_cast<long long>(blockIdx.x) * kStage +
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

// The stage pass over one workgroup's fragments from `src`.
template <class Source>
__device__ __forceinline__ void bin_stage(const Source& src, TileGrid g,
                                          int* __restrict__ runs_per_slot,
                                          int* __restrict__ counters,
                                          int* __restrict__ block_runs,
                                          unsigned* __restrict__ entries, int ovf_entries,
                                          BinRecord* __restrict__ records, int ovf_runs,
                                          int* keys, int* vals, unsigned* stage, int& spill,
                                          int& full, int& base_e, int& base_r, int* sh) {
    const long long fixed_e = static_cast<long long>(gridDim.x) * kStage;
    const long long fixed_r = static_cast<long long>(gridDim.x) * kBinRecs;
    const int group = static_cast<int>(blockIdx.x % kBinGroups);
    for (int s = threadIdx.x; s < kHash; s += kBlock) {
        keys[s] = -1;
        vals[s] = 0;
    }
    if (threadIdx.x == 0) {
        src.reset();
        spill = full = 0;
    }
    __syncthreads();
    // thread t's i-th staged fragment lives at i * kBlock + t: consecutive
    // lanes touch consecutive LDS words
    int cur = -1, slot = -1, len = 0, mine = 0;
    auto flush = [&]() {
        if (slot >= 0 && len > 0) atomicAdd(&vals[slot], len);
        len = 0;
    };
    src.walk([&](int b, int h, int w, int c) {
                   const int key = tile_of(g, b, h, w);
                   if (key != cur) {
                       flush();
                       cur = key;
                       slot = full ? -1 : hash_insert(keys, key);
                       if (slot < 0) full = 1;
                   }
                   ++len;
                   if (mine < kPerThread)
                       stage[mine * kBlock + threadIdx.x] =
                           (static_cast<unsigned>(slot) << 22) |
                           (static_cast<unsigned>(c) << g.local_bits) | local_of(g, h, w);
                   else
                       spill = 1;
                   ++mine;
               });
    flush();
    __syncthreads();
    if (full) {  // workgroup-uniform: one run per fragment, in the overflow areas
        if (threadIdx.x == 0) {
            src.reset();
            block_runs[blockIdx.x] = 0;
        }
        __syncthreads();
        src.walk([&](int b, int h, int w, int c) {
                       const int ks = tile_of(g, b, h, w) * kBinGroups + group;
                       const int j = atomicAdd(&counters[kEntriesOvf], 1);
                       const int r = atomicAdd(&counters[kRunsOvf], 1);
                       const int k = atomicAdd(&runs_per_slot[ks + 1], 1);
                       if (j < ovf_entries)
                           entries[fixed_e + j] =
                               (static_cast<unsigned>(c) << g.local_bits) | local_of(g, h, w);
                       if (r < ovf_runs)
                           records[fixed_r + r] =
                               BinRecord{ks, k, static_cast<int>(fixed_e + j), 1};
                   });
        return;
    }
    // slot offsets within the region (kHash / kBlock slots per thread)
    constexpr int per_thread = kHash / kBlock;
    int cnt[per_thread];
    int here = 0, used = 0;
#pragma unroll
    for (int j = 0; j < per_thread; ++j) {
        cnt[j] = vals[threadIdx.x * per_thread + j];
        here += cnt[j];
        used += keys[threadIdx.x * per_thread + j] >= 0;
    }
    int total;
    int run = block_excl_scan(here, sh, &total);
#pragma unroll
    for (int j = 0; j < per_thread; ++j) {
        vals[threadIdx.x * per_thread + j] = run;  // now the slot's offset
        run += cnt[j];
    }
    int nrec;
    int rec = block_excl_scan(used, sh, &nrec);  // this thread's first record
    const bool staged = spill == 0;  // workgroup-uniform
    if (threadIdx.x == 0) {
        base_e = staged ? static_cast<int>(blockIdx.x) * kStage
                        : static_cast<int>(fixed_e) + atomicAdd(&counters[kEntriesOvf], total);
        const bool in_fixed = nrec <= kBinRecs;
        base_r = in_fixed ? static_cast<int>(blockIdx.x) * kBinRecs
                          : static_cast<int>(fixed_r) + atomicAdd(&counters[kRunsOvf], nrec);
        block_runs[blockIdx.x] = in_fixed ? nrec : 0;
    }
    __syncthreads();
    const int be = base_e, br = base_r;
    const bool e_fits = staged || static_cast<long long>(be) + total <= fixed_e + ovf_entries;
    // run records; then the slots' keys become per-slot cursors
#pragma unroll
    for (int j = 0; j < per_thread; ++j) {
        const int s = threadIdx.x * per_thread + j;
        const int key = keys[s];
        if (key < 0) continue;
        const int ks = key * kBinGroups + group;
        const int k = atomicAdd(&runs_per_slot[ks + 1], 1);
        const long long r = static_cast<long long>(br) + rec++;
        if (r < fixed_r + ovf_runs) records[r] = BinRecord{ks, k, be + vals[s], cnt[j]};
    }
    if (!e_fits) return;
    if (staged) {
        __syncthreads();
#pragma unroll
        for (int j = 0; j < per_thread; ++j) keys[threadIdx.x * per_thread + j] = 0;
        // The thread's column in registers; after the barrier the stage is
        // free.  Each run of same-slot entries reserves its span at its last
        // entry, a backward pass puts every entry below its run's end in the
        // stage, now in slot order, and the workgroup copies the stage to
        // its region with coalesced 16-byte stores.
        unsigned v[kPerThread];
#pragma unroll
        for (int i = 0; i < kPerThread; ++i) v[i] = i < mine ? stage[i * kBlock + threadIdx.x] : 0u;
        __syncthreads();
        int end[kPerThread];
        int len = 0;
#pragma unroll
        for (int i = 0; i < kPerThread; ++i) {
            end[i] = 0;
            if (i < mine) {
                const unsigned sl = v[i] >> 22;
                ++len;
                const bool last = i + 1 == mine || (i + 1 < kPerThread && (v[i + 1] >> 22) != sl);
                if (last) {  // end > 0: the run holds at least this entry
                    end[i] = vals[sl] + atomicAdd(&keys[sl], len) + len;
                    len = 0;
                }
            }
        }
        int at = 0;
#pragma unroll
        for (int i = kPerThread - 1; i >= 0; --i) {
            if (i < mine) {
                if (end[i]) at = end[i];
                stage[--at] = v[i] & kEntryMask;
            }
        }
        __syncthreads();
        // total <= kStage here; the region is kStage entries from a 16-byte
        // aligned start (be is a multiple of kStage)
        uint4* dst = reinterpret_cast<uint4*>(entries + be);
        const uint4* src = reinterpret_cast<const uint4*>(stage);
        for (int q = threadIdx.x; q < (total + 3) / 4; q += kBlock) dst[q] = src[q];
        return;
    }
    // more fragments than the stage holds: walk again and place directly;
    // the stage words now serve as per-slot cursors
    __syncthreads();
    for (int s = threadIdx.x; s < kHash; s += kBlock) stage[s] = 0u;
    if (threadIdx.x == 0) src.reset();
    __syncthreads();
    cur = -1;
    slot = -1;
    src.walk([&](int b, int h, int w, int c) {
                   const int key = tile_of(g, b, h, w);
                   if (key != cur) {
                       cur = key;
                       slot = hash_find(keys, key);
                   }
                   if (slot < 0) return;
                   const int p = static_cast<int>(atomicAdd(&stage[slot], 1u));
                   entries[be + vals[slot] + p] =
                       (static_cast<unsigned>(c) << g.local_bits) | local_of(g, h, w);
               });
}

// LDS of a stage workgroup (25.6 KB with the walk's deferred list).
#define S2C_STAGE_LDS                                                                  \
    __shared__ int keys[kHash]; /* tile of each slot, then per-slot cursors */         \
    __shared__ int vals[kHash]; /* fragments per slot, then slot offsets */            \
    __shared__ unsigned stage[kStage];                                                 \
    __shared__ int spill, full, base_e, base_r;                                        \
    __shared__ int sh[kBlock / kWave + 1];

// 80 VGPRs (6 waves per SIMD, which the 25.6 KB of LDS also allows)
// instead of the 82 the compiler picks unconstrained (5 waves); no spills.
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(6, 8)))
dda_bin_stage_kernel(long long P, int H, int W, int V, const float* __restrict__ pos, int E,
                     const int* __restrict__ edges, const int* __restrict__ edge2cluster,
                     TileGrid g, int* __restrict__ runs_per_slot, int* __restrict__ counters,
                     int* __restrict__ block_runs, unsigned* __restrict__ entries, int ovf_entries,
                     BinRecord* __restrict__ records, int ovf_runs) {
    S2C_STAGE_LDS
    __shared__ int deferred[kDeferred];
    __shared__ int ndeferred;
    const EdgeSource src{P, H, W, V, pos, E, edges, edge2cluster, deferred, &ndeferred};
    bin_stage(src, g, runs_per_slot, counters, block_runs, entries, ovf_entries, records,
              ovf_runs, keys, vals, stage, spill, full, base_e, base_r, sh);
}

// The stage pass over a fragment list: kStage fragments per workgroup.
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(6, 8)))
list_bin_stage_kernel(long long N, int B, int H, int W, const int* __restrict__ frag_pix,
                      const float* __restrict__ frag_attrs, const int* __restrict__ prim2cluster,
                      unsigned kmax, TileGrid g, int* __restrict__ runs_per_slot,
                      int* __restrict__ counters, int* __restrict__ block_runs,
                      unsigned* __restrict__ entries, int ovf_entries,
                      BinRecord* __restrict__ records, int ovf_runs) {
    S2C_STAGE_LDS
    const ListSource src{N, B, H, W, frag_pix, frag_attrs, prim2cluster, kmax};
    bin_stage(src, g, runs_per_slot, counters, block_runs, entries, ovf_entries, records,
              ovf_runs, keys, vals, stage, spill, full, base_e, base_r, sh);
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

// The binned pass after its tile grid is chosen: workspace, counters,
// stage (stage(plan) launches the stage kernel over plan.blocks
// workgroups), resolve, and the device-side fallback (fallback(plan)
// launches a kernel that ORs every fragment into the zeroed bitset when the
// overflow areas were too small, and exits at once otherwise).
template <class Stage, class Fallback>
int binned_mask(BinPlan& plan, int H, int W, int num_slots, int* bitset, hipStream_t s,
                Stage&& stage, Fallback&& fallback) {
    std::lock_guard<std::mutex> held(BinWorkspace::lock());
    void* work = nullptr;
    if (int r = BinWorkspace::get(s, BinWorkspace::kPlanInts,
                                  sizeof(int) * bin_plan_ints(plan.g, plan.blocks), &work))
        return r;
    bin_carve(plan, static_cast<int*>(work));
    // No host wait: the overflow-area use of the last call on this stream
    // grows the running estimates once its copy has arrived, and a call
    // whose areas are too small finishes with direct ORs on the device.
    BinStreamCounters* cnt = nullptr;
    if (int r = BinWorkspace::counters(s, &cnt)) return r;
    if (cnt->pending) {
        const hipError_t q = hipEventQuery(cnt->done);
        if (q == hipSuccess) {
            long long ne = 0, nr = 0;
            if (bin_overflow_use(cnt->host, &ne, &nr)) {
                bin_entry_hint(ne + ne / 8 + 4 * kStage);
                bin_run_hint(nr + nr / 8 + 4 * kHash);
            }
            cnt->pending = false;
        } else if (q != hipErrorNotReady) {
            S2C_HIP(q);
        }
    }
    const long long want_e = dmin(bin_entry_hint(4 * kStage), static_cast<long long>(INT_MAX / 2));
    const long long want_r = dmin(bin_run_hint(4 * kHash), static_cast<long long>(INT_MAX / 2));
    plan.ovf_entries = static_cast<int>(want_e);
    plan.ovf_runs = static_cast<int>(want_r);
    void *ebuf = nullptr, *rbuf = nullptr, *obuf = nullptr;
    if (int r = BinWorkspace::get(
            s, BinWorkspace::kEntries,
            sizeof(unsigned) * static_cast<size_t>(plan.fixed_entries() + want_e), &ebuf))
        return r;
    if (int r = BinWorkspace::get(
            s, BinWorkspace::kRecords,
            sizeof(BinRecord) * static_cast<size_t>(plan.fixed_runs() + want_r), &rbuf))
        return r;
    if (int r = BinWorkspace::get(
            s, BinWorkspace::kRuns,
            sizeof(BinRun) * static_cast<size_t>(plan.fixed_runs() + want_r), &obuf))
        return r;
    plan.entries = static_cast<unsigned*>(ebuf);
    plan.records = static_cast<BinRecord*>(rbuf);
    plan.runs = static_cast<BinRun*>(obuf);
    if (int r = bin_clear(plan, s)) return r;
    stage(plan);
    S2C_LAUNCH_CHECK();
    if (int r = bin_counters_async(plan, cnt->host, s)) return r;
    S2C_HIP(hipEventRecord(cnt->done, s));
    cnt->pending = true;
    if (int r = bin_resolve(plan, H, W, num_slots, bitset, s)) return r;
    fallback(plan);
    S2C_LAUNCH_CHECK();
    return S2C_OK;
}

unsigned fallback_grid(long long blocks) {
    return static_cast<unsigned>(dmin(blocks, static_cast<long long>(kFallbackGrid)));
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
              "dda_cluster_mask: the requested tile bitset does not fit LDS");
    if (!tiled) {
        // Empty input, or a tile bitset too large for LDS: clear + direct ORs.
        if (int r = fill_u32(bitset, 0u, words, s)) return r;
        if (P == 0) return S2C_OK;
        dda_mask_direct_kernel<<<static_cast<unsigned>(plan.blocks), kBlock, 0, s>>>(
            P, H, W, V, pos, E, edges, edge2cluster, num_slots, bitset);
        S2C_LAUNCH_CHECK();
        return S2C_OK;
    }
    S2C_CHECK(plan.blocks * kStage < INT_MAX / 2 && plan.blocks * kBinRecs < INT_MAX / 2,
              "dda_cluster_mask: too many primitives for 32-bit entry offsets");
    return binned_mask(
        plan, H, W, num_slots, bitset, s,
        [&](const BinPlan& p) {
            dda_bin_stage_kernel<<<static_cast<unsigned>(p.blocks), kBlock, 0, s>>>(
                P, H, W, V, pos, E, edges, edge2cluster, p.g, p.runs_per_slot, p.counters,
                p.block_runs, p.entries, p.ovf_entries, p.records, p.ovf_runs);
        },
        [&](const BinPlan& p) {
            dda_mask_fallback_kernel<<<fallback_grid(p.blocks), kBlock, 0, s>>>(
                P, H, W, V, pos, E, edges, edge2cluster, num_slots, bitset, p.counters,
                p.ovf_entries, p.ovf_runs, p.blocks);
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
    plan.blocks = (N + kStage - 1) / kStage;
    const bool tiled = N > 0 && choose_tiles(B, H, W, num_slots, 4096, plan.g);
    if (!tiled) {
        if (int r = fill_u32(bitset, 0u, words, s)) return r;
        if (N == 0) return S2C_OK;
        list_mask_direct_kernel<<<grid_for(N, kBlock), kBlock, 0, s>>>(
            N, B, H, W, frag_pix, frag_attrs, prim2cluster, kmax, num_slots, bitset);
        S2C_LAUNCH_CHECK();
        return S2C_OK;
    }
    S2C_CHECK(plan.blocks * kStage < INT_MAX / 2 && plan.blocks * kBinRecs < INT_MAX / 2,
              "fragment_cluster_mask: too many fragments for 32-bit entry offsets");
    return binned_mask(
        plan, H, W, num_slots, bitset, s,
        [&](const BinPlan& p) {
            list_bin_stage_kernel<<<static_cast<unsigned>(p.blocks), kBlock, 0, s>>>(
                N, B, H, W, frag_pix, frag_attrs, prim2cluster, kmax, p.g, p.runs_per_slot,
                p.counters, p.block_runs, p.entries, p.ovf_entries, p.records, p.ovf_runs);
        },
        [&](const BinPlan& p) {
            list_mask_fallback_kernel<<<fallback_grid(p.blocks), kBlock, 0, s>>>(
                N, B, H, W, frag_pix, frag_attrs, prim2cluster, kmax, num_slots, bitset,
                p.counters, p.ovf_entries, p.ovf_runs);
        });
}

extern "C" int s2c_dda_cluster_mask_reset_estimates(void) {
    std::lock_guard<std::mutex> held(BinWorkspace::lock());
    bin_reset_hints();
    return S2C_OK;
}

extern "C" int s2c_dda_cluster_mask_release(void* stream) {
    std::lock_guard<std::mutex> held(BinWorkspace::lock());
    return BinWorkspace::release(static_cast<hipStream_t>(stream));
}
