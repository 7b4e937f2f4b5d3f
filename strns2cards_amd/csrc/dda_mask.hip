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
// edges.  `deferred` / `ndeferred` are LDS scratch (kPrimsPerBlock ints + 1,
// ndeferred zeroed and synchronized by the caller).  Contains
// __syncthreads(): every thread of the workgroup must call it.
template <class Emit>
__device__ __forceinline__ void walk_edges(long long P, int H, int W, int V,
                                           const float* __restrict__ pos, int E,
                                           const int* __restrict__ edges,
                                           const int* __restrict__ edge2cluster, int* deferred,
                                           int* ndeferred, Emit&& emit) {
    const long long q0 = static_cast<long long>(blockIdx.x) * kPrimsPerBlock +
                         static_cast<long long>(threadIdx.x) * kPrimsPerThread;
#pragma unroll 1
    for (int i = 0; i < kPrimsPerThread; ++i) {
        const long long q = q0 + i;
        if (q >= P) break;
        const EdgeView v = load_edge(q, H, W, V, pos, E, edges, edge2cluster);
        const int n = dmin(v.cnt, kSerialFrags);
        for (int k = 0; k < n; ++k) {
            int h, w;
            if (dda_mask_frag(H, W, v.sp, k, v.p0, v.p1, v.depth_ok, h, w)) emit(v.b, h, w, v.cluster);
        }
        if (v.cnt > kSerialFrags)
            deferred[atomicAdd(ndeferred, 1)] =
                static_cast<int>(q - static_cast<long long>(blockIdx.x) * kPrimsPerBlock);
    }
    __syncthreads();
    const int nd = *ndeferred;
    const long long qb = static_cast<long long>(blockIdx.x) * kPrimsPerBlock;
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
// tile bitsets that do not fit LDS.
__global__ void __launch_bounds__(kBlock)
dda_mask_direct_kernel(long long P, int H, int W, int V, const float* __restrict__ pos, int E,
                       const int* __restrict__ edges, const int* __restrict__ edge2cluster, int S,
                       int* __restrict__ bitset) {
    __shared__ int deferred[kPrimsPerBlock];
    __shared__ int ndeferred;
    if (threadIdx.x == 0) ndeferred = 0;
    __syncthreads();
    walk_edges(P, H, W, V, pos, E, edges, edge2cluster, deferred, &ndeferred,
               [&](int b, int h, int w, int c) {
                   const long long idx = ((static_cast<long long>(b) * H + h) * W + w) * S + (c >> 5);
                   atomicOr(&bitset[idx], static_cast<int>(1u << (c & 31)));
               });
}

// Binned form, pass 1: per-tile counts in an LDS table, one global atomic
// per (workgroup, tile); the table is saved as the workgroup's image.
__global__ void __launch_bounds__(kBlock)
dda_bin_count_kernel(long long P, int H, int W, int V, const float* __restrict__ pos, int E,
                     const int* __restrict__ edges, const int* __restrict__ edge2cluster,
                     TileGrid g, int* __restrict__ counts, int* __restrict__ image,
                     int* __restrict__ flags) {
    __shared__ int keys[kHash];
    __shared__ int vals[kHash];
    __shared__ int deferred[kPrimsPerBlock];
    __shared__ int ndeferred, overflow;
    for (int s = threadIdx.x; s < kHash; s += kBlock) {
        keys[s] = -1;
        vals[s] = 0;
    }
    if (threadIdx.x == 0) ndeferred = overflow = 0;
    __syncthreads();
    int cur = -1, len = 0;
    auto flush = [&]() {
        if (cur < 0 || overflow) return;
        const int s = hash_insert(keys, cur);
        if (s < 0) overflow = 1;
        else atomicAdd(&vals[s], len);
    };
    walk_edges(P, H, W, V, pos, E, edges, edge2cluster, deferred, &ndeferred,
               [&](int b, int h, int w, int) {
                   const int key = tile_of(g, b, h, w);
                   if (key == cur) {
                       ++len;
                   } else {
                       flush();
                       cur = key;
                       len = 1;
                   }
               });
    flush();
    __syncthreads();
    int* img = image + static_cast<long long>(blockIdx.x) * kImageInts;
    if (overflow) {  // workgroup-uniform: recount into the tiles' overflow slots
        if (threadIdx.x == 0) ndeferred = 0;
        __syncthreads();
        walk_edges(P, H, W, V, pos, E, edges, edge2cluster, deferred, &ndeferred,
                   [&](int b, int h, int w, int) {
                       atomicAdd(&counts[overflow_slot(tile_of(g, b, h, w)) + 1], 1);
                   });
        if (threadIdx.x == 0) flags[blockIdx.x] = 1;
        return;
    }
    for (int s = threadIdx.x; s < kHash; s += kBlock) {
        const int key = keys[s];
        int rel = 0;
        if (key >= 0) rel = atomicAdd(&counts[group_slot(key) + 1], vals[s]);
        img[s] = key;
        img[kHash + s] = rel;
        img[2 * kHash + s] = key >= 0 ? vals[s] : 0;
    }
    if (threadIdx.x == 0) flags[blockIdx.x] = 0;
}

// Binned form, pass 2: place every fragment's packed entry.  The
// workgroup's entries are staged in LDS grouped by tile (segment sizes come
// from the count pass's image) and then copied out as contiguous runs, so
// the global writes are coalesced; a workgroup with more entries than the
// stage holds writes each entry directly.
constexpr int kStage = 4096;

__global__ void __launch_bounds__(kBlock)
dda_bin_fill_kernel(long long P, int H, int W, int V, const float* __restrict__ pos, int E,
                    const int* __restrict__ edges, const int* __restrict__ edge2cluster,
                    TileGrid g, const int* __restrict__ offs, int* __restrict__ cursor,
                    const int* __restrict__ image, const int* __restrict__ flags,
                    unsigned* __restrict__ entries) {
    __shared__ int keys[kHash];
    __shared__ int base[kHash];
    __shared__ int used[kHash];
    __shared__ int soff[kHash + 1];
    __shared__ unsigned stage[kStage];
    __shared__ int deferred[kPrimsPerBlock];
    __shared__ int ndeferred;
    __shared__ int sh[kBlock / kWave + 1];
    const bool ovf = flags[blockIdx.x] != 0;
    const int* img = image + static_cast<long long>(blockIdx.x) * kImageInts;
    constexpr int per_thread = kHash / kBlock;
    int cnt[per_thread];
    int mine = 0;
#pragma unroll
    for (int j = 0; j < per_thread; ++j) {
        const int s = threadIdx.x * per_thread + j;
        const int key = ovf ? -1 : img[s];
        keys[s] = key;
        base[s] = key >= 0 ? offs[group_slot(key)] + img[kHash + s] : 0;
        used[s] = 0;
        cnt[j] = key >= 0 ? img[2 * kHash + s] : 0;
        mine += cnt[j];
    }
    int total;
    int run = block_excl_scan(mine, sh, &total);
#pragma unroll
    for (int j = 0; j < per_thread; ++j) {
        soff[threadIdx.x * per_thread + j] = run;
        run += cnt[j];
    }
    if (threadIdx.x == 0) {
        soff[kHash] = total;
        ndeferred = 0;
    }
    const bool staged = !ovf && total <= kStage;  // workgroup-uniform
    __syncthreads();
    int cur = -1, slot = -1;
    walk_edges(P, H, W, V, pos, E, edges, edge2cluster, deferred, &ndeferred,
               [&](int b, int h, int w, int c) {
                   const int key = tile_of(g, b, h, w);
                   const unsigned v = (static_cast<unsigned>(c) << g.local_bits) | local_of(g, h, w);
                   if (ovf) {
                       entries[atomicAdd(&cursor[overflow_slot(key)], 1)] = v;
                       return;
                   }
                   if (key != cur) {
                       cur = key;
                       slot = hash_find(keys, key);
                   }
                   if (slot < 0) return;
                   const int k = atomicAdd(&used[slot], 1);
                   if (staged) stage[soff[slot] + k] = v;
                   else entries[base[slot] + k] = v;
               });
    if (!staged) return;
    __syncthreads();
    for (int p = threadIdx.x; p < total; p += kBlock) {
        int lo = 0, hi = kHash;  // last slot with soff[slot] <= p
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (soff[mid] <= p) lo = mid;
            else hi = mid;
        }
        entries[base[lo] + (p - soff[lo])] = stage[p];
    }
}

}  // namespace s2c

using namespace s2c;

extern "C" int s2c_dda_cluster_mask(int B, int H, int W, int V, const float* pos, int E,
                                    const int* edges, const int* edge2cluster, int num_slots,
                                    int* bitset, void* stream) {
    S2C_CHECK(B >= 0 && H >= 0 && W >= 0 && V >= 0 && E >= 0 && num_slots >= 0,
              "dda_cluster_mask: negative size");
    hipStream_t s = static_cast<hipStream_t>(stream);
    const long long words = static_cast<long long>(B) * H * W * num_slots;
    const long long P = static_cast<long long>(B) * E;
    if (words == 0) return S2C_OK;
    BinPlan plan;
    plan.blocks = (P + kPrimsPerBlock - 1) / kPrimsPerBlock;
    if (P == 0 || !choose_tiles(B, H, W, num_slots, 4096, plan.g)) {
        // Empty input, or a tile bitset too large for LDS: clear + direct ORs.
        if (int r = fill_u32(bitset, 0u, words, s)) return r;
        if (P == 0) return S2C_OK;
        dda_mask_direct_kernel<<<static_cast<unsigned>(plan.blocks), kBlock, 0, s>>>(
            P, H, W, V, pos, E, edges, edge2cluster, num_slots, bitset);
        S2C_LAUNCH_CHECK();
        return S2C_OK;
    }
    Temp work;
    if (int r = work.alloc(sizeof(int) * bin_plan_ints(plan.g, plan.blocks), s)) return r;
    if (int r = bin_prepare(plan, work.as<int>(), s)) return r;
    const unsigned grid = static_cast<unsigned>(plan.blocks);
    dda_bin_count_kernel<<<grid, kBlock, 0, s>>>(P, H, W, V, pos, E, edges, edge2cluster, plan.g,
                                                 plan.counts, plan.image, plan.flags);
    S2C_LAUNCH_CHECK();
    int total = 0;
    if (int r = bin_offsets(plan, s, &total)) return r;
    Temp entries;
    if (int r = entries.alloc(sizeof(unsigned) * static_cast<size_t>(total), s)) return r;
    dda_bin_fill_kernel<<<grid, kBlock, 0, s>>>(P, H, W, V, pos, E, edges, edge2cluster, plan.g,
                                                plan.counts, plan.cursor, plan.image, plan.flags,
                                                entries.as<unsigned>());
    S2C_LAUNCH_CHECK();
    return bin_resolve(plan, entries.as<unsigned>(), H, W, num_slots, bitset, s);
}
