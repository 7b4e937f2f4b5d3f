// Record ordering and the per-tile LDS resolve of the binned cluster-bitset
// pass (see binning.hpp).
#include "binning.hpp"

#include <map>
#include <utility>

namespace s2c {

// A tile's resolve is a chain of dependent global round trips (order,
// offsets, records, masks, then the stores), so the kernel is bound by how
// many tiles are in flight: 512-thread workgroups let four tiles share a CU
// (LDS-bound at 32 KB of bitset each), and each thread keeps
// kResolveUnroll mask loads in flight.
constexpr int kResolveThreads = 512;
constexpr int kResolveUnroll = 4;

// ------------------------------------------------ counting sort by tile

// Records of sort block j's stage workgroups: wave w of the block takes
// workgroups first + w, first + w + waves, ..., lane l records l, l + 64, ...
// of each; f(r, (tile, info)) for every record r.
template <class F>
__device__ __forceinline__ void sort_records(long long blocks, long long j,
                                             const int* __restrict__ block_recs,
                                             const int2* __restrict__ rec_ti, F&& f) {
    const long long b1 = dmin((j + 1) * kSortChunk, blocks);
    for (long long b = j * kSortChunk + threadIdx.x / kWave; b < b1; b += kSortThreads / kWave) {
        const int n = block_recs[b];
        const long long r0 = b * kBinRecs;
#pragma unroll 1
        for (int i = lane_id(); i < n; i += 4 * kWave) {
            int2 ti[4];
#pragma unroll
            for (int u = 0; u < 4; ++u)
                ti[u] = i + u * kWave < n ? rec_ti[r0 + i + u * kWave] : make_int2(-1, 0);
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (ti[u].x >= 0) f(r0 + i + u * kWave, ti[u]);
        }
    }
}

// counts[j * nt + t] = records of sort block j in tile t (LDS counters,
// kSortTiles tiles per round; `cnt` is dynamic LDS of min(nt, kSortTiles)
// ints).
__global__ void __launch_bounds__(kSortThreads)
bin_count_kernel(long long nt, long long blocks, const int* __restrict__ block_recs,
                 const int2* __restrict__ rec_ti, int* __restrict__ counts) {
    extern __shared__ int cnt[];
    for (long long t0 = 0; t0 < nt; t0 += kSortTiles) {
        const int nr = static_cast<int>(dmin(static_cast<long long>(kSortTiles), nt - t0));
        for (int i = threadIdx.x; i < nr; i += kSortThreads) cnt[i] = 0;
        __syncthreads();
        sort_records(blocks, blockIdx.x, block_recs, rec_ti, [&](long long, int2 ti) {
            if (ti.x >= t0 && ti.x < t0 + nr) atomicAdd(&cnt[ti.x - t0], 1);
        });
        __syncthreads();
        int* row = counts + blockIdx.x * nt + t0;
        for (int i = threadIdx.x; i < nr; i += kSortThreads) row[i] = cnt[i];
        __syncthreads();
    }
}

// Rows [j0, j1) of the count matrix's segment `seg`.
__device__ __forceinline__ void sort_segment(long long nb, int seg, long long& j0, long long& j1) {
    const long long per = (nb + kSortSegs - 1) / kSortSegs;
    j0 = dmin(seg * per, nb);
    j1 = dmin(j0 + per, nb);
}

// segs[seg * nt + t] = sum of column t over the segment's rows.  Grid
// (tiles / kBlock, kSortSegs).
__global__ void __launch_bounds__(kBlock)
bin_colsum_kernel(long long nt, long long nb, const int* __restrict__ counts,
                  int* __restrict__ segs) {
    const long long t = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
    if (t >= nt) return;
    long long j0, j1;
    sort_segment(nb, blockIdx.y, j0, j1);
    int acc = 0;
    for (long long j = j0; j < j1; ++j) acc += counts[j * nt + t];
    segs[blockIdx.y * nt + t] = acc;
}

// Resolve order class of a tile with w records: powers of 4, heaviest
// first, empty tiles last.
__device__ __forceinline__ int order_class(int w) {
    return w <= 0 ? kOrderClasses - 1 : dmax(0, kOrderClasses - 2 - (31 - __clz(w)) / 2);
}

// In place: counts[j * nt + t] becomes the number of records of tile t in
// sort blocks before j.  The last segment writes the tile's total into
// offs[t], the sum of its block's totals into bsums[blockIdx.x], and its
// block's tiles per order class into classes[k * gridDim.x + blockIdx.x].
__global__ void __launch_bounds__(kBlock)
bin_colprefix_kernel(long long nt, long long nb, int* __restrict__ counts,
                     const int* __restrict__ segs, int* __restrict__ offs,
                     int* __restrict__ bsums, int* __restrict__ classes) {
    __shared__ int hist[kOrderClasses];
    __shared__ int sh[kBlock / kWave + 1];
    if (blockIdx.y != kSortSegs - 1) {
        const long long t = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
        if (t >= nt) return;
        long long j0, j1;
        sort_segment(nb, blockIdx.y, j0, j1);
        int acc = 0;
        for (int k = 0; k < static_cast<int>(blockIdx.y); ++k) acc += segs[k * nt + t];
        for (long long j = j0; j < j1; ++j) {
            const int c = counts[j * nt + t];
            counts[j * nt + t] = acc;
            acc += c;
        }
        return;
    }
    if (threadIdx.x < kOrderClasses) hist[threadIdx.x] = 0;
    __syncthreads();
    const long long t = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
    int acc = 0;
    if (t < nt) {
        long long j0, j1;
        sort_segment(nb, blockIdx.y, j0, j1);
        for (int k = 0; k < kSortSegs - 1; ++k) acc += segs[k * nt + t];
        for (long long j = j0; j < j1; ++j) {
            const int c = counts[j * nt + t];
            counts[j * nt + t] = acc;
            acc += c;
        }
        offs[t] = acc;
        atomicAdd(&hist[order_class(acc)], 1);
    }
    int total;
    block_excl_scan(acc, sh, &total);  // contains the barrier before hist is read
    if (threadIdx.x < kOrderClasses) classes[threadIdx.x * gridDim.x + blockIdx.x] = hist[threadIdx.x];
    if (threadIdx.x == 0) bsums[blockIdx.x] = total;
}

// Exclusive scans in place by one workgroup of the n class counts
// (class-major, so each (class, block) gets the first order position of its
// tiles) and of the nblk block sums of tile totals.
__global__ void __launch_bounds__(kWgScanThreads)
bin_scan_kernel(long long n, int* __restrict__ classes, long long nblk, int* __restrict__ bsums) {
    wg_exclusive_scan(classes, n);
    __syncthreads();
    wg_exclusive_scan(bsums, nblk);
}

// Tile offsets and the resolve order: offs[t] becomes the first record of
// tile t (its block's offset plus a block scan of the totals; offs[nt] the
// record count), and (tile, first record, record end) goes to its class's
// next position (order within a class arbitrary).  Same grid as the class
// counts.
__global__ void __launch_bounds__(kBlock)
bin_order_kernel(long long nt, int* __restrict__ offs, const int* __restrict__ bsums,
                 const int* __restrict__ classes, int4* __restrict__ order) {
    __shared__ int next[kOrderClasses];
    __shared__ int sh[kBlock / kWave + 1];
    if (threadIdx.x < kOrderClasses) next[threadIdx.x] = classes[threadIdx.x * gridDim.x + blockIdx.x];
    const long long t = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
    const int w = t < nt ? offs[t] : 0;
    int total;
    const int a = bsums[blockIdx.x] + block_excl_scan(w, sh, &total);  // barrier inside
    if (t >= nt) return;
    offs[t] = a;
    if (t == nt - 1) offs[nt] = a + w;
    const int pos = atomicAdd(&next[order_class(w)], 1);
    order[pos] = make_int4(static_cast<int>(t), a, a + w, 0);
}

// Sort block j places its records at offs[t] + (its first position in tile
// t) + (rank within the block, from an LDS counter), as (record, info).
__global__ void __launch_bounds__(kSortThreads)
bin_place_kernel(long long nt, long long blocks, const int* __restrict__ block_recs,
                 const int2* __restrict__ rec_ti, const int* __restrict__ firsts,
                 const int* __restrict__ offs, int2* __restrict__ runs) {
    extern __shared__ int cnt[];
    const int* row = firsts + blockIdx.x * nt;
    for (long long t0 = 0; t0 < nt; t0 += kSortTiles) {
        const int nr = static_cast<int>(dmin(static_cast<long long>(kSortTiles), nt - t0));
        for (int i = threadIdx.x; i < nr; i += kSortThreads) cnt[i] = 0;
        __syncthreads();
        sort_records(blocks, blockIdx.x, block_recs, rec_ti, [&](long long r, int2 ti) {
            if (ti.x < t0 || ti.x >= t0 + nr) return;
            const int rank = atomicAdd(&cnt[ti.x - t0], 1);
            runs[offs[ti.x] + row[ti.x] + rank] = make_int2(static_cast<int>(r), ti.y);
        });
        __syncthreads();
    }
}

// One workgroup per tile, in the resolve order.  Thread i takes word
// i % mask_words of record i / mask_words of each pass, so a record's mask
// is read with consecutive loads, and ORs the cluster's bit into the LDS
// bitset at every pixel the word sets.  The LDS word of (pixel l, word j)
// is l*S + (j ^ (l & smask)): with S a power of two, the pixels of one
// cluster's mask (same j) fall in different banks.
template <bool VEC4, int kLds, bool ONE_CELL>
__global__ void __launch_bounds__(kResolveThreads)
cell_resolve_kernel(TileGrid g, const int4* __restrict__ order, const int2* __restrict__ runs,
                    const uint4* __restrict__ rec_mask, int H, int W, int S,
                    int* __restrict__ bitset) {
    __shared__ __align__(16) unsigned tile[kLds / 4];
    const int4 ord = order[blockIdx.x];
    const int t = ord.x, r0 = ord.y, r1 = ord.z;
    const int side = 1 << g.ts_log2;
    const int words = side * side * S;  // a multiple of 4 (side >= 4)
    for (int i = threadIdx.x; i < words / 4; i += kResolveThreads)
        reinterpret_cast<uint4*>(tile)[i] = make_uint4(0u, 0u, 0u, 0u);
    __syncthreads();  // a tile without records goes straight to the write-out
    const unsigned smask = (S & (S - 1)) == 0 ? static_cast<unsigned>(S) - 1u : 0u;
    const int mwl = g.mw_log2;
    const int word = static_cast<int>(threadIdx.x) & ((1 << mwl) - 1);
    const int per = kResolveThreads >> mwl;  // records per pass
    const int cs = g.cs_log2;
    const int cpl = g.ts_log2 - cs;  // log2 of cells per tile row
    for (int rb = r0 + static_cast<int>(threadIdx.x >> mwl); rb < r1; rb += per * kResolveUnroll) {
        unsigned m[kResolveUnroll];
        int info[kResolveUnroll];
#pragma unroll
        for (int u = 0; u < kResolveUnroll; ++u) {
            const int r = rb + u * per;
            m[u] = 0u;
            info[u] = 0;
            if (r < r1) {
                const int2 run = runs[r];
                m[u] = reinterpret_cast<const unsigned*>(rec_mask + 2ll * run.x)[word];
                info[u] = run.y;
            }
        }
#pragma unroll
        for (int u = 0; u < kResolveUnroll; ++u) {
            unsigned v = m[u];
            if (v == 0u) continue;
            const unsigned c = static_cast<unsigned>(info[u]) & 0xFFFFu;
            const int cell = info[u] >> 16;
            const int oy = (cell >> cpl) << cs;
            const int ox = (cell & ((1 << cpl) - 1)) << cs;
            const unsigned bit = 1u << (c & 31u);
            const unsigned j = c >> 5;
            do {
                const unsigned p =  // pixel in the cell
                    cell_pixel(cs, static_cast<unsigned>(word), static_cast<unsigned>(__ffs(v) - 1));
                v &= v - 1u;
                // with one cell per tile the cell's pixel is the tile's
                const unsigned l = ONE_CELL ? p
                                            : static_cast<unsigned>(((oy + (p >> cs)) << g.ts_log2) |
                                                                    (ox + (p & ((1 << cs) - 1))));
                atomicOr(&tile[__umul24(l, static_cast<unsigned>(S)) + (j ^ (l & smask))], bit);
            } while (v != 0u);
        }
    }
    __syncthreads();
    const int per_b = g.tx * g.ty;
    const int b = t / per_b;
    const int rr = t - b * per_b;
    const int y0 = (rr / g.tx) * side;
    const int x0 = (rr % g.tx) * side;
    const int row_w = dmin(side, W - x0);  // pixels of this tile inside the image
    const int rows = dmin(side, H - y0);
    if (VEC4 && row_w == side && rows == side && (S & (S - 1)) == 0) {
        // whole tile, S a power of two: the index math is shifts and masks
        // (the general form below divides twice per store)
        const int lq = __ffs(S / 4) - 1;  // log2 of uint4 per pixel
        const int lr = g.ts_log2 + lq;    // log2 of uint4 per tile row
        const long long row0 = (static_cast<long long>(b) * H + y0) * W + x0;
        for (int i = threadIdx.x; i < (side << lr); i += kResolveThreads) {
            const int ly = i >> lr;
            const int rem = i & ((1 << lr) - 1);
            const unsigned l = static_cast<unsigned>((ly << g.ts_log2) | (rem >> lq));
            const int j0 = (rem & ((1 << lq) - 1)) << 2;
            const unsigned* lw = tile + l * S;
            const unsigned sw = l & smask;
            const uint4 v = make_uint4(lw[(j0 + 0) ^ sw], lw[(j0 + 1) ^ sw], lw[(j0 + 2) ^ sw],
                                       lw[(j0 + 3) ^ sw]);
            const long long dst = ((row0 + static_cast<long long>(ly) * W) << lq) + rem;
            reinterpret_cast<uint4*>(bitset)[dst] = v;
        }
    } else if (VEC4) {
        const int q = S / 4;
        const int per_row = row_w * q;
        for (int i = threadIdx.x; i < rows * per_row; i += kResolveThreads) {
            const int ly = i / per_row;
            const int rem = i - ly * per_row;
            const int px = rem / q;  // pixel within the row
            const unsigned l = static_cast<unsigned>(ly * side + px);
            const int j0 = (rem - px * q) * 4;
            const unsigned* lw = tile + l * S;
            const uint4 v = make_uint4(lw[(j0 + 0) ^ (l & smask)], lw[(j0 + 1) ^ (l & smask)],
                                       lw[(j0 + 2) ^ (l & smask)], lw[(j0 + 3) ^ (l & smask)]);
            const long long dst =
                ((static_cast<long long>(b) * H + y0 + ly) * W + x0) * q + rem;
            reinterpret_cast<uint4*>(bitset)[dst] = v;
        }
    } else {
        const int per_row = row_w * S;
        for (int i = threadIdx.x; i < rows * per_row; i += kResolveThreads) {
            const int ly = i / per_row;
            const int rem = i - ly * per_row;
            const int px = rem / S;
            const unsigned l = static_cast<unsigned>(ly * side + px);
            const int j = rem - px * S;
            const long long dst = ((static_cast<long long>(b) * H + y0 + ly) * W + x0) * S + rem;
            bitset[dst] = static_cast<int>(tile[l * S + (j ^ (l & smask))]);
        }
    }
}

size_t bin_plan_ints(const TileGrid& g, long long blocks) {
    BinPlan p;
    p.g = g;
    p.blocks = blocks;
    return static_cast<size_t>(4 * g.nt + (g.nt + 1) + blocks + kSortSegs * g.nt +
                               (kOrderClasses + 1) * p.class_blocks() + 2 +
                               p.sort_blocks() * g.nt + 2 * p.records() + 8);
}

void bin_carve(BinPlan& p, int* ints) {
    const long long nt = p.g.nt;
    long long off = 0;
    p.order = reinterpret_cast<int4*>(ints);  // 16-byte aligned
    off += 4 * nt;
    p.offs = ints + off;
    off += nt + 1;
    p.block_recs = ints + off;
    off += p.blocks;
    p.segs = ints + off;
    off += kSortSegs * nt;
    p.classes = ints + off;
    off += kOrderClasses * p.class_blocks() + 1;
    p.bsums = ints + off;
    off += p.class_blocks() + 1;
    p.counts = ints + off;
    off += p.sort_blocks() * nt;
    off = (off + 1) & ~1ll;
    p.rec_ti = reinterpret_cast<int2*>(ints + off);  // 8-byte aligned
}

int bin_resolve(const BinPlan& p, int H, int W, int S, int* bitset, hipStream_t s) {
    const long long nt = p.g.nt;
    const long long nb = p.sort_blocks();
    const size_t lds = sizeof(int) * static_cast<size_t>(dmin(nt, static_cast<long long>(kSortTiles)));
    bin_count_kernel<<<static_cast<unsigned>(nb), kSortThreads, lds, s>>>(
        nt, p.blocks, p.block_recs, p.rec_ti, p.counts);
    S2C_LAUNCH_CHECK();
    const unsigned cb = static_cast<unsigned>(p.class_blocks());
    bin_colsum_kernel<<<dim3(cb, kSortSegs), kBlock, 0, s>>>(nt, nb, p.counts, p.segs);
    S2C_LAUNCH_CHECK();
    bin_colprefix_kernel<<<dim3(cb, kSortSegs), kBlock, 0, s>>>(nt, nb, p.counts, p.segs, p.offs,
                                                                p.bsums, p.classes);
    S2C_LAUNCH_CHECK();
    bin_scan_kernel<<<1, kWgScanThreads, 0, s>>>(kOrderClasses * p.class_blocks(), p.classes,
                                                 p.class_blocks(), p.bsums);
    S2C_LAUNCH_CHECK();
    bin_order_kernel<<<cb, kBlock, 0, s>>>(nt, p.offs, p.bsums, p.classes, p.order);
    S2C_LAUNCH_CHECK();
    bin_place_kernel<<<static_cast<unsigned>(nb), kSortThreads, lds, s>>>(
        nt, p.blocks, p.block_recs, p.rec_ti, p.counts, p.offs, p.runs);
    S2C_LAUNCH_CHECK();
    const bool vec = (S % 4 == 0) && ((reinterpret_cast<uintptr_t>(bitset) & 15) == 0);
    const bool small = tile_bitset_bytes(p.g.ts_log2, S) <= kTileLdsBytes;
    const bool one = p.g.cs_log2 == p.g.ts_log2;
    const unsigned grid_t = static_cast<unsigned>(nt);
#define S2C_RESOLVE(V, L, O)                                                      \
    cell_resolve_kernel<V, L, O><<<grid_t, kResolveThreads, 0, s>>>(p.g, p.order, p.runs, \
                                                                  p.rec_mask, H, W, S, bitset)
    if (vec && small) {
        if (one) S2C_RESOLVE(true, kTileLdsBytes, true);
        else S2C_RESOLVE(true, kTileLdsBytes, false);
    } else if (vec) {
        if (one) S2C_RESOLVE(true, kTileLdsMax, true);
        else S2C_RESOLVE(true, kTileLdsMax, false);
    } else if (small) {
        if (one) S2C_RESOLVE(false, kTileLdsBytes, true);
        else S2C_RESOLVE(false, kTileLdsBytes, false);
    } else {
        if (one) S2C_RESOLVE(false, kTileLdsMax, true);
        else S2C_RESOLVE(false, kTileLdsMax, false);
    }
#undef S2C_RESOLVE
    S2C_LAUNCH_CHECK();
    return S2C_OK;
}

namespace {
struct WorkspaceBuffers {
    void* p[BinWorkspace::kSlots] = {};
    size_t n[BinWorkspace::kSlots] = {};
};
std::map<std::pair<int, hipStream_t>, WorkspaceBuffers>& workspaces() {
    static std::map<std::pair<int, hipStream_t>, WorkspaceBuffers> m;
    return m;
}
}  // namespace

std::mutex& BinWorkspace::lock() {
    static std::mutex mu;
    return mu;
}

int BinWorkspace::get(hipStream_t s, Slot slot, size_t bytes, void** out) {
    int dev = 0;
    S2C_HIP(hipGetDevice(&dev));
    WorkspaceBuffers& w = workspaces()[{dev, s}];
    if (w.n[slot] < bytes) {
        if (w.p[slot]) {
            // earlier calls on this stream may still read the old block
            S2C_HIP(hipStreamSynchronize(s));
            S2C_HIP(hipFree(w.p[slot]));
        }
        w.p[slot] = nullptr;
        w.n[slot] = 0;
        const size_t grown = bytes + bytes / 8;  // headroom for the next, larger call
        S2C_HIP(hipMalloc(&w.p[slot], grown));
        w.n[slot] = grown;
    }
    *out = w.p[slot];
    return S2C_OK;
}

int BinWorkspace::release(hipStream_t s) {
    int dev = 0;
    S2C_HIP(hipGetDevice(&dev));
    auto& m = workspaces();
    auto it = m.find({dev, s});
    if (it == m.end()) return S2C_OK;
    S2C_HIP(hipStreamSynchronize(s));
    for (int i = 0; i < kSlots; ++i) {
        if (it->second.p[i]) S2C_HIP(hipFree(it->second.p[i]));
        it->second.p[i] = nullptr;
        it->second.n[i] = 0;
    }
    m.erase(it);
    return S2C_OK;
}

}  // namespace s2c
