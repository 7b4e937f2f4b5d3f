// Screen-tile binning shared by the fused primitive -> cluster-bitset passes.
//
// Fragments are counted and scattered per (view, tile) as packed 32-bit
// entries (cluster << local_bits | local pixel); one workgroup per tile then
// ORs its entries into an LDS copy of the tile's bitset and writes the tile
// once with coalesced stores.  Within a wave, consecutive fragment slots
// mostly fall in the same tile (they walk along a primitive), so counts and
// cursors are advanced with one atomic per run of equal tiles instead of one
// per fragment.
#pragma once

#include "common.hpp"

namespace s2c {

constexpr int kTileLdsBytes = 32768;  // LDS bitset per tile workgroup

struct TileGrid {
    int ts_log2;   // tile side = 1 << ts_log2
    int tx, ty;    // tiles per row / column
    long long nt;  // B * tx * ty
    int local_bits;
};

// Largest power-of-two side (<= 64) whose bitset fits kTileLdsBytes, shrunk
// while the grid has fewer than `min_tiles` tiles.  Returns false when even
// 4 x 4 does not fit or the packed entry cannot hold the cluster ids.
inline bool choose_tiles(int B, int H, int W, int S, long long min_tiles, TileGrid& g) {
    int lg = 6;
    while (lg >= 2 && (static_cast<long long>(1) << (2 * lg)) * S * 4 > kTileLdsBytes) --lg;
    if (lg < 2) return false;
    auto grid = [&](int l) {
        const int side = 1 << l;
        g.ts_log2 = l;
        g.tx = (W + side - 1) / side;
        g.ty = (H + side - 1) / side;
        g.nt = static_cast<long long>(B) * g.tx * g.ty;
        g.local_bits = 2 * l;
    };
    grid(lg);
    while (g.nt < min_tiles && lg > 3) grid(--lg);
    const long long max_cluster = static_cast<long long>(S) * 32;
    if (max_cluster > (1ll << (32 - g.local_bits))) return false;
    return g.nt <= INT_MAX;
}

__device__ __forceinline__ int tile_of(const TileGrid& g, int b, int h, int w) {
    return (b * g.ty + (h >> g.ts_log2)) * g.tx + (w >> g.ts_log2);
}

__device__ __forceinline__ unsigned local_of(const TileGrid& g, int h, int w) {
    const int m = (1 << g.ts_log2) - 1;
    return static_cast<unsigned>(((h & m) << g.ts_log2) | (w & m));
}

// Run structure of `key` across the wave (key < 0 = inactive lane).
// `start` = first lane of this lane's run; at a run's first lane, `len` is
// the run length.
__device__ __forceinline__ void wave_runs(int key, int& start, int& len) {
    const int l = lane_id();
    const bool valid = key >= 0;
    const int prev = __shfl_up(key, 1, kWave);
    const bool first = valid && (l == 0 || prev != key);
    const unsigned long long starts = __ballot(first);
    const unsigned long long vmask = __ballot(valid);
    const unsigned long long le = (2ull << l) - 1;  // lanes <= l (wraps to all at l = 63)
    const unsigned long long mine = starts & le;
    start = mine ? 63 - __clzll(mine) : 0;
    const unsigned long long after = (starts | ~vmask) & ~le;
    const int next = after ? __ffsll(static_cast<long long>(after)) - 1 : kWave;
    len = next - l;
}

// Pass 1 visitor: counts per tile.
__device__ __forceinline__ void bin_count_visit(int key, int* __restrict__ counts) {
    int start, len;
    wave_runs(key, start, len);
    if (key >= 0 && lane_id() == start) atomicAdd(&counts[key], len);
}

// Pass 2 visitor: scatter packed entries.
__device__ __forceinline__ void bin_fill_visit(int key, unsigned value, int* __restrict__ cursor,
                                               unsigned* __restrict__ entries) {
    int start, len;
    wave_runs(key, start, len);
    int base = 0;
    if (key >= 0 && lane_id() == start) base = atomicAdd(&cursor[key], len);
    base = __shfl(base, start, kWave);
    if (key >= 0) entries[base + (lane_id() - start)] = value;
}

// Host side of the three-pass pipeline (binning.hip).
struct BinPlan {
    TileGrid g;
    int* counts = nullptr;  // nt + 1 ints (exclusive offsets after the scan)
    int* cursor = nullptr;  // nt ints
};

// Zeroes counts; after the caller's count pass, turns counts into offsets,
// seeds the cursors and returns the entry total (one host sync).
int bin_prepare(BinPlan& p, hipStream_t s);
int bin_offsets(BinPlan& p, hipStream_t s, int* total);
// One workgroup per tile: OR entries into LDS, write the tile bitset.
int bin_resolve(const BinPlan& p, const unsigned* entries, int H, int W, int S, int* bitset,
                hipStream_t s);

}  // namespace s2c
