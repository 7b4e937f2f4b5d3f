// Screen-tile binning shared by the fused primitive -> cluster-bitset passes.
//
// Fragments are scattered per (view, tile) as packed 32-bit entries
// (cluster << local_bits | local pixel); one workgroup per tile then ORs its
// entries into an LDS copy of the tile's bitset and writes the tile once
// with coalesced stores (no clear of the bitset, no global OR atomics).
//
// Device-scope atomics are the scarce resource (scattered ones run at a few
// tens of G/s), so the count and fill passes aggregate in LDS: a workgroup
// counts its fragments per tile in an LDS hash table, reserves each tile's
// share with ONE global atomic, and records the table (tile, offset) as an
// "image"; the fill pass reloads the image and places fragments with LDS
// atomics only.  A workgroup whose table overflows takes a global-atomic
// path in both passes (flagged per workgroup, so the passes agree).
//
// Per tile the entry segment is laid out as kBinGroups group slots
// (workgroups privatized by blockIdx % kBinGroups) followed by one overflow
// slot; the whole segment is contiguous.
#pragma once

#include "common.hpp"

namespace s2c {

constexpr int kTileLdsBytes = 32768;  // LDS bitset per tile workgroup
constexpr int kBinGroups = 8;
constexpr int kBinSlots = kBinGroups + 1;  // per tile: groups + overflow
constexpr int kHashLog2 = 10;
constexpr int kHash = 1 << kHashLog2;      // LDS table slots per workgroup
// Per-workgroup image of the count pass: slot keys, offsets within the
// tile's group slot, and fragment counts.
constexpr int kImageInts = 3 * kHash;

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
    return g.nt * kBinSlots < INT_MAX;
}

__device__ __forceinline__ int tile_of(const TileGrid& g, int b, int h, int w) {
    return (b * g.ty + (h >> g.ts_log2)) * g.tx + (w >> g.ts_log2);
}

__device__ __forceinline__ unsigned local_of(const TileGrid& g, int h, int w) {
    const int m = (1 << g.ts_log2) - 1;
    return static_cast<unsigned>(((h & m) << g.ts_log2) | (w & m));
}

__device__ __forceinline__ int group_slot(int tile) {
    return tile * kBinSlots + static_cast<int>(blockIdx.x % kBinGroups);
}

__device__ __forceinline__ int overflow_slot(int tile) { return tile * kBinSlots + kBinGroups; }

// ------------------------------------------------------------ LDS hash

__device__ __forceinline__ int hash_home(int key) {
    return static_cast<int>((static_cast<unsigned>(key) * 2654435761u) >> (32 - kHashLog2));
}

// Slot holding `key` (inserted if absent), or -1 when the table is full.
__device__ __forceinline__ int hash_insert(int* keys, int key) {
    int s = hash_home(key);
    for (int p = 0; p < kHash; ++p) {
        const int prev = atomicCAS(&keys[s], -1, key);
        if (prev == -1 || prev == key) return s;
        s = (s + 1) & (kHash - 1);
    }
    return -1;
}

// Slot holding `key`, or -1 if absent.
__device__ __forceinline__ int hash_find(const int* keys, int key) {
    int s = hash_home(key);
    for (int p = 0; p < kHash; ++p) {
        const int k = keys[s];
        if (k == key) return s;
        if (k == -1) return -1;
        s = (s + 1) & (kHash - 1);
    }
    return -1;
}

// ------------------------------------------------------------ host side

struct BinPlan {
    TileGrid g;
    long long blocks = 0;    // workgroups of the count / fill passes
    int* counts = nullptr;   // nt*kBinSlots + 1 ints, shifted by one; offsets after the scan
    int* cursor = nullptr;   // nt*kBinSlots ints (overflow placement)
    int* scratch = nullptr;  // scan scratch
    int* image = nullptr;    // blocks * kImageInts ints
    int* flags = nullptr;    // blocks ints (1 = overflow path)
    long long slots() const { return g.nt * kBinSlots; }
};

// Device ints a plan needs.
size_t bin_plan_ints(const TileGrid& g, long long blocks);
// Carves the plan's buffers out of `ints` and zeroes the counts.
int bin_prepare(BinPlan& p, int* ints, hipStream_t s);
// After the count pass: offsets, cursors and the entry total (host sync).
int bin_offsets(BinPlan& p, hipStream_t s, int* total);
// One workgroup per tile: OR entries into LDS, write the tile bitset.
int bin_resolve(const BinPlan& p, const unsigned* entries, int H, int W, int S, int* bitset,
                hipStream_t s);

}  // namespace s2c
