// Screen-tile binning for the fused primitive -> cluster-bitset pass.
//
// One walk over the primitives.  A workgroup ORs each fragment into an LDS
// bit mask of its (cell, cluster) key, where a cell is a square of up to
// 16 x 16 pixels inside a screen tile; the key's mask is found through an
// LDS hash, and a fragment that repeats a pixel and cluster already seen by
// the workgroup sets no new bit.  After the walk the workgroup writes one
// record per key (its tile, and its cluster, cell and mask) into its own
// region (no global atomics).
//
// The records are then put in tile order by a counting sort without global
// atomics: each sort block counts the records of its kSortChunk stage
// workgroups per tile in LDS, a two-level scan down the (sort block, tile)
// count matrix gives every sort block its first position within each tile,
// block scans of the tile totals give the tiles' offsets, and each sort
// block places its records, ranked within their tile by LDS atomics.  One workgroup per tile then ORs
// its records' mask pixels into an LDS copy of the tile's bitset and writes
// the tile once with coalesced stores (no clear of the bitset, no global OR
// atomics).
//
// A workgroup with more distinct keys than its table holds writes no record
// and marks its record count -1; a kernel behind the resolve walks the
// marked workgroups again and ORs their fragments directly into the written
// bitset.  Nothing is cleared before a pass and the host reads nothing
// back.
#pragma once

#include <mutex>

#include "common.hpp"

namespace s2c {

// LDS bitset of a tile workgroup: the default tile side fits kTileLdsBytes
// (several resolve workgroups per CU); an explicitly chosen side may use up
// to kTileLdsMax (one per CU, fewer records per tile).
constexpr int kTileLdsBytes = 32768;
constexpr int kTileLdsMax = 131072;
constexpr int kHashLog2 = 9;
constexpr int kHash = 1 << kHashLog2;  // (cell, cluster) keys per workgroup
constexpr int kHashProbes = 64;        // probes before a workgroup counts as full
constexpr int kCellLog2Max = 4;        // cells of at most 16 x 16 pixels
constexpr int kMaskWordsMax = 8;       // words of a 16 x 16 cell mask

struct TileGrid {
    int ts_log2;   // tile side = 1 << ts_log2
    int tx, ty;    // tiles per row / column
    long long nt;  // B * tx * ty
    int cs_log2;   // cell side = 1 << cs_log2 = min(tile side, 16)
    int cy;        // cell rows per view
    int col_bits;  // bits of a cell column in a key
    int kb;        // bits of a cluster id in a key
    int mw_log2;   // log2 of the 32-bit words of a cell mask
};

// Bytes of a tile bitset of side 1 << l with S words per pixel.
inline long long tile_bitset_bytes(int l, int S) {
    return (static_cast<long long>(1) << (2 * l)) * S * 4;
}

// Bits needed to write x >= 0.
inline int bit_width(long long x) {
    int n = 0;
    while (x > 0) {
        ++n;
        x >>= 1;
    }
    return n;
}

// Largest power-of-two side (<= 64) whose bitset fits kTileLdsBytes, shrunk
// while the grid has fewer than `min_tiles` tiles.  Returns false when even
// 4 x 4 does not fit or a (cell, cluster) key does not fit 31 bits.
// force_lg > 0 fixes the side to 1 << force_lg, whose bitset may use up to
// kTileLdsMax (false when it does not fit).
inline bool choose_tiles(int B, int H, int W, int S, long long min_tiles, TileGrid& g,
                         int force_lg = 0) {
    auto fits = [&](int l, long long cap) {
        return l >= 2 && l <= 6 && tile_bitset_bytes(l, S) <= cap;
    };
    int lg = 6;
    while (lg >= 2 && !fits(lg, kTileLdsBytes)) --lg;
    if (lg < 2) return false;
    if (force_lg > 0) {
        if (!fits(force_lg, kTileLdsMax)) return false;
        lg = force_lg;
        min_tiles = 0;
    }
    auto grid = [&](int l) {
        const int side = 1 << l;
        g.ts_log2 = l;
        g.tx = (W + side - 1) / side;
        g.ty = (H + side - 1) / side;
        g.nt = static_cast<long long>(B) * g.tx * g.ty;
        g.cs_log2 = l < kCellLog2Max ? l : kCellLog2Max;
        const int cs = 1 << g.cs_log2;
        g.cy = (H + cs - 1) / cs;
        g.col_bits = bit_width((W + cs - 1) / cs - 1);
        g.mw_log2 = 2 * g.cs_log2 > 5 ? 2 * g.cs_log2 - 5 : 0;
    };
    grid(lg);
    while (g.nt < min_tiles && lg > 3) grid(--lg);
    g.kb = bit_width(static_cast<long long>(S) * 32 - 1);
    // keys ((b * cy + cell row) << col_bits | cell column) << kb | cluster
    // are non-negative ints; the cluster and the cell's place in its tile
    // share one int of a record (cluster below 2^16)
    const int low = g.col_bits + g.kb;
    if (g.kb > 16 || low >= 31 || static_cast<long long>(B) * g.cy > (1ll << (31 - low)))
        return false;
    return g.nt + 2 < INT_MAX;
}

// ------------------------------------------------------------ cell keys

// ((b * cy + h / cs) << col_bits | w / cs) << kb | c.  The fields do not
// overlap, so the key is a sum whose first term is fixed along an edge.
__device__ __forceinline__ int cell_key(const TileGrid& g, int b, int h, int w, int c) {
    const int a = g.col_bits + g.kb;
    const int base = ((b * g.cy) << a) | c;
    return base + ((h >> g.cs_log2) << a) + ((w >> g.cs_log2) << g.kb);
}

// Word and bit of pixel (h, w) in its cell's mask.  A 16 x 16 cell spreads
// its pixels over the 8 words by (x XOR y) mod 8, bit 2y + x/8 (x, y within
// the cell), so pixels along a row, a column or most lines set bits in
// several words: the resolve takes the bits of a word one at a time, each
// wave as long as its fullest word, and this cuts those iterations by 43%
// against row pairs per word at the full workload.  Smaller cells are
// row-major.  For 16 x 16 cells (x ^ y) & 7 == (w ^ h) & 7 and the bit is
// (2h + ((w >> 3) & 1)) mod 32, so neither needs x or y.
__device__ __forceinline__ void cell_bit(const TileGrid& g, int h, int w, int& word,
                                         unsigned& bit) {
    if (g.cs_log2 == kCellLog2Max) {
        word = (w ^ h) & 7;
        bit = ((static_cast<unsigned>(h) << 1) | ((static_cast<unsigned>(w) >> 3) & 1u)) & 31u;
    } else {
        const int m = (1 << g.cs_log2) - 1;
        const int y = h & m, x = w & m;
        const unsigned l = static_cast<unsigned>((y << g.cs_log2) | x);
        word = static_cast<int>(l >> 5);
        bit = l & 31u;
    }
}

// The inverse of cell_bit: the cell's row-major pixel of (word, bit).
__device__ __forceinline__ unsigned cell_pixel(int cs_log2, unsigned word, unsigned bit) {
    if (cs_log2 == kCellLog2Max) return (bit << 3) | ((word ^ (bit >> 1)) & 7u);
    return (word << 5) | bit;
}

// The tile of a key's cell and the cell's row-major place in that tile.
__device__ __forceinline__ void cell_tile(const TileGrid& g, int key, int& tile, int& cell) {
    const int rc = key >> g.kb;
    const int col = rc & ((1 << g.col_bits) - 1);
    const int row = rc >> g.col_bits;
    const int b = row / g.cy;
    const int cr = row - b * g.cy;
    const int sh = g.ts_log2 - g.cs_log2;
    tile = (b * g.ty + (cr >> sh)) * g.tx + (col >> sh);
    const int f = (1 << sh) - 1;
    cell = ((cr & f) << sh) | (col & f);
}

// ------------------------------------------------------------ LDS hash

// Fibonacci hashing of the key's low 24 bits: the top kHashLog2 bits of
// key * floor(2^24 / golden ratio) mod 2^24, with a 24-bit multiply (full
// rate, where a 32-bit multiply is quarter rate).  Keys that differ only
// above bit 23 (in the high bits of the cell row, such as one cell seen in
// two views) share a home slot and are told apart by the probes.
__device__ __forceinline__ int hash_home(int key) {
    const unsigned p = __umul24(static_cast<unsigned>(key), 0x9E3779u);
    return static_cast<int>((p >> (24 - kHashLog2)) & (kHash - 1));
}

// Slot holding `key` (inserted if absent), or -1 after kHashProbes probes.
// The first probe is inline, the rest a rolled loop (the walk inlines this
// at every call site).
__device__ __forceinline__ int hash_insert(int* keys, int key) {
    int s = hash_home(key);
    int prev = atomicCAS(&keys[s], -1, key);
    if (prev == -1 || prev == key) return s;
#pragma unroll 1
    for (int p = 1; p < kHashProbes; ++p) {
        s = (s + 1) & (kHash - 1);
        prev = atomicCAS(&keys[s], -1, key);
        if (prev == -1 || prev == key) return s;
    }
    return -1;
}

// ------------------------------------------------------------ buffers

// A record is its tile and info (the cluster id | the cell's row-major
// place in its tile << 16) as an int2, and its cell mask (1 << mw_log2 of
// 8 words used) as two uint4; in tile order it is (record index, info).

// A stage workgroup writes at most kHash records (one per table slot) into
// its own region of kBinRecs.
constexpr int kBinRecs = kHash;

// Resolve order classes (tiles by record count, powers of 4, heaviest first).
constexpr int kOrderClasses = 8;

// Counting sort of the records by tile: sort block j takes the regions of
// stage workgroups [j * kSortChunk, (j + 1) * kSortChunk), one wave per
// workgroup at a time, and counts their records per tile in LDS,
// kSortTiles tiles at a time; the count matrix is scanned down its columns
// in kSortSegs segments.
constexpr int kSortThreads = 1024;
constexpr int kSortChunk = 64;
constexpr int kSortTiles = 16384;  // 64 KB of LDS counters
constexpr int kSortSegs = 16;

// Outputs of a stage workgroup.
struct BinOut {
    int* block_recs;  // records of each workgroup; -1 when its table overflowed
    int2* rec_ti;     // (tile, info) of each record
    uint4* rec_mask;  // two uint4 per record
};

struct BinPlan {
    TileGrid g;
    long long blocks = 0;       // workgroups of the stage pass
    int* offs = nullptr;        // nt + 1 ints: tile totals, then tile offsets
    int* block_recs = nullptr;  // `blocks` ints
    int4* order = nullptr;      // nt: resolve order (tile, first record, record end)
    int* segs = nullptr;        // kSortSegs * nt segment sums of the count matrix
    int* classes = nullptr;     // kOrderClasses * class_blocks() + 1: tiles per order class
    int* bsums = nullptr;       // class_blocks() + 1: record totals per class block
    int* counts = nullptr;      // sort_blocks() * nt: counts, then first positions
    int2* rec_ti = nullptr;     // records(): blocks regions of kBinRecs
    uint4* rec_mask = nullptr;  // 2 * records()
    int2* runs = nullptr;       // records(): (record, info) in tile order
    long long records() const { return blocks * kBinRecs; }
    long long sort_blocks() const { return (blocks + kSortChunk - 1) / kSortChunk; }
    long long class_blocks() const { return (g.nt + kBlock - 1) / kBlock; }
    BinOut out() const { return BinOut{block_recs, rec_ti, rec_mask}; }
};

// Device ints of a plan besides its masks and runs (tile offsets, resolve
// order, record counts, count matrix and its scans, record tiles).
size_t bin_plan_ints(const TileGrid& g, long long blocks);
// Carves them out of `ints`.
void bin_carve(BinPlan& p, int* ints);
// The records sorted into tile order, then one workgroup per tile ORs their
// masks into LDS and writes the tile bitset.
int bin_resolve(const BinPlan& p, int H, int W, int S, int* bitset, hipStream_t s);

// Device buffers of the binned pass kept per (device, stream) across calls,
// so that a call allocates nothing once its sizes have been seen.  Calls on
// one stream are ordered, so they can share the buffers; `lock` must be
// held from the first get() to the last launch of a call, so that two
// threads enqueueing on one stream cannot interleave (the lock is never held
// across a wait for the device, except to free a block that grows).  Buffers
// only grow: a larger size waits for the stream, frees the old block and
// allocates a new one.  They are kept until the process ends or release()
// (for the 50k-strand, 4-view 1024^2 workload about 1 GB per stream).
struct BinWorkspace {
    enum Slot { kPlanInts = 0, kMasks = 1, kRuns = 2, kSlots = 3 };
    static std::mutex& lock();
    // Device memory of at least `bytes` for `slot`, for calls on stream s of
    // the current device.
    static int get(hipStream_t s, Slot slot, size_t bytes, void** out);
    // Frees the buffers of stream s on the current device, after waiting for
    // the stream.
    static int release(hipStream_t s);
};

}  // namespace s2c
