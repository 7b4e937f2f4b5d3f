// Screen-tile binning for the fused primitive -> cluster-bitset pass.
//
// One walk over the primitives.  A workgroup stages its fragments in LDS
// as packed 32-bit entries (cluster << local_bits | local pixel) together
// with the LDS hash slot of their tile, counts them per tile, and writes
// them into its own fixed region of the entry buffer grouped by tile.  For
// every (workgroup, tile) pair it writes a run record (tile counter slot,
// index of the run within that slot, region offset, length) into its own
// fixed record region; the index comes from one global atomic on the
// tile's run counter, privatized kBinGroups ways.  A scan of the run
// counters and a scatter put the runs in tile order, and one workgroup per
// tile gathers its runs, ORs the entries into an LDS copy of the tile's
// bitset and writes the tile once with coalesced stores (no clear of the
// bitset, no global OR atomics).
//
// A workgroup that needs more than its fixed regions takes space from
// overflow areas; their sizes come from a running estimate, and the use is
// counted even past the capacity.  A pass whose areas were too small
// resolves to zeros and the same call finishes with one global OR per
// fragment (a kernel that exits at once when the areas sufficed); the host
// reads the counts of a call when they have arrived, without waiting, and
// grows the estimate for the calls after it.
#pragma once

#include <mutex>

#include "common.hpp"

namespace s2c {

// LDS bitset of a tile workgroup: the default tile side fits kTileLdsBytes
// (several resolve workgroups per CU); an explicitly chosen side may use up
// to kTileLdsMax (one per CU, fewer (workgroup, tile) runs).
constexpr int kTileLdsBytes = 32768;
constexpr int kTileLdsMax = 131072;
constexpr int kHashLog2 = 10;
constexpr int kHash = 1 << kHashLog2;  // LDS table slots per workgroup
constexpr int kStage = 4096;           // fragments staged in LDS per workgroup
constexpr int kBinGroups = 8;          // run-counter privatization per tile

struct TileGrid {
    int ts_log2;   // tile side = 1 << ts_log2
    int tx, ty;    // tiles per row / column
    long long nt;  // B * tx * ty
    int local_bits;
};

// Bytes of a tile bitset of side 1 << l with S words per pixel.
inline long long tile_bitset_bytes(int l, int S) {
    return (static_cast<long long>(1) << (2 * l)) * S * 4;
}

// Largest power-of-two side (<= 64) whose bitset fits kTileLdsBytes, shrunk
// while the grid has fewer than `min_tiles` tiles.  Returns false when even
// 4 x 4 does not fit or the packed entry cannot hold the cluster ids.
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
        g.local_bits = 2 * l;
    };
    grid(lg);
    while (g.nt < min_tiles && lg > 3) grid(--lg);
    // packed entries (cluster << local_bits | local pixel) stay below 2^22,
    // leaving 10 bits for a hash slot in the stage (always true while the
    // tile bitset fits kTileLdsMax)
    const long long max_cluster = static_cast<long long>(S) * 32;
    if (max_cluster > (1ll << (22 - g.local_bits))) return false;
    return g.nt * kBinGroups + 2 < INT_MAX;
}

__device__ __forceinline__ int tile_of(const TileGrid& g, int b, int h, int w) {
    return (b * g.ty + (h >> g.ts_log2)) * g.tx + (w >> g.ts_log2);
}

__device__ __forceinline__ unsigned local_of(const TileGrid& g, int h, int w) {
    const int m = (1 << g.ts_log2) - 1;
    return static_cast<unsigned>(((h & m) << g.ts_log2) | (w & m));
}

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

// One slot per active lane from an LDS or global counter, with one atomic
// per wave: the active lanes get consecutive indices in lane order.
__device__ __forceinline__ int wave_append(int* counter) {
    const unsigned long long mask = __ballot(1);
    const int leader = __ffsll(static_cast<long long>(mask)) - 1;
    int base = 0;
    if (lane_id() == leader) base = atomicAdd(counter, __popcll(mask));
    base = __shfl(base, leader, kWave);
    return base + __popcll(mask & lanemask_lt());
}

// ------------------------------------------------------------ buffers

// Run record of one (workgroup, tile) pair, as the stage pass writes it.
// `slot` is the tile's counter slot: tile * kBinGroups + group.
struct BinRecord {
    int slot, k, src, n;
};
// The same run in tile order: entries[src .. src + n).
struct BinRun {
    int src, n;
};

// Run counters are privatized per tile over kBinGroups workgroup groups
// (blockIdx % kBinGroups), so a hot tile's counter is not one address.
// Each workgroup owns a fixed region of kStage entries and kBinRecs run
// records; the rare workgroup that needs more takes space from the
// overflow areas behind the fixed regions.
constexpr int kBinRecs = 512;

// Global counters of the overflow areas.
enum BinCounter { kEntriesOvf = 0, kRunsOvf = 1, kBinCounters = 4 };

struct BinPlan {
    TileGrid g;
    long long blocks = 0;          // workgroups of the stage pass
    int* runs_per_slot = nullptr;  // nt*kBinGroups + 1 ints, shifted by one; offsets after the scan
    int* counters = nullptr;       // kBinCounters ints
    int* block_runs = nullptr;     // runs in each workgroup's fixed record region
    int4* order = nullptr;         // nt: resolve order (tile, first run, run end)
    int* scratch = nullptr;        // scan scratch
    unsigned* entries = nullptr;   // blocks*kStage fixed + ovf_entries
    BinRecord* records = nullptr;  // blocks*kBinRecs fixed + ovf_runs
    BinRun* runs = nullptr;        // same size as records
    int ovf_entries = 0, ovf_runs = 0;
    long long slots() const { return g.nt * kBinGroups; }
    long long fixed_entries() const { return blocks * kStage; }
    long long fixed_runs() const { return blocks * kBinRecs; }
};

// Device ints of the fixed part of a plan (counters, run counters, scan
// scratch, per-workgroup run counts).
size_t bin_plan_ints(const TileGrid& g, long long blocks);
// Carves the fixed part out of `ints`.
void bin_carve(BinPlan& p, int* ints);
// Zeroes the run counters and the global counters.
int bin_clear(const BinPlan& p, hipStream_t s);
// Queues the copy of the counters into `host` (pinned).
int bin_counters_async(const BinPlan& p, int* host, hipStream_t s);
// The overflow-area use of a stage pass from its copied counters; false
// when a count wrapped past 2^31.
bool bin_overflow_use(const int* host, long long* entries, long long* runs);
// Run offsets per tile, runs scattered into tile order, then one workgroup
// per tile ORs its entries into LDS and writes the tile bitset.  If the
// stage pass overflowed its areas, every tile is written as zeros instead.
int bin_resolve(const BinPlan& p, int H, int W, int S, int* bitset, hipStream_t s);

// Pinned copy of the last call's counters on a stream, the event after that
// copy, and whether the copy has not been read yet.
struct BinStreamCounters {
    int* host = nullptr;
    hipEvent_t done = nullptr;
    bool pending = false;
};

// Device buffers of the binned pass kept per (device, stream) across calls,
// so that a call allocates nothing once its sizes have been seen.  Calls on
// one stream are ordered, so they can share the buffers; `lock` must be
// held from the first get() to the last launch of a call, so that two
// threads enqueueing on one stream cannot interleave (the lock is never held
// across a wait for the device, except to free a block that grows).  Buffers only grow: a
// larger size waits for the stream, frees the old block and allocates a new
// one.  They are kept until the process ends or release() (for the
// 50k-strand, 4-view 1024^2 workload about 0.6 GB per stream).
struct BinWorkspace {
    enum Slot { kPlanInts = 0, kEntries = 1, kRecords = 2, kRuns = 3, kSlots = 4 };
    static std::mutex& lock();
    // Device memory of at least `bytes` for `slot`, for calls on stream s of
    // the current device.
    static int get(hipStream_t s, Slot slot, size_t bytes, void** out);
    // The pinned counter copy of stream s on the current device (allocated
    // on first use).
    static int counters(hipStream_t s, BinStreamCounters** out);
    // Frees the buffers of stream s on the current device, after waiting for
    // the stream.
    static int release(hipStream_t s);
};

// Running estimates of the overflow-area sizes, per process.
long long bin_entry_hint(long long want);
long long bin_run_hint(long long want);
// Back to the smallest estimates, forgetting counts not read yet (hold
// BinWorkspace::lock()).
void bin_reset_hints();

}  // namespace s2c
