// Run ordering and the per-tile LDS resolve of the binned cluster-bitset
// pass (see binning.hpp).
#include "binning.hpp"

#include <atomic>
#include <map>
#include <utility>

namespace s2c {

// A tile's resolve is a chain of dependent global round trips (order,
// offsets, runs, entries, then the stores), so the kernel is bound by how
// many tiles are in flight: 512-thread workgroups let four tiles share a CU
// (LDS-bound at 32 KB of bitset each), twice as many as 1024-thread ones,
// and each thread keeps kResolveUnroll entry loads in flight.
constexpr int kResolveThreads = 512;
constexpr int kResolveUnroll = 8;

// Resolve order: tiles by decreasing run count in classes of powers of 4,
// so the heaviest tiles start first and the light and empty ones fill the
// tail.  Each entry is (tile, first run, end of its runs), so a resolve
// workgroup starts with one load.  One workgroup; each thread classifies kOrderPer consecutive tiles
// from one round of loads, and counts / places them with wave scans and one
// LDS atomic per wave and class.  The order within a class is arbitrary.
constexpr int kOrderThreads = 1024;
constexpr int kOrderClasses = 8;
constexpr int kOrderPer = 16;

__device__ __forceinline__ void order_classes(long long t0, long long nt, const int* offs,
                                              int (&cls)[kOrderPer], int (&v)[kOrderPer + 1]) {
#pragma unroll
    for (int i = 0; i <= kOrderPer; ++i) v[i] = offs[dmin(t0 + i, nt) * kBinGroups];
#pragma unroll
    for (int i = 0; i < kOrderPer; ++i) {
        const int w = v[i + 1] - v[i];
        cls[i] = t0 + i >= nt ? -1
                 : w <= 0     ? kOrderClasses - 1
                              : dmax(0, kOrderClasses - 2 - (31 - __clz(w)) / 2);
    }
}

__device__ __forceinline__ void tile_order(long long nt, const int* __restrict__ offs,
                                           int4* __restrict__ order) {
    __shared__ int hist[kOrderClasses];
    if (threadIdx.x < kOrderClasses) hist[threadIdx.x] = 0;
    __syncthreads();
    const int lane = lane_id();
    constexpr long long step = static_cast<long long>(kOrderThreads) * kOrderPer;
    for (long long c0 = 0; c0 < nt; c0 += step) {
        int cls[kOrderPer], v[kOrderPer + 1];
        order_classes(c0 + threadIdx.x * kOrderPer, nt, offs, cls, v);
#pragma unroll
        for (int k = 0; k < kOrderClasses; ++k) {
            int n = 0;
#pragma unroll
            for (int i = 0; i < kOrderPer; ++i) n += cls[i] == k;
            const int incl = wave_incl_scan(n);
            if (lane == kWave - 1 && incl > 0) atomicAdd(&hist[k], incl);
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int k = 0; k < kOrderClasses; ++k) {
            const int v = hist[k];
            hist[k] = acc;
            acc += v;
        }
    }
    __syncthreads();
    for (long long c0 = 0; c0 < nt; c0 += step) {
        const long long t0 = c0 + threadIdx.x * kOrderPer;
        int cls[kOrderPer], v[kOrderPer + 1];
        order_classes(t0, nt, offs, cls, v);
#pragma unroll
        for (int k = 0; k < kOrderClasses; ++k) {
            int n = 0;
#pragma unroll
            for (int i = 0; i < kOrderPer; ++i) n += cls[i] == k;
            const int incl = wave_incl_scan(n);
            int b = 0;
            if (lane == kWave - 1 && incl > 0) b = atomicAdd(&hist[k], incl);
            b = __shfl(b, kWave - 1, kWave) + incl - n;
#pragma unroll
            for (int i = 0; i < kOrderPer; ++i)
                if (cls[i] == k) order[b++] = make_int4(static_cast<int>(t0 + i), v[i], v[i + 1], 0);
        }
    }
}

// Runs into tile order, and the resolve order, in one launch of
// kOrderThreads-thread workgroups: workgroup 0 orders the tiles while the
// others scatter the run records, each covering kOrderThreads / kBlock
// stage workgroups' fixed record regions (one kBlock-thread group each) or
// kOrderThreads records of the overflow area.
constexpr int kStagePerOrder = kOrderThreads / kBlock;

__global__ void __launch_bounds__(kOrderThreads)
bin_order_kernel(long long nt, long long blocks, const int* __restrict__ block_runs, int ovf_cap,
                 int ovf_entry_cap, const int* __restrict__ counters,
                 const BinRecord* __restrict__ recs, const int* __restrict__ offs,
                 BinRun* __restrict__ runs, int4* __restrict__ order) {
    // An attempt whose overflow areas were too small (the host relaunches
    // it) gets an order of empty tiles: its resolve writes zeros and reads
    // no run or entry.
    const int ce = counters[kEntriesOvf], cr = counters[kRunsOvf];
    if (ce < 0 || ce > ovf_entry_cap || cr < 0 || cr > ovf_cap) {
        if (blockIdx.x == 0)
            for (long long t = threadIdx.x; t < nt; t += kOrderThreads)
                order[t] = make_int4(static_cast<int>(t), 0, 0, 0);
        return;
    }
    if (blockIdx.x == 0) {
        tile_order(nt, offs, order);
        return;
    }
    const long long fixed_groups = (blocks + kStagePerOrder - 1) / kStagePerOrder;
    const long long w = blockIdx.x - 1;
    long long first, count;
    int tid = threadIdx.x, stride = kOrderThreads;
    if (w < fixed_groups) {
        const long long b = w * kStagePerOrder + threadIdx.x / kBlock;
        if (b >= blocks) return;
        first = b * kBinRecs;
        count = block_runs[b];
        tid = threadIdx.x % kBlock;
        stride = kBlock;
    } else {
        const long long o = (w - fixed_groups) * kOrderThreads;
        const long long used = dmin(static_cast<long long>(counters[kRunsOvf]),
                                    static_cast<long long>(ovf_cap));
        first = blocks * kBinRecs + o;
        count = dmax(dmin(used - o, static_cast<long long>(kOrderThreads)), 0ll);
    }
    for (long long i = tid; i < count; i += stride) {
        const BinRecord rec = recs[first + i];
        runs[offs[rec.slot] + rec.k] = BinRun{rec.src, rec.n};
    }
}

template <bool VEC4, int kLds>
__global__ void __launch_bounds__(kResolveThreads)
tile_resolve_kernel(TileGrid g, const int4* __restrict__ order,
                    const BinRun* __restrict__ runs,
                    const unsigned* __restrict__ entries, int H, int W, int S,
                    int* __restrict__ bitset) {
    __shared__ __align__(16) unsigned tile[kLds / 4];
    __shared__ int pre[kResolveThreads];
    __shared__ int src[kResolveThreads];
    __shared__ int sh[kResolveThreads / kWave + 1];
    const int4 ord = order[blockIdx.x];
    const int t = ord.x, r0 = ord.y, r1 = ord.z;
    // the first chunk's run is loaded before the clear of the LDS bitset
    BinRun next{0, 0};
    if (r0 + static_cast<int>(threadIdx.x) < r1) next = runs[r0 + threadIdx.x];
    const int side = 1 << g.ts_log2;
    const int words = side * side * S;  // a multiple of 4 (side >= 4)
    for (int i = threadIdx.x; i < words / 4; i += kResolveThreads)
        reinterpret_cast<uint4*>(tile)[i] = make_uint4(0u, 0u, 0u, 0u);
    __syncthreads();  // a tile without runs goes straight to the write-out
    const unsigned lmask = (1u << g.local_bits) - 1u;
    const unsigned kmax = static_cast<unsigned>(S) * 32u;
    // LDS word of (pixel l, word j) is l*S + (j ^ (l & smask)): with S a
    // power of two, consecutive entries of one cluster (same j) at
    // different pixels fall in different banks
    const unsigned smask = (S & (S - 1)) == 0 ? static_cast<unsigned>(S) - 1u : 0u;
    const int w = threadIdx.x / kWave;
    // Runs are taken kResolveThreads at a time: their lengths are scanned
    // into `pre`, then the chunk's entries are read as one flat range (each
    // lane finds its run by binary search), so consecutive lanes read
    // consecutive entries of a run and kResolveUnroll loads are in flight.
    for (int rb = r0; rb < r1; rb += kResolveThreads) {
        const int r = rb + static_cast<int>(threadIdx.x);
        BinRun run = next;
        if (rb != r0) run = r < r1 ? runs[r] : BinRun{0, 0};
        const int incl = wave_incl_scan(run.n);
        if (lane_id() == kWave - 1) sh[w] = incl;
        __syncthreads();
        if (threadIdx.x == 0) {
            int acc = 0;
            for (int i = 0; i < kResolveThreads / kWave; ++i) {
                const int v = sh[i];
                sh[i] = acc;
                acc += v;
            }
            sh[kResolveThreads / kWave] = acc;
        }
        __syncthreads();
        pre[threadIdx.x] = sh[w] + incl - run.n;
        src[threadIdx.x] = run.src;
        const int total = sh[kResolveThreads / kWave];
        __syncthreads();
        const int nr = dmin(r1 - rb, kResolveThreads);
        constexpr int step = kResolveThreads * kResolveUnroll;
        for (int e = threadIdx.x; e < total; e += step) {
            unsigned v[kResolveUnroll];
#pragma unroll
            for (int u = 0; u < kResolveUnroll; ++u) {
                const int j = e + u * kResolveThreads;
                v[u] = 0xFFFFFFFFu;
                if (j < total) {
                    int lo = 0, hi = nr;  // last run with pre <= j
                    while (hi - lo > 1) {
                        const int mid = (lo + hi) >> 1;
                        if (pre[mid] <= j) lo = mid;
                        else hi = mid;
                    }
                    v[u] = entries[src[lo] + (j - pre[lo])];
                }
            }
#pragma unroll
            for (int u = 0; u < kResolveUnroll; ++u) {
                const unsigned local = v[u] & lmask;
                const unsigned k = v[u] >> g.local_bits;
                if (k < kmax)
                    atomicOr(&tile[local * S + ((k >> 5) ^ (local & smask))], 1u << (k & 31));
            }
        }
        __syncthreads();  // pre / src / sh are rewritten by the next chunk
    }
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
    const long long n = g.nt * kBinGroups;
    return static_cast<size_t>(scan_scratch_ints(n + 1) + 4 + 4 * g.nt + kBinCounters + (n + 2) +
                               blocks + 4);
}

void bin_carve(BinPlan& p, int* ints) {
    const long long n = p.slots();
    p.scratch = ints;  // int64 values first: 8-byte aligned
    long long off = static_cast<long long>(scan_scratch_ints(n + 1));
    off = (off + 3) & ~3ll;
    p.order = reinterpret_cast<int4*>(ints + off);  // 16-byte aligned
    off += 4 * p.g.nt;
    p.counters = ints + off;  // counters then run counts: one contiguous clear
    p.runs_per_slot = ints + off + kBinCounters;
    p.block_runs = ints + off + kBinCounters + n + 2;
}

int bin_clear(const BinPlan& p, hipStream_t s) {
    return fill_u32(p.counters, 0u, kBinCounters + p.slots() + 1, s);
}

int bin_counters_async(const BinPlan& p, int* host, hipStream_t s) {
    S2C_HIP(hipMemcpyAsync(host, p.counters, sizeof(int) * kBinCounters, hipMemcpyDeviceToHost, s));
    return S2C_OK;
}

bool bin_overflow_use(const int* host, long long* entries, long long* runs) {
    // int32 atomics: a wrapped (negative) value means more than 2^31
    if (host[kEntriesOvf] < 0 || host[kRunsOvf] < 0) return false;
    *entries = host[kEntriesOvf];
    *runs = host[kRunsOvf];
    return true;
}

int bin_resolve(const BinPlan& p, int H, int W, int S, int* bitset, hipStream_t s) {
    if (int r = inclusive_scan_inplace(p.runs_per_slot, p.slots() + 1, p.scratch, s)) return r;
    const long long grid = 1 + (p.blocks + kStagePerOrder - 1) / kStagePerOrder +
                           (p.ovf_runs + kOrderThreads - 1) / kOrderThreads;
    bin_order_kernel<<<static_cast<unsigned>(grid), kOrderThreads, 0, s>>>(
        p.g.nt, p.blocks, p.block_runs, p.ovf_runs, p.ovf_entries, p.counters, p.records,
        p.runs_per_slot, p.runs, p.order);
    S2C_LAUNCH_CHECK();
    const bool vec = (S % 4 == 0) && ((reinterpret_cast<uintptr_t>(bitset) & 15) == 0);
    const bool small = tile_bitset_bytes(p.g.ts_log2, S) <= kTileLdsBytes;
    const unsigned grid_t = static_cast<unsigned>(p.g.nt);
    if (vec && small)
        tile_resolve_kernel<true, kTileLdsBytes><<<grid_t, kResolveThreads, 0, s>>>(
            p.g, p.order, p.runs, p.entries, H, W, S, bitset);
    else if (vec)
        tile_resolve_kernel<true, kTileLdsMax><<<grid_t, kResolveThreads, 0, s>>>(
            p.g, p.order, p.runs, p.entries, H, W, S, bitset);
    else if (small)
        tile_resolve_kernel<false, kTileLdsBytes><<<grid_t, kResolveThreads, 0, s>>>(
            p.g, p.order, p.runs, p.entries, H, W, S, bitset);
    else
        tile_resolve_kernel<false, kTileLdsMax><<<grid_t, kResolveThreads, 0, s>>>(
            p.g, p.order, p.runs, p.entries, H, W, S, bitset);
    S2C_LAUNCH_CHECK();
    return S2C_OK;
}

namespace {
struct WorkspaceBuffers {
    void* p[BinWorkspace::kSlots] = {};
    size_t n[BinWorkspace::kSlots] = {};
    BinStreamCounters counters;
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

int BinWorkspace::counters(hipStream_t s, BinStreamCounters** out) {
    int dev = 0;
    S2C_HIP(hipGetDevice(&dev));
    BinStreamCounters& c = workspaces()[{dev, s}].counters;
    if (!c.host) {
        S2C_HIP(hipHostMalloc(reinterpret_cast<void**>(&c.host), sizeof(int) * kBinCounters,
                              hipHostMallocPortable));
        if (hipEventCreateWithFlags(&c.done, hipEventDisableTiming) != hipSuccess) {
            (void)hipHostFree(c.host);
            c.host = nullptr;
            S2C_HIP(hipGetLastError());
            set_error("dda_cluster_mask: event creation failed");
            return S2C_ERR_ARG;
        }
        c.pending = false;
    }
    *out = &c;
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
    BinStreamCounters& c = it->second.counters;
    if (c.done) S2C_HIP(hipEventDestroy(c.done));
    if (c.host) S2C_HIP(hipHostFree(c.host));
    m.erase(it);
    return S2C_OK;
}

namespace {
std::atomic<long long> g_entry_hint{0}, g_run_hint{0};

long long raise_hint(std::atomic<long long>& h, long long want) {
    long long cur = h.load(std::memory_order_relaxed);
    while (want > cur && !h.compare_exchange_weak(cur, want, std::memory_order_relaxed)) {
    }
    return dmax(cur, want);
}
}  // namespace

long long bin_entry_hint(long long want) { return raise_hint(g_entry_hint, want); }
long long bin_run_hint(long long want) { return raise_hint(g_run_hint, want); }
void bin_reset_hints() {
    g_entry_hint.store(0, std::memory_order_relaxed);
    g_run_hint.store(0, std::memory_order_relaxed);
    // counts not read yet would raise the estimates again at the next call
    for (auto& kv : workspaces()) kv.second.counters.pending = false;
}

}  // namespace s2c
