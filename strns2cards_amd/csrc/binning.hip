// Tile offsets and the per-tile LDS resolve of the binned cluster-bitset
// passes (see binning.hpp).
#include "binning.hpp"

namespace s2c {

// A tile's entry count varies by more than an order of magnitude (strand
// roots crowd a few tiles), so the resolve uses 1024-thread workgroups and
// keeps kResolveUnroll loads in flight per thread: the hottest tiles set the
// kernel's duration.
constexpr int kResolveThreads = 1024;
constexpr int kResolveUnroll = 4;

template <bool VEC4>
__global__ void __launch_bounds__(kResolveThreads)
tile_resolve_kernel(TileGrid g, const int* __restrict__ offs, const unsigned* __restrict__ entries,
                    int H, int W, int S, int* __restrict__ bitset) {
    __shared__ unsigned tile[kTileLdsBytes / 4];
    const int t = blockIdx.x;
    const int side = 1 << g.ts_log2;
    const int words = side * side * S;
    for (int i = threadIdx.x; i < words; i += kResolveThreads) tile[i] = 0u;
    __syncthreads();
    const int e0 = offs[t * kBinSlots], e1 = offs[(t + 1) * kBinSlots];
    const unsigned lmask = (1u << g.local_bits) - 1u;
    const unsigned kmax = static_cast<unsigned>(S) * 32u;
    constexpr int step = kResolveThreads * kResolveUnroll;
    for (int e = e0 + threadIdx.x; e < e1; e += step) {
        unsigned v[kResolveUnroll];
#pragma unroll
        for (int u = 0; u < kResolveUnroll; ++u) {
            const int idx = e + u * kResolveThreads;
            v[u] = idx < e1 ? entries[idx] : 0xFFFFFFFFu;
        }
#pragma unroll
        for (int u = 0; u < kResolveUnroll; ++u) {
            const unsigned local = v[u] & lmask;
            const unsigned k = v[u] >> g.local_bits;
            if (k < kmax) atomicOr(&tile[local * S + (k >> 5)], 1u << (k & 31));
        }
    }
    __syncthreads();
    const int per_b = g.tx * g.ty;
    const int b = t / per_b;
    const int r = t - b * per_b;
    const int y0 = (r / g.tx) * side;
    const int x0 = (r % g.tx) * side;
    const int row_w = dmin(side, W - x0);  // pixels of this tile inside the image
    const int rows = dmin(side, H - y0);
    if (VEC4) {
        const int q = S / 4;
        const int per_row = row_w * q;
        const uint4* src = reinterpret_cast<const uint4*>(tile);
        for (int i = threadIdx.x; i < rows * per_row; i += kResolveThreads) {
            const int ly = i / per_row;
            const int rem = i - ly * per_row;
            const long long dst =
                ((static_cast<long long>(b) * H + y0 + ly) * W + x0) * q + rem;
            reinterpret_cast<uint4*>(bitset)[dst] = src[ly * side * q + rem];
        }
    } else {
        const int per_row = row_w * S;
        for (int i = threadIdx.x; i < rows * per_row; i += kResolveThreads) {
            const int ly = i / per_row;
            const int rem = i - ly * per_row;
            const long long dst = ((static_cast<long long>(b) * H + y0 + ly) * W + x0) * S + rem;
            bitset[dst] = static_cast<int>(tile[ly * side * S + rem]);
        }
    }
}

size_t bin_plan_ints(const TileGrid& g, long long blocks) {
    const long long n = g.nt * kBinSlots;
    return static_cast<size_t>((n + 2) + n + scan_scratch_ints(n + 1) +
                               kImageInts * blocks + blocks + 16);
}

int bin_prepare(BinPlan& p, int* ints, hipStream_t s) {
    const long long n = p.slots();
    p.scratch = ints;  // int64 values first: 8-byte aligned
    long long off = static_cast<long long>(scan_scratch_ints(n + 1));
    off = (off + 3) & ~3ll;  // counts 16-byte aligned for the vector fill
    p.counts = ints + off;
    off += n + 2;
    p.cursor = ints + off;
    off += n;
    p.image = ints + off;
    off += static_cast<long long>(kImageInts) * p.blocks;
    p.flags = ints + off;
    return fill_u32(p.counts, 0u, n + 1, s);
}

int bin_offsets(BinPlan& p, hipStream_t s, int* total) {
    const long long n = p.slots();
    if (int r = inclusive_scan_inplace(p.counts, n + 1, p.scratch, s)) return r;
    S2C_HIP(hipMemcpyAsync(p.cursor, p.counts, sizeof(int) * n, hipMemcpyDeviceToDevice, s));
    return read_scan_total(p.scratch, n + 1, s, total, "binning");
}

int bin_resolve(const BinPlan& p, const unsigned* entries, int H, int W, int S, int* bitset,
                hipStream_t s) {
    const bool vec = (S % 4 == 0) && ((reinterpret_cast<uintptr_t>(bitset) & 15) == 0);
    if (vec)
        tile_resolve_kernel<true><<<static_cast<unsigned>(p.g.nt), kResolveThreads, 0, s>>>(
            p.g, p.counts, entries, H, W, S, bitset);
    else
        tile_resolve_kernel<false><<<static_cast<unsigned>(p.g.nt), kResolveThreads, 0, s>>>(
            p.g, p.counts, entries, H, W, S, bitset);
    S2C_LAUNCH_CHECK();
    return S2C_OK;
}

}  // namespace s2c
