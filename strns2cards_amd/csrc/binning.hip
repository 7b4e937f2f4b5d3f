// Tile offsets and the per-tile LDS resolve of the binned cluster-bitset
// passes (see binning.hpp).
#include "binning.hpp"

namespace s2c {

// Exclusive scan of nt tile counts by one workgroup; offsets stay in
// counts[0..nt] (total at counts[nt]) and are copied to the cursors.
__global__ void __launch_bounds__(kBlock)
tile_offsets_kernel(int* __restrict__ counts, int* __restrict__ cursor, int nt) {
    __shared__ int wsum[kBlock / kWave];
    __shared__ int carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int c = 0; c < nt; c += kBlock) {
        const int i = c + threadIdx.x;
        const int v = (i < nt) ? counts[i] : 0;
        const int incl = wave_incl_scan(v);
        const int w = threadIdx.x / kWave;
        if (lane_id() == kWave - 1) wsum[w] = incl;
        __syncthreads();
        int before = carry;
        for (int k = 0; k < w; ++k) before += wsum[k];
        if (i < nt) {
            counts[i] = before + incl - v;
            cursor[i] = before + incl - v;
        }
        __syncthreads();
        if (threadIdx.x == kBlock - 1) carry = before + incl;
        __syncthreads();
    }
    if (threadIdx.x == 0) counts[nt] = carry;
}

template <bool VEC4>
__global__ void __launch_bounds__(kBlock)
tile_resolve_kernel(TileGrid g, const int* __restrict__ offs, const unsigned* __restrict__ entries,
                    int H, int W, int S, int* __restrict__ bitset) {
    __shared__ unsigned tile[kTileLdsBytes / 4];
    const int t = blockIdx.x;
    const int side = 1 << g.ts_log2;
    const int words = side * side * S;
    for (int i = threadIdx.x; i < words; i += kBlock) tile[i] = 0u;
    __syncthreads();
    const int e0 = offs[t], e1 = offs[t + 1];
    const unsigned lmask = (1u << g.local_bits) - 1u;
    for (int e = e0 + threadIdx.x; e < e1; e += kBlock) {
        const unsigned v = entries[e];
        const unsigned local = v & lmask;
        const unsigned k = v >> g.local_bits;
        atomicOr(&tile[local * S + (k >> 5)], 1u << (k & 31));
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
        for (int i = threadIdx.x; i < rows * per_row; i += kBlock) {
            const int ly = i / per_row;
            const int rem = i - ly * per_row;
            const long long dst =
                ((static_cast<long long>(b) * H + y0 + ly) * W + x0) * q + rem;
            reinterpret_cast<uint4*>(bitset)[dst] = src[ly * side * q + rem];
        }
    } else {
        const int per_row = row_w * S;
        for (int i = threadIdx.x; i < rows * per_row; i += kBlock) {
            const int ly = i / per_row;
            const int rem = i - ly * per_row;
            const long long dst = ((static_cast<long long>(b) * H + y0 + ly) * W + x0) * S + rem;
            bitset[dst] = static_cast<int>(tile[ly * side * S + rem]);
        }
    }
}

int bin_prepare(BinPlan& p, hipStream_t s) {
    return fill_u32(p.counts, 0u, p.g.nt + 1, s);
}

int bin_offsets(BinPlan& p, hipStream_t s, int* total) {
    tile_offsets_kernel<<<1, kBlock, 0, s>>>(p.counts, p.cursor, static_cast<int>(p.g.nt));
    S2C_LAUNCH_CHECK();
    return read_device_int(p.counts + p.g.nt, s, total);
}

int bin_resolve(const BinPlan& p, const unsigned* entries, int H, int W, int S, int* bitset,
                hipStream_t s) {
    const bool vec = (S % 4 == 0) && ((reinterpret_cast<uintptr_t>(bitset) & 15) == 0);
    if (vec)
        tile_resolve_kernel<true><<<static_cast<unsigned>(p.g.nt), kBlock, 0, s>>>(
            p.g, p.counts, entries, H, W, S, bitset);
    else
        tile_resolve_kernel<false><<<static_cast<unsigned>(p.g.nt), kBlock, 0, s>>>(
            p.g, p.counts, entries, H, W, S, bitset);
    S2C_LAUNCH_CHECK();
    return S2C_OK;
}

}  // namespace s2c
