// Library runtime pieces: error state, stream-ordered temporaries, device-wide
// scans, stable compaction and fills.  Replaces the reference's per-call
// cudaMalloc/cudaFree (dda.cu:101,120; rasterize.cu:70,268,347;
// compact_utils.cu:62) and Thrust inclusive_scan (dda.cu:111-115,
// rasterize.cu:77-81).
#include <cstdarg>
#include <mutex>

#include "common.hpp"

namespace s2c {

static thread_local std::string g_err;

void set_error(const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
}

// Keep the device's default stream-ordered pool from returning memory to
// the driver at every synchronization (release threshold = unlimited).
static void configure_pool_once() {
    static std::mutex mu;
    static bool done[64] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return;
    std::lock_guard<std::mutex> lk(mu);
    if (done[dev]) return;
    hipMemPool_t pool;
    if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
        uint64_t thr = UINT64_MAX;
        (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr);
    }
    (void)hipGetLastError();
    done[dev] = true;
}

int Temp::alloc(size_t bytes, hipStream_t s) {
    configure_pool_once();
    if (ptr) {  // re-sizing: release the previous block on its stream
        (void)hipFreeAsync(ptr, stream);
        ptr = nullptr;
    }
    stream = s;
    if (bytes < 256) bytes = 256;
    S2C_HIP(hipMallocAsync(&ptr, bytes, s));
    return S2C_OK;
}

Temp::~Temp() {
    if (ptr) (void)hipFreeAsync(ptr, stream);
}

int read_device_int(const int* dptr, hipStream_t stream, int* out) {
    int v = 0;
    S2C_HIP(hipMemcpyAsync(&v, dptr, sizeof(int), hipMemcpyDeviceToHost, stream));
    S2C_HIP(hipStreamSynchronize(stream));
    *out = v;
    return S2C_OK;
}

// ------------------------------------------------------------------ fills

__global__ void fill_u32_kernel(uint32_t* p, uint32_t v, long long n) {
    long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    const long long stride = static_cast<long long>(gridDim.x) * blockDim.x;
    // 16-byte stores for the aligned bulk, scalar tail.
    const long long n4 = n / 4;
    uint4* p4 = reinterpret_cast<uint4*>(p);
    const uint4 v4 = make_uint4(v, v, v, v);
    for (long long j = i; j < n4; j += stride) p4[j] = v4;
    for (long long j = n4 * 4 + i; j < n; j += stride) p[j] = v;
}

__global__ void fill_i64_kernel(long long* p, long long v, long long n) {
    long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    const long long stride = static_cast<long long>(gridDim.x) * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

int fill_u32(void* p, uint32_t v, long long n, hipStream_t s) {
    if (n <= 0) return S2C_OK;
    S2C_CHECK((reinterpret_cast<uintptr_t>(p) & 15) == 0, "fill_u32: pointer not 16-byte aligned");
    // one 16-byte store per thread up to 2^22 workgroups (16 GiB): a 512 MiB
    // fill runs at 6.8 TB/s this way, 4.9 TB/s with 8,192 grid-striding
    // workgroups
    long long blocks = (n / 4 + kBlock - 1) / kBlock;
    if (blocks > (1ll << 22)) blocks = 1ll << 22;
    if (blocks < 1) blocks = 1;
    fill_u32_kernel<<<static_cast<unsigned>(blocks), kBlock, 0, s>>>(
        static_cast<uint32_t*>(p), v, n);
    S2C_LAUNCH_CHECK();
    return S2C_OK;
}

// ------------------------------------------------------- inclusive scan
// Reduce-then-scan over tiles of kScanTile ints: (1) tile sums, (2) one
// workgroup scans the tile sums in int64, (3) each tile rescans itself and
// adds its offset.  No inter-workgroup waiting, so no forward-progress
// assumption.

constexpr int kScanItems = 8;
constexpr int kScanTile = kBlock * kScanItems;

size_t scan_scratch_ints(long long n) {
    // int64 per tile + 1 total, counted in ints.
    return static_cast<size_t>(((n + kScanTile - 1) / kScanTile + 1) * 2);
}

__global__ void __launch_bounds__(kBlock)
scan_tile_sums(const int* __restrict__ data, long long n, long long* __restrict__ sums) {
    __shared__ int sh[kBlock / kWave + 1];
    const long long base = static_cast<long long>(blockIdx.x) * kScanTile;
    int acc = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        const long long i = base + j * kBlock + threadIdx.x;
        if (i < n) acc += data[i];
    }
    int total;
    block_excl_scan(acc, sh, &total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// Exclusive scan of `tiles` int64 values in place; the grand total goes to
// sums[tiles].
__global__ void __launch_bounds__(kWgScanThreads)
scan_tile_offsets(long long* __restrict__ sums, long long tiles) {
    wg_exclusive_scan(sums, tiles);
}

__global__ void __launch_bounds__(kBlock)
scan_tile_apply(int* __restrict__ data, long long n, const long long* __restrict__ offs) {
    __shared__ int sh[kBlock / kWave + 1];
    const long long base = static_cast<long long>(blockIdx.x) * kScanTile;
    long long run = offs[blockIdx.x];
    // Striped passes keep loads/stores coalesced; each pass is a block scan.
#pragma unroll 1
    for (int j = 0; j < kScanItems; ++j) {
        const long long i = base + j * kBlock + threadIdx.x;
        const int v = (i < n) ? data[i] : 0;
        int total;
        const int ex = block_excl_scan(v, sh, &total);
        if (i < n) data[i] = static_cast<int>(run + ex + v);
        run += total;
    }
}

const long long* scan_total_ptr(const int* scratch, long long n) {
    return reinterpret_cast<const long long*>(scratch) + (n + kScanTile - 1) / kScanTile;
}

int read_scan_total(const int* scratch, long long n, hipStream_t s, int* out, const char* who) {
    *out = 0;
    if (n <= 0) return S2C_OK;
    long long total = 0;
    S2C_HIP(hipMemcpyAsync(&total, scan_total_ptr(scratch, n), sizeof(long long),
                           hipMemcpyDeviceToHost, s));
    S2C_HIP(hipStreamSynchronize(s));
    if (total > INT_MAX) {
        set_error("%s: fragment count %lld exceeds INT32_MAX", who, total);
        return S2C_ERR_OVERFLOW;
    }
    *out = static_cast<int>(total);
    return S2C_OK;
}

int inclusive_scan_inplace(int* data, long long n, int* scratch, hipStream_t s) {
    if (n <= 0) return S2C_OK;
    long long* sums = reinterpret_cast<long long*>(scratch);
    const long long tiles = (n + kScanTile - 1) / kScanTile;
    scan_tile_sums<<<static_cast<unsigned>(tiles), kBlock, 0, s>>>(data, n, sums);
    S2C_LAUNCH_CHECK();
    scan_tile_offsets<<<1, kWgScanThreads, 0, s>>>(sums, tiles);
    S2C_LAUNCH_CHECK();
    scan_tile_apply<<<static_cast<unsigned>(tiles), kBlock, 0, s>>>(data, n, sums);
    S2C_LAUNCH_CHECK();
    return S2C_OK;
}

// Exclusive scan of int tile counts in place; total written at counts[tiles].
__global__ void __launch_bounds__(kWgScanThreads)
scan_int_tiles(int* __restrict__ counts, int tiles) {
    wg_exclusive_scan(counts, tiles);
}

int exclusive_scan_tiles(int* counts, int tiles, hipStream_t s) {
    scan_int_tiles<<<1, kWgScanThreads, 0, s>>>(counts, tiles);
    S2C_LAUNCH_CHECK();
    return S2C_OK;
}

// ------------------------------------------------------ stable compaction
// Shared by filter_valid_fragments (rasterize.cu:280-365) and
// compact_valid_ints (compact_utils.cu:7-77).  Tile = kScanTile items,
// striped passes, block scans for the in-tile order.

template <int PIX>
__global__ void __launch_bounds__(kBlock)
compact_count(const int* __restrict__ keyed, long long n, int* __restrict__ counts) {
    __shared__ int sh[kBlock / kWave + 1];
    const long long base = static_cast<long long>(blockIdx.x) * kScanTile;
    int acc = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        const long long i = base + j * kBlock + threadIdx.x;
        if (i < n && keyed[i * PIX] >= 0) ++acc;
    }
    int total;
    block_excl_scan(acc, sh, &total);
    if (threadIdx.x == 0) counts[blockIdx.x] = total;
}

// Moves items with keyed[i*PIX] >= 0; each item is PIX ints of `a` plus
// ATTR floats of `b` (ATTR may be 0).
template <int PIX, int ATTR>
__global__ void __launch_bounds__(kBlock)
compact_scatter(const int* __restrict__ a, const float* __restrict__ b, long long n,
                const int* __restrict__ offs, int* __restrict__ a_out,
                float* __restrict__ b_out) {
    __shared__ int sh[kBlock / kWave + 1];
    const long long base = static_cast<long long>(blockIdx.x) * kScanTile;
    long long run = offs[blockIdx.x];
#pragma unroll 1
    for (int j = 0; j < kScanItems; ++j) {
        const long long i = base + j * kBlock + threadIdx.x;
        const bool ok = (i < n) && a[i * PIX] >= 0;
        int total;
        const int ex = block_excl_scan(ok ? 1 : 0, sh, &total);
        if (ok) {
            const long long o = run + ex;
#pragma unroll
            for (int k = 0; k < PIX; ++k) a_out[o * PIX + k] = a[i * PIX + k];
            if (ATTR == 4) {
                reinterpret_cast<float4*>(b_out)[o] = reinterpret_cast<const float4*>(b)[i];
            }
        }
        run += total;
    }
}

}  // namespace s2c

using namespace s2c;

#ifndef S2C_BUILD_ID
#define S2C_BUILD_ID "unknown"
#endif
extern "C" int s2c_version(void) { return 100; }

extern "C" const char* s2c_build_id(void) { return S2C_BUILD_ID; }

extern "C" const char* s2c_last_error(void) { return g_err.c_str(); }

extern "C" int s2c_filter_valid_fragments(int num_frags, const int* frag_pix,
                                          const float* frag_attrs, int* frag_pix_out,
                                          float* frag_attrs_out, int* valid_count_out,
                                          void* stream) {
    g_err.clear();
    S2C_CHECK(valid_count_out != nullptr, "filter_valid_fragments: null count pointer");
    *valid_count_out = 0;
    S2C_CHECK(num_frags >= 0, "filter_valid_fragments: negative size");
    if (num_frags == 0) return S2C_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int tiles = static_cast<int>((static_cast<long long>(num_frags) + kScanTile - 1) / kScanTile);
    Temp counts;
    if (int r = counts.alloc(sizeof(int) * (tiles + 1), s)) return r;
    compact_count<3><<<tiles, kBlock, 0, s>>>(frag_pix, num_frags, counts.as<int>());
    S2C_LAUNCH_CHECK();
    if (int r = exclusive_scan_tiles(counts.as<int>(), tiles, s)) return r;
    compact_scatter<3, 4><<<tiles, kBlock, 0, s>>>(frag_pix, frag_attrs, num_frags,
                                                   counts.as<int>(), frag_pix_out,
                                                   frag_attrs_out);
    S2C_LAUNCH_CHECK();
    return read_device_int(counts.as<int>() + tiles, s, valid_count_out);
}

extern "C" int s2c_compact_valid_ints(int N, const int* input, int* output,
                                      int* valid_count_out, void* stream) {
    g_err.clear();
    S2C_CHECK(valid_count_out != nullptr, "compact_valid_ints: null count pointer");
    *valid_count_out = 0;
    S2C_CHECK(N >= 0, "compact_valid_ints: negative size");
    if (N == 0) return S2C_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int tiles = static_cast<int>((static_cast<long long>(N) + kScanTile - 1) / kScanTile);
    Temp counts;
    if (int r = counts.alloc(sizeof(int) * (tiles + 1), s)) return r;
    compact_count<1><<<tiles, kBlock, 0, s>>>(input, N, counts.as<int>());
    S2C_LAUNCH_CHECK();
    if (int r = exclusive_scan_tiles(counts.as<int>(), tiles, s)) return r;
    compact_scatter<1, 0><<<tiles, kBlock, 0, s>>>(input, nullptr, N, counts.as<int>(),
                                                   output, nullptr);
    S2C_LAUNCH_CHECK();
    return read_device_int(counts.as<int>() + tiles, s, valid_count_out);
}
