// Shared device/host helpers for the gfx950 HIP kernels of strns2cards_amd.
//
// Numerics policy: every translation unit is compiled with -ffp-contract=off
// and correctly rounded fp32 divide/sqrt, so each float expression below is
// evaluated as written, one binary32 operation at a time.  That is what
// makes pixel coordinates, ids and masks bit-exact against the CPU oracle.
#pragma once

#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/strns2cards_hip.h"

namespace s2c {

constexpr int kBlock = 256;          // threads per workgroup (4 wave64)
constexpr int kWave = 64;

// ---------------------------------------------------------------- errors

void set_error(const char* fmt, ...);

#define S2C_HIP(expr)                                                          \
    do {                                                                       \
        hipError_t e_ = (expr);                                                \
        if (e_ != hipSuccess) {                                                \
            ::s2c::set_error("%s failed: %s (%s:%d)", #expr,                   \
                             hipGetErrorString(e_), __FILE__, __LINE__);       \
            return static_cast<int>(e_);                                       \
        }                                                                      \
    } while (0)

#define S2C_CHECK(cond, ...)                                                   \
    do {                                                                       \
        if (!(cond)) {                                                         \
            ::s2c::set_error(__VA_ARGS__);                                     \
            return S2C_ERR_ARG;                                                \
        }                                                                      \
    } while (0)

#define S2C_LAUNCH_CHECK()                                                     \
    do {                                                                       \
        hipError_t e_ = hipGetLastError();                                     \
        if (e_ != hipSuccess) {                                                \
            ::s2c::set_error("kernel launch failed: %s (%s:%d)",               \
                             hipGetErrorString(e_), __FILE__, __LINE__);       \
            return static_cast<int>(e_);                                       \
        }                                                                      \
    } while (0)

// Stream-ordered temporary device memory (released on the same stream).
struct Temp {
    void* ptr = nullptr;
    hipStream_t stream = nullptr;
    Temp() = default;
    Temp(const Temp&) = delete;
    Temp& operator=(const Temp&) = delete;
    int alloc(size_t bytes, hipStream_t s);
    ~Temp();
    template <class T> T* as() const { return static_cast<T*>(ptr); }
};

// Reads one device int after the stream's pending work (a host sync).
int read_device_int(const int* dptr, hipStream_t stream, int* out);

inline unsigned grid_for(long long n, int per_block) {
    long long g = (n + per_block - 1) / per_block;
    return static_cast<unsigned>(g < 1 ? 1 : g);
}

// ------------------------------------------------------------- device math

// Saturating fp32 -> int32 truncation with NaN -> 0 (the semantics of the
// hardware conversion on both the reference's GPU and gfx950, written out so
// the compiler cannot exploit the undefined C++ overflow case).
// v_cvt_i32_f32 truncates toward zero, clamps to [INT_MIN, INT_MAX] and
// maps NaN to 0: exactly this conversion, in one instruction (a C++ cast of
// an out-of-range float would be undefined, hence the asm).
__device__ __forceinline__ int f2i(float x) {
    int r;
    asm("v_cvt_i32_f32 %0, %1" : "=v"(r) : "v"(x));
    return r;
}

// The reference's MIN / MAX macros (cuda_common.cuh:16-22), NaN behaviour
// included: the second operand wins when the comparison is false.
template <class T>
__host__ __device__ __forceinline__ T dmin(T a, T b) { return a < b ? a : b; }
template <class T>
__host__ __device__ __forceinline__ T dmax(T a, T b) { return a > b ? a : b; }

// ------------------------------------------------------- wave / block scans

__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ unsigned long long lanemask_lt() {
    const int l = lane_id();
    return (l == 0) ? 0ull : ((~0ull) >> (64 - l));
}

// Inclusive wave64 scan of an int.
__device__ __forceinline__ int wave_incl_scan(int v) {
    const int l = lane_id();
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        int t = __shfl_up(v, d, kWave);
        if (l >= d) v += t;
    }
    return v;
}

__device__ __forceinline__ long long wave_incl_scan64(long long v) {
    const int l = lane_id();
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        long long t = __shfl_up(v, d, kWave);
        if (l >= d) v += t;
    }
    return v;
}

__device__ __forceinline__ int wave_min(int v) {
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) v = dmin(v, __shfl_xor(v, d, kWave));
    return v;
}

__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) v = dmax(v, __shfl_xor(v, d, kWave));
    return v;
}

// Exclusive block scan over kBlock threads; returns the exclusive prefix and
// writes the block total.  `sh` needs kBlock/kWave + 1 ints of LDS.
__device__ __forceinline__ int block_excl_scan(int v, int* sh, int* total) {
    const int tid = threadIdx.x;
    const int w = tid / kWave;
    const int incl = wave_incl_scan(v);
    if (lane_id() == kWave - 1) sh[w] = incl;
    __syncthreads();
    if (tid == 0) {
        int acc = 0;
        for (int i = 0; i < kBlock / kWave; ++i) {
            int t = sh[i];
            sh[i] = acc;
            acc += t;
        }
        sh[kBlock / kWave] = acc;
    }
    __syncthreads();
    const int res = sh[w] + incl - v;
    *total = sh[kBlock / kWave];
    __syncthreads();
    return res;
}

// Exclusive scan of n values in place by one workgroup of kWgScanThreads;
// writes the grand total to data[n].  Each thread scans kWgScanItems
// consecutive values of a chunk in registers, so a chunk costs one block
// scan (a 256-thread form with one value per thread took 87 us for 29K
// values, one barrier round per 256).
constexpr int kWgScanThreads = 1024;
constexpr int kWgScanItems = 8;

template <typename T>
__device__ __forceinline__ T wave_incl_scan_t(T v) {
    if constexpr (sizeof(T) == 8) return wave_incl_scan64(v);
    else return wave_incl_scan(v);
}

template <typename T>
__device__ __forceinline__ void wg_exclusive_scan(T* __restrict__ data, long long n) {
    __shared__ T wsum[kWgScanThreads / kWave];
    __shared__ T carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    constexpr long long chunk = static_cast<long long>(kWgScanThreads) * kWgScanItems;
    const int w = threadIdx.x / kWave;
    for (long long c = 0; c < n; c += chunk) {
        const long long i0 = c + static_cast<long long>(threadIdx.x) * kWgScanItems;
        T v[kWgScanItems];
        T sum = 0;
#pragma unroll
        for (int k = 0; k < kWgScanItems; ++k) {
            v[k] = i0 + k < n ? data[i0 + k] : T(0);
            sum += v[k];
        }
        const T incl = wave_incl_scan_t(sum);
        if (lane_id() == kWave - 1) wsum[w] = incl;
        __syncthreads();
        T run = carry;
        for (int k = 0; k < w; ++k) run += wsum[k];
        run += incl - sum;
#pragma unroll
        for (int k = 0; k < kWgScanItems; ++k) {
            if (i0 + k < n) data[i0 + k] = run;
            run += v[k];
        }
        __syncthreads();
        if (threadIdx.x == kWgScanThreads - 1) carry = run;
        __syncthreads();
    }
    if (threadIdx.x == 0) data[n] = carry;
}

// ----------------------------------------------------- device-wide helpers
// (implemented in scan.hip)

// In-place inclusive scan of `n` int32 values.  `block_sums` scratch needs
// scan_scratch_ints(n) ints.  The last element is the total.
size_t scan_scratch_ints(long long n);
int inclusive_scan_inplace(int* data, long long n, int* scratch, hipStream_t s);
// Device address of the int64 grand total inside a scan's scratch.
const long long* scan_total_ptr(const int* scratch, long long n);
// Reads the int64 total of the last scan (a host sync); fails with
// S2C_ERR_OVERFLOW when it does not fit an int32.
int read_scan_total(const int* scratch, long long n, hipStream_t s, int* out, const char* who);

// Stable compaction helpers: per-tile counts of a predicate evaluated by the
// caller's kernel, then an exclusive scan of those counts (written in place,
// total appended at index `tiles`).
int exclusive_scan_tiles(int* counts, int tiles, hipStream_t s);

__global__ void fill_u32_kernel(uint32_t* p, uint32_t v, long long n);
__global__ void fill_i64_kernel(long long* p, long long v, long long n);
int fill_u32(void* p, uint32_t v, long long n, hipStream_t s);

}  // namespace s2c
