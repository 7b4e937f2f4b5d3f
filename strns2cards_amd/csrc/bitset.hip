// Per-pixel cluster-bitset reductions (reference:
// pystrns2cards/src/bitset_utils.cu).
#include "common.hpp"

namespace s2c {

// bitset_utils.cu:6-35.  Colors are added in ascending cluster id, the
// reference's order, so the fp32 sums are bit-exact with the oracle.
// Words are fetched 16 bytes at a time when S is a multiple of 4.
template <bool VEC4>
__global__ void __launch_bounds__(kBlock)
accumulate_rgb_kernel(long long npix, int S, const int* __restrict__ bitset,
                      const float* __restrict__ rgb, float* __restrict__ out) {
    const long long i = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= npix) return;
    float r = 0.f, g = 0.f, bl = 0.f;
    const int* row = bitset + i * S;
    auto add_word = [&](uint32_t word, int s) {
        while (word) {
            const int j = __builtin_ctz(word);
            word &= word - 1;
            const int k = s * 32 + j;
            r += rgb[k * 3 + 0];
            g += rgb[k * 3 + 1];
            bl += rgb[k * 3 + 2];
        }
    };
    if (VEC4) {
        for (int s = 0; s < S; s += 4) {
            const int4 w4 = *reinterpret_cast<const int4*>(row + s);
            add_word(static_cast<uint32_t>(w4.x), s + 0);
            add_word(static_cast<uint32_t>(w4.y), s + 1);
            add_word(static_cast<uint32_t>(w4.z), s + 2);
            add_word(static_cast<uint32_t>(w4.w), s + 3);
        }
    } else {
        for (int s = 0; s < S; ++s) add_word(static_cast<uint32_t>(row[s]), s);
    }
    out[i * 3 + 0] = r;
    out[i * 3 + 1] = g;
    out[i * 3 + 2] = bl;
}

// bitset_utils.cu:37-52.
template <bool VEC4>
__global__ void __launch_bounds__(kBlock)
popcount_kernel(long long npix, int S, const int* __restrict__ bitset, float* __restrict__ out) {
    const long long i = static_cast<long long>(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= npix) return;
    const int* row = bitset + i * S;
    int c = 0;
    if (VEC4) {
        for (int s = 0; s < S; s += 4) {
            const int4 w4 = *reinterpret_cast<const int4*>(row + s);
            c += __popc(static_cast<unsigned>(w4.x)) + __popc(static_cast<unsigned>(w4.y)) +
                 __popc(static_cast<unsigned>(w4.z)) + __popc(static_cast<unsigned>(w4.w));
        }
    } else {
        for (int s = 0; s < S; ++s) c += __popc(static_cast<unsigned>(row[s]));
    }
    out[i] = static_cast<float>(c);
}

}  // namespace s2c

using namespace s2c;

static bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; }

extern "C" int s2c_accumulate_bitset_rgb(int B, int H, int W, int num_slots, const int* bitset,
                                         const float* cluster_rgb, float* accum_color,
                                         void* stream) {
    S2C_CHECK(B >= 0 && H >= 0 && W >= 0 && num_slots >= 0,
              "accumulate_bitset_rgb: negative size");
    const long long n = static_cast<long long>(B) * H * W;
    if (n == 0) return S2C_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (num_slots % 4 == 0 && aligned16(bitset))
        accumulate_rgb_kernel<true><<<grid_for(n, kBlock), kBlock, 0, s>>>(n, num_slots, bitset,
                                                                           cluster_rgb, accum_color);
    else
        accumulate_rgb_kernel<false><<<grid_for(n, kBlock), kBlock, 0, s>>>(
            n, num_slots, bitset, cluster_rgb, accum_color);
    S2C_LAUNCH_CHECK();
    return S2C_OK;
}

extern "C" int s2c_popcount_bitset(int B, int H, int W, int num_slots, const int* bitset,
                                   float* count_image, void* stream) {
    S2C_CHECK(B >= 0 && H >= 0 && W >= 0 && num_slots >= 0, "popcount_bitset: negative size");
    const long long n = static_cast<long long>(B) * H * W;
    if (n == 0) return S2C_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (num_slots % 4 == 0 && aligned16(bitset))
        popcount_kernel<true><<<grid_for(n, kBlock), kBlock, 0, s>>>(n, num_slots, bitset,
                                                                     count_image);
    else
        popcount_kernel<false><<<grid_for(n, kBlock), kBlock, 0, s>>>(n, num_slots, bitset,
                                                                      count_image);
    S2C_LAUNCH_CHECK();
    return S2C_OK;
}
