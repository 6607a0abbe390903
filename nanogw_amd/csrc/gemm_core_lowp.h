// Shared pieces of the low-precision MFMA GEMMs for gfx950 -- FP32
// (v_mfma_f32_16x16x4f32), BF16 (v_mfma_f32_16x16x32_bf16) and BF16x3 (each
// operand split into a BF16 head and a BF16 tail, acc += a_hi b_hi + a_hi b_lo
// + a_lo b_hi, which keeps ~16 mantissa bits of every product), all with FP32
// accumulation: kinds, BF16 rounding, accumulator tiles and their layout.
//
// Result layout of the 16x16 FP32-accumulating MFMAs (checked on gfx950 by
// the library's GEMM tests): lane l, component r holds row 4 (l/16) + r,
// column l % 16.
#pragma once

#include "gemm_core_f64.h"

namespace ngw {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));

enum LowpKind { LP_F32 = 0, LP_BF16 = 1, LP_BF16X3 = 2, LP_F16 = 3 };

__device__ __forceinline__ unsigned short f32_to_bf16_bits(float f)
{
    unsigned int u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf16_bits_to_f32(unsigned short h)
{
    return __uint_as_float(((unsigned int)h) << 16);
}
__device__ __forceinline__ unsigned short f32_to_f16_bits(float f)
{
    return __builtin_bit_cast(unsigned short, (_Float16)f);  // v_cvt_f16_f32, round to nearest even
}

// Smallest power of two >= m (1 for m == 0): exact scale factors for the FP16
// operands, whose range ends at 65504.
__device__ __forceinline__ float pow2_ceil(float m)
{
    const unsigned u = __float_as_uint(m) & 0x7FFFFFFFu;
    if (u == 0u) return 1.0f;
    unsigned e = u >> 23;
    if (u & 0x7FFFFFu) ++e;
    return __uint_as_float(e << 23);
}

template <int BM, int BN>
struct AccLp {
    static constexpr int FM = BM / 32, FN = BN / 32;
    f32x4 v[FM][FN];
};

template <int BM, int BN, class F>
__device__ __forceinline__ void for_acc_lp(const AccLp<BM, BN>& acc, F&& f)
{
    constexpr int FM = BM / 32, FN = BN / 32;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = (wave & 1) * (BM / 2), wn = (wave >> 1) * (BN / 2);
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r)
                f(wm + 16 * i + 4 * (lane >> 4) + r, wn + 16 * j + (lane & 15), acc.v[i][j][r]);
}

}  // namespace ngw
