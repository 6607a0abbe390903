// Low-precision MFMA GEMM core for gfx950: FP32 (v_mfma_f32_16x16x4f32),
// BF16 (v_mfma_f32_16x16x32_bf16) and BF16x3 (each operand split into a
// BF16 head and a BF16 tail, acc += a_hi b_hi + a_hi b_lo + a_lo b_hi, which
// keeps ~16 mantissa bits of every product), all with FP32 accumulation.
//
// Same structure as the FP64 core: 256 threads in a 2 x 2 wave layout,
// k-contiguous LDS rows, double-buffered slabs with register prefetch and one
// barrier per slab, loader functors with separate raw fetch and combine.
// Operands are converted from the loader's FP32 value to the MFMA type when
// a slab is written to LDS.
//
// Result layout of the 16x16 FP32-accumulating MFMAs (checked on gfx950 by
// the library's GEMM tests): lane l, component r holds row 4 (l/16) + r,
// column l % 16.
#pragma once

#include "gemm_core_f64.h"

namespace ngw {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

enum LowpKind { LP_F32 = 0, LP_BF16 = 1, LP_BF16X3 = 2 };

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

// Per-kind slab geometry and LDS element type.
template <int KIND>
struct LpTraits;
template <>
struct LpTraits<LP_F32> {
    static constexpr int BK = 16, PAD = 4, NSPLIT = 1;
    typedef float elem;
};
template <>
struct LpTraits<LP_BF16> {
    static constexpr int BK = 32, PAD = 8, NSPLIT = 1;
    typedef unsigned short elem;
};
template <>
struct LpTraits<LP_BF16X3> {
    static constexpr int BK = 32, PAD = 8, NSPLIT = 2;
    typedef unsigned short elem;
};

template <int KIND, int BM, int BN>
struct SmemLp {
    typedef LpTraits<KIND> Tr;
    typename Tr::elem a[2][Tr::NSPLIT][BM][Tr::BK + Tr::PAD];
    typename Tr::elem b[2][Tr::NSPLIT][BN][Tr::BK + Tr::PAD];
};

template <int KIND, class L, int MN>
struct SlabLoaderLp {
    typedef LpTraits<KIND> Tr;
    static constexpr int BK = Tr::BK;
    static constexpr int PER = MN * BK / GNT;
    static constexpr int NH = L::kContig ? PER : 1;
    typename L::Handle h[NH];
    bool ok[NH];
    int kk0;

    __device__ __forceinline__ void init(const L& ld, int mn0)
    {
        const int t = threadIdx.x;
        if (L::kContig) {
            kk0 = t % BK;
#pragma unroll
            for (int i = 0; i < PER; ++i) {
                int mn = mn0 + t / BK + (GNT / BK) * i;
                ok[i] = mn < ld.MN;
                h[i] = ld.row(ok[i] ? mn : 0);
            }
        } else {
            kk0 = t / MN;
            int mn = mn0 + t % MN;
            ok[0] = mn < ld.MN;
            h[0] = ld.row(ok[0] ? mn : 0);
        }
    }
    __device__ __forceinline__ int kof(int k0, int i) const
    {
        return L::kContig ? k0 + kk0 : k0 + kk0 + (GNT / MN) * i;
    }
    __device__ __forceinline__ void fetch(const L& ld, int k0, int K,
                                          typename L::Raw (&raw)[PER][L::NRAW]) const
    {
#pragma unroll
        for (int i = 0; i < PER; ++i) ld.fetch(h[L::kContig ? i : 0], min(kof(k0, i), K - 1), raw[i]);
    }
    __device__ __forceinline__ void store(const L& ld, int k0, int K,
                                          typename Tr::elem (*dst)[MN][BK + Tr::PAD],
                                          const typename L::Raw (&raw)[PER][L::NRAW]) const
    {
        const int t = threadIdx.x;
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const bool valid = ok[L::kContig ? i : 0] && kof(k0, i) < K;
            const float v = valid ? ld.combine(raw[i]) : 0.0f;
            const int r = L::kContig ? t / BK + (GNT / BK) * i : t % MN;
            const int c = L::kContig ? kk0 : kk0 + (GNT / MN) * i;
            if constexpr (KIND == LP_F32) {
                dst[0][r][c] = v;
            } else if constexpr (KIND == LP_BF16) {
                dst[0][r][c] = f32_to_bf16_bits(v);
            } else {
                unsigned short hi = f32_to_bf16_bits(v);
                dst[0][r][c] = hi;
                dst[1][r][c] = f32_to_bf16_bits(v - bf16_bits_to_f32(hi));
            }
        }
    }
};

template <int BM, int BN>
struct AccLp {
    static constexpr int FM = BM / 32, FN = BN / 32;
    f32x4 v[FM][FN];
};

template <int KIND, int BM, int BN, class AL, class BL>
__device__ __forceinline__ void gemm_lp_tile(const AL& al, const BL& bl, int m0, int n0, int K,
                                             AccLp<BM, BN>& acc, SmemLp<KIND, BM, BN>& sm)
{
    typedef LpTraits<KIND> Tr;
    constexpr int FM = BM / 32, FN = BN / 32, BK = Tr::BK;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = (wave & 1) * (BM / 2), wn = (wave >> 1) * (BN / 2);
    const int r16 = lane & 15, g = lane >> 4;
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc.v[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    SlabLoaderLp<KIND, AL, BM> la;
    SlabLoaderLp<KIND, BL, BN> lb;
    la.init(al, m0);
    lb.init(bl, n0);
    typename AL::Raw ra[SlabLoaderLp<KIND, AL, BM>::PER][AL::NRAW];
    typename BL::Raw rb[SlabLoaderLp<KIND, BL, BN>::PER][BL::NRAW];
    const int nslab = (K + BK - 1) / BK;
    la.fetch(al, 0, K, ra);
    lb.fetch(bl, 0, K, rb);
    la.store(al, 0, K, sm.a[0], ra);
    lb.store(bl, 0, K, sm.b[0], rb);
    __syncthreads();
    for (int t = 0; t < nslab; ++t) {
        const int cur = t & 1;
        if (t + 1 < nslab) {
            la.fetch(al, (t + 1) * BK, K, ra);
            lb.fetch(bl, (t + 1) * BK, K, rb);
        }
        if constexpr (KIND == LP_F32) {
            // lane group g, sub-step s uses k = 4 g + s
            float af[FM][4], bf[FN][4];
#pragma unroll
            for (int i = 0; i < FM; ++i) {
                float4 x = *reinterpret_cast<const float4*>(&sm.a[cur][0][wm + 16 * i + r16][4 * g]);
                af[i][0] = x.x; af[i][1] = x.y; af[i][2] = x.z; af[i][3] = x.w;
            }
#pragma unroll
            for (int j = 0; j < FN; ++j) {
                float4 x = *reinterpret_cast<const float4*>(&sm.b[cur][0][wn + 16 * j + r16][4 * g]);
                bf[j][0] = x.x; bf[j][1] = x.y; bf[j][2] = x.z; bf[j][3] = x.w;
            }
#pragma unroll
            for (int s = 0; s < 4; ++s)
#pragma unroll
                for (int i = 0; i < FM; ++i)
#pragma unroll
                    for (int j = 0; j < FN; ++j)
                        acc.v[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[i][s], bf[j][s], acc.v[i][j], 0, 0, 0);
        } else {
            // lane group g supplies k = 8 g .. 8 g + 7 of the 32-deep slab
            bf16x8 ah[FM], bh[FN], alo[KIND == LP_BF16X3 ? FM : 1], blo[KIND == LP_BF16X3 ? FN : 1];
#pragma unroll
            for (int i = 0; i < FM; ++i) {
                ah[i] = *reinterpret_cast<const bf16x8*>(&sm.a[cur][0][wm + 16 * i + r16][8 * g]);
                if constexpr (KIND == LP_BF16X3)
                    alo[i] = *reinterpret_cast<const bf16x8*>(&sm.a[cur][1][wm + 16 * i + r16][8 * g]);
            }
#pragma unroll
            for (int j = 0; j < FN; ++j) {
                bh[j] = *reinterpret_cast<const bf16x8*>(&sm.b[cur][0][wn + 16 * j + r16][8 * g]);
                if constexpr (KIND == LP_BF16X3)
                    blo[j] = *reinterpret_cast<const bf16x8*>(&sm.b[cur][1][wn + 16 * j + r16][8 * g]);
            }
#pragma unroll
            for (int i = 0; i < FM; ++i)
#pragma unroll
                for (int j = 0; j < FN; ++j) {
                    if constexpr (KIND == LP_BF16X3) {
                        acc.v[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alo[i], bh[j], acc.v[i][j], 0, 0, 0);
                        acc.v[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], blo[j], acc.v[i][j], 0, 0, 0);
                    }
                    acc.v[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], acc.v[i][j], 0, 0, 0);
                }
        }
        if (t + 1 < nslab) {
            la.store(al, (t + 1) * BK, K, sm.a[cur ^ 1], ra);
            lb.store(bl, (t + 1) * BK, K, sm.b[cur ^ 1], rb);
        }
        __syncthreads();
    }
}

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
