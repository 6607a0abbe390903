// Streamlined low-precision MFMA GEMM core for gfx950 (FP32 16x16x4, BF16
// 16x16x32, BF16x3 = head/tail split with three MFMAs per product).
//
// Both operands are pre-converted, k-contiguous row arrays in global memory
// (row r, k at base[r * ld + k], ld a multiple of the slab depth, the padding
// zero-filled), so a slab is moved with 16-byte vector loads and 16-byte LDS
// stores and no per-element work.  Rows beyond the operand are clamped to
// row 0 on load and zeroed on store.  Slabs are double-buffered in LDS with
// register prefetch and one barrier per slab; 256 threads in a 2 x 2 wave
// layout; FP32 accumulators (layout: lane l, component r -> row 4 (l/16) + r,
// column l % 16).
#pragma once

#include "gemm_core_lowp.h"

namespace ngw {

template <int KIND>
struct V2Traits;
template <>
struct V2Traits<LP_F32> {
    typedef float elem;
    static constexpr int BK = 16, PADV = 4, NSPLIT = 1;
};
template <>
struct V2Traits<LP_BF16> {
    typedef unsigned short elem;
    static constexpr int BK = 32, PADV = 8, NSPLIT = 1;
};
template <>
struct V2Traits<LP_BF16X3> {
    typedef unsigned short elem;
    static constexpr int BK = 32, PADV = 8, NSPLIT = 2;
};

// A pre-converted operand: rows x ld elements (hi part; lo part for BF16x3).
template <int KIND>
struct OperandV2 {
    const typename V2Traits<KIND>::elem* hi;
    const typename V2Traits<KIND>::elem* lo;
    long ld;
    int rows;
};

template <int KIND, int BM, int BN>
struct SmemV2 {
    typedef V2Traits<KIND> Tr;
    typename Tr::elem a[2][Tr::NSPLIT][BM][Tr::BK + Tr::PADV];
    typename Tr::elem b[2][Tr::NSPLIT][BN][Tr::BK + Tr::PADV];
};

// One operand slab: MN rows x 64 bytes = 4 x 16-byte chunks per row.
template <int KIND, int MN>
struct SlabV2 {
    typedef V2Traits<KIND> Tr;
    static constexpr int CH = 64 / 16;             // 16-byte chunks per row
    static constexpr int PER = MN * CH / GNT;      // chunks per thread
    static constexpr int EPC = 16 / sizeof(typename Tr::elem);  // elements per chunk
    uint4 reg[Tr::NSPLIT][PER];
    const char* base[Tr::NSPLIT][PER];
    bool ok[PER];

    __device__ __forceinline__ void init(const OperandV2<KIND>& op, int mn0)
    {
        const int t = threadIdx.x;
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int v = t + GNT * i;
            const int row = mn0 + v / CH, q = v % CH;
            ok[i] = row < op.rows;
            const long off = (long)(ok[i] ? row : 0) * op.ld * (long)sizeof(typename Tr::elem) + q * 16;
            base[0][i] = reinterpret_cast<const char*>(op.hi) + off;
            if constexpr (Tr::NSPLIT == 2) base[1][i] = reinterpret_cast<const char*>(op.lo) + off;
        }
    }
    __device__ __forceinline__ void fetch(int k0)
    {
        const long koff = (long)k0 * (long)sizeof(typename Tr::elem);
#pragma unroll
        for (int sp = 0; sp < Tr::NSPLIT; ++sp)
#pragma unroll
            for (int i = 0; i < PER; ++i) reg[sp][i] = *reinterpret_cast<const uint4*>(base[sp][i] + koff);
    }
    __device__ __forceinline__ void store(typename Tr::elem (*dst)[MN][Tr::BK + Tr::PADV]) const
    {
        const int t = threadIdx.x;
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int v = t + GNT * i;
            const int row = v / CH, q = v % CH;
#pragma unroll
            for (int sp = 0; sp < Tr::NSPLIT; ++sp) {
                uint4 x = ok[i] ? reg[sp][i] : make_uint4(0u, 0u, 0u, 0u);
                *reinterpret_cast<uint4*>(&dst[sp][row][q * EPC]) = x;
            }
        }
    }
};

template <int KIND, int BM, int BN>
__device__ __forceinline__ void gemm_v2_tile(const OperandV2<KIND>& A, const OperandV2<KIND>& B, int m0, int n0,
                                             int Kpad, AccLp<BM, BN>& acc, SmemV2<KIND, BM, BN>& sm)
{
    typedef V2Traits<KIND> Tr;
    constexpr int FM = BM / 32, FN = BN / 32, BK = Tr::BK;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = (wave & 1) * (BM / 2), wn = (wave >> 1) * (BN / 2);
    const int r16 = lane & 15, g = lane >> 4;
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc.v[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    SlabV2<KIND, BM> la;
    SlabV2<KIND, BN> lb;
    la.init(A, m0);
    lb.init(B, n0);
    const int nslab = Kpad / BK;
    la.fetch(0);
    lb.fetch(0);
    la.store(sm.a[0]);
    lb.store(sm.b[0]);
    __syncthreads();
    for (int t = 0; t < nslab; ++t) {
        const int cur = t & 1;
        if (t + 1 < nslab) {
            la.fetch((t + 1) * BK);
            lb.fetch((t + 1) * BK);
        }
        if constexpr (KIND == LP_F32) {
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
            bf16x8 ah[FM], bh[FN], alo[Tr::NSPLIT == 2 ? FM : 1], blo[Tr::NSPLIT == 2 ? FN : 1];
#pragma unroll
            for (int i = 0; i < FM; ++i) {
                ah[i] = *reinterpret_cast<const bf16x8*>(&sm.a[cur][0][wm + 16 * i + r16][8 * g]);
                if constexpr (Tr::NSPLIT == 2)
                    alo[i] = *reinterpret_cast<const bf16x8*>(&sm.a[cur][1][wm + 16 * i + r16][8 * g]);
            }
#pragma unroll
            for (int j = 0; j < FN; ++j) {
                bh[j] = *reinterpret_cast<const bf16x8*>(&sm.b[cur][0][wn + 16 * j + r16][8 * g]);
                if constexpr (Tr::NSPLIT == 2)
                    blo[j] = *reinterpret_cast<const bf16x8*>(&sm.b[cur][1][wn + 16 * j + r16][8 * g]);
            }
#pragma unroll
            for (int i = 0; i < FM; ++i)
#pragma unroll
                for (int j = 0; j < FN; ++j) {
                    if constexpr (Tr::NSPLIT == 2) {
                        acc.v[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alo[i], bh[j], acc.v[i][j], 0, 0, 0);
                        acc.v[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], blo[j], acc.v[i][j], 0, 0, 0);
                    }
                    acc.v[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], acc.v[i][j], 0, 0, 0);
                }
        }
        if (t + 1 < nslab) {
            la.store(sm.a[cur ^ 1]);
            lb.store(sm.b[cur ^ 1]);
        }
        __syncthreads();
    }
}

// FP32 value -> stored element(s) of a pre-converted operand.
template <int KIND>
__device__ __forceinline__ void put_v2(typename V2Traits<KIND>::elem* hi, typename V2Traits<KIND>::elem* lo, long i,
                                       float v)
{
    if constexpr (KIND == LP_F32) {
        hi[i] = v;
    } else {
        unsigned short h = f32_to_bf16_bits(v);
        hi[i] = h;
        if constexpr (KIND == LP_BF16X3) lo[i] = f32_to_bf16_bits(v - bf16_bits_to_f32(h));
    }
}

}  // namespace ngw
