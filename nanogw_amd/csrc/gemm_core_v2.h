// Streamlined low-precision MFMA GEMM core for gfx950 (FP32 16x16x4, BF16
// and FP16 16x16x32, BF16x3 = head/tail split with three MFMAs per product).
//
// Both operands are pre-converted, k-contiguous row arrays in global memory
// (row r, k at base[r * ld + k], ld a multiple of the slab depth, the padding
// zero-filled), so a slab is moved with 16-byte vector loads and 16-byte LDS
// stores and no per-element work.  Rows beyond the operand read its all-zero
// row (see OperandV2).  Slabs are double-buffered in LDS with
// register prefetch and one barrier per slab; 256 threads in a 2 x 2 wave
// layout; FP32 accumulators (layout: lane l, component r -> row 4 (l/16) + r,
// column l % 16).
#pragma once

#include "gemm_core_lowp.h"

#include <type_traits>

namespace ngw {

// LDS row layout of a slab (four 16-byte chunks of k per row), chosen per
// kernel instance:
//   SWZ 0: rows padded to 80 bytes (chunk c of row r at byte 80 r + 16 c);
//   SWZ 1: 64-byte rows, chunk c of row r stored at chunk c ^ ((r >> 1) & 3).
// gfx950 serves a ds_read_b128 / ds_write_b128 eight lanes (128 bytes, all
// 32 banks) at a time.  With SWZ 1, the eight lanes of a fragment read (rows
// r..r+7, one chunk) and of a slab store (rows r, r+1, chunks 0..3) each
// cover every bank once; measured on the FP16 K3 GEMM, bank-conflict cycles
// drop from 48% of the LDS-active cycles (SWZ 0) to none.
template <int SWZ>
__device__ __forceinline__ int lds_chunk(int row, int chunk)
{
    if constexpr (SWZ == 1) return chunk ^ ((row >> 1) & 3);
    else return chunk;
}

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
template <>
struct V2Traits<LP_F16> {
    typedef unsigned short elem;
    static constexpr int BK = 32, PADV = 8, NSPLIT = 1;
};

// FP32 value -> 16-bit element of a BF16 / BF16x3-head / FP16 operand.
template <int KIND>
__device__ __forceinline__ unsigned short to_e16(float v)
{
    if constexpr (KIND == LP_F16) return f32_to_f16_bits(v);
    else return f32_to_bf16_bits(v);
}

// A pre-converted operand: rows x ld elements (hi part; lo part for BF16x3),
// followed by one all-zero row (row `rows`), which tiles read for their rows
// past the operand, so no masking sits between a load and its LDS store.
template <int KIND>
struct OperandV2 {
    const typename V2Traits<KIND>::elem* hi;
    const typename V2Traits<KIND>::elem* lo;
    long ld;
    int rows;
};

// LDS slabs.  Rows are 16-byte aligned (pitch 80 bytes for SWZ 0, 64 for
// SWZ 1) so operand reads are single ds_read_b128s.
template <int KIND, int BM, int BN, int SWZ = 0>
struct alignas(16) SmemV2 {
    typedef V2Traits<KIND> Tr;
    static constexpr int PITCH = Tr::BK + (SWZ ? 0 : Tr::PADV);
    alignas(16) typename Tr::elem a[2][Tr::NSPLIT][BM][PITCH];
    alignas(16) typename Tr::elem b[2][Tr::NSPLIT][BN][PITCH];
};

// Loads through global-address-space pointers: global_load instructions count
// only in vmcnt, so the compiler can wait for an older slab while a newer
// one is still in flight (generic pointers give flat loads, which also count
// in lgkmcnt and force full waits).
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

template <class T>
__device__ __forceinline__ T gload(const void* p)
{
    return *(const __attribute__((address_space(1))) T*)(p);
}

// Compile-time loop: f(std::integral_constant<int, 0>) ... f(<N - 1>).
template <int I, int N, class F>
__device__ __forceinline__ void static_for_impl(F&& f)
{
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        static_for_impl<I + 1, N>(f);
    }
}
template <int N, class F>
__device__ __forceinline__ void static_for(F&& f)
{
    static_for_impl<0, N>(f);
}

// Padding of the operands' k extent (a multiple of every slab depth).
constexpr int LP_KPAD = 64;

// One operand slab: MN rows x 64 bytes = 4 x 16-byte chunks per row.  The
// loader keeps addresses; the fetched bytes live in a Raw register set so
// that several slabs can be in flight (see gemm_v2_run).
template <int KIND, int MN>
struct SlabV2 {
    typedef V2Traits<KIND> Tr;
    static constexpr int CH = 64 / 16;             // 16-byte chunks per row
    static constexpr int PER = MN * CH / GNT;      // chunks per thread
    static constexpr int EPC = 16 / sizeof(typename Tr::elem);  // elements per chunk
    static_assert(PER >= 1 && MN * CH == PER * GNT, "slab split");
    struct Raw {
        u32x4 r[Tr::NSPLIT][PER];
    };
    const char* base[Tr::NSPLIT];
    long rowstep;  // bytes between the rows of consecutive chunks of a thread
    int row0, rows_left;  // rows_left = the zero row, relative to the tile

    __device__ __forceinline__ void init(const OperandV2<KIND>& op, int mn0)
    {
        const int t = threadIdx.x;
        row0 = t / CH;
        rows_left = op.rows - mn0;  // local rows < rows_left are inside the operand
        const long es = (long)sizeof(typename Tr::elem);
        const long off = (long)mn0 * op.ld * es + (t % CH) * 16;
        base[0] = reinterpret_cast<const char*>(op.hi) + off;
        if constexpr (Tr::NSPLIT == 2) base[1] = reinterpret_cast<const char*>(op.lo) + off;
        rowstep = op.ld * es;
    }
    __device__ __forceinline__ void fetch(int k0, Raw& raw) const
    {
        const long koff = (long)k0 * (long)sizeof(typename Tr::elem);
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            int row = row0 + (GNT / CH) * i;
            row = row < rows_left ? row : rows_left;  // rows past the operand read the zero row
#pragma unroll
            for (int sp = 0; sp < Tr::NSPLIT; ++sp) raw.r[sp][i] = gload<u32x4>(base[sp] + (long)row * rowstep + koff);
        }
    }
    template <int SWZ, int PITCH>
    __device__ __forceinline__ void store(const Raw& raw, typename Tr::elem (*dst)[MN][PITCH]) const
    {
        const int q = threadIdx.x % CH;
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int row = row0 + (GNT / CH) * i;
#pragma unroll
            for (int sp = 0; sp < Tr::NSPLIT; ++sp)
                *reinterpret_cast<u32x4*>(&dst[sp][row][lds_chunk<SWZ>(row, q) * EPC]) = raw.r[sp][i];
        }
    }
};

// B-operand slab of W(j, k) = p(v, k) u(b, k), j = v + nv b, formed while
// loading: p (nv rows) and u (rows b; row `zero_row` all zeros) are FP32 row
// arrays of leading dimension ld, small enough to stay in L2, so W never
// exists in memory.  Rows j >= rows read u's zero row.  Each thread handles
// CHK consecutive k of one row; the product and the conversion happen at
// LDS-store time.
template <int KIND, int MN>
struct SlabW {
    typedef V2Traits<KIND> Tr;
    static constexpr int CHK = MN * Tr::BK / GNT;   // elements per thread per slab
    static constexpr int TPR = Tr::BK / CHK;        // threads per row
    static_assert(CHK % 8 == 0 && TPR * CHK == Tr::BK && MN * TPR == GNT, "slab split");
    static_assert(KIND != LP_F16, "FP16 W needs its per-vector scale: formed by wprep");
    struct Raw {
        f32x4 pv[CHK / 4], pu[CHK / 4];
    };
    const float* prow;
    const float* urow;
    int row, kq;

    __device__ __forceinline__ void init(const float* p, const float* u, long ld, int nv, int rows, int zero_row,
                                         int mn0)
    {
        row = threadIdx.x / TPR;
        kq = (threadIdx.x % TPR) * CHK;
        const int j = mn0 + row;
        const int b = j < rows ? j / nv : zero_row, v = j < rows ? j - (j / nv) * nv : 0;
        prow = p + (long)v * ld + kq;
        urow = u + (long)b * ld + kq;
    }
    __device__ __forceinline__ void fetch(int k0, Raw& raw) const
    {
#pragma unroll
        for (int i = 0; i < CHK / 4; ++i) {
            raw.pv[i] = gload<f32x4>(prow + k0 + 4 * i);
            raw.pu[i] = gload<f32x4>(urow + k0 + 4 * i);
        }
    }
    template <int SWZ, int PITCH>
    __device__ __forceinline__ void store(const Raw& raw, typename Tr::elem (*dst)[MN][PITCH]) const
    {
        float w[CHK];
#pragma unroll
        for (int i = 0; i < CHK / 4; ++i) {
            const f32x4 q = raw.pv[i] * raw.pu[i];
            w[4 * i] = q.x;
            w[4 * i + 1] = q.y;
            w[4 * i + 2] = q.z;
            w[4 * i + 3] = q.w;
        }
        if constexpr (KIND == LP_F32) {
#pragma unroll
            for (int i = 0; i < CHK / 4; ++i)
                *reinterpret_cast<f32x4*>(&dst[0][row][4 * lds_chunk<SWZ>(row, kq / 4 + i)]) =
                    f32x4{w[4 * i], w[4 * i + 1], w[4 * i + 2], w[4 * i + 3]};
        } else {
#pragma unroll
            for (int i = 0; i < CHK / 8; ++i) {
                unsigned hi[4], lo[4];
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const float a = w[8 * i + 2 * e], c = w[8 * i + 2 * e + 1];
                    const unsigned short ha = f32_to_bf16_bits(a), hc = f32_to_bf16_bits(c);
                    hi[e] = (unsigned)ha | ((unsigned)hc << 16);
                    if constexpr (Tr::NSPLIT == 2) {
                        const unsigned short la = f32_to_bf16_bits(a - bf16_bits_to_f32(ha));
                        const unsigned short lc = f32_to_bf16_bits(c - bf16_bits_to_f32(hc));
                        lo[e] = (unsigned)la | ((unsigned)lc << 16);
                    }
                }
                *reinterpret_cast<u32x4*>(&dst[0][row][8 * lds_chunk<SWZ>(row, kq / 8 + i)]) = u32x4{hi[0], hi[1], hi[2], hi[3]};
                if constexpr (Tr::NSPLIT == 2)
                    *reinterpret_cast<u32x4*>(&dst[1][row][8 * lds_chunk<SWZ>(row, kq / 8 + i)]) = u32x4{lo[0], lo[1], lo[2], lo[3]};
            }
        }
    }
};

// MFMA work on one LDS slab: the wave's (BM/2) x (BN/2) block of the tile.
// SM is SmemV2 or SmemDma (buffers a[cur][plane][row][k], b likewise).
template <int KIND, int BM, int BN, int SWZ, class SM>
__device__ __forceinline__ void mfma_slab(const SM& sm, int cur, AccLp<BM, BN>& acc)
{
    typedef V2Traits<KIND> Tr;
    constexpr int FM = BM / 32, FN = BN / 32;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = (wave & 1) * (BM / 2), wn = (wave >> 1) * (BN / 2);
    const int r16 = lane & 15, g = lds_chunk<SWZ>(r16, lane >> 4);  // rows wm + 16 i + r16 share r16's swizzle
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
    } else if constexpr (KIND == LP_F16) {
        f16x8 ah[FM], bh[FN];
#pragma unroll
        for (int i = 0; i < FM; ++i) ah[i] = *reinterpret_cast<const f16x8*>(&sm.a[cur][0][wm + 16 * i + r16][8 * g]);
#pragma unroll
        for (int j = 0; j < FN; ++j) bh[j] = *reinterpret_cast<const f16x8*>(&sm.b[cur][0][wn + 16 * j + r16][8 * g]);
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j)
                acc.v[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[i], bh[j], acc.v[i][j], 0, 0, 0);
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
}

// C tile (BM x BN) = A B^T over Kpad (a multiple of the slab depth).  LA / LB
// expose fetch(k0, raw) (global -> registers) and store(raw, dst) (registers
// -> LDS).  D register sets keep D slabs in flight: slab t + D is requested
// while slab t is multiplied, and slab t + 1 (requested D - 1 slabs earlier)
// is written to the other LDS buffer; one barrier per slab.  The steady loop
// runs whole groups of D slabs and is branch-free (requests past the last
// slab re-read it); the last nslab mod D slabs follow in an unrolled tail.
template <int KIND, int BM, int BN, int D, int SWZ, class LA, class LB>
__device__ __forceinline__ void gemm_v2_run(const LA& la, const LB& lb, int Kpad, AccLp<BM, BN>& acc,
                                            SmemV2<KIND, BM, BN, SWZ>& sm)
{
    typedef V2Traits<KIND> Tr;
    constexpr int FM = BM / 32, FN = BN / 32, BK = Tr::BK;
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc.v[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int nslab = Kpad / BK;
    const int last = (nslab - 1) * BK;
    typename LA::Raw ra[D];
    typename LB::Raw rb[D];
#pragma unroll
    for (int s = 0; s < D; ++s) {
        const int k = min(s * BK, last);
        la.fetch(k, ra[s]);
        lb.fetch(k, rb[s]);
    }
    la.template store<SWZ>(ra[0], sm.a[0]);
    lb.template store<SWZ>(rb[0], sm.b[0]);
    __syncthreads();
    // one slab; u = t mod D is a compile-time register-set index
    auto slab = [&](int t, auto uc) {
        constexpr int u = decltype(uc)::value;
        const int k = min((t + D) * BK, last);
        la.fetch(k, ra[u]);
        lb.fetch(k, rb[u]);
        mfma_slab<KIND, BM, BN, SWZ>(sm, t & 1, acc);
        la.template store<SWZ>(ra[(u + 1) % D], sm.a[(t + 1) & 1]);
        lb.template store<SWZ>(rb[(u + 1) % D], sm.b[(t + 1) & 1]);
        __syncthreads();
    };
    const int nfull = nslab - nslab % D;
    int t0 = 0;
    for (; t0 < nfull; t0 += D) {
        static_for<D>([&](auto uc) { slab(t0 + decltype(uc)::value, uc); });
    }
    const int rem = nslab - nfull;
    static_for<D - 1>([&](auto uc) {
        if (decltype(uc)::value < rem) slab(t0 + decltype(uc)::value, uc);
    });
}

template <int KIND, int BM, int BN, int D, int SWZ>
__device__ __forceinline__ void gemm_v2_tile(const OperandV2<KIND>& A, const OperandV2<KIND>& B, int m0, int n0,
                                             int Kpad, AccLp<BM, BN>& acc, SmemV2<KIND, BM, BN, SWZ>& sm)
{
    SlabV2<KIND, BM> la;
    SlabV2<KIND, BN> lb;
    la.init(A, m0);
    lb.init(B, n0);
    gemm_v2_run<KIND, BM, BN, D>(la, lb, Kpad, acc, sm);
}

// FP32 value -> stored element(s) of a pre-converted operand.
template <int KIND>
__device__ __forceinline__ void put_v2(typename V2Traits<KIND>::elem* hi, typename V2Traits<KIND>::elem* lo, long i,
                                       float v)
{
    if constexpr (KIND == LP_F32) {
        hi[i] = v;
    } else if constexpr (KIND == LP_F16) {
        hi[i] = f32_to_f16_bits(v);
    } else {
        unsigned short h = f32_to_bf16_bits(v);
        hi[i] = h;
        if constexpr (KIND == LP_BF16X3) lo[i] = f32_to_bf16_bits(v - bf16_bits_to_f32(h));
    }
}

// ---------------------------------------------------------------- direct-to-LDS slabs
// gfx950's global_load_lds_dwordx4 moves 16 bytes per lane from global memory
// straight into LDS: the 64 lanes of one instruction fill one 1 KB block, lane
// l at byte 16 l (the block base goes in M0).  With 64-byte rows and the SWZ 1
// chunk swizzle, block bi of an operand slab holds rows 16 bi .. 16 bi + 15,
// and lane l fills row 16 bi + l / 4 at position l % 4, so it loads chunk
// (l % 4) ^ ((row >> 1) & 3) of that row (lds_chunk is its own inverse).  No
// register holds slab data and no ds_write is issued.  A slab is ready once
// the loads of every wave have passed vmcnt and the workgroup has met at a
// barrier; NB buffers keep NB - 1 slabs in flight.
template <int KIND, int BM, int BN, int NB>
struct alignas(16) SmemDma {
    typedef V2Traits<KIND> Tr;
    alignas(16) typename Tr::elem a[NB][Tr::NSPLIT][BM][Tr::BK];
    alignas(16) typename Tr::elem b[NB][Tr::NSPLIT][BN][Tr::BK];
};

template <int KIND, int MN>
struct DmaSlab {
    typedef V2Traits<KIND> Tr;
    static constexpr int PERW = MN / 64;            // instructions per wave per plane
    static constexpr int NLOAD = PERW * Tr::NSPLIT; // per thread per slab
    static_assert(MN % 64 == 0, "direct-to-LDS slab: rows a multiple of 64");
    const char* src[Tr::NSPLIT][PERW];

    __device__ __forceinline__ void init(const OperandV2<KIND>& op, int mn0)
    {
        const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
        const long es = (long)sizeof(typename Tr::elem);
        const int rows_left = op.rows - mn0;  // rows past the operand read its zero row
#pragma unroll
        for (int i = 0; i < PERW; ++i) {
            const int row = 16 * (4 * i + wave) + (lane >> 2);
            const int r = row < rows_left ? row : rows_left;
            const long off = ((long)mn0 + r) * op.ld * es + 16 * lds_chunk<1>(row, lane & 3);
            src[0][i] = reinterpret_cast<const char*>(op.hi) + off;
            if constexpr (Tr::NSPLIT == 2) src[1][i] = reinterpret_cast<const char*>(op.lo) + off;
        }
    }
    // loads of the slab at k0 into dst[plane][row][k]
    __device__ __forceinline__ void issue(int k0, typename Tr::elem (*dst)[MN][Tr::BK]) const
    {
        const int wave = threadIdx.x >> 6;
        const long koff = (long)k0 * (long)sizeof(typename Tr::elem);
#pragma unroll
        for (int sp = 0; sp < Tr::NSPLIT; ++sp)
#pragma unroll
            for (int i = 0; i < PERW; ++i)
                __builtin_amdgcn_global_load_lds(
                    (const __attribute__((address_space(1))) void*)(src[sp][i] + koff),
                    (__attribute__((address_space(3))) void*)&dst[sp][16 * (4 * i + wave)][0], 16, 0, 0);
    }
};

// s_waitcnt vmcnt(N) lgkmcnt(0) (gfx9 encoding: vmcnt bits 3:0 and 15:14,
// expcnt 6:4 left at its maximum, lgkmcnt 11:8).
template <int N>
__device__ __forceinline__ void wait_vm_lgkm0()
{
    static_assert(N >= 0 && N < 64, "vmcnt range");
    __builtin_amdgcn_s_waitcnt((N & 15) | (7 << 4) | ((N >> 4) << 14));
}

// Workgroup barrier that keeps the compiler from moving memory accesses
// across it and adds no waits of its own (the caller has waited).
__device__ __forceinline__ void slab_barrier()
{
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
}

// C tile (BM x BN) = A B^T over Kpad with direct-to-LDS slabs.  Slab t is read
// from buffer t mod NB; at its turn each thread waits until only the newer
// NB - 2 slabs' loads are outstanding (and its own LDS reads of the previous
// slab are done), the workgroup meets, and the loads of slab t + NB - 1 go
// into the buffer the previous slab occupied.  Requests past the last slab
// re-read it.  On return no load is in flight and the LDS is free.
template <int KIND, int BM, int BN, int NB>
__device__ __forceinline__ void gemm_dma_tile(const OperandV2<KIND>& A, const OperandV2<KIND>& B, int m0, int n0,
                                              int Kpad, AccLp<BM, BN>& acc, SmemDma<KIND, BM, BN, NB>& sm)
{
    static_assert(NB >= 2, "at least two buffers");
    typedef V2Traits<KIND> Tr;
    constexpr int FM = BM / 32, FN = BN / 32, BK = Tr::BK;
    constexpr int L = DmaSlab<KIND, BM>::NLOAD + DmaSlab<KIND, BN>::NLOAD;
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc.v[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    DmaSlab<KIND, BM> la;
    DmaSlab<KIND, BN> lb;
    la.init(A, m0);
    lb.init(B, n0);
    const int nslab = Kpad / BK;
    const int last = (nslab - 1) * BK;
    static_for<NB - 1>([&](auto sc) {
        constexpr int u = decltype(sc)::value;
        const int k = min(u * BK, last);
        la.issue(k, sm.a[u]);
        lb.issue(k, sm.b[u]);
    });
    auto slab = [&](int t, auto uc) {
        constexpr int u = decltype(uc)::value;  // t mod NB
        constexpr int w = (u + NB - 1) % NB;    // the previous slab's buffer
        wait_vm_lgkm0<(NB - 2) * L>();
        slab_barrier();
        const int k = min((t + NB - 1) * BK, last);
        la.issue(k, sm.a[w]);
        lb.issue(k, sm.b[w]);
        mfma_slab<KIND, BM, BN, 1>(sm, u, acc);
    };
    const int nfull = nslab - nslab % NB;
    int t0 = 0;
    for (; t0 < nfull; t0 += NB) {
        static_for<NB>([&](auto uc) { slab(t0 + decltype(uc)::value, uc); });
    }
    const int rem = nslab - nfull;
    static_for<NB - 1>([&](auto uc) {
        if (decltype(uc)::value < rem) slab(t0 + decltype(uc)::value, uc);
    });
    wait_vm_lgkm0<0>();
    slab_barrier();
}

}  // namespace ngw
