// FP64 MFMA GEMM core for gfx950 (v_mfma_f64_16x16x4f64).
//
// C tile (BM x BN) of C = A B, K reduced in slabs of BK = 16:
//  * 256 threads = 4 waves in a 2 x 2 layout, each wave owns (BM/2) x (BN/2);
//  * LDS holds A and B slabs k-contiguous ([mn][BK + 2]) so that each lane
//    reads its 4 consecutive k of a fragment with two 16-byte loads; the k
//    order inside a slab is permuted (lane group g of MFMA sub-step s uses
//    k = 4 g + s), which leaves the sum unchanged;
//  * double-buffered LDS with the next slab prefetched into registers while
//    the current one is multiplied: one barrier per slab;
//  * operands come through loader functors: a thread keeps a per-row handle
//    (pointer arithmetic done once per tile) and reads element k on demand.
//
// MFMA f64 16x16x4 lane layout (measured on gfx950):
//   A/B operand: lane l supplies row/col l%16 at k = l/16;
//   result     : lane l, component r holds row l/16 + 4r, col l%16.
#pragma once

#include "ngw_kernels.h"

namespace ngw {

typedef double f64x4 __attribute__((ext_vector_type(4)));

constexpr int GK = 16;     // K slab
constexpr int GPAD = 2;    // LDS row padding (doubles); keeps rows 16-byte aligned
constexpr int GNT = 256;   // threads per workgroup

// Workgroup -> tile with each XCD owning a contiguous range of tiles (WGs are
// dispatched round-robin over the 8 XCDs); inside a range the M index runs
// fastest so neighbouring tiles share their B slab in that XCD's L2.
__device__ __forceinline__ int xcd_tile(int bid, int ntiles)
{
    const int x = bid & 7, local = bid >> 3;
    const int q = ntiles >> 3, r = ntiles & 7;
    const int start = x * q + (x < r ? x : r);
    return start + local;
}

template <int BM, int BN>
struct SmemF64 {
    double a[2][BM][GK + GPAD];
    double b[2][BN][GK + GPAD];
};

// Per-thread slab loader for one operand tile of MN rows.
// kContig operands: thread t reads k = t % 16 for rows t/16 + 16 i.
// otherwise       : thread t reads row t % MN for k = t/MN + (256/MN) i.
// Loader functors split an element into NRAW raw values fetched by
// unconditional loads from clamped (always valid) addresses, and a combine()
// applied when the slab is written to LDS, so that the prefetch of the next
// slab never has to complete before the current slab's MFMAs.
template <class L, int MN>
struct SlabLoader {
    static constexpr int PER = MN * GK / GNT;  // elements per thread
    static constexpr int NH = L::kContig ? PER : 1;
    typename L::Handle h[NH];
    bool ok[NH];
    int kk0;

    __device__ __forceinline__ void init(const L& ld, int mn0)
    {
        const int t = threadIdx.x;
        if (L::kContig) {
            kk0 = t % GK;
#pragma unroll
            for (int i = 0; i < PER; ++i) {
                int mn = mn0 + t / GK + (GNT / GK) * i;
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

    __device__ __forceinline__ void fetch(const L& ld, int k0, int K, double (&raw)[PER][L::NRAW]) const
    {
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            int k = min(kof(k0, i), K - 1);
            ld.fetch(h[L::kContig ? i : 0], k, raw[i]);
        }
    }

    __device__ __forceinline__ void store(const L& ld, int k0, int K, double (*dst)[GK + GPAD],
                                          const double (&raw)[PER][L::NRAW]) const
    {
        const int t = threadIdx.x;
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const bool valid = ok[L::kContig ? i : 0] && kof(k0, i) < K;
            const double v = valid ? ld.combine(raw[i]) : 0.0;
            if (L::kContig)
                dst[t / GK + (GNT / GK) * i][kk0] = v;
            else
                dst[t % MN][kk0 + (GNT / MN) * i] = v;
        }
    }
};

template <int BM, int BN>
struct AccF64 {
    static constexpr int FM = BM / 32, FN = BN / 32;
    f64x4 v[FM][FN];
};

template <int BM, int BN, class AL, class BL>
__device__ __forceinline__ void gemm_f64_tile(const AL& al, const BL& bl, int m0, int n0, int K,
                                              AccF64<BM, BN>& acc, SmemF64<BM, BN>& sm)
{
    constexpr int FM = BM / 32, FN = BN / 32;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = (wave & 1) * (BM / 2), wn = (wave >> 1) * (BN / 2);
    const int r16 = lane & 15, k4 = lane >> 4;
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc.v[i][j] = f64x4{0.0, 0.0, 0.0, 0.0};
    SlabLoader<AL, BM> la;
    SlabLoader<BL, BN> lb;
    la.init(al, m0);
    lb.init(bl, n0);
    double ra[SlabLoader<AL, BM>::PER][AL::NRAW], rb[SlabLoader<BL, BN>::PER][BL::NRAW];
    const int nslab = (K + GK - 1) / GK;
    la.fetch(al, 0, K, ra);
    lb.fetch(bl, 0, K, rb);
    la.store(al, 0, K, sm.a[0], ra);
    lb.store(bl, 0, K, sm.b[0], rb);
    __syncthreads();
    for (int t = 0; t < nslab; ++t) {
        const int cur = t & 1;
        if (t + 1 < nslab) {
            la.fetch(al, (t + 1) * GK, K, ra);
            lb.fetch(bl, (t + 1) * GK, K, rb);
        }
        double af[FM][4], bf[FN][4];
#pragma unroll
        for (int i = 0; i < FM; ++i) {
            const double2* p = reinterpret_cast<const double2*>(&sm.a[cur][wm + 16 * i + r16][4 * k4]);
            double2 x0 = p[0], x1 = p[1];
            af[i][0] = x0.x; af[i][1] = x0.y; af[i][2] = x1.x; af[i][3] = x1.y;
        }
#pragma unroll
        for (int j = 0; j < FN; ++j) {
            const double2* p = reinterpret_cast<const double2*>(&sm.b[cur][wn + 16 * j + r16][4 * k4]);
            double2 x0 = p[0], x1 = p[1];
            bf[j][0] = x0.x; bf[j][1] = x0.y; bf[j][2] = x1.x; bf[j][3] = x1.y;
        }
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
            for (int i = 0; i < FM; ++i)
#pragma unroll
                for (int j = 0; j < FN; ++j)
                    acc.v[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i][s], bf[j][s], acc.v[i][j], 0, 0, 0);
        if (t + 1 < nslab) {
            la.store(al, (t + 1) * GK, K, sm.a[cur ^ 1], ra);
            lb.store(bl, (t + 1) * GK, K, sm.b[cur ^ 1], rb);
        }
        __syncthreads();
    }
}

// Calls f(row, col, value) for every accumulator value of this thread
// (row, col relative to the tile origin).
template <int BM, int BN, class F>
__device__ __forceinline__ void for_acc_f64(const AccF64<BM, BN>& acc, F&& f)
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
                f(wm + 16 * i + (lane >> 4) + 4 * r, wn + 16 * j + (lane & 15), acc.v[i][j][r]);
}

}  // namespace ngw
