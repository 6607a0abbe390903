// FP64 MFMA GEMM core and the ISDF-factored Casida operator kernels (gfx950).
//
// The operator of NanoGW's Lanczos path (shared/apply_matvec_BLOCK.F90:25-208)
//   y = d o x' + beta s o C^T M C x',  x' = s o x (full) or x (TDA)
// with C(mu, (v,c)) = Psi_v(mu) Psi_c(mu) is applied WITHOUT forming C:
//   K1  T(mu, (v,b)) = sum_c PsiC(mu,c) x'_b(v,c)        GEMM Nmu x (nv B) x nc
//       t(mu, b)     = sum_v PsiV(mu,v) T(mu,(v,b))      (epilogue, LDS reduce)
//   K2  u(mu, b)     = sum_nu M(mu,nu) t(nu, b)          GEMM Nmu x B x Nmu
//   K3  Y(c, (v,b))  = sum_mu PsiC(mu,c) PsiV(mu,v) u(mu,b)  GEMM nc x (nv B) x Nmu
//       y = alpha (d o x' + beta s o Y) + gamma v0 [+ acc]   (epilogue)
// The epilogue also carries one Horner stage of p(H)
// (apply_matvec_BLOCK.F90:250-265) and the Lanczos "w = w + p(H) q" update
// (lanczos_spectra_BLOCK.F90:91-95).
//
// K1's reduction over v writes per-tile partial sums to a slot buffer instead
// of using atomics, so every result is bitwise independent of how many other
// vectors share the launch.

#include "ngw_kernels.h"

namespace ngw {

typedef double f64x4 __attribute__((ext_vector_type(4)));

constexpr int BM = 64, BN = 64, BK = 16, NT = 256, PAD = 1;

struct SmemAB {
    double a[BK][BM + PAD];
    double b[BK][BN + PAD];
};

// Loads one BM x BK (or BN x BK) tile through a loader functor into LDS in
// [k][mn] layout.  kContig loaders are read with consecutive threads along k.
template <class L, int MN>
__device__ __forceinline__ void load_tile(const L& ld, int mn0, int k0, int K, double (*dst)[MN + PAD])
{
    const int tid = threadIdx.x;
#pragma unroll
    for (int i = 0; i < (MN * BK) / NT; ++i) {
        int e = tid + i * NT;
        int mm, kk;
        if (L::kContig) {
            kk = e % BK;
            mm = e / BK;
        } else {
            mm = e % MN;
            kk = e / MN;
        }
        int gm = mn0 + mm, gk = k0 + kk;
        dst[kk][mm] = (gm < ld.MN && gk < K) ? ld(gm, gk) : 0.0;
    }
}

// acc[i][j] holds the 16x16 block (rows wm+16i.., cols wn+16j..) of this wave;
// lane l, component r <-> row l/16 + 4r, col l%16 (v_mfma_f64_16x16x4f64 output
// layout, measured on gfx950; A/B operands: lane l holds row/col l%16, k = l/16).
template <class AL, class BL>
__device__ __forceinline__ void mainloop(const AL& al, const BL& bl, int m0, int n0, int K,
                                         f64x4 (&acc)[2][2], SmemAB* sm)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
    const int r16 = lane & 15, k4 = lane >> 4;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f64x4{0.0, 0.0, 0.0, 0.0};
    for (int k0 = 0; k0 < K; k0 += BK) {
        load_tile<AL, BM>(al, m0, k0, K, sm->a);
        load_tile<BL, BN>(bl, n0, k0, K, sm->b);
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
            double a0 = sm->a[kk + k4][wm + r16];
            double a1 = sm->a[kk + k4][wm + 16 + r16];
            double b0 = sm->b[kk + k4][wn + r16];
            double b1 = sm->b[kk + k4][wn + 16 + r16];
            acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
        }
        __syncthreads();
    }
}

// Calls f(row, col, value) for the 16 accumulator values this thread holds.
template <class F>
__device__ __forceinline__ void for_acc(const f64x4 (&acc)[2][2], F&& f)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = (wave & 1) * 32, wn = (wave >> 1) * 32;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r)
                f(wm + 16 * i + (lane >> 4) + 4 * r, wn + 16 * j + (lane & 15), acc[i][j][r]);
}

// ---------------------------------------------------------------- plain GEMM
struct ColMajorA {  // A(m, k) = a[m + k*ld], contiguous along m
    static constexpr bool kContig = false;
    const double* a; long ld; int MN;
    __device__ double operator()(int m, int k) const { return a[m + (long)k * ld]; }
};
struct ColMajorB {  // B(k, n) = b[k + n*ld], contiguous along k
    static constexpr bool kContig = true;
    const double* b; long ld; int MN;
    __device__ double operator()(int n, int k) const { return b[k + (long)n * ld]; }
};

__global__ __launch_bounds__(NT) void gemm_f64_kernel(int m, int n, int k, const double* a, long lda,
                                                      const double* b, long ldb, double* c, long ldc,
                                                      int mtiles)
{
    __shared__ SmemAB sm;
    const int m0 = (blockIdx.x % mtiles) * BM, n0 = (blockIdx.x / mtiles) * BN;
    f64x4 acc[2][2];
    mainloop(ColMajorA{a, lda, m}, ColMajorB{b, ldb, n}, m0, n0, k, acc, &sm);
    for_acc(acc, [&](int row, int col, double val) {
        int gm = m0 + row, gn = n0 + col;
        if (gm < m && gn < n) c[gm + (long)gn * ldc] = val;
    });
}

int gemm_f64(hipStream_t s, int m, int n, int k, const double* a, long lda, const double* b,
             long ldb, double* c, long ldc)
{
    if (m <= 0 || n <= 0) return NGW_OK;
    int mt = (m + BM - 1) / BM, nt = (n + BN - 1) / BN;
    hipLaunchKernelGGL(gemm_f64_kernel, dim3(mt * nt), dim3(NT), 0, s, m, n, k, a, lda, b, ldb, c,
                       ldc, mt);
    NGW_HIP(hipGetLastError());
    return NGW_OK;
}

// C(m, n) = alpha A(k, m)^T B(k, n): A read along k (zeta^T V zeta of the V build)
struct ColMajorAT {
    static constexpr bool kContig = true;
    const double* a; long ld; int MN;
    __device__ double operator()(int m, int k) const { return a[k + (long)m * ld]; }
};

__global__ __launch_bounds__(NT) void gemm_f64_tn_kernel(int m, int n, int k, double alpha, const double* a,
                                                         long lda, const double* b, long ldb, double* c,
                                                         long ldc, int mtiles)
{
    __shared__ SmemAB sm;
    const int m0 = (blockIdx.x % mtiles) * BM, n0 = (blockIdx.x / mtiles) * BN;
    f64x4 acc[2][2];
    mainloop(ColMajorAT{a, lda, m}, ColMajorB{b, ldb, n}, m0, n0, k, acc, &sm);
    for_acc(acc, [&](int row, int col, double val) {
        int gm = m0 + row, gn = n0 + col;
        if (gm < m && gn < n) c[gm + (long)gn * ldc] = alpha * val;
    });
}

int gemm_f64_tn(hipStream_t s, int m, int n, int k, double alpha, const double* a, long lda,
                const double* b, long ldb, double* c, long ldc)
{
    if (m <= 0 || n <= 0) return NGW_OK;
    int mt = (m + BM - 1) / BM, nt = (n + BN - 1) / BN;
    hipLaunchKernelGGL(gemm_f64_tn_kernel, dim3(mt * nt), dim3(NT), 0, s, m, n, k, alpha, a, lda, b, ldb,
                       c, ldc, mt);
    NGW_HIP(hipGetLastError());
    return NGW_OK;
}

// ---------------------------------------------------------------- K1
struct K1A {  // PsiC(mu, c)
    static constexpr bool kContig = false;
    const double* psic; long ld; int MN;
    __device__ double operator()(int mu, int c) const { return psic[mu + (long)c * ld]; }
};
struct K1B {  // x'(c, j = v + nv*b) = s(v,c) x_b(v,c)
    static constexpr bool kContig = false;
    const double* x; const double* s; long ldx; int nv; int MN;
    __device__ double operator()(int j, int c) const
    {
        int b = j / nv, v = j - b * nv;
        long idx = v + (long)c * nv;
        double xv = x[idx + (long)b * ldx];
        return s ? s[idx] * xv : xv;
    }
};

__global__ __launch_bounds__(NT) void k1_f64_kernel(K1A al, K1B bl, const double* psiv, long ldp,
                                                    int nmu, int nc, int nv, int nvec, double* tpart,
                                                    long ldt, int mtiles)
{
    __shared__ union {
        SmemAB ab;
        double red[BM][BN + PAD];
    } sm;
    const int tm = blockIdx.x % mtiles, tn = blockIdx.x / mtiles;
    const int m0 = tm * BM, n0 = tn * BN;
    const int N = nv * nvec;
    f64x4 acc[2][2];
    mainloop(al, bl, m0, n0, nc, acc, &sm.ab);
    // acc -> LDS, multiplied by PsiV(mu, v(j))
    for_acc(acc, [&](int row, int col, double val) {
        int gm = m0 + row, gn = n0 + col;
        double p = 0.0;
        if (gm < nmu && gn < N) p = val * psiv[gm + (long)(gn % nv) * ldp];
        sm.red[row][col] = p;
    });
    __syncthreads();
    const int nend = min(n0 + BN, N);
    const int b_lo = n0 / nv, b_hi = (nend - 1) / nv;
    const int nbt = b_hi - b_lo + 1;
    for (int e = threadIdx.x; e < BM * nbt; e += NT) {
        int row = e % BM, b = b_lo + e / BM;
        int c0 = max(n0, b * nv) - n0, c1 = min(nend, (b + 1) * nv) - n0;
        double s = 0.0;
        for (int cc = c0; cc < c1; ++cc) s += sm.red[row][cc];
        int gm = m0 + row;
        int slot = tn - (b * nv) / BN;
        if (gm < nmu) tpart[((long)slot * nvec + b) * ldt + gm] = s;
    }
}

// ---------------------------------------------------------------- K2
struct K2A {  // M(mu, nu)
    static constexpr bool kContig = false;
    const double* m; long ld; int MN;
    __device__ double operator()(int mu, int nu) const { return m[mu + (long)nu * ld]; }
};
struct K2BPart {  // t(nu, b) = sum over the K1 slots that cover vector b
    static constexpr bool kContig = true;
    const double* tpart; long ldt; int nvec; int nv; int MN;
    __device__ double operator()(int b, int nu) const
    {
        int first = (b * nv) / BN, last = ((b + 1) * nv - 1) / BN;
        double s = 0.0;
        for (int sl = 0; sl <= last - first; ++sl) s += tpart[((long)sl * nvec + b) * ldt + nu];
        return s;
    }
};
struct K2BPair {  // c_mn(nu) = Psi(nu, m_b) Psi(nu, n_b)  (P-vector build)
    static constexpr bool kContig = true;
    const double* psi; long ld; const int* mband; const int* nband; int MN;
    __device__ double operator()(int b, int nu) const
    {
        return psi[nu + (long)mband[b] * ld] * psi[nu + (long)nband[b] * ld];
    }
};

template <class BL>
__global__ __launch_bounds__(NT) void k2_f64_kernel(K2A al, BL bl, int nmu, int nvec, double* u,
                                                    long ldu, int mtiles)
{
    __shared__ SmemAB sm;
    const int m0 = (blockIdx.x % mtiles) * BM, n0 = (blockIdx.x / mtiles) * BN;
    f64x4 acc[2][2];
    mainloop(al, bl, m0, n0, nmu, acc, &sm);
    for_acc(acc, [&](int row, int col, double val) {
        int gm = m0 + row, gn = n0 + col;
        if (gm < nmu && gn < nvec) u[gm + (long)gn * ldu] = val;
    });
}

// ---------------------------------------------------------------- K3
struct K3A {  // PsiC(mu, c) read as A(c, mu), contiguous along mu
    static constexpr bool kContig = true;
    const double* psic; long ld; int MN;
    __device__ double operator()(int c, int mu) const { return psic[mu + (long)c * ld]; }
};
struct K3B {  // W(mu, j = v + nv*b) = PsiV(mu, v) u(mu, b)
    static constexpr bool kContig = true;
    const double* psiv; long ldp; const double* u; long ldu; int nv; int MN;
    __device__ double operator()(int j, int mu) const
    {
        int b = j / nv, v = j - b * nv;
        return psiv[mu + (long)v * ldp] * u[mu + (long)b * ldu];
    }
};

__global__ __launch_bounds__(NT) void k3_f64_kernel(K3A al, K3B bl, int nmu, int nc, int nv, int nvec,
                                                    OperatorEpilogue ep, int mtiles)
{
    __shared__ SmemAB sm;
    const int m0 = (blockIdx.x % mtiles) * BM, n0 = (blockIdx.x / mtiles) * BN;
    const int N = nv * nvec;
    f64x4 acc[2][2];
    mainloop(al, bl, m0, n0, nmu, acc, &sm);
    for_acc(acc, [&](int row, int col, double val) {
        int c = m0 + row, j = n0 + col;
        if (c < nc && j < N) {
            int b = j / nv, v = j - b * nv;
            ep.apply(v + (long)c * nv, b, val);
        }
    });
}

// ---------------------------------------------------------------- launchers
int k1_slots(int nv) { return (nv - 1) / BN + 2; }

int operator_f64(ngw_context* ctx, const OperatorArgs& a)
{
    const int nmu = ctx->n_intp, nv = ctx->nval, nc = ctx->ncond, B = a.nvec;
    if (B <= 0) return NGW_OK;
    const double* psi = ctx->psi.as<double>();
    const long ldp = ctx->ld_psi;
    const double* psiv = psi;
    const double* psic = psi + (long)nv * ldp;
    const int S = k1_slots(nv);
    NGW_TRY(ctx->tpart.ensure(sizeof(double) * (size_t)S * B * nmu));
    NGW_TRY(ctx->ubuf.ensure(sizeof(double) * (size_t)B * nmu));
    double* tpart = ctx->tpart.as<double>();
    double* u = ctx->ubuf.as<double>();
    hipStream_t s = ctx->stream;
    const double* sx = a.tda ? nullptr : ctx->sqrtr.as<double>();
    {
        ProfScope ps(ctx, KC_K1);
        int mt = (nmu + BM - 1) / BM, nt = (nv * B + BN - 1) / BN;
        hipLaunchKernelGGL(k1_f64_kernel, dim3(mt * nt), dim3(NT), 0, s,
                           K1A{psic, ldp, nmu}, K1B{a.x, sx, a.ldx, nv, nv * B}, psiv, ldp, nmu, nc,
                           nv, B, tpart, (long)nmu, mt);
        NGW_HIP(hipGetLastError());
    }
    {
        ProfScope ps(ctx, KC_K2);
        int mt = (nmu + BM - 1) / BM, nt = (B + BN - 1) / BN;
        hipLaunchKernelGGL(k2_f64_kernel<K2BPart>, dim3(mt * nt), dim3(NT), 0, s,
                           K2A{ctx->mmat.as<double>(), ctx->ld_m, nmu},
                           K2BPart{tpart, (long)nmu, B, nv, B}, nmu, B, u, (long)nmu, mt);
        NGW_HIP(hipGetLastError());
    }
    {
        ProfScope ps(ctx, KC_K3);
        OperatorEpilogue ep = a.epilogue(ctx);
        int mt = (nc + BM - 1) / BM, nt = (nv * B + BN - 1) / BN;
        hipLaunchKernelGGL(k3_f64_kernel, dim3(mt * nt), dim3(NT), 0, s, K3A{psic, ldp, nc},
                           K3B{psiv, ldp, u, (long)nmu, nv, nv * B}, nmu, nc, nv, B, ep, mt);
        NGW_HIP(hipGetLastError());
    }
    return NGW_OK;
}

// P vectors: v0(:, b) = s o C^T (M c_{m_b n_b})   (s = 1 for TDA),
// SIGMA/calculate_sigma_lanczos_BLOCK.F90z:286-358.
int pvectors_f64(ngw_context* ctx, const int* d_mband, const int* d_nband, int nvec, int tda,
                 double* d_p, long ldp_out)
{
    const int nmu = ctx->n_intp, nv = ctx->nval, nc = ctx->ncond, B = nvec;
    const double* psi = ctx->psi.as<double>();
    const long ldp = ctx->ld_psi;
    NGW_TRY(ctx->ubuf.ensure(sizeof(double) * (size_t)B * nmu));
    double* u = ctx->ubuf.as<double>();
    hipStream_t s = ctx->stream;
    ProfScope ps(ctx, KC_PVEC);
    {
        int mt = (nmu + BM - 1) / BM, nt = (B + BN - 1) / BN;
        hipLaunchKernelGGL(k2_f64_kernel<K2BPair>, dim3(mt * nt), dim3(NT), 0, s,
                           K2A{ctx->mmat.as<double>(), ctx->ld_m, nmu},
                           K2BPair{psi, ldp, d_mband, d_nband, B}, nmu, B, u, (long)nmu, mt);
        NGW_HIP(hipGetLastError());
    }
    {
        OperatorEpilogue ep{};
        ep.out = d_p;
        ep.ldo = ldp_out;
        ep.s = tda ? nullptr : ctx->sqrtr.as<double>();
        ep.beta = 1.0;
        ep.alpha = 1.0;
        int mt = (nc + BM - 1) / BM, nt = (nv * B + BN - 1) / BN;
        hipLaunchKernelGGL(k3_f64_kernel, dim3(mt * nt), dim3(NT), 0, s,
                           K3A{psi + (long)nv * ldp, ldp, nc}, K3B{psi, ldp, u, (long)nmu, nv, nv * B},
                           nmu, nc, nv, B, ep, mt);
        NGW_HIP(hipGetLastError());
    }
    return NGW_OK;
}

}  // namespace ngw
