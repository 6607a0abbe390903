// FP64 ISDF-factored Casida operator kernels (gfx950 MFMA).
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
// of using atomics; a small kernel sums the slots in a fixed order, so every
// result is bitwise independent of how many other vectors share the launch.

#include "operator_loaders.h"

namespace ngw {

// ---------------------------------------------------------------- plain GEMMs
template <class AL, class BL>
__global__ __launch_bounds__(GNT) void gemm_plain_kernel(AL al, BL bl, int m, int n, int k, double alpha,
                                                         double* c, long ldc, int mtiles, int ntiles)
{
    __shared__ SmemF64<128, 128> sm;
    const int tile = xcd_tile(blockIdx.x, ntiles);
    const int m0 = (tile % mtiles) * 128, n0 = (tile / mtiles) * 128;
    AccF64<128, 128> acc;
    gemm_f64_tile<128, 128>(al, bl, m0, n0, k, acc, sm);
    for_acc_f64<128, 128>(acc, [&](int row, int col, double val) {
        int gm = m0 + row, gn = n0 + col;
        if (gm < m && gn < n) c[gm + (long)gn * ldc] = alpha * val;
    });
}

int gemm_f64(hipStream_t s, int m, int n, int k, const double* a, long lda, const double* b,
             long ldb, double* c, long ldc)
{
    if (m <= 0 || n <= 0) return NGW_OK;
    int mt = (m + 127) / 128, nt = (n + 127) / 128;
    hipLaunchKernelGGL((gemm_plain_kernel<ColMajorA, ColMajorB>), dim3(mt * nt), dim3(GNT), 0, s,
                       ColMajorA{a, lda, m}, ColMajorB{b, ldb, n}, m, n, k, 1.0, c, ldc, mt, mt * nt);
    NGW_HIP(hipGetLastError());
    return NGW_OK;
}

int gemm_f64_tn(hipStream_t s, int m, int n, int k, double alpha, const double* a, long lda,
                const double* b, long ldb, double* c, long ldc)
{
    if (m <= 0 || n <= 0) return NGW_OK;
    int mt = (m + 127) / 128, nt = (n + 127) / 128;
    hipLaunchKernelGGL((gemm_plain_kernel<ColMajorAT, ColMajorB>), dim3(mt * nt), dim3(GNT), 0, s,
                       ColMajorAT{a, lda, m}, ColMajorB{b, ldb, n}, m, n, k, alpha, c, ldc, mt, mt * nt);
    NGW_HIP(hipGetLastError());
    return NGW_OK;
}

// C = A B (column-major A m x k, B k x n) with the k sum split into ranges
// when the tiles alone are too few for the GPU (n of a few vectors): grid
// tiles x S, range partials added in range order by split_sum_kernel.
constexpr int GSM = 128, GSN = 64;
__global__ __launch_bounds__(GNT) void gemm_split_kernel(const double* a, long lda, const double* b, long ldb, int m,
                                                         int n, int k, int kc, int mtiles, int ntiles, double* part)
{
    __shared__ SmemF64<GSM, GSN> sm;
    const int split = blockIdx.y;
    const int tile = xcd_tile(blockIdx.x, mtiles * ntiles);
    const int m0 = (tile % mtiles) * GSM, n0 = (tile / mtiles) * GSN;
    const int k0 = split * kc, K = min(kc, k - k0);
    AccF64<GSM, GSN> acc;
    gemm_f64_tile<GSM, GSN>(ColMajorA{a + (long)k0 * lda, lda, m}, ColMajorB{b + k0, ldb, n}, m0, n0, K, acc, sm);
    const long mn = (long)m * n;
    for_acc_f64<GSM, GSN>(acc, [&](int row, int col, double val) {
        const int gm = m0 + row, gn = n0 + col;
        if (gm < m && gn < n) part[(long)split * mn + gm + (long)gn * m] = val;
    });
}

__global__ void split_sum_kernel(const double* part, int S, int m, int n, double* c, long ldc)
{
    const long mn = (long)m * n;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < mn; t += (long)gridDim.x * blockDim.x) {
        double s = 0.0;
        for (int p = 0; p < S; ++p) s += part[(long)p * mn + t];
        const long gn = t / m, gm = t - gn * m;
        c[gm + gn * ldc] = s;
    }
}

// Splits of a sum of length k over `tiles` tiles: about 1024 workgroups in
// all, each range at least 64 long (four 16-deep slabs).
static int sum_splits(int tiles, int k)
{
    const int want = std::max(1, 1024 / std::max(tiles, 1));
    return std::max(1, std::min(want, (k + 63) / 64));
}

int gemm_f64_split(hipStream_t s, int m, int n, int k, const double* a, long lda, const double* b, long ldb,
                   double* c, long ldc, DevBuf& part)
{
    if (m <= 0 || n <= 0) return NGW_OK;
    const int mt = (m + GSM - 1) / GSM, nt = (n + GSN - 1) / GSN;
    int S = sum_splits(mt * nt, k);
    const int kc = (k + S - 1) / S;
    S = (k + kc - 1) / kc;
    NGW_TRY(part.ensure(sizeof(double) * (size_t)S * m * n));
    hipLaunchKernelGGL(gemm_split_kernel, dim3(mt * nt, S), dim3(GNT), 0, s, a, lda, b, ldb, m, n, k, kc, mt, nt,
                       part.as<double>());
    const long mn = (long)m * n;
    hipLaunchKernelGGL(split_sum_kernel, dim3((unsigned)std::min<long>((mn + 255) / 256, 4096)), dim3(256), 0, s,
                       part.as<const double>(), S, m, n, c, ldc);
    NGW_HIP(hipGetLastError());
    return NGW_OK;
}

// ---------------------------------------------------------------- K1
constexpr int K1M = 128, K1N = 64;

__global__ __launch_bounds__(GNT) void k1_f64_kernel(ColMajorA al, XprimeB bl, const double* psiv, long ldp,
                                                     int nmu, int nc, int nv, int nvec, double* tpart,
                                                     long ldt, int mtiles, int ntiles)
{
    __shared__ union {
        SmemF64<K1M, K1N> ab;
        double red[K1M][K1N + 2];
    } sm;
    const int tile = xcd_tile(blockIdx.x, ntiles);
    const int tm = tile % mtiles, tn = tile / mtiles;
    const int m0 = tm * K1M, n0 = tn * K1N;
    const int N = nv * nvec;
    AccF64<K1M, K1N> acc;
    gemm_f64_tile<K1M, K1N>(al, bl, m0, n0, nc, acc, sm.ab);
    // T(mu, j) PsiV(mu, v(j)) into LDS, then per-(mu, b) sums over the tile's v range
    for_acc_f64<K1M, K1N>(acc, [&](int row, int col, double val) {
        int gm = m0 + row, gn = n0 + col;
        double p = 0.0;
        if (gm < nmu && gn < N) p = val * psiv[gm + (long)(gn % nv) * ldp];
        sm.red[row][col] = p;
    });
    __syncthreads();
    const int nend = min(n0 + K1N, N);
    const int b_lo = n0 / nv, b_hi = (nend - 1) / nv;
    const int nbt = b_hi - b_lo + 1;
    for (int e = threadIdx.x; e < K1M * nbt; e += GNT) {
        int row = e % K1M, b = b_lo + e / K1M;
        int c0 = max(n0, b * nv) - n0, c1 = min(nend, (b + 1) * nv) - n0;
        double s = 0.0;
        for (int cc = c0; cc < c1; ++cc) s += sm.red[row][cc];
        int gm = m0 + row;
        int slot = tn - (b * nv) / K1N;
        if (gm < nmu) tpart[((long)slot * nvec + b) * ldt + gm] = s;
    }
}

int k1_slots(int nv) { return (nv - 1) / K1N + 2; }

// t(nu, b) = sum over the K1 tiles that covered vector b, in tile order.
__global__ void k1_reduce_kernel(const double* tpart, long ldt, int nmu, int nvec, int nv, double* t)
{
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= (long)nmu * nvec) return;
    const int b = (int)(i / nmu), nu = (int)(i - (long)b * nmu);
    const int first = (b * nv) / K1N, last = ((b + 1) * nv - 1) / K1N;
    double s = 0.0;
    for (int sl = 0; sl <= last - first; ++sl) s += tpart[((long)sl * nvec + b) * ldt + nu];
    t[nu + (long)b * nmu] = s;
}

int k1_reduce(hipStream_t s, const double* tpart, long ldt, int nmu, int nvec, int nv, double* t)
{
    long total = (long)nmu * nvec;
    hipLaunchKernelGGL(k1_reduce_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, tpart, ldt, nmu,
                       nvec, nv, t);
    NGW_HIP(hipGetLastError());
    return NGW_OK;
}

// ---------------------------------------------------------------- K2
template <class BL>
__global__ __launch_bounds__(GNT) void k2_f64_kernel(ColMajorA al, BL bl, int nmu, int nvec, double* u,
                                                     long ldu, int mtiles, int ntiles)
{
    __shared__ SmemF64<64, 64> sm;
    const int tile = xcd_tile(blockIdx.x, ntiles);
    const int m0 = (tile % mtiles) * 64, n0 = (tile / mtiles) * 64;
    AccF64<64, 64> acc;
    gemm_f64_tile<64, 64>(al, bl, m0, n0, nmu, acc, sm);
    for_acc_f64<64, 64>(acc, [&](int row, int col, double val) {
        int gm = m0 + row, gn = n0 + col;
        if (gm < nmu && gn < nvec) u[gm + (long)gn * ldu] = val;
    });
}

// ---------------------------------------------------------------- K3
constexpr int K3M = 64, K3N = 128;

__global__ __launch_bounds__(GNT) void k3_f64_kernel(ColMajorAT al, WeightedPsiVB bl, int nmu, int nc, int nv,
                                                     int nvec, OperatorEpilogue ep, int mtiles, int ntiles)
{
    __shared__ SmemF64<K3M, K3N> sm;
    const int tile = xcd_tile(blockIdx.x, ntiles);
    const int m0 = (tile % mtiles) * K3M, n0 = (tile / mtiles) * K3N;
    const int N = nv * nvec;
    AccF64<K3M, K3N> acc;
    gemm_f64_tile<K3M, K3N>(al, bl, m0, n0, nmu, acc, sm);
    for_acc_f64<K3M, K3N>(acc, [&](int row, int col, double val) {
        int c = m0 + row, j = n0 + col;
        if (c < nc && j < N) {
            int b = j / nv, v = j - b * nv;
            ep.apply(v + (long)c * nv, b, val);
        }
    });
}

// ---------------------------------------------------------------- one vector
// The operator on one vector (NanoGW's single-vector routines,
// apply_matvec_UltraLowMem.F90:25-222): K1 and K3 have only N = nv columns,
// far too few tiles to fill the GPU, so the sum of each tile is split into
// S ranges (grid = tiles x S); a second pass adds the S range partials in
// range order -- no atomics, so the result does not depend on scheduling.
constexpr int OV1M = 128, OV1N = 64;  // K1: mu x v tile, sum over c
constexpr int OV3M = 32, OV3N = 128;  // K3: c x v tile, sum over mu

// part[(split ntn + tn) nmu + mu] = sum_{v in tile} PsiV(mu, v) sum_{c in range} PsiC(mu, c) xs(v, c)
__global__ __launch_bounds__(GNT) void ov_k1_kernel(const double* psi, long ldp, const double* xs, int nmu, int nc,
                                                    int nv, int kc, int mtiles, int ntn, double* part)
{
    __shared__ union {
        SmemF64<OV1M, OV1N> ab;
        double red[OV1M][OV1N + 2];
    } sm;
    const int split = blockIdx.y;
    const int tile = xcd_tile(blockIdx.x, mtiles * ntn);
    const int tm = tile % mtiles, tn = tile / mtiles;
    const int m0 = tm * OV1M, n0 = tn * OV1N;
    const int c0 = split * kc, K = min(kc, nc - c0);
    AccF64<OV1M, OV1N> acc;
    gemm_f64_tile<OV1M, OV1N>(ColMajorA{psi + (long)(nv + c0) * ldp, ldp, nmu},
                              XprimeB{xs + (long)c0 * nv, nullptr, 0, nv, nv}, m0, n0, K, acc, sm.ab);
    for_acc_f64<OV1M, OV1N>(acc, [&](int row, int col, double val) {
        const int gm = m0 + row, v = n0 + col;
        sm.red[row][col] = (gm < nmu && v < nv) ? val * psi[gm + (long)v * ldp] : 0.0;
    });
    __syncthreads();
    if (threadIdx.x < OV1M) {
        const int row = threadIdx.x, gm = m0 + row;
        double s = 0.0;
        for (int cc = 0; cc < OV1N; ++cc) s += sm.red[row][cc];
        if (gm < nmu) part[((long)split * ntn + tn) * nmu + gm] = s;
    }
}

// t(mu) = sum over the nparts partials, in order
__global__ void ov_sum_kernel(const double* part, int nparts, long n, double* t)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        double s = 0.0;
        for (int p = 0; p < nparts; ++p) s += part[(long)p * n + i];
        t[i] = s;
    }
}

// part[split ncv + v + nv c] = sum_{mu in range} PsiC(mu, c) PsiV(mu, v) u(mu)
__global__ __launch_bounds__(GNT) void ov_k3_kernel(const double* psi, long ldp, const double* u, int nmu, int nc,
                                                    int nv, int kmu, int mtiles, int ntn, double* part)
{
    __shared__ SmemF64<OV3M, OV3N> sm;
    const int split = blockIdx.y;
    const int tile = xcd_tile(blockIdx.x, mtiles * ntn);
    const int m0 = (tile % mtiles) * OV3M, n0 = (tile / mtiles) * OV3N;
    const int mu0 = split * kmu, K = min(kmu, nmu - mu0);
    AccF64<OV3M, OV3N> acc;
    gemm_f64_tile<OV3M, OV3N>(ColMajorAT{psi + (long)nv * ldp + mu0, ldp, nc},
                              WeightedPsiVB{psi + mu0, ldp, u + mu0, 0, nv, nv}, m0, n0, K, acc, sm);
    const long ncv = (long)nv * nc;
    for_acc_f64<OV3M, OV3N>(acc, [&](int row, int col, double val) {
        const int c = m0 + row, v = n0 + col;
        if (c < nc && v < nv) part[(long)split * ncv + v + (long)c * nv] = val;
    });
}

// y(idx) = sum of the S partials (range order), then the operator epilogue
__global__ void ov_epilogue_kernel(OperatorEpilogue ep, const double* part, int S, long ncv)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < ncv; i += (long)gridDim.x * blockDim.x) {
        double y = 0.0;
        for (int c = 0; c < S; ++c) y += part[(long)c * ncv + i];
        ep.apply(i, 0, y);
    }
}

__global__ void ov_scale_kernel(const double* x, const double* s, long ncv, double* xs)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < ncv; i += (long)gridDim.x * blockDim.x)
        xs[i] = s ? s[i] * x[i] : x[i];
}

int operator_f64_one(ngw_context* ctx, const OperatorArgs& a)
{
    hipStream_t s = ctx->stream;
    const int nmu = ctx->n_intp, nv = ctx->nval, nc = ctx->ncond;
    const long ncv = ctx->ncv, ldp = ctx->ld_psi;
    const double* psi = ctx->psi.as<double>();
    const int mt1 = (nmu + OV1M - 1) / OV1M, nt1 = (nv + OV1N - 1) / OV1N;
    const int mt3 = (nc + OV3M - 1) / OV3M, nt3 = (nv + OV3N - 1) / OV3N;
    int S1 = sum_splits(mt1 * nt1, nc), S3 = sum_splits(mt3 * nt3, nmu);
    const int kc = (nc + S1 - 1) / S1, kmu = (nmu + S3 - 1) / S3;
    S1 = (nc + kc - 1) / kc;
    S3 = (nmu + kmu - 1) / kmu;
    NGW_TRY(ctx->ex_x.ensure(sizeof(double) * (size_t)ncv));
    NGW_TRY(ctx->ex_t.ensure(sizeof(double) * ((size_t)S1 * nt1 * nmu + 2 * (size_t)nmu)));
    NGW_TRY(ctx->ex_s.ensure(sizeof(double) * (size_t)S3 * ncv));
    double* xs = ctx->ex_x.as<double>();
    double* p1 = ctx->ex_t.as<double>();
    double* t = p1 + (size_t)S1 * nt1 * nmu;
    double* u = t + nmu;
    double* p3 = ctx->ex_s.as<double>();
    const int gs = (int)std::min<long>((ncv + 255) / 256, 2048);
    hipLaunchKernelGGL(ov_scale_kernel, dim3(gs), dim3(256), 0, s, a.x, a.tda ? nullptr : ctx->sqrtr.as<double>(),
                       ncv, xs);
    {
        ProfScope ps(ctx, KC_K1);
        hipLaunchKernelGGL(ov_k1_kernel, dim3(mt1 * nt1, S1), dim3(GNT), 0, s, psi, ldp, xs, nmu, nc, nv, kc, mt1,
                           nt1, p1);
        hipLaunchKernelGGL(ov_sum_kernel, dim3((nmu + 255) / 256), dim3(256), 0, s, p1, S1 * nt1, (long)nmu, t);
    }
    {
        ProfScope ps(ctx, KC_K2);
        NGW_TRY(gemv_wide(s, ctx->mmat.as<double>(), ctx->ld_m, nmu, nmu, t, u, ctx->gv_part));
    }
    {
        ProfScope ps(ctx, KC_K3);
        hipLaunchKernelGGL(ov_k3_kernel, dim3(mt3 * nt3, S3), dim3(GNT), 0, s, psi, ldp, u, nmu, nc, nv, kmu, mt3,
                           nt3, p3);
        hipLaunchKernelGGL(ov_epilogue_kernel, dim3(gs), dim3(256), 0, s, a.epilogue(ctx), p3, S3, ncv);
    }
    NGW_HIP(hipGetLastError());
    return NGW_OK;
}

// ---------------------------------------------------------------- launchers
static int launch_k1(ngw_context* ctx, const double* x, long ldx, const double* sx, int B, double* tpart,
                     double* t)
{
    const int nmu = ctx->n_intp, nv = ctx->nval, nc = ctx->ncond;
    const double* psi = ctx->psi.as<double>();
    const long ldp = ctx->ld_psi;
    hipStream_t s = ctx->stream;
    {
        ProfScope ps(ctx, KC_K1);
        int mt = (nmu + K1M - 1) / K1M, nt = (nv * B + K1N - 1) / K1N;
        hipLaunchKernelGGL(k1_f64_kernel, dim3(mt * nt), dim3(GNT), 0, s, ColMajorA{psi + (long)nv * ldp, ldp, nmu},
                           XprimeB{x, sx, ldx, nv, nv * B}, psi, ldp, nmu, nc, nv, B, tpart, (long)nmu, mt,
                           mt * nt);
        NGW_HIP(hipGetLastError());
    }
    ProfScope ps(ctx, KC_CONVERT);
    return k1_reduce(s, tpart, (long)nmu, nmu, B, nv, t);
}

template <class BL>
static int launch_k2(ngw_context* ctx, const BL& bl, int B, double* u)
{
    const int nmu = ctx->n_intp;
    int mt = (nmu + 63) / 64, nt = (B + 63) / 64;
    hipLaunchKernelGGL(k2_f64_kernel<BL>, dim3(mt * nt), dim3(GNT), 0, ctx->stream,
                       ColMajorA{ctx->mmat.as<double>(), ctx->ld_m, nmu}, bl, nmu, B, u, (long)nmu, mt, mt * nt);
    NGW_HIP(hipGetLastError());
    return NGW_OK;
}

static int launch_k3(ngw_context* ctx, const double* u, int B, const OperatorEpilogue& ep)
{
    const int nmu = ctx->n_intp, nv = ctx->nval, nc = ctx->ncond;
    const double* psi = ctx->psi.as<double>();
    const long ldp = ctx->ld_psi;
    int mt = (nc + K3M - 1) / K3M, nt = (nv * B + K3N - 1) / K3N;
    hipLaunchKernelGGL(k3_f64_kernel, dim3(mt * nt), dim3(GNT), 0, ctx->stream,
                       ColMajorAT{psi + (long)nv * ldp, ldp, nc}, WeightedPsiVB{psi, ldp, u, (long)nmu, nv, nv * B},
                       nmu, nc, nv, B, ep, mt, mt * nt);
    NGW_HIP(hipGetLastError());
    return NGW_OK;
}

int operator_f64(ngw_context* ctx, const OperatorArgs& a)
{
    const int nmu = ctx->n_intp, nv = ctx->nval, B = a.nvec;
    if (B <= 0) return NGW_OK;
    if (B == 1) return operator_f64_one(ctx, a);
    const int S = k1_slots(nv);
    NGW_TRY(ctx->tpart.ensure(sizeof(double) * ((size_t)S + 1) * B * nmu));
    NGW_TRY(ctx->ubuf.ensure(sizeof(double) * (size_t)B * nmu));
    double* tpart = ctx->tpart.as<double>();
    double* t = tpart + (size_t)S * B * nmu;
    double* u = ctx->ubuf.as<double>();
    const double* sx = a.tda ? nullptr : ctx->sqrtr.as<double>();
    NGW_TRY(launch_k1(ctx, a.x, a.ldx, sx, B, tpart, t));
    {
        ProfScope ps(ctx, KC_K2);
        NGW_TRY(launch_k2(ctx, ColMajorB{t, (long)nmu, B}, B, u));
    }
    {
        ProfScope ps(ctx, KC_K3);
        NGW_TRY(launch_k3(ctx, u, B, a.epilogue(ctx)));
    }
    return NGW_OK;
}

// P vectors: v0(:, b) = s o C^T (M c_{m_b n_b})   (s = 1 for TDA),
// SIGMA/calculate_sigma_lanczos_BLOCK.F90z:286-358.
int pvectors_f64(ngw_context* ctx, const int* d_mband, const int* d_nband, int nvec, int tda,
                 double* d_p, long ldp_out)
{
    const int nmu = ctx->n_intp;
    NGW_TRY(ctx->ubuf.ensure(sizeof(double) * (size_t)nvec * nmu));
    double* u = ctx->ubuf.as<double>();
    ProfScope ps(ctx, KC_PVEC);
    NGW_TRY(launch_k2(ctx, PairB{ctx->psi.as<double>(), ctx->ld_psi, d_mband, d_nband, nvec}, nvec, u));
    OperatorEpilogue ep{};
    ep.out = d_p;
    ep.ldo = ldp_out;
    ep.s = tda ? nullptr : ctx->sqrtr.as<double>();
    ep.beta = 1.0;
    ep.alpha = 1.0;
    return launch_k3(ctx, u, nvec, ep);
}

}  // namespace ngw
