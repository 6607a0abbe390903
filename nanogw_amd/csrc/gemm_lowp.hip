// Mixed-precision ISDF operator (gfx950): the three GEMMs of the operator in
// FP32 / BF16 / BF16x3 on MFMA with FP32 accumulation, everything else
// (vectors in HBM, epilogues, Horner/Clenshaw combination, Lanczos) in FP64.
//
//   K1  T = PsiC x'        A = Psi_lp(:, c) (FP32 copy of Psi / sigma_c),
//                          B = sigma_c s(v,c) x_b(v,c) (FP64 -> MFMA type)
//   K2  u = M t            FP32 MFMA in every low-precision mode
//   K3  Y = PsiC^T W       A = Psi_lp, B = PsiV(mu,v) u(mu,b); Y *= sigma_c
//
// sigma_n = 2^round(log2 max_mu |Psi(mu, n)|) is a per-band power-of-two scale
// (exact, so the product is unchanged; it keeps every band's values near 1).

#include "operator_loaders.h"

#include <cmath>
#include <vector>

namespace ngw {

// ---------------------------------------------------------------- kernels
constexpr int LK1M = 128, LK1N = 64;

template <int KIND>
__global__ __launch_bounds__(GNT) void k1_lp_kernel(LpColA al, LpXprime bl, const double* psiv, long ldp, int nmu,
                                                    int nc, int nv, int nvec, double* tpart, long ldt, int mtiles,
                                                    int ntiles)
{
    __shared__ union {
        SmemLp<KIND, LK1M, LK1N> ab;
        double red[LK1M][LK1N + 2];
    } sm;
    const int tile = xcd_tile(blockIdx.x, ntiles);
    const int tm = tile % mtiles, tn = tile / mtiles;
    const int m0 = tm * LK1M, n0 = tn * LK1N;
    const int N = nv * nvec;
    AccLp<LK1M, LK1N> acc;
    gemm_lp_tile<KIND, LK1M, LK1N>(al, bl, m0, n0, nc, acc, sm.ab);
    for_acc_lp<LK1M, LK1N>(acc, [&](int row, int col, float val) {
        int gm = m0 + row, gn = n0 + col;
        double p = 0.0;
        if (gm < nmu && gn < N) p = (double)val * psiv[gm + (long)(gn % nv) * ldp];
        sm.red[row][col] = p;
    });
    __syncthreads();
    const int nend = min(n0 + LK1N, N);
    const int b_lo = n0 / nv, b_hi = (nend - 1) / nv;
    const int nbt = b_hi - b_lo + 1;
    for (int e = threadIdx.x; e < LK1M * nbt; e += GNT) {
        int row = e % LK1M, b = b_lo + e / LK1M;
        int c0 = max(n0, b * nv) - n0, c1 = min(nend, (b + 1) * nv) - n0;
        double s = 0.0;
        for (int cc = c0; cc < c1; ++cc) s += sm.red[row][cc];
        int gm = m0 + row;
        int slot = tn - (b * nv) / LK1N;
        if (gm < nmu) tpart[((long)slot * nvec + b) * ldt + gm] = s;
    }
}

__global__ __launch_bounds__(GNT) void k2_lp_kernel(LpColA al, LpColBd bl, int nmu, int nvec, double* u, long ldu,
                                                    int mtiles, int ntiles)
{
    __shared__ SmemLp<LP_F32, 64, 64> sm;
    const int tile = xcd_tile(blockIdx.x, ntiles);
    const int m0 = (tile % mtiles) * 64, n0 = (tile / mtiles) * 64;
    AccLp<64, 64> acc;
    gemm_lp_tile<LP_F32, 64, 64>(al, bl, m0, n0, nmu, acc, sm);
    for_acc_lp<64, 64>(acc, [&](int row, int col, float val) {
        int gm = m0 + row, gn = n0 + col;
        if (gm < nmu && gn < nvec) u[gm + (long)gn * ldu] = (double)val;
    });
}

constexpr int LK3M = 64, LK3N = 128;

template <int KIND>
__global__ __launch_bounds__(GNT) void k3_lp_kernel(LpColAT al, LpWeighted bl, const float* sigma_c, int nmu, int nc,
                                                    int nv, int nvec, OperatorEpilogue ep, int mtiles, int ntiles)
{
    __shared__ SmemLp<KIND, LK3M, LK3N> sm;
    const int tile = xcd_tile(blockIdx.x, ntiles);
    const int m0 = (tile % mtiles) * LK3M, n0 = (tile / mtiles) * LK3N;
    const int N = nv * nvec;
    AccLp<LK3M, LK3N> acc;
    gemm_lp_tile<KIND, LK3M, LK3N>(al, bl, m0, n0, nmu, acc, sm);
    for_acc_lp<LK3M, LK3N>(acc, [&](int row, int col, float val) {
        int c = m0 + row, j = n0 + col;
        if (c < nc && j < N) {
            int b = j / nv, v = j - b * nv;
            ep.apply(v + (long)c * nv, b, (double)val * (double)sigma_c[c]);
        }
    });
}

// Plain C = A B through the low-precision core (FP64 in/out, column-major):
// used by the library's GEMM tests to pin the MFMA lane layouts.
struct LpPlainA {
    static constexpr bool kContig = false;
    static constexpr int NRAW = 1;
    typedef double Raw;
    typedef const double* Handle;
    const double* a; long ld; int MN;
    __device__ Handle row(int m) const { return a + m; }
    __device__ void fetch(Handle h, int k, double (&r)[1]) const { r[0] = h[(long)k * ld]; }
    __device__ float combine(const double (&r)[1]) const { return (float)r[0]; }
};

template <int KIND>
__global__ __launch_bounds__(GNT) void gemm_lp_plain_kernel(LpPlainA al, LpColBd bl, int m, int n, int k, double* c,
                                                            long ldc, int mtiles, int ntiles)
{
    __shared__ SmemLp<KIND, 64, 64> sm;
    const int tile = xcd_tile(blockIdx.x, ntiles);
    const int m0 = (tile % mtiles) * 64, n0 = (tile / mtiles) * 64;
    AccLp<64, 64> acc;
    gemm_lp_tile<KIND, 64, 64>(al, bl, m0, n0, k, acc, sm);
    for_acc_lp<64, 64>(acc, [&](int row, int col, float val) {
        int gm = m0 + row, gn = n0 + col;
        if (gm < m && gn < n) c[gm + (long)gn * ldc] = (double)val;
    });
}

// ---------------------------------------------------------------- host side
__global__ void to_lowp_kernel(const double* psi, long ldp, int nmu, int nb, const float* sigma, float* out)
{
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t >= (long)nmu * nb) return;
    const int n = (int)(t / nmu), mu = (int)(t - (long)n * nmu);
    out[t] = (float)(psi[mu + (long)n * ldp] / (double)sigma[n]);
}

__global__ void to_f32_kernel(const double* in, long n, float* out)
{
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t < n) out[t] = (float)in[t];
}

int lowp_prepare(ngw_context* ctx)
{
    if (ctx->lowp_ready) return NGW_OK;
    const int nmu = ctx->n_intp, nb = ctx->nband;
    hipStream_t s = ctx->stream;
    // per-band power-of-two scales from the host copy of Psi's column maxima
    std::vector<double> hpsi((size_t)nmu * nb);
    NGW_HIP(hipMemcpyAsync(hpsi.data(), ctx->psi.ptr, sizeof(double) * hpsi.size(), hipMemcpyDeviceToHost, s));
    NGW_HIP(hipStreamSynchronize(s));
    std::vector<float> sig(nb);
    for (int n = 0; n < nb; ++n) {
        double mx = 0.0;
        for (int mu = 0; mu < nmu; ++mu) mx = std::max(mx, std::fabs(hpsi[mu + (size_t)n * nmu]));
        sig[n] = mx > 0.0 ? std::ldexp(1.0f, (int)std::lround(std::log2(mx))) : 1.0f;
    }
    NGW_TRY(ctx->psi_scale.ensure(sizeof(float) * nb));
    NGW_TRY(ctx->psi32.ensure(sizeof(float) * (size_t)nmu * nb));
    NGW_TRY(ctx->m32.ensure(sizeof(float) * (size_t)nmu * nmu));
    NGW_HIP(hipMemcpyAsync(ctx->psi_scale.ptr, sig.data(), sizeof(float) * nb, hipMemcpyHostToDevice, s));
    long n1 = (long)nmu * nb, n2 = (long)nmu * nmu;
    hipLaunchKernelGGL(to_lowp_kernel, dim3((unsigned)((n1 + 255) / 256)), dim3(256), 0, s, ctx->psi.as<double>(),
                       (long)ctx->ld_psi, nmu, nb, ctx->psi_scale.as<float>(), ctx->psi32.as<float>());
    NGW_HIP(hipGetLastError());
    hipLaunchKernelGGL(to_f32_kernel, dim3((unsigned)((n2 + 255) / 256)), dim3(256), 0, s, ctx->mmat.as<double>(), n2,
                       ctx->m32.as<float>());
    NGW_HIP(hipGetLastError());
    NGW_HIP(hipStreamSynchronize(s));
    ctx->lowp_ready = true;
    return NGW_OK;
}

template <int KIND>
static int operator_lowp_kind(ngw_context* ctx, const OperatorArgs& a)
{
    const int nmu = ctx->n_intp, nv = ctx->nval, nc = ctx->ncond, B = a.nvec;
    const int S = k1_slots(nv);
    NGW_TRY(ctx->tpart.ensure(sizeof(double) * ((size_t)S + 1) * B * nmu));
    NGW_TRY(ctx->ubuf.ensure(sizeof(double) * (size_t)B * nmu));
    double* tpart = ctx->tpart.as<double>();
    double* t = tpart + (size_t)S * B * nmu;
    double* u = ctx->ubuf.as<double>();
    const double* psi = ctx->psi.as<double>();
    const float* plp = ctx->psi32.as<float>();
    const float* sig = ctx->psi_scale.as<float>();
    const long ldp = ctx->ld_psi;
    hipStream_t s = ctx->stream;
    const double* sx = a.tda ? nullptr : ctx->sqrtr.as<double>();
    {
        ProfScope ps(ctx, KC_K1);
        int mt = (nmu + LK1M - 1) / LK1M, nt = (nv * B + LK1N - 1) / LK1N;
        hipLaunchKernelGGL(k1_lp_kernel<KIND>, dim3(mt * nt), dim3(GNT), 0, s,
                           LpColA{plp + (long)nv * nmu, (long)nmu, nmu},
                           LpXprime{a.x, sx, sig + nv, a.ldx, nv, nv * B}, psi, ldp, nmu, nc, nv, B, tpart,
                           (long)nmu, mt, mt * nt);
        NGW_HIP(hipGetLastError());
    }
    NGW_TRY(k1_reduce(s, tpart, (long)nmu, nmu, B, nv, t));
    {
        ProfScope ps(ctx, KC_K2);
        int mt = (nmu + 63) / 64, nt = (B + 63) / 64;
        hipLaunchKernelGGL(k2_lp_kernel, dim3(mt * nt), dim3(GNT), 0, s, LpColA{ctx->m32.as<float>(), (long)nmu, nmu},
                           LpColBd{t, (long)nmu, B}, nmu, B, u, (long)nmu, mt, mt * nt);
        NGW_HIP(hipGetLastError());
    }
    {
        ProfScope ps(ctx, KC_K3);
        int mt = (nc + LK3M - 1) / LK3M, nt = (nv * B + LK3N - 1) / LK3N;
        hipLaunchKernelGGL(k3_lp_kernel<KIND>, dim3(mt * nt), dim3(GNT), 0, s,
                           LpColAT{plp + (long)nv * nmu, (long)nmu, nc},
                           LpWeighted{psi, ldp, u, (long)nmu, nv, nv * B}, sig + nv, nmu, nc, nv, B, a.epilogue(ctx),
                           mt, mt * nt);
        NGW_HIP(hipGetLastError());
    }
    return NGW_OK;
}

int operator_lowp(ngw_context* ctx, const OperatorArgs& a)
{
    if (a.nvec <= 0) return NGW_OK;
    NGW_TRY(lowp_prepare(ctx));
    switch (ctx->precision) {
        case PREC_F32: return operator_lowp_kind<LP_F32>(ctx, a);
        case PREC_BF16: return operator_lowp_kind<LP_BF16>(ctx, a);
        case PREC_BF16X3: return operator_lowp_kind<LP_BF16X3>(ctx, a);
        default: return fail(NGW_ERR_STATE, "unknown precision");
    }
}

int gemm_lowp(hipStream_t s, int kind, int m, int n, int k, const double* a, long lda, const double* b, long ldb,
              double* c, long ldc)
{
    if (m <= 0 || n <= 0) return NGW_OK;
    int mt = (m + 63) / 64, nt = (n + 63) / 64;
    LpPlainA al{a, lda, m};
    LpColBd bl{b, ldb, n};
    switch (kind) {
        case LP_F32:
            hipLaunchKernelGGL(gemm_lp_plain_kernel<LP_F32>, dim3(mt * nt), dim3(GNT), 0, s, al, bl, m, n, k, c, ldc,
                               mt, mt * nt);
            break;
        case LP_BF16:
            hipLaunchKernelGGL(gemm_lp_plain_kernel<LP_BF16>, dim3(mt * nt), dim3(GNT), 0, s, al, bl, m, n, k, c, ldc,
                               mt, mt * nt);
            break;
        case LP_BF16X3:
            hipLaunchKernelGGL(gemm_lp_plain_kernel<LP_BF16X3>, dim3(mt * nt), dim3(GNT), 0, s, al, bl, m, n, k, c,
                               ldc, mt, mt * nt);
            break;
        default: return fail(NGW_ERR_ARG, "bad low-precision kind");
    }
    NGW_HIP(hipGetLastError());
    return NGW_OK;
}

}  // namespace ngw
