// Per-band scales of the low-precision operator and a plain low-precision
// GEMM on the same MFMA core as the operator (used by the library's tests to
// pin the gfx950 lane layouts of that core).
//
// sigma_n = 2^round(log2 max_mu |Psi(mu, n)|) is a per-band power-of-two scale
// (exact, so the product is unchanged; it keeps every band's values near 1).

#include "gemm_core_v2.h"
#include "ngw_kernels.h"

#include <cmath>
#include <vector>

namespace ngw {

int lowp_prepare(ngw_context* ctx)
{
    if (ctx->lowp_ready) return NGW_OK;
    const int nmu = ctx->n_intp, nb = ctx->nband;
    hipStream_t s = ctx->stream;
    // per-band power-of-two scales from the host copy of Psi's column maxima
    std::vector<double> hpsi((size_t)nmu * nb);
    NGW_HIP(hipMemcpy2DAsync(hpsi.data(), sizeof(double) * nmu, ctx->psi.ptr, sizeof(double) * ctx->ld_psi,
                             sizeof(double) * nmu, nb, hipMemcpyDeviceToHost, s));
    NGW_HIP(hipStreamSynchronize(s));
    std::vector<float> sig(nb);
    for (int n = 0; n < nb; ++n) {
        double mx = 0.0;
        for (int mu = 0; mu < nmu; ++mu) mx = std::max(mx, std::fabs(hpsi[mu + (size_t)n * nmu]));
        sig[n] = mx > 0.0 ? std::ldexp(1.0f, (int)std::lround(std::log2(mx))) : 1.0f;
    }
    NGW_TRY(ctx->psi_scale.ensure(sizeof(float) * nb));
    NGW_HIP(hipMemcpyAsync(ctx->psi_scale.ptr, sig.data(), sizeof(float) * nb, hipMemcpyHostToDevice, s));
    NGW_HIP(hipStreamSynchronize(s));
    ctx->lowp_ready = true;
    return NGW_OK;
}

// src(r, k) = base[r * rs + k * ks] (FP64) -> dst[r][kpad] in the MFMA type,
// zero for k >= K and for the extra row r = rows.
template <int KIND>
__global__ void to_rows_kernel(const double* base, long rs, long ks, int rows, int K, int kpad, void* hi, void* lo)
{
    typedef typename V2Traits<KIND>::elem E;
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t >= (long)(rows + 1) * kpad) return;
    const int r = (int)(t / kpad), k = (int)(t - (long)r * kpad);
    const float v = (k < K && r < rows) ? (float)base[(long)r * rs + (long)k * ks] : 0.0f;
    put_v2<KIND>((E*)hi, (E*)lo, t, v);
}

template <int KIND>
__global__ __launch_bounds__(GNT) void gemm_plain_v2_kernel(OperandV2<KIND> A, OperandV2<KIND> B, int m, int n,
                                                            int kpad, double* c, long ldc, int mtiles)
{
    __shared__ SmemV2<KIND, 64, 64, 1> sm;
    const int tile = xcd_tile((int)blockIdx.x, (int)gridDim.x);
    const int m0 = (tile % mtiles) * 64, n0 = (tile / mtiles) * 64;
    AccLp<64, 64> acc;
    gemm_v2_tile<KIND, 64, 64, 2>(A, B, m0, n0, kpad, acc, sm);  // swizzled LDS rows
    for_acc_lp<64, 64>(acc, [&](int row, int col, float val) {
        int gm = m0 + row, gn = n0 + col;
        if (gm < m && gn < n) c[gm + (long)gn * ldc] = (double)val;
    });
}

template <int KIND>
static int gemm_plain(hipStream_t s, int m, int n, int k, const double* a, long lda, const double* b, long ldb,
                      double* c, long ldc)
{
    typedef typename V2Traits<KIND>::elem E;
    const int kpad = (k + LP_KPAD - 1) / LP_KPAD * LP_KPAD;
    const int ns = V2Traits<KIND>::NSPLIT;
    const size_t na = (size_t)(m + 1) * kpad, nbb = (size_t)(n + 1) * kpad;  // + zero row
    DevBuf wa, wb;
    NGW_TRY(wa.ensure(sizeof(E) * na * ns));
    NGW_TRY(wb.ensure(sizeof(E) * nbb * ns));
    E* ah = wa.as<E>();
    E* bh = wb.as<E>();
    E* al = ns == 2 ? ah + na : nullptr;
    E* bl = ns == 2 ? bh + nbb : nullptr;
    // A(m, k) = a[m + k lda]: row m, stride 1 between rows, lda along k
    hipLaunchKernelGGL(to_rows_kernel<KIND>, dim3((unsigned)((na + 255) / 256)), dim3(256), 0, s, a, 1L, lda, m, k,
                       kpad, (void*)ah, (void*)al);
    // B(k, n) = b[k + n ldb]: row n, ldb between rows, stride 1 along k
    hipLaunchKernelGGL(to_rows_kernel<KIND>, dim3((unsigned)((nbb + 255) / 256)), dim3(256), 0, s, b, ldb, 1L, n, k,
                       kpad, (void*)bh, (void*)bl);
    const int mt = (m + 63) / 64, nt = (n + 63) / 64;
    hipLaunchKernelGGL(gemm_plain_v2_kernel<KIND>, dim3(mt * nt), dim3(GNT), 0, s,
                       OperandV2<KIND>{ah, al, kpad, m}, OperandV2<KIND>{bh, bl, kpad, n}, m, n, kpad, c, ldc, mt);
    NGW_HIP(hipGetLastError());
    NGW_HIP(hipStreamSynchronize(s));  // the conversion buffers are freed on return
    return NGW_OK;
}

int gemm_lowp(hipStream_t s, int kind, int m, int n, int k, const double* a, long lda, const double* b, long ldb,
              double* c, long ldc)
{
    if (m <= 0 || n <= 0) return NGW_OK;
    if (k <= 0) return fail(NGW_ERR_ARG, "k must be positive");
    switch (kind) {
        case LP_F32: return gemm_plain<LP_F32>(s, m, n, k, a, lda, b, ldb, c, ldc);
        case LP_BF16: return gemm_plain<LP_BF16>(s, m, n, k, a, lda, b, ldb, c, ldc);
        case LP_BF16X3: return gemm_plain<LP_BF16X3>(s, m, n, k, a, lda, b, ldb, c, ldc);
        case LP_F16: return gemm_plain<LP_F16>(s, m, n, k, a, lda, b, ldb, c, ldc);
    }
    return fail(NGW_ERR_ARG, "bad low-precision kind");
}

}  // namespace ngw
