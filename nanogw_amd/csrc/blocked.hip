// The irrep-blocked Casida operator of NanoGW's lowcomsym routines
// (shared/apply_matvec_lowcomsym.F90:155-247 for C x, :249-395 for C^T u,
// :396-412 for the TDA / full combination), applied in factored form.
//
// The pairs of symmetry representation nmrep come in blocks: for each
// valence irrep the conduction irrep whose product with it is nmrep
// (:158-161); block b pairs its nv_b valence states (outer) with its nc_b
// conduction states (inner), pair index p = p_b + v nc_b + c.  The caller
// gathers the blocks' orbitals at the ISDF points into PsiV (n_intp x sum
// nv_b) and PsiC (n_intp x sum nc_b), block after block.  Per block the
// operator's two pair products are the factored K1 / K3 of the full-pair
// operator on (PsiV_b, PsiC_b); the blocks' K1 sums are added into one t,
// and K2 = M t is shared:
//   K1_b  t += sum_v PsiV_b(mu, v) sum_c PsiC_b(mu, c) x'(p_b + v nc_b + c)
//   K2    u  = M t
//   K3_b  Y(p_b + v nc_b + c) = sum_mu PsiC_b(mu, c) PsiV_b(mu, v) u(mu)
// with the same epilogue as the other operator forms.  The pair matrix
// (n_intp x ncv, 6 GB at 5000 points x 150,000 pairs) is never formed; the
// resident data are n_intp x (nv + nc) orbitals and M.  Blocks run one after
// another on the stream and t is summed in block order, so results do not
// depend on scheduling.  FP64 only.

#include "ngw_kernels.h"
#include "operator_loaders.h"

#include <algorithm>
#include <vector>

namespace ngw {

namespace {

// x'(c, j = v + nv b) = s(p) x_b(p), p = p0 + v nc + c; s == null means 1 (TDA).
// The sum index c runs along contiguous memory.
struct XprimeBlockB {
    static constexpr bool kContig = true;
    static constexpr int NRAW = 2;
    struct Handle { const double* x; const double* s; };
    const double* x; const double* s; long ldx; long p0; int nv; int nc; int MN;
    __device__ Handle row(int j) const
    {
        const int b = j / nv, v = j - b * nv;
        const long p = p0 + (long)v * nc;
        return {x + (long)b * ldx + p, s ? s + p : nullptr};
    }
    __device__ void fetch(const Handle& h, int c, double (&r)[2]) const
    {
        r[0] = h.x[c];
        r[1] = h.s ? h.s[c] : 1.0;
    }
    __device__ double combine(const double (&r)[2]) const { return r[1] * r[0]; }
};

constexpr int BK1M = 128, BK1N = 64, BK3M = 64, BK3N = 128;

int blk_slots(int nv) { return (nv - 1) / BK1N + 2; }

// K1 of one block: tile (mu, j = v + nv b) of PsiC_b x', times PsiV_b(mu, v),
// summed over the tile's v range of each vector b into slot partials
// tpart[(slot nvec + b) nmu + mu].
__global__ __launch_bounds__(GNT) void blk_k1_kernel(ColMajorA al, XprimeBlockB bl, const double* psiv, long ldp,
                                                     int nmu, int nc, int nv, int nvec, double* tpart, int mtiles,
                                                     int ntiles)
{
    __shared__ union {
        SmemF64<BK1M, BK1N> ab;
        double red[BK1M][BK1N + 2];
    } sm;
    const int tile = xcd_tile(blockIdx.x, ntiles);
    const int tm = tile % mtiles, tn = tile / mtiles;
    const int m0 = tm * BK1M, n0 = tn * BK1N;
    const int N = nv * nvec;
    AccF64<BK1M, BK1N> acc;
    gemm_f64_tile<BK1M, BK1N>(al, bl, m0, n0, nc, acc, sm.ab);
    for_acc_f64<BK1M, BK1N>(acc, [&](int row, int col, double val) {
        const int gm = m0 + row, gn = n0 + col;
        double p = 0.0;
        if (gm < nmu && gn < N) p = val * psiv[gm + (long)(gn % nv) * ldp];
        sm.red[row][col] = p;
    });
    __syncthreads();
    const int nend = min(n0 + BK1N, N);
    const int b_lo = n0 / nv, b_hi = (nend - 1) / nv;
    const int nbt = b_hi - b_lo + 1;
    for (int e = threadIdx.x; e < BK1M * nbt; e += GNT) {
        const int row = e % BK1M, b = b_lo + e / BK1M;
        const int c0 = max(n0, b * nv) - n0, c1 = min(nend, (b + 1) * nv) - n0;
        double s = 0.0;
        for (int cc = c0; cc < c1; ++cc) s += sm.red[row][cc];
        const int gm = m0 + row;
        const int slot = tn - (b * nv) / BK1N;
        if (gm < nmu) tpart[((long)slot * nvec + b) * nmu + gm] = s;
    }
}

// t(mu, b) (+)= the slot partials of vector b in slot order; the first block
// stores, later blocks add, so t is summed in block order.
__global__ void blk_reduce_kernel(const double* tpart, int nmu, int nvec, int nv, int add, double* t)
{
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= (long)nmu * nvec) return;
    const int b = (int)(i / nmu), mu = (int)(i - (long)b * nmu);
    const int first = (b * nv) / BK1N, last = ((b + 1) * nv - 1) / BK1N;
    double s = 0.0;
    for (int sl = 0; sl <= last - first; ++sl) s += tpart[((long)sl * nvec + b) * nmu + mu];
    t[i] = add ? t[i] + s : s;
}

// K3 of one block with the operator epilogue on pairs p0 + v nc + c.
__global__ __launch_bounds__(GNT) void blk_k3_kernel(ColMajorAT al, WeightedPsiVB bl, int nmu, int nc, int nv,
                                                     int nvec, long p0, OperatorEpilogue ep, int mtiles, int ntiles)
{
    __shared__ SmemF64<BK3M, BK3N> sm;
    const int tile = xcd_tile(blockIdx.x, ntiles);
    const int m0 = (tile % mtiles) * BK3M, n0 = (tile / mtiles) * BK3N;
    const int N = nv * nvec;
    AccF64<BK3M, BK3N> acc;
    gemm_f64_tile<BK3M, BK3N>(al, bl, m0, n0, nmu, acc, sm);
    for_acc_f64<BK3M, BK3N>(acc, [&](int row, int col, double val) {
        const int c = m0 + row, j = n0 + col;
        if (c < nc && j < N) {
            const int b = j / nv, v = j - b * nv;
            ep.apply(p0 + (long)v * nc + c, b, val);
        }
    });
}

}  // namespace

int operator_blocked(ngw_context* ctx, const OperatorArgs& a)
{
    const int B = a.nvec;
    if (B <= 0) return NGW_OK;
    if (a.flags) return fail(NGW_ERR_STATE, "irrep-blocked operator: FP64 vector blocks only");
    hipStream_t s = ctx->stream;
    const int nmu = ctx->n_intp;
    const int nblk = (int)ctx->blk_nv.size();
    int nv_tot = 0, smax = 1;
    for (int k = 0; k < nblk; ++k) {
        nv_tot += ctx->blk_nv[k];
        if (ctx->blk_nv[k] > 0) smax = std::max(smax, blk_slots(ctx->blk_nv[k]));
    }
    NGW_TRY(ctx->tpart.ensure(sizeof(double) * ((size_t)smax + 1) * B * nmu));
    NGW_TRY(ctx->ubuf.ensure(sizeof(double) * (size_t)B * nmu));
    double* tpart = ctx->tpart.as<double>();
    double* t = tpart + (size_t)smax * B * nmu;
    double* u = ctx->ubuf.as<double>();
    const double* psiv_all = ctx->psi.as<double>();
    const double* psic_all = psiv_all + (size_t)nv_tot * nmu;
    const long ldp = nmu;
    const double* sx = a.tda ? nullptr : ctx->sqrtr.as<double>();
    // K1 per block, t summed in block order (blocks without pairs add nothing)
    bool first = true;
    {
        long p0 = 0;
        int v0 = 0, c0 = 0;
        for (int k = 0; k < nblk; ++k) {
            const int nv = ctx->blk_nv[k], nc = ctx->blk_nc[k];
            if (nv > 0 && nc > 0) {
                const double* pv = psiv_all + (size_t)v0 * nmu;
                const double* pc = psic_all + (size_t)c0 * nmu;
                {
                    ProfScope ps(ctx, KC_K1);
                    const int mt = (nmu + BK1M - 1) / BK1M, nt = (nv * B + BK1N - 1) / BK1N;
                    hipLaunchKernelGGL(blk_k1_kernel, dim3(mt * nt), dim3(GNT), 0, s, ColMajorA{pc, ldp, nmu},
                                       XprimeBlockB{a.x, sx, a.ldx, p0, nv, nc, nv * B}, pv, ldp, nmu, nc, nv, B,
                                       tpart, mt, mt * nt);
                }
                ProfScope ps(ctx, KC_CONVERT);
                const long total = (long)nmu * B;
                hipLaunchKernelGGL(blk_reduce_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, tpart,
                                   nmu, B, nv, first ? 0 : 1, t);
                NGW_HIP(hipGetLastError());
                first = false;
            }
            p0 += (long)nv * nc;
            v0 += nv;
            c0 += nc;
        }
    }
    if (first) return fail(NGW_ERR_STATE, "irrep-blocked operator without pairs");
    {
        ProfScope ps(ctx, KC_K2);
        NGW_TRY(gemm_f64(s, nmu, B, nmu, ctx->mmat.as<double>(), ctx->ld_m, t, nmu, u, nmu));
    }
    const OperatorEpilogue ep = a.epilogue(ctx);
    long p0 = 0;
    int v0 = 0, c0 = 0;
    for (int k = 0; k < nblk; ++k) {
        const int nv = ctx->blk_nv[k], nc = ctx->blk_nc[k];
        if (nv > 0 && nc > 0) {
            ProfScope ps(ctx, KC_K3);
            const double* pv = psiv_all + (size_t)v0 * nmu;
            const double* pc = psic_all + (size_t)c0 * nmu;
            const int mt = (nc + BK3M - 1) / BK3M, nt = (nv * B + BK3N - 1) / BK3N;
            hipLaunchKernelGGL(blk_k3_kernel, dim3(mt * nt), dim3(GNT), 0, s, ColMajorAT{pc, ldp, nc},
                               WeightedPsiVB{pv, ldp, u, (long)nmu, nv, nv * B}, nmu, nc, nv, B, p0, ep, mt,
                               mt * nt);
            NGW_HIP(hipGetLastError());
        }
        p0 += (long)nv * nc;
        v0 += nv;
        c0 += nc;
    }
    return NGW_OK;
}

}  // namespace ngw

using namespace ngw;

extern "C" int ngw_set_isdf_blocks(ngw_context* ctx, int n_intp, int nblock, const int* nv_blk, const int* nc_blk,
                                   const double* psiv, int ld_v, const double* psic, int ld_c, const double* mmtrx,
                                   int ld_m, const double* sqrtr)
{
    if (!ctx) return fail(NGW_ERR_ARG, "null context");
    if (n_intp <= 0 || nblock <= 0 || !nv_blk || !nc_blk) return fail(NGW_ERR_ARG, "need n_intp > 0 and nblock > 0");
    if (ld_v < n_intp || ld_c < n_intp || ld_m < n_intp) return fail(NGW_ERR_ARG, "leading dimension < n_intp");
    if (!psiv || !psic || !mmtrx || !sqrtr) return fail(NGW_ERR_ARG, "null input array");
    long ncv = 0;
    int nv_tot = 0, nc_tot = 0;
    for (int k = 0; k < nblock; ++k) {
        if (nv_blk[k] < 0 || nc_blk[k] < 0) return fail(NGW_ERR_ARG, "negative block size");
        ncv += (long)nv_blk[k] * nc_blk[k];
        nv_tot += nv_blk[k];
        nc_tot += nc_blk[k];
    }
    if (ncv <= 0) return fail(NGW_ERR_ARG, "the blocks hold no (v, c) pairs");
    if (ncv > 0x7fffffffL) return fail(NGW_ERR_ARG, "more than 2^31 - 1 pairs");
    for (long i = 0; i < ncv; ++i)
        if (!(sqrtr[i] > 0.0)) return fail(NGW_ERR_ARG, "sqrtR must be positive");
    NGW_HIP(hipSetDevice(ctx->device));
    explicit_release(ctx);
    hipStream_t s = ctx->stream;
    ctx->n_intp = n_intp;
    ctx->ncv = ncv;
    ctx->nband = 0;
    ctx->nval = 0;
    ctx->ncond = 0;
    ctx->ld_psi = n_intp;
    ctx->ld_m = n_intp;
    ctx->lowp_ready = false;
    ctx->lp2_ready = false;
    ctx->precision = PREC_F64;
    const size_t cols = (size_t)nv_tot + nc_tot;
    NGW_TRY(ctx->psi.ensure(sizeof(double) * (size_t)n_intp * cols));
    double* dpsi = ctx->psi.as<double>();
    if (nv_tot > 0)
        NGW_HIP(hipMemcpy2DAsync(dpsi, sizeof(double) * n_intp, psiv, sizeof(double) * ld_v, sizeof(double) * n_intp,
                                 nv_tot, hipMemcpyHostToDevice, s));
    if (nc_tot > 0)
        NGW_HIP(hipMemcpy2DAsync(dpsi + (size_t)nv_tot * n_intp, sizeof(double) * n_intp, psic, sizeof(double) * ld_c,
                                 sizeof(double) * n_intp, nc_tot, hipMemcpyHostToDevice, s));
    NGW_TRY(ctx->mmat.ensure(sizeof(double) * (size_t)n_intp * n_intp));
    NGW_HIP(hipMemcpy2DAsync(ctx->mmat.ptr, sizeof(double) * n_intp, mmtrx, sizeof(double) * ld_m,
                             sizeof(double) * n_intp, n_intp, hipMemcpyHostToDevice, s));
    std::vector<double> d3(ncv), d2(ncv);
    for (long i = 0; i < ncv; ++i) {
        d3[i] = sqrtr[i] * sqrtr[i] * sqrtr[i];
        d2[i] = sqrtr[i] * sqrtr[i];
    }
    NGW_TRY(ctx->sqrtr.ensure(sizeof(double) * ncv));
    NGW_TRY(ctx->diag_full.ensure(sizeof(double) * ncv));
    NGW_TRY(ctx->diag_tda.ensure(sizeof(double) * ncv));
    NGW_HIP(hipMemcpyAsync(ctx->sqrtr.ptr, sqrtr, sizeof(double) * ncv, hipMemcpyHostToDevice, s));
    NGW_HIP(hipMemcpyAsync(ctx->diag_full.ptr, d3.data(), sizeof(double) * ncv, hipMemcpyHostToDevice, s));
    NGW_HIP(hipMemcpyAsync(ctx->diag_tda.ptr, d2.data(), sizeof(double) * ncv, hipMemcpyHostToDevice, s));
    NGW_HIP(hipStreamSynchronize(s));
    ctx->blk_nv.assign(nv_blk, nv_blk + nblock);
    ctx->blk_nc.assign(nc_blk, nc_blk + nblock);
    ctx->blocked = true;
    return NGW_OK;
}
