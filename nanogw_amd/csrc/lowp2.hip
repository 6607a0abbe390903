// Mixed-precision ISDF operator on pre-converted operands (gfx950).
//
// Same algebra as the FP64 path (shared/apply_matvec_BLOCK.F90:25-208, the
// operator applied in ISDF-factored form), with the three GEMMs on BF16 /
// BF16x3 / FP32 MFMA:
//   per system (once):  pct = (Psi_c / sigma_c)^T  [mu][c]   K1 A operand
//                       pc  =  Psi_c / sigma_c     [c][mu]   K3 A operand
//                       m32t = M^T FP32            [mu][nu]  K2 A operand
//   per operator application:
//     xprep  x'(j, c) = sigma_c s(v,c) x_b(v,c)   (transpose through LDS)
//     K1     T = pct x'^T, epilogue * Psi_v and reduce over v (LDS)
//     red    t(b, nu) = sum of K1 slots  -> FP32 K2 B operand
//     K2     u = M t (FP32 MFMA)
//     wprep  W(j, mu) = Psi_v(mu, v) u(mu, b)
//     K3     Y = pc W^T, epilogue * sigma_c, then the FP64 combination
//            (diagonal term, Horner/Clenshaw stage, Lanczos accumulate).
// Every kernel reads one request descriptor (by value for a single context,
// from a device table for the grouped batch engine), so molecules of
// different sizes share a launch; tiles never mix requests.

#include "gemm_core_v2.h"
#include "operator_loaders.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace ngw {

__device__ __forceinline__ int find_pre(const int* pre, int nreq, int bid)
{
    if (!pre) return 0;
    int lo = 0, hi = nreq;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (pre[mid] <= bid) lo = mid; else hi = mid;
    }
    return lo;
}

template <int KIND>
struct K1Tile {
    static constexpr int BM = 128, BN = (KIND == LP_BF16X3) ? 64 : 128;
};
constexpr int L2K3M = 64, L2K3N = 128, L2K2 = 64;

// K1 tile width for a precision mode (NGW_PREC_* / Precision value).
static int k1_bn_of_prec(int prec) { return prec == PREC_BF16X3 ? K1Tile<LP_BF16X3>::BN : K1Tile<LP_BF16>::BN; }

// ---------------------------------------------------------------- per-system preparation
template <int KIND>
__global__ void prep_psi_kernel(const double* psi, long ldp, int nmu, int nv, int nc, const float* sig, int ldc1,
                                int ldm1, void* pct_hi, void* pct_lo, void* pc_hi, void* pc_lo)
{
    typedef typename V2Traits<KIND>::elem E;
    const long total = (long)nmu * ldc1 + (long)nc * ldm1;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
        if (t < (long)nmu * ldc1) {
            const int mu = (int)(t / ldc1), c = (int)(t - (long)mu * ldc1);
            float v = c < nc ? (float)(psi[mu + (long)(nv + c) * ldp] / (double)sig[nv + c]) : 0.0f;
            put_v2<KIND>((E*)pct_hi, (E*)pct_lo, t, v);
        } else {
            const long u = t - (long)nmu * ldc1;
            const int c = (int)(u / ldm1), mu = (int)(u - (long)c * ldm1);
            float v = mu < nmu ? (float)(psi[mu + (long)(nv + c) * ldp] / (double)sig[nv + c]) : 0.0f;
            put_v2<KIND>((E*)pc_hi, (E*)pc_lo, u, v);
        }
    }
}

__global__ void prep_m_kernel(const double* m, long ldm, int nmu, int ldm1, const double* psi, long ldp, int nv,
                              float* m32t, float* psiv32)
{
    const long total = (long)nmu * ldm1 + (long)nmu * nv;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
        if (t < (long)nmu * ldm1) {
            const int mu = (int)(t / ldm1), nu = (int)(t - (long)mu * ldm1);
            m32t[t] = nu < nmu ? (float)m[mu + (long)nu * ldm] : 0.0f;  // row mu of M, k = nu
        } else {
            const long u = t - (long)nmu * ldm1;
            const int v = (int)(u / nmu), mu = (int)(u - (long)v * nmu);
            psiv32[u] = (float)psi[mu + (long)v * ldp];
        }
    }
}

static int pad32(int n) { return (n + 31) & ~31; }

int lp2_prepare(ngw_context* ctx)
{
    if (ctx->lp2_ready) return NGW_OK;
    NGW_TRY(lowp_prepare(ctx));  // per-band scales
    const int nmu = ctx->n_intp, nv = ctx->nval, nc = ctx->ncond;
    ctx->ldc1 = pad32(nc);
    ctx->ldm1 = pad32(nmu);
    const size_t n1 = (size_t)nmu * ctx->ldc1, n2 = (size_t)nc * ctx->ldm1;
    NGW_TRY(ctx->pct32.ensure(4 * n1));
    NGW_TRY(ctx->pc32.ensure(4 * n2));
    NGW_TRY(ctx->pctb.ensure(2 * 2 * n1));  // hi then lo
    NGW_TRY(ctx->pcb.ensure(2 * 2 * n2));
    NGW_TRY(ctx->m32t.ensure(4 * (size_t)nmu * ctx->ldm1));
    NGW_TRY(ctx->psiv32.ensure(4 * (size_t)nmu * nv));
    hipStream_t s = ctx->stream;
    const double* psi = ctx->psi.as<double>();
    const float* sig = ctx->psi_scale.as<float>();
    unsigned short* pctb = ctx->pctb.as<unsigned short>();
    unsigned short* pcb = ctx->pcb.as<unsigned short>();
    hipLaunchKernelGGL(prep_psi_kernel<LP_F32>, dim3(2048), dim3(256), 0, s, psi, (long)ctx->ld_psi, nmu, nv, nc, sig,
                       ctx->ldc1, ctx->ldm1, (void*)ctx->pct32.ptr, (void*)nullptr, (void*)ctx->pc32.ptr, (void*)nullptr);
    hipLaunchKernelGGL(prep_psi_kernel<LP_BF16X3>, dim3(2048), dim3(256), 0, s, psi, (long)ctx->ld_psi, nmu, nv, nc,
                       sig, ctx->ldc1, ctx->ldm1, (void*)pctb, (void*)(pctb + n1), (void*)pcb, (void*)(pcb + n2));
    hipLaunchKernelGGL(prep_m_kernel, dim3(2048), dim3(256), 0, s, ctx->mmat.as<double>(), (long)ctx->ld_m, nmu,
                       ctx->ldm1, psi, (long)ctx->ld_psi, nv, ctx->m32t.as<float>(), ctx->psiv32.as<float>());
    NGW_HIP(hipGetLastError());
    NGW_HIP(hipStreamSynchronize(s));
    ctx->lp2_ready = true;
    return NGW_OK;
}

#define LP_PICK_OP()                                                     \
    const int r = find_pre(pre, nreq, blockIdx.x);                       \
    const LpOp op = ops ? ops[r] : op1;                                  \
    const int local = pre ? (int)blockIdx.x - pre[r] : (int)blockIdx.x;

// ---------------------------------------------------------------- conversion passes
// x'(j = v + nv b, c) = sigma_c s(v,c) x_b(v,c); 32 x 32 (v, c) tiles through LDS.
template <int KIND>
__global__ __launch_bounds__(256) void xprep_kernel(const LpOp* ops, LpOp op1, const LpStage* st, LpStage st1,
                                                    int nstage, int s, const int* pre, int nreq)
{
    typedef typename V2Traits<KIND>::elem E;
    __shared__ float tile[32][33];
    LP_PICK_OP();
    const double* x = ops ? st[r * nstage + s].x : st1.x;
    const int nvt = (op.nv + 31) / 32, nct = op.ldc1 / 32;
    const int b = local / (nvt * nct);
    const int rem = local - b * nvt * nct;
    const int v0 = (rem / nct) * 32, c0 = (rem % nct) * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int v = v0 + tx, c = c0 + ty + 8 * i;
        float val = 0.0f;
        if (v < op.nv && c < op.nc) {
            const long idx = v + (long)c * op.nv;
            double xv = x[idx + (long)b * op.ld];
            double xs = op.tda ? xv : op.sqrtr[idx] * xv;
            val = (float)(xs * (double)op.sig[op.nv + c]);
        }
        tile[ty + 8 * i][tx] = val;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int c = c0 + tx, v = v0 + ty + 8 * i;
        if (v < op.nv)
            put_v2<KIND>((E*)op.xlp_hi, (E*)op.xlp_lo, (long)(v + op.nv * b) * op.ldc1 + c, tile[tx][ty + 8 * i]);
    }
}

// W(j = v + nv b, mu) = Psi_v(mu, v) u(mu, b)
template <int KIND>
__global__ __launch_bounds__(256) void wprep_kernel(const LpOp* ops, LpOp op1, const int* pre, int nreq)
{
    typedef typename V2Traits<KIND>::elem E;
    LP_PICK_OP();
    const int per_row = op.ldm1 / 256 + ((op.ldm1 % 256) ? 1 : 0);
    const int j = local / per_row;
    const int mu = (local - j * per_row) * 256 + threadIdx.x;
    if (mu >= op.ldm1) return;
    const int b = j / op.nv, v = j - b * op.nv;
    float val = 0.0f;
    if (mu < op.nmu) val = (float)(op.psi[mu + (long)v * op.ldp] * op.u[mu + (long)b * op.nmu]);
    put_v2<KIND>((E*)op.wlp_hi, (E*)op.wlp_lo, (long)j * op.ldm1 + mu, val);
}

// t32(b, nu) = sum of the K1 slots covering vector b (fixed order); zero padding.
__global__ __launch_bounds__(256) void red2_kernel(const LpOp* ops, LpOp op1, const int* pre, int nreq, int bn)
{
    LP_PICK_OP();
    const long i = (long)local * 256 + threadIdx.x;
    if (i >= (long)op.nvec * op.ldm1) return;
    const int b = (int)(i / op.ldm1), nu = (int)(i - (long)b * op.ldm1);
    float out = 0.0f;
    if (nu < op.nmu) {
        const int first = (b * op.nv) / bn, last = ((b + 1) * op.nv - 1) / bn;
        double acc = 0.0;
        for (int sl = 0; sl <= last - first; ++sl) acc += op.tpart[((long)sl * op.nvec + b) * op.nmu + nu];
        out = (float)acc;
    }
    op.t32[i] = out;
}

// ---------------------------------------------------------------- GEMMs
template <int KIND>
__global__ __launch_bounds__(GNT) void k1v2_kernel(const LpOp* ops, LpOp op1, const int* pre, int nreq)
{
    constexpr int BM = K1Tile<KIND>::BM, BN = K1Tile<KIND>::BN;
    typedef typename V2Traits<KIND>::elem E;
    __shared__ union {
        SmemV2<KIND, BM, BN> ab;
        float red[BM][BN + 1];
    } sm;
    LP_PICK_OP();
    const int nv = op.nv, N = nv * op.nvec;
    const int mt = (op.nmu + BM - 1) / BM;
    const int tm = local % mt, tn = local / mt;
    const int m0 = tm * BM, n0 = tn * BN;
    AccLp<BM, BN> acc;
    gemm_v2_tile<KIND, BM, BN>(OperandV2<KIND>{(const E*)op.pct_hi, (const E*)op.pct_lo, op.ldc1, op.nmu},
                               OperandV2<KIND>{(const E*)op.xlp_hi, (const E*)op.xlp_lo, op.ldc1, N}, m0, n0, op.ldc1,
                               acc, sm.ab);
    for_acc_lp<BM, BN>(acc, [&](int row, int col, float val) {
        int gm = m0 + row, gn = n0 + col;
        float p = 0.0f;
        if (gm < op.nmu && gn < N) p = val * op.psiv32[gm + (long)(gn % nv) * op.nmu];
        sm.red[row][col] = p;
    });
    __syncthreads();
    const int nend = min(n0 + BN, N);
    const int b_lo = n0 / nv, b_hi = (nend - 1) / nv;
    const int nbt = b_hi - b_lo + 1;
    for (int e = threadIdx.x; e < BM * nbt; e += GNT) {
        int row = e % BM, b = b_lo + e / BM;
        int c0 = max(n0, b * nv) - n0, c1 = min(nend, (b + 1) * nv) - n0;
        float s = 0.0f;
        for (int cc = c0; cc < c1; ++cc) s += sm.red[row][cc];
        int gm = m0 + row;
        int slot = tn - (b * nv) / BN;
        if (gm < op.nmu) op.tpart[((long)slot * op.nvec + b) * op.nmu + gm] = (double)s;
    }
}

__global__ __launch_bounds__(GNT) void k2v2_kernel(const LpOp* ops, LpOp op1, const int* pre, int nreq)
{
    __shared__ SmemV2<LP_F32, L2K2, L2K2> sm;
    LP_PICK_OP();
    const int mt = (op.nmu + L2K2 - 1) / L2K2;
    const int m0 = (local % mt) * L2K2, n0 = (local / mt) * L2K2;
    AccLp<L2K2, L2K2> acc;
    gemm_v2_tile<LP_F32, L2K2, L2K2>(OperandV2<LP_F32>{op.m32t, nullptr, op.ldm1, op.nmu},
                                     OperandV2<LP_F32>{op.t32, nullptr, op.ldm1, op.nvec}, m0, n0, op.ldm1, acc, sm);
    for_acc_lp<L2K2, L2K2>(acc, [&](int row, int col, float val) {
        int gm = m0 + row, gn = n0 + col;
        if (gm < op.nmu && gn < op.nvec) op.u[gm + (long)gn * op.nmu] = (double)val;
    });
}

template <int KIND>
__global__ __launch_bounds__(GNT) void k3v2_kernel(const LpOp* ops, LpOp op1, const LpStage* st, LpStage st1,
                                                   int nstage, int s, const int* pre, int nreq)
{
    typedef typename V2Traits<KIND>::elem E;
    __shared__ SmemV2<KIND, L2K3M, L2K3N> sm;
    LP_PICK_OP();
    const LpStage sg = ops ? st[r * nstage + s] : st1;
    const int nv = op.nv, N = nv * op.nvec;
    const int mt = (op.nc + L2K3M - 1) / L2K3M;
    const int m0 = (local % mt) * L2K3M, n0 = (local / mt) * L2K3N;
    AccLp<L2K3M, L2K3N> acc;
    gemm_v2_tile<KIND, L2K3M, L2K3N>(OperandV2<KIND>{(const E*)op.pc_hi, (const E*)op.pc_lo, op.ldm1, op.nc},
                                     OperandV2<KIND>{(const E*)op.wlp_hi, (const E*)op.wlp_lo, op.ldm1, N}, m0, n0,
                                     op.ldm1, acc, sm);
    OperatorEpilogue ep;
    ep.x = sg.x; ep.ldx = op.ld;
    ep.d = op.tda ? op.dtda : op.dfull;
    ep.s = op.tda ? nullptr : op.sqrtr;
    ep.beta = op.tda ? 2.0 : 4.0;
    ep.alpha = sg.alpha;
    ep.bx = sg.bx;
    ep.y = sg.y; ep.ldy = op.ld; ep.by = sg.by;
    ep.v0 = op.q; ep.ldv0 = op.ld; ep.gamma = sg.gamma;
    const bool acc_on = sg.accumulate == 1 || (sg.accumulate == 2 && *op.kstep > 1);
    ep.acc = acc_on ? sg.out : nullptr; ep.ldacc = op.ld;
    ep.out = sg.out; ep.ldo = op.ld;
    const float* sigc = op.sig + nv;
    for_acc_lp<L2K3M, L2K3N>(acc, [&](int row, int col, float val) {
        int c = m0 + row, j = n0 + col;
        if (c < op.nc && j < N) {
            int b = j / nv, v = j - b * nv;
            ep.apply(v + (long)c * nv, b, (double)val * (double)sigc[c]);
        }
    });
}

// ---------------------------------------------------------------- launch plumbing
// Number of workgroups of each kernel for one request.
struct LpGrid {
    int xprep, k1, red, k2, wprep, k3;
};

static LpGrid lp_grid(int kind, int nmu, int nv, int nc, int nvec, int ldc1, int ldm1)
{
    LpGrid g;
    const int bn = k1_bn_of_prec(kind);
    g.xprep = nvec * ((nv + 31) / 32) * (ldc1 / 32);
    g.k1 = ((nmu + 127) / 128) * ((nv * nvec + bn - 1) / bn);
    g.red = (int)(((long)nvec * ldm1 + 255) / 256);
    g.k2 = ((nmu + L2K2 - 1) / L2K2) * ((nvec + L2K2 - 1) / L2K2);
    g.wprep = nv * nvec * (ldm1 / 256 + ((ldm1 % 256) ? 1 : 0));
    g.k3 = ((nc + L2K3M - 1) / L2K3M) * ((nv * nvec + L2K3N - 1) / L2K3N);
    return g;
}

template <int KIND>
static int lp_launch(hipStream_t st, const LpOp* ops, const LpOp& op1, const LpStage* stg, const LpStage& st1,
                     int nstage, int s, const int* const* pre, const int* tot, int nreq, ngw_context* prof)
{
    {
        ProfScope ps(prof, KC_CONVERT);
        hipLaunchKernelGGL(xprep_kernel<KIND>, dim3(tot[0]), dim3(256), 0, st, ops, op1, stg, st1, nstage, s, pre[0],
                           nreq);
    }
    {
        ProfScope ps(prof, KC_K1);
        hipLaunchKernelGGL(k1v2_kernel<KIND>, dim3(tot[1]), dim3(GNT), 0, st, ops, op1, pre[1], nreq);
    }
    {
        ProfScope ps(prof, KC_CONVERT);
        hipLaunchKernelGGL(red2_kernel, dim3(tot[2]), dim3(256), 0, st, ops, op1, pre[2], nreq, K1Tile<KIND>::BN);
    }
    {
        ProfScope ps(prof, KC_K2);
        hipLaunchKernelGGL(k2v2_kernel, dim3(tot[3]), dim3(GNT), 0, st, ops, op1, pre[3], nreq);
    }
    {
        ProfScope ps(prof, KC_CONVERT);
        hipLaunchKernelGGL(wprep_kernel<KIND>, dim3(tot[4]), dim3(256), 0, st, ops, op1, pre[4], nreq);
    }
    {
        ProfScope ps(prof, KC_K3);
        hipLaunchKernelGGL(k3v2_kernel<KIND>, dim3(tot[5]), dim3(GNT), 0, st, ops, op1, stg, st1, nstage, s, pre[5],
                           nreq);
    }
    NGW_HIP(hipGetLastError());
    return NGW_OK;
}

int lp2_launch_any(int kind, hipStream_t st, const LpOp* ops, const LpOp& op1, const LpStage* stg, const LpStage& st1,
                   int nstage, int s, const int* const* pre, const int* tot, int nreq, ngw_context* prof)
{
    switch (kind) {
        case PREC_F32: return lp_launch<LP_F32>(st, ops, op1, stg, st1, nstage, s, pre, tot, nreq, prof);
        case PREC_BF16: return lp_launch<LP_BF16>(st, ops, op1, stg, st1, nstage, s, pre, tot, nreq, prof);
        case PREC_BF16X3: return lp_launch<LP_BF16X3>(st, ops, op1, stg, st1, nstage, s, pre, tot, nreq, prof);
    }
    return fail(NGW_ERR_STATE, "unknown precision");
}

// Bytes of the per-request low-precision workspaces.
size_t lp2_work_bytes(int kind, int nmu, int nv, int nc, int nvec)
{
    const size_t es = kind == PREC_F32 ? 4 : 2, ns = kind == PREC_BF16X3 ? 2 : 1;
    const size_t ldc1 = pad32(nc), ldm1 = pad32(nmu), N = (size_t)nv * nvec;
    return es * ns * N * (ldc1 + ldm1) + 4 * (size_t)nvec * ldm1 + 256 * 4;
}

// Fill the descriptor of one request from its operator and workspaces.
void lp2_fill(const ngw_context* op, int kind, int nvec, long ld, int tda, char* work, double* tpart, double* u,
              LpOp& d)
{
    const size_t es = kind == PREC_F32 ? 4 : 2;
    const size_t n1 = (size_t)op->n_intp * op->ldc1, n2 = (size_t)op->ncond * op->ldm1;
    if (kind == PREC_F32) {
        d.pct_hi = op->pct32.ptr; d.pct_lo = nullptr;
        d.pc_hi = op->pc32.ptr; d.pc_lo = nullptr;
    } else {
        d.pct_hi = op->pctb.ptr; d.pct_lo = op->pctb.as<unsigned short>() + n1;
        d.pc_hi = op->pcb.ptr; d.pc_lo = op->pcb.as<unsigned short>() + n2;
    }
    d.m32t = op->m32t.as<float>(); d.psiv32 = op->psiv32.as<float>();
    d.psi = op->psi.as<double>(); d.ldp = op->ld_psi;
    d.sig = op->psi_scale.as<float>();
    d.sqrtr = op->sqrtr.as<double>(); d.dfull = op->diag_full.as<double>(); d.dtda = op->diag_tda.as<double>();
    d.nmu = op->n_intp; d.nv = op->nval; d.nc = op->ncond; d.nvec = nvec; d.ldc1 = op->ldc1; d.ldm1 = op->ldm1;
    d.ld = ld; d.tda = tda;
    const size_t N = (size_t)op->nval * nvec;
    const size_t xsz = es * N * op->ldc1, wsz = es * N * op->ldm1;
    d.xlp_hi = work;
    d.xlp_lo = kind == PREC_BF16X3 ? work + xsz : nullptr;
    char* w0 = work + xsz * (kind == PREC_BF16X3 ? 2 : 1);
    d.wlp_hi = w0;
    d.wlp_lo = kind == PREC_BF16X3 ? w0 + wsz : nullptr;
    char* t0 = w0 + wsz * (kind == PREC_BF16X3 ? 2 : 1);
    t0 = reinterpret_cast<char*>((reinterpret_cast<uintptr_t>(t0) + 255) & ~uintptr_t(255));
    d.t32 = reinterpret_cast<float*>(t0);
    d.tpart = tpart; d.u = u;
    d.q = nullptr; d.w = nullptr; d.a = nullptr; d.b = nullptr; d.niter = 0; d.kstep = nullptr;
}

void lp2_grid_counts(int kind, const LpOp& d, int* tot)
{
    LpGrid g = lp_grid(kind, d.nmu, d.nv, d.nc, d.nvec, d.ldc1, d.ldm1);
    tot[0] = g.xprep; tot[1] = g.k1; tot[2] = g.red; tot[3] = g.k2; tot[4] = g.wprep; tot[5] = g.k3;
}

// Single-context operator application in a low-precision mode.
int operator_lp2(ngw_context* ctx, const OperatorArgs& a)
{
    const int B = a.nvec;
    if (B <= 0) return NGW_OK;
    NGW_TRY(lp2_prepare(ctx));
    const int kind = ctx->precision, nmu = ctx->n_intp, nv = ctx->nval;
    const int S = std::max(k1_slots(nv), (nv - 1) / k1_bn_of_prec(kind) + 2);
    NGW_TRY(ctx->tpart.ensure(sizeof(double) * (size_t)S * B * nmu));
    NGW_TRY(ctx->ubuf.ensure(sizeof(double) * (size_t)B * nmu));
    NGW_TRY(ctx->lpwork.ensure(lp2_work_bytes(kind, nmu, nv, ctx->ncond, B)));
    LpOp d;
    lp2_fill(ctx, kind, B, a.ldx, a.tda, ctx->lpwork.as<char>(), ctx->tpart.as<double>(), ctx->ubuf.as<double>(), d);
    if (a.ldo != a.ldx || (a.v0 && a.ldv0 != a.ldx) || (a.y && a.ldy != a.ldx) || (a.acc && a.acc != a.out))
        return fail(NGW_ERR_ARG, "low-precision operator needs one leading dimension and acc == out");
    d.q = const_cast<double*>(a.v0);
    LpStage st{a.x, a.out, a.y, a.alpha, a.bx, a.by, a.v0 ? a.gamma : 0.0, a.acc ? 1 : 0};
    int tot[6];
    lp2_grid_counts(kind, d, tot);
    const int* pre[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    return lp2_launch_any(kind, ctx->stream, nullptr, d, nullptr, st, 1, 0, pre, tot, 1, ctx);
}

}  // namespace ngw
