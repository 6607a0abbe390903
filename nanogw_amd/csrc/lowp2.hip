// Mixed-precision ISDF operator on pre-converted operands (gfx950).
//
// Same algebra as the FP64 path (shared/apply_matvec_BLOCK.F90:25-208, the
// operator applied in ISDF-factored form), with the three GEMMs on BF16 /
// FP16 / BF16x3 / FP32 MFMA:
//   per system (once):  pct = (Psi_c / sigma_c)^T  [mu][c]   K1 A operand
//                       pc  =  Psi_c / sigma_c     [c][mu]   K3 A operand
//                       m32t = M^T FP32            [mu][nu]  K2 A operand
//                       psiv32 = Psi_v^T FP32      [v][mu]   wprep (K3 B factor), K1 epilogue
//   per operator application:
//     xprep  x'(j, c) = sigma_c s(v,c) x_b(v,c)   (transpose through LDS)
//     K1     T = pct x'^T, epilogue * Psi_v and reduce over v (LDS)
//     red    t(b, nu) = sum of K1 slots  -> FP32 K2 B operand
//     K2     u = M t (FP32 MFMA), stored FP32 [b][mu]
//     wprep  W(j, mu) = Psi_v(mu, v) u(mu, b)  (8 elements per thread)
//     K3     Y = pc W^T, epilogue * sigma_c, then the FP64 combination
//            (diagonal term, Horner/Clenshaw stage, Lanczos accumulate).
// All [row][k] operand arrays have k padded with zeros to a multiple of
// LP_KPAD and one extra all-zero row after the last (see OperandV2).
// FP16 (range 6e-8 .. 65504) adds exact power-of-two scales: pct and pc carry
// a factor 2^12; x' is scaled per vector so that its largest entry is at most
// 2^14 (xscale pass), W per (v, b) by the Psi_v band scale and max |u(:, b)|
// (taken in K2's epilogue); the K1 and K3 epilogues undo them.
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

constexpr int L2K2 = 64;

// Tile configurations of the K1 and K3 GEMMs: {BM, BN, D, PM} with D the
// number of slabs in flight (register prefetch depth) and PM the number of
// M-groups the 8 XCDs are split into.  XCD x owns the tile rectangle
// (M-group x mod PM) x (N-group x / PM): with PM > 1 each XCD touches only its
// share of the A operand, which then stays in that XCD's L2, while the tiles
// of one N column (M fastest inside the rectangle) share their B slab.  Index
// 0 is the default; the autotuner (ngw_autotune) times all of them on the
// caller's shape.
struct LpCfg {
    int bm, bn, d, pm;
    int fw = 0;  // K3 only: 1 = W formed in the slab loader (no wprep pass)
    int sw = 0;  // slabs: 0 padded LDS rows, 1 XOR-swizzled rows (both register-prefetched),
                 // 2 swizzled rows filled by direct-to-LDS loads into D buffers
};
constexpr int LP_NCFG = NGW_NTILE_CONFIGS;
constexpr LpCfg K1_CFG[3][LP_NCFG] = {
    {{128, 128, 2, 8, 0, 1}, {128, 128, 2, 8}, {128, 128, 2, 4, 0, 1}, {256, 128, 2, 8, 0, 1},
     {128, 128, 2, 2, 0, 1}, {128, 128, 3, 8, 0, 1}},                                             // FP32, BF16
    {{128, 128, 2, 8, 0, 1}, {128, 128, 2, 8}, {128, 64, 2, 8, 0, 1}, {128, 128, 1, 8, 0, 1},
     {128, 128, 2, 2, 0, 1}, {128, 64, 2, 2, 0, 1}},                                              // BF16x3
    {{128, 128, 2, 8, 0, 1}, {128, 128, 3, 8, 0, 2}, {128, 128, 2, 4, 0, 1}, {128, 128, 4, 8, 0, 2},
     {128, 256, 3, 8, 0, 2}, {128, 128, 3, 8, 0, 1}}};                                            // FP16
// PM = 1 or 2 (indices 4, 5 of the FP32/BF16 and BF16x3 sets): each XCD covers
// all or half of the c rows, so W is read once or twice instead of once per
// M-group.  FP16 K1 index 4 and K3 indices 4 and 5 are 128 x 256 tiles: per
// slab a workgroup moves 24 KB into LDS and its four 64 x 128 wave tiles read
// 48 KB for 128 MFMAs, 576 LDS bytes per MFMA against 768 for 128 x 128 and
// 640 for 128 x 192.  Which tile wins depends on the shape; K3's result does
// not (each element sums its K range in the same order in every tile), K1's
// does through its per-BN partial-sum slots.
constexpr LpCfg K3_CFG[3][LP_NCFG] = {
    {{128, 128, 2, 4}, {128, 128, 2, 4, 0, 1}, {128, 128, 2, 4, 1}, {128, 128, 2, 8, 1, 1},
     {128, 128, 2, 2, 0, 1}, {128, 128, 2, 1, 0, 1}},
    {{128, 128, 2, 4}, {128, 128, 2, 4, 0, 1}, {128, 128, 2, 4, 1}, {64, 128, 2, 8},
     {128, 128, 2, 2, 0, 1}, {128, 128, 2, 1, 0, 1}},
    {{128, 128, 2, 4}, {128, 192, 3, 1, 0, 2}, {128, 192, 3, 4, 0, 2}, {128, 192, 2, 4, 0, 1},
     {128, 256, 2, 4, 0, 2}, {128, 256, 3, 1, 0, 2}}};  // FP16: W from wprep
constexpr int NXCD = 8;
__host__ __device__ constexpr int cfg_set(int kind) { return kind == LP_BF16X3 ? 1 : kind == LP_F16 ? 2 : 0; }
static int cfg_set_prec(int prec) { return prec == PREC_BF16X3 ? 1 : prec == PREC_F16 ? 2 : 0; }

// FP16 operand scales (powers of two, so exact): operator operands x 2^12
// (|Psi_c / sigma_c| <= 2^0.5), vector operands up to 2^14; a K1 or K3
// product then carries 2^26.
constexpr float F16_OPS = 4096.0f, F16_VEC = 16384.0f;
constexpr double F16_UNSCALE = 1.0 / 67108864.0;

// Workgroups of one request's GEMM: NXCD rectangles of the largest size
// (workgroups of a smaller rectangle beyond its tiles exit at once).
__host__ __device__ inline int xcd_grid(int mt, int nt, int pm)
{
    const int pn = NXCD / pm;
    return NXCD * ((mt + pm - 1) / pm) * ((nt + pn - 1) / pn);
}

// Tile of a request-local workgroup index (the request's workgroups start at
// a multiple of NXCD, so local % NXCD is the XCD of round-robin dispatch);
// false for a workgroup past its rectangle.
__device__ __forceinline__ bool xcd_rect_tile(int local, int mt, int nt, int pm, int& tm, int& tn)
{
    const int pn = NXCD / pm;
    const int x = local % NXCD, i = local / NXCD;
    const int mg = x % pm, ng = x / pm;
    const int m_lo = mg * mt / pm, m_hi = (mg + 1) * mt / pm;
    const int n_lo = ng * nt / pn, n_hi = (ng + 1) * nt / pn;
    const int rows = m_hi - m_lo;
    if (rows <= 0 || i >= rows * (n_hi - n_lo)) return false;
    tm = m_lo + i % rows;
    tn = n_lo + i / rows;
    return true;
}

// Narrowest K1 tile of any configuration: sizes the slot partial sums.
constexpr int K1_BN_MIN = 64;

// ---------------------------------------------------------------- per-system preparation
template <int KIND>
__global__ void prep_psi_kernel(const double* psi, long ldp, int nmu, int nv, int nc, const float* sig, int ldc1,
                                int ldm1, void* pct_hi, void* pct_lo, void* pc_hi, void* pc_lo)
{
    typedef typename V2Traits<KIND>::elem E;
    const long n1 = (long)(nmu + 1) * ldc1, total = n1 + (long)(nc + 1) * ldm1;  // + zero rows
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
        if (t < n1) {
            const int mu = (int)(t / ldc1), c = (int)(t - (long)mu * ldc1);
            float v = (c < nc && mu < nmu) ? (float)(psi[mu + (long)(nv + c) * ldp] / (double)sig[nv + c]) : 0.0f;
            if constexpr (KIND == LP_F16) v *= F16_OPS;
            put_v2<KIND>((E*)pct_hi, (E*)pct_lo, t, v);
        } else {
            const long u = t - n1;
            const int c = (int)(u / ldm1), mu = (int)(u - (long)c * ldm1);
            float v = (mu < nmu && c < nc) ? (float)(psi[mu + (long)(nv + c) * ldp] / (double)sig[nv + c]) : 0.0f;
            if constexpr (KIND == LP_F16) v *= F16_OPS;
            put_v2<KIND>((E*)pc_hi, (E*)pc_lo, u, v);
        }
    }
}

__global__ void prep_m_kernel(const double* m, long ldm, int nmu, int ldm1, const double* psi, long ldp, int nv,
                              float* m32t, float* psiv32)
{
    const long n1 = (long)(nmu + 1) * ldm1, total = n1 + (long)nv * ldm1;  // m32t + zero row, psiv32
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
        if (t < n1) {
            const int mu = (int)(t / ldm1), nu = (int)(t - (long)mu * ldm1);
            m32t[t] = (nu < nmu && mu < nmu) ? (float)m[mu + (long)nu * ldm] : 0.0f;  // row mu of M, k = nu
        } else {
            const long u = t - n1;
            const int v = (int)(u / ldm1), mu = (int)(u - (long)v * ldm1);
            psiv32[u] = mu < nmu ? (float)psi[mu + (long)v * ldp] : 0.0f;
        }
    }
}

// Diagonal factors of the K3 epilogue as FP32 pairs (d', s') per pair index:
// full mode (d_full sqrtR, sqrtR), TDA (d_tda, 1), so that
// hx = d' x + beta s' Y.  The low-precision modes carry x and the Clenshaw
// intermediates in FP32, so FP32 factors add nothing to their error.
__global__ void diag_pairs_kernel(const double* sqrtr, const double* dfull, const double* dtda, long ncv,
                                  float2* full, float2* tda)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < ncv; i += (long)gridDim.x * blockDim.x) {
        full[i] = float2{(float)(dfull[i] * sqrtr[i]), (float)sqrtr[i]};
        tda[i] = float2{(float)dtda[i], 1.0f};
    }
}

// Maximum of one value per thread over a 256-thread workgroup (all threads).
__device__ __forceinline__ float block_max256(float m)
{
    __shared__ float part[4];
    m = fmaxf(m, __shfl_xor(m, 32));
    m = fmaxf(m, __shfl_xor(m, 16));
    m = fmaxf(m, __shfl_xor(m, 8));
    m = fmaxf(m, __shfl_xor(m, 4));
    m = fmaxf(m, __shfl_xor(m, 2));
    m = fmaxf(m, __shfl_xor(m, 1));
    __syncthreads();
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = m;
    __syncthreads();
    return fmaxf(fmaxf(part[0], part[1]), fmaxf(part[2], part[3]));
}

// xsig(v + nv c) = sigma_c sqrtR(v, c) (full) and sigma_c (TDA, second half).
__global__ void xsig_kernel(const double* sqrtr, const float* sig, int nv, long ncv, float* xsig)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < ncv; i += (long)gridDim.x * blockDim.x) {
        const double sc = (double)sig[nv + (int)(i / nv)];
        xsig[i] = (float)(sqrtr[i] * sc);
        xsig[ncv + i] = (float)sc;
    }
}

// vsc(v) = power of two >= max_mu |Psi_v(mu)| (one workgroup per band).
__global__ __launch_bounds__(256) void psiv_scale_kernel(const float* psiv32, int ldm1, int nmu, float* vsc)
{
    const float* p = psiv32 + (long)blockIdx.x * ldm1;
    float m = 0.0f;
    for (int mu = threadIdx.x; mu < nmu; mu += 256) m = fmaxf(m, fabsf(p[mu]));
    m = block_max256(m);
    if (threadIdx.x == 0) vsc[blockIdx.x] = pow2_ceil(m);
}

static int padk(int n) { return (n + LP_KPAD - 1) / LP_KPAD * LP_KPAD; }

int lp2_prepare(ngw_context* ctx)
{
    if (ctx->lp2_ready) return NGW_OK;
    NGW_TRY(lowp_prepare(ctx));  // per-band scales
    const int nmu = ctx->n_intp, nv = ctx->nval, nc = ctx->ncond;
    ctx->ldc1 = padk(nc);
    ctx->ldm1 = padk(nmu);
    const size_t n1 = (size_t)(nmu + 1) * ctx->ldc1, n2 = (size_t)(nc + 1) * ctx->ldm1;  // + zero rows
    NGW_TRY(ctx->pct32.ensure(4 * n1));
    NGW_TRY(ctx->pc32.ensure(4 * n2));
    NGW_TRY(ctx->pctb.ensure(2 * 2 * n1));  // hi then lo
    NGW_TRY(ctx->pcb.ensure(2 * 2 * n2));
    NGW_TRY(ctx->pcth.ensure(2 * n1));
    NGW_TRY(ctx->pch.ensure(2 * n2));
    NGW_TRY(ctx->psiv_scale.ensure(sizeof(float) * nv));
    NGW_TRY(ctx->xsig.ensure(sizeof(float) * 2 * (size_t)ctx->ncv));
    NGW_TRY(ctx->m32t.ensure(4 * (size_t)(nmu + 1) * ctx->ldm1));
    NGW_TRY(ctx->psiv32.ensure(4 * (size_t)nv * ctx->ldm1));
    NGW_TRY(ctx->dpairs.ensure(2 * sizeof(float2) * (size_t)ctx->ncv));
    hipStream_t s = ctx->stream;
    const double* psi = ctx->psi.as<double>();
    const float* sig = ctx->psi_scale.as<float>();
    unsigned short* pctb = ctx->pctb.as<unsigned short>();
    unsigned short* pcb = ctx->pcb.as<unsigned short>();
    hipLaunchKernelGGL(prep_psi_kernel<LP_F32>, dim3(2048), dim3(256), 0, s, psi, (long)ctx->ld_psi, nmu, nv, nc, sig,
                       ctx->ldc1, ctx->ldm1, (void*)ctx->pct32.ptr, (void*)nullptr, (void*)ctx->pc32.ptr, (void*)nullptr);
    hipLaunchKernelGGL(prep_psi_kernel<LP_BF16X3>, dim3(2048), dim3(256), 0, s, psi, (long)ctx->ld_psi, nmu, nv, nc,
                       sig, ctx->ldc1, ctx->ldm1, (void*)pctb, (void*)(pctb + n1), (void*)pcb, (void*)(pcb + n2));
    hipLaunchKernelGGL(prep_psi_kernel<LP_F16>, dim3(2048), dim3(256), 0, s, psi, (long)ctx->ld_psi, nmu, nv, nc, sig,
                       ctx->ldc1, ctx->ldm1, (void*)ctx->pcth.ptr, (void*)nullptr, (void*)ctx->pch.ptr, (void*)nullptr);
    hipLaunchKernelGGL(prep_m_kernel, dim3(2048), dim3(256), 0, s, ctx->mmat.as<double>(), (long)ctx->ld_m, nmu,
                       ctx->ldm1, psi, (long)ctx->ld_psi, nv, ctx->m32t.as<float>(), ctx->psiv32.as<float>());
    hipLaunchKernelGGL(psiv_scale_kernel, dim3(nv), dim3(256), 0, s, ctx->psiv32.as<float>(), ctx->ldm1, nmu,
                       ctx->psiv_scale.as<float>());
    hipLaunchKernelGGL(xsig_kernel, dim3(1024), dim3(256), 0, s, ctx->sqrtr.as<double>(), sig, nv, ctx->ncv,
                       ctx->xsig.as<float>());
    hipLaunchKernelGGL(diag_pairs_kernel, dim3(1024), dim3(256), 0, s, ctx->sqrtr.as<double>(),
                       ctx->diag_full.as<double>(), ctx->diag_tda.as<double>(), ctx->ncv, ctx->dpairs.as<float2>(),
                       ctx->dpairs.as<float2>() + ctx->ncv);
    NGW_HIP(hipGetLastError());
    NGW_HIP(hipStreamSynchronize(s));
    ctx->lp2_ready = true;
    return NGW_OK;
}

#define LP_PICK_OP()                                                     \
    const int r = find_pre(pre, nreq, blockIdx.x);                       \
    const LpOp op = ops ? ops[r] : op1;                                  \
    const int local = pre ? (int)blockIdx.x - pre[r] : (int)blockIdx.x;

// GEMM kernels: the workgroup index is first mapped so that each XCD owns a
// contiguous range of tiles (M fastest inside a request), then to a request.
#define LP_PICK_TILE()                                                   \
    const int bid = xcd_tile((int)blockIdx.x, (int)gridDim.x);           \
    const int r = find_pre(pre, nreq, bid);                              \
    const LpOp op = ops ? ops[r] : op1;                                  \
    const int local = pre ? bid - pre[r] : bid;

// ---------------------------------------------------------------- conversion passes
// FP16: partial maxima of |x'| = |xsig(v,c) x_b(v,c)| over XS_SPLIT ranges
// of each vector (XS_SPLIT workgroups per vector, eight loads in flight per
// thread); xprep turns them into the vector's scale xsc(b).
constexpr int XS_SPLIT = 8;
// workgroups per vector: about one per 16384 elements, at most XS_SPLIT
__host__ __device__ inline int xs_split(long ncv)
{
    const long k = (ncv + 16383) / 16384;
    return k < 1 ? 1 : (k > XS_SPLIT ? XS_SPLIT : (int)k);
}
__global__ __launch_bounds__(256) void xscale_kernel(const LpOp* ops, LpOp op1, const LpStage* st, LpStage st1,
                                                     int nstage, int s, const int* pre, int nreq)
{
    LP_PICK_OP();
    const long ncv = (long)op.nv * op.nc;
    const int split = xs_split(ncv);
    const int b = local / split, part = local - b * split;
    const LpStage& sg = ops ? st[r * nstage + s] : st1;
    const bool x32 = sg.flags & ST_X32;
    const long xoff = (long)b * op.ld;
    const long chunk = ((ncv + split - 1) / split + 3) & ~3L;  // a multiple of four elements
    const long lo = part * chunk, hi = min(ncv, lo + chunk);
    double m = 0.0;
    auto scan = [&](auto* xp) {
        for (long base = lo + threadIdx.x; base < hi; base += 2048) {
            double v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const long idx = base + 256 * u;
                v[u] = idx < hi ? fabs((double)xp[xoff + idx]) * (double)op.xsig[idx] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) m = fmax(m, v[u]);
        }
    };
    // FP32 input with 16-byte aligned starts: four elements per load
    auto scan4 = [&](const float* xp) {
        const float4* x4 = reinterpret_cast<const float4*>(xp + xoff + lo);
        const float4* s4 = reinterpret_cast<const float4*>(op.xsig + lo);
        const long nq = hi > lo ? (hi - lo) >> 2 : 0;
        for (long q0 = threadIdx.x; q0 < nq; q0 += 256 * 4) {
            float4 xv[4], sv[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const long q = q0 + 256 * u;
                if (q < nq) {
                    xv[u] = x4[q];
                    sv[u] = s4[q];
                } else {
                    xv[u] = float4{0.0f, 0.0f, 0.0f, 0.0f};
                    sv[u] = xv[u];
                }
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                m = fmax(m, fabs((double)xv[u].x) * (double)sv[u].x);
                m = fmax(m, fabs((double)xv[u].y) * (double)sv[u].y);
                m = fmax(m, fabs((double)xv[u].z) * (double)sv[u].z);
                m = fmax(m, fabs((double)xv[u].w) * (double)sv[u].w);
            }
        }
        for (long idx = lo + 4 * nq + threadIdx.x; idx < hi; idx += 256)
            m = fmax(m, fabs((double)xp[xoff + idx]) * (double)op.xsig[idx]);
    };
    if (x32) {
        const float* xf = reinterpret_cast<const float*>(sg.x);
        if ((((xoff + lo) & 3) == 0) && ((reinterpret_cast<uintptr_t>(xf) & 15) == 0) &&
            ((reinterpret_cast<uintptr_t>(op.xsig) & 15) == 0))
            scan4(xf);
        else
            scan(xf);
    } else {
        scan(sg.x);
    }
    const float mb = block_max256((float)m);
    if (threadIdx.x == 0) op.xpart[b * XS_SPLIT + part] = mb;
}

// x'(j = v + nv b, c) = xsig(v,c) x_b(v,c), xsig = sigma_c s(v,c) (sigma_c
// for TDA).  A workgroup converts one vector's block of VT = min(nv, 128)
// consecutive v by CT = 64 k consecutive c, k the largest power of two whose
// LDS tile CT (VT + 1) fits XP_TILE floats (and CT <= ldc1): 4-8k elements
// whatever nv is.  For
// nv <= 128 the block is one contiguous run of x, read flat and fully
// coalesced (FP32 input four elements per load); it is transposed through
// LDS and each thread writes eight consecutive c in one 16- or 32-byte store.
constexpr int XP_V = 128, XP_TILE = 10240;
__host__ __device__ inline int xp_vt(int nv) { return nv < XP_V ? nv : XP_V; }
__host__ __device__ inline int xp_ct(int nv, int ldc1)
{
    int k = 1;
    while (2 * k * 64 * (xp_vt(nv) + 1) <= XP_TILE && 2 * k * 64 <= ldc1) k *= 2;
    return 64 * k;
}
template <int KIND>
__global__ __launch_bounds__(256) void xprep_kernel(const LpOp* ops, LpOp op1, const LpStage* st, LpStage st1,
                                                    int nstage, int s, const int* pre, int nreq)
{
    typedef typename V2Traits<KIND>::elem E;
    __shared__ float tile[XP_TILE];  // [c][v], pitch vt + 1
    LP_PICK_OP();
    const LpStage& sg = ops ? st[r * nstage + s] : st1;
    const bool x32 = sg.flags & ST_X32;
    const bool wq32 = !x32 && (sg.flags & ST_QW32) && op.q32;  // x is q: keep an FP32 copy
    const int vt = xp_vt(op.nv), ct = xp_ct(op.nv, op.ldc1);
    const int nvt = (op.nv + vt - 1) / vt, nct = (op.ldc1 + ct - 1) / ct;
    const int b = local / (nvt * nct);
    const int rem = local - b * nvt * nct;
    const int v0 = (rem / nct) * vt, c0 = (rem % nct) * ct;
    const int vn = min(vt, op.nv - v0), cn = max(0, min(ct, op.nc - c0));
    const int pitch = vt + 1;
    if (local == 0)  // the zero row after the last x' row
        for (int c = threadIdx.x; c < op.ldc1; c += 256)
            put_v2<KIND>((E*)op.xlp_hi, (E*)op.xlp_lo, (long)op.nv * op.nvec * op.ldc1 + c, 0.0f);
    const long xoff = (long)b * op.ld;
    double f16s = 1.0;
    if constexpr (KIND == LP_F16) {
        float m = 0.0f;
        const int split = xs_split((long)op.nv * op.nc);
        for (int k = 0; k < split; ++k) m = fmaxf(m, op.xpart[b * XS_SPLIT + k]);
        const float xsc = pow2_ceil(m);
        f16s = (double)F16_VEC / (double)xsc;
        if (rem == 0 && threadIdx.x == 0) op.xsc[b] = xsc;  // for K1's epilogue
    }
    // flat walk e = v + vn c over the block; (c, v) from a float reciprocal
    // with an exact correction
    const int tot = vn * cn;
    const long base = v0 + (long)c0 * op.nv;
    const float inv = 1.0f / (float)vn;
    auto walk = [&](auto* xp) {
        for (int e0 = threadIdx.x; e0 < tot; e0 += 256 * 8) {
            double xv[8], sv[8];
            int cc[8], vv[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int e = e0 + 256 * u;
                int c = (int)((float)e * inv);
                if (c * vn > e) --c;
                if ((c + 1) * vn <= e) ++c;
                cc[u] = c;
                vv[u] = e - c * vn;
                const long idx = base + vv[u] + (long)c * op.nv;
                xv[u] = e < tot ? (double)xp[xoff + idx] : 0.0;
                sv[u] = e < tot ? (double)op.xsig[idx] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (e0 + 256 * u < tot) tile[cc[u] * pitch + vv[u]] = (float)(xv[u] * (sv[u] * f16s));
            if (wq32) {
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    if (e0 + 256 * u < tot)
                        op.q32[(long)b * op.nv * op.nc + base + vv[u] + (long)cc[u] * op.nv] = (float)xv[u];
            }
        }
    };
    // FP32 input over whole v columns (vn == nv) is one contiguous run of x
    // and of xsig: four elements per 16-byte load when both starts are
    // 16-byte aligned; the last tot mod 4 elements go element-wise.
    auto walk4 = [&](const float* xp) {
        const float4* x4 = reinterpret_cast<const float4*>(xp + xoff + base);
        const float4* s4 = reinterpret_cast<const float4*>(op.xsig + base);
        const int nq = tot >> 2;
        for (int q0 = threadIdx.x; q0 < nq; q0 += 256 * 4) {
            float4 xv[4], sv[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int q = q0 + 256 * u;
                if (q < nq) {
                    xv[u] = x4[q];
                    sv[u] = s4[q];
                }
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int q = q0 + 256 * u;
                if (q >= nq) continue;
                const float xa[4] = {xv[u].x, xv[u].y, xv[u].z, xv[u].w};
                const float sa[4] = {sv[u].x, sv[u].y, sv[u].z, sv[u].w};
                const int e = 4 * q;
                int c = (int)((float)e * inv);
                if (c * vn > e) --c;
                if ((c + 1) * vn <= e) ++c;
                int v = e - c * vn;
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    tile[c * pitch + v] = (float)((double)xa[k] * ((double)sa[k] * f16s));
                    if (++v == vn) {
                        v = 0;
                        ++c;
                    }
                }
            }
        }
        for (int e = 4 * nq + threadIdx.x; e < tot; e += 256) {
            const int c = e / vn, v = e - c * vn;
            tile[c * pitch + v] = (float)((double)xp[xoff + base + e] * ((double)op.xsig[base + e] * f16s));
        }
    };
    if (x32) {
        const float* xf = reinterpret_cast<const float*>(sg.x);
        const bool aligned = (((xoff + base) & 3) == 0) && ((base & 3) == 0) &&
                             ((reinterpret_cast<uintptr_t>(xf) & 15) == 0) &&
                             ((reinterpret_cast<uintptr_t>(op.xsig) & 15) == 0);
        if (vn == op.nv && aligned) walk4(xf);
        else walk(xf);
    } else {
        walk(sg.x);
    }
    for (int i = cn * pitch + threadIdx.x; i < ct * pitch; i += 256) tile[i] = 0.0f;  // k padding rows
    __syncthreads();
    const int cpr = ct / 8;  // 8-wide chunks per row (a power of two)
    const int sh = __ffs(cpr) - 1;
    const int lim = min(ct, op.ldc1 - c0);
    for (int t = threadIdx.x; t < vn * cpr; t += 256) {
        const int vl = t >> sh, cq = (t & (cpr - 1)) * 8;
        if (cq >= lim) continue;
        float w[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) w[e] = tile[(cq + e) * pitch + vl];
        const long o = (long)(v0 + vl + op.nv * b) * op.ldc1 + c0 + cq;
        if constexpr (KIND == LP_F32) {
            f32x4* out = reinterpret_cast<f32x4*>(reinterpret_cast<float*>(op.xlp_hi) + o);
            out[0] = f32x4{w[0], w[1], w[2], w[3]};
            out[1] = f32x4{w[4], w[5], w[6], w[7]};
        } else {
            unsigned hi[4], lo[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const unsigned short ha = to_e16<KIND>(w[2 * e]), hb = to_e16<KIND>(w[2 * e + 1]);
                hi[e] = (unsigned)ha | ((unsigned)hb << 16);
                if constexpr (V2Traits<KIND>::NSPLIT == 2) {
                    const unsigned short la = f32_to_bf16_bits(w[2 * e] - bf16_bits_to_f32(ha));
                    const unsigned short lb = f32_to_bf16_bits(w[2 * e + 1] - bf16_bits_to_f32(hb));
                    lo[e] = (unsigned)la | ((unsigned)lb << 16);
                }
            }
            *reinterpret_cast<u32x4*>(reinterpret_cast<E*>(op.xlp_hi) + o) = u32x4{hi[0], hi[1], hi[2], hi[3]};
            if constexpr (V2Traits<KIND>::NSPLIT == 2)
                *reinterpret_cast<u32x4*>(reinterpret_cast<E*>(op.xlp_lo) + o) = u32x4{lo[0], lo[1], lo[2], lo[3]};
        }
    }
}

// t32(b, nu) = sum of the K1 slots covering vector b (fixed order); zero
// padding and zero row b = nvec (also written into u, whose rows b < nvec
// come from K2).
__global__ __launch_bounds__(256) void red2_kernel(const LpOp* ops, LpOp op1, const int* pre, int nreq, int bn)
{
    LP_PICK_OP();
    const long i = (long)local * 256 + threadIdx.x;
    if (i >= (long)(op.nvec + 1) * op.ldm1) return;
    const int b = (int)(i / op.ldm1), nu = (int)(i - (long)b * op.ldm1);
    float out = 0.0f;
    if (nu < op.nmu && b < op.nvec) {
        const int first = (b * op.nv) / bn, last = ((b + 1) * op.nv - 1) / bn;
        double acc = 0.0;
        for (int sl = 0; sl <= last - first; ++sl) acc += op.tpart[((long)sl * op.nvec + b) * op.nmu + nu];
        out = (float)acc;
    }
    op.t32[i] = out;
    if (b == op.nvec) op.u32[i] = 0.0f;
    if (op.umax && nu == 0 && b < op.nvec) op.umax[b] = 0u;  // K2 takes the maximum
}

// ---------------------------------------------------------------- GEMMs
// LDS of a K1 / K3 tile: SW 0 / 1 register-prefetched slabs (SmemV2, two
// buffers, D slabs in flight), SW 2 direct-to-LDS slabs in D buffers with the
// SWZ 1 row layout (SmemDma).
template <int KIND, int BM, int BN, int D, int SW>
using LpSmem = std::conditional_t<SW == 2, SmemDma<KIND, BM, BN, D>, SmemV2<KIND, BM, BN, SW == 2 ? 1 : SW>>;

template <int KIND, int BM, int BN, int D, int SW>
__device__ __forceinline__ void lp_gemm(const OperandV2<KIND>& A, const OperandV2<KIND>& B, int m0, int n0, int Kpad,
                                        AccLp<BM, BN>& acc, LpSmem<KIND, BM, BN, D, SW>& sm)
{
    if constexpr (SW == 2) gemm_dma_tile<KIND, BM, BN, D>(A, B, m0, n0, Kpad, acc, sm);
    else gemm_v2_tile<KIND, BM, BN, D>(A, B, m0, n0, Kpad, acc, sm);
}

template <int KIND, int BM, int BN, int D, int PM, int SW>
__global__ __launch_bounds__(GNT) void k1v2_kernel(const LpOp* ops, LpOp op1, const int* pre, int nreq)
{
    typedef typename V2Traits<KIND>::elem E;
    constexpr int HALF = BM < 64 ? BM : 64;  // rows reduced per LDS pass (keeps the buffer within the slabs' LDS)
    __shared__ union {
        LpSmem<KIND, BM, BN, D, SW> ab;
        float red[HALF][BN + 1];
    } sm;
    LP_PICK_OP();
    const int nv = op.nv, N = nv * op.nvec;
    int tm, tn;
    if (!xcd_rect_tile(local, (op.nmu + BM - 1) / BM, (N + BN - 1) / BN, PM, tm, tn)) return;
    const int m0 = tm * BM, n0 = tn * BN;
    AccLp<BM, BN> acc;
    lp_gemm<KIND, BM, BN, D, SW>(OperandV2<KIND>{(const E*)op.pct_hi, (const E*)op.pct_lo, op.ldc1, op.nmu},
                                 OperandV2<KIND>{(const E*)op.xlp_hi, (const E*)op.xlp_lo, op.ldc1, N}, m0, n0,
                                 op.ldc1, acc, sm.ab);
    const int nend = min(n0 + BN, N);
    const int b_lo = n0 / nv, b_hi = (nend - 1) / nv;
    const int nbt = b_hi - b_lo + 1;
    // T(mu, j) Psi_v(mu, v(j)): a lane's accumulator fragment holds four
    // consecutive mu of one column j, so Psi_v comes as one 16-byte load from
    // the [v][mu] copy (zero for nmu <= mu < ldm1); v(j) once per column
    constexpr int FM = BM / 32, FN = BN / 32;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = (wave & 1) * (BM / 2), wn = (wave >> 1) * (BN / 2);
    int vrow[FN];
    bool cok[FN];
#pragma unroll
    for (int j = 0; j < FN; ++j) {
        const int gn = n0 + wn + 16 * j + (lane & 15);
        cok[j] = gn < N;
        vrow[j] = cok[j] ? gn % nv : 0;
    }
#pragma unroll
    for (int pass = 0; pass < BM / HALF; ++pass) {
        if (pass > 0) __syncthreads();
#pragma unroll
        for (int i = 0; i < FM; ++i) {
            const int row0 = wm + 16 * i + 4 * (lane >> 4);  // a fragment lies in one pass
            if (row0 < pass * HALF || row0 >= (pass + 1) * HALF) continue;
            const int gm0 = m0 + row0;
#pragma unroll
            for (int j = 0; j < FN; ++j) {
                f32x4 pv = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
                if (cok[j] && gm0 < op.ldm1)
                    pv = *reinterpret_cast<const f32x4*>(op.psiv32 + (long)vrow[j] * op.ldm1 + gm0);
                const int col = wn + 16 * j + (lane & 15);
#pragma unroll
                for (int r = 0; r < 4; ++r) sm.red[row0 + r - pass * HALF][col] = acc.v[i][j][r] * pv[r];
            }
        }
        __syncthreads();
        for (int e = threadIdx.x; e < HALF * nbt; e += GNT) {
            int row = e % HALF, b = b_lo + e / HALF;
            int c0 = max(n0, b * nv) - n0, c1 = min(nend, (b + 1) * nv) - n0;
            float sum = 0.0f;
            for (int cc = c0; cc < c1; ++cc) sum += sm.red[row][cc];
            int gm = m0 + pass * HALF + row;
            int slot = tn - (b * nv) / BN;
            double t = (double)sum;
            if constexpr (KIND == LP_F16) t *= (double)op.xsc[b] * (double)(1.0f / F16_VEC) * (double)(1.0f / F16_OPS);
            if (gm < op.nmu) op.tpart[((long)slot * op.nvec + b) * op.nmu + gm] = t;
        }
    }
}

template <bool UMAX>
__global__ __launch_bounds__(GNT) void k2v2_kernel(const LpOp* ops, LpOp op1, const int* pre, int nreq)
{
    __shared__ SmemV2<LP_F32, L2K2, L2K2, 1> sm;
    LP_PICK_TILE();
    const int mt = (op.nmu + L2K2 - 1) / L2K2;
    const int m0 = (local % mt) * L2K2, n0 = (local / mt) * L2K2;
    AccLp<L2K2, L2K2> acc;
    gemm_v2_tile<LP_F32, L2K2, L2K2, 1>(OperandV2<LP_F32>{op.m32t, nullptr, op.ldm1, op.nmu},
                                     OperandV2<LP_F32>{op.t32, nullptr, op.ldm1, op.nvec}, m0, n0, op.ldm1, acc, sm);
    // rows nmu <= mu < ldm1 come out as exact zeros (zero A rows): the K3
    // loader reads u over the whole padded slab
    for_acc_lp<L2K2, L2K2>(acc, [&](int row, int col, float val) {
        int gm = m0 + row, gn = n0 + col;
        if (gm < op.ldm1 && gn < op.nvec) op.u32[gm + (long)gn * op.ldm1] = val;
    });
    if constexpr (UMAX) {
        // umax(b) = max_mu |u(mu, b)|: per column over the wave's rows (lanes
        // l, l + 16, l + 32, l + 48 hold one column), then one atomic per column
        constexpr int FM = L2K2 / 32, FN = L2K2 / 32;
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        const int wn = (wave >> 1) * (L2K2 / 2);
#pragma unroll
        for (int j = 0; j < FN; ++j) {
            float m = 0.0f;
#pragma unroll
            for (int i = 0; i < FM; ++i)
#pragma unroll
                for (int q = 0; q < 4; ++q) m = fmaxf(m, fabsf(acc.v[i][j][q]));
            m = fmaxf(m, __shfl_xor(m, 16));
            m = fmaxf(m, __shfl_xor(m, 32));
            const int gn = n0 + wn + 16 * j + (lane & 15);
            if (lane < 16 && gn < op.nvec) atomicMax(op.umax + gn, __float_as_uint(m));
        }
    }
}

// W(j = v + nv b, mu) = Psi_v(mu, v) u(mu, b) for rows j <= N (row N zero),
// in chunks of eight consecutive mu from the FP32 psiv32 / u rows.  A
// workgroup takes one vector b, WP_V consecutive v and 256 chunks of mu: each
// thread loads its u chunk once and writes it into the WP_V rows v + nv b
// (Psi_v rows come from L2), so every store instruction of a wave writes 1 KB
// of one row.  Row N is the zero row: b = nvec, v = 0, u row nvec is zero.
constexpr int WP_V = 8;
__host__ __device__ constexpr int wp_grid(int nv, int nvec, int ldm1)
{
    return (nvec + 1) * ((nv + WP_V - 1) / WP_V) * ((ldm1 / 8 + 255) / 256);
}
template <int KIND>
__global__ __launch_bounds__(256) void wprep_kernel(const LpOp* ops, LpOp op1, const int* pre, int nreq)
{
    typedef typename V2Traits<KIND>::elem E;
    LP_PICK_OP();
    const int per_row = op.ldm1 / 8, nmc = (per_row + 255) / 256, nvg = (op.nv + WP_V - 1) / WP_V;
    const int mc = local % nmc, rest = local / nmc;
    const int vg = rest % nvg, b = rest / nvg;
    const int chunk = mc * 256 + threadIdx.x;
    const int v0 = vg * WP_V;
    const bool zero_row = b == op.nvec;  // only its v = 0 row exists (j = N)
    if (chunk >= per_row || b > op.nvec || (zero_row && v0 > 0)) return;
    const int mu = chunk * 8;
    const float* u = op.u32 + (long)b * op.ldm1 + mu;
    const f32x4 u0 = gload<f32x4>(u), u1 = gload<f32x4>(u + 4);
    float usc = 1.0f;
    if constexpr (KIND == LP_F16) usc = zero_row ? 1.0f : pow2_ceil(__uint_as_float(op.umax[b]));
    const int vend = zero_row ? 1 : min(WP_V, op.nv - v0);
#pragma unroll
    for (int dv = 0; dv < WP_V; ++dv) {
        if (dv >= vend) break;
        const int v = v0 + dv, j = v + op.nv * b;
        const float* p = op.psiv32 + (long)v * op.ldm1 + mu;
        f32x4 w0 = gload<f32x4>(p) * u0;
        f32x4 w1 = gload<f32x4>(p + 4) * u1;
        if constexpr (KIND == LP_F16) {
            const float sc = zero_row ? 1.0f : op.vsc[v] * usc;
            const float f = F16_VEC / sc;
            w0 = w0 * f;
            w1 = w1 * f;
            if (mu == 0 && !zero_row) op.wsc[j] = sc;
        }
        const long o = (long)j * op.ldm1 + mu;
        if constexpr (KIND == LP_F32) {
            f32x4* out = reinterpret_cast<f32x4*>(reinterpret_cast<float*>(op.wlp_hi) + o);
            out[0] = w0;
            out[1] = w1;
        } else {
            const float w[8] = {w0.x, w0.y, w0.z, w0.w, w1.x, w1.y, w1.z, w1.w};
            unsigned hi[4], lo[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const unsigned short ha = to_e16<KIND>(w[2 * e]), hb = to_e16<KIND>(w[2 * e + 1]);
                hi[e] = (unsigned)ha | ((unsigned)hb << 16);
                if constexpr (V2Traits<KIND>::NSPLIT == 2) {
                    const unsigned short la = f32_to_bf16_bits(w[2 * e] - bf16_bits_to_f32(ha));
                    const unsigned short lb = f32_to_bf16_bits(w[2 * e + 1] - bf16_bits_to_f32(hb));
                    lo[e] = (unsigned)la | ((unsigned)lb << 16);
                }
            }
            *reinterpret_cast<u32x4*>(reinterpret_cast<E*>(op.wlp_hi) + o) = u32x4{hi[0], hi[1], hi[2], hi[3]};
            if constexpr (V2Traits<KIND>::NSPLIT == 2)
                *reinterpret_cast<u32x4*>(reinterpret_cast<E*>(op.wlp_lo) + o) = u32x4{lo[0], lo[1], lo[2], lo[3]};
        }
    }
}

// 16-byte accesses of four consecutive elements (FP32 or FP64), widened to
// or narrowed from FP64.
__device__ __forceinline__ bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; }
__device__ __forceinline__ void ld4(const float* p, double (&o)[4])
{
    const f32x4 v = *reinterpret_cast<const f32x4*>(p);
    o[0] = (double)v.x; o[1] = (double)v.y; o[2] = (double)v.z; o[3] = (double)v.w;
}
__device__ __forceinline__ void ld4(const double* p, double (&o)[4])
{
    const double2 a = reinterpret_cast<const double2*>(p)[0], b = reinterpret_cast<const double2*>(p)[1];
    o[0] = a.x; o[1] = a.y; o[2] = b.x; o[3] = b.y;
}
__device__ __forceinline__ void st4(float* p, const double (&r)[4])
{
    *reinterpret_cast<f32x4*>(p) = f32x4{(float)r[0], (float)r[1], (float)r[2], (float)r[3]};
}
__device__ __forceinline__ void st4(double* p, const double (&r)[4])
{
    reinterpret_cast<double2*>(p)[0] = double2{r[0], r[1]};
    reinterpret_cast<double2*>(p)[1] = double2{r[2], r[3]};
}

// 128 x 128 tiles keep three workgroups per CU: the register budget is 168
// VGPRs (the swizzled-LDS variant otherwise lands at 170, two waves).
template <int KIND, int BM, int BN, int D, int PM, int FW, int SW>
__global__ __launch_bounds__(GNT) __attribute__((amdgpu_waves_per_eu(BM * BN <= 128 * 128 && D <= 2 && FW == 0 && KIND != LP_BF16X3 ? 3 : 1)))
void k3v2_kernel(const LpOp* ops, LpOp op1, const LpStage* st, LpStage st1,
                                                   int nstage, int s, const int* pre, int nreq)
{
    typedef typename V2Traits<KIND>::elem E;
    __shared__ LpSmem<KIND, BM, BN, D, SW> sm;
    LP_PICK_OP();
    const LpStage sg = ops ? st[r * nstage + s] : st1;
    const int nv = op.nv, N = nv * op.nvec;
    int tm, tn;
    if (!xcd_rect_tile(local, (op.nc + BM - 1) / BM, (N + BN - 1) / BN, PM, tm, tn)) return;
    const int m0 = tm * BM, n0 = tn * BN;
    AccLp<BM, BN> acc;
    if constexpr (FW) {
        SlabV2<KIND, BM> la;
        SlabW<KIND, BN> lb;
        la.init(OperandV2<KIND>{(const E*)op.pc_hi, (const E*)op.pc_lo, op.ldm1, op.nc}, m0);
        lb.init(op.psiv32, op.u32, op.ldm1, nv, N, op.nvec, n0);
        gemm_v2_run<KIND, BM, BN, D>(la, lb, op.ldm1, acc, sm);
    } else {
        lp_gemm<KIND, BM, BN, D, SW>(OperandV2<KIND>{(const E*)op.pc_hi, (const E*)op.pc_lo, op.ldm1, op.nc},
                                     OperandV2<KIND>{(const E*)op.wlp_hi, (const E*)op.wlp_lo, op.ldm1, N}, m0,
                                     n0, op.ldm1, acc, sm);
    }
    // epilogue: OperatorEpilogue's stage (diagonal term, Horner / Clenshaw
    // combination, Lanczos accumulate), written out for one leading
    // dimension shared by every vector block of the request
    const float2* dp = op.dpair;
    const double beta = op.tda ? 2.0 : 4.0;
    const bool acc_on = sg.accumulate == 1 || (sg.accumulate == 2 && *op.kstep > 1);
    const bool use_q = op.q != nullptr && sg.gamma != 0.0;
    // FP32 output stages read the first stage's FP32 copy of q
    const float* q32 = (sg.flags & ST_Q32) && (sg.flags & ST_O32) ? op.q32 : nullptr;
    const long ncv = (long)nv * op.nc;
    const float* sigc = op.sig + nv;
    // element types of x, y, out fixed at compile time per flag combination
    auto epi = [&](auto xt, auto yt, auto ot) {
        typedef decltype(xt) TX;
        typedef decltype(yt) TY;
        typedef decltype(ot) TO;
        const TX* xp = reinterpret_cast<const TX*>(sg.x);
        const TY* yp = reinterpret_cast<const TY*>(sg.y);
        TO* op_ = reinterpret_cast<TO*>(sg.out);
        // four consecutive pairs icv = v..v+3 of one c share a b when nv is a
        // multiple of four (tiles start at multiples of four): then every
        // array is read and written with 16-byte accesses
        const bool vec = (nv & 3) == 0 && (op.ld & 3) == 0 && aligned16(xp) && aligned16(yp) && aligned16(op_) &&
                         aligned16(use_q ? op.q : nullptr) && aligned16(q32) && aligned16(dp) &&
                         (KIND != LP_F16 || aligned16(op.wsc)) && (ncv & 3) == 0;
        // The tile passes through LDS one accumulator fragment row at a time:
        // in pass q every wave writes its fragment row q (16 tile rows, so
        // 32 rows in all: local rows 0-15 from the upper half of the tile,
        // 16-31 from the lower), then every thread takes runs of four
        // consecutive columns of one row (a wave covers whole rows, so each
        // access of the wave is one contiguous run of x, y, q and out per
        // pair block), or single elements where the arrays do not allow
        // 16-byte accesses.  The fragment index is a compile-time constant,
        // so a pass frees its accumulators.  Pitch = BN + 16 floats: the 16
        // lanes of an accumulator row and the next row's 16 fall on disjoint
        // banks.
        constexpr int FM = BM / 32, EPR = 32, PITCH = BN + 16, G = BN / 4, PER = EPR * G / GNT;
        static_assert(EPR * G % GNT == 0, "epilogue row blocks");
        static_assert(sizeof(sm) >= sizeof(float) * EPR * PITCH, "epilogue staging exceeds the slab LDS");
        float(*stg)[PITCH] = reinterpret_cast<float(*)[PITCH]>(&sm);
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        const int hrow = (wave & 1) * 16, wn = (wave >> 1) * (BN / 2);
        static_for<FM>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            if (q > 0) __syncthreads();
#pragma unroll
            for (int jf = 0; jf < BN / 32; ++jf)
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    stg[hrow + 4 * (lane >> 4) + r][wn + 16 * jf + (lane & 15)] = acc.v[q][jf][r];
            __syncthreads();
            if (!vec) {
#pragma unroll 1
                for (int e = threadIdx.x; e < EPR * BN; e += GNT) {
                    const int rl = e / BN, col = e - rl * BN;
                    const int c = m0 + (rl < 16 ? 16 * q + rl : BM / 2 + 16 * q + rl - 16), j = n0 + col;
                    if (c >= op.nc || j >= N) continue;
                    const int b = j / nv, v = j - b * nv;
                    const long idx = v + (long)c * nv, i = idx + (long)b * op.ld;
                    double Y = (double)stg[rl][col] * (double)sigc[c];
                    if constexpr (KIND == LP_F16) Y *= (double)op.wsc[j] * F16_UNSCALE;
                    const float2 f = dp[idx];
                    const double xv = (double)xp[i];
                    const double hx = (double)f.x * xv + (beta * (double)f.y) * Y;
                    double r = sg.alpha * hx + sg.bx * xv;
                    if (yp) r = r + sg.by * (double)yp[i];
                    if (use_q) r = r + sg.gamma * (q32 ? (double)q32[idx + (long)b * ncv] : op.q[i]);
                    if (acc_on) r = (double)op_[i] + r;
                    op_[i] = (TO)r;
                }
                return;
            }
            // two groups in flight for the 128 x 192 tile (two waves per SIMD,
            // with register room), one for 128 x 128 (three waves per SIMD)
            // and for 128 x 256 (its accumulators leave no room)
            constexpr int GUNROLL = BM * BN <= 128 * 128 || BM * BN > 128 * 192 ? 1 : 2;
#pragma unroll GUNROLL
            for (int k = 0; k < PER; ++k) {
                const int e = threadIdx.x + GNT * k;
                const int rl = e / G, cg = e - rl * G;
                const int c = m0 + (rl < 16 ? 16 * q + rl : BM / 2 + 16 * q + rl - 16), j0 = n0 + 4 * cg;
                if (c >= op.nc || j0 >= N) continue;
                const f32x4 v4 = *reinterpret_cast<const f32x4*>(&stg[rl][4 * cg]);
                const int b = j0 / nv, v0 = j0 - b * nv;
                const long idx = v0 + (long)c * nv, i = idx + (long)b * op.ld;
                // the stage's terms are added in the order of the element-wise
                // form above, each array loaded just before it is needed
                const double sc = (double)sigc[c];
                const float va[4] = {v4.x, v4.y, v4.z, v4.w};
                double res[4], xv[4], t4[4];
                {
                    float ws[4] = {1.0f, 1.0f, 1.0f, 1.0f};
                    if constexpr (KIND == LP_F16) {
                        const f32x4 w4 = *reinterpret_cast<const f32x4*>(op.wsc + j0);
                        ws[0] = w4.x; ws[1] = w4.y; ws[2] = w4.z; ws[3] = w4.w;
                    }
                    ld4(xp + i, xv);
                    const f32x4 d01 = *reinterpret_cast<const f32x4*>(dp + idx);
                    const f32x4 d23 = *reinterpret_cast<const f32x4*>(dp + idx + 2);
                    const float dd4[4] = {d01.x, d01.z, d23.x, d23.z}, ss4[4] = {d01.y, d01.w, d23.y, d23.w};
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        double Y = (double)va[u] * sc;
                        if constexpr (KIND == LP_F16) Y *= (double)ws[u] * F16_UNSCALE;
                        const double hx = (double)dd4[u] * xv[u] + (beta * (double)ss4[u]) * Y;
                        res[u] = sg.alpha * hx + sg.bx * xv[u];
                    }
                }
                if (yp) {
                    ld4(yp + i, t4);
#pragma unroll
                    for (int u = 0; u < 4; ++u) res[u] = res[u] + sg.by * t4[u];
                }
                if (use_q) {
                    if (q32) ld4(q32 + idx + (long)b * ncv, t4);
                    else ld4(op.q + i, t4);
#pragma unroll
                    for (int u = 0; u < 4; ++u) res[u] = res[u] + sg.gamma * t4[u];
                }
                if (acc_on) {
                    ld4(op_ + i, t4);
#pragma unroll
                    for (int u = 0; u < 4; ++u) res[u] = t4[u] + res[u];
                }
                st4(op_ + i, res);
            }
        });
    };
    switch (sg.flags & (ST_X32 | ST_Y32 | ST_O32)) {
        case 0: epi(double(), double(), double()); break;
        case ST_O32: epi(double(), double(), float()); break;
        case ST_X32 | ST_O32: epi(float(), double(), float()); break;
        case ST_X32 | ST_Y32 | ST_O32: epi(float(), float(), float()); break;
        case ST_X32 | ST_Y32: epi(float(), float(), double()); break;
        default: epi(float(), float(), double()); break;  // not produced by the stage planner
    }
}

// ---------------------------------------------------------------- launch plumbing
// Number of workgroups of each kernel for one request.
struct LpGrid {
    int xscale, xprep, k1, red, k2, wprep, k3;
};

static LpGrid lp_grid(int prec, int c1, int c3, int nmu, int nv, int nc, int nvec, int ldc1, int ldm1)
{
    const LpCfg k1 = K1_CFG[cfg_set_prec(prec)][c1], k3 = K3_CFG[cfg_set_prec(prec)][c3];
    LpGrid g;
    g.xscale = prec == PREC_F16 ? nvec * xs_split((long)nv * nc) : 0;
    g.xprep = nvec * ((nv + xp_vt(nv) - 1) / xp_vt(nv)) * ((ldc1 + xp_ct(nv, ldc1) - 1) / xp_ct(nv, ldc1));
    g.k1 = xcd_grid((nmu + k1.bm - 1) / k1.bm, (nv * nvec + k1.bn - 1) / k1.bn, k1.pm);
    g.red = (int)(((long)(nvec + 1) * ldm1 + 255) / 256);
    g.k2 = ((nmu + L2K2 - 1) / L2K2) * ((nvec + L2K2 - 1) / L2K2);
    g.wprep = wp_grid(nv, nvec, ldm1);
    g.k3 = xcd_grid((nc + k3.bm - 1) / k3.bm, (nv * nvec + k3.bn - 1) / k3.bn, k3.pm);
    return g;
}

template <int KIND, int C>
static void launch_k1(hipStream_t st, int nblk, const LpOp* ops, const LpOp& op1, const int* pre, int nreq)
{
    constexpr LpCfg c = K1_CFG[cfg_set(KIND)][C];
    hipLaunchKernelGGL((k1v2_kernel<KIND, c.bm, c.bn, c.d, c.pm, c.sw>), dim3(nblk), dim3(GNT), 0, st, ops, op1, pre,
                       nreq);
}

template <int KIND, int C>
static void launch_k3(hipStream_t st, int nblk, const LpOp* ops, const LpOp& op1, const LpStage* stg,
                      const LpStage& st1, int nstage, int s, const int* pre, int nreq)
{
    constexpr LpCfg c = K3_CFG[cfg_set(KIND)][C];
    hipLaunchKernelGGL((k3v2_kernel<KIND, c.bm, c.bn, c.d, c.pm, c.fw, c.sw>), dim3(nblk), dim3(GNT), 0, st, ops, op1,
                       stg, st1, nstage, s, pre, nreq);
}

template <int KIND>
static int lp_launch(hipStream_t st, int c1, int c3, const LpOp* ops, const LpOp& op1, const LpStage* stg,
                     const LpStage& st1, int nstage, int s, const int* const* pre, const int* tot, int nreq,
                     ngw_context* prof)
{
    if constexpr (KIND == LP_F16) {
        ProfScope ps(prof, KC_CONVERT);
        hipLaunchKernelGGL(xscale_kernel, dim3(tot[0]), dim3(256), 0, st, ops, op1, stg, st1, nstage, s, pre[0], nreq);
    }
    {
        ProfScope ps(prof, KC_CONVERT);
        hipLaunchKernelGGL(xprep_kernel<KIND>, dim3(tot[1]), dim3(256), 0, st, ops, op1, stg, st1, nstage, s, pre[1],
                           nreq);
    }
    {
        ProfScope ps(prof, KC_K1);
        switch (c1) {
            case 0: launch_k1<KIND, 0>(st, tot[2], ops, op1, pre[2], nreq); break;
            case 1: launch_k1<KIND, 1>(st, tot[2], ops, op1, pre[2], nreq); break;
            case 2: launch_k1<KIND, 2>(st, tot[2], ops, op1, pre[2], nreq); break;
            case 3: launch_k1<KIND, 3>(st, tot[2], ops, op1, pre[2], nreq); break;
            case 4: launch_k1<KIND, 4>(st, tot[2], ops, op1, pre[2], nreq); break;
            default: launch_k1<KIND, 5>(st, tot[2], ops, op1, pre[2], nreq); break;
        }
    }
    {
        ProfScope ps(prof, KC_CONVERT);
        hipLaunchKernelGGL(red2_kernel, dim3(tot[3]), dim3(256), 0, st, ops, op1, pre[3], nreq,
                           K1_CFG[cfg_set(KIND)][c1].bn);
    }
    {
        ProfScope ps(prof, KC_K2);
        hipLaunchKernelGGL(k2v2_kernel<KIND == LP_F16>, dim3(tot[4]), dim3(GNT), 0, st, ops, op1, pre[4], nreq);
    }
    if (!K3_CFG[cfg_set(KIND)][c3].fw) {
        ProfScope ps(prof, KC_CONVERT);
        hipLaunchKernelGGL(wprep_kernel<KIND>, dim3(tot[5]), dim3(256), 0, st, ops, op1, pre[5], nreq);
    }
    {
        ProfScope ps(prof, KC_K3);
        switch (c3) {
            case 0: launch_k3<KIND, 0>(st, tot[6], ops, op1, stg, st1, nstage, s, pre[6], nreq); break;
            case 1: launch_k3<KIND, 1>(st, tot[6], ops, op1, stg, st1, nstage, s, pre[6], nreq); break;
            case 2: launch_k3<KIND, 2>(st, tot[6], ops, op1, stg, st1, nstage, s, pre[6], nreq); break;
            case 3: launch_k3<KIND, 3>(st, tot[6], ops, op1, stg, st1, nstage, s, pre[6], nreq); break;
            case 4: launch_k3<KIND, 4>(st, tot[6], ops, op1, stg, st1, nstage, s, pre[6], nreq); break;
            default: launch_k3<KIND, 5>(st, tot[6], ops, op1, stg, st1, nstage, s, pre[6], nreq); break;
        }
    }
    NGW_HIP(hipGetLastError());
    return NGW_OK;
}

int lp2_launch_any(int kind, hipStream_t st, int c1, int c3, const LpOp* ops, const LpOp& op1, const LpStage* stg,
                   const LpStage& st1, int nstage, int s, const int* const* pre, const int* tot, int nreq,
                   ngw_context* prof)
{
    if (c1 < 0 || c1 >= LP_NCFG || c3 < 0 || c3 >= LP_NCFG) return fail(NGW_ERR_ARG, "tile configuration index");
    switch (kind) {
        case PREC_F32: return lp_launch<LP_F32>(st, c1, c3, ops, op1, stg, st1, nstage, s, pre, tot, nreq, prof);
        case PREC_BF16: return lp_launch<LP_BF16>(st, c1, c3, ops, op1, stg, st1, nstage, s, pre, tot, nreq, prof);
        case PREC_BF16X3:
            return lp_launch<LP_BF16X3>(st, c1, c3, ops, op1, stg, st1, nstage, s, pre, tot, nreq, prof);
        case PREC_F16: return lp_launch<LP_F16>(st, c1, c3, ops, op1, stg, st1, nstage, s, pre, tot, nreq, prof);
    }
    return fail(NGW_ERR_STATE, "unknown precision");
}

// K1 partial-sum slots per vector for any tile configuration.
int lp2_slots(int nv) { return std::max(k1_slots(nv), (nv - 1) / K1_BN_MIN + 2); }

// Bytes of the per-request low-precision workspaces: x' ([j][c]) and W
// ([j][mu]) in the MFMA type, t and u (FP32, [b][mu]), the FP16 per-vector
// scales; operand arrays carry their zero row; each part starts 256-byte
// aligned.
static size_t align256(size_t n) { return (n + 255) & ~size_t(255); }

size_t lp2_work_bytes(int kind, int nmu, int nv, int nc, int nvec)
{
    const size_t es = kind == PREC_F32 ? 4 : 2, ns = kind == PREC_BF16X3 ? 2 : 1;
    const size_t ldc1 = padk(nc), ldm1 = padk(nmu), N = (size_t)nv * nvec;
    return ns * (align256(es * (N + 1) * ldc1) + align256(es * (N + 1) * ldm1)) +
           2 * align256(4 * (size_t)(nvec + 1) * ldm1) + 2 * align256(4 * (size_t)nvec) + align256(4 * N) +
           align256(4 * (size_t)nvec * XS_SPLIT) + align256(4 * (size_t)nv * nc * nvec);
}

// Fill the descriptor of one request from its operator and workspaces.
void lp2_fill(const ngw_context* op, int kind, int nvec, long ld, int tda, char* work, double* tpart, LpOp& d)
{
    const size_t es = kind == PREC_F32 ? 4 : 2;
    const size_t n1 = (size_t)(op->n_intp + 1) * op->ldc1, n2 = (size_t)(op->ncond + 1) * op->ldm1;
    if (kind == PREC_F32) {
        d.pct_hi = op->pct32.ptr; d.pct_lo = nullptr;
        d.pc_hi = op->pc32.ptr; d.pc_lo = nullptr;
    } else if (kind == PREC_F16) {
        d.pct_hi = op->pcth.ptr; d.pct_lo = nullptr;
        d.pc_hi = op->pch.ptr; d.pc_lo = nullptr;
    } else {
        d.pct_hi = op->pctb.ptr; d.pct_lo = op->pctb.as<unsigned short>() + n1;
        d.pc_hi = op->pcb.ptr; d.pc_lo = op->pcb.as<unsigned short>() + n2;
    }
    d.m32t = op->m32t.as<float>(); d.psiv32 = op->psiv32.as<float>();
    d.psi = op->psi.as<double>(); d.ldp = op->ld_psi;
    d.sig = op->psi_scale.as<float>();
    d.xsig = op->xsig.as<float>() + (tda ? op->ncv : 0);
    d.dpair = op->dpairs.as<float2>() + (tda ? op->ncv : 0);
    d.nmu = op->n_intp; d.nv = op->nval; d.nc = op->ncond; d.nvec = nvec; d.ldc1 = op->ldc1; d.ldm1 = op->ldm1;
    d.ld = ld; d.tda = tda;
    const size_t N = (size_t)op->nval * nvec;
    const size_t ns = kind == PREC_BF16X3 ? 2 : 1;
    const size_t xsz = align256(es * (N + 1) * op->ldc1), wsz = align256(es * (N + 1) * op->ldm1);
    const size_t tsz = align256(4 * (size_t)(nvec + 1) * op->ldm1);
    d.xlp_hi = work;
    d.xlp_lo = ns == 2 ? work + xsz : nullptr;
    char* w0 = work + xsz * ns;
    d.wlp_hi = w0;
    d.wlp_lo = ns == 2 ? w0 + wsz : nullptr;
    char* t0 = w0 + wsz * ns;
    d.t32 = reinterpret_cast<float*>(t0);
    d.u32 = reinterpret_cast<float*>(t0 + tsz);
    char* s0 = t0 + 2 * tsz;
    const bool f16 = kind == PREC_F16;
    d.xsc = f16 ? reinterpret_cast<float*>(s0) : nullptr;
    d.umax = f16 ? reinterpret_cast<unsigned*>(s0 + align256(4 * (size_t)nvec)) : nullptr;
    d.vsc = f16 ? op->psiv_scale.as<float>() : nullptr;
    d.wsc = f16 ? reinterpret_cast<float*>(s0 + 2 * align256(4 * (size_t)nvec)) : nullptr;
    d.xpart = f16 ? reinterpret_cast<float*>(s0 + 2 * align256(4 * (size_t)nvec) + align256(4 * N)) : nullptr;
    d.q32 = reinterpret_cast<float*>(s0 + 2 * align256(4 * (size_t)nvec) + align256(4 * N) +
                                     align256(4 * (size_t)nvec * XS_SPLIT));
    d.tpart = tpart;
    d.q = nullptr; d.w = nullptr; d.a = nullptr; d.b = nullptr; d.niter = 0; d.kstep = nullptr;
}

void lp2_grid_counts(int kind, int c1, int c3, const LpOp& d, int* tot)
{
    LpGrid g = lp_grid(kind, c1, c3, d.nmu, d.nv, d.nc, d.nvec, d.ldc1, d.ldm1);
    tot[0] = g.xscale; tot[1] = g.xprep; tot[2] = g.k1; tot[3] = g.red; tot[4] = g.k2; tot[5] = g.wprep;
    tot[6] = g.k3;
}

// Single-context operator application in a low-precision mode.
int operator_lp2(ngw_context* ctx, const OperatorArgs& a)
{
    const int B = a.nvec;
    if (B <= 0) return NGW_OK;
    NGW_TRY(lp2_prepare(ctx));
    const int kind = ctx->precision, nmu = ctx->n_intp, nv = ctx->nval;
    const int S = lp2_slots(nv);
    NGW_TRY(ctx->tpart.ensure(sizeof(double) * (size_t)S * B * nmu));
    NGW_TRY(ctx->lpwork.ensure(lp2_work_bytes(kind, nmu, nv, ctx->ncond, B)));
    LpOp d;
    lp2_fill(ctx, kind, B, a.ldx, a.tda, ctx->lpwork.as<char>(), ctx->tpart.as<double>(), d);
    if (a.ldo != a.ldx || (a.v0 && a.ldv0 != a.ldx) || (a.y && a.ldy != a.ldx) || (a.acc && a.acc != a.out))
        return fail(NGW_ERR_ARG, "low-precision operator needs one leading dimension and acc == out");
    d.q = const_cast<double*>(a.v0);
    LpStage st{a.x, a.out, a.y, a.alpha, a.bx, a.by, a.v0 ? a.gamma : 0.0, a.acc ? 1 : 0, a.flags};
    const int c1 = ctx->lp_cfg[kind][0], c3 = ctx->lp_cfg[kind][1];
    int tot[LP_NKERNEL];
    lp2_grid_counts(kind, c1, c3, d, tot);
    const int* pre[LP_NKERNEL] = {};
    return lp2_launch_any(kind, ctx->stream, c1, c3, nullptr, d, nullptr, st, 1, 0, pre, tot, 1, ctx);
}

}  // namespace ngw
