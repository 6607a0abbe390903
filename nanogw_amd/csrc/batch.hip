// Batched / continuously served Lanczos GW on one GPU.
//
// Many independent GW requests -- each a molecule (its own ISDF operator) and
// a set of QP states -- advance their Lanczos recurrences together.  One
// Lanczos step of every active request is a fixed sequence of grouped
// launches: for each of the npoly polynomial stages one grouped K1 (all
// requests' tiles in one launch), one slot reduction, one grouped K2, one
// grouped K3 with the Horner/Clenshaw epilogue; then one grouped alpha/beta
// kernel and a step-counter advance.  Each WG finds its request from a tile
// prefix table, so molecules of different (Nmu, nv, nc) share every launch.
// The step index lives on the device, so the whole step is captured once in
// a hipGraph and replayed; the graph is re-captured only when requests join
// or leave.  Per-column arithmetic does not depend on the other requests in
// a launch (fixed tile shapes, per-request tiles, no atomics): a request's
// result is bitwise the same alone and in any mix.
//
// The per-request work follows the same reference routines as the single
// operator path: apply_matvec_BLOCK.F90:25-268, lanczos_spectra_BLOCK.F90:7-134
// and calculate_sigma_lanczos_BLOCK.F90z:201-444.

#include "operator_loaders.h"

#include <algorithm>
#include <map>
#include <memory>
#include <vector>

namespace ngw {

// Device descriptor of one request.
struct GOp {
    const double* psi; long ldp;
    const double* m; long ldm;
    const double* sqrtr; const double* dfull; const double* dtda;
    int nmu, nv, nc, nvec;
    long ld;
    int tda;
    double* tpart; double* t; double* u;
    double* q; double* w;
    double* a; double* b;
    int niter;
    int* kstep;
};

// Device descriptor of one polynomial stage of one request.
struct GStage {
    const double* x; double* out; const double* y;
    double alpha, bx, by, gamma;
    int accumulate;
};

__device__ __forceinline__ int find_req(const int* pre, int nreq, int bid)
{
    int lo = 0, hi = nreq;  // pre[lo] <= bid < pre[hi]
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (pre[mid] <= bid) lo = mid; else hi = mid;
    }
    return lo;
}

constexpr int GK1M = 128, GK1N = 64, GK3M = 64, GK3N = 128, GK2 = 64;
// Workgroup prefix tables of a step: one per kernel of the operator (FP64
// uses the first four, the low-precision path LP_NKERNEL), then alpha/beta.
constexpr int NTAB_AB = LP_NKERNEL > 4 ? LP_NKERNEL : 4;
__global__ __launch_bounds__(GNT) void gk1_kernel(const GOp* ops, const GStage* st, int nstage, int s, const int* pre,
                                                  int nreq)
{
    __shared__ union {
        SmemF64<GK1M, GK1N> ab;
        double red[GK1M][GK1N + 2];
    } sm;
    const int r = find_req(pre, nreq, blockIdx.x);
    const GOp op = ops[r];
    const GStage sg = st[r * nstage + s];
    const int local = blockIdx.x - pre[r];
    const int mt = (op.nmu + GK1M - 1) / GK1M;
    const int tm = local % mt, tn = local / mt;
    const int m0 = tm * GK1M, n0 = tn * GK1N;
    const int nv = op.nv, N = nv * op.nvec;
    const double* sx = op.tda ? nullptr : op.sqrtr;
    auto red_store = [&](int row, int col, double val) {
        int gm = m0 + row, gn = n0 + col;
        double p = 0.0;
        if (gm < op.nmu && gn < N) p = val * op.psi[gm + (long)(gn % nv) * op.ldp];
        sm.red[row][col] = p;
    };
    AccF64<GK1M, GK1N> acc;
    gemm_f64_tile<GK1M, GK1N>(ColMajorA{op.psi + (long)nv * op.ldp, op.ldp, op.nmu}, XprimeB{sg.x, sx, op.ld, nv, N},
                              m0, n0, op.nc, acc, sm.ab);
    for_acc_f64<GK1M, GK1N>(acc, red_store);
    __syncthreads();
    const int nend = min(n0 + GK1N, N);
    const int b_lo = n0 / nv, b_hi = (nend - 1) / nv;
    const int nbt = b_hi - b_lo + 1;
    for (int e = threadIdx.x; e < GK1M * nbt; e += GNT) {
        int row = e % GK1M, b = b_lo + e / GK1M;
        int c0 = max(n0, b * nv) - n0, c1 = min(nend, (b + 1) * nv) - n0;
        double acc = 0.0;
        for (int cc = c0; cc < c1; ++cc) acc += sm.red[row][cc];
        int gm = m0 + row;
        int slot = tn - (b * nv) / GK1N;
        if (gm < op.nmu) op.tpart[((long)slot * op.nvec + b) * op.nmu + gm] = acc;
    }
}

__global__ void gred_kernel(const GOp* ops, const int* pre, int nreq)
{
    const int bid = blockIdx.x;
    const int r = find_req(pre, nreq, bid);
    const GOp op = ops[r];
    const long i = (long)(bid - pre[r]) * blockDim.x + threadIdx.x;
    if (i >= (long)op.nmu * op.nvec) return;
    const int b = (int)(i / op.nmu), nu = (int)(i - (long)b * op.nmu);
    const int first = (b * op.nv) / GK1N, last = ((b + 1) * op.nv - 1) / GK1N;
    double acc = 0.0;
    for (int sl = 0; sl <= last - first; ++sl) acc += op.tpart[((long)sl * op.nvec + b) * op.nmu + nu];
    op.t[nu + (long)b * op.nmu] = acc;
}

__global__ __launch_bounds__(GNT) void gk2_kernel(const GOp* ops, const int* pre, int nreq)
{
    const int r = find_req(pre, nreq, blockIdx.x);
    const GOp op = ops[r];
    const int local = blockIdx.x - pre[r];
    const int mt = (op.nmu + GK2 - 1) / GK2;
    const int m0 = (local % mt) * GK2, n0 = (local / mt) * GK2;
    auto store = [&](int row, int col, double val) {
        int gm = m0 + row, gn = n0 + col;
        if (gm < op.nmu && gn < op.nvec) op.u[gm + (long)gn * op.nmu] = val;
    };
    __shared__ SmemF64<GK2, GK2> sm;
    AccF64<GK2, GK2> acc;
    gemm_f64_tile<GK2, GK2>(ColMajorA{op.m, op.ldm, op.nmu}, ColMajorB{op.t, (long)op.nmu, op.nvec}, m0, n0, op.nmu,
                            acc, sm);
    for_acc_f64<GK2, GK2>(acc, store);
}

__global__ __launch_bounds__(GNT) void gk3_kernel(const GOp* ops, const GStage* st, int nstage, int s, const int* pre,
                                                  int nreq)
{
    const int r = find_req(pre, nreq, blockIdx.x);
    const GOp op = ops[r];
    const GStage sg = st[r * nstage + s];
    const int local = blockIdx.x - pre[r];
    const int mt = (op.nc + GK3M - 1) / GK3M;
    const int m0 = (local % mt) * GK3M, n0 = (local / mt) * GK3N;
    const int nv = op.nv, N = nv * op.nvec;
    OperatorEpilogue ep;
    ep.x = sg.x; ep.ldx = op.ld;
    ep.d = op.tda ? op.dtda : op.dfull;
    ep.s = op.tda ? nullptr : op.sqrtr;
    ep.beta = op.tda ? 2.0 : 4.0;
    ep.alpha = sg.alpha;
    ep.bx = sg.bx;
    ep.y = sg.y; ep.ldy = op.ld; ep.by = sg.by;
    ep.v0 = op.q; ep.ldv0 = op.ld; ep.gamma = sg.gamma;
    const bool acc_on = sg.accumulate && (*op.kstep > 1);
    ep.acc = acc_on ? sg.out : nullptr; ep.ldacc = op.ld;
    ep.out = sg.out; ep.ldo = op.ld;
    const double* psic = op.psi + (long)nv * op.ldp;
    __shared__ SmemF64<GK3M, GK3N> sm;
    AccF64<GK3M, GK3N> acc;
    gemm_f64_tile<GK3M, GK3N>(ColMajorAT{psic, op.ldp, op.nc}, WeightedPsiVB{op.psi, op.ldp, op.u, (long)op.nmu, nv, N},
                              m0, n0, op.nmu, acc, sm);
    for_acc_f64<GK3M, GK3N>(acc, [&](int row, int col, double val) {
        int c = m0 + row, j = n0 + col;
        if (c < op.nc && j < N) {
            int b = j / nv, v = j - b * nv;
            ep.apply(v + (long)c * nv, b, val);
        }
    });
}

constexpr int GRT = 256;

__device__ __forceinline__ double gblock_sum(double v, double* sh)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[wave] = v;
    __syncthreads();
    return ((sh[0] + sh[1]) + sh[2]) + sh[3];
}

// Lanczos state of one request, shared by the FP64 and low-precision modes.
struct GLz {
    double* q; double* w; double* a; double* b;
    long ld; long n; int niter; int* kstep;
};

// Lanczos alpha/beta of step *kstep for every column of every request
// (lanczos_spectra_BLOCK.F90:63-107, rotation of :81-88 applied here).
__global__ __launch_bounds__(GRT) void gab_kernel(const GLz* lz, const int* pre, int nreq)
{
    __shared__ double sh[4];
    const int r = find_req(pre, nreq, blockIdx.x);
    const GLz op = lz[r];
    const long col = blockIdx.x - pre[r];
    const int k = *op.kstep;
    const long n = op.n;
    double* qc = op.q + col * op.ld;
    double* wc = op.w + col * op.ld;
    double s = 0.0;
    for (long i = threadIdx.x; i < n; i += GRT) s += wc[i] * qc[i];
    const double al = gblock_sum(s, sh);
    if (threadIdx.x == 0) {
        op.a[(k - 1) + col * op.niter] = al;
        if (k == 1) op.b[col * op.niter] = 1.0;
    }
    if (k + 1 > op.niter) return;
    double s2 = 0.0;
    for (long i = threadIdx.x; i < n; i += GRT) {
        double wi = wc[i] + (-al) * qc[i];
        wc[i] = wi;
        s2 += wi * wi;
    }
    const double be = sqrt(gblock_sum(s2, sh));
    if (threadIdx.x == 0) op.b[k + col * op.niter] = be;
    for (long i = threadIdx.x; i < n; i += GRT) {
        double wi = wc[i], qi = qc[i];
        qc[i] = wi / be;
        wc[i] = -qi * be;
    }
}

__global__ void gadvance_kernel(const GLz* lz, int nreq)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < nreq) *lz[r].kstep += 1;
}

__global__ void set_int_kernel(int* p, int v)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

}  // namespace ngw

using namespace ngw;

struct ngw_batch {
    ngw_context* ctx = nullptr;  // device, stream and precision
    struct System {
        std::unique_ptr<ngw_context> op;
    };
    struct Request {
        int id = 0, sys = 0;
        std::vector<int> qp;
        int nen = 0, niter = 0, npoly = 0, tda = 0;
        double deltae = 0, ecuts = 0;
        std::vector<double> poly;
        int B = 0;
        DevBuf vec, work, lpwork, ab, zz, spec, misc, kstep, sig;
        int k_host = 1;  // next Lanczos step to run
        bool done = false;
        bool held = false;  // kept out of the steps (ngw_batch_hold), state intact
        std::vector<double> sigma;
        size_t bytes = 0;
    };
    std::map<int, System> systems;
    std::map<int, Request> reqs;
    int next_sys = 1, next_req = 1;
    // grouped launch tables
    DevBuf gops, gstages, gpre, glz;
    int lp_tot[LP_NKERNEL] = {};
    int lp_c1 = 0, lp_c3 = 0;  // tile configurations of the current group
    std::vector<int> active;  // request ids in the current group
    int nstage = 0;
    int t_k1 = 0, t_red = 0, t_k2 = 0, t_k3 = 0, t_ab = 0;
    bool dirty = true;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    long steps_run = 0, graph_captures = 0;
};

namespace {

void drop_graph(ngw_batch* b)
{
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->gexec = nullptr;
    b->graph = nullptr;
}

size_t request_bytes(const ngw_context* op, int nqp, int niter, int nen)
{
    const size_t B = (size_t)nqp * op->nband;
    const size_t ncv = op->ncv;
    const size_t S = k1_slots(op->nval);
    size_t lp = 0;
    if (op->precision != PREC_F64)
        lp = lp2_work_bytes(op->precision, op->n_intp, op->nval, op->ncond, (int)B);
    return 8 * (5 * ncv * B + (S + 2) * B * op->n_intp + 2 * (size_t)niter * B + 4 * (size_t)(nen + 1) * B + 4 * B +
                2 * (size_t)nen * nqp) + lp + 64;
}

// One Lanczos step of every active request.  prof: the context whose kernel
// timers bracket each kernel class (not used under graph capture).
int launch_step_f64(ngw_batch* b, hipStream_t s, ngw_context* prof)
{
    const GOp* ops = b->gops.as<GOp>();
    const GStage* st = b->gstages.as<GStage>();
    const int nreq = (int)b->active.size();
    const int* pre = b->gpre.as<int>();
    const int* p_k1 = pre;
    const int* p_red = pre + (nreq + 1);
    const int* p_k2 = pre + 2 * (nreq + 1);
    const int* p_k3 = pre + 3 * (nreq + 1);
    for (int stg = 0; stg < b->nstage; ++stg) {
        {
            ProfScope ps(prof, KC_K1);
            hipLaunchKernelGGL(gk1_kernel, dim3(b->t_k1), dim3(GNT), 0, s, ops, st, b->nstage, stg, p_k1, nreq);
        }
        {
            ProfScope ps(prof, KC_CONVERT);
            hipLaunchKernelGGL(gred_kernel, dim3(b->t_red), dim3(256), 0, s, ops, p_red, nreq);
        }
        {
            ProfScope ps(prof, KC_K2);
            hipLaunchKernelGGL(gk2_kernel, dim3(b->t_k2), dim3(GNT), 0, s, ops, p_k2, nreq);
        }
        {
            ProfScope ps(prof, KC_K3);
            hipLaunchKernelGGL(gk3_kernel, dim3(b->t_k3), dim3(GNT), 0, s, ops, st, b->nstage, stg, p_k3, nreq);
        }
    }
    NGW_HIP(hipGetLastError());
    return NGW_OK;
}

int launch_any(ngw_batch* b, hipStream_t s, bool capturing)
{
    const int nreq = (int)b->active.size();
    const int* pre = b->gpre.as<int>();
    ngw_context* prof = capturing ? nullptr : b->ctx;
    if (b->ctx->precision == PREC_F64) {
        NGW_TRY(launch_step_f64(b, s, prof));
    } else {
        const int* p[LP_NKERNEL];
        for (int k = 0; k < LP_NKERNEL; ++k) p[k] = pre + k * (nreq + 1);
        LpOp dummy{};
        LpStage dst{};
        for (int stg = 0; stg < b->nstage; ++stg)
            NGW_TRY(lp2_launch_any(b->ctx->precision, s, b->lp_c1, b->lp_c3, b->gops.as<LpOp>(), dummy,
                                   b->gstages.as<LpStage>(), dst, b->nstage, stg, p, b->lp_tot, nreq, prof));
    }
    const int* p_ab = pre + NTAB_AB * (nreq + 1);
    {
        ProfScope ps(prof, KC_LANCZOS);
        hipLaunchKernelGGL(gab_kernel, dim3(b->t_ab), dim3(GRT), 0, s, b->glz.as<GLz>(), p_ab, nreq);
        hipLaunchKernelGGL(gadvance_kernel, dim3((nreq + 63) / 64), dim3(64), 0, s, b->glz.as<GLz>(), nreq);
    }
    NGW_HIP(hipGetLastError());
    return NGW_OK;
}

// Rebuild the descriptor tables for the requests that still have steps to run.
int rebuild_group(ngw_batch* b)
{
    drop_graph(b);
    b->active.clear();
    for (auto& kv : b->reqs)
        if (!kv.second.done && !kv.second.held) b->active.push_back(kv.first);
    const int nreq = (int)b->active.size();
    b->dirty = false;
    if (nreq == 0) return NGW_OK;
    int npoly = -1;
    for (int id : b->active) {
        const auto& rq = b->reqs[id];
        if (npoly < 0) npoly = rq.npoly;
        if (rq.npoly != npoly) return fail(NGW_ERR_ARG, "all requests in a batch must share npoly");
    }
    b->nstage = npoly;
    const int prec = b->ctx->precision;
    if (prec != PREC_F64) {
        b->lp_c1 = b->ctx->lp_cfg[prec][0];
        b->lp_c3 = b->ctx->lp_cfg[prec][1];
    }
    std::vector<GOp> ops(nreq);
    std::vector<GStage> stages((size_t)nreq * npoly);
    std::vector<LpOp> lops(nreq);
    std::vector<LpStage> lstages((size_t)nreq * npoly);
    std::vector<GLz> lz(nreq);
    std::vector<int> pre((NTAB_AB + 1) * (nreq + 1), 0);
    auto P = [&](int k, int i) -> int& { return pre[(size_t)k * (nreq + 1) + i]; };
    for (int i = 0; i < nreq; ++i) {
        auto& rq = b->reqs[b->active[i]];
        ngw_context* op = b->systems[rq.sys].op.get();
        const long ncv = op->ncv;
        const int B = rq.B, S = lp2_slots(op->nval);
        double* vec = rq.vec.as<double>();
        double* q = vec;
        double* w = vec + (size_t)ncv * B;
        double* h0 = vec + 2 * (size_t)ncv * B;
        double* h1 = vec + 3 * (size_t)ncv * B;
        double* h2 = vec + 4 * (size_t)ncv * B;
        double* work = rq.work.as<double>();
        double* tpart = work;
        double* t = work + (size_t)S * B * op->n_intp;
        double* u = t + (size_t)B * op->n_intp;
        GLz z;
        z.q = q; z.w = w; z.a = rq.ab.as<double>(); z.b = z.a + (size_t)rq.niter * B;
        z.ld = ncv; z.n = ncv; z.niter = rq.niter; z.kstep = rq.kstep.as<int>();
        lz[i] = z;
        const std::vector<PolyStage> plan = plan_poly(op, prec, q, w, h0, h1, h2, rq.poly.data(), npoly, true);
        const int nmu = op->n_intp, nv = op->nval, nc = op->ncond;
        if (prec == PREC_F64) {
            GOp g;
            g.psi = op->psi.as<double>(); g.ldp = op->ld_psi;
            g.m = op->mmat.as<double>(); g.ldm = op->ld_m;
            g.sqrtr = op->sqrtr.as<double>(); g.dfull = op->diag_full.as<double>(); g.dtda = op->diag_tda.as<double>();
            g.nmu = nmu; g.nv = nv; g.nc = nc; g.nvec = B;
            g.ld = ncv; g.tda = rq.tda;
            g.tpart = tpart; g.t = t; g.u = u;
            g.q = q; g.w = w; g.a = z.a; g.b = z.b; g.niter = rq.niter; g.kstep = z.kstep;
            ops[i] = g;
            for (int sgi = 0; sgi < npoly; ++sgi) {
                const PolyStage& p = plan[sgi];
                stages[(size_t)i * npoly + sgi] = GStage{p.x, p.out, p.y, p.alpha, p.bx, p.by, p.gamma,
                                                         p.accumulate ? 1 : 0};
            }
            P(0, i + 1) = P(0, i) + ((nmu + GK1M - 1) / GK1M) * ((nv * B + GK1N - 1) / GK1N);
            P(1, i + 1) = P(1, i) + (int)(((long)nmu * B + 255) / 256);
            P(2, i + 1) = P(2, i) + ((nmu + GK2 - 1) / GK2) * ((B + GK2 - 1) / GK2);
            P(3, i + 1) = P(3, i) + ((nc + GK3M - 1) / GK3M) * ((nv * B + GK3N - 1) / GK3N);
        } else {
            NGW_TRY(lp2_prepare(op));
            NGW_TRY(rq.lpwork.ensure(lp2_work_bytes(prec, nmu, nv, nc, B)));
            LpOp d;
            lp2_fill(op, prec, B, ncv, rq.tda, rq.lpwork.as<char>(), tpart, d);
            d.q = q; d.w = w; d.a = z.a; d.b = z.b; d.niter = rq.niter; d.kstep = z.kstep;
            lops[i] = d;
            for (int sgi = 0; sgi < npoly; ++sgi) {
                const PolyStage& p = plan[sgi];
                lstages[(size_t)i * npoly + sgi] = LpStage{p.x, p.out, p.y, p.alpha, p.bx, p.by, p.gamma,
                                                           p.accumulate ? 2 : 0, p.flags};
            }
            int tot[LP_NKERNEL];
            lp2_grid_counts(prec, b->lp_c1, b->lp_c3, d, tot);
            for (int k = 0; k < LP_NKERNEL; ++k) P(k, i + 1) = P(k, i) + tot[k];
        }
        P(NTAB_AB, i + 1) = P(NTAB_AB, i) + B;
    }
    hipStream_t s = b->ctx->stream;
    if (prec == PREC_F64) {
        b->t_k1 = P(0, nreq); b->t_red = P(1, nreq); b->t_k2 = P(2, nreq); b->t_k3 = P(3, nreq);
        NGW_TRY(b->gops.ensure(sizeof(GOp) * ops.size()));
        NGW_TRY(b->gstages.ensure(sizeof(GStage) * stages.size()));
        NGW_HIP(hipMemcpyAsync(b->gops.ptr, ops.data(), sizeof(GOp) * ops.size(), hipMemcpyHostToDevice, s));
        NGW_HIP(hipMemcpyAsync(b->gstages.ptr, stages.data(), sizeof(GStage) * stages.size(), hipMemcpyHostToDevice, s));
    } else {
        for (int k = 0; k < LP_NKERNEL; ++k) b->lp_tot[k] = P(k, nreq);
        NGW_TRY(b->gops.ensure(sizeof(LpOp) * lops.size()));
        NGW_TRY(b->gstages.ensure(sizeof(LpStage) * lstages.size()));
        NGW_HIP(hipMemcpyAsync(b->gops.ptr, lops.data(), sizeof(LpOp) * lops.size(), hipMemcpyHostToDevice, s));
        NGW_HIP(hipMemcpyAsync(b->gstages.ptr, lstages.data(), sizeof(LpStage) * lstages.size(),
                               hipMemcpyHostToDevice, s));
    }
    b->t_ab = P(NTAB_AB, nreq);
    NGW_TRY(b->glz.ensure(sizeof(GLz) * lz.size()));
    NGW_TRY(b->gpre.ensure(sizeof(int) * pre.size()));
    NGW_HIP(hipMemcpyAsync(b->glz.ptr, lz.data(), sizeof(GLz) * lz.size(), hipMemcpyHostToDevice, s));
    NGW_HIP(hipMemcpyAsync(b->gpre.ptr, pre.data(), sizeof(int) * pre.size(), hipMemcpyHostToDevice, s));
    NGW_HIP(hipStreamSynchronize(s));
    return NGW_OK;
}

// Continued fraction, Sigma_c and host copy for a request whose niter steps ran.
int finalize(ngw_batch* b, ngw_batch::Request& rq)
{
    ngw_context* op = b->systems[rq.sys].op.get();
    hipStream_t s = b->ctx->stream;
    const int B = rq.B, nz = rq.nen + 1, nqp = (int)rq.qp.size();
    double* a = rq.ab.as<double>();
    double* bb = a + (size_t)rq.niter * B;
    NGW_TRY(continued_fraction(s, a, bb, rq.niter, B, rq.zz.as<double>(), nz, rq.spec.as<double>()));
    double* pn2 = rq.misc.as<double>();
    NGW_TRY(sigma_assemble(s, rq.spec.as<double>(), rq.zz.as<double>(), pn2, rq.nen, nqp, op->nband, 0, op->nval, rq.tda,
                           rq.sig.as<double>()));
    rq.sigma.resize(2 * (size_t)rq.nen * nqp);
    NGW_HIP(hipMemcpyAsync(rq.sigma.data(), rq.sig.ptr, sizeof(double) * rq.sigma.size(), hipMemcpyDeviceToHost, s));
    NGW_HIP(hipStreamSynchronize(s));
    rq.done = true;
    rq.vec.release();
    rq.work.release();
    rq.lpwork.release();
    return NGW_OK;
}

}  // namespace

extern "C" {

int ngw_batch_create(ngw_context* ctx, ngw_batch** out)
{
    if (!ctx || !out) return fail(NGW_ERR_ARG, "null argument");
    ngw_batch* b = new ngw_batch();
    b->ctx = ctx;
    *out = b;
    return NGW_OK;
}

int ngw_batch_destroy(ngw_batch* b)
{
    if (!b) return NGW_OK;
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    drop_graph(b);
    for (auto& kv : b->systems) ngw_destroy(kv.second.op.release());
    delete b;
    return NGW_OK;
}

int ngw_batch_add_system(ngw_batch* b, int n_intp, int nband, int nval, const double* psi_intp, int ld_psi,
                         const double* mmtrx, int ld_m, const double* e1, int* sys_id)
{
    if (!b || !sys_id) return fail(NGW_ERR_ARG, "null argument");
    ngw_context* op = new ngw_context();
    op->device = b->ctx->device;
    op->stream = b->ctx->stream;
    op->shared_stream = true;
    int st = ngw_set_isdf(op, n_intp, nband, nval, psi_intp, ld_psi, mmtrx, ld_m, e1);
    if (st == NGW_OK && b->ctx->precision != PREC_F64) st = lp2_prepare(op);
    if (st != NGW_OK) {
        ngw_destroy(op);
        return st;
    }
    op->precision = b->ctx->precision;
    const int id = b->next_sys++;
    b->systems[id].op.reset(op);
    *sys_id = id;
    return NGW_OK;
}

int ngw_batch_remove_system(ngw_batch* b, int sys_id)
{
    if (!b) return fail(NGW_ERR_ARG, "null batch");
    auto it = b->systems.find(sys_id);
    if (it == b->systems.end()) return fail(NGW_ERR_ARG, "unknown system");
    for (auto& kv : b->reqs)
        if (kv.second.sys == sys_id && !kv.second.done) return fail(NGW_ERR_STATE, "system has running requests");
    NGW_HIP(hipStreamSynchronize(b->ctx->stream));
    ngw_destroy(it->second.op.release());
    b->systems.erase(it);
    return NGW_OK;
}

int ngw_batch_request_bytes(ngw_batch* b, int sys_id, int nqp, int niter, int nen, long* bytes)
{
    if (!b || !bytes) return fail(NGW_ERR_ARG, "null argument");
    auto it = b->systems.find(sys_id);
    if (it == b->systems.end()) return fail(NGW_ERR_ARG, "unknown system");
    *bytes = (long)request_bytes(it->second.op.get(), nqp, niter, nen);
    return NGW_OK;
}

int ngw_batch_submit(ngw_batch* b, int sys_id, int nqp, const int* qp_band, int nen, double deltae, double ecuts,
                     int niter, const double* polynomial, int npoly, int tamm_d, int* req_id)
{
    if (!b || !req_id || !qp_band || !polynomial) return fail(NGW_ERR_ARG, "null argument");
    auto it = b->systems.find(sys_id);
    if (it == b->systems.end()) return fail(NGW_ERR_ARG, "unknown system");
    ngw_context* op = it->second.op.get();
    if (nqp <= 0 || nen < 2 || niter < 1 || npoly < 1) return fail(NGW_ERR_ARG, "bad nqp, nen, niter or npoly");
    for (int i = 0; i < nqp; ++i)
        if (qp_band[i] < 0 || qp_band[i] >= op->nband) return fail(NGW_ERR_ARG, "qp_band out of range");
    NGW_HIP(hipSetDevice(b->ctx->device));
    op->precision = b->ctx->precision;
    hipStream_t s = b->ctx->stream;
    const int id = b->next_req++;
    ngw_batch::Request& rq = b->reqs[id];
    rq.id = id; rq.sys = sys_id;
    rq.qp.assign(qp_band, qp_band + nqp);
    rq.nen = nen; rq.deltae = deltae; rq.ecuts = ecuts; rq.niter = niter; rq.npoly = npoly; rq.tda = tamm_d ? 1 : 0;
    rq.poly.assign(polynomial, polynomial + npoly + 1);
    const int nb = op->nband, B = nqp * nb, nz = nen + 1;
    const long ncv = op->ncv;
    rq.B = B;
    rq.bytes = request_bytes(op, nqp, niter, nen);
    const int S = lp2_slots(op->nval);
    int st = NGW_OK;
    auto fail_req = [&](int code) {
        b->reqs.erase(id);
        return code;
    };
    if ((st = rq.vec.ensure(sizeof(double) * 5 * (size_t)ncv * B)) != NGW_OK) return fail_req(st);
    if ((st = rq.work.ensure(sizeof(double) * ((size_t)S + 2) * B * op->n_intp)) != NGW_OK) return fail_req(st);
    if ((st = rq.ab.ensure(sizeof(double) * 2 * (size_t)niter * B)) != NGW_OK) return fail_req(st);
    if ((st = rq.zz.ensure(sizeof(double) * 2 * (size_t)nz * B)) != NGW_OK) return fail_req(st);
    if ((st = rq.spec.ensure(sizeof(double) * 2 * (size_t)nz * B)) != NGW_OK) return fail_req(st);
    if ((st = rq.misc.ensure(sizeof(double) * 2 * (size_t)B + sizeof(int) * 2 * (size_t)B)) != NGW_OK) return fail_req(st);
    if ((st = rq.kstep.ensure(sizeof(int))) != NGW_OK) return fail_req(st);
    if ((st = rq.sig.ensure(sizeof(double) * 2 * (size_t)nen * nqp)) != NGW_OK) return fail_req(st);
    std::vector<int> mband(B), nband(B);
    for (int iq = 0; iq < nqp; ++iq)
        for (int n = 0; n < nb; ++n) {
            mband[(size_t)iq * nb + n] = qp_band[iq];
            nband[(size_t)iq * nb + n] = n;
        }
    double* pn2 = rq.misc.as<double>();
    double* pn = pn2 + B;
    int* dmb = reinterpret_cast<int*>(pn + B);
    int* dnb = dmb + B;
    NGW_HIP(hipMemcpyAsync(dmb, mband.data(), sizeof(int) * B, hipMemcpyHostToDevice, s));
    NGW_HIP(hipMemcpyAsync(dnb, nband.data(), sizeof(int) * B, hipMemcpyHostToDevice, s));
    double* q = rq.vec.as<double>();
    // P vectors straight into the Lanczos start vectors (BLOCK driver :286-358)
    NGW_TRY(pvectors_f64(op, dmb, dnb, B, rq.tda, q, ncv));
    NGW_TRY(column_norms(s, q, ncv, ncv, B, pn2, pn));
    NGW_TRY(scale_columns(s, q, q, ncv, ncv, B, pn));
    NGW_TRY(zz_grid(s, op->e1.as<double>(), dmb, dnb, op->nval, B, nen, deltae, ecuts, rq.zz.as<double>()));
    hipLaunchKernelGGL(set_int_kernel, dim3(1), dim3(1), 0, s, rq.kstep.as<int>(), 1);
    NGW_HIP(hipGetLastError());
    NGW_HIP(hipStreamSynchronize(s));
    rq.k_host = 1;
    b->dirty = true;
    *req_id = id;
    return NGW_OK;
}

int ngw_batch_step(ngw_batch* b, int nsteps, int use_graph, int* n_finished)
{
    if (!b) return fail(NGW_ERR_ARG, "null batch");
    NGW_HIP(hipSetDevice(b->ctx->device));
    hipStream_t s = b->ctx->stream;
    int finished = 0;
    for (int it = 0; it < nsteps; ++it) {
        if (b->dirty) NGW_TRY(rebuild_group(b));
        if (b->active.empty()) break;
        if (use_graph) {
            if (!b->gexec) {
                NGW_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                int st = launch_any(b, s, true);
                hipGraph_t g = nullptr;
                hipError_t e = hipStreamEndCapture(s, &g);
                if (st != NGW_OK) {
                    if (g) (void)hipGraphDestroy(g);
                    return st;
                }
                if (e != hipSuccess) return fail(NGW_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
                b->graph = g;
                NGW_HIP(hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0));
                ++b->graph_captures;
            }
            NGW_HIP(hipGraphLaunch(b->gexec, s));
        } else {
            NGW_TRY(launch_any(b, s, false));
        }
        ++b->steps_run;
        for (int id : b->active) {
            auto& rq = b->reqs[id];
            if (++rq.k_host > rq.niter) {
                NGW_TRY(finalize(b, rq));
                ++finished;
                b->dirty = true;
            }
        }
    }
    NGW_HIP(hipStreamSynchronize(s));
    if (n_finished) *n_finished = finished;
    return NGW_OK;
}

int ngw_batch_active(ngw_batch* b, int* n_active)
{
    if (!b || !n_active) return fail(NGW_ERR_ARG, "null argument");
    int n = 0;
    for (auto& kv : b->reqs)
        if (!kv.second.done) ++n;
    *n_active = n;
    return NGW_OK;
}

int ngw_batch_result(ngw_batch* b, int req_id, double* sigma_c, int* done)
{
    if (!b || !done) return fail(NGW_ERR_ARG, "null argument");
    auto it = b->reqs.find(req_id);
    if (it == b->reqs.end()) return fail(NGW_ERR_ARG, "unknown request");
    *done = it->second.done ? 1 : 0;
    if (it->second.done && sigma_c)
        std::copy(it->second.sigma.begin(), it->second.sigma.end(), sigma_c);
    return NGW_OK;
}

int ngw_batch_release(ngw_batch* b, int req_id)
{
    if (!b) return fail(NGW_ERR_ARG, "null batch");
    auto it = b->reqs.find(req_id);
    if (it == b->reqs.end()) return fail(NGW_ERR_ARG, "unknown request");
    NGW_HIP(hipStreamSynchronize(b->ctx->stream));
    if (!it->second.done) b->dirty = true;
    b->reqs.erase(it);
    return NGW_OK;
}

int ngw_batch_hold(ngw_batch* b, int req_id, int hold)
{
    if (!b) return fail(NGW_ERR_ARG, "null batch");
    auto it = b->reqs.find(req_id);
    if (it == b->reqs.end()) return fail(NGW_ERR_ARG, "unknown request");
    const bool h = hold != 0;
    if (it->second.held != h && !it->second.done) b->dirty = true;
    it->second.held = h;
    return NGW_OK;
}

int ngw_batch_stats(ngw_batch* b, long* steps_run, long* graph_captures)
{
    if (!b) return fail(NGW_ERR_ARG, "null batch");
    if (steps_run) *steps_run = b->steps_run;
    if (graph_captures) *graph_captures = b->graph_captures;
    return NGW_OK;
}

}  // extern "C"
