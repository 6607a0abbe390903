// extern "C" entry points of libngw (declared in include/ngw.h).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "ngw_kernels.h"

namespace ngw {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const std::string& msg)
{
    g_last_error = msg;
    return code;
}

int DevBuf::ensure(size_t need)
{
    if (need <= bytes && ptr) return NGW_OK;
    release();
    if (need == 0) need = 256;
    need = (need + 255) & ~size_t(255);
    hipError_t e = hipMalloc(&ptr, need);
    if (e != hipSuccess) {
        ptr = nullptr;
        bytes = 0;
        (void)hipGetLastError();
        return fail(NGW_ERR_NOMEM, "hipMalloc of " + std::to_string(need) + " bytes failed: " +
                                       hipGetErrorString(e));
    }
    bytes = need;
    return NGW_OK;
}

void DevBuf::release()
{
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

ProfScope::ProfScope(ngw_context* ctx, int k) : c(ctx), kind(k)
{
    if (!c || !c->prof_on) return;
    for (hipEvent_t* e : {&a, &b}) {
        if (!c->prof_pool.empty()) {
            *e = c->prof_pool.back();
            c->prof_pool.pop_back();
        } else if (hipEventCreate(e) != hipSuccess) {
            *e = nullptr;
        }
    }
    if (a) (void)hipEventRecord(a, c->stream);
}

static void prof_drain(ngw_context* c);

ProfScope::~ProfScope()
{
    if (!c || !c->prof_on || !a || !b) return;
    (void)hipEventRecord(b, c->stream);
    c->prof_recs.push_back({kind, a, b});
    // bound the number of live events on long runs (one sync per 4096 launches)
    if (c->prof_recs.size() >= 4096) prof_drain(c);
}

static void prof_drain(ngw_context* c)
{
    if (c->prof_recs.empty()) return;
    (void)hipStreamSynchronize(c->stream);
    for (auto& r : c->prof_recs) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
            c->prof_ms[r.kind] += ms;
            c->prof_cnt[r.kind] += 1;
        }
        c->prof_pool.push_back(r.a);
        c->prof_pool.push_back(r.b);
    }
    c->prof_recs.clear();
}

int operator_apply(ngw_context* ctx, const OperatorArgs& a)
{
    if (ctx->explicit_c) return operator_explicit(ctx, a);
    if (ctx->blocked) return operator_blocked(ctx, a);
    if (ctx->precision == PREC_F64) return operator_f64(ctx, a);
    return operator_lp2(ctx, a);
}

static int check_ctx(ngw_context* ctx)
{
    if (!ctx) return fail(NGW_ERR_ARG, "null context");
    if (ctx->n_intp <= 0 || ctx->ncv <= 0) return fail(NGW_ERR_STATE, "ISDF operator not uploaded (ngw_set_isdf)");
    NGW_HIP(hipSetDevice(ctx->device));
    return NGW_OK;
}

// Entries that read the orbitals at the ISDF points, not only the operator.
static int check_factored(ngw_context* ctx)
{
    NGW_TRY(check_ctx(ctx));
    if (ctx->explicit_c)
        return fail(NGW_ERR_STATE, "operator given by an explicit C matrix: this entry needs ngw_set_isdf");
    if (ctx->blocked)
        return fail(NGW_ERR_STATE, "operator given by irrep blocks: this entry needs ngw_set_isdf");
    return NGW_OK;
}

static int upload(ngw_context* ctx, DevBuf& buf, const void* src, size_t bytes)
{
    NGW_TRY(buf.ensure(bytes));
    NGW_HIP(hipMemcpyAsync(buf.ptr, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    return NGW_OK;
}

// ------------------------------------------------------------ host polyfit
// Equilibrated LU solve with refinement, the algorithm of LAPACK dgesvx with
// fact = 'E' that shared/linear_solver.f90:20-28 calls.
static void solve_equilibrated(int n, const std::vector<double>& Ain, const std::vector<double>& bin,
                               double* x)
{
    std::vector<double> A(Ain), LU, r(n), c(n), b(n), res(n), dx(n);
    std::vector<int> piv(n);
    double rmin = 1e300, rmax = 0.0, amax = 0.0;
    for (int i = 0; i < n; ++i) {
        double m = 0.0;
        for (int j = 0; j < n; ++j) m = std::max(m, std::fabs(A[i + j * n]));
        r[i] = m;
        amax = std::max(amax, m);
        rmin = std::min(rmin, m);
        rmax = std::max(rmax, m);
    }
    for (int i = 0; i < n; ++i) r[i] = 1.0 / r[i];
    const double rowcnd = rmin / rmax;
    double cmin = 1e300, cmax = 0.0;
    for (int j = 0; j < n; ++j) {
        double m = 0.0;
        for (int i = 0; i < n; ++i) m = std::max(m, std::fabs(A[i + j * n]) * r[i]);
        c[j] = m;
        cmin = std::min(cmin, m);
        cmax = std::max(cmax, m);
    }
    for (int j = 0; j < n; ++j) c[j] = 1.0 / c[j];
    const double colcnd = cmin / cmax;
    const double small = 2.2250738585072014e-308 / 1.1102230246251565e-16, large = 1.0 / small;
    if (rowcnd >= 0.1 && amax >= small && amax <= large) std::fill(r.begin(), r.end(), 1.0);
    if (colcnd >= 0.1) std::fill(c.begin(), c.end(), 1.0);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) A[i + j * n] *= r[i] * c[j];
    for (int i = 0; i < n; ++i) b[i] = bin[i] * r[i];
    LU = A;
    for (int k = 0; k < n; ++k) {
        int p = k;
        for (int i = k + 1; i < n; ++i)
            if (std::fabs(LU[i + k * n]) > std::fabs(LU[p + k * n])) p = i;
        piv[k] = p;
        if (p != k)
            for (int j = 0; j < n; ++j) std::swap(LU[k + j * n], LU[p + j * n]);
        const double d = LU[k + k * n];
        for (int i = k + 1; i < n; ++i) LU[i + k * n] /= d;
        for (int j = k + 1; j < n; ++j)
            for (int i = k + 1; i < n; ++i) LU[i + j * n] -= LU[i + k * n] * LU[k + j * n];
    }
    auto lu_solve = [&](double* v) {
        for (int k = 0; k < n; ++k)
            if (piv[k] != k) std::swap(v[k], v[piv[k]]);
        for (int j = 0; j < n; ++j)
            for (int i = j + 1; i < n; ++i) v[i] -= LU[i + j * n] * v[j];
        for (int j = n - 1; j >= 0; --j) {
            v[j] /= LU[j + j * n];
            for (int i = 0; i < j; ++i) v[i] -= LU[i + j * n] * v[j];
        }
    };
    std::copy(b.begin(), b.end(), x);
    lu_solve(x);
    const double eps = 1.1102230246251565e-16;
    double lastberr = 1e300;
    for (int it = 0; it < 5; ++it) {
        double berr = 0.0;
        for (int i = 0; i < n; ++i) {
            double s = b[i], den = std::fabs(b[i]);
            for (int j = 0; j < n; ++j) {
                s -= A[i + j * n] * x[j];
                den += std::fabs(A[i + j * n] * x[j]);
            }
            res[i] = s;
            if (den > 0) berr = std::max(berr, std::fabs(s) / den);
        }
        if (!(berr > eps && 2.0 * berr <= lastberr)) break;
        dx = res;
        lu_solve(dx.data());
        for (int i = 0; i < n; ++i) x[i] += dx[i];
        lastberr = berr;
    }
    for (int i = 0; i < n; ++i) x[i] *= c[i];
}

}  // namespace ngw

using namespace ngw;

extern "C" {

const char* ngw_version(void) { return "ngw-mi355x 0.1 (gfx950)"; }

const char* ngw_last_error(void) { return g_last_error.c_str(); }

int ngw_create(ngw_context** out, int device)
{
    if (!out) return fail(NGW_ERR_ARG, "null output pointer");
    int ndev = 0;
    NGW_HIP(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) return fail(NGW_ERR_ARG, "device index out of range");
    NGW_HIP(hipSetDevice(device));
    NGW_TRY(noalloc_setup(device));
    ngw_context* c = new ngw_context();
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return fail(NGW_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    *out = c;
    return NGW_OK;
}

int ngw_destroy(ngw_context* ctx)
{
    if (!ctx) return NGW_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    DevBuf* bufs[] = {&ctx->psi, &ctx->mmat, &ctx->e1, &ctx->sqrtr, &ctx->diag_full, &ctx->diag_tda,
                      &ctx->psi_scale,
                      &ctx->tpart, &ctx->ubuf, &ctx->w0, &ctx->w1, &ctx->w2, &ctx->w3, &ctx->w4,
                      &ctx->ab, &ctx->zz, &ctx->spec, &ctx->misc, &ctx->misc2, &ctx->staging,
                      &ctx->pct32, &ctx->pc32, &ctx->pctb, &ctx->pcb, &ctx->pcth, &ctx->pch, &ctx->m32t,
                      &ctx->psiv32, &ctx->psiv_scale, &ctx->xsig, &ctx->dpairs, &ctx->lpwork, &ctx->gv_part};
    for (DevBuf* b : bufs) b->release();
    explicit_release(ctx);
    for (auto& r : ctx->prof_recs) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    for (hipEvent_t e : ctx->prof_pool) (void)hipEventDestroy(e);
    if (!ctx->shared_stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return NGW_OK;
}

int ngw_synchronize(ngw_context* ctx)
{
    if (!ctx) return fail(NGW_ERR_ARG, "null context");
    NGW_HIP(hipStreamSynchronize(ctx->stream));
    return NGW_OK;
}

int ngw_set_isdf(ngw_context* ctx, int n_intp, int nband, int nval, const double* psi_intp,
                 int ld_psi, const double* mmtrx, int ld_m, const double* e1)
{
    if (!ctx) return fail(NGW_ERR_ARG, "null context");
    if (n_intp <= 0 || nband <= 0 || nval <= 0 || nval >= nband)
        return fail(NGW_ERR_ARG, "need n_intp > 0 and 0 < nval < nband");
    if (ld_psi < n_intp || ld_m < n_intp) return fail(NGW_ERR_ARG, "leading dimension < n_intp");
    if (!psi_intp || !mmtrx || !e1) return fail(NGW_ERR_ARG, "null input array");
    NGW_HIP(hipSetDevice(ctx->device));
    explicit_release(ctx);
    const int nc = nband - nval;
    const long ncv = (long)nval * nc;
    for (int ic = 0; ic < nc; ++ic)
        for (int iv = 0; iv < nval; ++iv)
            if (!(e1[nval + ic] - e1[iv] > 0.0))
                return fail(NGW_ERR_ARG, "e_c - e_v must be positive for every pair");
    ctx->n_intp = n_intp;
    ctx->nband = nband;
    ctx->nval = nval;
    ctx->ncond = nc;
    ctx->ncv = ncv;
    ctx->ld_psi = n_intp;
    ctx->ld_m = n_intp;
    ctx->lowp_ready = false;
    ctx->lp2_ready = false;
    // compact copies (ld = n_intp)
    std::vector<double> tmp((size_t)n_intp * std::max(nband, n_intp));
    for (int j = 0; j < nband; ++j)
        std::memcpy(&tmp[(size_t)j * n_intp], psi_intp + (size_t)j * ld_psi, sizeof(double) * n_intp);
    NGW_TRY(upload(ctx, ctx->psi, tmp.data(), sizeof(double) * (size_t)n_intp * nband));
    NGW_HIP(hipStreamSynchronize(ctx->stream));
    for (int j = 0; j < n_intp; ++j)
        std::memcpy(&tmp[(size_t)j * n_intp], mmtrx + (size_t)j * ld_m, sizeof(double) * n_intp);
    NGW_TRY(upload(ctx, ctx->mmat, tmp.data(), sizeof(double) * (size_t)n_intp * n_intp));
    NGW_TRY(upload(ctx, ctx->e1, e1, sizeof(double) * nband));
    ctx->e1_host.assign(e1, e1 + nband);
    // sqrtR(icv) = (e_c - e_v)^0.5 (calculate_sigma_lanczos_BLOCK.F90z:204),
    // sqrtR^3 and sqrtR^2 for the diagonal terms (apply_matvec_BLOCK.F90:198,
    // apply_matvec_lowcomsym.F90:399).
    std::vector<double> s(ncv), d3(ncv), d2(ncv);
    for (int ic = 0; ic < nc; ++ic)
        for (int iv = 0; iv < nval; ++iv) {
            long icv = iv + (long)nval * ic;
            double v = std::pow(e1[nval + ic] - e1[iv], 0.5);
            s[icv] = v;
            d3[icv] = v * v * v;
            d2[icv] = v * v;
        }
    NGW_TRY(upload(ctx, ctx->sqrtr, s.data(), sizeof(double) * ncv));
    NGW_TRY(upload(ctx, ctx->diag_full, d3.data(), sizeof(double) * ncv));
    NGW_TRY(upload(ctx, ctx->diag_tda, d2.data(), sizeof(double) * ncv));
    NGW_HIP(hipStreamSynchronize(ctx->stream));
    return NGW_OK;
}

int ngw_set_precision(ngw_context* ctx, int precision)
{
    if (!ctx) return fail(NGW_ERR_ARG, "null context");
    if (precision < NGW_PREC_F64 || precision > NGW_PREC_F16) return fail(NGW_ERR_ARG, "bad precision");
    if (ctx->explicit_c && precision != NGW_PREC_F64)
        return fail(NGW_ERR_STATE, "the explicit-C operator is FP64 only");
    if (ctx->blocked && precision != NGW_PREC_F64)
        return fail(NGW_ERR_STATE, "the irrep-blocked operator is FP64 only");
    ctx->precision = precision;
    if (precision != NGW_PREC_F64 && ctx->n_intp > 0) {
        NGW_HIP(hipSetDevice(ctx->device));
        NGW_TRY(lowp_prepare(ctx));
    }
    return NGW_OK;
}

int ngw_set_reorthogonalization(ngw_context* ctx, int mode)
{
    if (!ctx) return fail(NGW_ERR_ARG, "null context");
    if (mode != NGW_REORTH_NONE && mode != NGW_REORTH_FULL) return fail(NGW_ERR_ARG, "bad reorthogonalisation mode");
    ctx->reorth = mode;
    if (mode == NGW_REORTH_NONE) ctx->qbasis.release();
    return NGW_OK;
}

int ngw_matvec_isdf_block_dev(ngw_context* ctx, const double* d_vec, double* d_hvec, long ld,
                              int blksz, int tamm_d)
{
    NGW_TRY(check_ctx(ctx));
    if (blksz < 0 || ld < ctx->ncv) return fail(NGW_ERR_ARG, "bad blksz or leading dimension");
    if (blksz == 0) return NGW_OK;
    if (d_vec == d_hvec) return fail(NGW_ERR_ARG, "vec and hvec must not alias");
    OperatorArgs a;
    a.x = d_vec; a.ldx = ld; a.nvec = blksz; a.tda = tamm_d ? 1 : 0;
    a.out = d_hvec; a.ldo = ld;
    return operator_apply(ctx, a);
}

int ngw_autotune(ngw_context* ctx, int nvec, int reps, int* k1_cfg, int* k3_cfg)
{
    NGW_TRY(check_factored(ctx));
    if (nvec <= 0 || reps <= 0) return fail(NGW_ERR_ARG, "need nvec > 0 and reps > 0");
    const int prec = ctx->precision;
    if (prec == PREC_F64) {
        if (k1_cfg) *k1_cfg = 0;
        if (k3_cfg) *k3_cfg = 0;
        return NGW_OK;
    }
    NGW_TRY(lp2_prepare(ctx));
    hipStream_t s = ctx->stream;
    const size_t bytes = sizeof(double) * (size_t)ctx->ncv * nvec;
    DevBuf x, y;
    NGW_TRY(x.ensure(bytes));
    NGW_TRY(y.ensure(bytes));
    // any finite values time the same; 0x3f bytes make every double 4.8e-4
    NGW_HIP(hipMemsetAsync(x.ptr, 0x3f, bytes, s));
    prof_drain(ctx);
    const bool was_on = ctx->prof_on;
    double keep_ms[KC_COUNT];
    long keep_cnt[KC_COUNT];
    std::copy(ctx->prof_ms, ctx->prof_ms + KC_COUNT, keep_ms);
    std::copy(ctx->prof_cnt, ctx->prof_cnt + KC_COUNT, keep_cnt);
    ctx->prof_on = true;
    OperatorArgs a;
    a.x = x.as<double>(); a.ldx = ctx->ncv; a.nvec = nvec; a.out = y.as<double>(); a.ldo = ctx->ncv;
    double best1 = 1e300, best3 = 1e300;
    int b1 = 0, b3 = 0, st = NGW_OK;
    for (int c = 0; c < NGW_NTILE_CONFIGS && st == NGW_OK; ++c) {
        ctx->lp_cfg[prec][0] = c;
        ctx->lp_cfg[prec][1] = c;
        st = operator_apply(ctx, a);  // first use of a configuration
        prof_drain(ctx);
        std::fill(ctx->prof_ms, ctx->prof_ms + KC_COUNT, 0.0);
        for (int r = 0; r < reps && st == NGW_OK; ++r) st = operator_apply(ctx, a);
        prof_drain(ctx);
        if (ctx->prof_ms[KC_K1] < best1) { best1 = ctx->prof_ms[KC_K1]; b1 = c; }
        if (ctx->prof_ms[KC_K3] < best3) { best3 = ctx->prof_ms[KC_K3]; b3 = c; }
    }
    ctx->lp_cfg[prec][0] = b1;
    ctx->lp_cfg[prec][1] = b3;
    std::copy(keep_ms, keep_ms + KC_COUNT, ctx->prof_ms);
    std::copy(keep_cnt, keep_cnt + KC_COUNT, ctx->prof_cnt);
    ctx->prof_on = was_on;
    NGW_HIP(hipStreamSynchronize(s));
    if (st != NGW_OK) return st;
    if (k1_cfg) *k1_cfg = b1;
    if (k3_cfg) *k3_cfg = b3;
    return NGW_OK;
}

int ngw_set_tile_config(ngw_context* ctx, int k1_cfg, int k3_cfg)
{
    if (!ctx) return fail(NGW_ERR_ARG, "null context");
    if (k1_cfg < 0 || k1_cfg >= NGW_NTILE_CONFIGS || k3_cfg < 0 || k3_cfg >= NGW_NTILE_CONFIGS)
        return fail(NGW_ERR_ARG, "configuration index 0.." + std::to_string(NGW_NTILE_CONFIGS - 1));
    if (ctx->precision == PREC_F64) return NGW_OK;
    ctx->lp_cfg[ctx->precision][0] = k1_cfg;
    ctx->lp_cfg[ctx->precision][1] = k3_cfg;
    return NGW_OK;
}

int ngw_matvec_isdf_block(ngw_context* ctx, const double* vec, double* hvec, int ncv, int blksz,
                          int tamm_d)
{
    NGW_TRY(check_ctx(ctx));
    if (ncv != ctx->ncv) return fail(NGW_ERR_ARG, "ncv does not match the operator (nval*ncond, or the columns of C)");
    if (blksz < 0 || (blksz > 0 && (!vec || !hvec))) return fail(NGW_ERR_ARG, "bad vectors");
    if (blksz == 0) return NGW_OK;
    const size_t bytes = sizeof(double) * (size_t)ncv * blksz;
    NGW_TRY(ctx->w0.ensure(bytes));
    NGW_TRY(ctx->w1.ensure(bytes));
    NGW_HIP(hipMemcpyAsync(ctx->w0.ptr, vec, bytes, hipMemcpyHostToDevice, ctx->stream));
    NGW_TRY(ngw_matvec_isdf_block_dev(ctx, ctx->w0.as<double>(), ctx->w1.as<double>(), ncv, blksz, tamm_d));
    NGW_HIP(hipMemcpyAsync(hvec, ctx->w1.ptr, bytes, hipMemcpyDeviceToHost, ctx->stream));
    NGW_HIP(hipStreamSynchronize(ctx->stream));
    return NGW_OK;
}

int ngw_polynomial_matvec_isdf_block(ngw_context* ctx, const double* vec, double* hvec, int ncv,
                                     const double* polynomial, int npoly, int blksz, int tamm_d)
{
    NGW_TRY(check_ctx(ctx));
    if (ncv != ctx->ncv) return fail(NGW_ERR_ARG, "ncv does not match the operator (nval*ncond, or the columns of C)");
    if (npoly < 1 || !polynomial) return fail(NGW_ERR_ARG, "npoly must be >= 1");
    if (blksz < 0 || (blksz > 0 && (!vec || !hvec))) return fail(NGW_ERR_ARG, "bad vectors");
    if (blksz == 0) return NGW_OK;
    const size_t bytes = sizeof(double) * (size_t)ncv * blksz;
    NGW_TRY(ctx->w0.ensure(bytes));
    NGW_TRY(ctx->w1.ensure(bytes));
    NGW_TRY(ctx->w2.ensure(bytes));
    NGW_TRY(ctx->w3.ensure(bytes));
    NGW_TRY(ctx->w4.ensure(bytes));
    double* q = ctx->w0.as<double>();
    double* out = ctx->w1.as<double>();
    NGW_HIP(hipMemcpyAsync(q, vec, bytes, hipMemcpyHostToDevice, ctx->stream));
    NGW_TRY(poly_apply(ctx, q, out, ctx->w2.as<double>(), ctx->w3.as<double>(), ctx->w4.as<double>(), ncv, blksz,
                       polynomial, npoly, tamm_d ? 1 : 0, false));
    NGW_HIP(hipMemcpyAsync(hvec, out, bytes, hipMemcpyDeviceToHost, ctx->stream));
    NGW_HIP(hipStreamSynchronize(ctx->stream));
    return NGW_OK;
}

int ngw_lanczos_spectra_isdf_block(ngw_context* ctx, const double* v0, const double* v0_norm,
                                   int ncv, const double* zz, int nz, int niter,
                                   const double* polynomial, int npoly, double* spectra, int blksz,
                                   int tamm_d, double* a_elmt, double* b_elmt)
{
    NGW_TRY(check_ctx(ctx));
    if (ncv != ctx->ncv) return fail(NGW_ERR_ARG, "ncv does not match the operator (nval*ncond, or the columns of C)");
    if (npoly < 1 || !polynomial) return fail(NGW_ERR_ARG, "npoly must be >= 1");
    if (niter < 1 || nz < 0 || blksz < 0) return fail(NGW_ERR_ARG, "bad niter, nz or blksz");
    if (blksz == 0) return NGW_OK;
    if (!v0 || !v0_norm || (nz > 0 && (!zz || !spectra))) return fail(NGW_ERR_ARG, "null array");
    hipStream_t s = ctx->stream;
    const size_t vb = sizeof(double) * (size_t)ncv * blksz;
    NGW_TRY(ctx->w0.ensure(vb));
    NGW_TRY(ctx->w1.ensure(vb));
    NGW_TRY(ctx->w2.ensure(vb));
    NGW_TRY(ctx->w3.ensure(vb));
    NGW_TRY(ctx->w4.ensure(vb));
    NGW_TRY(ctx->ab.ensure(sizeof(double) * 2 * (size_t)niter * blksz));
    NGW_TRY(ctx->zz.ensure(sizeof(double) * 2 * (size_t)std::max(nz, 1) * blksz));
    NGW_TRY(ctx->spec.ensure(sizeof(double) * 2 * (size_t)std::max(nz, 1) * blksz));
    NGW_TRY(ctx->misc.ensure(sizeof(double) * blksz));
    double* q = ctx->w0.as<double>();
    double* w = ctx->w1.as<double>();
    double* da = ctx->ab.as<double>();
    double* db = da + (size_t)niter * blksz;
    NGW_HIP(hipMemcpyAsync(q, v0, vb, hipMemcpyHostToDevice, s));
    NGW_HIP(hipMemcpyAsync(ctx->misc.ptr, v0_norm, sizeof(double) * blksz, hipMemcpyHostToDevice, s));
    if (nz > 0)
        NGW_HIP(hipMemcpyAsync(ctx->zz.ptr, zz, sizeof(double) * 2 * (size_t)nz * blksz,
                               hipMemcpyHostToDevice, s));
    // vec = v0 / v0_norm  (lanczos_spectra_BLOCK.F90:40-42)
    NGW_TRY(scale_columns(s, q, q, ncv, ncv, blksz, ctx->misc.as<double>()));
    NGW_TRY(lanczos_run(ctx, q, w, ctx->w2.as<double>(), ctx->w3.as<double>(), ctx->w4.as<double>(), ncv, blksz,
                        niter, polynomial, npoly, tamm_d ? 1 : 0, da, db));
    if (nz > 0) {
        NGW_TRY(continued_fraction(s, da, db, niter, blksz, ctx->zz.as<double>(), nz,
                                   ctx->spec.as<double>()));
        NGW_HIP(hipMemcpyAsync(spectra, ctx->spec.ptr, sizeof(double) * 2 * (size_t)nz * blksz,
                               hipMemcpyDeviceToHost, s));
    }
    if (a_elmt)
        NGW_HIP(hipMemcpyAsync(a_elmt, da, sizeof(double) * (size_t)niter * blksz, hipMemcpyDeviceToHost, s));
    if (b_elmt)
        NGW_HIP(hipMemcpyAsync(b_elmt, db, sizeof(double) * (size_t)niter * blksz, hipMemcpyDeviceToHost, s));
    NGW_HIP(hipStreamSynchronize(s));
    return NGW_OK;
}

int ngw_sigma_lanczos(ngw_context* ctx, int nqp, const int* qp_band, int nen, double deltae,
                      double ecuts, int niter, const double* polynomial, int npoly, int tamm_d,
                      int max_cols, double* sigma_c)
{
    NGW_TRY(check_factored(ctx));
    return ngw_sigma_lanczos_bands(ctx, nqp, qp_band, 0, ctx->nband, nen, deltae, ecuts, niter, polynomial, npoly,
                                   tamm_d, max_cols, sigma_c);
}

int ngw_sigma_lanczos_bands(ngw_context* ctx, int nqp, const int* qp_band, int band_lo, int band_hi, int nen,
                            double deltae, double ecuts, int niter, const double* polynomial, int npoly, int tamm_d,
                            int max_cols, double* sigma_c)
{
    NGW_TRY(check_factored(ctx));
    if (nqp <= 0 || !qp_band || !sigma_c) return fail(NGW_ERR_ARG, "need nqp > 0 and arrays");
    if (nen < 2 || niter < 1 || npoly < 1 || !polynomial) return fail(NGW_ERR_ARG, "bad nen, niter or npoly");
    if (band_lo < 0 || band_hi > ctx->nband || band_lo >= band_hi)
        return fail(NGW_ERR_ARG, "need 0 <= band_lo < band_hi <= nband");
    for (int i = 0; i < nqp; ++i)
        if (qp_band[i] < 0 || qp_band[i] >= ctx->nband) return fail(NGW_ERR_ARG, "qp_band out of range");
    hipStream_t s = ctx->stream;
    const int nb = band_hi - band_lo;
    const long ncv = ctx->ncv;
    const long total = (long)nqp * nb;
    const int nz = nen + 1;
    // batch size: keep the four Krylov vectors of a batch under ~40% of free HBM
    long cols = max_cols > 0 ? max_cols : total;
    if (max_cols <= 0) {
        size_t freeb = 0, totalb = 0;
        NGW_HIP(hipMemGetInfo(&freeb, &totalb));
        const double per_col = 8.0 * (5.0 * ncv + (k1_slots(ctx->nval) + 2.0) * ctx->n_intp);
        long fit = (long)(0.4 * (double)freeb / per_col);
        cols = std::max(1L, std::min(total, fit));
    }
    cols = std::min(cols, total);
    // column -> (m, n) maps for all columns
    std::vector<int> mband(total), nband(total);
    for (int iq = 0; iq < nqp; ++iq)
        for (int n = 0; n < nb; ++n) {
            mband[(long)iq * nb + n] = qp_band[iq];
            nband[(long)iq * nb + n] = band_lo + n;
        }
    const size_t vb = sizeof(double) * (size_t)ncv * cols;
    NGW_TRY(ctx->w0.ensure(vb));
    NGW_TRY(ctx->w1.ensure(vb));
    NGW_TRY(ctx->w2.ensure(vb));
    NGW_TRY(ctx->w3.ensure(vb));
    NGW_TRY(ctx->w4.ensure(vb));
    NGW_TRY(ctx->ab.ensure(sizeof(double) * 2 * (size_t)niter * cols));
    NGW_TRY(ctx->zz.ensure(sizeof(double) * 2 * (size_t)nz * total));
    NGW_TRY(ctx->spec.ensure(sizeof(double) * 2 * (size_t)nz * total));
    NGW_TRY(ctx->misc.ensure(sizeof(double) * 2 * total + sizeof(int) * 2 * total));
    NGW_TRY(ctx->misc2.ensure(sizeof(double) * 2 * (size_t)nen * nqp));
    double* pn2 = ctx->misc.as<double>();
    double* pn = pn2 + total;
    int* dmb = reinterpret_cast<int*>(pn + total);
    int* dnb = dmb + total;
    NGW_HIP(hipMemcpyAsync(dmb, mband.data(), sizeof(int) * total, hipMemcpyHostToDevice, s));
    NGW_HIP(hipMemcpyAsync(dnb, nband.data(), sizeof(int) * total, hipMemcpyHostToDevice, s));
    double* zzd = ctx->zz.as<double>();
    double* spec = ctx->spec.as<double>();
    NGW_TRY(zz_grid(s, ctx->e1.as<double>(), dmb, dnb, ctx->nval, (int)total, nen, deltae, ecuts, zzd));
    double* q = ctx->w0.as<double>();
    double* w = ctx->w1.as<double>();
    double* da = ctx->ab.as<double>();
    for (long c0 = 0; c0 < total; c0 += cols) {
        const int nc = (int)std::min(cols, total - c0);
        double* db = da + (size_t)niter * nc;
        // P vectors (BLOCK driver :286-358) straight into the Lanczos start
        NGW_TRY(pvectors_f64(ctx, dmb + c0, dnb + c0, nc, tamm_d ? 1 : 0, q, ncv));
        NGW_TRY(column_norms(s, q, ncv, ncv, nc, pn2 + c0, pn + c0));
        NGW_TRY(scale_columns(s, q, q, ncv, ncv, nc, pn + c0));
        NGW_TRY(lanczos_run(ctx, q, w, ctx->w2.as<double>(), ctx->w3.as<double>(), ctx->w4.as<double>(), ncv, nc,
                            niter, polynomial, npoly, tamm_d ? 1 : 0, da, db));
        ProfScope ps(ctx, KC_CFRAC);
        NGW_TRY(continued_fraction(s, da, db, niter, nc, zzd + 2 * (size_t)nz * c0, nz,
                                   spec + 2 * (size_t)nz * c0));
    }
    {
        ProfScope ps(ctx, KC_SIGMA);
        NGW_TRY(sigma_assemble(s, spec, zzd, pn2, nen, nqp, nb, band_lo, ctx->nval, tamm_d ? 1 : 0,
                               ctx->misc2.as<double>()));
    }
    NGW_HIP(hipMemcpyAsync(sigma_c, ctx->misc2.ptr, sizeof(double) * 2 * (size_t)nen * nqp,
                           hipMemcpyDeviceToHost, s));
    NGW_HIP(hipStreamSynchronize(s));
    return NGW_OK;
}

int ngw_sigma_x_isdf(ngw_context* ctx, int nqp, const int* qp_band, double* sigx)
{
    NGW_TRY(check_factored(ctx));
    if (nqp <= 0 || !qp_band || !sigx) return fail(NGW_ERR_ARG, "need nqp > 0 and arrays");
    for (int i = 0; i < nqp; ++i)
        if (qp_band[i] < 0 || qp_band[i] >= ctx->nband) return fail(NGW_ERR_ARG, "qp_band out of range");
    hipStream_t s = ctx->stream;
    NGW_TRY(ctx->staging.ensure(sizeof(double) * nqp + sizeof(int) * nqp + 256));
    double* dsig = ctx->staging.as<double>();
    int* dqp = reinterpret_cast<int*>(dsig + nqp);
    NGW_HIP(hipMemcpyAsync(dqp, qp_band, sizeof(int) * nqp, hipMemcpyHostToDevice, s));
    NGW_TRY(sigma_x_isdf(ctx, nqp, dqp, dsig));
    NGW_HIP(hipMemcpyAsync(sigx, dsig, sizeof(double) * nqp, hipMemcpyDeviceToHost, s));
    NGW_HIP(hipStreamSynchronize(s));
    return NGW_OK;
}

int ngw_qp_newton(ngw_context* ctx, int nqp, int nen, double deltae_ev, const double* edft,
                  const double* en0, const double* sigx, const double* vxc, const double* sigma_c,
                  double* eqp, double* zfac, double* sigma0, double* imsig, int* ierr)
{
    if (!ctx) return fail(NGW_ERR_ARG, "null context");
    if (nqp <= 0 || nen < 4) return fail(NGW_ERR_ARG, "need nqp > 0 and nen >= 4");
    if (!edft || !en0 || !sigx || !vxc || !sigma_c || !eqp || !zfac || !sigma0 || !imsig || !ierr)
        return fail(NGW_ERR_ARG, "null array");
    NGW_HIP(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    std::vector<double> in(4 * (size_t)nqp + 2 * (size_t)nen * nqp);
    for (int i = 0; i < nqp; ++i) {
        in[4 * i] = edft[i];
        in[4 * i + 1] = en0[i];
        in[4 * i + 2] = sigx[i];
        in[4 * i + 3] = vxc[i];
    }
    std::memcpy(&in[4 * (size_t)nqp], sigma_c, sizeof(double) * 2 * (size_t)nen * nqp);
    NGW_TRY(ctx->staging.ensure(sizeof(double) * (in.size() + 6 * (size_t)nqp + 5 * (size_t)nen * nqp)));
    double* din = ctx->staging.as<double>();
    double* dout = din + in.size();
    double* scr = dout + 6 * (size_t)nqp;
    NGW_HIP(hipMemcpyAsync(din, in.data(), sizeof(double) * in.size(), hipMemcpyHostToDevice, s));
    NGW_TRY(qp_newton(s, nqp, nen, deltae_ev, din, dout, scr));
    std::vector<double> out(6 * (size_t)nqp);
    NGW_HIP(hipMemcpyAsync(out.data(), dout, sizeof(double) * out.size(), hipMemcpyDeviceToHost, s));
    NGW_HIP(hipStreamSynchronize(s));
    for (int i = 0; i < nqp; ++i) {
        eqp[i] = out[6 * i];
        zfac[i] = out[6 * i + 1];
        sigma0[2 * i] = out[6 * i + 2];
        sigma0[2 * i + 1] = out[6 * i + 3];
        imsig[i] = out[6 * i + 4];
        ierr[i] = (int)out[6 * i + 5];
    }
    return NGW_OK;
}

int ngw_polyfit_sqrt(double xmin, double xmax, int npoly, int nx, int chebyshev, double* beta,
                     double* fit_error)
{
    if (!beta || npoly < 0 || nx < npoly + 1 || nx < 2) return fail(NGW_ERR_ARG, "bad npoly/nx");
    if (xmin >= xmax) std::swap(xmin, xmax);
    if (xmin <= 0.0 || xmin <= 1e-4) return fail(NGW_ERR_RANGE, "Polyfit_sqrt xmin is too small");
    const int np = npoly + 1;
    std::vector<double> xx(nx), yy(nx), pxx((size_t)nx * np), M((size_t)np * np), y(np);
    const double pi = 3.1415926535897932384626433832795;
    const double emid = 0.5 * (xmin + xmax), scal = 0.5 * (xmax - xmin);
    const double dx = chebyshev ? pi / (nx - 1) : (xmax - xmin) / (nx - 1);
    for (int ix = 0; ix < nx; ++ix) {
        xx[ix] = chebyshev ? emid - scal * std::cos(dx * ix) : xmin + dx * ix;
        yy[ix] = std::sqrt(xx[ix]);
        pxx[ix] = 1.0;
        for (int ip = 1; ip < np; ++ip) pxx[ix + (size_t)ip * nx] = std::pow(xx[ix], (double)ip);
    }
    for (int i = 0; i < np; ++i) {
        for (int j = 0; j < np; ++j) {
            double sacc = 0.0;
            for (int ix = 0; ix < nx; ++ix) sacc += pxx[ix + (size_t)i * nx] * pxx[ix + (size_t)j * nx];
            M[i + (size_t)j * np] = sacc;
        }
        double sacc = 0.0;
        for (int ix = 0; ix < nx; ++ix) sacc += pxx[ix + (size_t)i * nx] * yy[ix];
        y[i] = sacc;
    }
    solve_equilibrated(np, M, y, beta);
    if (fit_error) {
        const int mx = 2000;
        double fe = 0.0;
        for (int ix = 0; ix < mx; ++ix) {
            double tx = xmin + ix * (xmax - xmin) / (mx - 1), tp = 1.0, fy = 0.0;
            for (int ip = 0; ip < np; ++ip) {
                fy += tp * beta[ip];
                if (ip < np - 1) tp *= tx;
            }
            fe += std::fabs((std::sqrt(tx) - fy) / std::sqrt(tx)) / (double)mx;
        }
        *fit_error = fe;
    }
    return NGW_OK;
}

int ngw_gemm_f64_dev(ngw_context* ctx, int m, int n, int k, const double* d_a, long lda,
                     const double* d_b, long ldb, double* d_c, long ldc)
{
    if (!ctx) return fail(NGW_ERR_ARG, "null context");
    if (m < 0 || n < 0 || k < 0 || lda < m || ldb < k || ldc < m) return fail(NGW_ERR_ARG, "bad GEMM shape");
    NGW_HIP(hipSetDevice(ctx->device));
    NGW_TRY(gemm_f64(ctx->stream, m, n, k, d_a, lda, d_b, ldb, d_c, ldc));
    NGW_HIP(hipStreamSynchronize(ctx->stream));
    return NGW_OK;
}

int ngw_gemm_lowp_dev(ngw_context* ctx, int kind, int m, int n, int k, const double* d_a, long lda,
                      const double* d_b, long ldb, double* d_c, long ldc)
{
    if (!ctx) return fail(NGW_ERR_ARG, "null context");
    if (m < 0 || n < 0 || k < 0 || lda < m || ldb < k || ldc < m) return fail(NGW_ERR_ARG, "bad GEMM shape");
    NGW_HIP(hipSetDevice(ctx->device));
    NGW_TRY(gemm_lowp(ctx->stream, kind, m, n, k, d_a, lda, d_b, ldb, d_c, ldc));
    NGW_HIP(hipStreamSynchronize(ctx->stream));
    return NGW_OK;
}

int ngw_profile_enable(ngw_context* ctx, int on)
{
    if (!ctx) return fail(NGW_ERR_ARG, "null context");
    ctx->prof_on = on != 0;
    return NGW_OK;
}

int ngw_profile_reset(ngw_context* ctx)
{
    if (!ctx) return fail(NGW_ERR_ARG, "null context");
    prof_drain(ctx);
    for (int k = 0; k < KC_COUNT; ++k) {
        ctx->prof_ms[k] = 0.0;
        ctx->prof_cnt[k] = 0;
    }
    return NGW_OK;
}

int ngw_profile_read(ngw_context* ctx, int kernel_class, double* total_ms, long* launches)
{
    if (!ctx || !total_ms || !launches) return fail(NGW_ERR_ARG, "null argument");
    if (kernel_class < 0 || kernel_class >= KC_COUNT) return fail(NGW_ERR_ARG, "bad kernel class");
    prof_drain(ctx);
    *total_ms = ctx->prof_ms[kernel_class];
    *launches = ctx->prof_cnt[kernel_class];
    return NGW_OK;
}

}  // extern "C"
