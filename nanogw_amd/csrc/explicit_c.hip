// The Casida operator with the pair matrix C given explicitly, the form of
// NanoGW's single-vector routines matvec_isdf (shared/apply_matvec.f90:25-140)
// and the polynomial / Lanczos drivers built on it: the caller passes
// tmpCmtrx(n_intp, ncv) with C(mu, icv) = Psi_v(mu) Psi_c(mu), and
//   y = R^3/2 (R^1/2 x) + 4 R^1/2 C^T M C (R^1/2 x)       (full, :126-131)
//   y = R x + 2 C^T M C x                                 (TDA)
// The three products are plain library GEMMs (rocBLAS dgemm: for one vector
// they are the reference's dgemv calls); the diagonal term and the Horner /
// Lanczos combination are the same epilogue as the factored operator.
// FP64 only.

#include "ngw_kernels.h"

#include <rocblas/rocblas.h>

#include <cmath>
#include <vector>

namespace ngw {

namespace {

__global__ void scale_in_kernel(const double* x, long ldx, const double* s, long ncv, int nvec, double* xs)
{
    const long total = ncv * nvec;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
        const int b = (int)(t / ncv);
        const long i = t - (long)b * ncv;
        const double v = x[i + (long)b * ldx];
        xs[t] = s ? s[i] * v : v;
    }
}

__global__ void explicit_epilogue_kernel(OperatorEpilogue ep, const double* y, long ncv, int nvec)
{
    const long total = ncv * nvec;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
        const int b = (int)(t / ncv);
        const long i = t - (long)b * ncv;
        ep.apply(i, b, y[t]);
    }
}

int rb(rocblas_status st, const char* what)
{
    if (st != rocblas_status_success)
        return fail(NGW_ERR_HIP, std::string(what) + " failed (rocblas status " + std::to_string((int)st) + ")");
    return NGW_OK;
}

}  // namespace

int operator_explicit(ngw_context* ctx, const OperatorArgs& a)
{
    const int B = a.nvec;
    if (B <= 0) return NGW_OK;
    if (a.flags) return fail(NGW_ERR_STATE, "explicit-C operator: FP64 vector blocks only");
    hipStream_t s = ctx->stream;
    const long ncv = ctx->ncv;
    const int nmu = ctx->n_intp;
    if (!ctx->rb_handle) {
        rocblas_handle h = nullptr;
        NGW_TRY(rb(rocblas_create_handle(&h), "rocblas_create_handle"));
        ctx->rb_handle = h;
    }
    rocblas_handle h = static_cast<rocblas_handle>(ctx->rb_handle);
    NGW_TRY(rb(rocblas_set_stream(h, s), "rocblas_set_stream"));
    NGW_TRY(ctx->ex_x.ensure(sizeof(double) * (size_t)ncv * B));
    NGW_TRY(ctx->ex_y.ensure(sizeof(double) * (size_t)ncv * B));
    NGW_TRY(ctx->ex_t.ensure(sizeof(double) * 2 * (size_t)nmu * B));
    double* xs = ctx->ex_x.as<double>();
    double* y = ctx->ex_y.as<double>();
    double* t = ctx->ex_t.as<double>();
    double* u = t + (size_t)nmu * B;
    hipLaunchKernelGGL(scale_in_kernel, dim3(2048), dim3(256), 0, s, a.x, a.ldx,
                       a.tda ? nullptr : ctx->sqrtr.as<double>(), ncv, B, xs);
    NGW_HIP(hipGetLastError());
    const double one = 1.0, zero = 0.0;
    const double* C = ctx->cexp.as<double>();
    // t = C xs, u = M t, y = C^T u
    NGW_TRY(rb(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, nmu, B, (int)ncv, &one, C, nmu, xs,
                             (int)ncv, &zero, t, nmu), "dgemm C x"));
    NGW_TRY(rb(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, nmu, B, nmu, &one,
                             ctx->mmat.as<double>(), nmu, t, nmu, &zero, u, nmu), "dgemm M t"));
    NGW_TRY(rb(rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, (int)ncv, B, nmu, &one, C, nmu,
                             u, nmu, &zero, y, (int)ncv), "dgemm C^T u"));
    hipLaunchKernelGGL(explicit_epilogue_kernel, dim3(2048), dim3(256), 0, s, a.epilogue(ctx), y, ncv, B);
    NGW_HIP(hipGetLastError());
    return NGW_OK;
}

void explicit_release(ngw_context* ctx)
{
    if (ctx->rb_handle) (void)rocblas_destroy_handle(static_cast<rocblas_handle>(ctx->rb_handle));
    ctx->rb_handle = nullptr;
    ctx->cexp.release();
    ctx->ex_x.release();
    ctx->ex_y.release();
    ctx->ex_t.release();
    ctx->explicit_c = false;
}

}  // namespace ngw

using namespace ngw;

extern "C" int ngw_set_explicit_c(ngw_context* ctx, int n_intp, int ncv, const double* cmtrx, int ld_c,
                                  const double* mmtrx, int ld_m, const double* sqrtr)
{
    if (!ctx) return fail(NGW_ERR_ARG, "null context");
    if (n_intp <= 0 || ncv <= 0 || ld_c < n_intp || ld_m < n_intp || !cmtrx || !mmtrx || !sqrtr)
        return fail(NGW_ERR_ARG, "bad sizes or null arrays");
    for (int i = 0; i < ncv; ++i)
        if (!(sqrtr[i] > 0.0)) return fail(NGW_ERR_ARG, "sqrtR must be positive");
    NGW_HIP(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    ctx->n_intp = n_intp;
    ctx->ncv = ncv;
    ctx->nband = 0;
    ctx->nval = 0;
    ctx->ncond = 0;
    ctx->ld_m = n_intp;
    ctx->lowp_ready = false;
    ctx->lp2_ready = false;
    ctx->precision = PREC_F64;
    NGW_TRY(ctx->cexp.ensure(sizeof(double) * (size_t)n_intp * ncv));
    NGW_HIP(hipMemcpy2DAsync(ctx->cexp.ptr, sizeof(double) * n_intp, cmtrx, sizeof(double) * ld_c,
                             sizeof(double) * n_intp, ncv, hipMemcpyHostToDevice, s));
    NGW_TRY(ctx->mmat.ensure(sizeof(double) * (size_t)n_intp * n_intp));
    NGW_HIP(hipMemcpy2DAsync(ctx->mmat.ptr, sizeof(double) * n_intp, mmtrx, sizeof(double) * ld_m,
                             sizeof(double) * n_intp, n_intp, hipMemcpyHostToDevice, s));
    std::vector<double> d3(ncv), d2(ncv);
    for (int i = 0; i < ncv; ++i) {
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
    ctx->explicit_c = true;
    return NGW_OK;
}
