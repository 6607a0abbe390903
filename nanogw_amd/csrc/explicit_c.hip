// The Casida operator with the pair matrix C given explicitly, the form of
// NanoGW's single-vector routines matvec_isdf (shared/apply_matvec.f90:25-140)
// and the polynomial / Lanczos drivers built on it: the caller passes
// tmpCmtrx(n_intp, ncv) with C(mu, icv) = Psi_v(mu) Psi_c(mu), and
//   y = R^3/2 (R^1/2 x) + 4 R^1/2 C^T M C (R^1/2 x)       (full, :126-131)
//   y = R x + 2 C^T M C x                                 (TDA)
// Also the factored operator on one vector (operator_f64_one, below).
// The three products are library GEMMs (rocBLAS dgemm) for several vectors;
// for one vector (the reference's three dgemv calls) C x and M t run on the
// column-split GEMV kernels below and C^T u on rocBLAS dgemv.  The diagonal
// term and the Horner / Lanczos combination are the same epilogue as the
// factored operator.  FP64 only.

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

}  // namespace

// y = A x for a column-major m x n matrix A with m small against n (C x
// and M t of one vector; M t is also the factored operator's one-vector K2): workgroup (row block rb, column chunk cc) sums its 256 rows over
// its columns, reading each column segment as one coalesced run; part(cc, i)
// holds the chunk sums, added in chunk order by gemv_wide_reduce (the
// result does not depend on scheduling).
constexpr int GV_MAX_CHUNKS = 1024;
__global__ __launch_bounds__(256) void gemv_wide_partial(const double* A, long lda, int m, int n, const double* x,
                                                         int cw, double* part)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int cc = blockIdx.y;
    const int j0 = cc * cw, j1 = min(n, j0 + cw);
    if (i >= m) return;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int j = j0;
    for (; j + 3 < j1; j += 4) {
        s0 += A[i + (long)j * lda] * x[j];
        s1 += A[i + (long)(j + 1) * lda] * x[j + 1];
        s2 += A[i + (long)(j + 2) * lda] * x[j + 2];
        s3 += A[i + (long)(j + 3) * lda] * x[j + 3];
    }
    for (; j < j1; ++j) s0 += A[i + (long)j * lda] * x[j];
    part[(long)cc * m + i] = (s0 + s1) + (s2 + s3);
}

// 64 rows per workgroup; wave g sums chunks g, g + 4, g + 8, ... (four
// loads in flight per thread), and the four wave sums are added in LDS.
__global__ __launch_bounds__(256) void gemv_wide_reduce(const double* part, int m, int nchunk, double* y)
{
    __shared__ double red[4][64];
    const int r = threadIdx.x & 63, g = threadIdx.x >> 6;
    const int i = blockIdx.x * 64 + r;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    if (i < m) {
        int c = g;
        for (; c + 12 < nchunk; c += 16) {
            s0 += part[(long)c * m + i];
            s1 += part[(long)(c + 4) * m + i];
            s2 += part[(long)(c + 8) * m + i];
            s3 += part[(long)(c + 12) * m + i];
        }
        for (; c < nchunk; c += 4) s0 += part[(long)c * m + i];
    }
    red[g][r] = (s0 + s1) + (s2 + s3);
    __syncthreads();
    if (g == 0 && i < m) y[i] = (red[0][r] + red[1][r]) + (red[2][r] + red[3][r]);
}

int gemv_wide(hipStream_t s, const double* A, long lda, int m, int n, const double* x, double* y, DevBuf& part)
{
    const int rblk = (m + 255) / 256;
    const int cap = std::max(1, std::min(GV_MAX_CHUNKS, 2048 / rblk));  // depends on m only: one size per operator
    int nchunk = std::max(1, std::min(cap, (n + 31) / 32));
    const int cw = (n + nchunk - 1) / nchunk;
    nchunk = (n + cw - 1) / cw;
    NGW_TRY(part.ensure(sizeof(double) * (size_t)cap * m));
    hipLaunchKernelGGL(gemv_wide_partial, dim3(rblk, nchunk), dim3(256), 0, s, A, lda, m, n, x, cw,
                       part.as<double>());
    hipLaunchKernelGGL(gemv_wide_reduce, dim3((m + 63) / 64), dim3(256), 0, s, part.as<const double>(), m, nchunk, y);
    NGW_HIP(hipGetLastError());
    return NGW_OK;
}

namespace {

int rb(rocblas_status st, const char* what)
{
    if (st != rocblas_status_success)
        return fail(NGW_ERR_HIP, std::string(what) + " failed (rocblas status " + std::to_string((int)st) + ")");
    return NGW_OK;
}

// The context's rocBLAS handle (created on first use) bound to its stream.
// y(idx) = sum of the S chunk products (chunk order), then the epilogue
__global__ void chunk_sum_epilogue_kernel(OperatorEpilogue ep, const double* part, int S, long ncv)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < ncv; i += (long)gridDim.x * blockDim.x) {
        double y = 0.0;
        for (int c = 0; c < S; ++c) y += part[(long)c * ncv + i];
        ep.apply(i, 0, y);
    }
}

rocblas_handle blas(ngw_context* ctx, int& st)
{
    st = NGW_OK;
    if (!ctx->rb_handle) {
        rocblas_handle h = nullptr;
        st = rb(rocblas_create_handle(&h), "rocblas_create_handle");
        if (st != NGW_OK) return nullptr;
        ctx->rb_handle = h;
    }
    rocblas_handle h = static_cast<rocblas_handle>(ctx->rb_handle);
    st = rb(rocblas_set_stream(h, ctx->stream), "rocblas_set_stream");
    return h;
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
    int st = NGW_OK;
    rocblas_handle h = blas(ctx, st);
    NGW_TRY(st);
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
    // t = C xs, u = M t, y = C^T u: GEMV for one vector (the library's GEMM
    // kernels tile n = 1 as a full 128-wide block; its plain GEMV splits only
    // the n_intp rows of the wide C x over workgroups), GEMM otherwise
    const double* M = ctx->mmat.as<double>();
    if (B == 1) {
        NGW_TRY(gemv_wide(s, C, nmu, nmu, (int)ncv, xs, t, ctx->gv_part));
        NGW_TRY(gemv_wide(s, M, nmu, nmu, nmu, t, u, ctx->gv_part));
        NGW_TRY(rb(rocblas_dgemv(h, rocblas_operation_transpose, nmu, (int)ncv, &one, C, nmu, u, 1, &zero, y, 1),
                   "dgemv C^T u"));
    } else {
        NGW_TRY(rb(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, nmu, B, (int)ncv, &one, C, nmu,
                                 xs, (int)ncv, &zero, t, nmu), "dgemm C x"));
        NGW_TRY(rb(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, nmu, B, nmu, &one, M, nmu, t,
                                 nmu, &zero, u, nmu), "dgemm M t"));
        NGW_TRY(rb(rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, (int)ncv, B, nmu, &one, C,
                                 nmu, u, nmu, &zero, y, (int)ncv), "dgemm C^T u"));
    }
    hipLaunchKernelGGL(explicit_epilogue_kernel, dim3(2048), dim3(256), 0, s, a.epilogue(ctx), y, ncv, B);
    NGW_HIP(hipGetLastError());
    return NGW_OK;
}

namespace {

// t(mu) = sum_v Psi_v(mu, v) T1(mu, v)
__global__ void rowdot_kernel(const double* psi, long ldp, const double* t1, int nmu, int nv, double* t)
{
    const int mu = blockIdx.x * 256 + threadIdx.x;
    if (mu >= nmu) return;
    double s = 0.0;
    for (int v = 0; v < nv; ++v) s += psi[mu + (long)v * ldp] * t1[mu + (long)v * nmu];
    t[mu] = s;
}

// W(mu, v) = Psi_v(mu, v) u(mu)
__global__ void weight_kernel(const double* psi, long ldp, const double* u, int nmu, int nv, double* w)
{
    const long n = (long)nmu * nv;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
        const int v = (int)(e / nmu), mu = (int)(e - (long)v * nmu);
        w[e] = psi[mu + (long)v * ldp] * u[mu];
    }
}

}  // namespace

// The factored operator on one vector.  The MFMA tiles of operator_f64 would
// run nearly empty at N = nv columns, so the products are regrouped into
// library GEMMs with the pair index split as (v, c), icv = v + nv c:
//   T1 = Psi_c X^T          (n_intp x nv;  X(v, c) = (s o x)(icv))
//   t(mu) = sum_v Psi_v(mu, v) T1(mu, v)             (= C (s o x))
//   u = M t                                           (column-split GEMV)
//   Y^T = (Psi_v o u)^T Psi_c   (nv x nc, Y^T(v, c) = (C^T u)(icv)),
//         split over n_intp into S chunks (one strided-batched GEMM: the
//         nv x nc output alone is too few tiles to fill the GPU)
// then the chunk sum and the shared epilogue.
int operator_f64_one(ngw_context* ctx, const OperatorArgs& a)
{
    hipStream_t s = ctx->stream;
    const int nmu = ctx->n_intp, nv = ctx->nval, nc = ctx->ncond;
    const long ncv = ctx->ncv, ldp = ctx->ld_psi;
    int st = NGW_OK;
    rocblas_handle h = blas(ctx, st);
    NGW_TRY(st);
    NGW_TRY(ctx->ex_x.ensure(sizeof(double) * (size_t)ncv));
    NGW_TRY(ctx->ex_t.ensure(sizeof(double) * (2 * (size_t)nmu * nv + 2 * (size_t)nmu)));
    double* xs = ctx->ex_x.as<double>();
    double* t1 = ctx->ex_t.as<double>();  // n_intp x nv
    double* t = t1 + (size_t)nmu * nv;
    double* u = t + nmu;
    double* w = u + nmu;
    const double* psi = ctx->psi.as<double>();
    const double* psic = psi + (size_t)nv * ldp;
    const double one = 1.0, zero = 0.0;
    hipLaunchKernelGGL(scale_in_kernel, dim3(1024), dim3(256), 0, s, a.x, a.ldx,
                       a.tda ? nullptr : ctx->sqrtr.as<double>(), ncv, 1, xs);
    {
        ProfScope ps(ctx, KC_K1);
        NGW_TRY(rb(rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_transpose, nmu, nv, nc, &one, psic,
                                 (int)ldp, xs, nv, &zero, t1, nmu), "dgemm Psi_c X^T"));
        hipLaunchKernelGGL(rowdot_kernel, dim3((nmu + 255) / 256), dim3(256), 0, s, psi, ldp, t1, nmu, nv, t);
    }
    {
        ProfScope ps(ctx, KC_K2);
        NGW_TRY(gemv_wide(s, ctx->mmat.as<double>(), ctx->ld_m, nmu, nmu, t, u, ctx->gv_part));
    }
    {
        ProfScope ps(ctx, KC_K3);
        hipLaunchKernelGGL(weight_kernel, dim3(std::min((nmu * (long)nv + 255) / 256, 2048L)), dim3(256), 0, s, psi,
                           ldp, u, nmu, nv, w);
        const int S = std::max(1, std::min(16, nmu / 256));
        const int kc = nmu / S, rem = nmu - S * kc;
        NGW_TRY(ctx->ex_s.ensure(sizeof(double) * (size_t)S * ncv));
        double* part = ctx->ex_s.as<double>();
        NGW_TRY(rb(rocblas_dgemm_strided_batched(h, rocblas_operation_transpose, rocblas_operation_none, nv, nc, kc,
                                                 &one, w, nmu, (rocblas_stride)kc, psic, (int)ldp, (rocblas_stride)kc,
                                                 &zero, part, nv, (rocblas_stride)ncv, S),
                   "dgemm W^T Psi_c (chunks)"));
        if (rem > 0)
            NGW_TRY(rb(rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, nv, nc, rem, &one,
                                     w + (size_t)S * kc, nmu, psic + (size_t)S * kc, (int)ldp, &one,
                                     part + (size_t)(S - 1) * ncv, nv), "dgemm W^T Psi_c (rest)"));
        hipLaunchKernelGGL(chunk_sum_epilogue_kernel, dim3(1024), dim3(256), 0, s, a.epilogue(ctx), part, S, ncv);
    }
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
    ctx->ex_s.release();
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
