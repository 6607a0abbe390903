// The Casida operator with the pair matrix C given explicitly, the form of
// NanoGW's single-vector routines matvec_isdf (shared/apply_matvec.f90:25-140)
// and the polynomial / Lanczos drivers built on it: the caller passes
// tmpCmtrx(n_intp, ncv) with C(mu, icv) = Psi_v(mu) Psi_c(mu), and
//   y = R^3/2 (R^1/2 x) + 4 R^1/2 C^T M C (R^1/2 x)       (full, :126-131)
//   y = R x + 2 C^T M C x                                 (TDA)
// All products are this library's FP64 kernels: for one vector (the
// reference's three dgemv calls) C x and M t on the column-split GEMV and
// C^T u on a one-wave-per-column GEMV; for several vectors the FP64 MFMA
// GEMMs (C x and M t with their sums split into ranges, the range partials
// added in order).  The diagonal term and the Horner / Lanczos combination
// are the same epilogue as the factored operator.  FP64 only.

#include "ngw_kernels.h"

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

// y(j) = sum_i A(i, j) x(i) for column-major A (m x n), one wave per column
// (a coalesced run down the column), lanes combined by a fixed butterfly.
__global__ __launch_bounds__(256) void gemv_t_kernel(const double* A, long lda, int m, long n, const double* x,
                                                     double* y)
{
    const int lane = threadIdx.x & 63;
    const long nw = (long)gridDim.x * 4;
    for (long j = blockIdx.x * 4L + (threadIdx.x >> 6); j < n; j += nw) {
        const double* a = A + j * lda;
        double s0 = 0.0, s1 = 0.0;
        int i = lane;
        for (; i + 64 < m; i += 128) {
            s0 += a[i] * x[i];
            s1 += a[i + 64] * x[i + 64];
        }
        if (i < m) s0 += a[i] * x[i];
        double t = s0 + s1;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
        if (lane == 0) y[j] = t;
    }
}

int operator_explicit(ngw_context* ctx, const OperatorArgs& a)
{
    const int B = a.nvec;
    if (B <= 0) return NGW_OK;
    if (a.flags) return fail(NGW_ERR_STATE, "explicit-C operator: FP64 vector blocks only");
    hipStream_t s = ctx->stream;
    const long ncv = ctx->ncv;
    const int nmu = ctx->n_intp;
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
    const double* C = ctx->cexp.as<double>();
    const double* M = ctx->mmat.as<double>();
    if (B == 1) {
        NGW_TRY(gemv_wide(s, C, nmu, nmu, (int)ncv, xs, t, ctx->gv_part));
        NGW_TRY(gemv_wide(s, M, nmu, nmu, nmu, t, u, ctx->gv_part));
        hipLaunchKernelGGL(gemv_t_kernel, dim3((unsigned)std::min<long>((ncv + 3) / 4, 4096)), dim3(256), 0, s, C,
                           (long)nmu, nmu, ncv, u, y);
    } else {
        NGW_TRY(gemm_f64_split(s, nmu, B, (int)ncv, C, nmu, xs, ncv, t, nmu, ctx->ex_s));
        NGW_TRY(gemm_f64_split(s, nmu, B, nmu, M, nmu, t, nmu, u, nmu, ctx->ex_s));
        NGW_TRY(gemm_f64_tn(s, (int)ncv, B, nmu, 1.0, C, nmu, u, nmu, y, ncv));
    }
    hipLaunchKernelGGL(explicit_epilogue_kernel, dim3(2048), dim3(256), 0, s, a.epilogue(ctx), y, ncv, B);
    NGW_HIP(hipGetLastError());
    return NGW_OK;
}

void explicit_release(ngw_context* ctx)
{
    ctx->cexp.release();
    ctx->ex_x.release();
    ctx->ex_y.release();
    ctx->ex_t.release();
    ctx->ex_s.release();
    ctx->explicit_c = false;
    ctx->blocked = false;
    ctx->blk_nv.clear();
    ctx->blk_nc.clear();
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
