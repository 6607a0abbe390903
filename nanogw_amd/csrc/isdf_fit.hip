// ISDF interpolation vectors on the device: the zeta solve of
// isdf_parallel_sym_UltraLowMem (shared/isdf_parallel_sym_UltraLowMem.F90:
// 448-642), where the reference forms P and Q by GEMMs, A = P o Q at the
// interpolation points and B = P o Q on the grid, and solves A zeta^T = B^T
// with ScaLAPACK pdgesv.
//
// Here, with Psi(r, n) on the grid (device, column-major) and the point
// indices r_mu:
//   G(n, mu)  = Psi(r_mu, n)                       gather, nband x n_intp
//   A         = (G_v^T G_v) o (G^T G)  (+ lambda tr(A)/n_intp I)
//   A^-1      by Cholesky (rocSOLVER potrf + potri; LU getrf + getri when A
//             is not numerically positive definite)
//   per block of grid rows R:
//     B(R, :) = (Psi_v(R, :) G_v) o (Psi(R, :) G)
//     zeta(R, :) = B(R, :) A^-1
// The products are the library's FP64 MFMA GEMM; the factorization is a
// plain library call on an n_intp x n_intp matrix (n_intp^3 / 3 FLOP).

#include "ngw_kernels.h"

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <vector>

namespace ngw {

namespace {

// G[n + mu nband] = psi[intp[mu] + n ld]
__global__ void gather_points_kernel(const double* psi, long ld, const int* intp, int nband, int n_intp, double* g)
{
    const long total = (long)nband * n_intp;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
        const int mu = (int)(t / nband), n = (int)(t - (long)mu * nband);
        g[t] = psi[(long)intp[mu] + (long)n * ld];
    }
}

// a = p o q (n x n, ld n); diag(a) -> d
__global__ void hadamard_diag_kernel(const double* p, const double* q, double* a, int n, double* d)
{
    const long total = (long)n * n;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
        const double v = p[t] * q[t];
        a[t] = v;
        const int j = (int)(t / n), i = (int)(t - (long)j * n);
        if (i == j) d[i] = v;
    }
}

__global__ void add_diag_kernel(double* a, int n, double shift)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) a[(long)i * n + i] += shift;
}

// a(i, j) = a(j, i) for i < j (the inverse from potri is in the lower triangle)
__global__ void symmetrize_lower_kernel(double* a, int n)
{
    const long total = (long)n * n;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
        const int j = (int)(t / n), i = (int)(t - (long)j * n);
        if (i < j) a[t] = a[(long)i * n + j];
    }
}

// x(r, c) *= y(r, c), both (rows x cols) with leading dimension rows
__global__ void hadamard_inplace_kernel(double* x, const double* y, long total)
{
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x)
        x[t] *= y[t];
}

int rb_check(rocblas_status st, const char* what)
{
    if (st != rocblas_status_success) return fail(NGW_ERR_HIP, std::string(what) + " failed (rocblas status " +
                                                              std::to_string((int)st) + ")");
    return NGW_OK;
}

}  // namespace

// zeta (ld_zeta, n_intp) on the device; see ngw_isdf_zeta_dev in ngw.h.
int isdf_zeta_dev(ngw_context* ctx, int ngrid, int nband, int nval, const double* d_psi, long ld_psi, int n_intp,
                  const int* intp, double ridge, double* d_zeta, long ld_zeta, int* used_lu)
{
    hipStream_t s = ctx->stream;
    const int n = n_intp;
    DevBuf dint, g, p, q, a, diag, info, b1, b2, ipiv;
    NGW_TRY(dint.ensure(sizeof(int) * (size_t)n));
    NGW_TRY(g.ensure(sizeof(double) * (size_t)nband * n));
    NGW_TRY(p.ensure(sizeof(double) * (size_t)n * n));
    NGW_TRY(q.ensure(sizeof(double) * (size_t)n * n));
    NGW_TRY(a.ensure(sizeof(double) * (size_t)n * n));
    NGW_TRY(diag.ensure(sizeof(double) * (size_t)n));
    NGW_TRY(info.ensure(sizeof(int)));
    NGW_HIP(hipMemcpyAsync(dint.ptr, intp, sizeof(int) * (size_t)n, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(gather_points_kernel, dim3(1024), dim3(256), 0, s, d_psi, ld_psi, dint.as<int>(), nband, n,
                       g.as<double>());
    NGW_HIP(hipGetLastError());
    // P = G_v^T G_v (k = nval), Q = G^T G (k = nband)
    NGW_TRY(gemm_f64_tn(s, n, n, nval, 1.0, g.as<double>(), nband, g.as<double>(), nband, p.as<double>(), n));
    NGW_TRY(gemm_f64_tn(s, n, n, nband, 1.0, g.as<double>(), nband, g.as<double>(), nband, q.as<double>(), n));
    hipLaunchKernelGGL(hadamard_diag_kernel, dim3(1024), dim3(256), 0, s, p.as<double>(), q.as<double>(),
                       a.as<double>(), n, diag.as<double>());
    NGW_HIP(hipGetLastError());
    double shift = 0.0;
    if (ridge > 0.0) {
        std::vector<double> hd(n);
        NGW_HIP(hipMemcpyAsync(hd.data(), diag.ptr, sizeof(double) * n, hipMemcpyDeviceToHost, s));
        NGW_HIP(hipStreamSynchronize(s));
        double tr = 0.0;
        for (double v : hd) tr += v;
        shift = ridge * tr / n;
        hipLaunchKernelGGL(add_diag_kernel, dim3((n + 255) / 256), dim3(256), 0, s, a.as<double>(), n, shift);
        NGW_HIP(hipGetLastError());
    }
    // keep A (in p) for the LU fallback
    NGW_HIP(hipMemcpyAsync(p.ptr, a.ptr, sizeof(double) * (size_t)n * n, hipMemcpyDeviceToDevice, s));
    rocblas_handle hb = nullptr;
    struct HandleGuard {
        rocblas_handle* h;
        ~HandleGuard() { if (*h) rocblas_destroy_handle(*h); }
    } hguard{&hb};
    NGW_TRY(rb_check(rocblas_create_handle(&hb), "rocblas_create_handle"));
    NGW_TRY(rb_check(rocblas_set_stream(hb, s), "rocblas_set_stream"));
    int hinfo = 0;
    NGW_TRY(rb_check(rocsolver_dpotrf(hb, rocblas_fill_lower, n, a.as<double>(), n, info.as<int>()), "potrf"));
    NGW_HIP(hipMemcpyAsync(&hinfo, info.ptr, sizeof(int), hipMemcpyDeviceToHost, s));
    NGW_HIP(hipStreamSynchronize(s));
    if (used_lu) *used_lu = 0;
    if (hinfo == 0) {
        NGW_TRY(rb_check(rocsolver_dpotri(hb, rocblas_fill_lower, n, a.as<double>(), n, info.as<int>()), "potri"));
        NGW_HIP(hipMemcpyAsync(&hinfo, info.ptr, sizeof(int), hipMemcpyDeviceToHost, s));
        NGW_HIP(hipStreamSynchronize(s));
        if (hinfo != 0) return fail(NGW_ERR_HIP, "potri: singular Cholesky factor");
        hipLaunchKernelGGL(symmetrize_lower_kernel, dim3(1024), dim3(256), 0, s, a.as<double>(), n);
        NGW_HIP(hipGetLastError());
    } else {
        // not numerically positive definite: LU with partial pivoting
        if (used_lu) *used_lu = 1;
        NGW_TRY(ipiv.ensure(sizeof(int) * (size_t)n));
        NGW_HIP(hipMemcpyAsync(a.ptr, p.ptr, sizeof(double) * (size_t)n * n, hipMemcpyDeviceToDevice, s));
        NGW_TRY(rb_check(rocsolver_dgetrf(hb, n, n, a.as<double>(), n, ipiv.as<int>(), info.as<int>()), "getrf"));
        NGW_TRY(rb_check(rocsolver_dgetri(hb, n, a.as<double>(), n, ipiv.as<int>(), info.as<int>()), "getri"));
        NGW_HIP(hipMemcpyAsync(&hinfo, info.ptr, sizeof(int), hipMemcpyDeviceToHost, s));
        NGW_HIP(hipStreamSynchronize(s));
        if (hinfo != 0) return fail(NGW_ERR_HIP, "getrf/getri: A is singular");
    }
    q.release();
    p.release();
    // zeta, block of R grid rows at a time
    const int R = (int)std::min<long>(ngrid, 32768);
    NGW_TRY(b1.ensure(sizeof(double) * (size_t)R * n));
    NGW_TRY(b2.ensure(sizeof(double) * (size_t)R * n));
    for (int r0 = 0; r0 < ngrid; r0 += R) {
        const int rows = std::min(R, ngrid - r0);
        NGW_TRY(gemm_f64(s, rows, n, nval, d_psi + r0, ld_psi, g.as<double>(), nband, b1.as<double>(), rows));
        NGW_TRY(gemm_f64(s, rows, n, nband, d_psi + r0, ld_psi, g.as<double>(), nband, b2.as<double>(), rows));
        hipLaunchKernelGGL(hadamard_inplace_kernel, dim3(1024), dim3(256), 0, s, b1.as<double>(), b2.as<double>(),
                           (long)rows * n);
        NGW_HIP(hipGetLastError());
        NGW_TRY(gemm_f64(s, rows, n, n, b1.as<double>(), rows, a.as<double>(), n, d_zeta + r0, ld_zeta));
    }
    NGW_HIP(hipStreamSynchronize(s));  // workspaces are freed on return
    return NGW_OK;
}

}  // namespace ngw

using namespace ngw;

extern "C" int ngw_isdf_zeta_dev(ngw_context* ctx, int ngrid, int nband, int nval, const double* d_psi, long ld_psi,
                                 int n_intp, const int* intp, double ridge, double* d_zeta, long ld_zeta,
                                 int* used_lu)
{
    if (!ctx) return fail(NGW_ERR_ARG, "null context");
    if (ngrid <= 0 || nband <= 0 || nval <= 0 || nval > nband || n_intp <= 0 || n_intp > ngrid)
        return fail(NGW_ERR_ARG, "bad sizes");
    if (ld_psi < ngrid || ld_zeta < ngrid || !d_psi || !d_zeta || !intp) return fail(NGW_ERR_ARG, "bad arrays");
    if (ridge < 0.0) return fail(NGW_ERR_ARG, "ridge must be >= 0");
    for (int i = 0; i < n_intp; ++i)
        if (intp[i] < 0 || intp[i] >= ngrid) return fail(NGW_ERR_ARG, "interpolation point outside the grid");
    NGW_HIP(hipSetDevice(ctx->device));
    return isdf_zeta_dev(ctx, ngrid, nband, nval, d_psi, ld_psi, n_intp, intp, ridge, d_zeta, ld_zeta, used_lu);
}

extern "C" int ngw_isdf_zeta(ngw_context* ctx, int ngrid, int nband, int nval, const double* psi, long ld_psi,
                             int n_intp, const int* intp, double ridge, double* zeta, long ld_zeta, int* used_lu)
{
    if (!ctx) return fail(NGW_ERR_ARG, "null context");
    if (ngrid <= 0 || nband <= 0 || n_intp <= 0 || ld_psi < ngrid || ld_zeta < ngrid || !psi || !zeta)
        return fail(NGW_ERR_ARG, "bad sizes or arrays");
    NGW_HIP(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    DevBuf dpsi, dzeta;
    NGW_TRY(dpsi.ensure(sizeof(double) * (size_t)ngrid * nband));
    NGW_TRY(dzeta.ensure(sizeof(double) * (size_t)ngrid * n_intp));
    NGW_HIP(hipMemcpy2DAsync(dpsi.ptr, sizeof(double) * ngrid, psi, sizeof(double) * ld_psi, sizeof(double) * ngrid,
                             nband, hipMemcpyHostToDevice, s));
    NGW_TRY(ngw_isdf_zeta_dev(ctx, ngrid, nband, nval, dpsi.as<double>(), ngrid, n_intp, intp, ridge,
                              dzeta.as<double>(), ngrid, used_lu));
    NGW_HIP(hipMemcpy2DAsync(zeta, sizeof(double) * ld_zeta, dzeta.ptr, sizeof(double) * ngrid, sizeof(double) * ngrid,
                             n_intp, hipMemcpyDeviceToHost, s));
    NGW_HIP(hipStreamSynchronize(s));
    return NGW_OK;
}
