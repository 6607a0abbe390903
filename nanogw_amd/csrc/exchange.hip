// Bare exchange Sigma_x of QP states in the ISDF representation:
//   Sigma_x(m) = - sum_{v occupied} (m v | v m)
//              = - sum_v sum_{mu,nu} c_mv(mu) M(mu, nu) c_mv(nu),
//   c_mv(mu) = Psi(mu, m) Psi(mu, v),
// the ISDF form of the exchange matrix element NanoGW evaluates with FFT pair
// densities in Zfock_exchange (SIGMA/fock_exchange.F90z:12-229); M is the same
// Coulomb kernel the Lanczos operator uses, so Sigma_x and Sigma_c share one
// representation.  Energies in Ry (e^2 = 2).

#include "ngw_kernels.h"

namespace ngw {

// C(mu, (v, q)) = Psi(mu, m_q) Psi(mu, v), column v + nv q.
__global__ void pair_cols_kernel(const double* psi, long ldp, int nmu, int nv, const int* qp, int nqp, double* c)
{
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t >= (long)nmu * nv * nqp) return;
    const int mu = (int)(t % nmu);
    const long col = t / nmu;
    const int v = (int)(col % nv), q = (int)(col / nv);
    c[t] = psi[mu + (long)qp[q] * ldp] * psi[mu + (long)v * ldp];
}

// sigx(q) = - sum over (v, mu) of C * (M C), one workgroup per state, fixed
// summation order.
__global__ __launch_bounds__(256) void exchange_reduce_kernel(const double* c, const double* mc, int nmu, int nv,
                                                              double* sigx)
{
    __shared__ double sh[256];
    const int q = blockIdx.x;
    const long n = (long)nmu * nv, base = (long)q * n;
    double acc = 0.0;
    for (long i = threadIdx.x; i < n; i += 256) acc += c[base + i] * mc[base + i];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) sigx[q] = -sh[0];
}

int sigma_x_isdf(ngw_context* ctx, int nqp, const int* d_qp, double* d_sigx)
{
    const int nmu = ctx->n_intp, nv = ctx->nval;
    const long cols = (long)nv * nqp;
    NGW_TRY(ctx->w0.ensure(sizeof(double) * (size_t)nmu * cols));
    NGW_TRY(ctx->w1.ensure(sizeof(double) * (size_t)nmu * cols));
    hipStream_t s = ctx->stream;
    double* c = ctx->w0.as<double>();
    double* mc = ctx->w1.as<double>();
    const long tot = (long)nmu * cols;
    hipLaunchKernelGGL(pair_cols_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, ctx->psi.as<double>(),
                       (long)ctx->ld_psi, nmu, nv, d_qp, nqp, c);
    NGW_HIP(hipGetLastError());
    NGW_TRY(gemm_f64(s, nmu, (int)cols, nmu, ctx->mmat.as<double>(), ctx->ld_m, c, nmu, mc, nmu));
    hipLaunchKernelGGL(exchange_reduce_kernel, dim3(nqp), dim3(256), 0, s, c, mc, nmu, nv, d_sigx);
    NGW_HIP(hipGetLastError());
    return NGW_OK;
}

}  // namespace ngw
