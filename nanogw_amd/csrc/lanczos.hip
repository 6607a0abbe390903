// FP64 Lanczos recurrence, continued fraction, Sigma_c assembly and QP Newton
// kernels (gfx950).  These stay in FP64 in every precision mode.

#include "ngw_kernels.h"

#include <cmath>
#include <vector>

namespace ngw {

constexpr int RT = 256;  // threads of the per-column reduction kernels

// Deterministic sum over the 256 threads of a block (fixed shuffle pattern,
// then the four wave sums in wave order).
__device__ __forceinline__ double block_sum(double v, double* sh)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[wave] = v;
    __syncthreads();
    double t = ((sh[0] + sh[1]) + sh[2]) + sh[3];
    return t;
}

// One Lanczos step's scalar work for column blockIdx.x, following
// shared/lanczos_spectra_BLOCK.F90:63-107 (alpha, w -= alpha q, beta) and the
// start of the next step (:81-88: w /= beta, q = -q beta, swap), which is
// applied here so that the next p(H) q reads q in place.
__global__ __launch_bounds__(RT) void lanczos_ab_kernel(double* q, double* w, long ld, long n,
                                                        double* a, double* b, int niter, int k)
{
    __shared__ double sh[4];
    const long col = blockIdx.x;
    double* qc = q + col * ld;
    double* wc = w + col * ld;
    double s = 0.0;
    for (long i = threadIdx.x; i < n; i += RT) s += wc[i] * qc[i];
    const double al = block_sum(s, sh);
    if (threadIdx.x == 0) {
        a[(k - 1) + col * niter] = al;
        if (k == 1) b[col * niter] = 1.0;
    }
    if (k + 1 > niter) return;
    double s2 = 0.0;
    for (long i = threadIdx.x; i < n; i += RT) {
        double wi = wc[i] + (-al) * qc[i];
        wc[i] = wi;
        s2 += wi * wi;
    }
    const double be = sqrt(block_sum(s2, sh));
    if (threadIdx.x == 0) b[k + col * niter] = be;
    for (long i = threadIdx.x; i < n; i += RT) {
        double wi = wc[i], qi = qc[i];
        qc[i] = wi / be;
        wc[i] = -qi * be;
    }
}

// ---------------------------------------------------------------- full reorthogonalisation
// Optional (off in the reference path): the step above split in two around
// two classical Gram-Schmidt passes of w against every stored Lanczos
// vector q_1..q_k of its column.  The basis of column c is Q + (c niter + j)
// ld, j = 0..niter-1.  All sums are in a fixed order, so results do not
// depend on the launch.
//
// Part 1: alpha = w.q, w -= alpha q, q_k stored as basis vector k - 1.
__global__ __launch_bounds__(RT) void lanczos_a_kernel(const double* q, double* w, long ld, long n, double* a,
                                                       double* b, int niter, int k, double* basis)
{
    __shared__ double sh[4];
    const long col = blockIdx.x;
    const double* qc = q + col * ld;
    double* wc = w + col * ld;
    double* bc = basis + ((long)col * niter + (k - 1)) * ld;
    double s = 0.0;
    for (long i = threadIdx.x; i < n; i += RT) s += wc[i] * qc[i];
    const double al = block_sum(s, sh);
    if (threadIdx.x == 0) {
        a[(k - 1) + col * niter] = al;
        if (k == 1) b[col * niter] = 1.0;
    }
    for (long i = threadIdx.x; i < n; i += RT) {
        const double qi = qc[i];
        wc[i] = wc[i] + (-al) * qi;
        bc[i] = qi;
    }
}

// h(j, col) = q_j . w for j < k: one workgroup per (j, col).
__global__ __launch_bounds__(RT) void reorth_dot_kernel(const double* basis, const double* w, long ld, long n,
                                                        int niter, int k, double* h)
{
    __shared__ double sh[4];
    const int j = blockIdx.x;
    const long col = blockIdx.y;
    const double* qj = basis + ((long)col * niter + j) * ld;
    const double* wc = w + col * ld;
    double s = 0.0;
    for (long i = threadIdx.x; i < n; i += RT) s += qj[i] * wc[i];
    const double t = block_sum(s, sh);
    if (threadIdx.x == 0) h[j + col * niter] = t;
}

// w(i, col) -= sum_{j<k} q_j(i) h(j, col), j ascending.
__global__ __launch_bounds__(RT) void reorth_update_kernel(const double* basis, double* w, long ld, long n,
                                                           int niter, int k, const double* h)
{
    const long col = blockIdx.y;
    const double* hc = h + (long)col * niter;
    const double* bc = basis + (long)col * niter * ld;
    double* wc = w + col * ld;
    for (long i = blockIdx.x * (long)RT + threadIdx.x; i < n; i += (long)gridDim.x * RT) {
        double s = 0.0;
        for (int j = 0; j < k; ++j) s += bc[(long)j * ld + i] * hc[j];
        wc[i] -= s;
    }
}

// Part 2: beta = |w|, then the rotation into the next step (q = w / beta,
// w = -beta q_k), as in lanczos_ab_kernel.
__global__ __launch_bounds__(RT) void lanczos_b_kernel(double* q, double* w, long ld, long n, double* b, int niter,
                                                       int k)
{
    __shared__ double sh[4];
    const long col = blockIdx.x;
    double* qc = q + col * ld;
    double* wc = w + col * ld;
    double s2 = 0.0;
    for (long i = threadIdx.x; i < n; i += RT) s2 += wc[i] * wc[i];
    const double be = sqrt(block_sum(s2, sh));
    if (threadIdx.x == 0) b[k + col * niter] = be;
    for (long i = threadIdx.x; i < n; i += RT) {
        const double wi = wc[i], qi = qc[i];
        qc[i] = wi / be;
        wc[i] = -qi * be;
    }
}

__global__ __launch_bounds__(RT) void column_norm_kernel(const double* v, long ld, long n,
                                                         double* norm2, double* norm)
{
    __shared__ double sh[4];
    const long col = blockIdx.x;
    const double* vc = v + col * ld;
    double s = 0.0;
    for (long i = threadIdx.x; i < n; i += RT) s += vc[i] * vc[i];
    const double t = block_sum(s, sh);
    if (threadIdx.x == 0) {
        if (norm2) norm2[col] = t;
        if (norm) norm[col] = sqrt(t);
    }
}

__global__ void scale_columns_kernel(const double* in, double* out, long ld, long n,
                                     const double* norm)
{
    const long col = blockIdx.y;
    const double nr = norm[col];
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        out[i + col * ld] = in[i + col * ld] / nr;
}

int scale_columns(hipStream_t s, const double* in, double* out, long ld, long n, int nvec,
                  const double* d_norm)
{
    if (nvec <= 0) return NGW_OK;
    int gx = (int)std::min<long>((n + 255) / 256, 1024);
    hipLaunchKernelGGL(scale_columns_kernel, dim3(gx, nvec), dim3(256), 0, s, in, out, ld, n, d_norm);
    NGW_HIP(hipGetLastError());
    return NGW_OK;
}

int column_norms(hipStream_t s, const double* v, long ld, long n, int nvec, double* d_norm2,
                 double* d_norm)
{
    if (nvec <= 0) return NGW_OK;
    hipLaunchKernelGGL(column_norm_kernel, dim3(nvec), dim3(RT), 0, s, v, ld, n, d_norm2, d_norm);
    NGW_HIP(hipGetLastError());
    return NGW_OK;
}

// Chebyshev coefficients c_0..c_n of the monomial polynomial
// poly = (a_n, ..., a_0) on [lo, hi]: p(x) = sum_k c_k T_k(t),
// x = ((hi - lo) t + (hi + lo)) / 2.  Horner in the Chebyshev basis in long
// double (the monomial terms cancel heavily, the result does not).
static std::vector<double> chebyshev_from_monomial(const double* poly, int npoly, double lo, double hi)
{
    typedef long double ld;
    const ld al = ((ld)hi - (ld)lo) / 2, be = ((ld)hi + (ld)lo) / 2;
    std::vector<ld> c(1, (ld)poly[0]);
    for (int k = 1; k <= npoly; ++k) {
        // c <- (al t + be) c + a_{n-k}
        std::vector<ld> r(c.size() + 1, 0.0L);
        for (size_t j = 0; j < c.size(); ++j) {
            r[j] += be * c[j];
            if (j == 0) {
                r[1] += al * c[0];
            } else {
                r[j + 1] += al * c[j] / 2;
                r[j - 1] += al * c[j] / 2;
            }
        }
        r[0] += (ld)poly[k];
        c.swap(r);
    }
    std::vector<double> out(c.size());
    for (size_t j = 0; j < c.size(); ++j) out[j] = (double)c[j];
    return out;
}

// Spectral interval of the driver's sqrt fit, calculate_sigma_lanczos_BLOCK.F90z:166-171:
// e_min = (e_{nv+1} - e_nv)^2 / 1.25, e_max = (e_nb - e_1)^2 * 1.25 (Ry^2).
static void fit_interval(const ngw_context* ctx, double& lo, double& hi)
{
    lo = std::pow(ctx->e1_host[ctx->nval] - ctx->e1_host[ctx->nval - 1], 2.0) / 1.25;
    hi = std::pow(ctx->e1_host[ctx->nband - 1] - ctx->e1_host[0], 2.0) * 1.25;
    if (!(hi > lo)) {
        lo = 0.0;
        hi = 1.0;
    }
}

std::vector<PolyStage> plan_poly(const ngw_context* ctx, int precision, const double* q, double* out, double* h0,
                                 double* h1, double* h2, const double* poly, int npoly, bool accumulate)
{
    std::vector<PolyStage> plan;
    if (precision == PREC_F64 || npoly == 1) {
        // Horner, apply_matvec_BLOCK.F90:250-265: h = a_n H q + a_{n-1} q; h = H h + a_k q
        const double* in = q;
        double* bufs[2] = {h0, h1};
        for (int k = 1; k <= npoly; ++k) {
            PolyStage st;
            st.x = in;
            st.alpha = (k == 1) ? poly[0] : 1.0;
            st.gamma = poly[k];
            const bool last = (k == npoly);
            st.out = last ? out : bufs[k & 1];
            st.accumulate = last && accumulate;
            plan.push_back(st);
            in = st.out;
        }
        return plan;
    }
    // Clenshaw on p(x) = sum c_k T_k(t), t = (x - be)/al:
    //   b_{n+1} = b_{n+2} = 0, b_k = 2 t(b_{k+1}) - b_{k+2} + c_k q, p = t(b_1) - b_2 + c_0 q.
    // b_n = c_n q is kept implicit (a multiple of q).  The intermediates
    // b_k are stored as FP32 in the h buffers (the GEMM operands are at most
    // FP32 anyway); q and the result stay FP64.
    double lo, hi;
    fit_interval(ctx, lo, hi);
    const std::vector<double> c = chebyshev_from_monomial(poly, npoly, lo, hi);
    const double al = (hi - lo) / 2, be = (hi + lo) / 2;
    struct V { const double* p; double qc; };  // p, or qc * q when p == null
    V bk1{nullptr, c[npoly]}, bk2{nullptr, 0.0};
    double* bufs[3] = {h0, h1, h2};
    int nb = 0;
    for (int k = npoly - 1; k >= 0; --k) {
        const bool last = (k == 0);
        const double f = last ? 1.0 : 2.0;  // t(b) weight: the final stage uses t, the others 2t
        PolyStage st;
        double gam = c[k];
        if (bk1.p) {
            st.x = bk1.p; st.alpha = f / al; st.bx = -f * be / al;
            st.flags |= ST_X32;
        } else {
            st.x = q; st.alpha = f * bk1.qc / al; gam += -f * be / al * bk1.qc;
            st.flags |= ST_QW32;  // the first stage: its x' pass keeps q as FP32
        }
        if (bk2.p) {
            st.y = bk2.p; st.by = -1.0;
            st.flags |= ST_Y32;
        } else {
            gam += -bk2.qc;
        }
        st.gamma = gam;
        double* o;
        if (last) {
            o = out;
            st.accumulate = accumulate;
        } else {
            // a buffer that is neither b_{k+1} nor b_{k+2}
            o = bufs[nb];
            nb = (nb + 1) % 3;
            if (o == bk1.p || o == bk2.p) { o = bufs[nb]; nb = (nb + 1) % 3; }
            if (o == bk1.p || o == bk2.p) { o = bufs[nb]; nb = (nb + 1) % 3; }
            st.flags |= ST_O32 | ST_Q32;  // an FP32 b_k: its gamma q term reads q in FP32
        }
        st.out = o;
        plan.push_back(st);
        bk2 = bk1;
        bk1 = V{o, 0.0};
    }
    return plan;
}

int poly_apply(ngw_context* ctx, const double* q, double* out, double* h0, double* h1, double* h2, long ld,
               int nvec, const double* poly, int npoly, int tda, bool accumulate)
{
    const std::vector<PolyStage> plan = plan_poly(ctx, ctx->precision, q, out, h0, h1, h2, poly, npoly, accumulate);
    for (const PolyStage& st : plan) {
        OperatorArgs a;
        a.x = st.x; a.ldx = ld; a.nvec = nvec; a.tda = tda;
        a.alpha = st.alpha; a.bx = st.bx;
        a.y = st.y; a.ldy = ld; a.by = st.by;
        a.v0 = q; a.ldv0 = ld; a.gamma = st.gamma;
        a.out = st.out; a.ldo = ld;
        if (st.accumulate) { a.acc = st.out; a.ldacc = ld; }
        a.flags = st.flags;
        NGW_TRY(operator_apply(ctx, a));
    }
    return NGW_OK;
}

int lanczos_run(ngw_context* ctx, double* d_q, double* d_w, double* d_h0, double* d_h1, double* d_h2, long ld,
                int nvec, int niter, const double* poly, int npoly, int tda, double* d_a, double* d_b)
{
    hipStream_t s = ctx->stream;
    const long n = ctx->ncv;
    double* basis = nullptr;
    double* h = nullptr;
    if (ctx->reorth) {
        const double need = 8.0 * ((double)niter * ld * nvec + (double)niter * nvec);
        if (ctx->qbasis.ensure((size_t)need) != NGW_OK)
            return fail(NGW_ERR_NOMEM, "reorthogonalisation keeps every Lanczos vector: " +
                                           std::to_string(need / 1e9) + " GB for " + std::to_string(nvec) +
                                           " columns x " + std::to_string(niter) + " steps does not fit");
        basis = ctx->qbasis.as<double>();
        h = basis + (size_t)niter * ld * nvec;
    }
    for (int k = 1; k <= niter; ++k) {
        NGW_TRY(poly_apply(ctx, d_q, d_w, d_h0, d_h1, d_h2, ld, nvec, poly, npoly, tda, k > 1));
        ProfScope ps(ctx, KC_LANCZOS);
        if (!basis) {
            hipLaunchKernelGGL(lanczos_ab_kernel, dim3(nvec), dim3(RT), 0, s, d_q, d_w, ld, n, d_a, d_b, niter, k);
        } else {
            hipLaunchKernelGGL(lanczos_a_kernel, dim3(nvec), dim3(RT), 0, s, d_q, d_w, ld, n, d_a, d_b, niter, k,
                               basis);
            if (k < niter) {
                const int gx = (int)std::min<long>((n + RT - 1) / RT, 64);
                for (int pass = 0; pass < 2; ++pass) {
                    hipLaunchKernelGGL(reorth_dot_kernel, dim3(k, nvec), dim3(RT), 0, s, basis, d_w, ld, n, niter,
                                       k, h);
                    hipLaunchKernelGGL(reorth_update_kernel, dim3(gx, nvec), dim3(RT), 0, s, basis, d_w, ld, n,
                                       niter, k, h);
                }
                hipLaunchKernelGGL(lanczos_b_kernel, dim3(nvec), dim3(RT), 0, s, d_q, d_w, ld, n, d_b, niter, k);
            }
        }
        NGW_HIP(hipGetLastError());
    }
    return NGW_OK;
}

// ------------------------------------------------------------ complex helpers
struct cplx {
    double re, im;
};
__device__ __forceinline__ cplx cdiv(cplx x, cplx y)
{
    // Smith's algorithm for x / y.
    if (fabs(y.re) >= fabs(y.im)) {
        double r = y.im / y.re, d = y.re + y.im * r;
        return {(x.re + x.im * r) / d, (x.im - x.re * r) / d};
    } else {
        double r = y.re / y.im, d = y.re * r + y.im;
        return {(x.re * r + x.im) / d, (x.im * r - x.re) / d};
    }
}
__device__ __forceinline__ cplx cmul(cplx x, cplx y)
{
    return {x.re * y.re - x.im * y.im, x.re * y.im + x.im * y.re};
}

// S(z) continued fraction, lanczos_spectra_BLOCK.F90:117-130; one thread per
// (z, column).
__global__ void cfrac_kernel(const double* a, const double* b, int niter, int nvec,
                             const double* zz, int nz, double* spec)
{
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t >= (long)nz * nvec) return;
    const long col = t / nz;
    const int iz = (int)(t - col * nz);
    const cplx z = {zz[2 * (iz + col * nz)], zz[2 * (iz + col * nz) + 1]};
    const double* ac = a + col * niter;
    const double* bc = b + col * niter;
    cplx sv = {0.0, 0.0};
    for (int j = niter; j >= 1; --j) {
        double en = (j == 1) ? 1.0 : bc[j - 1] * bc[j - 1];
        cplx de = {z.re - ac[j - 1] - sv.re, z.im - sv.im};
        sv = cdiv(cplx{en, 0.0}, de);
    }
    spec[2 * t] = sv.re;
    spec[2 * t + 1] = sv.im;
}

int continued_fraction(hipStream_t s, const double* d_a, const double* d_b, int niter, int nvec,
                       const double* d_zz, int nz, double* d_spec)
{
    long total = (long)nz * nvec;
    if (total <= 0) return NGW_OK;
    hipLaunchKernelGGL(cfrac_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, d_a, d_b,
                       niter, nvec, d_zz, nz, d_spec);
    NGW_HIP(hipGetLastError());
    return NGW_OK;
}

// zeta grid of every (QP state, band) column:
//   z_j = e_m - deltae/2 + j deltae/(nen-1)            (BLOCK driver :259-263)
//   zeta_j = -(z_j - e_n) + i gamma (occupied n)       (:366-368)
//          =  (z_j - e_n) + i gamma (empty n)          (:393-395)
//   zeta_nen = 0                                       (:371)
__global__ void zz_grid_kernel(const double* e1, const int* mband, const int* nband, int nval,
                               int nvec, int nen, double deltae, double ecuts, double* zz)
{
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    const int nz = nen + 1;
    if (t >= (long)nz * nvec) return;
    const long col = t / nz;
    const int i = (int)(t - col * nz);
    double re = 0.0, im = 0.0;
    if (i < nen) {
        const double de = deltae / (nen - 1);
        const double en0 = e1[mband[col]] - deltae * 0.5;
        const double z = de * i + en0;
        const double en = e1[nband[col]];
        re = (nband[col] < nval) ? -(z - en) : (z - en);
        im = ecuts;
    }
    zz[2 * t] = re;
    zz[2 * t + 1] = im;
}

int zz_grid(hipStream_t s, const double* d_e1, const int* d_mband, const int* d_nband, int nval,
            int nvec, int nen, double deltae, double ecuts, double* d_zz)
{
    long total = (long)(nen + 1) * nvec;
    hipLaunchKernelGGL(zz_grid_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, d_e1,
                       d_mband, d_nband, nval, nvec, nen, deltae, ecuts, d_zz);
    NGW_HIP(hipGetLastError());
    return NGW_OK;
}

// Sigma_c(z_i) for QP state iq = sum over bands n (in band order) of
//   full, occupied: 2|P|^2 / (-zeta) (S(zeta) - S(0))   (BLOCK driver :382-383)
//   full, empty   : 2|P|^2 / zeta (S(zeta) - S(0))      (:408-409)
//   TDA           : -S(zeta) eta sqrt(2) |P|^2          (lowcomsym :738-740; the
//                   reference's 2.0**0.5 is single precision)
// Columns are laid out as col = iq * nband + n for the bands band0 + n,
// n < nband (a contiguous band range; the full sum is band0 = 0, nband = N_b).
__global__ void sigma_assemble_kernel(const double* spec, const double* zz, const double* pn2,
                                      int nen, int nqp, int nband, int band0, int nval, int tda, double* sigma)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nen * nqp) return;
    const int iq = t / nen, i = t - iq * nen;
    const int nz = nen + 1;
    const double sqrt2_f32 = (double)sqrtf(2.0f);
    cplx acc = {0.0, 0.0};
    for (int n = 0; n < nband; ++n) {
        const long col = (long)iq * nband + n;
        const long o = 2 * (i + col * nz), o0 = 2 * (nen + col * nz);
        const cplx si = {spec[o], spec[o + 1]}, s0 = {spec[o0], spec[o0 + 1]};
        const double p2 = pn2[col];
        const bool occ = band0 + n < nval;
        cplx inc;
        if (tda) {
            const double eta = occ ? 1.0 : -1.0;
            inc = {-si.re * eta * sqrt2_f32 * p2, -si.im * eta * sqrt2_f32 * p2};
        } else {
            cplx ez = {zz[o], zz[o + 1]};
            if (occ) ez = {-ez.re, -ez.im};
            cplx f = cdiv(cplx{2.0 * p2, 0.0}, ez);
            inc = cmul(f, cplx{-s0.re + si.re, -s0.im + si.im});
        }
        acc.re = acc.re + inc.re;
        acc.im = acc.im + inc.im;
    }
    sigma[2 * t] = acc.re;
    sigma[2 * t + 1] = acc.im;
}

int sigma_assemble(hipStream_t s, const double* d_spec, const double* d_zz, const double* d_pn2,
                   int nen, int nqp, int nband, int band0, int nval, int tda, double* d_sigma)
{
    int total = nen * nqp;
    hipLaunchKernelGGL(sigma_assemble_kernel, dim3((total + 127) / 128), dim3(128), 0, s, d_spec, d_zz,
                       d_pn2, nen, nqp, nband, band0, nval, tda, d_sigma);
    NGW_HIP(hipGetLastError());
    return NGW_OK;
}

// ------------------------------------------------------------ QP Newton
// Cubic spline of SIGMA/spline.f90:20-77 on one real table (nn points).
__device__ void spline_dev(const double* xx, const double* yy, int nn, double yp1, double ypn,
                           double* y2, double* uu)
{
    const double HUGEV = 1.0e10;
    if (fabs(yp1) > HUGEV) {
        y2[0] = 0.0;
        uu[0] = 0.0;
    } else {
        y2[0] = -0.5;
        uu[0] = (3.0 / (xx[1] - xx[0])) * ((yy[1] - yy[0]) / (xx[1] - xx[0]) - yp1);
    }
    for (int i = 1; i < nn - 1; ++i) {
        double sig = (xx[i] - xx[i - 1]) / (xx[i + 1] - xx[i - 1]);
        double p = sig * y2[i - 1] + 2.0;
        y2[i] = (sig - 1.0) / p;
        uu[i] = (6.0 * ((yy[i + 1] - yy[i]) / (xx[i + 1] - xx[i]) - (yy[i] - yy[i - 1]) / (xx[i] - xx[i - 1])) /
                     (xx[i + 1] - xx[i - 1]) - sig * uu[i - 1]) / p;
    }
    double qn, un;
    if (ypn > HUGEV) {
        qn = 0.0;
        un = 0.0;
    } else {
        qn = 0.5;
        un = (3.0 / (xx[nn - 1] - xx[nn - 2])) * (ypn - (yy[nn - 1] - yy[nn - 2]) / (xx[nn - 1] - xx[nn - 2]));
    }
    y2[nn - 1] = (un - qn * uu[nn - 2]) / (qn * y2[nn - 2] + 1.0);
    for (int k = nn - 2; k >= 0; --k) y2[k] = y2[k] * y2[k + 1] + uu[k];
}

// SIGMA/splint.f90:12-53.
__device__ double splint_dev(const double* xx, const double* yy, const double* y2, int nn, double x0)
{
    int klo = 1, khi = nn;
    while (khi - klo > 1) {
        int kk = (khi + klo) / 2;
        if (xx[kk - 1] > x0) khi = kk; else klo = kk;
    }
    double hh = xx[khi - 1] - xx[klo - 1];
    double aa = (xx[khi - 1] - x0) / hh;
    double bb = (x0 - xx[klo - 1]) / hh;
    return aa * yy[klo - 1] + bb * yy[khi - 1] +
           ((aa * aa * aa - aa) * y2[klo - 1] + (bb * bb * bb - bb) * y2[khi - 1]) * (hh * hh) / 6.0;
}

// SIGMA/spline_evaluate.f90:13-51 for a real table.
__device__ void spline_eval_dev(int nt, const double* xt, const double* ft, double x, double* fx,
                                double* df, double* y2, double* uu)
{
    const int nn = nt - 2;
    double yp1 = (ft[1] - ft[0]) / (xt[1] - xt[0]);
    double ypn = (ft[nt - 1] - ft[nt - 2]) / (xt[nt - 1] - xt[nt - 2]);
    spline_dev(xt + 1, ft + 1, nn, yp1, ypn, y2, uu);
    const double DELTAX = 1.0e-2;
    *fx = splint_dev(xt + 1, ft + 1, y2, nn, x);
    double ym = splint_dev(xt + 1, ft + 1, y2, nn, x - DELTAX);
    double yp = splint_dev(xt + 1, ft + 1, y2, nn, x + DELTAX);
    *df = (yp - ym) / DELTAX * 0.5;
}

// One QP state per thread, SIGMA/printsigma.F90:97-128 with newtonr
// (SIGMA/newtonr.f90:18-58, EN_ACC = 0.01 eV, 20 iterations).
// in  (per state): [edft, en0, sigx, vxc]  at in[4*iq], then Sigma_c(nen) complex
//                  at in[4*nqp + 2*nen*iq]
// out (per state): [eqp, zfac, sigma0_re, sigma0_im, imsig, ierr]
__global__ void qp_newton_kernel(int nqp, int nen, double deltae_ev, const double* in, double* out,
                                 double* scratch)
{
    const int iq = blockIdx.x * blockDim.x + threadIdx.x;
    if (iq >= nqp) return;
    double* xt = scratch + (long)iq * 5 * nen;
    double* ft = xt + nen;
    double* it = ft + nen;
    double* y2 = it + nen;
    double* uu = y2 + nen;
    const double edft = in[4 * iq], en0 = in[4 * iq + 1], sigx = in[4 * iq + 2], vxc = in[4 * iq + 3];
    const double* sc = in + 4 * (long)nqp + 2 * (long)nen * iq;
    for (int ii = 0; ii < nen; ++ii) {
        xt[ii] = en0 - deltae_ev * 0.5 + deltae_ev / (double)(nen - 1) * (double)ii;
        ft[ii] = sigx + sc[2 * ii] - vxc + edft - xt[ii];
        it[ii] = sc[2 * ii + 1];
    }
    double e = en0, fx = 0.0, df = 0.0;
    int ierr = 21;
    for (int jj = 1; jj <= 20; ++jj) {
        spline_eval_dev(nen, xt, ft, e, &fx, &df, y2, uu);
        double dx = fx / df;
        e = e - dx;
        if (fabs(dx) < 0.01) {
            spline_eval_dev(nen, xt, ft, e, &fx, &df, y2, uu);
            ierr = 0;
            break;
        }
    }
    double s0, zf, xtmp, si;
    spline_eval_dev(nen, xt, ft, e, &s0, &zf, y2, uu);
    spline_eval_dev(nen, xt, ft, edft, &s0, &xtmp, y2, uu);
    s0 = s0 + vxc;
    zf = -1.0 / zf;
    spline_eval_dev(nen, xt, it, e, &si, &xtmp, y2, uu);
    double* o = out + 6 * (long)iq;
    o[0] = e;
    o[1] = zf;
    o[2] = s0;
    o[3] = 0.0;
    o[4] = si;
    o[5] = (double)ierr;
}

int qp_newton(hipStream_t s, int nqp, int nen, double deltae_ev, const double* d_in, double* d_out,
              double* d_scratch)
{
    hipLaunchKernelGGL(qp_newton_kernel, dim3((nqp + 63) / 64), dim3(64), 0, s, nqp, nen, deltae_ev,
                       d_in, d_out, d_scratch);
    NGW_HIP(hipGetLastError());
    return NGW_OK;
}

}  // namespace ngw
