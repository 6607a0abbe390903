// Internal kernel interfaces of libngw.
#pragma once

#include "ngw_common.h"

namespace ngw {

// Vector blocks of the polynomial evaluation are FP64, or FP32 for the
// Clenshaw intermediates of the low-precision modes (flag set): the element
// i of a block stored as float is ((const float*)p)[i].
__device__ __forceinline__ double vload(const double* p, bool f32, long i)
{
    return f32 ? (double)reinterpret_cast<const float*>(p)[i] : p[i];
}
__device__ __forceinline__ void vstore(double* p, bool f32, long i, double v)
{
    if (f32) reinterpret_cast<float*>(p)[i] = (float)v;
    else p[i] = v;
}

// Stage flags: which vector blocks of a stage are stored as FP32.  ST_QW32:
// the stage's input x is the polynomial's argument q (FP64), and the
// low-precision x' pass also writes q as FP32 (LpOp::q32); ST_Q32: the
// epilogue's gamma q term reads that FP32 copy (an FP32 output stage).
enum { ST_X32 = 1, ST_Y32 = 2, ST_O32 = 4, ST_QW32 = 8, ST_Q32 = 16 };

// Epilogue of the operator's last GEMM (K3), applied per output element
// (pair index idx of vector b, GEMM value Y = (C^T M C x')(idx)):
//   hx  = d(idx) x'(idx) + beta s(idx) Y        (x' = s o x, or x if s == null)
//   r   = alpha hx + bx x(idx) + by y(idx) + gamma v0(idx)
//                                                (one Horner or Clenshaw stage)
//   out = acc(idx) + r                           (Lanczos w += p(H) q)
// x, y and out (and acc, which is out) may be FP32 blocks (flags).
struct OperatorEpilogue {
    const double* x = nullptr; long ldx = 0;
    const double* d = nullptr;
    const double* s = nullptr;
    double beta = 0.0, alpha = 1.0;
    const double* v0 = nullptr; long ldv0 = 0; double gamma = 0.0;
    const double* acc = nullptr; long ldacc = 0;
    double* out = nullptr; long ldo = 0;
    double bx = 0.0;
    const double* y = nullptr; long ldy = 0; double by = 0.0;
    int flags = 0;

    __device__ __forceinline__ void apply(long idx, int b, double Y) const
    {
        double sv = s ? s[idx] : 1.0;
        double hx = (beta * sv) * Y;
        double xv = 0.0;
        if (x) {
            xv = vload(x, flags & ST_X32, idx + (long)b * ldx);
            double xp = s ? sv * xv : xv;
            hx = d[idx] * xp + hx;
        }
        double r = alpha * hx;
        if (bx != 0.0) r = r + bx * xv;
        if (y) r = r + by * vload(y, flags & ST_Y32, idx + (long)b * ldy);
        if (v0) r = r + gamma * v0[idx + (long)b * ldv0];
        if (acc) r = vload(acc, flags & ST_O32, idx + (long)b * ldacc) + r;
        vstore(out, flags & ST_O32, idx + (long)b * ldo, r);
    }
};

// One operator application on nvec device vectors (the matvec of
// apply_matvec_BLOCK.F90 with a fused Horner / Lanczos epilogue).
struct OperatorArgs {
    const double* x = nullptr; long ldx = 0;
    int nvec = 0;
    int tda = 0;
    double* out = nullptr; long ldo = 0;
    double alpha = 1.0;
    const double* v0 = nullptr; long ldv0 = 0; double gamma = 0.0;
    const double* acc = nullptr; long ldacc = 0;
    double bx = 0.0;
    const double* y = nullptr; long ldy = 0; double by = 0.0;
    int flags = 0;  // ST_*: FP32 vector blocks (low-precision Clenshaw intermediates)

    OperatorEpilogue epilogue(const ngw_context* ctx) const
    {
        OperatorEpilogue e;
        e.x = x; e.ldx = ldx;
        e.d = tda ? ctx->diag_tda.as<double>() : ctx->diag_full.as<double>();
        e.s = tda ? nullptr : ctx->sqrtr.as<double>();
        e.beta = tda ? 2.0 : 4.0;
        e.alpha = alpha;
        e.v0 = v0; e.ldv0 = ldv0; e.gamma = gamma;
        e.acc = acc; e.ldacc = ldacc;
        e.out = out; e.ldo = ldo;
        e.bx = bx; e.y = y; e.ldy = ldy; e.by = by;
        e.flags = flags;
        return e;
    }
};

int gemm_f64(hipStream_t s, int m, int n, int k, const double* a, long lda, const double* b,
             long ldb, double* c, long ldc);
int gemm_f64_tn(hipStream_t s, int m, int n, int k, double alpha, const double* a, long lda,
                const double* b, long ldb, double* c, long ldc);
int k1_slots(int nv);
int k1_reduce(hipStream_t s, const double* tpart, long ldt, int nmu, int nvec, int nv, double* t);
int gemm_lowp(hipStream_t s, int kind, int m, int n, int k, const double* a, long lda, const double* b, long ldb,
              double* c, long ldc);
// One operator application of a polynomial evaluation:
//   out = alpha H(x) + bx x + by y + gamma q   (+ out when accumulate)
struct PolyStage {
    const double* x = nullptr;
    double* out = nullptr;
    const double* y = nullptr;
    double alpha = 1.0, bx = 0.0, by = 0.0, gamma = 0.0;
    bool accumulate = false;
    int flags = 0;  // ST_*
};
std::vector<PolyStage> plan_poly(const ngw_context* ctx, int precision, const double* q, double* out, double* h0,
                                 double* h1, double* h2, const double* poly, int npoly, bool accumulate);

// out = p(H) q (+ out when accumulate): Horner in FP64 mode
// (apply_matvec_BLOCK.F90:250-265), Clenshaw on the Chebyshev form otherwise.
// h0, h1, h2: workspaces of the same shape as q.
int poly_apply(ngw_context* ctx, const double* q, double* out, double* h0, double* h1, double* h2, long ld,
               int nvec, const double* poly, int npoly, int tda, bool accumulate);
int operator_f64(ngw_context* ctx, const OperatorArgs& a);
// operator_f64 on one vector: library GEMMs on the (v, c) split of the pairs
int operator_f64_one(ngw_context* ctx, const OperatorArgs& a);
int pvectors_f64(ngw_context* ctx, const int* d_mband, const int* d_nband, int nvec, int tda,
                 double* d_p, long ldp_out);

// y = A x for a column-major m x n A, split over column chunks, summed in a
// fixed order; part is the chunk-sum workspace.
int gemv_wide(hipStream_t s, const double* A, long lda, int m, int n, const double* x, double* y, DevBuf& part);
// C = A B (FP64, column-major) with the k sum split into ranges when n is a
// few vectors; partials in `part`, added in range order.
int gemm_f64_split(hipStream_t s, int m, int n, int k, const double* a, long lda, const double* b, long ldb,
                   double* c, long ldc, DevBuf& part);
// The operator with an explicit pair matrix C (FP64).
int operator_explicit(ngw_context* ctx, const OperatorArgs& a);
// Releases the explicit-C form and leaves the irrep-blocked form.
void explicit_release(ngw_context* ctx);
// The irrep-blocked operator of the lowcomsym routines in factored form (FP64).
int operator_blocked(ngw_context* ctx, const OperatorArgs& a);

// Dispatch on the operator form and ctx->precision.
int operator_apply(ngw_context* ctx, const OperatorArgs& a);
int lowp_prepare(ngw_context* ctx);
// Low-precision request descriptor (the operator of one system plus one
// block of vectors) for the pre-converted-operand kernels.
struct LpOp {
    const void* pct_hi; const void* pct_lo;
    const void* pc_hi; const void* pc_lo;
    const float* m32t; const float* psiv32;
    const double* psi; long ldp;
    const float* sig;
    const float* xsig;  // x' = xsig o x: sigma_c sqrtR(v,c), or sigma_c for TDA (ncv)
    const float2* dpair;  // K3 epilogue diagonal factors (d', s') of the request's mode (FP32)
    int nmu, nv, nc, nvec, ldc1, ldm1;
    long ld; int tda;
    void* xlp_hi; void* xlp_lo; void* wlp_hi; void* wlp_lo;
    float* t32; float* u32; double* tpart;
    // FP16 only: per-vector power-of-two scales of x' (xsc, from the partial
    // maxima xpart) and of u (from umax = max |u(:, b)| as float bits),
    // per-band scales of Psi_v (vsc), scale of W row j = v + nv b
    // (wsc = vsc(v) x scale of u(:, b))
    float* xsc; unsigned* umax; const float* vsc; float* wsc; float* xpart;
    // FP32 copy of q (pairs x vectors, leading dimension nv nc) for the
    // gamma q term of the Clenshaw stages whose outputs are FP32
    float* q32;
    double* q; double* w; double* a; double* b; int niter; int* kstep;
};

// Stage of a polynomial evaluation (see PolyStage).  accumulate: 0 no,
// 1 yes, 2 only once the request's Lanczos step counter is past 1.
struct LpStage {
    const double* x; double* out; const double* y;
    double alpha, bx, by, gamma;
    int accumulate;
    int flags;  // ST_*
};

int lp2_prepare(ngw_context* ctx);
int lp2_launch_any(int kind, hipStream_t st, int c1, int c3, const LpOp* ops, const LpOp& op1, const LpStage* stg,
                   const LpStage& st1, int nstage, int s, const int* const* pre, const int* tot, int nreq,
                   ngw_context* prof);
int lp2_slots(int nv);
size_t lp2_work_bytes(int kind, int nmu, int nv, int nc, int nvec);
void lp2_fill(const ngw_context* op, int kind, int nvec, long ld, int tda, char* work, double* tpart, LpOp& d);
// Kernels per low-precision operator application (x' scale (FP16 only),
// xprep, K1, slot reduction, K2, wprep, K3): the number of workgroup prefix
// tables a grouped launch needs.
constexpr int LP_NKERNEL = 7;
void lp2_grid_counts(int kind, int c1, int c3, const LpOp& d, int* tot);
int operator_lp2(ngw_context* ctx, const OperatorArgs& a);

// Lanczos pieces (lanczos.hip).
int lanczos_run(ngw_context* ctx, double* d_q, double* d_w, double* d_h0, double* d_h1, double* d_h2,
                long ld, int nvec, int niter, const double* poly, int npoly, int tda, double* d_a,
                double* d_b);
int scale_columns(hipStream_t s, const double* in, double* out, long ld, long n, int nvec,
                  const double* d_norm);
int column_norms(hipStream_t s, const double* v, long ld, long n, int nvec, double* d_norm2,
                 double* d_norm);
int continued_fraction(hipStream_t s, const double* d_a, const double* d_b, int niter, int nvec,
                       const double* d_zz, int nz, double* d_spec);
int zz_grid(hipStream_t s, const double* d_e1, const int* d_mband, const int* d_nband, int nval,
            int nvec, int nen, double deltae, double ecuts, double* d_zz);
int sigma_assemble(hipStream_t s, const double* d_spec, const double* d_zz, const double* d_pn2, int nen, int nqp,
                   int nband, int band0, int nval, int tda, double* d_sigma);
int qp_newton(hipStream_t s, int nqp, int nen, double deltae_ev, const double* d_in,
              double* d_out, double* d_scratch);
int sigma_x_isdf(ngw_context* ctx, int nqp, const int* d_qp, double* d_sigx);
// Per-device address of the no-alloc entry's work array (noalloc.hip);
// called by ngw_create with the device current.
int noalloc_setup(int device);

}  // namespace ngw
