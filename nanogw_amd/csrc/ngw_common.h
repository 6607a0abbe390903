// Shared definitions of the gfx950 ISDF-Lanczos library (libngw).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/ngw.h"

namespace ngw {

// Thread-local last-error text returned by ngw_last_error().
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

#define NGW_HIP(call)                                                                    \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess)                                                            \
            return ::ngw::fail(NGW_ERR_HIP, std::string(#call " -> ") + hipGetErrorString(e_)); \
    } while (0)

#define NGW_TRY(call)                 \
    do {                              \
        int s_ = (call);              \
        if (s_ != NGW_OK) return s_;  \
    } while (0)

// Device buffer that grows on demand and is freed with its owner (movable,
// not copyable).
struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : ptr(o.ptr), bytes(o.bytes) { o.ptr = nullptr; o.bytes = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept
    {
        if (this != &o) {
            release();
            ptr = o.ptr; bytes = o.bytes;
            o.ptr = nullptr; o.bytes = 0;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    int ensure(size_t need);
    void release();
    template <class T>
    T* as() const { return static_cast<T*>(ptr); }
};

// Kernel classes timed by the optional per-launch HIP-event profiler.
enum KernelClass { KC_K1 = 0, KC_K2, KC_K3, KC_LANCZOS, KC_CFRAC, KC_PVEC, KC_SIGMA, KC_CONVERT, KC_COUNT };

// Precision of the three GEMM stages of the operator.
enum Precision { PREC_F64 = 0, PREC_F32 = 1, PREC_BF16 = 2, PREC_BF16X3 = 3, PREC_F16 = 4, PREC_COUNT = 5 };

}  // namespace ngw

// Device-resident state of one ISDF operator (what type(ISDF) carries for the
// Lanczos path): Psi at the ISDF points, the Coulomb kernel M, energies.
struct ngw_context {
    int device = 0;
    hipStream_t stream = nullptr;
    bool shared_stream = false;  // stream owned by a batch, not by this context
    // problem
    int n_intp = 0, nband = 0, nval = 0, ncond = 0;
    long ncv = 0;
    int precision = ngw::PREC_F64;
    // Lanczos reorthogonalisation (NGW_REORTH_*): off in the reference path;
    // full keeps every Lanczos vector in qbasis (niter x ncv x columns FP64)
    int reorth = 0;
    ngw::DevBuf qbasis;
    ngw::DevBuf psi;      // n_intp x nband, column-major, ld = ld_psi
    int ld_psi = 0;
    ngw::DevBuf mmat;     // n_intp x n_intp, ld = ld_m
    int ld_m = 0;
    ngw::DevBuf e1;       // nband energies (Ry)
    std::vector<double> e1_host;
    ngw::DevBuf sqrtr;    // ncv: (e_c - e_v)^1/2
    ngw::DevBuf diag_full;  // ncv: sqrtR^3   (applied to sqrtR*x)
    ngw::DevBuf diag_tda;   // ncv: sqrtR^2   (applied to x)
    // low-precision operator copies (FP32 / BF16 with per-band scales)
    ngw::DevBuf psi_scale;
    bool lowp_ready = false;
    // pre-converted operands of the streamlined low-precision path
    bool lp2_ready = false;
    int ldc1 = 0, ldm1 = 0;
    ngw::DevBuf pct32, pc32, pctb, pcb, pcth, pch, m32t, psiv32, psiv_scale, xsig, dpairs, lpwork;
    // chosen {K1, K3} tile configuration per precision mode (ngw_autotune)
    int lp_cfg[ngw::PREC_COUNT][2] = {};
    // explicit pair matrix C (n_intp x ncv, ngw_set_explicit_c) in place of
    // the factored Psi_v o Psi_c, and the workspaces of the explicit and
    // one-vector products (scaled input, partial sums of split sums)
    bool explicit_c = false;
    ngw::DevBuf cexp, ex_x, ex_y, ex_t, ex_s;
    // irrep-blocked pairs (ngw_set_isdf_blocks): psi holds the blocks'
    // valence orbitals, then their conduction orbitals; block k pairs
    // blk_nv[k] valence with blk_nc[k] conduction states
    bool blocked = false;
    std::vector<int> blk_nv, blk_nc;
    ngw::DevBuf gv_part;  // chunk sums of the one-vector GEMV (gemv_wide)
    // workspaces
    ngw::DevBuf tpart, ubuf, w0, w1, w2, w3, w4, ab, zz, spec, misc, misc2, staging;
    // host staging for drop-in calls
    std::vector<double> hbuf;
    // per-launch event profiler (off unless ngw_profile_enable)
    bool prof_on = false;
    std::vector<hipEvent_t> prof_pool;
    struct ProfRec { int kind; hipEvent_t a, b; };
    std::vector<ProfRec> prof_recs;
    double prof_ms[ngw::KC_COUNT] = {0};
    long prof_cnt[ngw::KC_COUNT] = {0};
};

namespace ngw {
// Records HIP events around one kernel launch on ctx->stream when profiling.
struct ProfScope {
    ngw_context* c; int kind; hipEvent_t a = nullptr, b = nullptr;
    ProfScope(ngw_context* ctx, int k);
    ~ProfScope();
};
}  // namespace ngw
