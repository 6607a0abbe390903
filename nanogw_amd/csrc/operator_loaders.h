// Operand loaders of the FP64 ISDF operator GEMMs.  row(mn) returns a per-row
// handle; fetch() issues the raw loads of element (mn, k) from clamped
// addresses; combine() forms the element.
#pragma once

#include "gemm_core_lowp.h"

namespace ngw {

// ---------------------------------------------------------------- FP64 loaders
struct ColMajorA {  // A(m, k) = a[m + k*ld]
    static constexpr bool kContig = false;
    static constexpr int NRAW = 1;
    typedef const double* Handle;
    const double* a; long ld; int MN;
    __device__ Handle row(int m) const { return a + m; }
    __device__ void fetch(Handle h, int k, double (&r)[1]) const { r[0] = h[(long)k * ld]; }
    __device__ double combine(const double (&r)[1]) const { return r[0]; }
};
struct ColMajorB {  // B(k, n) = b[k + n*ld]
    static constexpr bool kContig = true;
    static constexpr int NRAW = 1;
    typedef const double* Handle;
    const double* b; long ld; int MN;
    __device__ Handle row(int n) const { return b + (long)n * ld; }
    __device__ void fetch(Handle h, int k, double (&r)[1]) const { r[0] = h[k]; }
    __device__ double combine(const double (&r)[1]) const { return r[0]; }
};
struct ColMajorAT {  // A(m, k) = a[k + m*ld]
    static constexpr bool kContig = true;
    static constexpr int NRAW = 1;
    typedef const double* Handle;
    const double* a; long ld; int MN;
    __device__ Handle row(int m) const { return a + (long)m * ld; }
    __device__ void fetch(Handle h, int k, double (&r)[1]) const { r[0] = h[k]; }
    __device__ double combine(const double (&r)[1]) const { return r[0]; }
};
struct XprimeB {  // x'(c, j = v + nv b) = s(v,c) x_b(v,c); s == null means 1 (TDA)
    static constexpr bool kContig = false;
    static constexpr int NRAW = 2;
    struct Handle { const double* x; const double* s; };
    const double* x; const double* s; long ldx; int nv; int MN;
    __device__ Handle row(int j) const
    {
        int b = j / nv, v = j - b * nv;
        return {x + (long)b * ldx + v, s ? s + v : nullptr};
    }
    __device__ void fetch(const Handle& h, int c, double (&r)[2]) const
    {
        long o = (long)c * nv;
        r[0] = h.x[o];
        r[1] = h.s ? h.s[o] : 1.0;
    }
    __device__ double combine(const double (&r)[2]) const { return r[1] * r[0]; }
};
struct WeightedPsiVB {  // W(mu, j = v + nv b) = PsiV(mu, v) u(mu, b)
    static constexpr bool kContig = true;
    static constexpr int NRAW = 2;
    struct Handle { const double* pv; const double* ub; };
    const double* psiv; long ldp; const double* u; long ldu; int nv; int MN;
    __device__ Handle row(int j) const
    {
        int b = j / nv, v = j - b * nv;
        return {psiv + (long)v * ldp, u + (long)b * ldu};
    }
    __device__ void fetch(const Handle& h, int mu, double (&r)[2]) const
    {
        r[0] = h.pv[mu];
        r[1] = h.ub[mu];
    }
    __device__ double combine(const double (&r)[2]) const { return r[0] * r[1]; }
};
struct PairB {  // c_mn(nu) = Psi(nu, m_b) Psi(nu, n_b)  (P-vector build)
    static constexpr bool kContig = true;
    static constexpr int NRAW = 2;
    struct Handle { const double* pm; const double* pn; };
    const double* psi; long ld; const int* mband; const int* nband; int MN;
    __device__ Handle row(int b) const { return {psi + (long)mband[b] * ld, psi + (long)nband[b] * ld}; }
    __device__ void fetch(const Handle& h, int nu, double (&r)[2]) const
    {
        r[0] = h.pm[nu];
        r[1] = h.pn[nu];
    }
    __device__ double combine(const double (&r)[2]) const { return r[0] * r[1]; }
};

}  // namespace ngw
