"""ctypes front end of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg
import this module, and only to check or time the product; the product
package ``nanogw_amd`` never imports it.

The C restatement lives in ``ngw_oracle.c`` (each function cites the NanoGW
routine it follows).  The FFT Poisson solve and the ISDF Coulomb-kernel build
are restated here with numpy's pocketfft (``poisson``, ``isdf_kernel``)
following shared/fft_module.F90z:122-177, 628-691 and
shared/isdf_parallel_sym_UltraLowMem.F90:686-875.

Pinning (tests/test_oracle_*.py):
  * lanczos/continued fraction: H_aux_v0.txt known answer, exact resolvent,
    moment identity;
  * operator, Horner, Sigma_c: dense closed forms built independently with
    numpy eigh, sum over states;
  * poisson: analytic potential of a Gaussian charge (erf form);
    isdf_kernel: absolute scale (ntrans/h^3, e^2 = 2) from the pair Coulomb
    energies of Gaussian zeta columns at two grid spacings and two ntrans,
    symmetry, positive semi-definiteness;
    sigma_x_direct: -sum_n (mn|nm) of normalised Gaussian orbitals in closed
    form (erf-form self energies of the Gaussian pair densities); in the
    exact-ISDF limit (every grid point an interpolation point) the ISDF
    exchange -sum_v C^T M C equals it to rounding;
  * spline/Newton: the reference's own newtonr / spline_evaluate compiled
    from its sources (oracle/_ref/libngw_ref_spline.so, make target
    ref-spline) agree to a few ulps; exactness on a cubic table away from the
    ends (the end slopes are finite differences), one-step Newton on linear f;
  * polyfit: numpy.linalg.lstsq on well-conditioned cases.  The npoly = 12
    LAPACK dgesvx coefficients are parity unpinned; they are pinned inputs.
  * cvt_lattice: a brute-force Lloyd step (full distance matrix, weighted
    means, nearest free grid point by a scan of all points).  The point set
    of the reference's own CVT depends on gfortran's random_number and is
    parity unpinned; points are shared inputs.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

RYD = 13.60569193  # shared/myconstants.F90:41

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def build() -> str:
    """Compile ngw_oracle.c into libngw_oracle.so (gcc + OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return os.path.join(_HERE, "libngw_oracle.so")


def build_ref() -> bool:
    """Where the reference sources are present, compile NanoGW's own Fortran
    type modules into oracle/_ref/fortran and the drop-in against them (make
    target dropin-types), and the reference's QP root finder into
    oracle/_ref/libngw_ref_spline.so (target ref-spline); False where the
    sources or the Fortran compiler are absent."""
    if not os.path.exists(os.environ.get("NGW_FC", "/opt/rocm/bin/amdflang")):
        return False
    r = subprocess.run(["make", "-s", "-C", _HERE, "ref-present"], capture_output=True, text=True)
    if r.stdout.strip() != "yes":
        return False
    subprocess.run(["make", "-s", "-C", _HERE, "dropin-types", "ref-spline"], check=True)
    return True


_REF = None


def ref_lib():
    """The reference's own newtonr / spline_evaluate (SIGMA/newtonr.f90,
    spline_evaluate.f90, spline.f90, splint.f90 compiled as they are, by the
    ref-spline make target), or None where that library was not built.
    Loaded with lazy binding: its mpi_module objects name BLAS routines that
    this path never calls."""
    global _REF
    if _REF is None:
        path = os.path.join(_HERE, "_ref", "libngw_ref_spline.so")
        if not os.path.exists(path):
            return None
        _REF = C.CDLL(path, mode=os.RTLD_LAZY | os.RTLD_LOCAL)
        _REF.newtonr_.restype = None
        _REF.spline_evaluate_.restype = None
    return _REF


def ref_spline_evaluate(xtable, ftable, xx):
    """spline_evaluate(ntable, xtable, ftable, xx, fx, df) of the reference."""
    xtable, ftable = _f64(xtable), _c128(ftable)
    n, x, df = C.c_int(len(xtable)), C.c_double(xx), C.c_double()
    fx = np.zeros(1, np.complex128)
    ref_lib().spline_evaluate_(C.byref(n), _p(xtable), _p(ftable), C.byref(x), _p(fx), C.byref(df))
    return complex(fx[0]), df.value


def ref_newtonr(xtable, ftable, x0, xacc=0.01):
    """newtonr(ntable, xtable, ftable, x_zero, xacc, fx, ierr) of the reference."""
    xtable, ftable = _f64(xtable), _c128(ftable)
    n, x, xa, ierr = C.c_int(len(xtable)), C.c_double(x0), C.c_double(xacc), C.c_int()
    fx = np.zeros(1, np.complex128)
    ref_lib().newtonr_(C.byref(n), _p(xtable), _p(ftable), C.byref(x), C.byref(xa), _p(fx), C.byref(ierr))
    return x.value, complex(fx[0]), ierr.value


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libngw_oracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.orc_matvec_block.argtypes = [C.c_int, C.c_int, C.c_int, _dp, C.c_long, _dp, C.c_long,
                                       _dp, _dp, _dp, C.c_long, C.c_int, C.c_int, _dp]
        L.orc_polynomial_matvec_block.argtypes = [C.c_int, C.c_int, C.c_int, _dp, C.c_long, _dp,
                                                  C.c_long, _dp, _dp, _dp, C.c_int, _dp, C.c_int,
                                                  C.c_int, _dp]
        L.orc_continued_fraction.argtypes = [C.c_int, _dp, _dp, C.c_int, _dp, _dp]
        L.orc_lanczos_dense.argtypes = [_dp, C.c_long, _dp, C.c_double, _dp, C.c_int, C.c_int,
                                        _dp, _dp, _dp]
        L.orc_matvec_explicit.argtypes = [C.c_int, C.c_long, _dp, C.c_long, _dp, C.c_long, _dp, _dp, _dp]
        L.orc_polynomial_matvec_explicit.argtypes = [C.c_int, C.c_long, _dp, C.c_long, _dp, C.c_long, _dp,
                                                     _dp, _dp, _dp, C.c_int]
        L.orc_lanczos_spectra_explicit.argtypes = [C.c_int, C.c_long, _dp, C.c_long, _dp, C.c_long, _dp,
                                                   _dp, C.c_double, _dp, C.c_int, C.c_int, _dp, C.c_int,
                                                   _dp, _dp, _dp, _ip]
        L.orc_lanczos_spectra_block.argtypes = [C.c_int, C.c_int, C.c_int, _dp, C.c_long, _dp,
                                                C.c_long, _dp, _dp, _dp, _dp, C.c_int, C.c_int,
                                                _dp, C.c_int, _dp, C.c_int, C.c_int, _dp, _dp,
                                                C.c_int]
        L.orc_pvectors.argtypes = [C.c_int, C.c_int, C.c_int, _dp, C.c_long, _dp, C.c_long, _dp,
                                   C.c_int, C.c_int, C.c_int, C.c_int, _dp, _dp, _dp]
        L.orc_sigma_lanczos.argtypes = [C.c_int, C.c_int, C.c_int, _dp, C.c_long, _dp, C.c_long,
                                        _dp, C.c_int, _ip, C.c_int, C.c_double, C.c_double,
                                        C.c_int, _dp, C.c_int, C.c_int, C.c_int, _dp]
        L.orc_sigma_lanczos_bands.argtypes = [C.c_int, C.c_int, C.c_int, _dp, C.c_long, _dp, C.c_long,
                                              _dp, C.c_int, _ip, C.c_int, C.c_int, C.c_int, C.c_double,
                                              C.c_double, C.c_int, _dp, C.c_int, C.c_int, C.c_int, _dp]
        L.orc_spline_evaluate.argtypes = [C.c_int, _dp, _dp, C.c_double, _dp, _dp]
        L.orc_newtonr.argtypes = [C.c_int, _dp, _dp, _dp, C.c_double, _dp]
        L.orc_newtonr.restype = C.c_int
        L.orc_qp_solve.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                   C.c_double, _dp, _dp, _dp, _dp, _dp]
        L.orc_qp_solve.restype = C.c_int
        L.orc_polyfit_sqrt.argtypes = [C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, _dp, _dp]
        L.orc_polyfit_sqrt.restype = C.c_int
        L.orc_num_threads.restype = C.c_int
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(_dp)


def _f64(a):
    return np.require(a, np.float64, ["F_CONTIGUOUS", "ALIGNED"])


def _c128(a):
    return np.require(a, np.complex128, ["F_CONTIGUOUS", "ALIGNED"])


def num_threads() -> int:
    return lib().orc_num_threads()


def sqrt_r(e1, nv, nc):
    """sqrtR(icv) = (e_c - e_v)^1/2, icv = iv + nv*ic (calculate_sigma_lanczos_BLOCK.F90z:201-208)."""
    e1 = np.asarray(e1, np.float64)
    return np.sqrt((e1[nv:nv + nc][:, None] - e1[:nv][None, :]).reshape(-1))


def matvec_block(psi, mmat, sqrtR, vec, nv, nc, tda=False):
    psi, mmat, sqrtR = _f64(psi), _f64(mmat), _f64(sqrtR)
    vec = _f64(vec.reshape(nv * nc, -1, order="F"))
    out = np.zeros_like(vec, order="F")
    nmu = psi.shape[0]
    lib().orc_matvec_block(nmu, nv, nc, _p(psi), psi.shape[0], _p(mmat), mmat.shape[0], _p(sqrtR),
                           _p(vec), _p(out), vec.shape[0], vec.shape[1], int(tda), None)
    return out


def polynomial_matvec_block(psi, mmat, sqrtR, vec, nv, nc, poly, tda=False):
    psi, mmat, sqrtR, poly = _f64(psi), _f64(mmat), _f64(sqrtR), _f64(poly)
    vec = _f64(vec.reshape(nv * nc, -1, order="F"))
    out = np.zeros_like(vec, order="F")
    nmu = psi.shape[0]
    lib().orc_polynomial_matvec_block(nmu, nv, nc, _p(psi), psi.shape[0], _p(mmat), mmat.shape[0],
                                      _p(sqrtR), _p(vec), _p(out), vec.shape[1], _p(poly),
                                      len(poly) - 1, int(tda), None)
    return out


def continued_fraction(a, b, zz):
    a, b, zz = _f64(a), _f64(b), _c128(zz)
    out = np.zeros(len(zz), np.complex128)
    lib().orc_continued_fraction(len(a), _p(a), _p(b), len(zz), _p(zz), _p(out))
    return out


def lanczos_dense(h, v0, zz, niter):
    h, v0, zz = _f64(h), _f64(v0), _c128(zz)
    n = h.shape[0]
    spec = np.zeros(len(zz), np.complex128)
    a = np.zeros(niter)
    b = np.zeros(niter)
    lib().orc_lanczos_dense(_p(h), n, _p(v0), float(np.linalg.norm(v0)), _p(zz), len(zz), niter,
                            _p(spec), _p(a), _p(b))
    return spec, a, b


def lanczos_reorth(apply, v0, zz, niter, reorth=True):
    """Lanczos with full reorthogonalisation (numpy, one column at a time).

    NanoGW has no reorthogonalisation; this is the textbook scheme (Parlett,
    The Symmetric Eigenvalue Problem, ch. 13) in the recurrence order of
    shared/lanczos_spectra_BLOCK.F90:79-108: w = A q_k - beta_k q_{k-1};
    alpha_k = w.q_k; w -= alpha_k q_k; then, with reorth, two classical
    Gram-Schmidt passes w -= Q (Q^T w) against Q = [q_1 .. q_k] ("twice is
    enough", Giraud, Langou, Rozloznik 2005); beta_{k+1} = |w|,
    q_{k+1} = w / beta_{k+1}.  b(1) = 1 as in the reference.  `apply`
    maps an (n,) vector to A times it.  Returns spectra (the reference's
    continued fraction of a, b at zz), a, b and Q (n, niter)."""
    v0 = np.asarray(v0, np.float64)
    n = v0.shape[0]
    a = np.zeros(niter)
    b = np.zeros(niter)
    Q = np.zeros((n, niter))
    q = v0 / np.linalg.norm(v0)
    qold = np.zeros(n)
    beta = 0.0
    for k in range(niter):
        w = apply(q) - beta * qold
        al = float(w @ q)
        a[k] = al
        if k == 0:
            b[0] = 1.0
        Q[:, k] = q
        if k + 1 == niter:
            break
        w = w - al * q
        if reorth:
            for _ in range(2):
                w = w - Q[:, :k + 1] @ (Q[:, :k + 1].T @ w)
        beta = float(np.linalg.norm(w))
        b[k + 1] = beta
        qold, q = q, w / beta
    return continued_fraction(a, b, zz), a, b, Q


def lanczos_spectra_block(psi, mmat, sqrtR, v0, v0_norm, zz, niter, poly, nv, nc, tda=False,
                          rebuild=False):
    """zz: (nz, blksz) complex.  Returns spectra (nz, blksz), a, b (niter, blksz)."""
    psi, mmat, sqrtR, poly = _f64(psi), _f64(mmat), _f64(sqrtR), _f64(poly)
    v0 = _f64(v0.reshape(nv * nc, -1, order="F"))
    blksz = v0.shape[1]
    v0_norm = _f64(np.asarray(v0_norm, np.float64).reshape(blksz))
    zz = _c128(np.asarray(zz).reshape(-1, blksz, order="F"))
    nz = zz.shape[0]
    spec = np.zeros((nz, blksz), np.complex128, order="F")
    a = np.zeros((niter, blksz), order="F")
    b = np.zeros((niter, blksz), order="F")
    nmu = psi.shape[0]
    lib().orc_lanczos_spectra_block(nmu, nv, nc, _p(psi), psi.shape[0], _p(mmat), mmat.shape[0],
                                    _p(sqrtR), _p(v0), _p(v0_norm), _p(zz), nz, niter, _p(poly),
                                    len(poly) - 1, _p(spec), blksz, int(tda), _p(a), _p(b),
                                    int(rebuild))
    return spec, a, b


def matvec_explicit(cm, mmat, sqrtR, vec):
    """matvec_isdf on a caller-built Cmtrx (nmu, ncv): one vector."""
    cm, mmat, sqrtR, vec = _f64(cm), _f64(mmat), _f64(sqrtR), _f64(vec)
    nmu, ncv = cm.shape
    out = np.zeros(ncv)
    lib().orc_matvec_explicit(nmu, ncv, _p(cm), nmu, _p(mmat), mmat.shape[0], _p(sqrtR), _p(vec), _p(out))
    return out


def polynomial_matvec_explicit(cm, mmat, sqrtR, vec, poly):
    cm, mmat, sqrtR, vec, poly = _f64(cm), _f64(mmat), _f64(sqrtR), _f64(vec), _f64(poly)
    nmu, ncv = cm.shape
    out = np.zeros(ncv)
    lib().orc_polynomial_matvec_explicit(nmu, ncv, _p(cm), nmu, _p(mmat), mmat.shape[0], _p(sqrtR), _p(vec),
                                         _p(out), _p(poly), len(poly) - 1)
    return out


def lanczos_spectra_explicit(cm, mmat, sqrtR, v0, v0_norm, zz, niter, poly):
    """lanczos_spectra_isdf (single vector, breakdown exit).  Returns
    spectra (nz,), a, b (niter,), and the number of fraction levels used."""
    cm, mmat, sqrtR, v0, poly = _f64(cm), _f64(mmat), _f64(sqrtR), _f64(v0), _f64(poly)
    zz = _c128(np.asarray(zz).reshape(-1))
    nmu, ncv = cm.shape
    spec = np.zeros(len(zz), np.complex128)
    a, b = np.zeros(niter), np.zeros(niter)
    nk = C.c_int(0)
    lib().orc_lanczos_spectra_explicit(nmu, ncv, _p(cm), nmu, _p(mmat), mmat.shape[0], _p(sqrtR), _p(v0),
                                       float(v0_norm), _p(zz), len(zz), int(niter), _p(poly), len(poly) - 1,
                                       _p(spec), _p(a), _p(b), C.byref(nk))
    return spec, a, b, nk.value


def pair_cmtrx(psi, nv, nc):
    """Cmtrx(mu, icv) = psi(mu, c) psi(mu, v) in the pair order icv = iv + nv*ic."""
    psi = np.asarray(psi, np.float64)
    nmu = psi.shape[0]
    return np.asfortranarray((psi[:, nv:nv + nc][:, :, None] * psi[:, :nv][:, None, :]).reshape(nmu, nc * nv))


def pvectors(psi, mmat, sqrtR, m, n0, blksz, nv, nc, tda=False):
    psi, mmat, sqrtR = _f64(psi), _f64(mmat), _f64(sqrtR)
    nmu = psi.shape[0]
    p = np.zeros((nv * nc, blksz), order="F")
    pn2 = np.zeros(blksz)
    # cm(mu, icv), icv = iv + nv*ic: the (nmu, nc, nv) product reshaped in C order
    cm = np.asfortranarray((psi[:, nv:nv + nc][:, :, None] * psi[:, :nv][:, None, :]).reshape(nmu, nc * nv))
    lib().orc_pvectors(nmu, nv, nc, _p(psi), psi.shape[0], _p(mmat), mmat.shape[0], _p(sqrtR),
                       int(m), int(n0), int(blksz), int(tda), _p(cm), _p(p), _p(pn2))
    return p, pn2


def sigma_lanczos(psi, mmat, e1, nv, nc, qp, nen, deltae, ecuts, niter, poly, jnblock, tda=False, bands=None):
    """Sigma_c(omega) (nen, nqp) complex in Ry for 0-based QP band indices `qp`;
    bands = (lo, hi) keeps only Green's-function bands lo <= n < hi."""
    psi, mmat, e1, poly = _f64(psi), _f64(mmat), _f64(e1), _f64(poly)
    qp = np.ascontiguousarray(qp, np.int32)
    out = np.zeros((nen, len(qp)), np.complex128, order="F")
    nmu = psi.shape[0]
    if bands is None:
        lib().orc_sigma_lanczos(nmu, nv, nc, _p(psi), psi.shape[0], _p(mmat), mmat.shape[0], _p(e1),
                                len(qp), qp.ctypes.data_as(_ip), nen, float(deltae), float(ecuts),
                                int(niter), _p(poly), len(poly) - 1, int(jnblock), int(tda), _p(out))
    else:
        lib().orc_sigma_lanczos_bands(nmu, nv, nc, _p(psi), psi.shape[0], _p(mmat), mmat.shape[0], _p(e1),
                                      len(qp), qp.ctypes.data_as(_ip), int(bands[0]), int(bands[1]), nen,
                                      float(deltae), float(ecuts), int(niter), _p(poly), len(poly) - 1,
                                      int(jnblock), int(tda), _p(out))
    return out


def spline_evaluate(xtable, ftable, xx):
    xtable, ftable = _f64(xtable), _c128(ftable)
    fx = np.zeros(1, np.complex128)
    df = np.zeros(1)
    lib().orc_spline_evaluate(len(xtable), _p(xtable), _p(ftable), float(xx), _p(fx), _p(df))
    return complex(fx[0]), float(df[0])


def newtonr(xtable, ftable, x0, xacc=0.01):
    xtable, ftable = _f64(xtable), _c128(ftable)
    x = np.array([x0], np.float64)
    fx = np.zeros(1, np.complex128)
    ierr = lib().orc_newtonr(len(xtable), _p(xtable), _p(ftable), _p(x), float(xacc), _p(fx))
    return float(x[0]), complex(fx[0]), ierr


def qp_solve(sigc_ev, deltae_ev, edft, en0, sigx, vxc):
    """printsigma.F90:97-128 for one state; all energies in eV."""
    sigc = _c128(sigc_ev)
    out = np.zeros(3)
    s0 = np.zeros(1, np.complex128)
    e = np.zeros(1); z = np.zeros(1); im = np.zeros(1)
    ierr = lib().orc_qp_solve(len(sigc), float(deltae_ev), float(edft), float(en0), float(sigx),
                              float(vxc), _p(sigc), _p(e), _p(z), _p(s0), _p(im))
    del out
    return {"eqp": float(e[0]), "zfac": float(z[0]), "sigma0": complex(s0[0]),
            "imsig": float(im[0]), "ierr": int(ierr)}


def polyfit_sqrt(xmin, xmax, npoly, nx=20, cheb=False):
    """Returns (beta low->high, fit_error); raises on the reference's die() ranges."""
    beta = np.zeros(npoly + 1)
    fe = np.zeros(1)
    rc = lib().orc_polyfit_sqrt(float(xmin), float(xmax), int(npoly), int(nx), int(cheb), _p(beta),
                                _p(fe))
    if rc != 0:
        raise ValueError("polyfit_sqrt: xmin too small")
    return beta, float(fe[0])


# ---------------------------------------------------------------------------
# FFT Poisson solve and ISDF Coulomb kernel (numpy restatement)
# ---------------------------------------------------------------------------

QZERO = 1.0e-6  # shared/fft_module.F90z:63


def coulomb_0d(nfft, h, trunc):
    """v(G) on the rfftn half box, shared/fft_module.F90z:628-691 (Ry, e^2 = 2)."""
    n1, n2, n3 = nfft
    b = [2.0 * np.pi / (n * h) for n in nfft]  # parsec_wfn.F90 bvec(ii,ii) = 2 pi/(nfft h)
    j1 = np.fft.fftfreq(n1, 1.0 / n1)[: n1 // 2 + 1]
    j1 = np.abs(j1)
    j2 = np.fft.fftfreq(n2, 1.0 / n2)
    j3 = np.fft.fftfreq(n3, 1.0 / n3)
    g2 = ((b[0] * j1)[None, None, :] ** 2 + (b[1] * j2)[None, :, None] ** 2 +
          (b[2] * j3)[:, None, None] ** 2)
    # axes ordered (z, y, x) so that rfftn's last axis is x
    with np.errstate(divide="ignore", invalid="ignore"):
        if trunc > 0:
            v = np.where(g2 < QZERO * QZERO, trunc ** 2 / 2.0,
                         (1.0 - np.cos(trunc * np.sqrt(g2))) / g2)
        else:
            v = np.where(g2 < QZERO * QZERO, 0.0, 1.0 / g2)
    return 8.0 * np.pi * v


def poisson(rho, gidx, nfft, h, trunc, vcoul=None):
    """V(r) = IFFT(FFT(rho) v(G)) / N on the grid points (fft_module.F90z:122-177).

    rho: (ngrid,) or (ngrid, nvec); gidx: (ngrid, 3) integer coordinates in
    [-nfft/2, nfft/2).
    """
    if vcoul is None:
        vcoul = coulomb_0d(nfft, h, trunc)
    rho = np.asarray(rho, np.float64)
    single = rho.ndim == 1
    if single:
        rho = rho[:, None]
    n1, n2, n3 = nfft
    ix = np.mod(gidx[:, 0], n1)
    iy = np.mod(gidx[:, 1], n2)
    iz = np.mod(gidx[:, 2], n3)
    out = np.empty_like(rho)
    for k in range(rho.shape[1]):
        box = np.zeros((n3, n2, n1))
        box[iz, iy, ix] = rho[:, k]
        f = np.fft.rfftn(box)
        f *= vcoul
        box = np.fft.irfftn(f, s=(n3, n2, n1), axes=(0, 1, 2))
        out[:, k] = box[iz, iy, ix]
    return out[:, 0] if single else out


def isdf_kernel(zeta, gidx, nfft, h, trunc, ntrans=1):
    """M = (ntrans/h^3) zeta^T (V zeta), isdf_parallel_sym_UltraLowMem.F90:312, 749, 865-867."""
    vz = poisson(zeta, gidx, nfft, h, trunc)
    return (ntrans / h ** 3) * (zeta.T @ vz)


TOL_OCC = 0.01  # shared/myconstants.F90:51


def sigma_x_direct(psi, occ, gidx, nfft, h, trunc, qp, ntrans=1, nk=1):
    """Bare exchange on the grid, SIGMA/fock_exchange.F90z (0D, real orbitals,
    one spin and k-point): the occupied states are those with occ1 > tol_occ
    (:107-111); for each state m and occupied n the pair density
    rho = psi_n psi_m goes through Zpoisson (:117-158), and the kernel
    -<psi_m | psi_n V[rho]> ntrans/h^3 (fac, :51, :174) is added with weight
    occ1(n)/nk (:175-176).  psi: (ngrid, nb) with sum_r psi^2 = 1 (the
    reference's wave-function normalisation); occ: (nb,).  Returns
    Sigma_x(m) in Ry for m in qp."""
    vc = coulomb_0d(nfft, h, trunc)
    occ = np.asarray(occ, np.float64)
    nocc = [n for n in range(psi.shape[1]) if occ[n] > TOL_OCC]
    out = []
    for m in qp:
        rho = psi[:, nocc] * psi[:, [m]]
        v = poisson(rho, gidx, nfft, h, trunc, vcoul=vc)
        out.append(-(ntrans / h ** 3) * float(np.sum((rho * v).sum(axis=0) * occ[nocc] / nk)))
    return np.array(out)


def _snap_one(table, lo, taken, c):
    """Nearest free grid point to centroid c (lattice units): cubic shells
    s = 0, 1, ... around the rounded centroid, each scanned in (z, y, x)
    order, first point at the smallest distance kept (cvt.F90:104-128 picks
    the nearest unused grid point; this fixes the order among ties)."""
    r = [int(np.floor(v + 0.5)) if v >= 0 else -int(np.floor(-v + 0.5)) for v in c]  # C lround
    n = table.shape[::-1]
    best, bd = -1, 1e300
    s = 0
    while True:
        for dk in range(-s, s + 1):
            for dj in range(-s, s + 1):
                for di in range(-s, s + 1):
                    if max(abs(di), abs(dj), abs(dk)) != s:
                        continue
                    i, j, k = r[0] + di - lo[0], r[1] + dj - lo[1], r[2] + dk - lo[2]
                    if i < 0 or j < 0 or k < 0 or i >= n[0] or j >= n[1] or k >= n[2]:
                        continue
                    g = int(table[k, j, i])
                    if g < 0 or taken[g]:
                        continue
                    ex, ey, ez = (r[0] + di) - c[0], (r[1] + dj) - c[1], (r[2] + dk) - c[2]
                    d = (ex * ex + ey * ey) + ez * ez
                    if d < bd:
                        bd, best = d, g
        if best >= 0 and s + 0.5 > np.sqrt(bd):
            return best
        if s > 4 * sum(n):
            return best
        s += 1


def cvt_lattice(gidx, rho, init, h, tol=1e-3, max_iter=300):
    """rho-weighted Lloyd iteration of shared/cvt.F90:8-195 in lattice units:
    nearest-centroid assignment (:77-88; lowest index on ties), weighted
    centroids with the old position kept below weight 1e-10 (:92-103), snap
    in index order to the nearest free grid point (:104-128), stop when the
    summed displacement (bohr) is below tol (:131-135).  Small grids only
    (dense distance blocks, Python snap loop).  Returns (intp, iterations)."""
    g = np.asarray(gidx, np.int64)
    gf = g.astype(np.float64)
    rho = np.asarray(rho, np.float64)
    lo = g.min(axis=0)
    n = g.max(axis=0) - lo + 1
    table = -np.ones((n[2], n[1], n[0]), np.int64)
    table[g[:, 2] - lo[2], g[:, 1] - lo[1], g[:, 0] - lo[0]] = np.arange(len(g))
    cen = np.asarray(init, np.float64) / h
    old = cen.copy()
    nc = len(cen)
    intp = np.zeros(nc, np.int64)
    it = 0
    for it in range(1, max_iter + 1):
        owner = np.empty(len(g), np.int64)
        for r0 in range(0, len(g), 4096):
            p = gf[r0:r0 + 4096]
            dx = p[:, None, 0] - cen[None, :, 0]
            dy = p[:, None, 1] - cen[None, :, 1]
            dz = p[:, None, 2] - cen[None, :, 2]
            owner[r0:r0 + 4096] = np.argmin((dx * dx + dy * dy) + dz * dz, axis=1)
        w = np.bincount(owner, weights=rho, minlength=nc)
        wx = [np.bincount(owner, weights=rho * gf[:, k], minlength=nc) for k in range(3)]
        taken = np.zeros(len(g), bool)
        diff = 0.0
        new = np.empty_like(cen)
        for c in range(nc):
            cand = [wx[k][c] / w[c] for k in range(3)] if w[c] >= 1e-10 else list(old[c])
            gi = _snap_one(table, lo, taken, cand)
            taken[gi] = True
            intp[c] = gi
            new[c] = gf[gi]
            diff += np.sqrt(((new[c] - old[c]) ** 2).sum()) * h
        old = new.copy()
        cen = new
        if diff < tol:
            break
    return intp, it
