"""Host-side mirror of NanoGW's Lanczos GW interfaces on top of libngw.

Names, argument meaning and order follow the Fortran routines each method
replaces; arrays are numpy, Fortran (column-major) order, energies in Ry.

    ctx = IsdfContext(device=0)
    ctx.set_isdf(psi_intp, mmtrx, e1, nval)           # type(ISDF) fields
    hvec = ctx.matvec_isdf_BLOCK(vec)                  # apply_matvec_BLOCK.F90:25
    hvec = ctx.polynomial_matvec_isdf_BLOCK(vec, polynomial)      # :210
    spectra, a, b = ctx.lanczos_spectra_isdf_BLOCK(v0, v0_norm, zz, niter, polynomial)
    sigma_c = ctx.calculate_sigma_lanczos(qp, nen, deltae, ecuts, niter, polynomial)
    qp = ctx.qp_newton(...)                            # printsigma.F90:97-128

Errors raise ``NgwError`` (NanoGW calls die()).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib

PREC = {"f64": 0, "f32": 1, "bf16": 2, "bf16x3": 3, "f16": 4}


def _f64(a):
    return np.require(a, np.float64, ["F_CONTIGUOUS", "ALIGNED"])


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def polyfit_sqrt(xmin, xmax, npoly, nx=20, chebyshev=False):
    """shared/polyfit.f90:1 (chebyshev: polyfitcheb.f90:1).  Returns (beta low->high, fit_error)."""
    beta = np.zeros(npoly + 1)
    fe = np.zeros(1)
    _lib.check(_lib.load().ngw_polyfit_sqrt(float(xmin), float(xmax), int(npoly), int(nx),
                                            int(bool(chebyshev)), _p(beta), _p(fe)), "polyfit_sqrt")
    return beta, float(fe[0])


def lanczos_polynomial(e1, nval, npoly=12, nx=20, chebyshev=False):
    """The driver's sqrt polynomial, highest order first
    (calculate_sigma_lanczos_BLOCK.F90z:161-177)."""
    e1 = np.asarray(e1, np.float64)
    nb = len(e1)
    emin = (e1[nval] - e1[nval - 1]) ** 2.0 / 1.25
    emax = (e1[nb - 1] - e1[0]) ** 2.0 * 1.25
    beta, _ = polyfit_sqrt(emin, emax, npoly, nx, chebyshev)
    return beta[::-1].copy()


class IsdfContext:
    """Device-resident ISDF operator and Lanczos engine (one HIP stream)."""

    def __init__(self, device: int = 0):
        self._lib = _lib.load()
        h = C.c_void_p()
        _lib.check(self._lib.ngw_create(C.byref(h), int(device)), "ngw_create")
        self._h = h
        self.nval = self.ncond = self.n_intp = self._ncv = 0

    def close(self):
        if getattr(self, "_h", None):
            self._lib.ngw_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        return self._h

    # ------------------------------------------------------------ operator
    def set_isdf(self, psi_intp, mmtrx, e1, nval):
        psi_intp, mmtrx, e1 = _f64(psi_intp), _f64(mmtrx), _f64(e1)
        n_intp, nband = psi_intp.shape
        _lib.check(self._lib.ngw_set_isdf(self._h, n_intp, nband, int(nval), _p(psi_intp), n_intp,
                                          _p(mmtrx), mmtrx.shape[0], _p(e1)), "ngw_set_isdf")
        self.n_intp, self.nval, self.ncond = n_intp, int(nval), nband - int(nval)
        self.nband = nband
        self.e1 = e1.copy()
        self._ncv = self.nval * self.ncond

    def set_explicit_c(self, cmtrx, mmtrx, sqrt_r):
        """The operator of matvec_isdf (shared/apply_matvec.f90:25): the pair
        matrix C (n_intp, ncv), Mmtrx and sqrtR given explicitly."""
        cmtrx, mmtrx, sqrt_r = _f64(cmtrx), _f64(mmtrx), _f64(sqrt_r)
        n_intp, ncv = cmtrx.shape
        if sqrt_r.shape != (ncv,) or mmtrx.shape != (n_intp, n_intp):
            raise ValueError("C is (n_intp, ncv), Mmtrx (n_intp, n_intp), sqrtR (ncv,)")
        _lib.check(self._lib.ngw_set_explicit_c(self._h, n_intp, ncv, _p(cmtrx), n_intp, _p(mmtrx), n_intp,
                                                _p(sqrt_r)), "ngw_set_explicit_c")
        self.n_intp, self._ncv = n_intp, ncv
        self.nval = self.ncond = self.nband = 0
        self.e1 = None

    def set_isdf_blocks(self, psiv, psic, nv_blk, nc_blk, mmtrx, sqrt_r):
        """The irrep-blocked operator of matvec_isdf_lowcomsym
        (shared/apply_matvec_lowcomsym.F90:155-412) in factored form: psiv
        (n_intp, sum nv_blk) and psic (n_intp, sum nc_blk) hold the blocks'
        orbitals block after block; pairs are numbered block after block,
        valence outer, conduction inner; sqrt_r (ncv,) in that order."""
        psiv, psic, mmtrx, sqrt_r = _f64(psiv), _f64(psic), _f64(mmtrx), _f64(sqrt_r)
        nv_blk = np.ascontiguousarray(nv_blk, dtype=np.int32)
        nc_blk = np.ascontiguousarray(nc_blk, dtype=np.int32)
        n_intp = psiv.shape[0]
        ncv = int(np.dot(nv_blk.astype(np.int64), nc_blk.astype(np.int64)))
        if (nv_blk.shape != nc_blk.shape or psiv.shape != (n_intp, int(nv_blk.sum()))
                or psic.shape != (n_intp, int(nc_blk.sum())) or mmtrx.shape != (n_intp, n_intp)
                or sqrt_r.shape != (ncv,)):
            raise ValueError("psiv (n_intp, sum nv_blk), psic (n_intp, sum nc_blk), Mmtrx (n_intp, n_intp), "
                             "sqrtR (sum nv_blk nc_blk,)")
        ip = C.POINTER(C.c_int)
        _lib.check(self._lib.ngw_set_isdf_blocks(self._h, n_intp, len(nv_blk), nv_blk.ctypes.data_as(ip),
                                                 nc_blk.ctypes.data_as(ip), _p(psiv), n_intp, _p(psic), n_intp,
                                                 _p(mmtrx), n_intp, _p(sqrt_r)), "ngw_set_isdf_blocks")
        self.n_intp, self._ncv = n_intp, ncv
        self.nval = self.ncond = self.nband = 0
        self.e1 = None

    def set_precision(self, precision):
        p = PREC[precision] if isinstance(precision, str) else int(precision)
        _lib.check(self._lib.ngw_set_precision(self._h, p), "ngw_set_precision")

    def set_reorthogonalization(self, full: bool):
        """Full reorthogonalisation of the Lanczos recurrences (no NanoGW
        counterpart; off = the reference algorithm)."""
        _lib.check(self._lib.ngw_set_reorthogonalization(self._h, 1 if full else 0),
                   "ngw_set_reorthogonalization")

    @property
    def ncv(self):
        return self._ncv

    def _vecs(self, vec):
        v = _f64(np.asarray(vec, np.float64).reshape(self.ncv, -1, order="F"))
        return v, v.shape[1]

    def matvec_isdf_BLOCK(self, vec, tamm_d=False):
        v, b = self._vecs(vec)
        out = np.zeros_like(v, order="F")
        _lib.check(self._lib.ngw_matvec_isdf_block(self._h, _p(v), _p(out), self.ncv, b, int(tamm_d)),
                   "ngw_matvec_isdf_block")
        return out

    def polynomial_matvec_isdf_BLOCK(self, vec, polynomial, tamm_d=False):
        v, b = self._vecs(vec)
        poly = _f64(polynomial)
        out = np.zeros_like(v, order="F")
        _lib.check(self._lib.ngw_polynomial_matvec_isdf_block(self._h, _p(v), _p(out), self.ncv,
                                                              _p(poly), len(poly) - 1, b,
                                                              int(tamm_d)),
                   "ngw_polynomial_matvec_isdf_block")
        return out

    def lanczos_spectra_isdf_BLOCK(self, v0, v0_norm, zz, niter, polynomial, tamm_d=False):
        """Returns spectra (nz, blksz) complex and the a, b coefficients (niter, blksz)."""
        v, b = self._vecs(v0)
        vn = _f64(np.asarray(v0_norm, np.float64).reshape(b))
        zz = np.require(np.asarray(zz, np.complex128).reshape(-1, b, order="F"), np.complex128,
                        ["F_CONTIGUOUS"])
        nz = zz.shape[0]
        poly = _f64(polynomial)
        spec = np.zeros((nz, b), np.complex128, order="F")
        a = np.zeros((niter, b), order="F")
        bb = np.zeros((niter, b), order="F")
        _lib.check(self._lib.ngw_lanczos_spectra_isdf_block(
            self._h, _p(v), _p(vn), self.ncv, zz.ctypes.data_as(C.POINTER(C.c_double)), nz,
            int(niter), _p(poly), len(poly) - 1, spec.ctypes.data_as(C.POINTER(C.c_double)), b,
            int(tamm_d), _p(a), _p(bb)), "ngw_lanczos_spectra_isdf_block")
        return spec, a, bb

    def calculate_sigma_lanczos(self, qp, nen, deltae, ecuts, niter, polynomial, tamm_d=False,
                                max_cols=0, bands=None):
        """Sigma_c(z) (nen, nqp) complex in Ry for 0-based QP bands `qp`
        (calculate_sigma_lanczos_BLOCK.F90z:201-444).  bands = (lo, hi)
        restricts the sum to Green's-function bands lo <= n < hi."""
        qp = np.ascontiguousarray(qp, np.int32)
        poly = _f64(polynomial)
        out = np.zeros((nen, len(qp)), np.complex128, order="F")
        qpp = qp.ctypes.data_as(C.POINTER(C.c_int))
        outp = out.ctypes.data_as(C.POINTER(C.c_double))
        if bands is None:
            _lib.check(self._lib.ngw_sigma_lanczos(
                self._h, len(qp), qpp, int(nen), float(deltae), float(ecuts), int(niter), _p(poly),
                len(poly) - 1, int(tamm_d), int(max_cols), outp), "ngw_sigma_lanczos")
        else:
            lo, hi = bands
            _lib.check(self._lib.ngw_sigma_lanczos_bands(
                self._h, len(qp), qpp, int(lo), int(hi), int(nen), float(deltae), float(ecuts), int(niter),
                _p(poly), len(poly) - 1, int(tamm_d), int(max_cols), outp), "ngw_sigma_lanczos_bands")
        return out

    def sigma_x(self, qp):
        """Bare exchange (Ry) of 0-based QP bands in the ISDF representation
        (the diagonal of fock_exchange.F90z:12-229 with the operator's M)."""
        qp = np.ascontiguousarray(qp, np.int32)
        out = np.zeros(len(qp))
        _lib.check(self._lib.ngw_sigma_x_isdf(self._h, len(qp), qp.ctypes.data_as(C.POINTER(C.c_int)), _p(out)),
                   "ngw_sigma_x_isdf")
        return out

    def qp_newton(self, sigma_c_ev, deltae_ev, edft_ev, en0_ev, sigx_ev, vxc_ev):
        """printsigma.F90:97-128 on the device; all energies eV; sigma_c (nen, nqp)."""
        sc = np.require(np.asarray(sigma_c_ev, np.complex128), np.complex128, ["F_CONTIGUOUS"])
        nen, nqp = sc.shape
        vals = [_f64(np.asarray(x, np.float64).reshape(nqp)) for x in (edft_ev, en0_ev, sigx_ev, vxc_ev)]
        eqp, zfac, imsig = np.zeros(nqp), np.zeros(nqp), np.zeros(nqp)
        s0 = np.zeros(nqp, np.complex128)
        ierr = np.zeros(nqp, np.int32)
        _lib.check(self._lib.ngw_qp_newton(
            self._h, nqp, nen, float(deltae_ev), *[_p(v) for v in vals],
            sc.ctypes.data_as(C.POINTER(C.c_double)), _p(eqp), _p(zfac),
            s0.ctypes.data_as(C.POINTER(C.c_double)), _p(imsig),
            ierr.ctypes.data_as(C.POINTER(C.c_int))), "ngw_qp_newton")
        return {"eqp": eqp, "zfac": zfac, "sigma0": s0, "imsig": imsig, "ierr": ierr}

    # ------------------------------------------------------------ V build
    def isdf_coulomb_kernel(self, zeta, gidx, nfft, h, trunc, ntrans=1.0):
        """M = (ntrans/h^3) zeta^T V zeta (isdf_parallel_sym_UltraLowMem.F90:686-875)."""
        zeta = _f64(zeta)
        ngrid, n_intp = zeta.shape
        grid = np.ascontiguousarray(np.asarray(gidx, np.int32).reshape(ngrid, 3))
        nf = np.ascontiguousarray(np.asarray(nfft, np.int32).reshape(3))
        out = np.zeros((n_intp, n_intp), order="F")
        _lib.check(self._lib.ngw_isdf_coulomb_kernel(
            self._h, ngrid, n_intp, _p(zeta), zeta.shape[0], grid.ctypes.data_as(C.POINTER(C.c_int)),
            nf.ctypes.data_as(C.POINTER(C.c_int)), float(h), float(trunc), float(ntrans), _p(out),
            n_intp), "ngw_isdf_coulomb_kernel")
        return out

    def isdf_coulomb_kernel_dev(self, d_zeta: int, ld_zeta: int, ngrid: int, n_intp: int, gidx, nfft, h, trunc,
                                ntrans=1.0):
        """Same V build on a device zeta (ld_zeta, n_intp), e.g. from isdf_zeta_dev."""
        grid = np.ascontiguousarray(np.asarray(gidx, np.int32).reshape(ngrid, 3))
        nf = np.ascontiguousarray(np.asarray(nfft, np.int32).reshape(3))
        out = np.zeros((n_intp, n_intp), order="F")
        _lib.check(self._lib.ngw_isdf_coulomb_kernel_dev(
            self._h, int(ngrid), int(n_intp), int(d_zeta), int(ld_zeta), grid.ctypes.data_as(C.POINTER(C.c_int)),
            nf.ctypes.data_as(C.POINTER(C.c_int)), float(h), float(trunc), float(ntrans), _p(out),
            n_intp), "ngw_isdf_coulomb_kernel_dev")
        return out

    def cvt_points(self, gidx, rho, init, h, tol=1e-3, max_iter=300):
        """ISDF points by rho-weighted CVT (shared/cvt.F90:8-195) on the GPU.
        gidx (ngrid, 3) lattice coordinates, rho (ngrid), init (n_intp, 3) bohr.
        Returns (0-based grid indices, iterations run)."""
        grid = np.ascontiguousarray(np.asarray(gidx, np.int32).reshape(-1, 3))
        r = np.ascontiguousarray(np.asarray(rho, np.float64))
        ini = np.ascontiguousarray(np.asarray(init, np.float64).reshape(-1, 3))
        out = np.zeros(ini.shape[0], np.int32)
        it = C.c_int(0)
        _lib.check(self._lib.ngw_cvt_points(
            self._h, grid.shape[0], grid.ctypes.data_as(C.POINTER(C.c_int)), _p(r), ini.shape[0], _p(ini),
            float(h), float(tol), int(max_iter), out.ctypes.data_as(C.POINTER(C.c_int)), C.byref(it)),
            "ngw_cvt_points")
        return out.astype(np.int64), it.value

    def isdf_zeta(self, psi, intp, nv, ridge=0.0):
        """zeta (ngrid, n_intp) = B A^-1 on the device from host orbitals psi
        (ngrid, nband) and 0-based point indices (isdf_parallel_sym_UltraLowMem.F90:448-642).
        Returns (zeta, used_lu)."""
        psi = _f64(psi)
        ngrid, nband = psi.shape
        ii = np.ascontiguousarray(np.asarray(intp, np.int32))
        zeta = np.zeros((ngrid, len(ii)), order="F")
        lu = C.c_int(0)
        _lib.check(self._lib.ngw_isdf_zeta(
            self._h, ngrid, nband, int(nv), _p(psi), ngrid, len(ii), ii.ctypes.data_as(C.POINTER(C.c_int)),
            float(ridge), _p(zeta), ngrid, C.byref(lu)), "ngw_isdf_zeta")
        return zeta, bool(lu.value)

    def isdf_zeta_dev(self, d_psi: int, ld_psi: int, ngrid: int, nband: int, nv: int, intp, ridge, d_zeta: int,
                      ld_zeta: int):
        """Device form of isdf_zeta; returns whether the LU fallback was used."""
        ii = np.ascontiguousarray(np.asarray(intp, np.int32))
        lu = C.c_int(0)
        _lib.check(self._lib.ngw_isdf_zeta_dev(
            self._h, int(ngrid), int(nband), int(nv), int(d_psi), int(ld_psi), len(ii),
            ii.ctypes.data_as(C.POINTER(C.c_int)), float(ridge), int(d_zeta), int(ld_zeta), C.byref(lu)),
            "ngw_isdf_zeta_dev")
        return bool(lu.value)

    # ------------------------------------------------------------ device pointers
    def matvec_dev(self, d_vec: int, d_hvec: int, ld: int, blksz: int, tamm_d=False):
        _lib.check(self._lib.ngw_matvec_isdf_block_dev(self._h, d_vec, d_hvec, int(ld), int(blksz),
                                                       int(tamm_d)), "ngw_matvec_isdf_block_dev")

    def gemm_lowp_dev(self, kind, m, n, k, d_a, lda, d_b, ldb, d_c, ldc):
        _lib.check(self._lib.ngw_gemm_lowp_dev(self._h, kind, m, n, k, d_a, lda, d_b, ldb, d_c, ldc),
                   "ngw_gemm_lowp_dev")

    def gemm_f64_dev(self, m, n, k, d_a, lda, d_b, ldb, d_c, ldc):
        _lib.check(self._lib.ngw_gemm_f64_dev(self._h, m, n, k, d_a, lda, d_b, ldb, d_c, ldc),
                   "ngw_gemm_f64_dev")

    def autotune(self, nvec, reps=2):
        """Time the K1/K3 tile configurations of the current precision on nvec
        vectors and keep the fastest; returns (k1_index, k3_index)."""
        a, b = C.c_int(), C.c_int()
        _lib.check(self._lib.ngw_autotune(self._h, int(nvec), int(reps), C.byref(a), C.byref(b)), "ngw_autotune")
        return a.value, b.value

    def set_tile_config(self, k1, k3):
        _lib.check(self._lib.ngw_set_tile_config(self._h, int(k1), int(k3)), "ngw_set_tile_config")

    # ------------------------------------------------------------ profiling
    KERNEL_CLASSES = ("K1", "K2", "K3", "lanczos_ab", "cfrac", "pvectors", "sigma", "convert")

    def profile_enable(self, on=True):
        _lib.check(self._lib.ngw_profile_enable(self._h, int(bool(on))), "ngw_profile_enable")

    def profile_reset(self):
        _lib.check(self._lib.ngw_profile_reset(self._h), "ngw_profile_reset")

    def profile_read(self):
        """{class: (total_ms, launches)} since the last reset (synchronizes)."""
        out = {}
        for k, name in enumerate(self.KERNEL_CLASSES):
            ms = C.c_double()
            n = C.c_long()
            _lib.check(self._lib.ngw_profile_read(self._h, k, C.byref(ms), C.byref(n)), "ngw_profile_read")
            out[name] = (ms.value, n.value)
        return out

    def synchronize(self):
        _lib.check(self._lib.ngw_synchronize(self._h), "ngw_synchronize")
