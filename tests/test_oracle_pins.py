"""Pins of the CPU oracle (no GPU): the reference's own known-answer data,
closed forms and invariants that a dropped term, wrong sign or transposed
operand would break."""
import os

import numpy as np
import pytest

from conftest import GOLDEN, rel_err

RYD = 13.60569193


@pytest.fixture(scope="module")
def O():
    from oracle import oracle

    oracle.lib()
    return oracle


@pytest.fixture(scope="module")
def tiny():
    """nv = 2, nc = 5: ncv = 10, small enough for Lanczos to span the space."""
    from nanogw_amd import synth

    return synth.make_molecule(3.5, 0.55, 2, 5, 16, seed=99)


def _dense_h(mol, tda=False):
    nv, nc = mol.nv, mol.nc
    psi = mol.psi_intp
    C = (psi[:, nv:][:, :, None] * psi[:, :nv][:, None, :]).reshape(psi.shape[0], -1)
    s = np.sqrt((mol.e1[nv:][:, None] - mol.e1[:nv][None, :]).reshape(-1))
    K = C.T @ mol.mmtrx @ C
    K = 0.5 * (K + K.T)
    if tda:
        return np.diag(s ** 2) + 2.0 * K, s, C
    return np.diag(s ** 4) + 4.0 * (s[:, None] * K * s[None, :]), s, C


# --------------------------------------------------------------- known answer
def test_kat_h_aux_v0(O):
    """shared/test_lanczos/H_aux_v0.txt reproduced at niter = 4 (print precision 1e-8).

    The file's header says niter = 5, which is off by one; niter = 5 misses by ~0.2.
    """
    d = np.load(os.path.join(GOLDEN, "H_aux_v0.npz"))
    h, v0, tab, eta = d["h_aux"], d["v0"], d["table"], float(d["eta"])
    zz = np.concatenate([tab[:, 0] + 1j * eta, [0.0]])
    spec, a, b = O.lanczos_dense(h, v0, zz, 4)
    assert np.abs(spec[:-1].real - tab[:, 2]).max() < 2e-7
    assert np.abs(spec[:-1].imag - tab[:, 3]).max() < 2e-7
    # column 2 is -<v|H^-1|v>, i.e. S(0)
    assert abs(spec[-1].real - tab[0, 1]) < 2e-7
    spec5, _, _ = O.lanczos_dense(h, v0, zz, 5)
    assert np.abs(spec5[:-1].real - tab[:, 2]).max() > 1e-2


def test_lanczos_moments_and_exact_limit(O):
    """<v|H^k|v> = (T^k)_11 for k < 2 niter; niter = n gives the exact resolvent."""
    rng = np.random.default_rng(1)
    n = 12
    a = rng.normal(size=(n, n))
    h = a + a.T
    v = rng.normal(size=n)
    vn = v / np.linalg.norm(v)
    zz = np.array([0.3 + 0.2j, -1.0 + 0.05j, 2.0 + 1.0j, 0.0 + 0.3j])
    niter = 5
    _, al, be = O.lanczos_dense(h, v, zz, niter)
    T = np.diag(al) + np.diag(be[1:], 1) + np.diag(be[1:], -1)
    for k in range(2 * niter):
        ref = vn @ np.linalg.matrix_power(h, k) @ vn
        got = np.linalg.matrix_power(T, k)[0, 0]
        assert abs(got - ref) <= 1e-9 * max(1.0, abs(ref)), k
    spec, _, _ = O.lanczos_dense(h, v, zz, n)
    w, Z = np.linalg.eigh(h)
    wt = (Z.T @ vn) ** 2
    exact = (wt[None, :] / (zz[:, None] - w[None, :])).sum(1)
    assert rel_err(spec, exact) < 1e-9


def _separated_extremes(n, seed):
    """Symmetric matrix whose spectrum has three well-separated top eigenvalues
    over a dense bulk: the case where the plain recurrence loses orthogonality
    once the top Ritz values converge and then repeats them (ghosts)."""
    rng = np.random.default_rng(seed)
    lam = np.concatenate([np.linspace(0.0, 1.0, n - 3), [3.0, 5.0, 8.0]])
    Z, _ = np.linalg.qr(rng.normal(size=(n, n)))
    h = (Z * lam) @ Z.T
    return 0.5 * (h + h.T), lam, rng.normal(size=n)


def test_lanczos_reorth_orthogonal_ritz_exact_no_ghosts(O):
    """Full reorthogonalisation (no NanoGW counterpart, textbook scheme):
    Q^T Q = I and Q^T H Q = T (a on the diagonal, b(k+1) coupling q_k and
    q_k+1) at niter = n; the Ritz values are the eigenvalues, each once; the
    continued fraction is the exact resolvent.  The same run without
    reorthogonalisation loses orthogonality (> 0.1) and repeats the top
    eigenvalue, so the case tests what it claims."""
    n = 60
    h, lam, v = _separated_extremes(n, 5)
    vn = v / np.linalg.norm(v)
    zz = np.array([0.5 + 0.05j, 3.0 + 0.01j, 7.0 + 0.2j, -0.5 + 0.1j])
    spec, a, b, Q = O.lanczos_reorth(lambda x: h @ x, v, zz, n, reorth=True)
    T = np.diag(a) + np.diag(b[1:], 1) + np.diag(b[1:], -1)
    assert np.abs(Q.T @ Q - np.eye(n)).max() < 1e-13
    assert np.abs(Q.T @ h @ Q - T).max() < 1e-12
    assert np.abs(np.linalg.eigvalsh(T) - np.sort(lam)).max() < 1e-12
    w, Z = np.linalg.eigh(h)
    exact = (((Z.T @ vn) ** 2)[None, :] / (zz[:, None] - w[None, :])).sum(1)
    assert rel_err(spec, exact) < 1e-10
    _, a0, b0, Q0 = O.lanczos_reorth(lambda x: h @ x, v, zz, n, reorth=False)
    assert np.abs(Q0.T @ Q0 - np.eye(n)).max() > 0.1
    T0 = np.diag(a0) + np.diag(b0[1:], 1) + np.diag(b0[1:], -1)
    assert np.sum(np.abs(np.linalg.eigvalsh(T0) - 8.0) < 1e-8) > 1


def test_lanczos_reorth_recurrence_is_the_reference_one(O):
    """Before orthogonality is lost the reorthogonalised run equals the plain
    one, and the plain numpy restatement equals the C restatement of
    lanczos_spectra_BLOCK.F90 (same recurrence order and b(1) = 1)."""
    h, _, v = _separated_extremes(40, 8)
    zz = np.array([0.5 + 0.05j, 3.0 + 0.01j])
    s1, a1, b1, _ = O.lanczos_reorth(lambda x: h @ x, v, zz, 6, reorth=True)
    s0, a0, b0, _ = O.lanczos_reorth(lambda x: h @ x, v, zz, 6, reorth=False)
    sc, ac, bc = O.lanczos_dense(h, v, zz, 6)
    assert np.abs(a1 - a0).max() < 1e-12 and np.abs(b1 - b0).max() < 1e-12
    assert np.abs(a0 - ac).max() < 1e-12 and np.abs(b0 - bc).max() < 1e-12
    assert rel_err(s0, sc) < 1e-12 and rel_err(s1, sc) < 1e-11


# --------------------------------------------------------------- operator
@pytest.mark.parametrize("tda", [False, True])
def test_operator_dense_closed_form(O, tiny, tda):
    H, s, _ = _dense_h(tiny, tda)
    x = np.random.default_rng(2).normal(size=(tiny.nv * tiny.nc, 3))
    y = O.matvec_block(tiny.psi_intp, tiny.mmtrx, s, x, tiny.nv, tiny.nc, tda)
    assert rel_err(y, H @ x) < 1e-9
    # symmetry x^T H y = y^T H x
    u, v = x[:, 0], x[:, 1]
    hu = O.matvec_block(tiny.psi_intp, tiny.mmtrx, s, u, tiny.nv, tiny.nc, tda)[:, 0]
    hv = O.matvec_block(tiny.psi_intp, tiny.mmtrx, s, v, tiny.nv, tiny.nc, tda)[:, 0]
    assert abs(v @ hu - u @ hv) < 1e-9 * abs(v @ hu)


def test_horner_is_p_of_h(O, tiny):
    H, s, _ = _dense_h(tiny)
    poly = np.array([0.01, -0.2, 1.3, 0.4])  # highest order first
    w, Z = np.linalg.eigh(H)
    pH = Z @ np.diag(np.polyval(poly, w)) @ Z.T
    x = np.random.default_rng(3).normal(size=(tiny.nv * tiny.nc, 2))
    y = O.polynomial_matvec_block(tiny.psi_intp, tiny.mmtrx, s, x, tiny.nv, tiny.nc, poly)
    assert rel_err(y, pH @ x) < 1e-9


def _sigma_sum_over_states(mol, qp, nen, deltae, ecuts, poly, tda=False, bands=None):
    """Sigma_c from dense eigh: S(z) = sum_s |<P|Z_s>|^2 / (z - p(Omega_s^2));
    bands = (lo, hi) keeps Green's-function bands lo <= n < hi."""
    H, s, C = _dense_h(mol, tda)
    w, Z = np.linalg.eigh(H)
    lam = np.polyval(poly, w)
    nv, nb = mol.nv, mol.nb
    out = np.zeros((nen, len(qp)), complex)
    for iq, m in enumerate(qp):
        z = deltae / (nen - 1) * np.arange(nen) + mol.e1[m] - deltae / 2
        for n in range(nb) if bands is None else range(*bands):
            c = mol.psi_intp[:, m] * mol.psi_intp[:, n]
            P = C.T @ (mol.mmtrx @ c)
            if not tda:
                P = P * s
            p2 = P @ P
            wt = (Z.T @ P) ** 2 / p2
            eta = 1.0 if n < nv else -1.0
            zeta = eta * (mol.e1[n] - z) + 1j * ecuts
            S = (wt[None, :] / (zeta[:, None] - lam[None, :])).sum(1)
            S0 = (wt / (0.0 - lam)).sum()
            if tda:
                out[:, iq] += -S * eta * float(np.float32(np.sqrt(2.0))) * p2
            else:
                # occupied: 2|P|^2/(-zeta) (S - S0); empty: 2|P|^2/zeta (S - S0)
                out[:, iq] += 2.0 * p2 / (-eta * zeta) * (S - S0)
    return out


@pytest.mark.parametrize("tda", [False, True])
def test_sigma_equals_sum_over_states(O, tiny, tda):
    """At niter = ncv the Lanczos Sigma_c equals the sum over Casida states."""
    poly = np.array([1.0, 0.0]) if tda else np.array([-0.004, 0.09, 0.6, 0.35])
    qp = [tiny.nv - 1, tiny.nv]
    ncv = tiny.nv * tiny.nc
    sig = O.sigma_lanczos(tiny.psi_intp, tiny.mmtrx, tiny.e1, tiny.nv, tiny.nc, qp, 31, 3.0, 0.05,
                          ncv, poly, 3, tda=tda)
    ref = _sigma_sum_over_states(tiny, qp, 31, 3.0, 0.05, poly, tda)
    assert rel_err(sig, ref) < 1e-7


@pytest.mark.parametrize("bands", [(0, 2), (1, 5), (2, 7)])
def test_sigma_band_range_equals_sum_over_states(O, tiny, bands):
    """The band-range restriction (a multi-device driver's unit of work):
    Sigma_c over Green's-function bands lo <= n < hi equals the sum over
    Casida states restricted to those bands, and ranges add up to the whole."""
    poly = np.array([-0.004, 0.09, 0.6, 0.35])
    qp = [tiny.nv - 1, tiny.nv]
    ncv = tiny.nv * tiny.nc
    args = (tiny.psi_intp, tiny.mmtrx, tiny.e1, tiny.nv, tiny.nc, qp, 31, 3.0, 0.05, ncv, poly, 3)
    sig = O.sigma_lanczos(*args, bands=bands)
    ref = _sigma_sum_over_states(tiny, qp, 31, 3.0, 0.05, poly, bands=bands)
    assert rel_err(sig, ref) < 1e-7
    rest = O.sigma_lanczos(*args, bands=(0, bands[0])) + O.sigma_lanczos(*args, bands=(bands[1], tiny.nb))
    assert rel_err(sig + rest, O.sigma_lanczos(*args)) < 1e-12


def test_pvectors_match_definition(O, tiny):
    _, s, C = _dense_h(tiny)
    m, n0, b = 1, 0, 4
    p, pn2 = O.pvectors(tiny.psi_intp, tiny.mmtrx, s, m, n0, b, tiny.nv, tiny.nc)
    for k in range(b):
        c = tiny.psi_intp[:, m] * tiny.psi_intp[:, n0 + k]
        ref = (C.T @ (tiny.mmtrx @ c)) * s
        assert rel_err(p[:, k], ref) < 1e-11
        assert abs(pn2[k] - ref @ ref) < 1e-10 * (ref @ ref)


# --------------------------------------------------------------- QP Newton
def test_newton_linear_table(O):
    """f(E) linear: the spline is exact, Newton lands on the root, zfac = -1/slope."""
    nen = 50
    x = np.linspace(-5.0, 5.0, nen)
    f = (0.7 - 1.6 * x) + 0.0j
    root, fx, ierr = O.newtonr(x, f, 0.0)
    assert ierr == 0 and abs(root - 0.7 / 1.6) < 1e-12
    fx2, df = O.spline_evaluate(x, f, 1.234)
    assert abs(fx2.real - (0.7 - 1.6 * 1.234)) < 1e-12 and abs(df + 1.6) < 1e-10


def test_spline_cubic_interior(O):
    """Cubic-spline interpolation of a smooth table converges (interior error O(h^4))."""
    for nen, tol in [(41, 2e-5), (161, 1e-7)]:
        x = np.linspace(-1.0, 1.0, nen)
        f = np.sin(2 * x) + 1j * np.cos(x)
        fx, _ = O.spline_evaluate(x, f, 0.123)
        assert abs(fx - (np.sin(0.246) + 1j * np.cos(0.123))) < tol


@pytest.mark.parametrize("n", [5, 21, 201])
def test_spline_and_newton_equal_the_reference_build(O, n):
    """The oracle's spline_evaluate / newtonr against the reference's own
    SIGMA/spline_evaluate.f90, spline.f90, splint.f90 and newtonr.f90,
    compiled from their sources into oracle/_ref (make -C oracle ref-spline):
    interpolated value and slope on uneven complex tables, and the Newton
    root, residual and error flag, including a table without a root (the
    reference returns ierr = MAX_IT + 1 = 21).  Skipped where that library
    was not built (no reference sources)."""
    if O.ref_lib() is None:
        pytest.skip("oracle/_ref/libngw_ref_spline.so not built (reference sources absent)")
    rng = np.random.default_rng(60 + n)
    x = np.sort(np.concatenate([[-2.0, 3.0], rng.uniform(-2.0, 3.0, n - 2)]))
    f = np.sin(1.3 * x) + 0.2 * x - 0.1 + 1j * (0.3 * np.cos(x) + 0.05 * rng.normal(size=n))
    eps = np.finfo(float).eps
    for xx in np.concatenate([rng.uniform(x[1], x[-2], 6), [x[2], x[-3]]]):
        fa, da = O.spline_evaluate(x, f, xx)
        fb, db = O.ref_spline_evaluate(x, f, xx)
        assert abs(fa - fb) <= 8 * eps * max(1.0, abs(fb)), (xx, fa, fb)
        assert abs(da - db) <= 1e3 * eps * max(1.0, abs(db)), (xx, da, db)
    for x0, acc in ((0.1, 1e-6), (-1.5, 1e-3)):
        ra, rb = O.newtonr(x, f, x0, acc), O.ref_newtonr(x, f, x0, acc)
        assert ra[2] == rb[2]
        assert abs(ra[0] - rb[0]) <= 1e-12 * max(1.0, abs(rb[0]))
        assert abs(ra[1] - rb[1]) <= 1e-12 * max(1.0, abs(rb[1]))
    g = 2.0 + 0.1 * x ** 2 + 0.0j  # no real root
    ra, rb = O.newtonr(x, g, 0.3, 1e-8), O.ref_newtonr(x, g, 0.3, 1e-8)
    assert rb[2] == 21 and ra[2] == rb[2]


def test_spline_exact_on_cubic_table_interior(O):
    """spline.f90 takes its boundary slopes from the end pairs, so a cubic
    table is not reproduced near the ends; the boundary error decays by about
    2 - sqrt(3) per interval, and in the interior of a 200-point table the
    spline equals the cubic to rounding (a wrong second-derivative system
    fails this).  The returned slope is the central difference of Re f over
    +-0.01 (printsigma's derivative step), which for the cubic term 0.3 E^3
    exceeds the derivative by exactly 0.3 * 0.01^2."""
    x = np.linspace(-3.0, 3.0, 200)

    def cub(t):
        return (0.3 * t ** 3 - 1.1 * t ** 2 + 0.5 * t - 2.0) + 1j * (-0.2 * t ** 3 + 0.7 * t)

    f = cub(x)
    for xx in (-1.0, 0.0, 0.37, 1.234, 2.0):
        fx, df = O.spline_evaluate(x, f, xx)
        assert abs(fx - cub(xx)) < 1e-12
        assert abs(df - (0.9 * xx ** 2 - 2.2 * xx + 0.5 + 0.3e-4)) < 1e-9


def test_qp_solve_matches_direct_root(O):
    """E_qp solves E = E_dft + Sigma_x + Re Sigma_c(E) - V_xc for a smooth Sigma_c."""
    nen = 200
    en0 = -6.0
    x = en0 - 13.6 + 27.2 / (nen - 1) * np.arange(nen)
    sig = (-1.0 + 0.1 * (x - en0)) + 0.05j * np.cos(x)
    r = O.qp_solve(sig, 27.2, en0, en0, -3.0, -2.0)
    # exact root of -6 - 3 + (-1 + 0.1 (E + 6)) + 2 - E = 0
    e_exact = (-6.0 - 3.0 - 1.0 + 0.6 + 2.0) / 0.9
    assert r["ierr"] == 0
    assert abs(r["eqp"] - e_exact) < 0.011  # Newton stops once |dx| < EN_ACC = 0.01 eV
    assert abs(r["zfac"] - 1.0 / 0.9) < 1e-9


# --------------------------------------------------------------- polyfit
@pytest.mark.parametrize("npoly,cheb", [(2, False), (3, True), (4, False)])
def test_polyfit_matches_lstsq(O, npoly, cheb):
    xmin, xmax, nx = 0.5, 4.0, 20
    beta, fe = O.polyfit_sqrt(xmin, xmax, npoly, nx, cheb)
    if cheb:
        xx = 0.5 * (xmin + xmax) - 0.5 * (xmax - xmin) * np.cos(np.pi / (nx - 1) * np.arange(nx))
    else:
        xx = np.linspace(xmin, xmax, nx)
    V = np.vander(xx, npoly + 1, increasing=True)
    ref, *_ = np.linalg.lstsq(V, np.sqrt(xx), rcond=None)
    assert rel_err(beta, ref) < 1e-8
    assert fe < 0.02


def test_polyfit_rejects_small_xmin(O):
    with pytest.raises(ValueError):
        O.polyfit_sqrt(1e-5, 1.0, 4)


# --------------------------------------------------------------- Poisson
def test_poisson_gaussian_charge(O):
    """V of a normalised Gaussian charge: V(r) = e^2 erf(sqrt(a) r) / r, e^2 = 2 (Ry)."""
    from math import erf

    from nanogw_amd import synth

    h, R, alpha = 0.25, 4.5, 2.0
    g = synth.sphere_grid(R, h)
    rmax, nfft, trunc = synth.fft_box(g, h)
    r = np.sqrt(((g * h) ** 2).sum(1))
    rho = (alpha / np.pi) ** 1.5 * np.exp(-alpha * r ** 2) * h ** 3  # charge per grid point
    V = O.poisson(rho / h ** 3, g, nfft, h, trunc)
    for rr in [0.5, 1.0, 2.0, 3.0]:
        i = np.argmin(np.abs(r - rr))
        ref = 2.0 * erf(np.sqrt(alpha) * r[i]) / r[i]
        assert abs(V[i] - ref) < 2e-3 * ref, (rr, V[i], ref)


def test_isdf_kernel_symmetric_psd(O, tiny):
    M = O.isdf_kernel(tiny.zeta, tiny.gidx, tiny.nfft, tiny.h, tiny.trunc)
    assert rel_err(M, M.T) < 1e-10
    assert np.linalg.eigvalsh(0.5 * (M + M.T)).min() > -1e-8 * np.abs(M).max()


@pytest.mark.parametrize("h", [0.25, 0.3])
@pytest.mark.parametrize("ntrans", [1, 2])
def test_isdf_kernel_scale_gaussian_pair_energies(O, h, ntrans):
    """Absolute scale of M = (ntrans/h^3) zeta^T V zeta
    (isdf_parallel_sym_UltraLowMem.F90:312, 865-867) with Gaussian columns
    zeta_mu(r) = exp(-a |r - R_mu|^2).  Sums over grid points are integrals /
    h^3, so M_mu,nu = (ntrans / h^6) E_mu,nu with the Coulomb energy of two
    Gaussian charges Q = (pi/a)^(3/2) (e^2 = 2, Ry):
    E = 2 Q^2 erf(sqrt(a/2) d) / d, and 2 Q^2 2 sqrt(a/2) / sqrt(pi) at d = 0.
    A dropped or doubled h^3, a missing ntrans or a wrong e^2 fails by a
    factor of at least 2; at two spacings the h dependence is fixed too."""
    from math import erf, pi, sqrt

    from nanogw_amd import synth

    a = 2.0
    g = synth.sphere_grid(4.5, h)
    _, nfft, trunc = synth.fft_box(g, h)
    r = g * h
    cen = np.array([[0.0, 0.0, 0.0], [1.2, 0.0, 0.0], [0.0, -0.9, 0.7], [-1.0, 0.5, -0.6]])
    zeta = np.asfortranarray(np.stack([np.exp(-a * ((r - c) ** 2).sum(1)) for c in cen], axis=1))
    M = O.isdf_kernel(zeta, g, nfft, h, trunc, ntrans=ntrans)
    q2 = (pi / a) ** 3
    ref = np.empty((4, 4))
    for i in range(4):
        for j in range(4):
            d = float(np.linalg.norm(cen[i] - cen[j]))
            ref[i, j] = 2.0 * q2 * (erf(sqrt(a / 2) * d) / d if d > 0 else 2.0 * sqrt(a / 2) / sqrt(pi))
    ref *= ntrans / h ** 6
    assert np.abs(M / ref - 1.0).max() < 1e-8


def _gaussian_orbitals(g, h, cen, alpha):
    r = g * h
    phi = np.stack([np.exp(-a * ((r - c) ** 2).sum(1)) for c, a in zip(cen, alpha)], axis=1)
    return np.asfortranarray(phi / np.sqrt((phi ** 2).sum(axis=0)))  # sum_r psi^2 = 1


@pytest.mark.parametrize("h", [0.25, 0.3])
def test_sigma_x_direct_gaussian_closed_form(O, h):
    """sigma_x_direct (SIGMA/fock_exchange.F90z) on normalised Gaussian
    orbitals phi_i = (2 a_i / pi)^(3/4) exp(-a_i |r - A_i|^2): the pair
    density phi_m phi_n is the Gaussian K exp(-p |r - P|^2), p = a_m + a_n,
    K = N_m N_n exp(-a_m a_n |A_m - A_n|^2 / p), whose Coulomb self energy is
    (mn|nm) = 2 K^2 (pi/p)^3 sqrt(2p/pi) (Ry, e^2 = 2).  Sigma_x(m) =
    -sum_n occ_n (mn|nm).  A dropped or doubled h^3, a wrong sign, an
    unoccupied state in the sum (occ below tol_occ) or pair densities of the
    wrong states fail it."""
    from math import pi, sqrt

    from nanogw_amd import synth

    g = synth.sphere_grid(4.5, h)
    _, nfft, trunc = synth.fft_box(g, h)
    cen = np.array([[0.0, 0.0, 0.0], [0.9, 0.3, 0.0], [-0.4, -0.8, 0.5], [0.3, 0.6, -0.9], [0.0, 0.0, 1.0]])
    alpha = np.array([1.6, 1.2, 2.0, 1.4, 1.1])
    psi = _gaussian_orbitals(g, h, cen, alpha)
    occ = np.array([1.0, 1.0, 0.6, 0.005, 0.0])  # the last two are not occupied (tol_occ = 0.01)
    qp = [0, 3, 4]
    got = O.sigma_x_direct(psi, occ, g, nfft, h, trunc, qp)
    ref = []
    for m in qp:
        s = 0.0
        for n in range(3):
            p = alpha[m] + alpha[n]
            nm, nn = (2 * alpha[m] / pi) ** 0.75, (2 * alpha[n] / pi) ** 0.75
            k = nm * nn * np.exp(-alpha[m] * alpha[n] * ((cen[m] - cen[n]) ** 2).sum() / p)
            s += occ[n] * 2.0 * k * k * (pi / p) ** 3 * sqrt(2.0 * p / pi)
        ref.append(-s)
    ref = np.array(ref)
    print("sigma_x direct vs closed form", got, ref, np.abs(got / ref - 1).max())
    assert np.abs(got / ref - 1.0).max() < 1e-7
    # ntrans and nk enter as ntrans / nk
    assert rel_err(O.sigma_x_direct(psi, occ, g, nfft, h, trunc, qp, ntrans=2, nk=4), 0.5 * got) < 1e-14


def test_isdf_exchange_equals_direct_in_exact_limit(O):
    """With every grid point an interpolation point (zeta = identity, so the
    ISDF fit is exact), the ISDF exchange -sum_v C_v^T M C_v, C_v = psi_m psi_v
    at the points and M = (1/h^3) zeta^T V zeta (synth.exchange_diag, the
    Sigma_x input of both paths; the device Sigma_x is checked against it),
    equals the reference's direct exchange (sigma_x_direct) to rounding."""
    from nanogw_amd import synth

    h = 0.4
    g = synth.sphere_grid(2.4, h)
    _, nfft, trunc = synth.fft_box(g, h)
    rng = np.random.default_rng(17)
    nv, nb = 3, 6
    psi = rng.normal(size=(g.shape[0], nb))
    psi = np.asfortranarray(psi / np.sqrt((psi ** 2).sum(axis=0)))
    mm = O.isdf_kernel(np.eye(g.shape[0]), g, nfft, h, trunc)
    qp = [0, 2, 3, 5]
    isdf = synth.exchange_diag(psi, mm, nv, qp)
    direct = O.sigma_x_direct(psi, [1.0] * nv + [0.0] * (nb - nv), g, nfft, h, trunc, qp)
    print("exact-limit ISDF exchange vs direct", isdf, direct)
    assert rel_err(isdf, direct) < 1e-11


def test_cvt_lattice_lloyd_step_by_brute_force(O):
    """cvt_lattice (the CVT restatement the device CVT is checked against),
    pinned by a brute-force Lloyd step: from the restatement's points after
    k - 1 iterations, nearest-centroid assignment over the full distance
    matrix, rho-weighted means, and for each centroid in order the nearest
    grid point not yet taken (by a scan of every grid point) give the
    restatement's points after k iterations."""
    from nanogw_amd import synth

    radius, h, nv, n_intp = 3.5, 0.5, 3, 20
    gidx = synth.sphere_grid(radius, h)
    rng = np.random.default_rng(9)
    psi = synth.orbitals(gidx, h, nv + 4, rng, radius)
    rho = 2.0 * np.sum(psi[:, :nv] ** 2, axis=1)
    pts = gidx.astype(np.float64)
    lo, hi = pts.min(axis=0) * h, pts.max(axis=0) * h
    init = lo[None, :] + rng.uniform(size=(n_intp, 3)) * 0.9 * (hi - lo)[None, :]
    for k in (1, 2, 3):
        prev, _ = O.cvt_lattice(gidx, rho, init, h, tol=-1.0, max_iter=k - 1) if k > 1 else (None, 0)
        cen = init / h if prev is None else pts[prev]
        got, _ = O.cvt_lattice(gidx, rho, init, h, tol=-1.0, max_iter=k)
        d = ((pts[:, None, :] - cen[None, :, :]) ** 2).sum(axis=2)
        owner = np.argmin(d, axis=1)
        taken = np.zeros(len(pts), bool)
        for c in range(n_intp):
            mine = owner == c
            w = rho[mine].sum()
            cand = (rho[mine, None] * pts[mine]).sum(axis=0) / w if w >= 1e-10 else cen[c]
            dist = ((pts - cand) ** 2).sum(axis=1)
            dist[taken] = np.inf
            best = dist.min()
            assert abs(dist[got[c]] - best) <= 1e-9 * max(1.0, best), (k, c)
            taken[got[c]] = True
        assert len(set(got.tolist())) == n_intp


# ------------------------------------------- single-vector explicit-C routines
def test_explicit_matvec_equals_block_and_permutes(O, tiny):
    """matvec_isdf on a caller-built Cmtrx equals the BLOCK operator on the
    orbital pair products, and a permuted pair list permutes Hvec."""
    H, s, _ = _dense_h(tiny)
    nv, nc = tiny.nv, tiny.nc
    cm = O.pair_cmtrx(tiny.psi_intp, nv, nc)
    x = np.random.default_rng(21).normal(size=nv * nc)
    y = O.matvec_explicit(cm, tiny.mmtrx, s, x)
    assert rel_err(y, H @ x) < 1e-9
    perm = np.random.default_rng(22).permutation(nv * nc)
    yp = O.matvec_explicit(np.asfortranarray(cm[:, perm]), tiny.mmtrx, s[perm], x[perm])
    assert rel_err(yp, y[perm]) < 1e-12
    poly = np.array([0.01, -0.2, 1.3, 0.4])
    w, Z = np.linalg.eigh(H)
    yh = O.polynomial_matvec_explicit(cm, tiny.mmtrx, s, x, poly)
    assert rel_err(yh, Z @ (np.polyval(poly, w) * (Z.T @ x))) < 1e-9


def test_explicit_lanczos_equals_block_and_breakdown(O, tiny):
    """Without breakdown the single-vector Lanczos equals the BLOCK one; from
    an eigenvector of p(H) it stops at k = 2 and S(z) = 1/(z - p(lambda))."""
    H, s, _ = _dense_h(tiny)
    nv, nc = tiny.nv, tiny.nc
    cm = O.pair_cmtrx(tiny.psi_intp, nv, nc)
    poly = np.array([0.02, 0.9, 0.1])
    v = np.random.default_rng(23).normal(size=nv * nc)
    zz = np.linspace(-1.0, 4.0, 9) + 0.05j
    spec, a, b, nk = O.lanczos_spectra_explicit(cm, tiny.mmtrx, s, v, np.linalg.norm(v), zz, 6, poly)
    sref, aref, bref = O.lanczos_spectra_block(tiny.psi_intp, tiny.mmtrx, s, v, [np.linalg.norm(v)], zz, 6, poly,
                                               nv, nc)
    assert nk == 6
    assert rel_err(a, aref[:, 0]) < 1e-12 and rel_err(b, bref[:, 0]) < 1e-12
    assert rel_err(spec, sref[:, 0]) < 1e-12
    w, Z = np.linalg.eigh(H)
    e = Z[:, 3]
    lam = np.polyval(poly, w[3])
    spec, a, b, nk = O.lanczos_spectra_explicit(cm, tiny.mmtrx, s, 2.0 * e, 2.0, zz, 6, poly)
    assert nk == 1 and abs(b[1]) < 1e-8 and np.all(a[1:] == 0.0)
    assert abs(a[0] - lam) < 1e-10 * abs(lam)
    assert rel_err(spec, 1.0 / (zz - lam)) < 1e-10


def test_operator_dominates_r_squared(O, tiny):
    """With M positive semi-definite, H - R^2 = 4 S C^T M C S is PSD, so
    x^T H x >= x^T R^2 x for every x (a wrong sign or a dropped S breaks it)."""
    H, s, _ = _dense_h(tiny)
    w = np.linalg.eigvalsh(0.5 * (tiny.mmtrx + tiny.mmtrx.T))
    assert w.min() > -1e-10 * abs(w).max()
    x = np.random.default_rng(41).normal(size=(tiny.nv * tiny.nc, 8))
    y = O.matvec_block(tiny.psi_intp, tiny.mmtrx, s, x, tiny.nv, tiny.nc)
    quad = np.einsum("ij,ij->j", x, y)
    base = np.einsum("ij,ij->j", x, (s ** 4)[:, None] * x)
    assert np.all(quad >= base * (1 - 1e-12))
    assert np.linalg.eigvalsh(H - np.diag(s ** 4)).min() > -1e-9 * np.abs(H).max()


def test_isdf_fit_residual_decreases_with_points():
    """The interpolative fit of the pair densities psi_v psi_c improves as
    interpolation points are added (points from the same CVT growth), and is
    exact at the points themselves."""
    from nanogw_amd import synth

    gidx = synth.sphere_grid(4.0, 0.5)
    rng = np.random.default_rng(7)
    nv, nc = 3, 6
    psi = synth.orbitals(gidx, 0.5, nv + nc, rng, 4.0)
    pairs = (psi[:, nv:][:, :, None] * psi[:, :nv][:, None, :]).reshape(len(gidx), -1)
    rho = 2.0 * np.sum(psi[:, :nv] ** 2, axis=1)
    res = []
    for n_intp in (6, 12, 18):
        intp = synth.cvt_points(gidx, 0.5, rho, n_intp, np.random.default_rng(11))
        zeta = synth.isdf_zeta(psi, intp, nv)
        fit = zeta @ pairs[intp, :]
        assert np.abs(fit[intp] - pairs[intp]).max() < 1e-8 * np.abs(pairs).max()
        res.append(np.linalg.norm(fit - pairs) / np.linalg.norm(pairs))
    print("ISDF relative residuals", res)
    assert res[0] > res[1] > res[2]
