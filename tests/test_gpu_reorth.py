"""Full reorthogonalisation of the Lanczos recurrence on the device (an FP64
option with no NanoGW counterpart; off in the parity path) against the
oracle's restatement (oracle.lanczos_reorth, pinned by tests in
test_oracle_pins.py).

The operator is p(H) of a well-conditioned molecule with ncv = 180; at
niter = 100 > ncv / 2 the plain recurrence has lost orthogonality (its
tridiagonal repeats the top eigenvalue), the reorthogonalised one has not.
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def wc():
    from nanogw_amd import synth

    return synth.make_molecule(5.0, 0.5, 6, 30, 60, seed=31337, ridge=1e-6)


@pytest.fixture(scope="module")
def dense_ph(wc):
    from nanogw_amd.lanczos import lanczos_polynomial
    from oracle import oracle as O

    poly = lanczos_polynomial(wc.e1, wc.nv, 12)
    s = O.sqrt_r(wc.e1, wc.nv, wc.nc)
    n = wc.nv * wc.nc
    hp = O.polynomial_matvec_block(wc.psi_intp, wc.mmtrx, s, np.eye(n), wc.nv, wc.nc, poly)
    return 0.5 * (hp + hp.T), poly


def test_reorth_lanczos_vs_oracle(ngw_ctx, wc, dense_ph):
    from oracle import oracle as O

    hp, poly = dense_ph
    n = hp.shape[0]
    niter, blk = 100, 3
    v0 = np.random.default_rng(2).normal(size=(n, blk))
    vn = np.linalg.norm(v0, axis=0)
    ev = np.linalg.eigvalsh(hp)
    zz1 = np.array([ev[0] + 0.02j, np.median(ev) + 0.02j, ev[-1] * (1 + 0.01j)])
    zz = np.repeat(zz1[:, None], blk, axis=1)
    ngw_ctx.set_isdf(wc.psi_intp, wc.mmtrx, wc.e1, wc.nv)
    ngw_ctx.set_precision("f64")
    ngw_ctx.set_reorthogonalization(True)
    try:
        spec, a, b = ngw_ctx.lanczos_spectra_isdf_BLOCK(v0, vn, zz, niter, poly)
    finally:
        ngw_ctx.set_reorthogonalization(False)
    _, a_plain, b_plain = ngw_ctx.lanczos_spectra_isdf_BLOCK(v0, vn, zz, niter, poly)
    for j in range(blk):
        s_ref, a_ref, b_ref, _ = O.lanczos_reorth(lambda x: hp @ x, v0[:, j], zz1, niter)
        print(f"column {j}: a {rel_err(a[:, j], a_ref):.2e} b {rel_err(b[:, j], b_ref):.2e} "
              f"spectra {rel_err(spec[:, j], s_ref):.2e}")
        assert rel_err(a[:, j], a_ref) < 1e-9 and rel_err(b[:, j], b_ref) < 1e-9
        assert rel_err(spec[:, j], s_ref) < 1e-8
        # Ritz values: each eigenvalue at most once (no ghosts), the top one found
        ritz = np.linalg.eigvalsh(np.diag(a[:, j]) + np.diag(b[1:, j], 1) + np.diag(b[1:, j], -1))
        assert np.sum(np.abs(ritz - ev[-1]) < 1e-8 * ev[-1]) == 1
        ritz0 = np.linalg.eigvalsh(np.diag(a_plain[:, j]) + np.diag(b_plain[1:, j], 1)
                                   + np.diag(b_plain[1:, j], -1))
        assert np.sum(np.abs(ritz0 - ev[-1]) < 1e-8 * ev[-1]) > 1  # the plain recurrence has ghosts here


def test_reorth_sigma_equals_sum_over_states(ngw_ctx):
    """With reorthogonalisation, Sigma_c at niter = ncv is the sum over states
    (dense eigh) to round-off, on a molecule where the plain recurrence has
    lost orthogonality by then."""
    from nanogw_amd import synth
    from test_oracle_pins import _sigma_sum_over_states

    mol = synth.make_molecule(4.0, 0.5, 3, 12, 30, seed=4711, ridge=1e-6)
    poly = np.array([-0.004, 0.09, 0.6, 0.35])
    qp = [mol.nv - 1, mol.nv]
    ncv = mol.nv * mol.nc
    ngw_ctx.set_isdf(mol.psi_intp, mol.mmtrx, mol.e1, mol.nv)
    ngw_ctx.set_precision("f64")
    ngw_ctx.set_reorthogonalization(True)
    try:
        sig = ngw_ctx.calculate_sigma_lanczos(qp, 31, 3.0, 0.05, ncv, poly)
    finally:
        ngw_ctx.set_reorthogonalization(False)
    plain = ngw_ctx.calculate_sigma_lanczos(qp, 31, 3.0, 0.05, ncv, poly)
    ref = _sigma_sum_over_states(mol, qp, 31, 3.0, 0.05, poly)
    print(f"Sigma_c vs sum over states: reorth {rel_err(sig, ref):.2e}, plain {rel_err(plain, ref):.2e}")
    assert rel_err(sig, ref) < 1e-8


def test_reorth_rejects_what_does_not_fit(ngw_ctx, wc):
    """A basis that does not fit HBM is NGW_ERR_NOMEM with a message, not a
    fault: 1 column x 262M steps x ncv = 180 x 8 B = 377 GB (the run stops
    at the allocation, before the first step)."""
    import ctypes as C

    from nanogw_amd import _lib

    ngw_ctx.set_isdf(wc.psi_intp, wc.mmtrx, wc.e1, wc.nv)
    ngw_ctx.set_reorthogonalization(True)
    try:
        n = wc.nv * wc.nc
        v0 = np.ones(n)
        vn = np.ones(1)
        poly = (C.c_double * 2)(1.0, 0.0)
        dp = C.POINTER(C.c_double)
        st = _lib.load().ngw_lanczos_spectra_isdf_block(ngw_ctx.handle, v0.ctypes.data_as(dp), vn.ctypes.data_as(dp),
                                                        n, None, 0, 4000 * 65536, poly, 1, None, 1, 0, None, None)
        assert st == -4, st
        assert "reorthogonalisation" in _lib.load().ngw_last_error().decode()
    finally:
        ngw_ctx.set_reorthogonalization(False)
