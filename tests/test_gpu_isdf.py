"""Device ISDF fit (zeta = B A^-1, isdf_parallel_sym_UltraLowMem.F90:448-642)
and the device-resident psi -> zeta -> M chain, against the numpy
restatement.  Pins from the mathematics: with an exactly solvable fit (no
ridge) the interpolation vectors are the Kronecker delta at the points,
zeta(r_nu, mu) = delta(mu, nu), since B's rows at the points are A's rows.
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small_fit():
    """Few points (n_intp = 16 < nv nc): A is well conditioned without a ridge."""
    from nanogw_amd import synth

    return synth.make_molecule(4.5, 0.5, 4, 12, 16, seed=4711, keep_psi=True, with_kernel=False, cvt_iter=60)


@pytest.fixture(scope="module")
def ridge_fit():
    from nanogw_amd import synth

    return synth.make_molecule(4.5, 0.5, 4, 12, 40, seed=12345, keep_psi=True, with_kernel=False)


def test_zeta_delta_at_points_and_numpy(ngw_ctx, small_fit):
    from nanogw_amd import synth

    m = small_fit
    zeta, used_lu = ngw_ctx.isdf_zeta(m.psi, m.intp, m.nv)
    assert not used_lu
    at_points = zeta[m.intp, :]
    print("zeta at points - I:", np.abs(at_points - np.eye(m.n_intp)).max())
    assert np.abs(at_points - np.eye(m.n_intp)).max() < 1e-9
    ref = synth.isdf_zeta(m.psi, m.intp, m.nv)
    print("zeta vs numpy:", rel_err(zeta, ref))
    assert rel_err(zeta, ref) < 1e-10


def test_zeta_ridge_vs_numpy(ngw_ctx, ridge_fit):
    from nanogw_amd import synth

    m = ridge_fit
    zeta, used_lu = ngw_ctx.isdf_zeta(m.psi, m.intp, m.nv, ridge=1e-6)
    ref = synth.isdf_zeta(m.psi, m.intp, m.nv, ridge=1e-6)
    print("ridge zeta vs numpy:", rel_err(zeta, ref), "LU" if used_lu else "Cholesky")
    assert rel_err(zeta, ref) < 1e-8


def test_device_chain_psi_zeta_m(gpu, ngw_ctx, ridge_fit):
    """psi on the device -> zeta on the device -> M, against the host chain
    (numpy zeta, then the oracle's FFT Poisson and zeta^T V zeta)."""
    import torch

    from nanogw_amd import synth
    from oracle import oracle as O

    m = ridge_fit
    ngrid, nband = m.psi.shape
    tpsi = torch.as_tensor(np.asfortranarray(m.psi).T.copy(), device=gpu)  # (nband, ngrid) = column-major psi
    tz = torch.zeros((m.n_intp, ngrid), dtype=torch.float64, device=gpu)
    torch.cuda.synchronize()
    ngw_ctx.isdf_zeta_dev(tpsi.data_ptr(), ngrid, ngrid, nband, m.nv, m.intp, 1e-6, tz.data_ptr(), ngrid)
    mm = ngw_ctx.isdf_coulomb_kernel_dev(tz.data_ptr(), ngrid, ngrid, m.n_intp, m.gidx, m.nfft, m.h, m.trunc)
    zref = synth.isdf_zeta(m.psi, m.intp, m.nv, ridge=1e-6)
    mref = O.isdf_kernel(zref, m.gidx, m.nfft, m.h, m.trunc)
    print("device chain M vs host chain:", rel_err(mm, mref))
    assert rel_err(mm, mref) < 1e-8
    assert rel_err(tz.cpu().numpy().T, zref) < 1e-8


def test_zeta_argument_checks(ngw_ctx, small_fit):
    from nanogw_amd._lib import NgwError

    m = small_fit
    bad = np.array(m.intp)
    bad[0] = m.psi.shape[0]
    with pytest.raises(NgwError):
        ngw_ctx.isdf_zeta(m.psi, bad, m.nv)
    with pytest.raises(NgwError):
        ngw_ctx.isdf_zeta(m.psi, m.intp, m.nv, ridge=-1.0)


def test_cvt_points_vs_oracle(ngw_ctx):
    """Device CVT against the numpy restatement with the same tie rules:
    identical point sets and iteration counts; points unique and on the grid."""
    from nanogw_amd import synth
    from oracle import oracle as O

    for (radius, h, nv, n_intp, seed) in [(4.0, 0.5, 3, 24, 5), (5.0, 0.5, 4, 60, 6)]:
        gidx = synth.sphere_grid(radius, h)
        rng = np.random.default_rng(seed)
        psi = synth.orbitals(gidx, h, nv + 6, rng, radius)
        rho = 2.0 * np.sum(psi[:, :nv] ** 2, axis=1)
        pts = gidx.astype(np.float64) * h
        lo, hi = pts.min(axis=0), pts.max(axis=0)
        init = lo[None, :] + rng.uniform(size=(n_intp, 3)) * 0.9 * (hi - lo)[None, :]
        got, it = ngw_ctx.cvt_points(gidx, rho, init, h, tol=1e-3, max_iter=100)
        ref, it_ref = O.cvt_lattice(gidx, rho, init, h, tol=1e-3, max_iter=100)
        print("cvt", radius, n_intp, "iterations", it, it_ref, "mismatched points", int(np.sum(got != ref)))
        assert len(set(got.tolist())) == n_intp and got.min() >= 0 and got.max() < len(gidx)
        assert it == it_ref
        assert np.array_equal(got, ref)
