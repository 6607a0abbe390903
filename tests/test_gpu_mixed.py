"""Mixed-precision operator (FP32 / BF16 / FP16 / BF16x3 GEMMs, FP64
recurrence) against the FP64 oracle.  Errors are stated per precision; the
bars: QP energies within 1 meV of the FP64 result for the precisions that
claim it (FP32, BF16x3, FP16), Sigma_c within 1e-4 relative on the QP window
(1e-3 for FP16).
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

RYD = 13.60569193


@pytest.fixture(scope="module")
def wellcond():
    """Small molecule with a Tikhonov-regularised zeta solve (well-conditioned M,
    like the Nmu << nv nc regime of the large configurations)."""
    from nanogw_amd import synth

    return synth.make_molecule(5.0, 0.5, 6, 30, 60, seed=31337, ridge=1e-6)


@pytest.mark.parametrize("kind,tol", [(0, 1e-5), (1, 2e-2), (2, 5e-5), (3, 4e-3)])
def test_gemm_lowp_lane_layout(gpu, ngw_ctx, kind, tol):
    import torch

    rng = np.random.default_rng(7)
    for (m, n, k) in [(16, 16, 32), (70, 45, 33), (64, 128, 96), (129, 3, 200)]:
        a = rng.normal(size=(m, k))
        b = rng.normal(size=(k, n))
        ta = torch.tensor(a.T.copy(), dtype=torch.float64, device=gpu)
        tb = torch.tensor(b.T.copy(), dtype=torch.float64, device=gpu)
        tc = torch.zeros((n, m), dtype=torch.float64, device=gpu)
        torch.cuda.synchronize()
        ngw_ctx.gemm_lowp_dev(kind, m, n, k, ta.data_ptr(), m, tb.data_ptr(), k, tc.data_ptr(), m)
        c = tc.cpu().numpy().T
        # relative to the scale of the products
        err = np.abs(c - a @ b).max() / np.sqrt(k)
        print("lowp gemm", kind, (m, n, k), err)
        assert err < tol, (kind, m, n, k, err)


@pytest.mark.parametrize("nvec", [5, 47])
@pytest.mark.parametrize("prec,tol", [("f32", 1e-5), ("bf16x3", 1e-4), ("bf16", 3e-2), ("f16", 5e-3)])
def test_matvec_lowp_vs_oracle(ngw_ctx, wellcond, prec, tol, nvec):
    """nvec = 47 spreads nv * nvec = 282 columns over several 64- and
    128-wide K1 tiles, ending in a ragged one; every column is checked."""
    from oracle import oracle as O

    m = wellcond
    ngw_ctx.set_isdf(m.psi_intp, m.mmtrx, m.e1, m.nv)
    ngw_ctx.set_precision(prec)
    try:
        x = np.random.default_rng(1).normal(size=(m.nv * m.nc, nvec))
        y = ngw_ctx.matvec_isdf_BLOCK(x)
        s = O.sqrt_r(m.e1, m.nv, m.nc)
        ref = O.matvec_block(m.psi_intp, m.mmtrx, s, x, m.nv, m.nc)
        err = max(rel_err(y[:, j], ref[:, j]) for j in range(nvec))
        print("matvec", prec, nvec, err)
        assert err < tol
    finally:
        ngw_ctx.set_precision("f64")


@pytest.fixture(scope="module")
def tall():
    """nc and Nmu above 256: several M tiles of every K1/K3 configuration,
    the last one ragged."""
    from nanogw_amd import synth

    return synth.make_molecule(6.0, 0.5, 6, 300, 300, seed=2718, ridge=1e-6, cvt_iter=40)


@pytest.mark.parametrize("prec,tol", [("f32", 1e-5), ("bf16x3", 1e-4), ("bf16", 3e-2), ("f16", 5e-3)])
def test_tile_configs_and_autotune_vs_oracle(ngw_ctx, tall, prec, tol):
    """Every K1/K3 tile configuration (shape, slabs in flight) gives the
    oracle's operator to the precision's tolerance, per column; the autotuner
    returns valid indices and its choice is also correct."""
    from oracle import oracle as O

    m = tall
    ngw_ctx.set_isdf(m.psi_intp, m.mmtrx, m.e1, m.nv)
    ngw_ctx.set_precision(prec)
    try:
        x = np.random.default_rng(4).normal(size=(m.nv * m.nc, 37))
        ref = O.matvec_block(m.psi_intp, m.mmtrx, O.sqrt_r(m.e1, m.nv, m.nc), x, m.nv, m.nc)
        outs = []
        for cfg in range(6):
            ngw_ctx.set_tile_config(cfg, cfg)
            y = ngw_ctx.matvec_isdf_BLOCK(x)
            err = max(rel_err(y[:, j], ref[:, j]) for j in range(x.shape[1]))
            print("tile config", prec, cfg, err)
            assert err < tol, (cfg, err)
            outs.append(y)
        if prec == "f16":
            # K1 configurations 1 and 3 fill their LDS slabs by direct-to-LDS
            # loads; configurations 5 and 2 have the same tile widths with
            # slabs that pass through registers.  K3 configuration 2 is direct
            # to LDS and 3 its register twin, and K3's products do not depend
            # on the tile (1 and 5 are 128 x 192 and 128 x 256, both direct to
            # LDS).  So the products are the same bit for bit
            for a, b in ((1, 5), (3, 2)):
                print("direct-to-LDS vs register slabs", a, b, np.abs(outs[a] - outs[b]).max())
                assert np.array_equal(outs[a], outs[b]), (a, b)
        k1, k3 = ngw_ctx.autotune(37, reps=1)
        assert 0 <= k1 < 6 and 0 <= k3 < 6
        y = ngw_ctx.matvec_isdf_BLOCK(x)
        assert max(rel_err(y[:, j], ref[:, j]) for j in range(x.shape[1])) < tol
    finally:
        ngw_ctx.set_tile_config(0, 0)
        ngw_ctx.set_precision("f64")


@pytest.mark.parametrize("prec,tol", [("f32", 1e-5), ("bf16x3", 1e-4), ("f16", 5e-3)])
def test_clenshaw_polynomial_vs_oracle(ngw_ctx, wellcond, prec, tol):
    """Low-precision p(H) (Chebyshev/Clenshaw form of the pinned monomial
    polynomial) against the oracle's FP64 Horner."""
    from nanogw_amd.lanczos import lanczos_polynomial
    from oracle import oracle as O

    m = wellcond
    poly = lanczos_polynomial(m.e1, m.nv, 12)
    ngw_ctx.set_isdf(m.psi_intp, m.mmtrx, m.e1, m.nv)
    ngw_ctx.set_precision(prec)
    try:
        x = np.random.default_rng(2).normal(size=(m.nv * m.nc, 3))
        y = ngw_ctx.polynomial_matvec_isdf_BLOCK(x, poly)
        s = O.sqrt_r(m.e1, m.nv, m.nc)
        ref = O.polynomial_matvec_block(m.psi_intp, m.mmtrx, s, x, m.nv, m.nc, poly)
        err = rel_err(y, ref)
        print("p(H)", prec, err)
        assert err < tol
    finally:
        ngw_ctx.set_precision("f64")


def test_f16_scaling_exact_and_range(ngw_ctx, wellcond):
    """FP16 operands carry exact power-of-two scales per vector: scaling the
    input by 2^k scales the output by exactly 2^k (bit for bit) far outside
    FP16's range (65504 .. 6e-8), and vectors of very different norms in one
    block each keep their own accuracy against the oracle."""
    from oracle import oracle as O

    m = wellcond
    ngw_ctx.set_isdf(m.psi_intp, m.mmtrx, m.e1, m.nv)
    ngw_ctx.set_precision("f16")
    try:
        x = np.random.default_rng(8).normal(size=(m.nv * m.nc, 4))
        y = ngw_ctx.matvec_isdf_BLOCK(x)
        for k in (60, -60):
            yk = ngw_ctx.matvec_isdf_BLOCK(np.ldexp(x, k))
            assert np.array_equal(yk, np.ldexp(y, k)), k
        mixed = x * np.array([1e-30, 1.0, 3e12, 7e-9])
        ym = ngw_ctx.matvec_isdf_BLOCK(mixed)
        ref = O.matvec_block(m.psi_intp, m.mmtrx, O.sqrt_r(m.e1, m.nv, m.nc), mixed, m.nv, m.nc)
        errs = [rel_err(ym[:, j], ref[:, j]) for j in range(4)]
        print("f16 mixed-norm block", errs)
        assert max(errs) < 5e-3
        z = ngw_ctx.matvec_isdf_BLOCK(np.zeros((m.nv * m.nc, 2)))
        assert not z.any()
    finally:
        ngw_ctx.set_precision("f64")


def test_clenshaw_fp64_equals_horner(ngw_ctx, wellcond):
    """The Chebyshev re-expression is the same polynomial: with FP32 GEMMs the
    difference to FP64 Horner is at the FP32 level, far below the monomial
    cancellation error FP32 Horner would give."""
    from nanogw_amd.lanczos import lanczos_polynomial

    m = wellcond
    poly = lanczos_polynomial(m.e1, m.nv, 12)
    ngw_ctx.set_isdf(m.psi_intp, m.mmtrx, m.e1, m.nv)
    x = np.random.default_rng(3).normal(size=(m.nv * m.nc, 2))
    y64 = ngw_ctx.polynomial_matvec_isdf_BLOCK(x, poly)
    ngw_ctx.set_precision("f32")
    try:
        y32 = ngw_ctx.polynomial_matvec_isdf_BLOCK(x, poly)
    finally:
        ngw_ctx.set_precision("f64")
    assert rel_err(y32, y64) < 1e-5


@pytest.mark.parametrize("prec,qp_tol_ev", [("f32", 1e-3), ("bf16x3", 1e-3), ("bf16", 2e-2), ("f16", 1e-3)])
def test_sigma_qp_mixed_vs_fp64(ngw_ctx, wellcond, prec, qp_tol_ev):
    """Sigma_c and QP energies of the mixed-precision path against the FP64
    oracle (niter = 60, all bands, 4 QP states)."""
    from nanogw_amd import synth
    from nanogw_amd.lanczos import lanczos_polynomial
    from oracle import oracle as O

    m = wellcond
    poly = lanczos_polynomial(m.e1, m.nv, 12)
    qp = synth.qp_states(m.nv, 4)
    sd = synth.sigma_defaults(m.nv)
    nen, deltae, ecuts, niter = sd["nen"], sd["deltae"], sd["ecuts"], 60
    ref = O.sigma_lanczos(m.psi_intp, m.mmtrx, m.e1, m.nv, m.nc, qp, nen, deltae, ecuts, niter, poly, 36)
    ngw_ctx.set_isdf(m.psi_intp, m.mmtrx, m.e1, m.nv)
    ngw_ctx.set_precision(prec)
    try:
        sig = ngw_ctx.calculate_sigma_lanczos(qp, nen, deltae, ecuts, niter, poly)
    finally:
        ngw_ctx.set_precision("f64")
    sigx = synth.exchange_diag(m.psi_intp, m.mmtrx, m.nv, qp)
    e = m.e1[qp] * RYD
    got = ngw_ctx.qp_newton(sig * RYD, deltae * RYD, e, e, sigx * RYD, np.zeros(len(qp)))["eqp"]
    exp = np.array([O.qp_solve(ref[:, i] * RYD, deltae * RYD, e[i], e[i], sigx[i] * RYD, 0.0)["eqp"]
                    for i in range(len(qp))])
    dq = np.abs(got - exp).max()
    print(f"{prec}: max |dE_qp| = {dq * 1e3:.4f} meV, Sigma_c rel err {rel_err(sig, ref):.2e}")
    assert dq < qp_tol_ev


@pytest.mark.parametrize("nv,nc", [(130, 20), (1, 70), (17, 200)])
@pytest.mark.parametrize("prec,tol", [("f16", 5e-3), ("bf16x3", 1e-4)])
def test_conversion_tiles_odd_shapes(ngw_ctx, nv, nc, prec, tol):
    """x' conversion blocks for nv above one 128-wide tile (two v blocks, the
    second ragged), for nv = 1 (the widest c block) and for a ragged c tail,
    on a random well-conditioned operator against the oracle."""
    from oracle import oracle as O

    rng = np.random.default_rng(nv * 1000 + nc)
    nmu = 72
    psi = np.asfortranarray(rng.normal(size=(nmu, nv + nc)) / np.sqrt(nmu))
    g = rng.normal(size=(nmu, nmu)) / nmu
    mm = np.asfortranarray(g @ g.T + np.eye(nmu))
    e1 = np.concatenate([np.sort(rng.uniform(-1.5, -0.5, nv)), np.sort(rng.uniform(-0.2, 3.0, nc))])
    ngw_ctx.set_isdf(psi, mm, e1, nv)
    ngw_ctx.set_precision(prec)
    try:
        x = rng.normal(size=(nv * nc, 9))
        y = ngw_ctx.matvec_isdf_BLOCK(x)
        ref = O.matvec_block(psi, mm, O.sqrt_r(e1, nv, nc), x, nv, nc)
        err = max(rel_err(y[:, j], ref[:, j]) for j in range(x.shape[1]))
        print("odd shape", prec, nv, nc, err)
        assert err < tol
    finally:
        ngw_ctx.set_precision("f64")
