"""GPU parity of the FP64 HIP path against the CPU oracle (through the C ABI).

Tolerances: the FP64 operator, Horner and Lanczos coefficients agree with the
oracle to round-off (relative 1e-10 or better, summation order differs);
Sigma_c(omega) within 1e-6 relative and QP energies within 1 meV, the bars
NanoGW's FP64 path must meet.
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

RYD = 13.60569193


def _poly(mol, npoly=12):
    from nanogw_amd.lanczos import lanczos_polynomial

    return lanczos_polynomial(mol.e1, mol.nv, npoly)


def test_gemm_f64_lane_layout(gpu, ngw_ctx):
    """The MFMA f64 16x16x4 GEMM reproduces A @ B on ragged shapes."""
    import torch

    rng = np.random.default_rng(0)
    for (m, n, k) in [(16, 16, 4), (70, 45, 33), (64, 64, 16), (129, 3, 200), (1, 1, 1)]:
        a = rng.normal(size=(m, k))
        b = rng.normal(size=(k, n))
        ta = torch.tensor(a.T.copy(), dtype=torch.float64, device=gpu)  # column-major a
        tb = torch.tensor(b.T.copy(), dtype=torch.float64, device=gpu)
        tc = torch.zeros((n, m), dtype=torch.float64, device=gpu)
        torch.cuda.synchronize()
        ngw_ctx.gemm_f64_dev(m, n, k, ta.data_ptr(), m, tb.data_ptr(), k, tc.data_ptr(), m)
        c = tc.cpu().numpy().T
        assert rel_err(c, a @ b) < 1e-13, (m, n, k)


@pytest.mark.parametrize("tda", [False, True])
@pytest.mark.parametrize("blksz", [1, 3, 17])
def test_matvec_block_vs_oracle(ngw_ctx, mini, tda, blksz):
    from oracle import oracle as O

    ngw_ctx.set_isdf(mini.psi_intp, mini.mmtrx, mini.e1, mini.nv)
    ncv = mini.nv * mini.nc
    x = np.random.default_rng(blksz).normal(size=(ncv, blksz))
    y = ngw_ctx.matvec_isdf_BLOCK(x, tamm_d=tda)
    s = O.sqrt_r(mini.e1, mini.nv, mini.nc)
    ref = O.matvec_block(mini.psi_intp, mini.mmtrx, s, x, mini.nv, mini.nc, tda)
    print("matvec", tda, blksz, rel_err(y, ref))
    assert rel_err(y, ref) < 1e-11


def test_polynomial_matvec_vs_oracle(ngw_ctx, mini):
    from oracle import oracle as O

    ngw_ctx.set_isdf(mini.psi_intp, mini.mmtrx, mini.e1, mini.nv)
    poly = _poly(mini)
    x = np.random.default_rng(5).normal(size=(mini.nv * mini.nc, 4))
    y = ngw_ctx.polynomial_matvec_isdf_BLOCK(x, poly)
    s = O.sqrt_r(mini.e1, mini.nv, mini.nc)
    ref = O.polynomial_matvec_block(mini.psi_intp, mini.mmtrx, s, x, mini.nv, mini.nc, poly)
    assert rel_err(y, ref) < 1e-9


def test_lanczos_spectra_vs_oracle(ngw_ctx, mini):
    from oracle import oracle as O

    ngw_ctx.set_isdf(mini.psi_intp, mini.mmtrx, mini.e1, mini.nv)
    poly = _poly(mini)
    ncv = mini.nv * mini.nc
    rng = np.random.default_rng(9)
    v0 = rng.normal(size=(ncv, 5))
    vn = np.linalg.norm(v0, axis=0)
    zz = (rng.uniform(-2, 4, size=(31, 5)) + 0.05j)
    zz[-1, :] = 0.0
    niter = 25
    spec, a, b = ngw_ctx.lanczos_spectra_isdf_BLOCK(v0, vn, zz, niter, poly)
    s = O.sqrt_r(mini.e1, mini.nv, mini.nc)
    rs, ra, rb = O.lanczos_spectra_block(mini.psi_intp, mini.mmtrx, s, v0, vn, zz, niter, poly,
                                         mini.nv, mini.nc)
    assert rel_err(a[:6], ra[:6]) < 1e-10
    assert rel_err(spec, rs) < 1e-6


def test_sigma_lanczos_mini_vs_oracle(ngw_ctx, mini):
    from oracle import oracle as O

    ngw_ctx.set_isdf(mini.psi_intp, mini.mmtrx, mini.e1, mini.nv)
    poly = _poly(mini)
    qp = [mini.nv - 1, mini.nv]
    nen, deltae, ecuts, niter = 41, 4.0, 0.02, 30
    sig = ngw_ctx.calculate_sigma_lanczos(qp, nen, deltae, ecuts, niter, poly)
    ref = O.sigma_lanczos(mini.psi_intp, mini.mmtrx, mini.e1, mini.nv, mini.nc, qp, nen, deltae, ecuts,
                          niter, poly, 7)
    print("sigma mini", rel_err(sig, ref))
    assert rel_err(sig, ref) < 1e-6


@pytest.mark.parametrize("tda", [False, True])
def test_sigma_band_ranges_add_up(ngw_ctx, mini, tda):
    """Sigma_c over band ranges (occupied part, two empty parts, a range that
    straddles the gap) sums to the full band sum: the per-device unit of a
    band-sharded run."""
    ngw_ctx.set_isdf(mini.psi_intp, mini.mmtrx, mini.e1, mini.nv)
    poly = np.array([1.0, 0.0]) if tda else _poly(mini)
    qp = [mini.nv - 1, mini.nv]
    args = (qp, 31, 4.0, 0.02, 20, poly)
    full = ngw_ctx.calculate_sigma_lanczos(*args, tamm_d=tda)
    nv, nb = mini.nv, mini.nv + mini.nc
    for cuts in ([0, nv, nv + 5, nb], [0, nv - 1, nv + 2, nb]):
        parts = [ngw_ctx.calculate_sigma_lanczos(*args, tamm_d=tda, bands=(lo, hi))
                 for lo, hi in zip(cuts[:-1], cuts[1:])]
        assert all(np.abs(p).max() > 0 for p in parts)
        assert rel_err(sum(parts), full) < 1e-12, cuts


def test_sigma_lanczos_tda_vs_oracle(ngw_ctx, mini_tda):
    from oracle import oracle as O

    m = mini_tda
    ngw_ctx.set_isdf(m.psi_intp, m.mmtrx, m.e1, m.nv)
    qp = [m.nv - 1, m.nv]
    poly = np.array([1.0, 0.0])  # TDA: p(x) = x (calculate_sigma_lanczos_lowcomsym.F90z:242-245)
    sig = ngw_ctx.calculate_sigma_lanczos(qp, 21, 3.0, 0.02, 20, poly, tamm_d=True)
    ref = O.sigma_lanczos(m.psi_intp, m.mmtrx, m.e1, m.nv, m.nc, qp, 21, 3.0, 0.02, 20, poly, 5,
                          tda=True)
    assert rel_err(sig, ref) < 1e-6


def test_sigma_c1_fp64_parity(ngw_ctx, c1_inputs, c1_sigma):
    """Config 1 (niter = 200, no reorthogonalisation, as the reference).

    Bars: QP energies within 1 meV of the oracle; Sigma_c within 1e-6
    (relative to max |Sigma_c| of the state) on the QP window |w - E_qp| <= 1 eV.
    Away from that window Sigma_c of the reference algorithm is not fixed by
    FP64 arithmetic at this niter: two oracle runs whose inputs differ by
    1e-15 relative differ by ~1.2% near the resonances (tests/golden: the
    sensitivity study is described with the parity results).  The full-grid
    difference is therefore only required to lie inside 3x that envelope.
    """
    d = c1_inputs
    ngw_ctx.set_isdf(d["psi_intp"], d["mmtrx"], d["e1"], int(d["nv"]))
    qp = [int(x) for x in d["qp"]]
    nen, deltae = int(d["nen"]), float(d["deltae"])
    sig = ngw_ctx.calculate_sigma_lanczos(qp, nen, deltae, float(d["ecuts"]),
                                          int(c1_sigma["niter"]), d["poly"])
    ref = c1_sigma["sigma"]
    e1 = d["e1"]
    res = ngw_ctx.qp_newton(sig * RYD, deltae * RYD, e1[qp] * RYD, e1[qp] * RYD,
                            d["sigx"] * RYD, np.zeros(len(qp)))
    ref_qp = c1_sigma["qp_results"]
    dq = np.abs(res["eqp"] - ref_qp[:, 0]).max()
    print("c1 QP |dE| (eV)", dq, res["eqp"])
    assert dq < 1e-3
    assert np.all(res["ierr"] == ref_qp[:, 4])
    for i, m in enumerate(qp):
        w = (deltae / (nen - 1) * np.arange(nen) + e1[m] - deltae / 2) * RYD
        win = np.abs(w - ref_qp[i, 0]) <= 1.0
        err = np.abs(sig[win, i] - ref[win, i]).max() / np.abs(ref[:, i]).max()
        full = np.abs(sig[:, i] - ref[:, i]).max() / np.abs(ref[:, i]).max()
        print(f"state {m}: QP-window rel err {err:.2e}, full-grid {full:.2e}")
        assert err < 1e-6
        assert full < 0.036


def test_qp_newton_vs_oracle(ngw_ctx):
    from oracle import oracle as O

    rng = np.random.default_rng(3)
    nen, nqp = 200, 5
    x = np.linspace(-1, 1, nen)
    sig = np.empty((nen, nqp), np.complex128)
    for i in range(nqp):
        c = rng.normal(size=4)
        sig[:, i] = (c[0] + c[1] * x + c[2] * np.sin(3 * x) + 0.3 * c[3] * x ** 3) + 1j * (0.1 * np.cos(x))
    edft = rng.uniform(-10, 0, nqp)
    sigx = rng.uniform(-5, 0, nqp)
    vxc = rng.uniform(-3, 0, nqp)
    res = ngw_ctx.qp_newton(sig * 13.6, 27.2, edft, edft, sigx, vxc)
    for i in range(nqp):
        r = O.qp_solve(sig[:, i] * 13.6, 27.2, edft[i], edft[i], sigx[i], vxc[i])
        assert abs(res["eqp"][i] - r["eqp"]) < 1e-9
        assert abs(res["zfac"][i] - r["zfac"]) < 1e-9
        assert abs(res["imsig"][i] - r["imsig"]) < 1e-9
        assert res["ierr"][i] == r["ierr"]


def test_qp_newton_vxc_from_sigma_mtxel(ngw_ctx, tmp_path):
    """V_xc read from sigma_mtxel.dat (SIGMA/vxc.F90z:56-111) enters the QP
    Newton solve on the device as printsigma uses it (f = E_dft + Sigma_x +
    Sigma_c(E) - V_xc - E); the oracle solves with the same values."""
    from nanogw_amd import nanogw_files as F
    from oracle import oracle as O

    rng = np.random.default_rng(8)
    nen, bands = 200, [4, 5, 6]
    x = np.linspace(-1, 1, nen)
    sig = np.stack([(rng.normal() + 0.4 * x + 0.2 * np.sin(2 * x)) + 0.05j * np.cos(x) for _ in bands], axis=1)
    path = tmp_path / "sigma_mtxel.dat"
    path.write_text("".join(f"1 1 {b} {b} {-11.0 + 0.7 * k:.6f}\n" for k, b in enumerate(bands)))
    vxc_ry, _ = F.read_vxc_matrix_elements(str(path), 1, 1, bands)
    vxc_ev = vxc_ry.real * F.RYD
    edft = rng.uniform(-10, 0, len(bands))
    sigx = rng.uniform(-5, 0, len(bands))
    res = ngw_ctx.qp_newton(sig * 13.6, 27.2, edft, edft, sigx, vxc_ev)
    res0 = ngw_ctx.qp_newton(sig * 13.6, 27.2, edft, edft, sigx, np.zeros(len(bands)))
    for i in range(len(bands)):
        r = O.qp_solve(sig[:, i] * 13.6, 27.2, edft[i], edft[i], sigx[i], vxc_ev[i])
        assert abs(res["eqp"][i] - r["eqp"]) < 1e-9 and res["ierr"][i] == r["ierr"]
        assert abs(res["eqp"][i] - res0["eqp"][i]) > 1.0  # V_xc of about -10 eV moves E_qp


def test_errors_raise(ngw_ctx, mini):
    """Size mismatches and bad indices come back as statuses, not faults."""
    import ctypes as C

    from nanogw_amd._lib import NgwError

    ngw_ctx.set_isdf(mini.psi_intp, mini.mmtrx, mini.e1, mini.nv)
    lib = ngw_ctx._lib
    v = np.ones(7)
    spec = np.zeros(6)
    dp = C.POINTER(C.c_double)
    st = lib.ngw_lanczos_spectra_isdf_block(ngw_ctx.handle, v.ctypes.data_as(dp), v.ctypes.data_as(dp), 7,
                                            spec.ctypes.data_as(dp), 3, 5, v.ctypes.data_as(dp), 1,
                                            spec.ctypes.data_as(dp), 1, 0, None, None)
    assert st == -1 and b"ncv" in lib.ngw_last_error()
    with pytest.raises(NgwError):
        ngw_ctx.calculate_sigma_lanczos([999], 10, 1.0, 0.01, 5, [1.0, 0.0])
    with pytest.raises(NgwError):
        ngw_ctx.isdf_coulomb_kernel(mini.zeta, mini.gidx, (4, 4, 4), mini.h, mini.trunc)


def test_isdf_coulomb_kernel_vs_oracle(ngw_ctx, mini):
    """Device V build (hipFFT Poisson + zeta^T V zeta) against the numpy restatement."""
    from oracle import oracle as O

    m_gpu = ngw_ctx.isdf_coulomb_kernel(mini.zeta, mini.gidx, mini.nfft, mini.h, mini.trunc)
    m_ref = O.isdf_kernel(mini.zeta, mini.gidx, mini.nfft, mini.h, mini.trunc)
    print("V build rel err", rel_err(m_gpu, m_ref))
    assert rel_err(m_gpu, m_ref) < 1e-10
    # ragged leading dimension and a non-cubic box
    z = np.asfortranarray(np.vstack([mini.zeta, np.zeros((5, mini.zeta.shape[1]))]))
    nf = (mini.nfft[0], mini.nfft[1] + 2, mini.nfft[2] + 4)
    m2 = ngw_ctx.isdf_coulomb_kernel(z[: mini.ngrid], mini.gidx, nf, mini.h, mini.trunc)
    r2 = O.isdf_kernel(mini.zeta, mini.gidx, nf, mini.h, mini.trunc)
    assert rel_err(m2, r2) < 1e-10


def test_sigma_x_isdf(ngw_ctx, mini):
    """Device bare exchange: the numpy expression on the same inputs, and the
    closed form sum_v sum_mu (Psi_m Psi_v)^2 when M is the identity."""
    from nanogw_amd import synth

    m = mini
    qp = [m.nv - 1, m.nv, m.nv + 2]
    ngw_ctx.set_isdf(m.psi_intp, m.mmtrx, m.e1, m.nv)
    got = ngw_ctx.sigma_x(qp)
    ref = synth.exchange_diag(m.psi_intp, m.mmtrx, m.nv, qp)
    assert rel_err(got, ref) < 1e-12
    eye = np.eye(m.n_intp)
    ngw_ctx.set_isdf(m.psi_intp, eye, m.e1, m.nv)
    got = ngw_ctx.sigma_x(qp)
    closed = [-np.sum((m.psi_intp[:, [q]] * m.psi_intp[:, :m.nv]) ** 2) for q in qp]
    assert rel_err(got, np.array(closed)) < 1e-13
