"""The irrep-blocked operator (ngw_set_isdf_blocks), the operator of NanoGW's
lowcomsym routines (shared/apply_matvec_lowcomsym.F90:155-412,
shared/lanczos_spectra_lowcomsym.F90:7) applied in factored form, against the
CPU oracle's explicit-C restatement on the pair matrix built from the same
blocks in the reference's pair order (block after block, valence outer,
conduction inner).  Tolerances as for the explicit form: operator and Horner
p(H) 1e-11 / 1e-10 relative, Lanczos coefficients 1e-10, spectra 1e-8."""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def bctx(gpu):
    from nanogw_amd.lanczos import IsdfContext

    ctx = IsdfContext(0)
    yield ctx
    ctx.close()


def _blocks(seed, nmu, nv_blk, nc_blk):
    rng = np.random.default_rng(seed)
    psiv = np.asfortranarray(rng.normal(size=(nmu, sum(nv_blk))) / np.sqrt(nmu))
    psic = np.asfortranarray(rng.normal(size=(nmu, sum(nc_blk))) / np.sqrt(nmu))
    g = rng.normal(size=(nmu, nmu)) / nmu
    mm = np.asfortranarray(g @ g.T + 0.1 * np.eye(nmu))
    ncv = int(np.dot(nv_blk, nc_blk))
    s = rng.uniform(0.6, 1.8, size=ncv)
    return psiv, psic, mm, s


def _pair_matrix(psiv, psic, nv_blk, nc_blk):
    """C(mu, p) = PsiC(mu, c) PsiV(mu, v), p = block after block, v outer, c inner
    (apply_matvec_lowcomsym.F90:225-233)."""
    cols = []
    v0 = c0 = 0
    for nv, nc in zip(nv_blk, nc_blk):
        for v in range(nv):
            for c in range(nc):
                cols.append(psic[:, c0 + c] * psiv[:, v0 + v])
        v0 += nv
        c0 += nc
    return np.asfortranarray(np.stack(cols, axis=1))


# four irrep blocks, one of them without valence states (no pairs)
NMU, NVB, NCB = 150, [3, 0, 5, 2], [20, 7, 13, 31]


def test_blocked_operator_matches_explicit_restatement(bctx):
    from nanogw_amd.lanczos import lanczos_polynomial
    from oracle import oracle as O

    psiv, psic, mm, s = _blocks(41, NMU, NVB, NCB)
    cm = _pair_matrix(psiv, psic, NVB, NCB)
    ncv = cm.shape[1]
    bctx.set_isdf_blocks(psiv, psic, NVB, NCB, mm, s)
    assert bctx.ncv == ncv == 3 * 20 + 5 * 13 + 2 * 31
    x = np.random.default_rng(42).normal(size=(ncv, 5))
    y = bctx.matvec_isdf_BLOCK(x)
    ref = np.stack([O.matvec_explicit(cm, mm, s, x[:, j]) for j in range(5)], axis=1)
    print("blocked matvec full", rel_err(y, ref))
    assert rel_err(y, ref) < 1e-11
    # TDA (apply_matvec_lowcomsym.F90:396-403): R x + 2 C^T M C x
    yt = bctx.matvec_isdf_BLOCK(x, tamm_d=True)
    reft = s[:, None] ** 2 * x + 2.0 * cm.T @ (mm @ (cm @ x))
    print("blocked matvec TDA", rel_err(yt, reft))
    assert rel_err(yt, reft) < 1e-11
    # one vector: the same kernels with N = nv_b columns per block
    y1 = bctx.matvec_isdf_BLOCK(x[:, 2])[:, 0]
    assert rel_err(y1, ref[:, 2]) < 1e-11
    e1 = np.concatenate([[-1.1, -0.8], np.linspace(0.4, 2.2, 9)])
    poly = lanczos_polynomial(e1, 2, 12)
    ph = bctx.polynomial_matvec_isdf_BLOCK(x[:, :2], poly)
    for j in range(2):
        assert rel_err(ph[:, j], O.polynomial_matvec_explicit(cm, mm, s, x[:, j], poly)) < 1e-10
    zz = np.linspace(-1.0, 4.0, 11) + 0.05j
    v = x[:, 0]
    spec, a, b = bctx.lanczos_spectra_isdf_BLOCK(v, [np.linalg.norm(v)], zz[:, None], 30, poly)
    sref, aref, bref, nk = O.lanczos_spectra_explicit(cm, mm, s, v, np.linalg.norm(v), zz, 30, poly)
    print("blocked lanczos a, b, spec", rel_err(a[:, 0], aref), rel_err(b[:, 0], bref),
          rel_err(spec[:, 0], sref))
    assert nk == 30
    assert rel_err(a[:, 0], aref) < 1e-10 and rel_err(b[:, 0], bref) < 1e-10
    assert rel_err(spec[:, 0], sref) < 1e-8


def test_one_block_is_the_factored_operator_with_pairs_transposed(bctx, mini):
    """One block holding every valence and conduction state is the full-pair
    operator of ngw_set_isdf with the pair index transposed (c inner instead
    of v inner), on a molecule's orbitals."""
    from oracle import oracle as O

    m = mini
    nv, nc = m.nv, m.nc
    s = O.sqrt_r(m.e1, nv, nc)  # pair iv + nv ic
    perm = np.array([iv + nv * ic for iv in range(nv) for ic in range(nc)])  # blocked p -> factored index
    psi = np.asfortranarray(m.psi_intp)
    bctx.set_isdf_blocks(psi[:, :nv], psi[:, nv:nv + nc], [nv], [nc], m.mmtrx, s[perm])
    x = np.random.default_rng(43).normal(size=(nv * nc, 4))
    for tda in (False, True):
        y = bctx.matvec_isdf_BLOCK(x[perm], tamm_d=tda)
        ref = O.matvec_block(m.psi_intp, m.mmtrx, s, x, nv, nc, tda)
        print("one block vs factored, tda", tda, rel_err(y, ref[perm]))
        assert rel_err(y, ref[perm]) < 1e-11


def test_blocked_mode_switching(bctx, mini):
    """Entries that need the factored operator refuse the blocked one, low
    precision is refused, bad blocks are refused, and ngw_set_isdf switches
    back."""
    from nanogw_amd._lib import NgwError
    from nanogw_amd.lanczos import lanczos_polynomial

    psiv, psic, mm, s = _blocks(44, NMU, NVB, NCB)
    bctx.set_isdf_blocks(psiv, psic, NVB, NCB, mm, s)
    m = mini
    poly = lanczos_polynomial(m.e1, m.nv, 12)
    with pytest.raises(NgwError):
        bctx.calculate_sigma_lanczos([0], 5, 0.5, 0.01, 10, poly)
    with pytest.raises(NgwError):
        bctx.set_precision("f16")
    with pytest.raises(NgwError):
        bctx.set_isdf_blocks(psiv, psic, NVB, NCB, mm, -s)
    with pytest.raises(NgwError):
        bctx.set_isdf_blocks(psiv[:, :0], psic, [0, 0, 0, 0], NCB, mm, s[:0])
    bctx.set_isdf(m.psi_intp, m.mmtrx, m.e1, m.nv)
    bctx.set_precision("f64")
    sig = bctx.calculate_sigma_lanczos([m.nv - 1], 5, 0.5, 0.01, 10, poly)
    assert np.all(np.isfinite(sig))
