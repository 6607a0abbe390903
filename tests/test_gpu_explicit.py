"""The explicit-C operator (ngw_set_explicit_c), the form of NanoGW's
single-vector matvec_isdf / polynomial_matvec_isdf / lanczos_spectra_isdf
(shared/apply_matvec.f90:25, :142; shared/lanczos_spectra.f90:7), against the
CPU oracle.  Tolerances: operator and Horner p(H) to round-off (1e-11 /
1e-10 relative), Lanczos coefficients 1e-10, spectra 1e-8."""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ectx(gpu):
    from nanogw_amd.lanczos import IsdfContext

    ctx = IsdfContext(0)
    yield ctx
    ctx.close()


def test_explicit_pairs_equal_factored_operator(ectx, mini):
    """C built from the orbital pair products: same operator as the factored
    one (full and TDA), and a permuted pair list permutes Hvec."""
    from oracle import oracle as O

    m = mini
    nv, nc = m.nv, m.nc
    s = O.sqrt_r(m.e1, nv, nc)
    cm = O.pair_cmtrx(m.psi_intp, nv, nc)
    ectx.set_explicit_c(cm, m.mmtrx, s)
    x = np.random.default_rng(31).normal(size=(nv * nc, 3))
    for tda in (False, True):
        y = ectx.matvec_isdf_BLOCK(x, tamm_d=tda)
        ref = O.matvec_block(m.psi_intp, m.mmtrx, s, x, nv, nc, tda)
        print("explicit matvec tda", tda, rel_err(y, ref))
        assert rel_err(y, ref) < 1e-11
    perm = np.random.default_rng(32).permutation(nv * nc)
    ectx.set_explicit_c(np.asfortranarray(cm[:, perm]), m.mmtrx, s[perm])
    yp = ectx.matvec_isdf_BLOCK(x[perm])
    ref = O.matvec_block(m.psi_intp, m.mmtrx, s, x, nv, nc)
    assert rel_err(yp, ref[perm]) < 1e-11


def test_explicit_general_c_single_vector(ectx):
    """A C that is no pair product (ncv not nval*ncond), one vector at a time,
    against the single-vector restatement."""
    from nanogw_amd.lanczos import lanczos_polynomial
    from oracle import oracle as O

    rng = np.random.default_rng(33)
    nmu, ncv = 37, 101
    cm = np.asfortranarray(rng.normal(size=(nmu, ncv)) / np.sqrt(nmu))
    g = rng.normal(size=(nmu, nmu)) / nmu
    mm = np.asfortranarray(g @ g.T + 0.1 * np.eye(nmu))
    s = rng.uniform(0.5, 1.5, size=ncv)
    ectx.set_explicit_c(cm, mm, s)
    v = rng.normal(size=ncv)
    y = ectx.matvec_isdf_BLOCK(v)[:, 0]
    assert rel_err(y, O.matvec_explicit(cm, mm, s, v)) < 1e-11
    e1 = np.concatenate([[-1.2, -0.9], np.linspace(0.3, 2.0, 8)])
    poly = lanczos_polynomial(e1, 2, 12)
    ph = ectx.polynomial_matvec_isdf_BLOCK(v, poly)[:, 0]
    assert rel_err(ph, O.polynomial_matvec_explicit(cm, mm, s, v, poly)) < 1e-10
    zz = np.linspace(-1.0, 4.0, 11) + 0.05j
    spec, a, b = ectx.lanczos_spectra_isdf_BLOCK(v, [np.linalg.norm(v)], zz[:, None], 30, poly)
    sref, aref, bref, nk = O.lanczos_spectra_explicit(cm, mm, s, v, np.linalg.norm(v), zz, 30, poly)
    print("explicit lanczos a, b, spec", rel_err(a[:, 0], aref), rel_err(b[:, 0], bref), rel_err(spec[:, 0], sref))
    assert nk == 30
    assert rel_err(a[:, 0], aref) < 1e-10 and rel_err(b[:, 0], bref) < 1e-10
    assert rel_err(spec[:, 0], sref) < 1e-8


def test_explicit_mode_switching(ectx, mini):
    """Entries that need the factored operator refuse the explicit one, low
    precision is refused, and ngw_set_isdf switches back."""
    from nanogw_amd._lib import NgwError
    from nanogw_amd.lanczos import lanczos_polynomial
    from oracle import oracle as O

    m = mini
    nv, nc = m.nv, m.nc
    s = O.sqrt_r(m.e1, nv, nc)
    ectx.set_explicit_c(O.pair_cmtrx(m.psi_intp, nv, nc), m.mmtrx, s)
    poly = lanczos_polynomial(m.e1, nv, 12)
    with pytest.raises(NgwError):
        ectx.calculate_sigma_lanczos([nv - 1], 5, 0.5, 0.01, 10, poly)
    with pytest.raises(NgwError):
        ectx.sigma_x([nv - 1])
    with pytest.raises(NgwError):
        ectx.set_precision("f16")
    with pytest.raises(NgwError):
        ectx.set_explicit_c(O.pair_cmtrx(m.psi_intp, nv, nc), m.mmtrx, -s)
    x = np.random.default_rng(34).normal(size=(nv * nc, 2))
    ye = ectx.matvec_isdf_BLOCK(x)
    ectx.set_isdf(m.psi_intp, m.mmtrx, m.e1, nv)
    ectx.set_precision("f64")
    yf = ectx.matvec_isdf_BLOCK(x)
    assert rel_err(ye, yf) < 1e-11
    sig = ectx.calculate_sigma_lanczos([nv - 1], 5, 0.5, 0.01, 10, poly)
    assert np.all(np.isfinite(sig))
