"""FP64 GPU path against the CPU oracle at the config-2 shapes (R = 10,
h = 0.35, nv = 40, nc = 400, Nmu = 2000; the plain-FP64 configuration).

A full c2 QP solve (440 band runs x 300 steps) is out of reach of the CPU
oracle in a test, so a sample is compared: the 50-step Lanczos recurrence
and continued fraction of one QP state (the HOMO) against one block of 8
Green's-function bands, started from the oracle's P vectors.  These are the
same runs a QP solve is made of, at the real operator shapes (Sigma_c and
the P vectors as a whole are compared at config 1).  Tolerances: Lanczos
coefficients to round-off (1e-10 relative), spectra 1e-8.
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c2_mol(gpu, ngw_ctx):
    from nanogw_amd import synth

    R, h, nv, nc, nmu, _, _ = synth.CONFIGS["c2"]
    return synth.make_molecule(R, h, nv, nc, nmu, synth.config_seed(2, 0), cvt_iter=40, ridge=synth.BENCH_RIDGE,
                               kernel_ctx=ngw_ctx, device=gpu, keep_zeta=False)


def test_c2_pvectors_lanczos_vs_oracle(ngw_ctx, c2_mol):
    from nanogw_amd.lanczos import lanczos_polynomial
    from oracle import oracle as O

    m = c2_mol
    nv, nc = m.nv, m.nc
    poly = lanczos_polynomial(m.e1, nv, 12)
    s = O.sqrt_r(m.e1, nv, nc)
    qp, n0, blk, niter = nv - 1, nv - 4, 8, 50
    p_ref, pn2_ref = O.pvectors(m.psi_intp, m.mmtrx, s, qp, n0, blk, nv, nc)
    ngw_ctx.set_isdf(m.psi_intp, m.mmtrx, m.e1, nv)
    ngw_ctx.set_precision("f64")
    vn = np.sqrt(pn2_ref)
    e = m.e1[qp]
    zz = np.array([[complex(e + d, 0.02) for _ in range(blk)] for d in (-0.5, -0.1, 0.0, 0.2, 0.7)])
    spec_ref, a_ref, b_ref = O.lanczos_spectra_block(m.psi_intp, m.mmtrx, s, p_ref, vn, zz, niter, poly, nv, nc)
    spec, a, b = ngw_ctx.lanczos_spectra_isdf_BLOCK(p_ref, vn, zz, niter, poly)
    print("c2 Lanczos a, b, spectra rel err", rel_err(a, a_ref), rel_err(b, b_ref), rel_err(spec, spec_ref))
    assert rel_err(a, a_ref) < 1e-10 and rel_err(b, b_ref) < 1e-10
    assert rel_err(spec, spec_ref) < 1e-8
