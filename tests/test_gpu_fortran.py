"""The Fortran drop-in (nanogw_amd/fortran/ngw_dropin.F90, NanoGW's
matvec_isdf_BLOCK / lanczos_spectra_isdf_BLOCK names and argument lists over
iso_c_binding) called from a Fortran program, against the CPU oracle."""
import os
import subprocess

import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu


def test_fortran_dropin_matches_oracle(gpu, mini, tmp_path):
    from nanogw_amd.build import DROPIN_CHECK
    from nanogw_amd.lanczos import lanczos_polynomial
    from oracle import oracle as O

    if not os.path.exists(DROPIN_CHECK):
        pytest.skip("Fortran drop-in check program not built (no amdflang)")
    m = mini
    nmu, nv, nc = m.n_intp, m.nv, m.nc
    ncv, blksz, niter = nv * nc, 3, 20
    poly = lanczos_polynomial(m.e1, nv, 12)
    rng = np.random.default_rng(11)
    vec = rng.normal(size=(ncv, blksz))
    zz = rng.uniform(-1, 3, size=(17, blksz)) + 0.03j
    zz[-1, :] = 0.0
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as f:
        np.array([nmu, nv, nc, blksz, zz.shape[0], niter, len(poly) - 1], np.int32).tofile(f)
        for a in (m.psi_intp, m.mmtrx, m.e1, vec, poly):
            np.asfortranarray(a, np.float64).T.tofile(f)  # column-major bytes
        np.asfortranarray(zz, np.complex128).T.tofile(f)
    r = subprocess.run([DROPIN_CHECK, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = np.fromfile(fout, np.float64)
    hvec = raw[: ncv * blksz].reshape(blksz, ncv).T
    spec = raw[ncv * blksz:].view(np.complex128).reshape(blksz, zz.shape[0]).T
    s = O.sqrt_r(m.e1, nv, nc)
    ref_h = O.matvec_block(m.psi_intp, m.mmtrx, s, vec, nv, nc)
    ref_s, _, _ = O.lanczos_spectra_block(m.psi_intp, m.mmtrx, s, vec, np.linalg.norm(vec, axis=0), zz, niter,
                                          poly, nv, nc)
    assert rel_err(hvec, ref_h) < 1e-11
    assert rel_err(spec, ref_s) < 1e-6
