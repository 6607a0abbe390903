"""The Fortran drop-in (nanogw_amd/fortran/ngw_dropin.F90, NanoGW's
matvec_isdf_BLOCK / lanczos_spectra_isdf_BLOCK, single-vector
matvec_isdf / polynomial_matvec_isdf / lanczos_spectra_isdf, their
UltraLowMem variants and the lowcomsym routines, with the reference names and
argument lists over iso_c_binding) called from a Fortran program, against the
CPU oracle."""
import os
import subprocess

import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu


def _write_inputs(fin, m, blksz, niter, poly, vec, zz, veig):
    """Input file of the check program: sizes, then column-major arrays."""
    with open(fin, "wb") as f:
        np.array([m.n_intp, m.nv, m.nc, blksz, zz.shape[0], niter, len(poly) - 1], np.int32).tofile(f)
        for a in (m.psi_intp, m.mmtrx, m.e1, vec, poly):
            np.asfortranarray(a, np.float64).T.tofile(f)  # column-major bytes
        np.asfortranarray(zz, np.complex128).T.tofile(f)
        veig.tofile(f)


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
    s = O.sqrt_r(m.e1, nv, nc)
    # an eigenvector of H in the check program's reversed pair order: the
    # single-vector Lanczos breaks down at k = 2, S(z) = 1/(z - p(lambda))
    hd = O.matvec_block(m.psi_intp, m.mmtrx, s, np.eye(ncv), nv, nc)
    w, Z = np.linalg.eigh(0.5 * (hd + hd.T))
    veig = Z[::-1, 5].copy()
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    _write_inputs(fin, m, blksz, niter, poly, vec, zz, veig)
    r = subprocess.run([DROPIN_CHECK, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = np.fromfile(fout, np.float64)
    nz = zz.shape[0]
    o1 = ncv * blksz
    o2 = o1 + 2 * nz * blksz
    hvec = raw[:o1].reshape(blksz, ncv).T
    spec = raw[o1:o2].view(np.complex128).reshape(blksz, nz).T
    h1, p1 = raw[o2:o2 + ncv], raw[o2 + ncv:o2 + 2 * ncv]
    o3 = o2 + 2 * ncv + 4 * nz
    spec1 = raw[o2 + 2 * ncv:o2 + 2 * ncv + 2 * nz].view(np.complex128)
    speig = raw[o2 + 2 * ncv + 2 * nz:o3].view(np.complex128)
    hu, pu = raw[o3:o3 + ncv], raw[o3 + ncv:o3 + 2 * ncv]
    specu = raw[o3 + 2 * ncv:o3 + 2 * ncv + 2 * nz].view(np.complex128)
    o4 = o3 + 2 * ncv + 4 * nz
    speigu = raw[o3 + 2 * ncv + 2 * nz:o4].view(np.complex128)
    # lowcomsym blocking of the check program: irrep 1 = valence 0..nv/2-1 and
    # even conduction indices, irrep 2 = the rest; pairs within one irrep,
    # valence outer, conduction inner
    nv1, nc1 = nv // 2, (nc + 1) // 2
    cl = list(range(0, nc, 2)) + list(range(1, nc, 2))
    pairs = [(v, cl[c]) for v in range(nv1) for c in range(nc1)]
    pairs += [(v, cl[c]) for v in range(nv1, nv) for c in range(nc1, nc)]
    nsub = len(pairs)
    hvec_e = raw[-ncv * blksz:].reshape(blksz, ncv).T
    blk = raw[o4:-ncv * blksz - 2 * nsub].reshape(-1)
    hs2 = raw[-ncv * blksz - 2 * nsub:-ncv * blksz].reshape(2, nsub).T
    hs = blk[:2 * nsub].reshape(2, nsub).T
    hst = blk[2 * nsub:4 * nsub].reshape(2, nsub).T
    ps = blk[4 * nsub:6 * nsub].reshape(2, nsub).T
    specs = blk[6 * nsub:6 * nsub + 4 * nz].view(np.complex128).reshape(2, nz).T
    spect = blk[6 * nsub + 4 * nz:].view(np.complex128).reshape(2, nz).T
    ref_h = O.matvec_block(m.psi_intp, m.mmtrx, s, vec, nv, nc)
    ref_s, _, _ = O.lanczos_spectra_block(m.psi_intp, m.mmtrx, s, vec, np.linalg.norm(vec, axis=0), zz, niter,
                                          poly, nv, nc)
    assert rel_err(hvec, ref_h) < 1e-11
    # wfn%e1 changed in place between calls: the drop-in rebuilt its operator
    e1e = m.e1.copy()
    e1e[nv:] += 0.1
    ref_e = O.matvec_block(m.psi_intp, m.mmtrx, O.sqrt_r(e1e, nv, nc), vec, nv, nc)
    assert rel_err(hvec_e, ref_e) < 1e-11
    assert rel_err(hvec_e, ref_h) > 1e-3
    assert rel_err(spec, ref_s) < 1e-6
    # single-vector routines on tmp_Cmtrx over the reversed pair list
    cm = np.asfortranarray(O.pair_cmtrx(m.psi_intp, nv, nc)[:, ::-1])
    sr, v1 = s[::-1].copy(), vec[::-1, 0].copy()
    assert rel_err(h1, O.matvec_explicit(cm, m.mmtrx, sr, v1)) < 1e-11
    assert rel_err(p1, O.polynomial_matvec_explicit(cm, m.mmtrx, sr, v1, poly)) < 1e-10
    ref1, _, _, nk = O.lanczos_spectra_explicit(cm, m.mmtrx, sr, v1, np.linalg.norm(v1), zz[:, 0], niter, poly)
    assert nk == niter
    assert rel_err(spec1, ref1) < 1e-6
    lam = np.polyval(poly, w[5])
    print("single-vector rel err", rel_err(h1, O.matvec_explicit(cm, m.mmtrx, sr, v1)), rel_err(spec1, ref1),
          "breakdown", rel_err(speig, 1.0 / (zz[:, 0] - lam)))
    assert rel_err(speig, 1.0 / (zz[:, 0] - lam)) < 1e-9
    # UltraLowMem variants on the factored operator, NanoGW's pair order
    v = vec[:, 0]
    assert rel_err(hu, O.matvec_block(m.psi_intp, m.mmtrx, s, v, nv, nc)[:, 0]) < 1e-11
    assert rel_err(pu, O.polynomial_matvec_block(m.psi_intp, m.mmtrx, s, v, nv, nc, poly)[:, 0]) < 1e-10
    refu, _, _ = O.lanczos_spectra_block(m.psi_intp, m.mmtrx, s, v, [np.linalg.norm(v)], zz[:, :1], niter, poly,
                                         nv, nc)
    assert rel_err(specu, refu[:, 0]) < 1e-6
    assert rel_err(speigu, 1.0 / (zz[:, 0] - lam)) < 1e-9
    # lowcomsym: the operator on a pair subset is the principal submatrix of
    # the all-pairs operator; Horner and Lanczos against the single-vector
    # restatements on the subset's Cmtrx; the TDA Lanczos runs on H itself
    idx = np.array([v + nv * c for v, c in pairs])
    xs = vec[:nsub, :2]

    def sub_apply(x, tda):
        xf = np.zeros((ncv, x.shape[1]))
        xf[idx] = x
        return O.matvec_block(m.psi_intp, m.mmtrx, s, xf, nv, nc, tda)[idx]

    assert rel_err(hs, sub_apply(xs, False)) < 1e-11
    assert rel_err(hst, sub_apply(xs, True)) < 1e-11
    # PsiC_intp_bl(3, 2) changed in place (conduction state 2, 1-based, is
    # band nv + 1 of Psi): the drop-in rebuilt the blocked operator
    psi2 = m.psi_intp.copy()
    psi2[2, nv + 1] += 0.5
    xf = np.zeros((ncv, 2))
    xf[idx] = xs
    ref2 = O.matvec_block(psi2, m.mmtrx, s, xf, nv, nc)[idx]
    print("lowcomsym after in-place change", rel_err(hs2, ref2), rel_err(hs2, hs))
    assert rel_err(hs2, ref2) < 1e-11 and rel_err(hs2, hs) > 1e-6
    cms = np.asfortranarray(O.pair_cmtrx(m.psi_intp, nv, nc)[:, idx])
    for k in range(2):
        assert rel_err(ps[:, k], O.polynomial_matvec_explicit(cms, m.mmtrx, s[idx], xs[:, k], poly)) < 1e-10
        ref, _, _, nk = O.lanczos_spectra_explicit(cms, m.mmtrx, s[idx], xs[:, k], np.linalg.norm(xs[:, k]),
                                                   zz[:, k], niter, poly)
        assert nk == niter and rel_err(specs[:, k], ref) < 1e-6
        htda = sub_apply(np.eye(nsub), True)
        reft, _, _ = O.lanczos_dense(0.5 * (htda + htda.T), xs[:, k], zz[:, k], niter)
        print("lowcomsym TDA lanczos rel err", rel_err(spect[:, k], reft))
        assert rel_err(spect[:, k], reft) < 1e-6
