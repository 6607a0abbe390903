"""The three large workloads on the GPU against the FP64 oracle.

R = 13 molecule (h = 0.33, 256k grid points, nv = 100, nc = 1500,
Nmu = 5000; the mixed-precision headline workload, built exactly as the
benchmark builds it: device V build, zeta ridge 1e-6):
  * the plain FP64 GPU path (lanczos_spectra_isdf_BLOCK) against the
    oracle's lanczos_spectra_block on a sample of the runs a QP solve is made
    of: the HOMO's P vectors for 8 Green's-function bands, 2 Lanczos steps of
    npoly = 12 (a, b to 1e-10 relative, spectra 1e-8);
  * the FP16 and BF16x3 operators per column against the oracle's FP64
    operator on those 8 vectors.  Bars 2e-2 (FP16) and 1e-3 (BF16x3) of the
    column's largest entry: the operand rounding (2^-11, about 2^-16) enters
    C^T M C, whose terms cancel by about an order of magnitude in these P
    vectors over Nmu = 5000 and nc = 1500 accumulations (measured 4e-3..9e-3
    and 7e-5..2.5e-4).  The QP energies, the quantity the 1 meV bar is
    stated for, differ from FP64 by 0.145 meV at most (benchmark line).
  Together with the FP64-vs-FP16 QP comparison of the benchmark this gives
  the chain FP16 -> FP64 GPU (QP energies) -> oracle (these runs).

Batched molecules (R ~ U[6, 11], h = 0.4, nv ~ U{8..60}, nc = 10 nv,
Nmu = 12 nv, 16 QP states each; reduced niter):
  * three molecules of the seeded batch spanning nv (smallest, median, and
    nv = 60) in one FP16 batch with hipGraph replay: each result bitwise
    equal to the same request run alone;
  * QP energies against the FP64 batch (nv = 9 and 38; nv = 60 is
    rounding-chaotic at 16 steps in every precision): BF16x3 within 1 meV for
    all 16 states; FP16 printed (it moves the valence states of the nv = 38
    molecule by up to tens of meV);
  * on the nv = 38 molecule, full reorthogonalisation brings FP16 within
    1 meV of FP64 on the same 16 states, so the FP16 valence error is the
    recurrence's loss of orthogonality driven by the operator's rounding;
  * the FP64 batch equal to the single-context path, and the single-context
    path equal to the oracle on a band sample (8 Green's-function bands),
    within 1e-8 or 3x the oracle's own rounding envelope (oracle runs on
    inputs 1e-15, -1e-15 and 3e-15 apart), whichever is larger.
  Without reorthogonalisation the reference recurrence is rounding-sensitive
  on these molecules: for the nv = 38 one the oracle's own full-grid Sigma_c
  envelope is 5e-11 up to 16 steps and 0.44 at 32, for nv = 60 it is 7e-5 at
  8 steps and 4e-2 at 12.  test_reorth_stabilises_batched_molecule shows the
  same molecule with full reorthogonalisation.

Served stream (Poisson arrivals at 20 req/s from a catalog of batched-workload
molecules, U{4..16} QP states each): every request finishes, and each result
is bitwise equal to the request run alone.
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

RYD = 13.60569193


def _log(*a):
    print(*a, flush=True)


# ---------------------------------------------------------------- R = 13, nv = 100
@pytest.fixture(scope="module")
def big_mol(gpu, ngw_ctx):
    from nanogw_amd import synth

    R, h, nv, nc, nmu, _, _ = synth.CONFIGS["c3"]
    _log(f"building R={R} h={h} nv={nv} nc={nc} Nmu={nmu} molecule")
    return synth.make_molecule(R, h, nv, nc, nmu, synth.config_seed(3, 0), cvt_iter=60, ridge=synth.BENCH_RIDGE,
                               kernel_ctx=ngw_ctx, device=gpu, keep_zeta=False)


@pytest.fixture(scope="module")
def big_pvec(big_mol):
    from oracle import oracle as O

    m = big_mol
    s = O.sqrt_r(m.e1, m.nv, m.nc)
    qp, n0, blk = m.nv - 1, m.nv - 4, 8
    p, pn2 = O.pvectors(m.psi_intp, m.mmtrx, s, qp, n0, blk, m.nv, m.nc)
    return s, p, pn2


def test_big_molecule_matches_benchmark_fingerprint(big_mol):
    """The molecule is the one whose FP64 QP energies the benchmark compares
    against (same seed, grid, points and kernel)."""
    import json
    import os

    from conftest import GOLDEN

    ref = json.load(open(os.path.join(GOLDEN, "c3_fp64_qp.json")))["fingerprint"]
    m = big_mol
    fp = {"sum_psi2": float(np.sum(m.psi_intp ** 2)), "trace_m": float(np.trace(m.mmtrx)),
          "sum_m2": float(np.sum(m.mmtrx ** 2)), "e1_sum": float(np.sum(m.e1))}
    for k in ref:
        assert abs(fp[k] - ref[k]) <= 1e-9 * max(1.0, abs(ref[k])), (k, fp[k], ref[k])


def test_big_fp64_lanczos_vs_oracle(ngw_ctx, big_mol, big_pvec):
    from nanogw_amd.lanczos import lanczos_polynomial
    from oracle import oracle as O

    m = big_mol
    s, p, pn2 = big_pvec
    poly = lanczos_polynomial(m.e1, m.nv, 12)
    niter, blk = 2, p.shape[1]
    e = m.e1[m.nv - 1]
    zz = np.array([[complex(e + d, 0.02) for _ in range(blk)] for d in (-0.5, -0.1, 0.0, 0.2, 0.7)])
    vn = np.sqrt(pn2)
    ngw_ctx.set_isdf(m.psi_intp, m.mmtrx, m.e1, m.nv)
    ngw_ctx.set_precision("f64")
    spec, a, b = ngw_ctx.lanczos_spectra_isdf_BLOCK(p, vn, zz, niter, poly)
    _log("oracle Lanczos, 8 bands x 2 steps")
    spec_ref, a_ref, b_ref = O.lanczos_spectra_block(m.psi_intp, m.mmtrx, s, p, vn, zz, niter, poly, m.nv, m.nc,
                                                     rebuild=True)
    print("R=13 FP64 Lanczos a, b, spectra rel err", rel_err(a, a_ref), rel_err(b, b_ref), rel_err(spec, spec_ref))
    assert rel_err(a, a_ref) < 1e-10 and rel_err(b, b_ref) < 1e-10
    assert rel_err(spec, spec_ref) < 1e-8


@pytest.mark.parametrize("prec,tol", [("f64", 1e-12), ("bf16x3", 1e-3), ("f16", 2e-2)])
def test_big_operator_per_column_vs_oracle(ngw_ctx, big_mol, big_pvec, prec, tol):
    from oracle import oracle as O

    m = big_mol
    s, p, _ = big_pvec
    ref = O.matvec_block(m.psi_intp, m.mmtrx, s, p, m.nv, m.nc)
    ngw_ctx.set_isdf(m.psi_intp, m.mmtrx, m.e1, m.nv)
    ngw_ctx.set_precision(prec)
    try:
        y = ngw_ctx.matvec_isdf_BLOCK(p)
    finally:
        ngw_ctx.set_precision("f64")
    errs = [rel_err(y[:, j], ref[:, j]) for j in range(p.shape[1])]
    print("R=13 operator", prec, "per-column rel err", ["%.2e" % e for e in errs])
    assert max(errs) < tol


def test_big_one_vector_operator_vs_oracle(ngw_ctx, big_mol, big_pvec):
    """The one-vector FP64 operator (the single-vector routines' call shape:
    split-sum MFMA kernels for Psi_c X^T and Psi_c^T W, column-split GEMV for
    M t) on single P vectors against the oracle, full and TDA."""
    from oracle import oracle as O

    m = big_mol
    s, p, _ = big_pvec
    ngw_ctx.set_isdf(m.psi_intp, m.mmtrx, m.e1, m.nv)
    ngw_ctx.set_precision("f64")
    for tda in (False, True):
        for j in (0, 5):
            y = ngw_ctx.matvec_isdf_BLOCK(p[:, j:j + 1], tamm_d=tda)
            ref = O.matvec_block(m.psi_intp, m.mmtrx, s, p[:, j:j + 1], m.nv, m.nc, tda=tda)
            print(f"R=13 one-vector operator (tda={tda}) column {j}: rel err {rel_err(y, ref):.2e}")
            assert rel_err(y, ref) < 1e-12


# ---------------------------------------------------------------- batched molecules
NITER_BATCH = 16


@pytest.fixture(scope="module")
def batch_mols(gpu, ngw_ctx):
    """Smallest-nv and median-nv molecules of the seeded 64-molecule batch and
    one with nv = 60 (the distribution's upper end)."""
    from nanogw_amd import synth

    specs = synth.c4_specs(64)
    by_nv = sorted(specs, key=lambda sp: sp["nv"])
    pick = [by_nv[0], by_nv[len(by_nv) // 2]]
    top = [sp for sp in specs if sp["nv"] == 60]
    if top:
        pick.append(top[0])
    else:  # the seeded draw has no nv = 60: same R / seed as the largest, nv set to 60
        sp = dict(by_nv[-1])
        sp.update(nv=60, nc=600, n_intp=720)
        pick.append(sp)
    mols = []
    for sp in pick:
        _log(f"building batched molecule R={sp['radius']:.2f} nv={sp['nv']} nc={sp['nc']} Nmu={sp['n_intp']}")
        mols.append(synth.make_molecule(sp["radius"], sp["h"], sp["nv"], sp["nc"], sp["n_intp"], sp["seed"],
                                        cvt_iter=20, kernel_ctx=ngw_ctx, device=gpu, keep_zeta=False,
                                        ridge=synth.BENCH_RIDGE))
    assert [m.nv for m in mols][-1] == 60
    return mols


def _run_alone(b, sid, qp, nen, deltae, ecuts, niter, poly, use_graph=True):
    rid = b.submit(sid, qp, nen, deltae, ecuts, niter, poly)
    while b.result(rid, nen, len(qp)) is None:
        b.step(1, use_graph)
    out = b.result(rid, nen, len(qp))
    b.release(rid)
    return out


def test_batched_f16_alone_equals_mixed_and_qp_vs_f64(ngw_ctx, batch_mols):
    """FP16 batch: bitwise equal alone and in the mix.  QP energies against
    the FP64 batch (nv = 9 and 38): BF16x3 within 1 meV for all 16 states.  FP16 is printed,
    not bounded: on the nv = 38 molecule at 16 steps it moves the valence
    states by 6-39 meV (HOMO 23 meV) while the conduction states stay within
    0.23 meV and BF16x3 within 0.03 meV everywhere; the batched and served
    workloads therefore state their QP accuracy for BF16x3."""
    from nanogw_amd import synth
    from nanogw_amd.batch import GwBatch
    from nanogw_amd.lanczos import lanczos_polynomial

    mols = batch_mols
    sd = synth.sigma_defaults(8)
    nen, deltae, ecuts = sd["nen"], sd["deltae"], sd["ecuts"]
    polys = [lanczos_polynomial(m.e1, m.nv, 12) for m in mols]
    qps = [synth.qp_window(m.nv, m.nb, 16) for m in mols]
    sig = {}
    for prec in ("f64", "bf16x3", "f16"):
        ngw_ctx.set_precision(prec)
        ngw_ctx.set_tile_config(0, 0)
        b = GwBatch(ngw_ctx)
        try:
            sids = [b.add_system(m.psi_intp, m.mmtrx, m.e1, m.nv) for m in mols]
            rids = [b.submit(sids[k], qps[k], nen, deltae, ecuts, NITER_BATCH, polys[k]) for k in range(len(mols))]
            while b.active():
                b.step(1, True)
            mixed = [b.result(r, nen, len(qps[k])) for k, r in enumerate(rids)]
            for r in rids:
                b.release(r)
            if prec == "f16":
                for k in range(len(mols)):
                    alone = _run_alone(b, sids[k], qps[k], nen, deltae, ecuts, NITER_BATCH, polys[k])
                    assert np.array_equal(alone, mixed[k]), f"nv={mols[k].nv}: batched result differs from alone"
            sig[prec] = mixed
        finally:
            b.close()
            ngw_ctx.set_precision("f64")
    # QP energies of nv = 9 and nv = 38 only: the nv = 60 molecule's
    # recurrence is rounding-chaotic past 8 steps already in FP64 (its
    # perturbation envelope is 7e-5 at 8 steps, 4e-2 at 12), so its
    # precisions cannot be compared at 16 steps
    for k, m in enumerate(mols[:2]):
        e = m.e1[qps[k]] * RYD
        sx = synth.exchange_diag(m.psi_intp, m.mmtrx, m.nv, qps[k]) * RYD
        q = {p: ngw_ctx.qp_newton(sig[p][k] * RYD, deltae * RYD, e, e, sx, np.zeros(len(e))) for p in sig}
        assert all(not np.any(q[p]["ierr"]) for p in q)
        d16 = 1000.0 * (q["f16"]["eqp"] - q["f64"]["eqp"])
        d3 = 1000.0 * (q["bf16x3"]["eqp"] - q["f64"]["eqp"])
        print(f"batched nv={m.nv}, {NITER_BATCH} steps, E_qp - E_qp(FP64) in meV per state {qps[k][0]}..{qps[k][-1]}:")
        print("  FP16   " + " ".join(f"{v:8.3f}" for v in d16))
        print("  BF16x3 " + " ".join(f"{v:8.3f}" for v in d3))
        assert np.abs(d3).max() < 1.0


def test_batched_f64_vs_single_and_oracle_band_sample(ngw_ctx, batch_mols):
    from nanogw_amd import synth
    from nanogw_amd.batch import GwBatch
    from nanogw_amd.lanczos import lanczos_polynomial
    from oracle import oracle as O

    sd = synth.sigma_defaults(8)
    nen, deltae, ecuts = 41, sd["deltae"], sd["ecuts"]
    ngw_ctx.set_precision("f64")
    for m in batch_mols:
        poly = lanczos_polynomial(m.e1, m.nv, 12)
        qp = [m.nv - 1, m.nv]
        b = GwBatch(ngw_ctx)
        try:
            sid = b.add_system(m.psi_intp, m.mmtrx, m.e1, m.nv)
            sb = _run_alone(b, sid, qp, nen, deltae, ecuts, NITER_BATCH, poly)
        finally:
            b.close()
        ngw_ctx.set_isdf(m.psi_intp, m.mmtrx, m.e1, m.nv)
        ss = ngw_ctx.calculate_sigma_lanczos(qp, nen, deltae, ecuts, NITER_BATCH, poly)
        print(f"nv={m.nv}: FP64 batch vs single-context Sigma_c rel err {rel_err(sb, ss):.2e}")
        assert rel_err(sb, ss) < 1e-10
        lo, hi = m.nv - 4, m.nv + 4
        sg = ngw_ctx.calculate_sigma_lanczos(qp, nen, deltae, ecuts, NITER_BATCH, poly, bands=(lo, hi))
        ref = O.sigma_lanczos(m.psi_intp, m.mmtrx, m.e1, m.nv, m.nc, qp, nen, deltae, ecuts, NITER_BATCH, poly, 8,
                              bands=(lo, hi))
        env = max(rel_err(O.sigma_lanczos(m.psi_intp * (1.0 + f), m.mmtrx, m.e1, m.nv, m.nc, qp, nen, deltae, ecuts,
                                          NITER_BATCH, poly, 8, bands=(lo, hi)), ref) for f in (1e-15, -1e-15, 3e-15))
        print(f"nv={m.nv}: bands {lo}..{hi - 1} Sigma_c rel err vs oracle {rel_err(sg, ref):.2e} "
              f"(oracle rounding envelope {env:.2e})")
        assert rel_err(sg, ref) < max(1e-8, 3.0 * env)


def test_reorth_stabilises_batched_molecule(ngw_ctx, batch_mols):
    """The nv = 38 molecule at 48 Lanczos steps, where the plain recurrence is
    rounding-chaotic: with full reorthogonalisation (FP64 recurrence) the QP
    energies of FP64 and of Psi x (1 + 1e-15) agree to 1e-6 meV, and FP16
    operator products stay within 1 meV of FP64."""
    from nanogw_amd import synth
    from nanogw_amd.lanczos import lanczos_polynomial

    m = batch_mols[1]
    sd = synth.sigma_defaults(8)
    nen, deltae, ecuts = sd["nen"], sd["deltae"], sd["ecuts"]
    poly = lanczos_polynomial(m.e1, m.nv, 12)
    qp = [m.nv - 1, m.nv]
    e = m.e1[qp] * RYD
    sx = synth.exchange_diag(m.psi_intp, m.mmtrx, m.nv, qp) * RYD
    res = {}
    for key, prec, scale, reorth in (("p64", "f64", 1.0, False), ("p64p", "f64", 1.0 + 1e-15, False),
                                     ("r64", "f64", 1.0, True), ("r64p", "f64", 1.0 + 1e-15, True),
                                     ("r16", "f16", 1.0, True)):
        ngw_ctx.set_isdf(m.psi_intp * scale, m.mmtrx, m.e1, m.nv)
        ngw_ctx.set_precision(prec)
        ngw_ctx.set_tile_config(0, 0)
        ngw_ctx.set_reorthogonalization(reorth)
        try:
            sig = ngw_ctx.calculate_sigma_lanczos(qp, nen, deltae, ecuts, 48, poly)
        finally:
            ngw_ctx.set_reorthogonalization(False)
            ngw_ctx.set_precision("f64")
        res[key] = ngw_ctx.qp_newton(sig * RYD, deltae * RYD, e, e, sx, np.zeros(len(e)))
    d = lambda a, b: 1000.0 * np.abs(res[a]["eqp"] - res[b]["eqp"]).max()
    print(f"nv={m.nv}, 48 steps: plain FP64 vs perturbed {d('p64', 'p64p'):.4f} meV; reorthogonalised FP64 vs "
          f"perturbed {d('r64', 'r64p'):.2e} meV, FP16 vs FP64 {d('r16', 'r64'):.4f} meV")
    assert all(not np.any(res[k]["ierr"]) for k in ("r64", "r64p", "r16"))
    assert d("r64", "r64p") < 1e-6
    assert d("r16", "r64") < 1.0

    # The batched workload's 16-state window at 16 steps: FP64's recurrence is
    # still stable there (perturbation envelope 5e-11), FP16's operator
    # rounding is not absorbed by it.  Printed: FP16 plain and reorthogonalised
    # against plain FP64 (measured: plain FP16 moves the eight valence states
    # by 6.8-38.9 meV, reorthogonalised FP16 stays within 0.13 meV, and
    # reorthogonalised FP64 equals plain FP64 to 1e-3 meV); bounded: FP16
    # reorthogonalised within 1 meV.
    qw = synth.qp_window(m.nv, m.nb, 16)
    e = m.e1[qw] * RYD
    sx = synth.exchange_diag(m.psi_intp, m.mmtrx, m.nv, qw) * RYD
    win = {}
    for key, prec, reorth in (("p64", "f64", False), ("p16", "f16", False), ("r16", "f16", True),
                              ("r64", "f64", True)):
        ngw_ctx.set_isdf(m.psi_intp, m.mmtrx, m.e1, m.nv)
        ngw_ctx.set_precision(prec)
        ngw_ctx.set_tile_config(0, 0)
        ngw_ctx.set_reorthogonalization(reorth)
        try:
            sig = ngw_ctx.calculate_sigma_lanczos(qw, nen, deltae, ecuts, 16, poly)
        finally:
            ngw_ctx.set_reorthogonalization(False)
            ngw_ctx.set_precision("f64")
        win[key] = ngw_ctx.qp_newton(sig * RYD, deltae * RYD, e, e, sx, np.zeros(len(e)))
        assert not np.any(win[key]["ierr"])
    for key in ("p16", "r16", "r64"):
        dv = 1000.0 * (win[key]["eqp"] - win["p64"]["eqp"])
        print(f"  16 steps, states {qw[0]}..{qw[-1]}, {key} - plain FP64 (meV): " + " ".join(f"{v:.3f}" for v in dv))
    assert 1000.0 * np.abs(win["r16"]["eqp"] - win["r64"]["eqp"]).max() < 1.0


# ---------------------------------------------------------------- served stream
def test_served_stream_all_finish_and_equal_alone(gpu, ngw_ctx):
    from nanogw_amd import synth
    from nanogw_amd.batch import GwBatch, serve_stream
    from nanogw_amd.lanczos import lanczos_polynomial

    reqs_raw, specs = synth.c5_stream(0.6, rate=20.0, large_frac=0.0, n_small=5)
    assert len(reqs_raw) >= 5
    used = sorted({r["mol"] for r in reqs_raw})
    mols = []
    for i in used:
        sp = specs[i]
        _log(f"building stream molecule nv={sp['nv']} nc={sp['nc']} Nmu={sp['n_intp']}")
        mols.append(synth.make_molecule(sp["radius"], sp["h"], sp["nv"], sp["nc"], sp["n_intp"], sp["seed"],
                                        cvt_iter=20, kernel_ctx=ngw_ctx, device=gpu, keep_zeta=False,
                                        ridge=synth.BENCH_RIDGE))
    index = {i: k for k, i in enumerate(used)}
    reqs = [{"t": r["t"], "mol": index[r["mol"]], "qp": synth.qp_window(mols[index[r["mol"]]].nv,
                                                                        mols[index[r["mol"]]].nb, r["nqp"])}
            for r in reqs_raw]
    sd = synth.sigma_defaults(8)
    nen, deltae, ecuts = 61, sd["deltae"], sd["ecuts"]
    niter = 16
    ngw_ctx.set_precision("f16")
    ngw_ctx.set_tile_config(0, 0)
    b = GwBatch(ngw_ctx)
    try:
        sids = [b.add_system(m.psi_intp, m.mmtrx, m.e1, m.nv) for m in mols]
        polys = [lanczos_polynomial(m.e1, m.nv, 12) for m in mols]
        recs, st = serve_stream(ngw_ctx, mols, reqs, niter, nen, deltae, ecuts, 40 << 30, batch=b, sids=sids,
                                polys=polys)
        assert all(r is not None and "finish" in r for r in recs), "a request did not finish"
        print(f"stream: {len(reqs)} requests, {st['batch_steps']} batched steps")
        for i, r in enumerate(recs):
            alone = _run_alone(b, sids[r["mol"]], r["qp"], nen, deltae, ecuts, niter, polys[r["mol"]])
            assert np.array_equal(alone, r["sigma"]), f"request {i}: served result differs from alone"
            assert not np.any(r["ierr"])
    finally:
        b.close()
        ngw_ctx.set_precision("f64")
