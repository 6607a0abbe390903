"""Batched / served engine: heterogeneous molecules in grouped launches,
hipGraph replay, continuous admission -- each result equals the single
request run alone (bitwise: fixed tiles, per-request tiles, no atomics) and
the FP64 oracle within the parity bars."""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mols():
    from nanogw_amd import synth

    return [synth.make_molecule(4.5, 0.5, 4, 12, 40, seed=12345),
            synth.make_molecule(4.0, 0.5, 3, 9, 30, seed=777),
            synth.make_molecule(5.0, 0.55, 5, 20, 56, seed=4242, ridge=1e-8)]


def test_held_request_equals_alone(ngw_ctx, mols):
    """A request held out of the steps for a while (ngw_batch_hold) and then
    resumed gives the same Sigma_c, bit for bit, as run alone; a held request
    does not step and the others run past it."""
    from nanogw_amd.batch import GwBatch
    from nanogw_amd.lanczos import lanczos_polynomial

    niter, nen, deltae, ecuts = 20, 21, 4.0, 0.02
    ngw_ctx.set_precision("f64")
    b = GwBatch(ngw_ctx)
    try:
        sids = [b.add_system(m.psi_intp, m.mmtrx, m.e1, m.nv) for m in mols[:2]]
        polys = [lanczos_polynomial(m.e1, m.nv, 12) for m in mols[:2]]
        qps = [[m.nv - 1, m.nv] for m in mols[:2]]
        alone = []
        for k in range(2):
            rid = b.submit(sids[k], qps[k], nen, deltae, ecuts, niter, polys[k])
            while b.result(rid, nen, 2) is None:
                b.step(1, True)
            alone.append(b.result(rid, nen, 2))
            b.release(rid)
        r0 = b.submit(sids[0], qps[0], nen, deltae, ecuts, niter, polys[0])
        r1 = b.submit(sids[1], qps[1], nen, deltae, ecuts, niter, polys[1])
        b.step(3, True)
        b.hold(r1, True)
        b.step(niter, True)          # r0 finishes while r1 waits
        assert b.result(r0, nen, 2) is not None and b.result(r1, nen, 2) is None
        b.hold(r1, False)
        while b.result(r1, nen, 2) is None:
            b.step(1, True)
        assert np.array_equal(b.result(r0, nen, 2), alone[0])
        assert np.array_equal(b.result(r1, nen, 2), alone[1])
    finally:
        b.close()


@pytest.mark.parametrize("use_graph", [False, True])
def test_batch_matches_oracle_and_alone(ngw_ctx, mols, use_graph):
    from nanogw_amd.batch import GwBatch
    from nanogw_amd.lanczos import lanczos_polynomial
    from oracle import oracle as O

    niter, nen, deltae, ecuts = 25, 31, 4.0, 0.02
    ngw_ctx.set_precision("f64")
    b = GwBatch(ngw_ctx)
    sids = [b.add_system(m.psi_intp, m.mmtrx, m.e1, m.nv) for m in mols]
    polys = [lanczos_polynomial(m.e1, m.nv, 12) for m in mols]
    qps = [[m.nv - 1, m.nv] for m in mols]
    # alone
    alone = []
    for k, m in enumerate(mols):
        rid = b.submit(sids[k], qps[k], nen, deltae, ecuts, niter, polys[k])
        while b.result(rid, nen, 2) is None:
            b.step(1, use_graph)
        alone.append(b.result(rid, nen, 2))
        b.release(rid)
    # together, with staggered joins (continuous batching)
    rids = [b.submit(sids[0], qps[0], nen, deltae, ecuts, niter, polys[0])]
    b.step(7, use_graph)
    rids.append(b.submit(sids[1], qps[1], nen, deltae, ecuts, niter, polys[1]))
    b.step(3, use_graph)
    rids.append(b.submit(sids[2], qps[2], nen, deltae, ecuts, niter, polys[2]))
    while b.active():
        b.step(1, use_graph)
    for k, m in enumerate(mols):
        mixed = b.result(rids[k], nen, 2)
        assert np.array_equal(mixed, alone[k]), f"molecule {k}: mixed result differs from alone"
        ref = O.sigma_lanczos(m.psi_intp, m.mmtrx, m.e1, m.nv, m.nc, qps[k], nen, deltae, ecuts, niter, polys[k], 7)
        assert rel_err(mixed, ref) < 1e-6, k
    st = b.stats()
    if use_graph:
        assert st["graph_captures"] >= 1
    b.close()


def test_batch_mixed_precision_same_alone_and_mixed(ngw_ctx, mols):
    from nanogw_amd.batch import GwBatch
    from nanogw_amd.lanczos import lanczos_polynomial

    niter, nen = 20, 21
    ngw_ctx.set_precision("bf16x3")
    try:
        b = GwBatch(ngw_ctx)
        sids = [b.add_system(m.psi_intp, m.mmtrx, m.e1, m.nv) for m in mols]
        polys = [lanczos_polynomial(m.e1, m.nv, 12) for m in mols]
        r0 = b.submit(sids[2], [mols[2].nv], nen, 4.0, 0.02, niter, polys[2])
        while b.active():
            b.step(1, True)
        alone = b.result(r0, nen, 1)
        rids = [b.submit(sids[k], [mols[k].nv], nen, 4.0, 0.02, niter, polys[k]) for k in range(3)]
        while b.active():
            b.step(1, True)
        assert np.array_equal(b.result(rids[2], nen, 1), alone)
        b.close()
    finally:
        ngw_ctx.set_precision("f64")


def test_serve_stream_admission_and_latency(ngw_ctx, mols):
    """Requests arriving while others run are admitted as memory frees up;
    results equal the single-request runs."""
    from nanogw_amd import synth
    from nanogw_amd.batch import serve_stream

    ngw_ctx.set_precision("f64")
    niter, nen, deltae, ecuts = 15, 21, 4.0, 0.02
    reqs = [{"t": 0.00, "mol": 0, "qp": [3, 4]}, {"t": 0.001, "mol": 1, "qp": [2]},
            {"t": 0.002, "mol": 2, "qp": [4, 5]}, {"t": 0.003, "mol": 0, "qp": [3, 4]}]
    sigx = [lambda qp, m=m: synth.exchange_diag(m.psi_intp, m.mmtrx, m.nv, qp) for m in mols]
    # small budget forces queueing
    recs, stats = serve_stream(ngw_ctx, mols, reqs, niter, nen, deltae, ecuts, hbm_budget_bytes=1, sigx=sigx)
    assert all(r is not None and "finish" in r for r in recs)
    assert np.array_equal(recs[0]["sigma"], recs[3]["sigma"])
    assert all(np.all(r["ierr"] == 0) for r in recs)
    recs2, _ = serve_stream(ngw_ctx, mols, reqs, niter, nen, deltae, ecuts, hbm_budget_bytes=1 << 40, sigx=sigx)
    for a, b in zip(recs, recs2):
        assert np.array_equal(a["sigma"], b["sigma"])
