"""The host side of the continuous-batching loop (nanogw_amd.batch.serve_stream)
on the CPU, against a recording stand-in for the device batch: FIFO admission
under the HBM budget and the active-request cap, the size policy (a request
whose step costs more than large_factor x the median of the others is held
except on every large_every-th step), completion of every request, release
of its memory, and the Newton call on each finished request.  The device
batch itself (ngw_batch_*) is tested on the GPU in test_gpu_batch.py."""
import types

import numpy as np

from nanogw_amd import batch as B


class FakeBatch:
    """Steps every submitted, unheld request by one Lanczos step per step()."""

    def __init__(self, niter, sizes):
        self.niter, self.sizes = niter, sizes
        self.reqs, self.next_id, self.steps_run = {}, 0, 0
        self.log = []          # (step index, {rid: held?})
        self.live = []         # bytes of the running requests per step, 0 when one runs alone
        self.peak_bytes = 0
        self.live_bytes = 0

    def request_bytes(self, sid, nqp, niter, nen):
        return self.sizes[sid] * nqp

    def submit(self, sid, qp, nen, deltae, ecuts, niter, poly):
        rid = self.next_id
        self.next_id += 1
        need = self.request_bytes(sid, len(qp), niter, nen)
        self.reqs[rid] = {"sid": sid, "nqp": len(qp), "done": 0, "held": False, "need": need}
        self.live_bytes += need
        self.peak_bytes = max(self.peak_bytes, self.live_bytes)
        return rid

    def hold(self, rid, hold=True):
        self.reqs[rid]["held"] = bool(hold)

    def step(self, n, use_graph):
        self.log.append((self.steps_run, {r: q["held"] for r, q in self.reqs.items()}))
        self.live.append(sum(q["need"] for q in self.reqs.values()) if len(self.reqs) > 1 else 0)
        for q in self.reqs.values():
            if not q["held"] and q["done"] < self.niter:
                q["done"] += 1
        self.steps_run += 1

    def result(self, rid, nen, nqp):
        q = self.reqs[rid]
        return np.full((nen, nqp), 0.01 * (rid + 1), complex) if q["done"] >= self.niter else None

    def release(self, rid):
        self.live_bytes -= self.reqs.pop(rid)["need"]

    def stats(self):
        return {"steps_run": self.steps_run}


class FakeCtx:
    def __init__(self):
        self.calls = 0

    def qp_newton(self, sig_ev, deltae_ev, edft, en0, sigx, vxc):
        self.calls += 1
        return {"eqp": np.asarray(edft) + 1.0, "zfac": np.ones(len(edft)), "ierr": np.zeros(len(edft), int)}


def _mol(nv, nmu):
    e1 = np.linspace(-1.0, 1.0, 4 * nv)
    return types.SimpleNamespace(nv=nv, nc=3 * nv, nb=4 * nv, n_intp=nmu, e1=e1,
                                 psi_intp=np.zeros((nmu, 4 * nv)), mmtrx=None)


def _serve(mols, reqs, niter, budget, **kw):
    fb = FakeBatch(niter, sizes=[10 * m.nv for m in mols])
    ctx = FakeCtx()
    recs, stats = B.serve_stream(ctx, mols, reqs, niter, 5, 0.1, 0.01, budget, use_graph=False, batch=fb,
                                 sids=list(range(len(mols))), polys=[np.ones(3)] * len(mols), **kw)
    return recs, stats, fb, ctx


def test_every_request_finishes_under_the_budget_fifo():
    mols = [_mol(4, 40), _mol(6, 60)]
    reqs = [{"t": 0.0, "mol": i % 2, "qp": [0, 1, 2][: 1 + i % 3]} for i in range(7)]
    budget = 150
    recs, stats, fb, ctx = _serve(mols, reqs, niter=3, budget=budget, large_every=0)
    assert all(r is not None and "finish" in r for r in recs)
    assert ctx.calls == len(reqs)
    assert fb.reqs == {} and fb.live_bytes == 0          # every request released
    # several requests share a step only within the footprint budget; one
    # larger than the budget (3 states of the second molecule, 180) runs alone
    assert max(fb.live) <= budget and fb.peak_bytes == 180
    admits = [r["admit"] for r in recs]
    assert admits == sorted(admits)                     # FIFO at equal arrival times
    for r in recs:
        assert np.allclose(r["eqp"], _mol(4 if r["mol"] == 0 else 6, 0).e1[r["qp"]] * B.RYD + 1.0)
    assert stats["lanczos_steps"] == stats["batch_steps"] == fb.steps_run


def test_a_request_larger_than_the_budget_runs_alone():
    mols = [_mol(4, 40), _mol(30, 300)]
    reqs = [{"t": 0.0, "mol": 1, "qp": [0, 1]}, {"t": 0.0, "mol": 0, "qp": [0]}]
    recs, stats, fb, _ = _serve(mols, reqs, niter=2, budget=100, large_every=0)
    assert all("finish" in r for r in recs)
    # the oversized request was admitted alone, the second one after it finished
    assert recs[1]["admit"] >= recs[0]["finish"] - 1e-9


def test_max_active_caps_the_batch():
    mols = [_mol(4, 40)]
    reqs = [{"t": 0.0, "mol": 0, "qp": [0]} for _ in range(5)]
    recs, stats, fb, _ = _serve(mols, reqs, niter=2, budget=10 ** 9, max_active=2, large_every=0)
    assert all("finish" in r for r in recs)
    assert max(sum(1 for h in held.values()) for _, held in fb.log) <= 2


def test_large_request_held_except_every_kth_step():
    """One request whose step costs far more than the others' median is held
    on every step except steps 0, k, 2k, ... while small requests run; once it
    runs alone it steps every time and finishes."""
    small, big = _mol(4, 48), _mol(60, 720)
    mols = [small, big]
    reqs = [{"t": 0.0, "mol": 1, "qp": [0, 1, 2, 3]}] + [{"t": 0.0, "mol": 0, "qp": [0]} for _ in range(3)]
    k, niter = 4, 10
    recs, stats, fb, _ = _serve(mols, reqs, niter=niter, budget=10 ** 9, large_every=k, large_factor=50.0)
    assert all("finish" in r for r in recs)
    big_rid = 0                                          # submitted first (FIFO)
    for s, held in fb.log:
        if len(held) > 1 and big_rid in held:            # small requests still running
            assert held[big_rid] == (s % k != 0), (s, held)
        elif big_rid in held:
            assert not held[big_rid]                     # alone: never held
    assert stats["steps_with_large_held"] > 0
    # the small requests finish first, after niter steps each
    assert max(r["finish"] for r in recs[1:]) <= recs[0]["finish"]
    # the large one progressed on exactly niter steps (every k-th step, then alone)
    assert sum(1 for _, h in fb.log if big_rid in h and not h[big_rid]) == niter


def test_policy_off_steps_everything():
    small, big = _mol(4, 48), _mol(60, 720)
    reqs = [{"t": 0.0, "mol": 1, "qp": [0]}, {"t": 0.0, "mol": 0, "qp": [0]}]
    recs, stats, fb, _ = _serve([small, big], [dict(r, mol=1 - r["mol"]) for r in reqs], niter=3,
                                budget=10 ** 9, large_every=0)
    assert stats["steps_with_large_held"] == 0
    assert not any(h for _, held in fb.log for h in held.values())
    assert fb.steps_run == 3
