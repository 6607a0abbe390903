"""Batched and served GW on top of libngw's batch engine (ngw_batch_*).

``GwBatch`` keeps several molecules' ISDF operators resident on one GPU and
advances many GW requests (molecule + QP states) together: every Lanczos step
of every active request is one set of grouped launches, captured in a
hipGraph.  ``serve_stream`` is the request loop: requests arrive on a
schedule, are admitted when their HBM footprint fits, join the running batch
at the next step boundary, and leave it when their niter steps are done.
"""
from __future__ import annotations

import ctypes as C
import heapq
import time

import numpy as np

from . import _lib
from .lanczos import IsdfContext, _f64, _p, lanczos_polynomial

RYD = 13.60569193


class GwBatch:
    """Resident operators + continuously batched Lanczos requests."""

    def __init__(self, ctx: IsdfContext):
        self.ctx = ctx
        self._lib = ctx._lib
        h = C.c_void_p()
        _lib.check(self._lib.ngw_batch_create(ctx.handle, C.byref(h)), "ngw_batch_create")
        self._h = h
        self.systems = {}

    def close(self):
        if getattr(self, "_h", None):
            self._lib.ngw_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_system(self, psi_intp, mmtrx, e1, nval) -> int:
        psi_intp, mmtrx, e1 = _f64(psi_intp), _f64(mmtrx), _f64(e1)
        n_intp, nband = psi_intp.shape
        sid = C.c_int()
        _lib.check(self._lib.ngw_batch_add_system(self._h, n_intp, nband, int(nval), _p(psi_intp), n_intp,
                                                  _p(mmtrx), mmtrx.shape[0], _p(e1), C.byref(sid)),
                   "ngw_batch_add_system")
        self.systems[sid.value] = {"e1": e1.copy(), "nval": int(nval), "nband": nband}
        return sid.value

    def remove_system(self, sid: int):
        _lib.check(self._lib.ngw_batch_remove_system(self._h, int(sid)), "ngw_batch_remove_system")
        self.systems.pop(sid, None)

    def request_bytes(self, sid, nqp, niter, nen) -> int:
        b = C.c_long()
        _lib.check(self._lib.ngw_batch_request_bytes(self._h, int(sid), int(nqp), int(niter), int(nen),
                                                     C.byref(b)), "ngw_batch_request_bytes")
        return b.value

    def submit(self, sid, qp, nen, deltae, ecuts, niter, polynomial, tamm_d=False) -> int:
        qp = np.ascontiguousarray(qp, np.int32)
        poly = _f64(polynomial)
        rid = C.c_int()
        _lib.check(self._lib.ngw_batch_submit(self._h, int(sid), len(qp), qp.ctypes.data_as(C.POINTER(C.c_int)),
                                              int(nen), float(deltae), float(ecuts), int(niter), _p(poly),
                                              len(poly) - 1, int(tamm_d), C.byref(rid)), "ngw_batch_submit")
        return rid.value

    def step(self, nsteps=1, use_graph=True) -> int:
        n = C.c_int()
        _lib.check(self._lib.ngw_batch_step(self._h, int(nsteps), int(bool(use_graph)), C.byref(n)),
                   "ngw_batch_step")
        return n.value

    def active(self) -> int:
        n = C.c_int()
        _lib.check(self._lib.ngw_batch_active(self._h, C.byref(n)), "ngw_batch_active")
        return n.value

    def result(self, rid, nen, nqp):
        out = np.zeros((nen, nqp), np.complex128, order="F")
        done = C.c_int()
        _lib.check(self._lib.ngw_batch_result(self._h, int(rid), out.ctypes.data_as(C.POINTER(C.c_double)),
                                              C.byref(done)), "ngw_batch_result")
        return out if done.value else None

    def release(self, rid):
        _lib.check(self._lib.ngw_batch_release(self._h, int(rid)), "ngw_batch_release")

    def hold(self, rid, hold=True):
        """Keep an unfinished request out of the steps (hold=True) or let it
        step again; its Lanczos state stays on the device."""
        _lib.check(self._lib.ngw_batch_hold(self._h, int(rid), 1 if hold else 0), "ngw_batch_hold")

    def stats(self):
        s, g = C.c_long(), C.c_long()
        _lib.check(self._lib.ngw_batch_stats(self._h, C.byref(s), C.byref(g)), "ngw_batch_stats")
        return {"steps_run": s.value, "graph_captures": g.value}


def step_cost(m, nqp):
    """Operator FLOP of one Lanczos step of a request (nqp states, all bands
    as columns, F_mv = 4 Nmu nv nc + 2 Nmu^2 per vector, npoly applications
    cost the same for every request and are left out)."""
    nmu, nb = m.psi_intp.shape[0], len(m.e1)
    nc = nb - m.nv
    return nqp * nb * (4.0 * nmu * m.nv * nc + 2.0 * nmu * nmu)


def large_requests(costs, factor):
    """Ids of the requests whose step cost exceeds `factor` times the median
    step cost of the other running requests (none when only one runs)."""
    if len(costs) < 2:
        return set()
    out = set()
    for rid, c in costs.items():
        others = [v for r, v in costs.items() if r != rid]
        if c > factor * float(np.median(others)):
            out.add(rid)
    return out


def serve_stream(ctx: IsdfContext, molecules, requests, niter, nen, deltae, ecuts, hbm_budget_bytes,
                 sigx=None, use_graph=True, max_active=None, log=None, batch=None, sids=None, polys=None,
                 large_factor=50.0, large_every=8):
    """Continuous-batching request loop.

    molecules: list of objects with psi_intp, mmtrx, e1, nv (resident systems)
    requests : list of dicts {"t": arrival time (s), "mol": index, "qp": [bands]}
    batch, sids, polys: an existing GwBatch with the molecules already added
               (system ids and sqrt polynomials per molecule), so that operator
               upload and preparation stay outside a timed serving window
    large_factor, large_every: a running request whose step costs more than
               large_factor x the median step cost of the other running
               requests is held out of the batched steps while smaller
               requests run, except on every large_every-th step (it keeps
               progressing and cannot starve); with only large requests
               running, they step every time.  large_every = 0 turns the
               policy off (every request in every step).
    Returns per-request records with arrival, admit, finish times, Sigma_c and
    QP energies (Newton on the device), and loop statistics.
    """
    own = batch is None
    if own:
        batch = GwBatch(ctx)
        sids = [batch.add_system(m.psi_intp, m.mmtrx, m.e1, m.nv) for m in molecules]
    if polys is None:
        polys = [lanczos_polynomial(m.e1, m.nv, 12) for m in molecules]
    pending = sorted(range(len(requests)), key=lambda i: requests[i]["t"])
    queue = []           # arrived, waiting for admission (FIFO by arrival)
    running = {}         # rid -> request index
    cost = {}            # rid -> operator FLOP per Lanczos step
    held = set()
    held_steps = 0
    used = 0
    recs = [None] * len(requests)
    s0 = batch.stats()["steps_run"]
    t0 = time.perf_counter()
    pi = 0
    steps = 0
    done = 0
    t_log = t0
    while pi < len(pending) or queue or running:
        now = time.perf_counter() - t0
        while pi < len(pending) and requests[pending[pi]]["t"] <= now:
            heapq.heappush(queue, (requests[pending[pi]]["t"], pending[pi]))
            pi += 1
        # admission by HBM footprint, FIFO
        while queue:
            _, i = queue[0]
            rq = requests[i]
            need = batch.request_bytes(sids[rq["mol"]], len(rq["qp"]), niter, nen)
            if running and (used + need > hbm_budget_bytes or (max_active and len(running) >= max_active)):
                break
            heapq.heappop(queue)
            rid = batch.submit(sids[rq["mol"]], rq["qp"], nen, deltae, ecuts, niter, polys[rq["mol"]])
            running[rid] = (i, need)
            cost[rid] = step_cost(molecules[rq["mol"]], len(rq["qp"]))
            used += need
            recs[i] = {"arrival": rq["t"], "admit": time.perf_counter() - t0, "mol": rq["mol"], "qp": rq["qp"]}
        if not running:
            if pi < len(pending):
                time.sleep(max(0.0, min(0.01, requests[pending[pi]]["t"] - now)))
            continue
        if large_every:
            large = large_requests({rid: cost[rid] for rid in running}, large_factor)
            run_large = len(large) == len(running) or steps % large_every == 0
            want = set() if run_large else large
            for rid in running:
                if (rid in want) != (rid in held):
                    batch.hold(rid, rid in want)
            held = want
            held_steps += 1 if want else 0
        batch.step(1, use_graph)
        steps += 1
        if log and time.perf_counter() - t_log > 20.0:
            t_log = time.perf_counter()
            log(f"[serve] {t_log - t0:.1f}s: {steps} batched steps, {len(running)} running, {len(queue)} queued, "
                f"{done} done")
        for rid in list(running):
            i, need = running[rid]
            sig = batch.result(rid, nen, len(requests[i]["qp"]))
            if sig is None:
                continue
            m = molecules[requests[i]["mol"]]
            qp = list(requests[i]["qp"])
            e = m.e1[qp] * RYD
            sx = np.zeros(len(qp)) if sigx is None else sigx[requests[i]["mol"]](qp) * RYD
            res = ctx.qp_newton(sig * RYD, deltae * RYD, e, e, sx, np.zeros(len(qp)))
            recs[i].update({"finish": time.perf_counter() - t0, "sigma": sig, "eqp": res["eqp"],
                            "zfac": res["zfac"], "ierr": res["ierr"]})
            batch.release(rid)
            del running[rid]
            cost.pop(rid, None)
            held.discard(rid)
            used -= need
            done += 1
            if log:
                log(f"[serve] request {i} done: latency {recs[i]['finish'] - recs[i]['arrival']:.3f}s")
    stats = batch.stats()
    stats["batch_steps"] = steps
    stats["steps_with_large_held"] = held_steps
    stats["lanczos_steps"] = stats["steps_run"] - s0
    stats["wall_s"] = time.perf_counter() - t0
    if own:
        batch.close()
    return recs, stats


def latency_stats(recs):
    """p50 / p99 / max time-to-QP-energies (s) of finished requests."""
    lat = np.array([r["finish"] - r["arrival"] for r in recs if r is not None and "finish" in r])
    if lat.size == 0:
        return {"p50_s": None, "p99_s": None, "max_s": None, "n": 0}
    return {"p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "max_s": float(lat.max()), "n": int(lat.size)}
