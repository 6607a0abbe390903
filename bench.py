"""Throughput of the ISDF-Lanczos GW hot path on one MI355X.

One step = the full GW quasiparticle solve of `--qp-per-step` QP states of a
synthetic molecule: for each state every band's Lanczos run (niter steps of
npoly = 12 operator applications each), the continued fraction on nen + 1
points, Sigma_c(omega) assembly and the Newton QP root -- all on the GPU,
with the operator (Psi at the ISDF points, M, energies) resident in HBM when
the timed region starts.

Prints one JSON line (metric, value = QP solves/s, roofline of the dominant
kernel from HIP events inside the library, CPU baseline from the oracle on a
bounded sample).  Only one GPU is supported.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"f64": 78.6e12, "f32": 157.3e12, "bf16": 2.5e15, "bf16x3": 2.5e15}
HBM_PEAK = 8.0e12
RYD = 13.60569193


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_workload(cfg, ctx, seed_index=0, cvt_iter=None):
    import torch

    from nanogw_amd import synth

    R, h, nv, nc, nmu, niter, nqp = synth.CONFIGS[cfg]
    cfgnum = int(cfg[1:])
    t = time.time()
    mol = synth.make_molecule(R, h, nv, nc, nmu, synth.config_seed(cfgnum, seed_index),
                              cvt_iter=cvt_iter or (300 if nmu <= 500 else 60), kernel_ctx=ctx,
                              device=torch.device("cuda:0"), keep_zeta=False)
    log(f"[bench] {cfg}: ngrid={mol.ngrid} nfft={mol.nfft} nv={nv} nc={nc} nmu={nmu} "
        f"built in {time.time() - t:.1f}s")
    return mol, niter, synth.qp_states(nv, nqp)


def cpu_baseline(mol, poly, niter, nb_sample=8, target_s=12.0):
    """Oracle (reference BLOCK algorithm: explicit C rebuilt in every matvec)
    timed on nb_sample bands for a few Lanczos steps, extrapolated linearly to
    one QP solve (N_b bands x niter steps; every step costs the same)."""
    from oracle import oracle as O

    nv, nc = mol.nv, mol.nc
    ncv = nv * nc
    s = O.sqrt_r(mol.e1, nv, nc)
    rng = np.random.default_rng(0)
    v0 = rng.normal(size=(ncv, nb_sample))
    vn = np.linalg.norm(v0, axis=0)
    zz = np.zeros((201, nb_sample), np.complex128) + 0.02j
    steps = 2
    t = time.time()
    O.lanczos_spectra_block(mol.psi_intp, mol.mmtrx, s, v0, vn, zz, steps, poly, nv, nc, rebuild=True)
    dt = time.time() - t
    if dt < target_s:
        steps = max(2, min(niter, int(steps * target_s / max(dt, 1e-3))))
        t = time.time()
        O.lanczos_spectra_block(mol.psi_intp, mol.mmtrx, s, v0, vn, zz, steps, poly, nv, nc,
                                rebuild=True)
        dt = time.time() - t
    per_qp = dt * (niter / steps) * (mol.nb / nb_sample)
    return {"value": 1.0 / per_qp, "unit": "QP solves/s", "cores": O.num_threads(), "kind": "port",
            "sample": f"oracle BLOCK Lanczos, {nb_sample} bands x {steps} of {niter} steps "
                      f"(npoly={len(poly) - 1}), {dt:.1f}s, extrapolated x{niter / steps:.1f} steps "
                      f"x{mol.nb / nb_sample:.1f} bands per QP solve"}


def fp64_sample(ctx, mol, poly, niter, steps, q):
    """Plain FP64 GPU path on the same batch shape (q states x N_b bands) for
    `steps` Lanczos steps; every step does identical work, so QP solves/s is
    extrapolated linearly in the step count."""
    nv, nc, nb = mol.nv, mol.nc, mol.nb
    B = q * nb
    rng = np.random.default_rng(5)
    v0 = rng.normal(size=(nv * nc, B))
    vn = np.linalg.norm(v0, axis=0)
    zz = np.zeros((2, B), np.complex128) + 0.02j
    ctx.set_precision("f64")
    ctx.lanczos_spectra_isdf_BLOCK(v0, vn, zz, 1, poly)  # allocate
    ctx.synchronize()
    t = time.perf_counter()
    ctx.lanczos_spectra_isdf_BLOCK(v0, vn, zz, steps, poly)
    dt = time.perf_counter() - t
    per_qp = dt * (niter / steps) / q
    return {"qp_solves_per_s": 1.0 / per_qp, "timed_lanczos_steps": steps, "of_niter": niter,
            "seconds": dt, "columns": B,
            "note": "FP64 GPU path timed on the full batch for a few Lanczos steps, extrapolated"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c2", choices=["c1", "c2", "c3"])
    ap.add_argument("--precision", default="f64", choices=["f64", "f32", "bf16", "bf16x3"])
    ap.add_argument("--qp-per-step", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-profile", action="store_true")
    ap.add_argument("--fp64-steps", type=int, default=2,
                    help="Lanczos steps of the plain FP64 GPU path timed on the same batch for comparison")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the ISDF-Lanczos path "
              "shards by independent (QP state, band) runs, see DESIGN.md", file=sys.stderr)
        sys.exit(2)

    import torch

    from nanogw_amd import synth
    from nanogw_amd.lanczos import IsdfContext, lanczos_polynomial

    ctx = IsdfContext(0)
    mol, niter, qps = build_workload(args.config, ctx)
    poly = lanczos_polynomial(mol.e1, mol.nv, 12)
    ctx.set_isdf(mol.psi_intp, mol.mmtrx, mol.e1, mol.nv)
    ctx.set_precision(args.precision)
    sd = synth.sigma_defaults(mol.nv)
    sigx = synth.exchange_diag(mol.psi_intp, mol.mmtrx, mol.nv, qps)
    nen, deltae, ecuts = sd["nen"], sd["deltae"], sd["ecuts"]
    q = args.qp_per_step

    def step(i):
        sel = [qps[(i * q + j) % len(qps)] for j in range(q)]
        sig = ctx.calculate_sigma_lanczos(sel, nen, deltae, ecuts, niter, poly)
        idx = [qps.index(m) for m in sel]
        e = mol.e1[sel] * RYD
        return ctx.qp_newton(sig * RYD, deltae * RYD, e, e, sigx[idx] * RYD, np.zeros(len(sel)))

    for i in range(args.warmup):
        t = time.time()
        step(i)
        log(f"[bench] warmup step {i}: {time.time() - t:.2f}s")
    if not args.no_profile:
        ctx.profile_enable(True)
        ctx.profile_reset()
    torch.cuda.synchronize()
    ctx.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        r = step(args.warmup + i)
        log(f"[bench] step {i}: E_qp = {np.round(r['eqp'], 4).tolist()} eV")
    ctx.synchronize()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    ms_per_step = 1000.0 * elapsed / args.steps
    solves = args.steps * q
    value = solves / elapsed

    nv, nc, nmu, nb = mol.nv, mol.nc, mol.n_intp, mol.nb
    npoly = len(poly) - 1
    f_mv = 4.0 * nmu * nv * nc + 2.0 * nmu * nmu
    matvecs = solves * nb * niter * npoly
    roof = None
    prof = {}
    if not args.no_profile:
        prof = ctx.profile_read()
        ctx.profile_enable(False)
        # per launch, B = q * nb vectors
        B = q * nb
        flops = {"K1": 2.0 * nmu * nc * nv * B, "K2": 2.0 * nmu * nmu * B, "K3": 2.0 * nc * nmu * nv * B}
        dom = max(("K1", "K2", "K3"), key=lambda k: prof[k][0])
        ms, cnt = prof[dom]
        avg_s = ms / 1000.0 / max(cnt, 1)
        ach = flops[dom] / avg_s
        peak = PEAK[args.precision]
        roof = {"bound": "mfma", "kernel": dom, "achieved": ach / 1e12, "peak": peak / 1e12,
                "unit": "TFLOP/s", "frac": ach / peak, "traffic": None,
                "avg_launch_us": avg_s * 1e6, "launches": cnt,
                "flop_per_launch": flops[dom]}
    fp64_ref = None
    if args.precision != "f64" and args.fp64_steps > 0:
        fp64_ref = fp64_sample(ctx, mol, poly, niter, args.fp64_steps, q)
        fp64_ref["speedup"] = value / fp64_ref["qp_solves_per_s"]
    cpu = None
    if not args.no_cpu_baseline:
        log("[bench] cpu baseline (oracle sample)...")
        cpu = cpu_baseline(mol, poly, niter)
    out = {
        "metric": "QP solves/s (ISDF-Lanczos full-frequency GW, all bands, niter x npoly matvecs)",
        "value": value,
        "unit": "QP solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.precision,
        "data": "synthetic (seeded Gaussian-Hermite molecule, CVT ISDF points, device V build)",
        "config": {"workload": f"{args.config}: R={mol.radius} h={mol.h} ngrid={mol.ngrid} nv={nv} "
                               f"nc={nc} Nmu={nmu} N_b={nb} niter={niter} npoly={npoly} nen={nen}, "
                               f"{q} QP state(s) per step",
                   "qp_states": qps, "parallelism": "single GPU"},
        "matvecs_per_s": matvecs / elapsed,
        "gflops": f_mv * matvecs / elapsed / 1e9,
        "operator_frac_of_peak": f_mv * matvecs / elapsed / PEAK[args.precision],
        "roofline": roof,
        "kernel_ms": {k: v[0] for k, v in prof.items()},
        "cpu_baseline": cpu,
        "fp64_gpu": fp64_ref,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
