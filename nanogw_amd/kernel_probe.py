"""Short fixed-shape driver of the operator kernels, for profiler runs.

Builds a random ISDF operator of the given sizes on the GPU (Psi at the ISDF
points ~ N(0, 1/Nmu), M a random symmetric positive matrix), applies the
operator to --nvec vectors --reps times in the chosen precision through the
device-pointer entry point (or, with --lanczos N, runs N Lanczos steps of
p(H) with the degree-12 sqrt polynomial, so every GEMM launch carries the
Horner / Clenshaw and Lanczos-accumulate epilogues of a GW run), then runs one
plain device-to-device copy of --copy-bytes (the reference kernel that
converts a hardware byte counter into a scale factor).  Kernel timings and counters come from the profiler
wrapped around this command; the script prints what it launched.

With --single-vector the probe instead times one-vector FP64 applications
(the shape of NanoGW's matvec_isdf / lanczos_spectra_isdf calls) in both
operator forms: the factored one (ngw_set_isdf) and the explicit pair matrix
C (ngw_set_explicit_c), --reps applications each, and prints the time per
application against the bytes each form has to read.

  python -m nanogw_amd.kernel_probe --nmu 5000 --nv 100 --nc 1500 \\
      --nvec 1600 --precision bf16 --reps 3
"""
from __future__ import annotations

import argparse
import sys
import time


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--nmu", type=int, default=5000)
    ap.add_argument("--nv", type=int, default=100)
    ap.add_argument("--nc", type=int, default=1500)
    ap.add_argument("--nvec", type=int, default=1600)
    ap.add_argument("--precision", default="f16", choices=["f64", "f32", "bf16", "bf16x3", "f16"])
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--copy-bytes", type=int, default=1 << 30)
    ap.add_argument("--tda", action="store_true")
    ap.add_argument("--lanczos", type=int, default=0,
                    help="run this many Lanczos steps (12 operator applications each) instead of --reps matvecs")
    ap.add_argument("--single-vector", action="store_true",
                    help="time FP64 one-vector applications, factored and explicit-C forms")
    ap.add_argument("--tiles", default="auto",
                    help="K1,K3 tile configuration indices (e.g. 0,0), or 'auto' to run the autotuner first")
    a = ap.parse_args(argv)

    import numpy as np
    import torch

    from .lanczos import IsdfContext

    rng = np.random.default_rng(1)
    nb = a.nv + a.nc
    psi = rng.normal(size=(a.nmu, nb)) / np.sqrt(a.nmu)
    g = rng.normal(size=(a.nmu, a.nmu)) / a.nmu
    mm = g @ g.T + np.eye(a.nmu)
    e1 = np.concatenate([np.sort(rng.uniform(-1.8, -0.5, a.nv)), np.sort(rng.uniform(-0.3, 4.0, a.nc))])
    ctx = IsdfContext(0)
    ctx.set_isdf(np.asfortranarray(psi), np.asfortranarray(mm), e1, a.nv)
    if a.single_vector:
        return single_vector_probe(a, ctx, psi, mm, e1)
    ctx.set_precision(a.precision)
    if a.precision != "f64":
        if a.tiles == "auto":
            k1, k3 = ctx.autotune(a.nvec, reps=2)
        else:
            k1, k3 = (int(v) for v in a.tiles.split(","))
            ctx.set_tile_config(k1, k3)
        print(f"[probe] tile configurations K1 {k1}, K3 {k3}", flush=True)
    ncv = a.nv * a.nc
    dev = torch.device("cuda:0")
    if a.lanczos > 0:
        return lanczos_probe(a, ctx, e1, dev)
    x = torch.randn((a.nvec, ncv), dtype=torch.float64, device=dev)
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    ctx.matvec_dev(x.data_ptr(), y.data_ptr(), ncv, a.nvec, a.tda)  # first call prepares the operands
    ctx.synchronize()
    ctx.profile_enable(True)
    ctx.profile_reset()
    t = time.perf_counter()
    for _ in range(a.reps):
        ctx.matvec_dev(x.data_ptr(), y.data_ptr(), ncv, a.nvec, a.tda)
    ctx.synchronize()
    dt = (time.perf_counter() - t) / a.reps
    prof = ctx.profile_read()
    ctx.profile_enable(False)
    fl = {"K1": 2.0 * a.nmu * a.nc * a.nv * a.nvec, "K2": 2.0 * a.nmu * a.nmu * a.nvec,
          "K3": 2.0 * a.nc * a.nmu * a.nv * a.nvec}
    per = "  ".join(f"{k} {prof[k][0] / max(prof[k][1], 1):.3f} ms ({fl[k] / (prof[k][0] / max(prof[k][1], 1)) / 1e9:.1f} "
                    f"TFLOP/s)" for k in ("K1", "K2", "K3") if prof[k][1])
    conv = prof["convert"][0] / a.reps
    print(f"[probe] operator {a.precision} Nmu={a.nmu} nv={a.nv} nc={a.nc} nvec={a.nvec}: "
          f"{a.reps + 1} applications, {dt * 1e3:.3f} ms each after the first; {per}; convert {conv:.3f} ms",
          flush=True)
    calibration_copy(a, dev)
    ctx.close()
    return 0


def calibration_copy(a, dev):
    import torch

    n = a.copy_bytes // 8
    src = torch.ones(n, dtype=torch.float64, device=dev)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    print(f"[probe] device copy of {n * 8} bytes (read {n * 8}, write {n * 8})", flush=True)


def single_vector_probe(a, ctx, psi, mm, e1):
    """FP64 one-vector applications: factored operator, then explicit C
    (C(mu, icv) = Psi_c(mu) Psi_v(mu), icv = iv + nv ic)."""
    import numpy as np
    import torch

    ncv = a.nv * a.nc
    dev = torch.device("cuda:0")
    x = torch.randn((ncv,), dtype=torch.float64, device=dev)
    y = torch.empty_like(x)
    s = np.sqrt((e1[a.nv:][:, None] - e1[:a.nv][None, :]).reshape(-1))
    rows = []
    for form in ("factored", "explicit"):
        if form == "factored":
            ctx.set_precision("f64")
            nbytes = 8.0 * (a.nmu * (a.nv + a.nc) + a.nmu * a.nmu + 3 * ncv)
        else:
            cm = np.asfortranarray((psi[:, a.nv:][:, :, None] * psi[:, :a.nv][:, None, :]).reshape(a.nmu, ncv))
            ctx.set_explicit_c(cm, mm, s)
            del cm
            nbytes = 8.0 * (2 * a.nmu * ncv + a.nmu * a.nmu + 3 * ncv)
        torch.cuda.synchronize()
        ctx.matvec_dev(x.data_ptr(), y.data_ptr(), ncv, 1, False)
        ctx.synchronize()
        t = time.perf_counter()
        for _ in range(a.reps):
            ctx.matvec_dev(x.data_ptr(), y.data_ptr(), ncv, 1, False)
        ctx.synchronize()
        dt = (time.perf_counter() - t) / a.reps
        rows.append((form, dt, nbytes))
        print(f"[probe] single vector FP64 {form} Nmu={a.nmu} nv={a.nv} nc={a.nc}: {a.reps} applications, "
              f"{dt * 1e6:.1f} us each; {nbytes / 1e6:.1f} MB to read, {nbytes / dt / 1e12:.2f} TB/s", flush=True)
    ctx.close()
    return 0


def lanczos_probe(a, ctx, e1, dev):
    """a.lanczos Lanczos steps on a.nvec random start vectors (the recurrence
    of ngw_lanczos_spectra_isdf_block, 12 operator applications per step)."""
    import numpy as np

    from .lanczos import lanczos_polynomial

    poly = lanczos_polynomial(e1, a.nv, 12)
    ncv = a.nv * a.nc
    v0 = np.asfortranarray(np.random.default_rng(2).normal(size=(ncv, a.nvec)))
    vn = np.linalg.norm(v0, axis=0)
    zz = np.zeros((2, a.nvec), np.complex128) + 0.02j
    ctx.lanczos_spectra_isdf_BLOCK(v0, vn, zz, 1, poly)  # allocates and prepares
    ctx.synchronize()
    ctx.profile_enable(True)
    ctx.profile_reset()
    t = time.perf_counter()
    ctx.lanczos_spectra_isdf_BLOCK(v0, vn, zz, a.lanczos, poly)
    ctx.synchronize()
    dt = time.perf_counter() - t
    prof = ctx.profile_read()
    ctx.profile_enable(False)
    napp = 12 * a.lanczos
    fl = {"K1": 2.0 * a.nmu * a.nc * a.nv * a.nvec, "K2": 2.0 * a.nmu * a.nmu * a.nvec,
          "K3": 2.0 * a.nc * a.nmu * a.nv * a.nvec}
    per = "  ".join(f"{k} {prof[k][0] / max(prof[k][1], 1):.3f} ms ({fl[k] / (prof[k][0] / max(prof[k][1], 1)) / 1e9:.1f} "
                    f"TFLOP/s)" for k in ("K1", "K2", "K3") if prof[k][1])
    print(f"[probe] lanczos {a.precision} Nmu={a.nmu} nv={a.nv} nc={a.nc} nvec={a.nvec}: {a.lanczos} steps "
          f"({napp} operator applications) in {dt:.3f} s including the host copies; {per}; convert "
          f"{prof['convert'][0] / napp:.3f} ms per application", flush=True)
    calibration_copy(a, dev)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
