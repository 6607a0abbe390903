"""Writes the golden fixtures of tests/golden/ (run from the repository root).

  H_aux_v0.npz   the data of NanoGW's own Lanczos known-answer file
                 shared/test_lanczos/H_aux_v0.txt (104 x 104 H_aux, v0, and the
                 printed (z, S(0), Re S(z+0.1i), Im S(z+0.1i)) table), converted
                 to npz; read from /root/reference when present.
  c1_inputs.npz  the config-1 synthetic molecule (R = 7, h = 0.45, nv = 8,
                 nc = 56, Nmu = 480): Psi at the ISDF points, M, energies, the
                 pinned sqrt polynomial (npoly = 12, polyfit_sqrt nx = 20) and
                 the fixed Sigma_x inputs.
  c1_sigma.npz   Sigma_c(omega) of the four QP states (HOMO-1..LUMO+1) from the
                 CPU oracle at niter = 200, and the oracle's QP energies.

Usage: python tests/golden/make_fixtures.py [kat] [c1]
"""
from __future__ import annotations

import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
OUT = os.path.dirname(os.path.abspath(__file__))
REF_KAT = "/root/reference/shared/test_lanczos/H_aux_v0.txt"


def make_kat():
    lines = open(REF_KAT).read().split("\n")
    n = int(lines[0].split()[1])
    h = np.zeros((n, n))
    k = 1
    for _ in range(n * n):
        i, j, v = lines[k].split()
        h[int(i), int(j)] = float(v)
        k += 1
    assert lines[k].startswith("v0")
    k += 1
    v0 = np.array([float(lines[k + i].split()[1]) for i in range(n)])
    k += n
    header = lines[k]
    k += 1
    rows = []
    while k < len(lines) and lines[k].strip():
        rows.append([float(x) for x in lines[k].split()])
        k += 1
    np.savez_compressed(os.path.join(OUT, "H_aux_v0.npz"), h_aux=h, v0=v0, table=np.array(rows),
                        eta=0.1, header=header)
    print("H_aux_v0.npz", h.shape, len(rows))


def make_c1():
    from nanogw_amd import synth
    from oracle import oracle as O

    t = time.time()
    mol = synth.make_molecule(7.0, 0.45, 8, 56, 480, synth.config_seed(1))
    nv, nc = mol.nv, mol.nc
    emin, emax = synth.polyfit_range(mol.e1, nv, mol.nb)
    beta, fe = O.polyfit_sqrt(emin, emax, 12, 20, False)
    poly = beta[::-1].copy()
    sd = synth.sigma_defaults(nv)
    sigx = synth.exchange_diag(mol.psi_intp, mol.mmtrx, nv, sd["qp"])
    np.savez_compressed(os.path.join(OUT, "c1_inputs.npz"), psi_intp=mol.psi_intp, mmtrx=mol.mmtrx,
                        e1=mol.e1, nv=nv, nc=nc, poly=poly, fit_error=fe, qp=np.array(sd["qp"]),
                        nen=sd["nen"], deltae=sd["deltae"], ecuts=sd["ecuts"], sigx=sigx,
                        intp=mol.intp, nfft=np.array(mol.nfft), h=mol.h, trunc=mol.trunc,
                        seed=mol.seed)
    print("c1 inputs", time.time() - t)
    t = time.time()
    niter = 200
    sig = O.sigma_lanczos(mol.psi_intp, mol.mmtrx, mol.e1, nv, nc, sd["qp"], sd["nen"], sd["deltae"],
                          sd["ecuts"], niter, poly, 64)
    print("c1 oracle sigma", time.time() - t)
    ryd = O.RYD
    qps = []
    for i, m in enumerate(sd["qp"]):
        r = O.qp_solve(sig[:, i] * ryd, sd["deltae"] * ryd, mol.e1[m] * ryd, mol.e1[m] * ryd,
                       sigx[i] * ryd, 0.0)
        qps.append([r["eqp"], r["zfac"], r["sigma0"].real, r["imsig"], r["ierr"]])
    np.savez_compressed(os.path.join(OUT, "c1_sigma.npz"), sigma=sig, niter=niter,
                        qp_results=np.array(qps))
    print(np.array(qps))


if __name__ == "__main__":
    what = sys.argv[1:] or ["kat", "c1"]
    if "kat" in what:
        make_kat()
    if "c1" in what:
        make_c1()
