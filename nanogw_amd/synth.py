"""Seeded synthetic closed-shell molecules for the ISDF-Lanczos GW path.

Host-side input preparation (numpy).  It produces the arrays NanoGW's Lanczos
drivers consume -- orbitals at the ISDF points (``Psi_intp``), orbital
energies, the ISDF interpolation vectors zeta and the Coulomb kernel
``Mmtrx`` -- in NanoGW's units and conventions:

* Rydberg energies (e^2 = 2); epsilon[Ry] = 2 epsilon[Ha];
* grid-normalised orbitals, sum_r psi^2 = 1 (shared/parsec_wfn.F90:723-734),
  psi = phi * sqrt(h^3) for phi normalised as integral |phi|^2 = 1;
* real-space sphere of radius R and spacing h, integer grid coordinates,
  r_max = max |r|, FFT box nfft = nint(4 r_max / h) made even
  (shared/parsec_wfn.F90:365-381), Coulomb truncation R_c = 2 r_max (:421-425);
* ISDF points by a density-weighted Lloyd (CVT) iteration restated from
  shared/cvt.F90:8-195 with this module's seeded PRNG (the reference seeds
  gfortran's generator, which is not reproducible here), with each centroid
  snapped to the nearest grid point not taken by an earlier centroid;
* zeta = A^{-1} B with P(mu,r) = sum_v psi_v(r_mu) psi_v(r),
  Q(mu,r) = sum_n psi_n(r_mu) psi_n(r), A = P_intp o Q_intp, B = P o Q
  (shared/isdf_parallel_sym_UltraLowMem.F90:448-642).

The Coulomb kernel M = (1/h^3) zeta^T V zeta is built by the device path
(``IsdfContext.isdf_coulomb_kernel``, the device V build) when a context is
passed; ``host_coulomb_kernel``
below is a small numpy version used only to prepare inputs of CPU-side tests
when no GPU is present.

Orbitals: 3-D Gaussian envelope (sigma = 2.5 bohr) times seeded random
Hermite-polynomial sums of total degree <= 4, each orbital centred at a
seeded random offset |d| <= R/2 (the offset keeps nv + nc > 35 orbitals
linearly independent), then QR-orthonormalised in FP64 with weight h^3.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

HA_TO_RY = 2.0
RYD_EV = 13.60569193  # shared/myconstants.F90:41


@dataclasses.dataclass
class Molecule:
    """One synthetic system; energies in Ry, orbitals grid-normalised."""

    radius: float
    h: float
    nv: int
    nc: int
    n_intp: int
    seed: int
    gidx: np.ndarray          # (ngrid, 3) int32 grid coordinates
    psi: np.ndarray           # (ngrid, nb) grid-normalised orbitals, or None
    e1: np.ndarray            # (nb,) Ry, sorted within v and c
    intp: np.ndarray          # (n_intp,) 0-based grid indices of ISDF points
    psi_intp: np.ndarray      # (n_intp, nb) Fortran order
    nfft: tuple
    rmax: float
    trunc: float
    zeta: np.ndarray = None   # (ngrid, n_intp) Fortran order
    mmtrx: np.ndarray = None  # (n_intp, n_intp)

    @property
    def nb(self) -> int:
        return self.nv + self.nc

    @property
    def ngrid(self) -> int:
        return self.gidx.shape[0]

    @property
    def hcub(self) -> float:
        return self.h ** 3


def sphere_grid(radius: float, h: float):
    """Integer coordinates of grid points with |r| <= radius (shift 0)."""
    n = int(math.floor(radius / h)) + 1
    ax = np.arange(-n, n + 1)
    x, y, z = np.meshgrid(ax, ax, ax, indexing="ij")
    pts = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)
    r2 = (pts.astype(np.float64) * h) ** 2
    keep = r2.sum(axis=1) <= radius * radius + 1e-12
    g = pts[keep].astype(np.int32)
    # order z-slowest like a row-major box walk, deterministic
    order = np.lexsort((g[:, 0], g[:, 1], g[:, 2]))
    return np.ascontiguousarray(g[order])


def fft_box(gidx: np.ndarray, h: float):
    """r_max, nfft (even) and R_c = 2 r_max, shared/parsec_wfn.F90:257-269, 365-381, 421-425."""
    rmax = float(np.sqrt(((gidx.astype(np.float64) * h) ** 2).sum(axis=1).max()))
    n = int(np.rint(4.0 * rmax / h))
    if n % 2:
        n += 1
    return rmax, (n, n, n), 2.0 * rmax


def _hermite(n: int, x):
    h0 = x * 0.0 + 1.0
    if n == 0:
        return h0
    h1 = 2.0 * x
    for k in range(1, n):
        h0, h1 = h1, 2.0 * x * h1 - 2.0 * k * h0
    return h1


_DEGS = [(a, b, c) for a in range(5) for b in range(5) for c in range(5) if a + b + c <= 4]


def _orbital_draws(nb, rng, radius, nterms):
    """Seeded per-orbital parameters: centre offset, Hermite terms, coefficients."""
    out = []
    for _ in range(nb):
        d = rng.normal(size=3)
        d *= rng.uniform(0.0, 0.5 * radius) / max(np.linalg.norm(d), 1e-12)
        pick = rng.choice(len(_DEGS), size=nterms, replace=False)
        coef = rng.normal(size=nterms)
        out.append((d, [_DEGS[i] for i in pick], coef))
    return out


def orbitals(gidx, h, nb, rng, radius, sigma=2.5, nterms=6, device=None):
    """(ngrid, nb) grid-normalised orthonormal orbitals (QR with weight h^3).

    device: a torch device on which the orbitals are evaluated and
    QR-factorised in FP64 (large systems); numpy otherwise.  The random
    draws are the same either way.
    """
    draws = _orbital_draws(nb, rng, radius, nterms)
    if device is not None:
        import torch

        r = torch.as_tensor(gidx.astype(np.float64) * h, device=device)
        phi = torch.empty((nb, r.shape[0]), dtype=torch.float64, device=device)
        for j, (d, degs, coef) in enumerate(draws):
            u = (r - torch.as_tensor(d, device=device)) / sigma
            env = torch.exp(-0.5 * (u * u).sum(dim=1))
            poly = torch.zeros_like(env)
            for (a, b, c), cf in zip(degs, coef):
                poly += float(cf) * _hermite(a, u[:, 0]) * _hermite(b, u[:, 1]) * _hermite(c, u[:, 2])
            phi[j] = env * poly
        q, rr = torch.linalg.qr(phi.T * math.sqrt(h ** 3))
        del phi
        q *= torch.sign(torch.diagonal(rr))[None, :]
        out = np.asfortranarray(q.cpu().numpy())
        del q, rr
        torch.cuda.empty_cache()
        return out
    r = gidx.astype(np.float64) * h
    ng = r.shape[0]
    phi = np.empty((ng, nb), order="F")
    for j, (d, degs, coef) in enumerate(draws):
        u = (r - d) / sigma
        env = np.exp(-0.5 * (u * u).sum(axis=1))
        poly = np.zeros(ng)
        for (a, b, c), cf in zip(degs, coef):
            poly += cf * _hermite(a, u[:, 0]) * _hermite(b, u[:, 1]) * _hermite(c, u[:, 2])
        phi[:, j] = env * poly
    q, rr = np.linalg.qr(phi * math.sqrt(h ** 3))
    q *= np.sign(np.diag(rr))[None, :]
    return np.asfortranarray(q)


def energies(nv, nc, rng):
    """epsilon_v ~ U[-0.9, -0.25] Ha, epsilon_c ~ U[-0.15, 2.0] Ha, sorted; returned in Ry."""
    ev = np.sort(rng.uniform(-0.9, -0.25, nv))
    ec = np.sort(rng.uniform(-0.15, 2.0, nc))
    return HA_TO_RY * np.concatenate([ev, ec])


def cvt_points(gidx, h, rho, n_intp, rng, tol=1e-3, max_iter=300):
    """Density-weighted Lloyd iteration, shared/cvt.F90:8-195.

    Initial points uniform in [lo, lo + 0.9 (hi - lo)] per axis (:66-72);
    per iteration: nearest-centroid assignment (:77-88), rho-weighted
    centroids with the old point kept when the weight is < 1e-10 (:92-103),
    snap to the nearest grid point not already taken by an earlier centroid
    (:104-128, here through a KD-tree instead of the O(Nmu^2 N) scan),
    stop when the summed displacement is < tol (:131-135).  max_iter caps
    the iteration count (the reference allows 1e5).
    """
    from scipy.spatial import cKDTree

    pts = gidx.astype(np.float64) * h
    tree_grid = cKDTree(pts)
    lo, hi = pts.min(axis=0), pts.max(axis=0)
    new = lo[None, :] + rng.uniform(size=(n_intp, 3)) * 0.9 * (hi - lo)[None, :]
    old = new.copy()
    intp = np.zeros(n_intp, np.int64)
    for _ in range(max_iter):
        _, owner = cKDTree(new).query(pts)
        w = np.bincount(owner, weights=rho, minlength=n_intp)
        cen = np.stack([np.bincount(owner, weights=rho * pts[:, k], minlength=n_intp)
                        for k in range(3)], axis=1)
        ok = w >= 1e-10
        cand = np.where(ok[:, None], cen / np.where(ok, w, 1.0)[:, None], old)
        _, near = tree_grid.query(cand)
        taken = np.zeros(pts.shape[0], bool)
        for ip in range(n_intp):
            g = int(near[ip])
            if taken[g]:
                k = 2
                while True:
                    _, idx = tree_grid.query(cand[ip], k=k)
                    free = [i for i in np.atleast_1d(idx) if not taken[i]]
                    if free:
                        g = int(free[0])
                        break
                    k *= 2
            taken[g] = True
            intp[ip] = g
        new = pts[intp]
        diff = np.sqrt(((new - old) ** 2).sum(axis=1)).sum()
        if diff < tol:
            break
        old = new.copy()
    return intp


def isdf_zeta(psi, intp, nv, ridge=0.0, device=None):
    """zeta (ngrid, n_intp) = (A^{-1} B)^T, isdf_parallel_sym_UltraLowMem.F90:448-642.

    device: a torch device to run the FP64 products and the solve on (input
    preparation of large synthetic systems); numpy otherwise.
    """
    if device is not None:
        import torch

        t = torch.as_tensor(np.asarray(psi), dtype=torch.float64, device=device)
        ii = torch.as_tensor(np.asarray(intp), dtype=torch.long, device=device)
        pv = t[:, :nv]
        pi_v = pv[ii, :]
        pi_n = t[ii, :]
        A = (pi_v @ pi_v.T) * (pi_n @ pi_n.T)
        if ridge > 0.0:
            A = A + ridge * torch.trace(A) / A.shape[0] * torch.eye(A.shape[0], dtype=A.dtype, device=device)
        B = (pv @ pi_v.T) * (t @ pi_n.T)
        del t, pv
        # A is symmetric: zeta = B A^{-1}.  The n_intp x n_intp inverse is formed
        # on the host (LU with partial pivoting, scipy); the large product is a
        # plain GPU GEMM (device triangular solves need a library workspace
        # that torch caps below what n_intp = 5000 asks for).
        import scipy.linalg as sla

        ainv = sla.inv(A.cpu().numpy(), check_finite=False)
        ainv = torch.as_tensor(0.5 * (ainv + ainv.T), device=device)
        out = np.empty((B.shape[0], A.shape[0]), order="F")
        for r0 in range(0, B.shape[0], 32768):
            zt = B[r0:r0 + 32768] @ ainv
            out[r0:r0 + zt.shape[0], :] = zt.cpu().numpy()
        del A, B, ainv
        torch.cuda.empty_cache()
        return out
    pv = psi[:, :nv]
    pi_v = pv[intp, :]
    pi_n = psi[intp, :]
    P_intp = pi_v @ pi_v.T
    Q_intp = pi_n @ pi_n.T
    A = P_intp * Q_intp
    Pg = pv @ pi_v.T            # (ngrid, n_intp) = P(mu, r)^T
    Qg = psi @ pi_n.T
    B = Pg * Qg                 # (ngrid, n_intp)
    if ridge > 0.0:
        A = A + ridge * np.trace(A) / A.shape[0] * np.eye(A.shape[0])
    import scipy.linalg as sla

    zt = sla.solve(A, B.T, assume_a="sym")
    return np.asfortranarray(zt.T)


def host_coulomb_kernel(zeta, gidx, nfft, h, trunc, ntrans=1):
    """Numpy M = (ntrans/h^3) zeta^T V zeta for CPU-side test inputs only."""
    n1, n2, n3 = nfft
    bx = [2.0 * np.pi / (n * h) for n in nfft]
    j1 = np.abs(np.fft.fftfreq(n1, 1.0 / n1)[: n1 // 2 + 1])
    j2 = np.fft.fftfreq(n2, 1.0 / n2)
    j3 = np.fft.fftfreq(n3, 1.0 / n3)
    g2 = ((bx[0] * j1)[None, None, :] ** 2 + (bx[1] * j2)[None, :, None] ** 2 +
          (bx[2] * j3)[:, None, None] ** 2)
    with np.errstate(divide="ignore", invalid="ignore"):
        v = np.where(g2 < 1e-12, trunc ** 2 / 2.0, (1.0 - np.cos(trunc * np.sqrt(g2))) / g2)
    v *= 8.0 * np.pi
    ix, iy, iz = (np.mod(gidx[:, k], nfft[k]) for k in range(3))
    vz = np.empty_like(zeta)
    for k in range(zeta.shape[1]):
        box = np.zeros((n3, n2, n1))
        box[iz, iy, ix] = zeta[:, k]
        f = np.fft.rfftn(box) * v
        vz[:, k] = np.fft.irfftn(f, s=(n3, n2, n1), axes=(0, 1, 2))[iz, iy, ix]
    return (ntrans / h ** 3) * (zeta.T @ vz)


def make_molecule(radius, h, nv, nc, n_intp, seed, with_kernel=True, keep_psi=False,
                  cvt_iter=300, ridge=0.0, kernel_ctx=None, device=None, keep_zeta=True):
    """Build one synthetic molecule; see the module docstring for conventions.

    kernel_ctx: an ``IsdfContext`` whose device V build computes M (otherwise
    the numpy ``host_coulomb_kernel``); device: torch device for the zeta solve.
    """
    rng = np.random.default_rng(seed)
    gidx = sphere_grid(radius, h)
    rmax, nfft, trunc = fft_box(gidx, h)
    nb = nv + nc
    psi = orbitals(gidx, h, nb, rng, radius, device=device)
    e1 = energies(nv, nc, rng)
    rho = 2.0 * (psi[:, :nv] ** 2).sum(axis=1)
    intp = cvt_points(gidx, h, rho, n_intp, rng, max_iter=cvt_iter)
    psi_intp = np.asfortranarray(psi[intp, :])
    mol = Molecule(radius, h, nv, nc, n_intp, seed, gidx, psi if keep_psi else None, e1, intp,
                   psi_intp, nfft, rmax, trunc)
    if with_kernel:
        zeta = isdf_zeta(psi, intp, nv, ridge=ridge, device=device)
        if kernel_ctx is not None:
            m = kernel_ctx.isdf_coulomb_kernel(zeta, gidx, nfft, h, trunc)
        else:
            m = host_coulomb_kernel(zeta, gidx, nfft, h, trunc)
        mol.mmtrx = np.asfortranarray(0.5 * (m + m.T))
        if keep_zeta:
            mol.zeta = zeta
    if keep_psi:
        mol.psi = psi
    return mol


CONFIGS = {
    # name: (R, h, nv, nc, n_intp, niter, number of QP states)
    "c1": (7.0, 0.45, 8, 56, 480, 200, 4),
    "c2": (10.0, 0.35, 40, 400, 2000, 300, 8),
    "c3": (13.0, 0.33, 100, 1500, 5000, 300, 8),
}

# Tikhonov ridge of the zeta solve, relative to tr(A)/n_intp, for the
# benchmark molecules (c2 and larger).  The unregularized fit of these smooth
# synthetic orbitals is rank-deficient at large n_intp (c1 already has
# cond(A) ~ 4e11 and max|zeta| ~ 700): M then carries entries ~1e12 whose
# contributions to C^T M C cancel over ~12 digits, which only FP64 GEMMs
# resolve.  With this ridge max|zeta| stays O(10) and the pair-density fit
# error is ~1 %, the size of the ISDF approximation itself.  Both the FP64
# and the mixed-precision paths consume the same regularized inputs.
BENCH_RIDGE = 1e-6


def c4_specs(n_mol=64, seed=None):
    """Molecule sizes of the batched workload: R ~ U[6, 11] bohr, h = 0.4,
    nv ~ U{8..60}, nc = 10 nv, Nmu = 12 nv, niter = 250, 16 QP states each.
    Returns a list of dicts with the make_molecule arguments (the zeta fit
    uses BENCH_RIDGE)."""
    rng = np.random.default_rng(config_seed(4, 0) if seed is None else seed)
    out = []
    for i in range(n_mol):
        nv = int(rng.integers(8, 61))
        out.append({"radius": float(rng.uniform(6.0, 11.0)), "h": 0.4, "nv": nv, "nc": 10 * nv,
                    "n_intp": 12 * nv, "seed": config_seed(4, i), "nqp": 16, "niter": 250})
    return out


def c5_stream(duration_s, rate=20.0, seed=None, large_frac=0.1, n_small=64):
    """Poisson request stream: exponential gaps at `rate` req/s over
    duration_s; each request picks one of the n_small c4-distributed
    molecules, or with probability large_frac the R = 13, nv = 100 molecule
    (index n_small), and asks for U{4..16} QP states around the gap.
    Returns (requests, specs): requests as {"t", "mol", "nqp"}."""
    rng = np.random.default_rng(config_seed(5, 0) if seed is None else seed)
    specs = c4_specs(n_small, seed=config_seed(5, 1) if seed is None else seed + 1)
    R, h, nv, nc, nmu, niter, _ = CONFIGS["c3"]
    specs.append({"radius": R, "h": h, "nv": nv, "nc": nc, "n_intp": nmu, "seed": config_seed(5, 999),
                  "nqp": 8, "niter": niter})
    reqs, t = [], 0.0
    while True:
        t += float(rng.exponential(1.0 / rate))
        if t > duration_s:
            break
        big = rng.uniform() < large_frac
        mol = n_small if big else int(rng.integers(0, n_small))
        reqs.append({"t": t, "mol": mol, "nqp": int(rng.integers(4, 17))})
    return reqs, specs


def qp_window(nv, nb, nqp):
    """nqp QP states around the gap (HOMO-(nqp/2-1) .. LUMO+(nqp - nqp/2 - 1)),
    clipped to the available bands."""
    lo = max(0, nv - nqp // 2)
    hi = min(nb, lo + nqp)
    lo = max(0, hi - nqp)
    return list(range(lo, hi))


def qp_states(nv, nqp):
    """nqp QP states centred on the gap: HOMO-(nqp/2-1) .. LUMO+(nqp/2-1) (0-based)."""
    half = nqp // 2
    return list(range(nv - half, nv + half))


def config_seed(config: int, index: int = 0) -> int:
    """Seed scheme 20261202 + 1000*config + molecule index."""
    return 20261202 + 1000 * config + index


def sigma_defaults(nv: int):
    """NanoGW input settings of the benchmark workloads c1-c5 (Ry):
    energy_range 2 Ha, 200 points, gamma = 0.01 Ha, QP states HOMO-1..LUMO+1."""
    return {
        "deltae": 2.0 * HA_TO_RY,
        "nen": 200,
        "ecuts": 0.01 * HA_TO_RY,
        "qp": [nv - 2, nv - 1, nv, nv + 1],
    }


def polyfit_range(e1, nv, nb):
    """e_min, e_max (Ry^2), calculate_sigma_lanczos_BLOCK.F90z:166-171."""
    emin = (e1[nv] - e1[nv - 1]) ** 2 / 1.25
    emax = (e1[nb - 1] - e1[0]) ** 2 * 1.25
    return emin, emax


def exchange_diag(psi_intp, mmtrx, nv, qp):
    """Fixed Sigma_x input (Ry) for both paths: -sum_v (mv|vm) through ISDF."""
    out = []
    for m in qp:
        c = psi_intp[:, [m]] * psi_intp[:, :nv]      # (n_intp, nv)
        out.append(-float(np.einsum("iv,ij,jv->", c, mmtrx, c)))
    return np.array(out)
