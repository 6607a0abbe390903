"""Reader of PARSEC's wavefunction file, the input NanoGW's Sigma run starts
from (shared/parsec_wfn.F90:20-785): ``parsec.dat``, or the older
``wfn.dat`` layout that parsec_wfn falls back to.

The file is Fortran sequential unformatted: every record is framed by its
byte length as a 4-byte integer before and after (the gfortran / flang
convention; records over 2 GiB are split into sub-records whose leading
length is negative when more follow).  Records are read in exactly the
order parsec_wfn reads them, skipped records included:

  parsec.dat                               wfn.dat
  date label (26 characters)               date label
  nspin, complex flag, periodicity         step, -, ndim, nwedge, ntrans, nspin
  [periodic: step, avec, Laplacian
   directions, k-points, weights]
  [cluster: step]
  ndim; shift; nwedge, ntrans              shift
  -; trans(3,3,ntrans); -; -; chi(ntrans, ntrans)
  gtable(3, nwedge)
  per spin and k-point: nstate, irep, e0, occ0; then -, rho(nwedge)
  per spin and k-point: 7 records skipped,  per spin: the nstate
  nn, iord(nn), then nn wave functions      wave functions
  of nwedge values

What parsec_wfn derives is derived the same way here:
  * periodicity code 0 / 1 / 2 / 3 in the file means cluster / bulk / wire /
    slab, i.e. gvec%per = 0 / 3 / 1 / 2 (:110-128);
  * r_max = max |(unfold(g) + shift) * step| over the wedge points and every
    symmetry operation, non-periodic directions only (:237-255);
  * FFT box nfft(i) = nint(4 r_max / step(i)) for non-periodic directions (the
    cell length for periodic ones), made even, nfft(4) = nfft(1) + 2
    (:365-381; adjust_fft is a no-op outside ESSL builds);
  * hcub = step^3 and the Coulomb truncation R_c = 2 r_max for a cluster
    (:386-425);
  * grid points are ordered by decreasing total density (:459-477);
    parsec_wfn's quicksort is not stable, this order is (numpy's stable
    sort), so points of exactly equal density may come in another order --
    every quantity of the GW path is invariant under that permutation;
  * densities below zero are stored as zero (:566-577);
  * a wave function whose norm sum |psi|^2 ntrans differs from 1 by more
    than 1e-6 is reported (parsec_wfn warns and goes on) (:723-734);
  * the scissors operator of nanogw.in, E_new = E_old + (E_old - E_ref)
    slope + const (eV inputs, read as jsp, imin, imax, const, E_ref, slope)
    gives e1 from e0 (shared/read_scissors.f90:25-73).
Energies stay in Rydberg and wave functions grid-normalised, as in the file.
"""
from __future__ import annotations

import dataclasses
import struct

import numpy as np

RYD = 13.60569193  # shared/myconstants.F90:41


class FortranRecords:
    """Sequential unformatted records of a byte string (4-byte markers)."""

    def __init__(self, data: bytes, endian: str = "<"):
        self.data = memoryview(data)
        self.pos = 0
        self.endian = endian
        self.index = 0

    def _marker(self, at):
        if at + 4 > len(self.data):
            raise ValueError(f"truncated record marker at byte {at}")
        return struct.unpack_from(self.endian + "i", self.data, at)[0]

    def record(self) -> bytes:
        parts = []
        while True:
            n = self._marker(self.pos)
            more = n < 0
            n = abs(n)
            start = self.pos + 4
            if start + n + 4 > len(self.data):
                raise ValueError(f"record {self.index} runs past the end of the file")
            tail = self._marker(start + n)
            if abs(tail) != n:
                raise ValueError(f"record {self.index}: leading length {n} and trailing length {abs(tail)} differ")
            parts.append(bytes(self.data[start:start + n]))
            self.pos = start + n + 4
            if not more:
                break
        self.index += 1
        return b"".join(parts)

    def skip(self, n: int = 1):
        for _ in range(n):
            self.record()

    def at_end(self) -> bool:
        return self.pos >= len(self.data)

    def ints(self, count=None):
        r = self.record()
        a = np.frombuffer(r, dtype=self.endian + "i4")
        if count is not None and a.size < count:
            raise ValueError(f"record {self.index - 1}: {a.size} integers, expected {count}")
        return a[:count] if count is not None else a

    def reals(self, count=None):
        r = self.record()
        a = np.frombuffer(r, dtype=self.endian + "f8")
        if count is not None and a.size < count:
            raise ValueError(f"record {self.index - 1}: {a.size} reals, expected {count}")
        return a[:count] if count is not None else a

    def complexes(self, count):
        r = self.record()
        a = np.frombuffer(r, dtype=self.endian + "c16")
        if a.size < count:
            raise ValueError(f"record {self.index - 1}: {a.size} complex values, expected {count}")
        return a[:count]


@dataclasses.dataclass
class SpinKState:
    """Orbitals of one spin channel and k-point (kpt%wfn(isp, ik))."""
    nstate: int
    irep: np.ndarray
    e0: np.ndarray
    e1: np.ndarray
    occ0: np.ndarray
    states: np.ndarray = None   # 1-based indices of the kept orbitals, in memory order
    wf: np.ndarray = None       # (nr, nmem) in decreasing-density grid order
    norms: np.ndarray = None    # sum |psi|^2 ntrans of every kept orbital


@dataclasses.dataclass
class ParsecData:
    datelabel: str
    wfn_dat: bool
    nspin: int
    lcplx: bool
    per: int
    step: np.ndarray
    avec: np.ndarray
    ndim: int
    nwedge: int
    ntrans: int
    shift: np.ndarray
    trans: np.ndarray           # (ntrans, 3, 3): trans[j] is Fortran trans(:, :, j)
    chi: np.ndarray             # (ntrans, ntrans) as Fortran chi
    kpoints: np.ndarray         # (nk, 3) Cartesian, as in the file
    kweights: np.ndarray
    gtable: np.ndarray          # (nwedge, 3) grid coordinates in file order
    rindex: np.ndarray          # 0-based: r[j] = gtable[rindex[j]]
    r: np.ndarray               # (nr, 3) in decreasing-density order
    rho: np.ndarray             # (nr, nspin) in that order, negatives zeroed
    rmax: float
    nfft: tuple                 # (n1, n2, n3, n1 + 2)
    hcub: float
    trunc: float                # Coulomb truncation length (-1: none)
    wfn: dict                   # (isp, ik) 1-based -> SpinKState
    warnings: list


def _unfold(g, trans, shift):
    """shared/unfold.f90:10-29: diagonal of trans applied, nint."""
    d = np.array([trans[0, 0], trans[1, 1], trans[2, 2]])
    return np.rint(d[None, :] * g + d[None, :] * shift[None, :] - shift[None, :]).astype(np.int64)


def parse_scissors(lines):
    """Rows of the nanogw.in 'scissors' block: (jsp, imin, imax, const, eref, slope),
    the order read_scissors.f90:51 reads them in."""
    out = []
    for line in lines:
        t = line.replace(",", " ").split()
        if len(t) < 6:
            raise ValueError("ERROR in reading scissors operator")
        out.append((int(t[0]), int(t[1]), int(t[2]), float(t[3].lower().replace("d", "e")),
                    float(t[4].lower().replace("d", "e")), float(t[5].lower().replace("d", "e"))))
    return out


def apply_scissors(e0, isp, scissors):
    """read_scissors.f90:46-66: E_new = E_old + (E_old - eref/ryd) slope + const/ryd
    for orbitals imin..imax of spin isp; every matching operator applies to E_old."""
    e1 = np.array(e0, np.float64, copy=True)
    idx = np.arange(1, len(e0) + 1)
    for jsp, imin, imax, const, eref, slope in scissors or ():
        if jsp != isp:
            continue
        sel = (idx >= imin) & (idx <= imax)
        e1[sel] = e0[sel] + (e0[sel] - eref / RYD) * slope + const / RYD
    return e1


def read_parsec_dat(path_or_bytes, keep=None, scissors=None, wfn_dat=None, endian="<") -> ParsecData:
    """Read parsec.dat (or wfn.dat) as parsec_wfn does.

    keep: 1-based orbital indices to keep in memory (NanoGW's wmap; None keeps
    every orbital, nmap = 0); an index beyond the file's states raises, as
    parsec_wfn dies.  scissors: rows of parse_scissors.  wfn_dat: force the
    layout (None: the wfn.dat layout when the file is named wfn.dat)."""
    if isinstance(path_or_bytes, (bytes, bytearray, memoryview)):
        data, name = bytes(path_or_bytes), ""
    else:
        name = str(path_or_bytes)
        with open(name, "rb") as fh:
            data = fh.read()
    if wfn_dat is None:
        wfn_dat = name.endswith("wfn.dat")
    f = FortranRecords(data, endian)
    warnings = []
    datelabel = f.record()[:26].decode("latin-1")
    kpoints = np.zeros((1, 3))
    kweights = np.ones(1)
    avec = np.eye(3)
    if wfn_dat:
        r = f.record()
        rtmp, _vt = struct.unpack_from(endian + "dd", r, 0)
        ndim, nwedge, ntrans, nspin = struct.unpack_from(endian + "iiii", r, 16)
        per, lcplx = 0, False
        step = np.full(3, rtmp)
    else:
        nspin, cflag, fper = (int(v) for v in f.ints(3))
        lcplx = cflag == 1
        per = {1: 3, 2: 1, 3: 2}.get(fper, fper)
        if per < 0 or per > 3:
            raise ValueError(f"boundary conditions not consistent: {per}")
        if per > 0:
            step = np.array(f.reals(3))
            avec = np.array(f.reals(9)).reshape(3, 3, order="F")
            f.ints(1)       # lap_dir_num
            f.ints(3)       # lap_dir
            f.reals(3)      # lap_dir_step
            nk = int(f.ints(1)[0])
            f.skip(3)
            kpoints = np.array(f.reals(3 * nk)).reshape(nk, 3)
            kweights = np.array(f.reals(nk))
        else:
            step = np.full(3, float(f.reals(1)[0]))
        ndim = int(f.ints(1)[0])
        shift = np.array(f.reals(3))
        nwedge, ntrans = (int(v) for v in f.ints(2))
    if wfn_dat:
        shift = np.array(f.reals(3))
    f.skip(1)
    trans = np.array(f.reals(9 * ntrans)).reshape(ntrans, 3, 3).transpose(0, 2, 1)
    f.skip(2)
    chi = np.array(f.ints(ntrans * ntrans)).reshape(ntrans, ntrans, order="F")
    gtable = np.array(f.ints(3 * nwedge)).reshape(nwedge, 3).astype(np.int64)
    # r_max over the unfolded grid, non-periodic directions
    rmax2 = 0.0
    for j in range(ntrans):
        g = _unfold(gtable, trans[j], shift)
        v = (g + shift[None, :]) * step[None, :]
        v[:, :per] = 0.0
        rmax2 = max(rmax2, float((v * v).sum(axis=1).max()))
    rmax = float(np.sqrt(rmax2))
    nk = kpoints.shape[0]
    wfn = {}
    rho_in = np.zeros((nwedge, 3))
    for isp in range(1, nspin + 1):
        if isp == 2 and wfn_dat:
            f.skip(wfn[(1, 1)].nstate)
        for ik in range(1, nk + 1):
            nst = int(f.ints(1)[0])
            irep = np.array(f.ints(nst))
            e0 = np.array(f.reals(nst))
            occ0 = np.array(f.reals(nst))
            wfn[(isp, ik)] = SpinKState(nst, irep, e0, apply_scissors(e0, isp, scissors), occ0)
        f.skip(1)
        rho_in[:, isp] = f.reals(nwedge)
    if nspin == 2 and wfn_dat:
        f = FortranRecords(data, endian)
        f.skip(15)
    rho_in[:, 0] = rho_in[:, 1] + rho_in[:, 2]
    # FFT box and cell (:365-425)
    nfft = []
    for i in range(3):
        if not wfn_dat and per >= i + 1:
            length = float(np.sqrt(avec[:, i] @ avec[:, i]))
        else:
            length = 4.0 * rmax
        n = int(np.rint(length / step[i]))
        nfft.append(n + (n % 2))
    nfft = (nfft[0], nfft[1], nfft[2], nfft[0] + 2)
    hcub = float(step[0] ** 3) if per == 0 else None
    if per > 0:
        m = avec.copy()
        for i in range(per):
            m[:, i] = step[i] * m[:, i] / np.sqrt((m[:, i] ** 2).sum())
        hcub = abs(float(np.linalg.det(m)))
        for i in range(per, 3):
            hcub *= step[i]
    trunc = 2.0 * rmax if per < 3 else -1.0
    # grid order: decreasing total density
    rindex = np.argsort(-rho_in[:, 0], kind="stable")
    r = gtable[rindex]
    rho = np.where(rho_in[rindex, 1:1 + nspin] > 0.0, rho_in[rindex, 1:1 + nspin], 0.0)
    # wave functions
    for isp in range(1, nspin + 1):
        for ik in range(1, nk + 1):
            st = wfn[(isp, ik)]
            want = list(range(1, st.nstate + 1)) if keep is None else [int(k) for k in keep]
            for k in want:
                if k > st.nstate:
                    raise ValueError(f"ERROR: You requested at least {k} states but the file contains only "
                                     f"{st.nstate}")
            slot = {k: i for i, k in enumerate(want)}
            if wfn_dat:
                if isp == 2:
                    f.skip(6)
                iord = np.arange(1, st.nstate + 1)
            else:
                f.skip(7)
                nn = int(f.ints(1)[0])
                if nn == 0:
                    raise ValueError("ERROR: parsec.dat contains no wavefunctions")
                iord = np.array(f.ints(nn))
            wf = np.zeros((nwedge, len(want)), np.complex128 if lcplx else np.float64, order="F")
            found = np.zeros(len(want), bool)
            norms = np.zeros(len(want))
            for o in iord:
                col = f.complexes(nwedge) if lcplx else f.reals(nwedge)
                if int(o) not in slot:
                    continue
                j = slot[int(o)]
                wf[:, j] = col[rindex]
                found[j] = True
                norms[j] = float(np.real(np.vdot(wf[:, j], wf[:, j]))) * ntrans
                if abs(norms[j] - 1.0) > 1e-6:
                    warnings.append(f"WARNING: The wavefuctions are not normalized {int(o)} {nwedge} "
                                    f"{norms[j]} {isp}")
            if not found.all():
                missing = [k for k, ok in zip(want, found) if not ok]
                raise ValueError(f"ERROR! wavefunction for state {missing[0]} not found.")
            st.states = np.array(want)
            st.wf = wf
            st.norms = norms
    return ParsecData(datelabel, wfn_dat, nspin, lcplx, per, step, avec, ndim, nwedge, ntrans, shift, trans, chi,
                      kpoints, kweights, gtable, rindex, r, rho, rmax, nfft, hcub, trunc, wfn, warnings)


def write_parsec_dat(path, datelabel, step, shift, gtable, states, rho, endian="<"):
    """Write a cluster (periodicity 0, real orbitals, one spin, no symmetry)
    parsec.dat in the record layout read_parsec_dat reads; states is a list of
    (e0, occ0, irep, psi(nwedge)) per orbital.  Used to prepare inputs of
    synthetic molecules for the Sigma path."""
    def rec(b):
        return struct.pack(endian + "i", len(b)) + b + struct.pack(endian + "i", len(b))

    def ints(a):
        return rec(np.asarray(a, endian + "i4").tobytes())

    def reals(a):
        return rec(np.asarray(a, endian + "f8").tobytes())

    g = np.asarray(gtable, np.int64)
    nw = g.shape[0]
    out = [rec(datelabel.ljust(26)[:26].encode("latin-1")), ints([1, 0, 0]), reals([step]), ints([nw]),
           reals(shift), ints([nw, 1]), rec(b""), reals(np.eye(3).ravel(order="F")), rec(b""), rec(b""),
           ints([1]), ints(g.reshape(-1))]
    nst = len(states)
    out += [ints([nst]), ints([s[2] for s in states]), reals([s[0] for s in states]),
            reals([s[1] for s in states]), rec(b""), reals(rho)]
    out += [rec(b"")] * 7 + [ints([nst]), ints(np.arange(1, nst + 1))]
    out += [reals(s[3]) for s in states]
    with open(path, "wb") as fh:
        fh.write(b"".join(out))
