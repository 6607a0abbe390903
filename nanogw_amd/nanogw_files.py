"""NanoGW's text formats on either side of the GW hot path.

Input: the ``nanogw.in`` keys that configure the ISDF-Lanczos Sigma run,
read with the rules of NanoGW's ESDF parser (shared/esdf_module.f90):
  * comments start at '#', ';' or '!' (:394);
  * tokens are separated by ' ', '=' or ':' (:395);
  * labels compare after lower-casing and removing '.' and '-' (esdf_reduce,
    :889-946): ``Lanczos_GW`` matches ``lanczos_gw``, ``lanczos-gw`` reads as
    ``lanczosgw``;
  * the first definition of a label wins (esdf_integer / esdf_physical read
    the first matching record);
  * physical values carry an optional unit (third token), converted with
    NanoGW's unit table (esdf_module.f90:176-320, energy entries here);
  * blocks are ``begin <label>`` ... ``end <label>``; the ``diag`` block
    lists 1-based orbital indices, one per line, or ``range lo hi``
    (esdf_parse_block, :1082-1135).
Keys and defaults follow SIGMA/input_s.F90:74-184 and shared/input_g.f90:
66-94.

V_xc: the diagonal (and off-diagonal) matrix elements of sigma_mtxel.dat
that Zvxc reads when ``read_vxc_matrix_elements`` is set (SIGMA/vxc.F90z:
56-111), list-directed, eV in the file, Ry returned.

Output: the self-energy table of printsigma (SIGMA/printsigma.F90:167-188,
file ``sigma_nostatic_<ik>_<it>``, ``_up_`` / ``_down_`` for spin) and its
``hmat_qp_nostatic`` line per state, with the same Fortran edit
descriptors.
"""
from __future__ import annotations

import dataclasses

import numpy as np

RYD = 13.60569193  # shared/myconstants.F90:41

_COMMENT = ("#", ";", "!")
_DIVIDE = (" ", "=", ":")
# energy units of esdf_module.f90 (phy_n / phy_u, dimension 'e'), in joule
_ENERGY_J = {"j": 1.0, "erg": 1.0e-7, "ev": 1.60219e-19, "mev": 1.60219e-22, "ry": 2.17991e-18,
             "mry": 2.17991e-21, "hartree": 4.35982e-18, "kcal/mol": 6.94780e-21, "mhartree": 4.35982e-21,
             "kj/mol": 1.6606e-21}


def esdf_reduce(s: str) -> str:
    """Lower case, '.' and '-' removed, blanks trimmed (esdf_reduce)."""
    return s.lower().replace(".", "").replace("-", "").strip()


def _tokens(line: str):
    out, cur = [], ""
    for ch in line:
        if ch in _DIVIDE:
            if cur:
                out.append(cur)
            cur = ""
        else:
            cur += ch
    if cur:
        out.append(cur)
    return out


class EsdfFile:
    """Records and tokens of an ESDF input file."""

    def __init__(self, text: str):
        self.records = []
        for raw in text.splitlines():
            cut = len(raw)
            for c in _COMMENT:
                i = raw.find(c)
                if i >= 0:
                    cut = min(cut, i)
            line = raw[:cut].strip()
            if line:
                self.records.append(line)
        self.tokens = [_tokens(r) for r in self.records]

    def _first(self, label):
        key = esdf_reduce(label)
        inblock = False
        for t in self.tokens:
            head = esdf_reduce(t[0])
            if head == "begin":
                inblock = True
                continue
            if head == "end":
                inblock = False
                continue
            if not inblock and head == key:
                return t
        return None

    def defined(self, label) -> bool:
        return self._first(label) is not None or self.block(label) is not None

    def integer(self, label, default: int) -> int:
        t = self._first(label)
        if t is None or len(t) < 2:
            return default
        return int(t[1])

    def double(self, label, default: float) -> float:
        t = self._first(label)
        if t is None or len(t) < 2:
            return default
        return float(t[1].lower().replace("d", "e"))

    def string(self, label, default: str) -> str:
        t = self._first(label)
        if t is None or len(t) < 2:
            return default
        return t[1]

    def physical(self, label, default: float, unit: str) -> float:
        """Value converted to `unit`; the default is already in `unit`."""
        t = self._first(label)
        if t is None or len(t) < 2:
            return default
        val = float(t[1].lower().replace("d", "e"))
        src = esdf_reduce(t[2]) if len(t) > 2 else esdf_reduce(unit)
        dst = esdf_reduce(unit)
        if src not in _ENERGY_J or dst not in _ENERGY_J:
            raise ValueError(f"units not recognised: {src!r} -> {dst!r}")
        return val * _ENERGY_J[src] / _ENERGY_J[dst]

    def block(self, label):
        """Lines of block `label`, or None when absent (esdf_block)."""
        key = esdf_reduce(label)
        for i, t in enumerate(self.tokens):
            if esdf_reduce(t[0]) == "begin" and len(t) > 1 and esdf_reduce(t[1]) == key:
                lines = []
                for j in range(i + 1, len(self.tokens)):
                    if esdf_reduce(self.tokens[j][0]) == "end":
                        if len(self.tokens[j]) < 2 or esdf_reduce(self.tokens[j][1]) != key:
                            raise ValueError(f'Block "{key}" not closed correctly')
                        return lines
                    lines.append(self.records[j])
                raise ValueError(f'Block "{key}" not closed correctly')
        return None


def parse_index_block(lines):
    """Integer list of a block: one value per line or 'range lo hi'
    (esdf_parse_block)."""
    out = []
    for line in lines:
        t = line.split()
        if esdf_reduce(t[0]) == "range":
            lo, hi = int(t[1]), int(t[2])
            out.extend(range(lo, hi + 1))
        else:
            out.append(int(t[0]))
    return out


@dataclasses.dataclass
class SigmaInput:
    """The nanogw.in settings of an ISDF-Lanczos Sigma run (energies in Ry)."""
    lanczos_gw: bool = False
    doisdf: bool = False
    lessmemory: int = 0
    num_isdf_points: int = -1
    intp_type: int = 1
    tamm_dancoff: bool = False
    lanczos_niter: int = 9
    lanczos_npoly: int = 12
    jnblock: int = 400
    deltae: float = 20.0 / RYD           # energy_range
    nen: int = 21                        # energy_data_points
    ecuts: float = 0.005                 # dynamic_energy_resolution
    max_number_states: int = -1
    qp_min: int = -1
    qp_max: int = -1
    diag: list = dataclasses.field(default_factory=list)  # 1-based orbitals
    read_vxc: bool = False
    exchange_correlation: str = "gw"
    scissors: list = dataclasses.field(default_factory=list)  # lines of the 'scissors' block


def read_nanogw_in(path_or_text, is_text=False) -> SigmaInput:
    """Settings of SIGMA/input_s.F90:74-184 and shared/input_g.f90:66-94 that
    drive the Lanczos path."""
    text = path_or_text if is_text else open(path_or_text).read()
    f = EsdfFile(text)
    s = SigmaInput()
    s.lanczos_gw = f.defined("lanczos_gw")
    s.doisdf = f.defined("doisdf")
    s.lessmemory = f.integer("lessmemory", 0)
    s.num_isdf_points = f.integer("num_isdf_points", -1)
    s.intp_type = f.integer("intp_type", 1)
    s.tamm_dancoff = f.defined("tamm_dancoff")
    s.lanczos_niter = f.integer("lanczos_niter", 9)
    s.lanczos_npoly = f.integer("lanczos_npoly", 12)
    s.jnblock = f.integer("jnblock", 400)
    s.deltae = f.physical("energy_range", 20.0, "eV") / RYD
    s.nen = f.integer("energy_data_points", 21)
    s.ecuts = f.physical("dynamic_energy_resolution", 0.005, "Ry")
    s.max_number_states = f.integer("max_number_states", -1)
    s.qp_min = f.integer("qp_min", -1)
    s.qp_max = f.integer("qp_max", -1)
    blk = f.block("diag")
    s.diag = parse_index_block(blk) if blk is not None else []
    s.read_vxc = f.defined("read_vxc_matrix_elements")
    s.exchange_correlation = esdf_reduce(f.string("exchange_correlation", "gw"))
    sc = f.block("scissors")
    s.scissors = list(sc) if sc is not None else []
    return s


def _list_directed_records(text):
    """Token lists of the records of a list-directed input file: values are
    separated by blanks or commas, 'd' exponents read as 'e', and r*c repeats
    expand to r copies of c (Fortran list-directed input)."""
    recs = []
    for line in text.splitlines():
        out = []
        for tok in line.replace(",", " ").split():
            if tok.startswith("/"):
                break
            if "*" in tok:
                r, c = tok.split("*", 1)
                out.extend([c] * int(r))
            else:
                out.append(tok)
        recs.append(out)
    return recs


def read_vxc_matrix_elements(path_or_text, ik, isp, diag_bands, offd_pairs=(), nstate=None, cplx=False,
                             is_text=False):
    """V_xc matrix elements from sigma_mtxel.dat, as Zvxc reads them with
    read_vxc_matrix_elements set (SIGMA/vxc.F90z:56-111): every statement
    ``read (23, *) ikp, jsp, ii, jj, re [, im]`` takes the next 5 (6 for
    complex wave functions) values, starting a new record and continuing
    over records as list-directed input does; lines of other k-points or
    spins are skipped; ii == jj is a diagonal element, stored when band ii
    is one of diag_bands; other pairs are stored when (ii, jj) is one of
    offd_pairs; a later line overwrites an earlier one.  Values are eV in
    the file and returned in Ry (/ ryd).  Bands are 1-based.  An element
    that stays undefined raises, as Zvxc dies ('undefined Vxc matrix
    elements').  Returns (vxc_diag, vxc_offd) complex arrays in the order
    of diag_bands and offd_pairs."""
    text = path_or_text if is_text else open(path_or_text).read()
    big = 9.0e9
    diag = np.full(len(diag_bands), big + 0j)
    offd = np.full(len(offd_pairs), big + 0j)
    inv_diag = {int(b): i for i, b in enumerate(diag_bands)}
    inv_offd = {(int(a), int(b)): i for i, (a, b) in enumerate(offd_pairs)}
    nval = 6 if cplx else 5
    recs = _list_directed_records(text)
    ri = 0
    while ri < len(recs):
        vals = []
        while len(vals) < nval and ri < len(recs):
            vals.extend(recs[ri])
            ri += 1
        if len(vals) < nval:
            break  # end of file inside a statement ends the loop (end=10)
        ikp, jsp, ii, jj = (int(v) for v in vals[:4])
        re = float(vals[4].lower().replace("d", "e"))
        im = float(vals[5].lower().replace("d", "e")) if cplx else 0.0
        if ikp != ik or jsp != isp:
            continue
        if ii == jj:
            if ii in inv_diag:
                diag[inv_diag[ii]] = complex(re, im) / RYD
        elif (ii, jj) in inv_offd:
            offd[inv_offd[(ii, jj)]] = complex(re, im) / RYD
    undefined = int(np.sum(diag.real == big) + np.sum(offd.real == big))
    if undefined:
        raise ValueError(f"ERROR! {undefined} undefined Vxc matrix elements for spin channel {isp}")
    return diag, offd


def _fortran_f(x: float, w: int, d: int) -> str:
    """Fortran Fw.d edit descriptor: asterisks when the value does not fit."""
    s = f"{x:{w}.{d}f}"
    if len(s) > w:
        if s.startswith("-0.") and len(s) == w + 1:
            return "-" + s[2:]  # Fortran may drop the leading zero
        if s.startswith("0.") and len(s) == w + 1:
            return s[1:]
        return "*" * w
    return s


def _fortran_i(n: int, w: int) -> str:
    s = f"{n:{w}d}"
    return "*" * w if len(s) > w else s


def sigma_file_name(ik: int, it_scf: int, spin: str = "") -> str:
    """sigma_nostatic_<ik:3>_<it:4> (calculate_sigma_lanczos.F90z:429-444);
    spin '' / 'up' / 'down'."""
    mid = f"{spin}_" if spin else ""
    return f"sigma_nostatic_{mid}{ik:03d}_{it_scf:04d}"


SIGMA_HEADER = [
    "-" * 65,
    "",
    " OUTPUT SELF-ENERGY DIAGONAL MATRIX ELEMENTS (eV)",
    "     n     occ       E_dft        V_xc     Sigma_x    Sigma_sx     Sigma_c     Sigma_g  Sigma-V_xc"
    "        zfac   Im[Sigma]         E_0        E_qp",
]


def sigma_row(n, occ, edft, vxc, sigx, sigsx, sigc, sigg, sig_m_vxc, zfac, imsig, e0, eqp) -> str:
    """One row, '(i6,f8.3,11(1x,f11.5))' (printsigma.F90:181-186)."""
    vals = (edft, vxc, sigx, sigsx, sigc, sigg, sig_m_vxc, zfac, imsig, e0, eqp)
    return _fortran_i(n, 6) + _fortran_f(occ, 8, 3) + "".join(" " + _fortran_f(v, 11, 5) for v in vals)


def hmat_row(n, eqp, isp, ik, edft, e0) -> str:
    """hmat_qp_nostatic line, '(2i8,f10.3,2i4,2f10.3)' (printsigma.F90:187-188)."""
    return (_fortran_i(n, 8) + _fortran_i(n, 8) + _fortran_f(eqp, 10, 3) + _fortran_i(isp, 4) + _fortran_i(ik, 4) +
            _fortran_f(edft, 10, 3) + _fortran_f(e0, 10, 3))


def write_sigma_table(path, rows):
    """Write the diagonal self-energy table; rows are dicts with the
    keys of sigma_row (energies in eV, n 1-based)."""
    lines = list(SIGMA_HEADER)
    lines += [sigma_row(**r) for r in rows]
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


def qp_rows(qp_bands, occ, edft_ev, vxc_ev, sigx_ev, newton):
    """Rows of the table for the Lanczos GW path from ngw_qp_newton's
    outputs, as printsigma fills them (printsigma.F90:97-128, 181-186):
    sigma0 = Sigma_x + Sigma_c(E_dft) (+ Sigma_g, zero here: no vertex), so
    the Sigma_c column is sigma0 - Sigma_x; Sigma_sx is zero (full-frequency
    Sigma_c); E_0 = E_dft (one-shot)."""
    rows = []
    for i, b in enumerate(qp_bands):
        s0 = complex(newton["sigma0"][i]).real
        rows.append({"n": int(b) + 1, "occ": float(occ[i]), "edft": float(edft_ev[i]), "vxc": float(vxc_ev[i]),
                     "sigx": float(sigx_ev[i]), "sigsx": 0.0, "sigc": s0 - float(sigx_ev[i]), "sigg": 0.0,
                     "sig_m_vxc": s0 - float(vxc_ev[i]), "zfac": float(newton["zfac"][i]),
                     "imsig": float(newton["imsig"][i]), "e0": float(edft_ev[i]), "eqp": float(newton["eqp"][i])})
    return rows
