"""The parsec.dat reader (nanogw_amd/parsec_dat.py; shared/parsec_wfn.F90)
on files written by a Fortran runtime (AMD flang's sequential unformatted
output, record for record in the order parsec_wfn reads them) and by the
module's own writer: grid, energies, occupations, density order, the FFT
box, hcub, the Coulomb truncation, the normalisation check, kept orbitals,
the scissors operator and the wfn.dat layout."""
import os
import struct
import subprocess

import numpy as np
import pytest

from nanogw_amd import parsec_dat as P

FORTRAN_WRITER = r"""
program write_parsec
  implicit none
  integer, parameter :: dp = kind(1.d0)
  integer :: u, i, j, k, n, nw, nst, ii
  integer, allocatable :: g(:, :), irep(:), iord(:)
  real(dp), allocatable :: rho(:), e0(:), occ(:), wf(:, :)
  real(dp) :: h, x, y, z, r2, trans(3, 3, 1), nrm
  character(len=26) :: date
  h = 0.5d0
  n = 0
  allocate (g(3, 2000))
  do k = -4, 4
    do j = -4, 4
      do i = -4, 4
        if ((i*i + j*j + k*k)*h*h <= 2.0d0**2 + 1d-12) then
          n = n + 1
          g(:, n) = [i, j, k]
        end if
      end do
    end do
  end do
  nw = n
  nst = 5
  allocate (rho(nw), e0(nst), occ(nst), irep(nst), iord(nst), wf(nw, nst))
  do ii = 1, nst
    e0(ii) = -0.8d0 + 0.3d0*ii
    occ(ii) = merge(2.d0, 0.d0, ii <= 2)
    irep(ii) = 1
    iord(ii) = nst + 1 - ii
  end do
  do i = 1, nw
    x = g(1, i)*h; y = g(2, i)*h; z = g(3, i)*h
    r2 = x*x + y*y + z*z
    do ii = 1, nst
      wf(i, ii) = exp(-0.5d0*r2*ii)*(1.d0 + 0.1d0*ii*x - 0.05d0*y*z)
    end do
  end do
  do ii = 1, nst
    nrm = sqrt(sum(wf(:, ii)**2))
    if (ii /= 4) wf(:, ii) = wf(:, ii)/nrm
  end do
  rho = 2.d0*(wf(:, 1)**2 + wf(:, 2)**2)
  trans = 0.d0
  trans(1, 1, 1) = 1.d0; trans(2, 2, 1) = 1.d0; trans(3, 3, 1) = 1.d0
  date = 'Tue Dec  1 12:00:00 2026  '
  open (newunit=u, file='parsec.dat', form='unformatted', access='sequential', status='replace')
  write (u) date
  write (u) 1, 0, 0
  write (u) h
  write (u) nw
  write (u) 0.d0, 0.d0, 0.d0
  write (u) nw, 1
  write (u) 'symmetry operations'
  write (u) trans
  write (u) 'x'
  write (u) 'y'
  write (u) 1
  write (u) ((g(i, j), i=1, 3), j=1, nw)
  write (u) nst
  write (u) irep
  write (u) e0
  write (u) occ
  write (u) 'density'
  write (u) rho
  do i = 1, 7
    write (u) i
  end do
  write (u) nst
  write (u) iord
  do ii = 1, nst
    write (u) wf(:, iord(ii))
  end do
  close (u)
end program write_parsec
"""


def _fortran_file(tmp_path):
    fc = os.environ.get("NGW_FC", "/opt/rocm/bin/amdflang")
    if not os.path.exists(fc):
        pytest.skip("amdflang not available")
    src = tmp_path / "w.f90"
    src.write_text(FORTRAN_WRITER)
    subprocess.run([fc, "-O0", "-o", str(tmp_path / "w"), str(src)], check=True, cwd=tmp_path, timeout=120)
    subprocess.run([str(tmp_path / "w")], check=True, cwd=tmp_path, timeout=60)
    return tmp_path / "parsec.dat"


def _expected(h=0.5):
    g = []
    for k in range(-4, 5):
        for j in range(-4, 5):
            for i in range(-4, 5):
                if (i * i + j * j + k * k) * h * h <= 4.0 + 1e-12:
                    g.append((i, j, k))
    g = np.array(g)
    x, y, z = (g[:, c] * h for c in range(3))
    r2 = x * x + y * y + z * z
    wf = np.stack([np.exp(-0.5 * r2 * ii) * (1 + 0.1 * ii * x - 0.05 * y * z) for ii in range(1, 6)], 1)
    for ii in range(5):
        if ii != 3:
            wf[:, ii] /= np.linalg.norm(wf[:, ii])
    return g, wf, 2.0 * (wf[:, 0] ** 2 + wf[:, 1] ** 2)


def test_reads_fortran_written_parsec_dat(tmp_path):
    d = P.read_parsec_dat(_fortran_file(tmp_path))
    g, wf, rho = _expected()
    assert d.datelabel.startswith("Tue Dec  1") and not d.wfn_dat
    assert (d.nspin, d.lcplx, d.per, d.ntrans, d.nwedge, d.ndim) == (1, False, 0, 1, len(g), len(g))
    assert np.array_equal(d.gtable, g)
    st = d.wfn[(1, 1)]
    assert st.nstate == 5 and np.allclose(st.e0, -0.8 + 0.3 * np.arange(1, 6)) and np.array_equal(st.e1, st.e0)
    assert np.array_equal(st.occ0, [2, 2, 0, 0, 0])
    # decreasing density order, orbitals reordered with the grid
    assert np.all(np.diff(d.rho[:, 0]) <= 0)
    assert np.array_equal(d.r, g[d.rindex]) and np.allclose(d.rho[:, 0], rho[d.rindex])
    assert np.array_equal(st.states, [1, 2, 3, 4, 5])
    assert np.allclose(st.wf, wf[d.rindex])
    # box: r_max = 2 bohr, nfft = nint(4 r_max / h) made even, hcub, R_c = 2 r_max
    rmax = float(np.sqrt(((g * 0.5) ** 2).sum(1).max()))
    assert abs(d.rmax - rmax) < 1e-12
    n = int(np.rint(4 * rmax / 0.5))
    n += n % 2
    assert d.nfft == (n, n, n, n + 2)
    assert d.hcub == 0.125 and abs(d.trunc - 2 * rmax) < 1e-12
    # orbital 4 is written unnormalised: reported, the others are not
    assert len(d.warnings) == 1 and " 4 " in d.warnings[0]
    assert np.allclose(st.norms[[0, 1, 2, 4]], 1.0)


def test_kept_orbitals_and_scissors(tmp_path):
    path = _fortran_file(tmp_path)
    sc = P.parse_scissors(["1 2 4 0.5 -10.0 0.1", "2 1 5 9.0 0.0 0.0"])
    d = P.read_parsec_dat(path, keep=[5, 2], scissors=sc)
    st = d.wfn[(1, 1)]
    g, wf, _ = _expected()
    assert np.array_equal(st.states, [5, 2]) and np.allclose(st.wf, wf[d.rindex][:, [4, 1]])
    e0 = st.e0
    ref = e0.copy()
    ref[1:4] = e0[1:4] + (e0[1:4] + 10.0 / P.RYD) * 0.1 + 0.5 / P.RYD   # spin-2 row ignored
    assert np.allclose(st.e1, ref)
    with pytest.raises(ValueError, match="requested"):
        P.read_parsec_dat(path, keep=[6])


def test_record_framing_errors(tmp_path):
    raw = _fortran_file(tmp_path).read_bytes()
    with pytest.raises(ValueError):
        P.read_parsec_dat(raw[:-7])
    bad = bytearray(raw)
    struct.pack_into("<i", bad, 4 + 26, 999)  # trailing marker of the first record
    with pytest.raises(ValueError, match="differ"):
        P.read_parsec_dat(bytes(bad))


def test_writer_round_trip_and_symmetry_unfold(tmp_path):
    """The module's writer produces what the reader reads; r_max follows
    unfold (diagonal of each operation, shift) over every operation."""
    from nanogw_amd import synth

    g = synth.sphere_grid(2.2, 0.4)
    rng = np.random.default_rng(3)
    psi = rng.normal(size=(len(g), 3))
    psi /= np.linalg.norm(psi, axis=0)
    states = [(-0.5 + 0.1 * k, 2.0 if k == 0 else 0.0, 1, psi[:, k]) for k in range(3)]
    p = tmp_path / "parsec.dat"
    P.write_parsec_dat(p, "synthetic", 0.4, [0.0, 0.0, 0.0], g, states, 2 * psi[:, 0] ** 2)
    d = P.read_parsec_dat(p)
    assert np.allclose(d.wfn[(1, 1)].wf, psi[d.rindex])
    assert d.nfft[:3] == synth.fft_box(g, 0.4)[1]
    # a mirror operation with a half-step shift moves the unfolded points
    sh = np.array([0.5, 0.0, 0.0])
    tr = np.diag([-1.0, 1.0, 1.0])
    u = P._unfold(g, tr, sh)
    assert np.array_equal(u[:, 0], np.rint(-g[:, 0] - 1.0).astype(int))
    v = (u + sh) * 0.4
    assert np.allclose(np.abs(v[:, 0]), np.abs((g[:, 0] + 0.5) * 0.4))


def test_wfn_dat_layout(tmp_path):
    """The older wfn.dat layout (parsec_wfn.F90:84-106, :216-217, :646-653)."""
    g, wf, rho = _expected()
    e = struct.pack

    def rec(b):
        return e("<i", len(b)) + b + e("<i", len(b))

    nw = len(g)
    blob = rec(b"wfn.dat test label        ")
    blob += rec(e("<ddiiii", 0.5, 0.0, nw, nw, 1, 1))
    blob += rec(np.zeros(3).tobytes())
    blob += rec(b"") + rec(np.eye(3).ravel().tobytes()) + rec(b"") + rec(b"") + rec(np.array([1], "<i4").tobytes())
    blob += rec(g.astype("<i4").reshape(-1).tobytes())
    blob += rec(np.array([5], "<i4").tobytes()) + rec(np.ones(5, "<i4").tobytes())
    blob += rec(np.linspace(-1, 0, 5).tobytes()) + rec(np.array([2, 2, 0, 0, 0], "<f8").tobytes())
    blob += rec(b"") + rec(rho.astype("<f8").tobytes())
    for k in range(5):
        blob += rec(wf[:, k].astype("<f8").tobytes())
    p = tmp_path / "wfn.dat"
    p.write_bytes(blob)
    d = P.read_parsec_dat(p)
    assert d.wfn_dat and d.per == 0 and d.nfft[3] == d.nfft[0] + 2
    assert np.allclose(d.wfn[(1, 1)].wf, wf[d.rindex]) and np.allclose(d.wfn[(1, 1)].e0, np.linspace(-1, 0, 5))
