"""nanogw.in reading (ESDF rules of shared/esdf_module.f90) and the
printsigma table / hmat_qp_nostatic writers (SIGMA/printsigma.F90:167-188)."""
import numpy as np
import pytest

from nanogw_amd import nanogw_files as F

SAMPLE = """
# ISDF Lanczos GW run
Lanczos_GW
doisdf                      ! ISDF on
lessmemory = 3
num_isdf_points: 480
intp_type 1
lanczos_niter   200
LANCZOS_NPOLY   10          ; labels are case-insensitive
jnblock 64
energy_range 54.4228 eV
energy_data_points 200
dynamic_energy_resolution 0.01 hartree
begin diag
  range 7 9
  10
end diag
lanczos_niter 15            # second definition: the first one is used
"""


def test_read_nanogw_in_keys_units_blocks():
    s = F.read_nanogw_in(SAMPLE, is_text=True)
    assert s.lanczos_gw and s.doisdf and not s.tamm_dancoff
    assert s.lessmemory == 3 and s.num_isdf_points == 480 and s.intp_type == 1
    assert s.lanczos_niter == 200 and s.lanczos_npoly == 10 and s.jnblock == 64
    assert s.nen == 200
    assert s.deltae == pytest.approx(54.4228 / F.RYD, rel=1e-12)
    # 0.01 Ha in Ry with the ESDF unit table (4.35982e-18 J / 2.17991e-18 J)
    assert s.ecuts == pytest.approx(0.01 * 4.35982e-18 / 2.17991e-18, rel=1e-14)
    assert s.diag == [7, 8, 9, 10]
    d = F.read_nanogw_in("", is_text=True)
    assert (d.lanczos_niter, d.lanczos_npoly, d.jnblock, d.nen) == (9, 12, 400, 21)
    assert d.deltae == pytest.approx(20.0 / F.RYD) and d.ecuts == 0.005 and d.diag == []


def test_esdf_label_reduction_and_errors():
    assert F.esdf_reduce(" Lanczos-GW. ") == "lanczosgw"
    assert F.esdf_reduce("Lanczos_GW") == "lanczos_gw"
    with pytest.raises(ValueError):
        F.read_nanogw_in("begin diag\n 3\n", is_text=True)
    with pytest.raises(ValueError):
        F.read_nanogw_in("energy_range 3 furlongs\n", is_text=True)


def test_sigma_row_and_hmat_formats():
    # (i6,f8.3,11(1x,f11.5)): 6 + 8 + 11 x 12 characters
    row = F.sigma_row(100, 2.0, -12.448464, 0.0, -14.1, 0.0, 1.2345678, 0.0, -12.8654322, 0.87654, -0.0123,
                      -12.448464, -12.45)
    assert len(row) == 6 + 8 + 11 * 12
    assert row[:14] == "   100   2.000"
    assert row[14:26] == "   -12.44846"
    assert row.endswith("   -12.45000")
    # (2i8,f10.3,2i4,2f10.3)
    h = F.hmat_row(100, -12.45, 1, 1, -12.448464, -12.448464)
    assert h == "     100     100   -12.450   1   1   -12.448   -12.448"
    # a value that does not fit its field prints as asterisks, like Fortran
    assert F.sigma_row(1, 0.0, 1e9, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0)[14:26] == " " + "*" * 11
    assert F.sigma_file_name(1, 0) == "sigma_nostatic_001_0000"
    assert F.sigma_file_name(2, 13, "up") == "sigma_nostatic_up_002_0013"


def test_qp_rows_follow_printsigma(tmp_path):
    newton = {"eqp": np.array([-16.76, -12.45]), "zfac": np.array([0.9, 0.8]),
              "sigma0": np.array([-15.0 + 0.1j, -13.0 + 0.0j]), "imsig": np.array([0.01, 0.02])}
    rows = F.qp_rows([99, 100], [2.0, 0.0], [-16.0, -12.0], [-10.0, -9.0], [-14.0, -11.0], newton)
    assert rows[0]["n"] == 100 and rows[1]["n"] == 101
    assert rows[0]["sigc"] == pytest.approx(-1.0) and rows[0]["sig_m_vxc"] == pytest.approx(-5.0)
    path = tmp_path / F.sigma_file_name(1, 0)
    F.write_sigma_table(path, rows)
    lines = path.read_text().splitlines()
    assert lines[:4] == F.SIGMA_HEADER and len(lines) == 6


FMT_PROGRAM = """program fmt_check
  implicit none
  integer :: n, i, nn, ik, isp
  real(8) :: v(12), h(3)
  character(len=512) :: fin
  call get_command_argument(1, fin)
  open (10, file=trim(fin), status='old')
  read (10, *) n
  do i = 1, n
    read (10, *) nn, v
    write (*, '(i6,f8.3,11(1x,f11.5))') nn, v
  end do
  read (10, *) n
  do i = 1, n
    read (10, *) nn, h(1), isp, ik, h(2), h(3)
    write (*, '(2i8,f10.3,2i4,2f10.3)') nn, nn, h(1), isp, ik, h(2), h(3)
  end do
end program fmt_check
"""


def test_formats_match_a_fortran_runtime(tmp_path):
    """sigma_row / hmat_row against the same edit descriptors written by a
    Fortran runtime (AMD flang), including rounding, negative values that
    round to zero and fields that overflow to asterisks."""
    import os
    import subprocess

    fc = os.environ.get("NGW_FC", "/opt/rocm/bin/amdflang")
    if not os.path.exists(fc):
        pytest.skip("no Fortran compiler")
    src, exe, data = tmp_path / "fmt.f90", tmp_path / "fmt", tmp_path / "in.txt"
    src.write_text(FMT_PROGRAM)
    subprocess.run([fc, "-O0", "-o", str(exe), str(src)], check=True, cwd=tmp_path, timeout=120)
    rng = np.random.default_rng(5)
    rows = []
    for k in range(40):
        vals = list(rng.normal(scale=10.0 ** rng.integers(-6, 6), size=11))
        rows.append((int(rng.integers(1, 99999)), float(rng.choice([0.0, 1.0, 2.0, 1.9995])), vals))
    rows.append((7, 2.0, [-0.0000049, 0.0000049, 99999.999994, -9999.999996, 123456.0, -1e-12, 0.5, -0.5,
                          1.234565, -1.234565, 0.0]))
    hm = [(int(rng.integers(1, 9999)), float(rng.normal(scale=30.0)), 1 + k % 2, 1 + k % 3,
           float(rng.normal(scale=30.0)), float(rng.normal(scale=30.0))) for k in range(20)]
    hm.append((12, -0.0004, 1, 1, 999999.9996, -99999.9996))
    with open(data, "w") as f:
        f.write(f"{len(rows)}\n")
        for n, occ, vals in rows:
            f.write(" ".join([str(n), repr(occ)] + [repr(float(v)) for v in vals]) + "\n")
        f.write(f"{len(hm)}\n")
        for n, eqp, isp, ik, edft, e0 in hm:
            f.write(f"{n} {eqp!r} {isp} {ik} {edft!r} {e0!r}\n")
    out = subprocess.run([str(exe), str(data)], capture_output=True, text=True, timeout=60, check=True).stdout
    lines = out.splitlines()
    assert len(lines) == len(rows) + len(hm)
    for line, (n, occ, vals) in zip(lines, rows):
        assert F.sigma_row(n, occ, *vals) == line, (F.sigma_row(n, occ, *vals), line)
    for line, (n, eqp, isp, ik, edft, e0) in zip(lines[len(rows):], hm):
        assert F.hmat_row(n, eqp, isp, ik, edft, e0) == line, (F.hmat_row(n, eqp, isp, ik, edft, e0), line)


def test_read_vxc_matrix_elements_list_directed():
    """sigma_mtxel.dat as Zvxc reads it (SIGMA/vxc.F90z:56-111): 5 values
    per read statement (6 for complex wave functions), continuing over
    records, the rest of the last record skipped; other k-points and spins
    ignored; later lines overwrite; eV -> Ry; undefined elements raise."""
    text = "\n".join([
        "1 1 3 3 -10.5",
        "1 1 4 4 -9.25d0 ignored trailing values",
        "2 1 3 3 -99.0",          # another k-point
        "1 2 3 3 -98.0",          # another spin
        "1, 1, 5, 5,",            # one statement over two records
        "  -8.0",
        "",
        "1 1 3 4 0.75",           # off-diagonal
        "1 1 4 4 -9.5",           # overwrites band 4
        "1 1 2*6 -7.0",           # repeat count: ii = jj = 6
    ])
    d, o = F.read_vxc_matrix_elements(text, 1, 1, [3, 4, 5, 6], [(3, 4)], is_text=True)
    assert np.allclose(d.real * F.RYD, [-10.5, -9.5, -8.0, -7.0]) and np.all(d.imag == 0)
    assert np.allclose(o.real * F.RYD, [0.75])
    with pytest.raises(ValueError, match="1 undefined"):
        F.read_vxc_matrix_elements(text, 1, 1, [3, 7], is_text=True)
    dc, _ = F.read_vxc_matrix_elements("1 1 2 2 -4.0 0.5\n", 1, 1, [2], cplx=True, is_text=True)
    assert np.allclose(dc * F.RYD, [-4.0 + 0.5j])


def test_scissors_block_of_nanogw_in():
    s = F.read_nanogw_in("begin scissors\n1 1 4 0.5 -10.0 0.1\nend scissors\n", is_text=True)
    assert s.scissors == ["1 1 4 0.5 -10.0 0.1"]
