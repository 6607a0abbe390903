"""The Fortran drop-in (nanogw_amd/fortran/ngw_dropin.F90) compiled the way
NanoGW would build it: without NGW_DROPIN_SELFTEST, against NanoGW's own
modules myconstants (shared/myconstants.F90), typedefs (shared/typedefs.F90)
and mpi_module (shared/mpi_module.F90z preprocessed without MPI, real pass
then complex pass, as NanoGW's build concatenates them).  Component names,
ranks and kinds of type(ISDF), type(wavefunction), type(gspace), peinf and
w_grp are then NanoGW's, not the stand-ins'.  The object must export the
linker symbols NanoGW's drivers call (lower case, trailing underscore).

This needs the NanoGW sources; it is skipped where they are not present
(the reference is read in place, nothing of it is kept)."""
import os
import shutil
import subprocess

import pytest

REF = "/root/reference/shared"
HERE = os.path.dirname(os.path.abspath(__file__))
FORTRAN = os.path.join(os.path.dirname(HERE), "nanogw_amd", "fortran")

SYMBOLS = ["matvec_isdf_block_", "polynomial_matvec_isdf_block_", "lanczos_spectra_isdf_block_",
           "matvec_isdf_", "polynomial_matvec_isdf_", "lanczos_spectra_isdf_",
           "matvec_isdf_ultralowmem_", "polynomial_matvec_isdf_ultralowmem_", "lanczos_spectra_isdf_ultralowmem_",
           "matvec_isdf_lowcomsym_", "polynomial_matvec_isdf_lowcomsym_", "lanczos_spectra_isdf_lowcomsym_"]


def test_dropin_compiles_against_nanogw_modules(tmp_path):
    fc = os.environ.get("NGW_FC", "/opt/rocm/bin/amdflang")
    if not os.path.exists(fc):
        pytest.skip("amdflang not available")
    if not all(os.path.exists(os.path.join(REF, f)) for f in ("typedefs.F90", "myconstants.F90", "mpi_module.F90z")):
        pytest.skip("NanoGW sources not present")
    cpp = shutil.which("cpp")
    if cpp is None:
        pytest.skip("cpp not available")
    t = str(tmp_path)

    def run(cmd, **kw):
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=300, **kw)
        assert r.returncode == 0, " ".join(cmd) + "\n" + r.stdout + r.stderr
        return r.stdout

    run([fc, "-c", os.path.join(REF, "myconstants.F90"), f"-J{t}", "-o", os.path.join(t, "myconstants.o")])
    run([fc, "-cpp", "-c", os.path.join(REF, "typedefs.F90"), f"-J{t}", "-o", os.path.join(t, "typedefs.o")])
    pp = run([cpp, "-traditional-cpp", "-P", f"-I{REF}", os.path.join(REF, "mpi_module.F90z")])
    pp += run([cpp, "-traditional-cpp", "-P", "-DCPLX", f"-I{REF}", os.path.join(REF, "mpi_module.F90z")])
    (tmp_path / "mpi_module_pp.f90").write_text(pp)
    run([fc, "-c", os.path.join(t, "mpi_module_pp.f90"), f"-J{t}", "-o", os.path.join(t, "mpi_module.o")])
    run([fc, "-O2", f"-J{t}", "-c", os.path.join(FORTRAN, "ngw_c.f90"), "-o", os.path.join(t, "ngw_c.o")])
    run([fc, "-O2", "-cpp", f"-I{t}", f"-J{t}", "-c", os.path.join(FORTRAN, "ngw_dropin.F90"), "-o",
         os.path.join(t, "dropin.o")])
    nm = shutil.which("nm")
    if nm is None:
        return
    out = run([nm, os.path.join(t, "dropin.o")])
    defined = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    missing = [s for s in SYMBOLS if s not in defined]
    assert not missing, missing
