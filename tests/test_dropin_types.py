"""The Fortran drop-in against NanoGW's own type modules.

`make -C oracle dropin-types` compiles the reference's shared/myconstants.F90,
shared/mpi_module.F90z and shared/typedefs.F90 from their sources (serial
build, the .F90z file preprocessed twice as SIGMA/Makefile:107-110 does) and
then nanogw_amd/fortran/ngw_dropin.F90 against them, so the component names,
ranks and kinds the drop-in uses (ISDF%Psi_intp_loc, Mmtrx_loc, n_intp_r,
PsiV_intp_bl, lvrep_bound, w_grp%npes, gvec%syms%prod, ...; typedefs.F90:24-72,
mpi_module.F90z:20-85) are checked by the compiler, and the object exports the
linker symbols NanoGW's drivers call.  A copy of the drop-in with
one component name changed must then fail to compile against the same
modules, which shows the check reads the real types.  Skipped where the
reference sources or the Fortran compiler are absent."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")
FC = os.environ.get("NGW_FC", "/opt/rocm/bin/amdflang")
DROPIN = os.path.join(ROOT, "nanogw_amd", "fortran", "ngw_dropin.F90")
REFMODS = os.path.join(ORACLE, "_ref", "fortran")


def _have_reference():
    if not (os.path.exists(FC) and shutil.which("cpp") and shutil.which("make")):
        return False
    r = subprocess.run(["make", "-s", "-C", ORACLE, "ref-present"], capture_output=True, text=True)
    return r.stdout.strip() == "yes"


pytestmark = pytest.mark.skipif(not _have_reference(),
                                reason="reference type modules or amdflang not available")


# linker symbols NanoGW's drivers call (lower case, trailing underscore)
SYMBOLS = ["matvec_isdf_block_", "polynomial_matvec_isdf_block_", "lanczos_spectra_isdf_block_",
           "matvec_isdf_", "polynomial_matvec_isdf_", "lanczos_spectra_isdf_",
           "matvec_isdf_ultralowmem_", "polynomial_matvec_isdf_ultralowmem_", "lanczos_spectra_isdf_ultralowmem_",
           "matvec_isdf_lowcomsym_", "polynomial_matvec_isdf_lowcomsym_", "lanczos_spectra_isdf_lowcomsym_"]


def _compile(src, outdir):
    return subprocess.run([FC, "-O2", "-cpp", f"-I{REFMODS}", f"-J{outdir}", "-c", src,
                           "-o", os.path.join(outdir, "dropin.o")],
                          capture_output=True, text=True)


def test_dropin_compiles_against_reference_types(tmp_path):
    r = subprocess.run(["make", "-B", "-C", ORACLE, "dropin-types"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert os.path.exists(os.path.join(REFMODS, "ngw_dropin.o"))
    for mod in ("typedefs.mod", "mpi_module.mod", "ngw_c.mod"):
        assert os.path.exists(os.path.join(REFMODS, mod))
    # the unmodified drop-in compiles from a second directory too
    assert _compile(DROPIN, str(tmp_path)).returncode == 0
    if shutil.which("nm"):
        out = subprocess.run(["nm", os.path.join(REFMODS, "ngw_dropin.o")], capture_output=True, text=True,
                             check=True).stdout
        defined = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
        assert not [sym for sym in SYMBOLS if sym not in defined]


@pytest.mark.parametrize("old,new", [("%Mmtrx_loc", "%Mmtrx_lox"), ("%lvrep_bound", "%lvrep_bnd")])
def test_changed_component_fails(tmp_path, old, new):
    if not os.path.exists(os.path.join(REFMODS, "typedefs.mod")):
        subprocess.run(["make", "-C", ORACLE, "dropin-types"], check=True, capture_output=True)
    with open(DROPIN) as f:
        src = f.read()
    assert old in src
    bad = tmp_path / "ngw_dropin.F90"
    bad.write_text(src.replace(old, new, 1))
    r = _compile(str(bad), str(tmp_path))
    assert r.returncode != 0
    assert new[1:].lower() in (r.stdout + r.stderr).lower()
