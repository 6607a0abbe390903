"""C-ABI checks that need no GPU: the library loads, exports every symbol the
header declares, and its host-only entry points behave like NanoGW's."""
import ctypes as C
import subprocess

import numpy as np
import pytest

from conftest import rel_err


def test_library_exports_every_declared_symbol():
    from nanogw_amd import _lib

    lib = _lib.load()
    declared = _lib.declared_symbols()
    assert len(declared) >= 18
    for name in declared:
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    assert set(declared) <= exported
    assert set(_lib._SIGS) == set(declared)


def test_version_and_error_text():
    from nanogw_amd import _lib

    lib = _lib.load()
    assert b"gfx950" in lib.ngw_version()
    st = lib.ngw_polyfit_sqrt(1e-6, 1.0, 4, 20, 0, (C.c_double * 5)(), None)
    assert st == -5
    assert b"xmin" in lib.ngw_last_error()


def test_polyfit_matches_oracle():
    from nanogw_amd.lanczos import polyfit_sqrt
    from oracle import oracle as O

    for npoly, nx, cheb in [(2, 20, False), (6, 20, False), (12, 20, False), (12, 300, True)]:
        b1, f1 = polyfit_sqrt(0.078, 38.47, npoly, nx, cheb)
        b2, f2 = O.polyfit_sqrt(0.078, 38.47, npoly, nx, cheb)
        assert rel_err(b1, b2) < 1e-12
        assert abs(f1 - f2) < 1e-12


def test_null_context_rejected():
    from nanogw_amd import _lib

    lib = _lib.load()
    assert lib.ngw_synchronize(None) == -1
    assert lib.ngw_matvec_isdf_block(None, None, None, 1, 1, 0) == -1
    assert lib.ngw_set_precision(None, 0) == -1
    assert lib.ngw_set_explicit_c(None, 4, 4, None, 4, None, 4, None) == -1


def test_header_documents_reference_interfaces():
    from nanogw_amd import _lib

    text = open(_lib.HEADER).read()
    for ref in ["apply_matvec_BLOCK.F90:25", "apply_matvec_BLOCK.F90:210",
                "lanczos_spectra_BLOCK.F90:7", "calculate_sigma_lanczos_BLOCK.F90z:201-444",
                "printsigma.F90:97-128", "polyfit.f90:1", "isdf_parallel_sym_UltraLowMem.F90",
                "fft_module.F90z:122-177", "apply_matvec.f90:25", "lanczos_spectra.f90:7"]:
        assert ref in text, ref
