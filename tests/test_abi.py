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


def test_header_is_plain_c_and_links(tmp_path):
    """include/ngw.h compiles as strict C99 and as C++ (the boundary carries
    only C types), and a C host that takes the address of every declared entry
    point links against libngw.so and runs its host-only calls."""
    import os
    import shutil

    from nanogw_amd import _lib

    if shutil.which("gcc") is None or shutil.which("g++") is None:
        pytest.skip("no host C compiler")
    _lib.load()
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    inc = os.path.join(root, "include")
    names = _lib.declared_symbols()
    src = ['#include <stdio.h>', '#include "ngw.h"', "typedef void (*entry)(void);", "int main(void) {",
           "    entry fp[] = {" + ", ".join(f"(entry)&{n}" for n in names) + "};",
           "    size_t i, n = sizeof fp / sizeof fp[0];",
           "    for (i = 0; i < n; ++i) if (!fp[i]) return 2;",
           '    printf("%s|%d|%s\\n", ngw_version(), ngw_synchronize(0), ngw_last_error());',
           "    return 0;", "}"]
    c = tmp_path / "host.c"
    c.write_text("\n".join(src) + "\n")
    exe = tmp_path / "host"
    libdir = os.path.dirname(_lib.LIB_PATH)
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-Werror", "-pedantic", f"-I{inc}", str(c), "-o", str(exe),
                    f"-L{libdir}", "-lngw", f"-Wl,-rpath,{libdir}"], check=True, capture_output=True, text=True)
    cpp = tmp_path / "host.cpp"
    cpp.write_text('#include "ngw.h"\nint main() { return ngw_version() ? 0 : 1; }\n')
    subprocess.run(["g++", "-std=c++11", "-Wall", "-Werror", f"-I{inc}", "-fsyntax-only", str(cpp)], check=True,
                   capture_output=True, text=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.strip()
    ver, st, err = out.split("|", 2)
    assert "gfx950" in ver
    assert int(st) == -1 and err
