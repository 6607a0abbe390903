"""Builds libngw.so (gfx950 HIP + C++ host) in-tree with hipcc."""
from __future__ import annotations

import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libngw.so")
ARCH = os.environ.get("NGW_OFFLOAD_ARCH", "gfx950")

# Packed FP32 instructions are disabled for every device compile.  MFMA
# accumulators are allowed in VGPRs (gfx90a and later): with accumulators
# pinned to AGPRs the compiler rotates them through v_accvgpr_mov copies in
# the unrolled GEMM loops.
HIPFLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-mllvm", "-amdgpu-mfma-vgpr-form",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
]


def sources():
    src = sorted(glob.glob(os.path.join(HERE, "csrc", "*.hip")))
    src += sorted(glob.glob(os.path.join(HERE, "csrc", "*.cpp")))
    return src


def build(verbose: bool = False, jobs: int = 8) -> str:
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    procs = []
    objs = []
    for s in sources():
        o = os.path.join(objdir, os.path.basename(s) + ".o")
        objs.append(o)
        lang = ["-x", "hip"] if s.endswith(".hip") else ["-x", "hip"]
        cmd = [hipcc, *HIPFLAGS, *lang, "-c", s, "-o", o, "-I", os.path.join(ROOT, "include")]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((s, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        while sum(p.poll() is None for _, p in procs) >= jobs:
            procs[[p.poll() is None for _, p in procs].index(True)][1].wait()
    failed = False
    for s, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        if p.returncode != 0:
            failed = True
            sys.stderr.write(f"--- {os.path.basename(s)} ---\n{text}\n")
        elif verbose and text.strip():
            sys.stderr.write(text)
    if failed:
        raise RuntimeError("libngw build failed")
    tmp = LIB + ".tmp"
    cmd = [hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs,
           "-L/opt/rocm/lib", "-lhipfft", "-lrocsolver", "-lrocblas", "-Wl,-rpath,/opt/rocm/lib"]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


FORTRAN_DIR = os.path.join(HERE, "fortran")
DROPIN_CHECK = os.path.join(FORTRAN_DIR, "_bin", "ngw_dropin_check")


def build_fortran(verbose: bool = False):
    """Compile the iso_c_binding module, the NanoGW drop-in routines (against
    the stand-in types) and the check program with amdflang; None if the
    Fortran compiler is absent."""
    fc = os.environ.get("NGW_FC", "/opt/rocm/bin/amdflang")
    if not os.path.exists(fc):
        sys.stderr.write("amdflang not found: Fortran drop-in not built\n")
        return None
    out = os.path.join(FORTRAN_DIR, "_bin")
    os.makedirs(out, exist_ok=True)
    steps = [
        ["-c", "ngw_c.f90", "-o", os.path.join(out, "ngw_c.o")],
        ["-c", "ngw_dropin_selftest_types.f90", "-o", os.path.join(out, "types.o")],
        ["-cpp", "-DNGW_DROPIN_SELFTEST", "-c", "ngw_dropin.F90", "-o", os.path.join(out, "dropin.o")],
        ["-c", "ngw_dropin_check.f90", "-o", os.path.join(out, "check.o")],
    ]
    for st in steps:
        cmd = [fc, "-O2", f"-J{out}", *st]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True, cwd=FORTRAN_DIR)
    tmp = DROPIN_CHECK + ".tmp"
    subprocess.run([fc, "-o", tmp, *(os.path.join(out, o) for o in ("check.o", "dropin.o", "types.o", "ngw_c.o")),
                    f"-L{HERE}", "-lngw", "-Wl,-rpath,$ORIGIN/../.."], check=True, cwd=FORTRAN_DIR)
    os.replace(tmp, DROPIN_CHECK)
    return DROPIN_CHECK


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
