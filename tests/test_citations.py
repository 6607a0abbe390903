"""Every file:line citation of a NanoGW source file in the header, the
sources, the oracle, the tests and the documents names a file that exists in
the reference tree and lines inside it.  Reads the reference sources as text;
skipped where they are not present."""
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.environ.get("NGW_REFERENCE", "/root/reference")
CITE = re.compile(r"([A-Za-z_0-9]+\.(?:F90z|F90|f90))((?::\d+(?:-\d+)?)(?:,\s*:?\d+(?:-\d+)?)*)")


def _reference_lengths():
    out = {}
    for p in glob.glob(os.path.join(REF, "**", "*"), recursive=True):
        if re.search(r"\.(F90z|F90|f90)$", p):
            with open(p, errors="replace") as f:
                n = sum(1 for _ in f)
            b = os.path.basename(p)
            out[b] = max(out.get(b, 0), n)
    return out


def _citing_files():
    pats = ["include/*.h", "*.md", "nanogw_amd/csrc/*", "nanogw_amd/*.py", "nanogw_amd/fortran/*.f90",
            "nanogw_amd/fortran/*.F90", "oracle/*.py", "oracle/*.c", "oracle/Makefile", "tests/*.py", "bench.py",
            "profiles/*.md"]
    for pat in pats:
        for p in sorted(glob.glob(os.path.join(ROOT, pat))):
            if os.path.basename(p) in ("SURVEY.md", "VERDICT.md", "ADVICE.md", "BASELINE.md"):
                continue
            yield p


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "shared")), reason="reference sources not present")
def test_reference_citations_point_inside_existing_files():
    lengths = _reference_lengths()
    bad, n = [], 0
    for path in _citing_files():
        with open(path, errors="replace") as f:
            text = f.read()
        for m in CITE.finditer(text):
            name = m.group(1)
            if name not in lengths:
                bad.append((os.path.relpath(path, ROOT), name, "no such reference file"))
                continue
            for lo, hi in re.findall(r":?(\d+)(?:-(\d+))?", m.group(2)):
                n += 1
                if int(lo) < 1 or int(hi or lo) > lengths[name] or (hi and int(hi) < int(lo)):
                    bad.append((os.path.relpath(path, ROOT), name + m.group(2), f"file has {lengths[name]} lines"))
    assert n > 200
    assert not bad, bad
