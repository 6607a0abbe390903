"""The Fortran interfaces (nanogw_amd/fortran/ngw_c.f90) against the C
prototypes of include/ngw.h: for every bind(C) function, the same number of
arguments in the same order, scalars passed by value with the C type's kind
(int <-> integer(c_int), long <-> integer(c_long), double <-> real(c_double)),
pointers passed by reference (arrays, intent(out) handles) or as
type(c_ptr), value; the result kind matches.  A mismatch here compiles on both
sides and corrupts the call at run time, so it is checked from the sources."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ngw.h")
FMOD = os.path.join(ROOT, "nanogw_amd", "fortran", "ngw_c.f90")


def _c_prototypes():
    text = re.sub(r"/\*.*?\*/", " ", open(HEADER).read(), flags=re.S)
    out = {}
    for m in re.finditer(r"^\s*((?:const\s+)?\w+\s*\**)\s*(ngw_\w+)\s*\(([^)]*)\)\s*;", text, re.M):
        ret, name, args = m.group(1), m.group(2), m.group(3).strip()
        params = []
        if args and args != "void":
            for a in args.split(","):
                a = " ".join(a.split())
                ptr = a.count("*")
                base = re.sub(r"\bconst\b", "", a).replace("*", " ").split()[0]
                params.append((base, ptr))
        out[name] = ("ptr" if "*" in ret else ret.strip(), params)
    return out


def _fortran_interfaces():
    lines = open(FMOD).read().splitlines()
    # join continuation lines and drop comments
    joined, cur = [], ""
    for ln in lines:
        ln = ln.split("!")[0].rstrip()
        if ln.endswith("&"):
            cur += ln[:-1] + " "
            continue
        cur += ln.lstrip(" &") if cur else ln
        joined.append(cur)
        cur = ""
    out, i = {}, 0
    head = re.compile(r"^\s*(integer\(c_int\)\s+)?function\s+(\w+)\s*\(([^)]*)\)\s*bind\(C,\s*name='(\w+)'\)"
                      r"(\s*result\((\w+)\))?", re.I)
    while i < len(joined):
        m = head.match(joined[i])
        if not m:
            i += 1
            continue
        fname, args, cname, resvar = m.group(2), m.group(3), m.group(4), m.group(6)
        names = [a.strip().lower() for a in args.split(",") if a.strip()]
        decl = {}
        i += 1
        while not re.match(r"^\s*end\s+function", joined[i], re.I):
            d = re.match(r"^\s*(integer\(c_\w+\)|real\(c_\w+\)|complex\(c_\w+\)|type\(c_ptr\))\s*(,[^:]*)?::(.*)$",
                         joined[i], re.I)
            if d:
                ftype, attrs = d.group(1).lower(), (d.group(2) or "").lower()
                for v in re.findall(r"(\w+)(\s*\([^)]*\))?", d.group(3)):
                    decl[v[0].lower()] = (ftype, "value" in attrs, bool(v[1]))
            i += 1
        if m.group(1):
            ret = "integer(c_int)"
        else:
            ret = decl.get((resvar or fname).lower(), ("?",))[0]
        out[cname] = (ret, [decl.get(n, ("?", False, False)) for n in names], names)
    return out


KIND = {"int": "integer(c_int)", "long": "integer(c_long)", "double": "real(c_double)"}
PTR_OK = {"double": ("real(c_double)", "complex(c_double_complex)"), "int": ("integer(c_int)",),
          "long": ("integer(c_long)",)}


def test_every_fortran_interface_matches_the_c_prototype():
    cp, fi = _c_prototypes(), _fortran_interfaces()
    assert len(fi) >= 19
    for cname, (fret, fargs, fnames) in fi.items():
        assert cname in cp, f"{cname} bound in ngw_c.f90 but not declared in ngw.h"
        cret, cargs = cp[cname]
        where = f"{cname}: "
        if cret == "int":
            assert fret == "integer(c_int)", where + "result"
        else:
            assert fret == "type(c_ptr)", where + "result"
        assert len(fargs) == len(cargs), where + f"{len(fargs)} Fortran vs {len(cargs)} C arguments"
        for k, ((ftype, by_value, _), (cbase, cptr), nm) in enumerate(zip(fargs, cargs, fnames)):
            w = where + f"argument {k + 1} ({nm})"
            assert ftype != "?", w + " undeclared"
            if cptr == 0:
                assert by_value and ftype == KIND[cbase], w + f": C {cbase} by value, Fortran {ftype}"
            elif ftype == "type(c_ptr)":
                # ngw_context* / ngw_batch* / a nullable array: the address by value;
                # ngw_context**: a c_ptr variable by reference
                assert by_value == (cptr == 1), w + ": c_ptr passing"
            else:
                assert cptr == 1 and not by_value, w + ": array or scalar by reference"
                assert ftype in PTR_OK.get(cbase, ()), w + f": C {cbase}* vs Fortran {ftype}"


def test_parsers_see_the_known_signatures():
    cp, fi = _c_prototypes(), _fortran_interfaces()
    assert cp["ngw_create"] == ("int", [("ngw_context", 2), ("int", 0)])
    assert cp["ngw_last_error"] == ("ptr", [])
    ret, args, names = fi["ngw_set_isdf"]
    assert names[:4] == ["ctx", "n_intp", "nband", "nval"]
    assert args[0] == ("type(c_ptr)", True, False) and args[4] == ("real(c_double)", False, True)
    # a planted mismatch is caught: a double passed where C takes int
    bad = dict(fi)
    r, a, n = bad["ngw_set_precision"]
    bad["ngw_set_precision"] = (r, [a[0], ("real(c_double)", True, False)], n)
    with pytest.raises(AssertionError):
        (ftype, by_value, _), (cbase, cptr) = bad["ngw_set_precision"][1][1], cp["ngw_set_precision"][1][1]
        assert by_value and ftype == KIND[cbase]
