"""The sb-alien shim (lisp/hgx.lisp) against the C ABI it binds.

SBCL is not in this image, so the shim is checked as text: it reads as
balanced S-expressions, and every DEFINE-ALIEN-ROUTINE names a function that
include/hgx.h declares and libhgx.so exports, with as many parameters as the
declaration and a foreign type of the same size and kind for each (pointers
and size_t as 64-bit words, int as int; an :out parameter where the C
parameter points to the value returned).  The drop-in functions it redefines
are the reference's MPZ-*/NTT and MPZ-SQUARE (src/ntt-multiply.lisp:101-254)."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SHIM = os.path.join(ROOT, "lisp", "hgx.lisp")
HEADER = os.path.join(ROOT, "include", "hgx.h")


def read_sexps(text):
    """Top-level forms of TEXT as nested lists of atoms (strings keep their
    quotes); comments dropped."""
    tokens = re.findall(r'"(?:\\.|[^"\\])*"|;[^\n]*|[()]|[^\s()";]+', text)
    stack = [[]]
    for t in tokens:
        if t.startswith(";"):
            continue
        if t == "(":
            stack.append([])
        elif t == ")":
            assert len(stack) > 1, "unbalanced )"
            form = stack.pop()
            stack[-1].append(form)
        else:
            stack[-1].append(t)
    assert len(stack) == 1, "unbalanced ("
    return stack[0]


def c_prototypes():
    text = re.sub(r"//[^\n]*", "", open(HEADER).read())
    protos = {}
    for ret, name, params in re.findall(r"\b(int|void|uint64_t|const char \*)\s*(hgx_\w+)\(([^)]*)\);", text):
        ps = [p.strip() for p in params.split(",") if p.strip() and p.strip() != "void"]
        protos[name] = (ret.strip(), ps)
    return protos


def routines():
    out = {}
    for form in read_sexps(open(SHIM).read()):
        if form and form[0] == "sb-alien:define-alien-routine":
            cname = form[1][0].strip('"')
            out[cname] = (form[1][1], form[2], form[3:])
    return out


WORD = {"sb-alien:unsigned-long", "sb-sys:system-area-pointer"}


def kind_ok(ctype: str, lisp: list) -> bool:
    ltype = lisp[1]
    out = len(lisp) > 2 and lisp[2] == ":out"
    pointer = "*" in ctype
    if out:
        # the C parameter points to the returned value
        assert pointer, ctype
        base = ctype.replace("const", "").strip().rsplit("*", 1)[0].strip()
        if base in ("size_t", "uint64_t", "hgx_ticket", "uint64_t *", "hgx_partial *"):
            return ltype == "sb-alien:unsigned-long"
        if base == "int":
            return ltype == "sb-alien:int"
        return False
    if isinstance(ltype, list):  # (* hgx-mul-desc), (* hgx-series)
        return pointer and ltype[0] == "*" and ltype[1].replace("-", "_") in ctype
    if pointer or ctype.split()[0] in ("size_t", "uint64_t", "hgx_ticket"):
        return ltype in WORD
    if ctype.split()[0] == "int":
        return ltype == "sb-alien:int"
    return False


def test_shim_reads_as_balanced_forms():
    forms = read_sexps(open(SHIM).read())
    names = [f[1] for f in forms if f and f[0] == "defun"]
    assert "mpz-*/ntt" in names and "mpz-square" in names
    assert ["in-package", "#:hypergeometrica"] in forms


def test_every_routine_matches_the_header():
    protos = c_prototypes()
    rs = routines()
    assert len(rs) >= 25
    for cname, (_lname, ret, params) in rs.items():
        assert cname in protos, f"{cname} is not declared in include/hgx.h"
        cret, cparams = protos[cname]
        assert len(params) == len(cparams), (cname, params, cparams)
        for cp, lp in zip(cparams, params):
            assert kind_ok(cp, lp), (cname, cp, lp)
        want = {"int": "sb-alien:int", "void": "sb-alien:void", "uint64_t": "sb-alien:unsigned-long",
                "const char *": "sb-alien:c-string"}[cret]
        assert ret == want, (cname, ret, cret)


def test_every_routine_is_exported(hgx_lib_path):
    import ctypes
    lib = ctypes.CDLL(hgx_lib_path)
    for cname in routines():
        assert hasattr(lib, cname), cname


@pytest.fixture
def hgx_lib_path():
    p = os.path.join(ROOT, "hypergeometrica_amd", "libhgx.so")
    if not os.path.exists(p):
        pytest.skip("libhgx.so not built")
    return p
