"""ctypes wrapper of tests/_bbp.c: frac(16^d pi) by BBP digit extraction, as
a 64-bit fixed-point fraction within 8 (d + 17) units of 2^-64 of the true
value (each floored term is low by less than one unit; its coefficient is
4, -2, -1 or -1).  Test and benchmark infrastructure only."""
from __future__ import annotations

import ctypes
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "_bbp.c")
LIB_PATH = os.path.join(HERE, "libhgx_bbp.so")
_lib = None


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["gcc", "-O3", "-fPIC", "-shared", "-std=c11", "-pthread", "-o", LIB_PATH + ".tmp", SRC],
                       check=True)
        os.replace(LIB_PATH + ".tmp", LIB_PATH)
    return LIB_PATH


def frac(d: int, threads: int = 8) -> int:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        _lib.hgxc_bbp_frac.argtypes = [ctypes.c_uint64, ctypes.c_int]
        _lib.hgxc_bbp_frac.restype = ctypes.c_uint64
    return int(_lib.hgxc_bbp_frac(d, threads))


def error_bound(d: int) -> int:
    """Units of 2^-64 by which frac() may differ from frac(16^d pi)."""
    return 8 * (d + 17)


def distance(d: int, g: int, threads: int = 8) -> int:
    """Distance mod 1, in units of 2^-64, between G (frac(16^d pi) as a
    64-bit fixed-point fraction from another computation) and BBP's value."""
    diff = (g - frac(d, threads)) % (1 << 64)
    return min(diff, (1 << 64) - diff)


def matches(d: int, g: int, threads: int = 8) -> bool:
    """G agrees with BBP within the error bound plus one unit of G's own
    truncation: about 60 - log2(d) leading bits of pi at hex position d."""
    return distance(d, g, threads) <= error_bound(d) + 1
