"""hypergeometrica_amd: exact NTT big-integer multiplication on AMD MI355X.

A from-scratch gfx950 implementation of Hypergeometrica's MPZ-*/NTT and
MPZ-SQUARE (src/ntt-multiply.lisp), exported through the C ABI in
include/hgx.h (libhgx.so) and mirrored here in Python.
"""
from .mpz import *  # noqa: F401,F403
from ._lib import HgxError, EXPORTS  # noqa: F401
