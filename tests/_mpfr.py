"""pi from MPFR (mpfr_const_pi through ctypes on libmpfr.so.6): an
independent value of pi for the high-precision pi checks.  Test and
benchmark infrastructure only; absent MPFR, available() is False."""
from __future__ import annotations

import ctypes
import ctypes.util

_m = None


class _Mpfr(ctypes.Structure):
    _fields_ = [("prec", ctypes.c_long), ("sign", ctypes.c_int), ("exp", ctypes.c_long), ("d", ctypes.c_void_p)]


def _load():
    global _m
    if _m is None:
        for name in ("libmpfr.so.6", ctypes.util.find_library("mpfr")):
            if not name:
                continue
            try:
                m = ctypes.CDLL(name)
            except OSError:
                continue
            m.mpfr_get_str.restype = ctypes.c_void_p
            m.mpfr_get_str.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_long), ctypes.c_int, ctypes.c_size_t,
                                       ctypes.POINTER(_Mpfr), ctypes.c_int]
            m.mpfr_free_str.argtypes = [ctypes.c_void_p]
            _m = m
            break
    return _m


def available() -> bool:
    return _load() is not None


def pi_hex(ndigits: int) -> int:
    """floor(pi * 16^(ndigits - 1)): pi's first NDIGITS hexadecimal digits as
    an integer (3 is the first), from mpfr_const_pi at 4 NDIGITS + 64 bits
    and mpfr_get_str toward zero."""
    m = _load()
    x = _Mpfr()
    m.mpfr_init2(ctypes.byref(x), ctypes.c_long(4 * ndigits + 64))
    try:
        m.mpfr_const_pi(ctypes.byref(x), 0)  # to nearest
        e = ctypes.c_long(0)
        p = m.mpfr_get_str(None, ctypes.byref(e), 16, ndigits, ctypes.byref(x), 1)  # toward zero
        try:
            s = ctypes.string_at(p).decode()
        finally:
            m.mpfr_free_str(p)
        assert e.value == 1, e.value  # pi = 0.3243F6... * 16
        return int(s, 16)
    finally:
        m.mpfr_clear(ctypes.byref(x))
        m.mpfr_free_cache()
