"""Exact products through GMP's mpz_mul (libgmp.so.10 via ctypes), used as
an independent pin of big products.  Returns None when GMP is absent."""
from __future__ import annotations

import ctypes
import ctypes.util

import numpy as np


class _Mpz(ctypes.Structure):
    _fields_ = [("alloc", ctypes.c_int), ("size", ctypes.c_int), ("d", ctypes.c_void_p)]


_gmp = None


def _load():
    global _gmp
    if _gmp is None:
        for name in ("libgmp.so.10", ctypes.util.find_library("gmp")):
            if not name:
                continue
            try:
                _gmp = ctypes.CDLL(name)
                break
            except OSError:
                continue
    return _gmp


def available() -> bool:
    return _load() is not None


def mul(a: np.ndarray, b: np.ndarray) -> np.ndarray | None:
    g = _load()
    if g is None:
        return None
    x, y, z = _Mpz(), _Mpz(), _Mpz()
    for v in (x, y, z):
        g.__gmpz_init(ctypes.byref(v))
    a = np.ascontiguousarray(a, dtype=np.uint64)
    b = np.ascontiguousarray(b, dtype=np.uint64)
    g.__gmpz_import(ctypes.byref(x), ctypes.c_size_t(a.size), -1, 8, 0, 0, a.ctypes.data_as(ctypes.c_void_p))
    g.__gmpz_import(ctypes.byref(y), ctypes.c_size_t(b.size), -1, 8, 0, 0, b.ctypes.data_as(ctypes.c_void_p))
    g.__gmpz_mul(ctypes.byref(z), ctypes.byref(x), ctypes.byref(y))
    out = np.zeros(a.size + b.size, dtype=np.uint64)
    cnt = ctypes.c_size_t(0)
    g.__gmpz_export(out.ctypes.data_as(ctypes.c_void_p), ctypes.byref(cnt), -1, 8, 0, 0, ctypes.byref(z))
    for v in (x, y, z):
        g.__gmpz_clear(ctypes.byref(v))
    return out
