"""ctypes wrapper of tests/_residue.c: residues of limb vectors modulo
q = 2^61 - 1, the size-independent check of products.  Test and benchmark
infrastructure only."""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "_residue.c")
LIB_PATH = os.path.join(HERE, "libhgx_residue.so")
Q = (1 << 61) - 1

_lib = None


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["gcc", "-O3", "-fPIC", "-shared", "-std=c11", "-o", LIB_PATH + ".tmp", SRC], check=True)
        os.replace(LIB_PATH + ".tmp", LIB_PATH)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        u64, vp, sz = ctypes.c_uint64, ctypes.c_void_p, ctypes.c_size_t
        L.hgxc_residue_at.argtypes = [vp, sz, u64]
        L.hgxc_residue_at.restype = u64
        L.hgxc_residue.argtypes = [vp, sz]
        L.hgxc_residue.restype = u64
        L.hgxc_prefix_residues.argtypes = [vp, sz, vp]
        L.hgxc_mulmod.argtypes = [u64, u64]
        L.hgxc_mulmod.restype = u64
        L.hgxc_slice_residue.argtypes = [vp, sz, sz]
        L.hgxc_slice_residue.restype = u64
        _lib = L
    return _lib


def residue(a: np.ndarray) -> int:
    """sum a[i] 2^(64 i) mod q (ctypes releases the GIL meanwhile)."""
    a = np.ascontiguousarray(a, dtype=np.uint64)
    return int(lib().hgxc_residue(a.ctypes.data, a.size))


def residue_ptr(ptr: int, n: int) -> int:
    """Residue of N limbs at host address PTR."""
    return int(lib().hgxc_residue(ptr, n))


def prefix_residues(a: np.ndarray) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.uint64)
    P = np.empty(a.size + 1, dtype=np.uint64)
    lib().hgxc_prefix_residues(a.ctypes.data, a.size, P.ctypes.data)
    return P


def slice_residue(P: np.ndarray, off: int, n: int) -> int:
    """Residue of a[off:off + n] from the prefix residues P of a."""
    return int(lib().hgxc_slice_residue(P.ctypes.data, off, n))


def mulmod(x: int, y: int) -> int:
    return (x * y) % Q
