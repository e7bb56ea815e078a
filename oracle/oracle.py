"""ctypes wrapper around the CPU oracle (oracle/libhgx_oracle.so).

TEST INFRASTRUCTURE ONLY.  The oracle is a C restatement of Hypergeometrica's
NTT multiply (src/ntt-multiply.lisp, src/number-theoretic-transform.lisp,
src/modular-arithmetic.lisp, src/moduli.lisp); see hgx_oracle.c for the
per-function citations.  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module; the product path never does.

Also holds small pure-Python references used as pins: a direct DFT over
Z/pZ (tests/ntt.lisp:62-89 computes the same thing), bit reversal, limb
packing, and SplitMix64.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhgx_oracle.so")

_u64 = ctypes.c_uint64
_u64p = ctypes.POINTER(ctypes.c_uint64)
_sz = ctypes.c_size_t

_lib = None


def build(force: bool = False) -> str:
    """Compile the oracle with gcc (make -C oracle)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-C", _HERE, "-B" if force else "-s"], check=True,
                       stdout=subprocess.DEVNULL)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.hgxo_primep.argtypes = [_u64]
        L.hgxo_primep.restype = ctypes.c_int
        for name in ("hgxo_m_add", "hgxo_m_sub", "hgxo_m_mul", "hgxo_calc_b_inv",
                     "hgxo_expt_mod", "hgxo_inv_mod", "hgxo_primitive_root"):
            getattr(L, name).restype = _u64
        L.hgxo_m_add.argtypes = [_u64, _u64, _u64]
        L.hgxo_m_sub.argtypes = [_u64, _u64, _u64]
        L.hgxo_m_mul.argtypes = [_u64, _u64, _u64]
        L.hgxo_calc_b_inv.argtypes = [_u64, _u64]
        L.hgxo_expt_mod.argtypes = [_u64, _u64, _u64]
        L.hgxo_inv_mod.argtypes = [_u64, _u64]
        L.hgxo_primitive_root.argtypes = [_u64, _u64, _u64]
        L.hgxo_m_mul_fast.argtypes = [_u64, _u64, _u64, _u64]
        L.hgxo_m_mul_fast.restype = _u64
        L.hgxo_mod128_fast.argtypes = [_u64, _u64, _u64, _u64]
        L.hgxo_mod128_fast.restype = _u64
        L.hgxo_m_mul_fast2.argtypes = [_u64, _u64, _u64, _u64]
        L.hgxo_m_mul_fast2.restype = _u64
        L.hgxo_find_generator.argtypes = [_u64]
        L.hgxo_find_generator.restype = _u64
        L.hgxo_m_mul_fast_vec.argtypes = [_u64p, _u64p, _u64p, _sz, _u64, _u64]
        L.hgxo_m_mul_fast2_vec.argtypes = [_u64p, _u64p, _u64p, _sz, _u64]
        L.hgxo_default_moduli.argtypes = [_u64p, ctypes.c_int]
        L.hgxo_default_moduli.restype = ctypes.c_int
        L.hgxo_convolution_bits.argtypes = [_u64]
        L.hgxo_convolution_bits.restype = ctypes.c_int
        L.hgxo_least_power_of_two_ge.argtypes = [_u64]
        L.hgxo_least_power_of_two_ge.restype = _u64
        L.hgxo_ntt_forward.argtypes = [_u64p, _sz, ctypes.c_int]
        L.hgxo_ntt_reverse.argtypes = [_u64p, _sz, ctypes.c_int]
        L.hgxo_mpz_mul_ntt.argtypes = [_u64p, _sz, _u64p, _sz, _u64p]
        L.hgxo_mpz_mul_ntt.restype = _sz
        L.hgxo_mpz_mul_ntt_parallel.argtypes = [_u64p, _sz, _u64p, _sz, _u64p, ctypes.c_int]
        L.hgxo_mpz_mul_ntt_parallel.restype = _sz
        L.hgxo_mpz_square.argtypes = [_u64p, _sz, _u64p]
        L.hgxo_mpz_square.restype = _sz
        L.hgxo_product_length.argtypes = [_sz, _sz]
        L.hgxo_product_length.restype = _sz
        L.hgxo_crt.argtypes = [_u64p, _u64p]
        L.hgxo_add_big_digit.argtypes = [_u64p, ctypes.c_int, _u64p, _sz, _sz]
        L.hgxo_add_big_digit.restype = ctypes.c_int
        L.hgxo_schoolbook.argtypes = [_u64p, _sz, _u64p, _sz, _u64p]
        L.hgxo_splitmix64.argtypes = [_u64, _u64p, _sz]
        _lib = L
    return _lib


def _p(a: np.ndarray):
    assert a.dtype == np.uint64 and a.flags.c_contiguous
    return a.ctypes.data_as(_u64p)


class Fault:
    """Planted-mistake switch of the oracle (hgxo_fault), for mutation tests."""
    NONE, CRT_DROP_TERM, TWIDDLE_INDEX, INVERSE_SCALE, BARRETT_FIX, CARRY = range(6)

    def __init__(self, kind: int):
        self.kind = kind

    def __enter__(self):
        ctypes.c_int.in_dll(lib(), "hgxo_fault").value = self.kind
        return self

    def __exit__(self, *exc):
        ctypes.c_int.in_dll(lib(), "hgxo_fault").value = 0
        return False


# ---------------------------------------------------------------- scheme

def default_moduli() -> list[int]:
    buf = np.zeros(16, dtype=np.uint64)
    n = lib().hgxo_default_moduli(_p(buf), 16)
    return [int(x) for x in buf[:n]]


MODULI = None


def moduli() -> list[int]:
    global MODULI
    if MODULI is None:
        MODULI = default_moduli()
    return MODULI


def generator(m: int) -> int:
    return int(lib().hgxo_find_generator(m))


def barrett_inverse(m: int) -> int:
    """scheme-inverses entry floor(2^126/m) (src/moduli.lisp:342-345)."""
    return (1 << 126) // m


def primitive_root(n: int, m: int) -> int:
    return int(lib().hgxo_primitive_root(generator(m), n, m))


# ---------------------------------------------------------------- arithmetic

def m_mul_fast(a: np.ndarray, b: np.ndarray, mod_num: int) -> np.ndarray:
    m = moduli()[mod_num]
    a = np.ascontiguousarray(a, dtype=np.uint64)
    b = np.ascontiguousarray(b, dtype=np.uint64)
    out = np.empty_like(a)
    lib().hgxo_m_mul_fast_vec(_p(a), _p(b), _p(out), a.size, m, barrett_inverse(m))
    return out


def m_mul_fast2(b: np.ndarray, a: np.ndarray, mod_num: int) -> np.ndarray:
    m = moduli()[mod_num]
    a = np.ascontiguousarray(a, dtype=np.uint64)
    b = np.ascontiguousarray(b, dtype=np.uint64)
    out = np.empty_like(a)
    lib().hgxo_m_mul_fast2_vec(_p(b), _p(a), _p(out), a.size, m)
    return out


# ---------------------------------------------------------------- transforms

def ntt_forward(a: np.ndarray, mod_num: int) -> np.ndarray:
    """ntt-forward (src/number-theoretic-transform.lisp:40-90) on a copy."""
    x = np.array(a, dtype=np.uint64, copy=True)
    lib().hgxo_ntt_forward(_p(x), x.size, mod_num)
    return x


def ntt_reverse(a: np.ndarray, mod_num: int) -> np.ndarray:
    """ntt-reverse (src/number-theoretic-transform.lisp:93-143) on a copy."""
    x = np.array(a, dtype=np.uint64, copy=True)
    lib().hgxo_ntt_reverse(_p(x), x.size, mod_num)
    return x


# ---------------------------------------------------------------- multiply

def product_length(nx: int, ny: int) -> int:
    return int(lib().hgxo_product_length(nx, ny))


def mpz_mul_ntt(x: np.ndarray, y: np.ndarray) -> np.ndarray:
    """mpz-*/ntt (src/ntt-multiply.lisp:174-254); returns all L limbs."""
    x = np.ascontiguousarray(x, dtype=np.uint64)
    y = np.ascontiguousarray(y, dtype=np.uint64)
    L = product_length(x.size, y.size)
    out = np.zeros(L, dtype=np.uint64)
    got = lib().hgxo_mpz_mul_ntt(_p(x), x.size, _p(y), y.size, _p(out))
    assert got == L
    return out


def mpz_mul_ntt_parallel(x: np.ndarray, y: np.ndarray, threads: int = 6) -> np.ndarray:
    """mpz-*/ntt with the reference's parallel tasks (src/ntt-multiply.lisp:60-86,
    :202-234), at most THREADS at once; returns all L limbs."""
    x = np.ascontiguousarray(x, dtype=np.uint64)
    y = np.ascontiguousarray(y, dtype=np.uint64)
    L = product_length(x.size, y.size)
    out = np.zeros(L, dtype=np.uint64)
    got = lib().hgxo_mpz_mul_ntt_parallel(_p(x), x.size, _p(y), y.size, _p(out), threads)
    assert got == L, "thread creation failed"
    return out


def mpz_square(x: np.ndarray) -> np.ndarray:
    """mpz-square (src/ntt-multiply.lisp:101-172); returns all L limbs."""
    x = np.ascontiguousarray(x, dtype=np.uint64)
    L = int(lib().hgxo_least_power_of_two_ge(2 * x.size))
    out = np.zeros(L, dtype=np.uint64)
    got = lib().hgxo_mpz_square(_p(x), x.size, _p(out))
    assert got == L
    return out


def schoolbook(x: np.ndarray, y: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.uint64)
    y = np.ascontiguousarray(y, dtype=np.uint64)
    out = np.zeros(x.size + y.size, dtype=np.uint64)
    lib().hgxo_schoolbook(_p(x), x.size, _p(y), y.size, _p(out))
    return out


def crt(residues) -> int:
    r = np.array(residues, dtype=np.uint64)
    out = np.zeros(4, dtype=np.uint64)
    lib().hgxo_crt(_p(r), _p(out))
    return limbs_to_int(out)


# ---------------------------------------------------------------- helpers

def splitmix64(seed: int, n: int) -> np.ndarray:
    """out[i] = mix(seed + (i+1)*0x9E3779B97F4A7C15), the SplitMix64 stream."""
    out = np.empty(n, dtype=np.uint64)
    if n:
        lib().hgxo_splitmix64(seed & (2**64 - 1), _p(out), n)
    return out


def splitmix64_py(seed: int, n: int) -> list[int]:
    """Pure-Python SplitMix64 (pins the C and HIP generators)."""
    M = 2**64 - 1
    out = []
    for i in range(n):
        z = (seed + (i + 1) * 0x9E3779B97F4A7C15) & M
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        out.append(z ^ (z >> 31))
    return out


def limbs_to_int(a) -> int:
    a = np.asarray(a, dtype=np.uint64)
    return int.from_bytes(a.astype("<u8").tobytes(), "little")


def int_to_limbs(x: int, n: int | None = None) -> np.ndarray:
    assert x >= 0
    if n is None:
        n = max(1, (x.bit_length() + 63) // 64)
    return np.frombuffer(x.to_bytes(8 * n, "little"), dtype="<u8").astype(np.uint64)


def bitrev(i: int, bits: int) -> int:
    r = 0
    for _ in range(bits):
        r = (r << 1) | (i & 1)
        i >>= 1
    return r


def direct_dft(v, m: int, w: int) -> list[int]:
    """X[k] = sum_j v[j] w^(jk) mod m (ntt-forward-direct, tests/ntt.lisp:62-74)."""
    n = len(v)
    out = []
    for k in range(n):
        wk = pow(w, k, m)
        s = 0
        acc = 1
        for j in range(n):
            s += int(v[j]) * acc
            acc = acc * wk % m
        out.append(s % m)
    return out
