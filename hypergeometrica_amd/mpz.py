"""Host-side mirror of Hypergeometrica's MPZ multiply interface.

The reference's seam is the generic MPZ-* (src/mpz-protocol.lisp:74) and its
mpz/ram method (src/multiply.lisp:12-49), which trims both operands, orders
them, and sends squares to MPZ-SQUARE (src/ntt-multiply.lisp:101-172) and
other products to MPZ-*/NTT (src/ntt-multiply.lisp:174-254).  This module
keeps those names, argument meanings and results, and runs the NTT work on
the GPU through libhgx.so.  There is no CPU multiply here: without the
library every call raises HgxError.

An MPZ is (sign, storage) as mpz/ram is (src/mpz-ram.lisp:19-24): sign in
{-1, 0, 1} and a little-endian uint64 limb vector.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import HgxError, MulDesc, check

NTT_MULTIPLY_THRESHOLD = 100
"""*ntt-multiply-threshold* (src/multiply.lisp:9).  The reference multiplies
operands of at most this many limbs on the host by schoolbook; the callers
in series.py and mpd.py keep that routing.  The mpz_* functions of this
module send every non-zero product to the GPU NTT, which is exact at every
size (as the reference's own test forces with the threshold set to 0,
tests/mpz.lisp:44-47)."""


def _u64(a) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.uint64)
    return a


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def trim(a: np.ndarray) -> np.ndarray:
    """vec-digit-length* / optimize-storage (src/mpz-ram.lisp:83-92): drop
    high zero limbs."""
    a = _u64(a)
    nz = np.flatnonzero(a)
    return a[: (int(nz[-1]) + 1) if nz.size else 0]


@dataclass
class MPZ:
    sign: int
    storage: np.ndarray

    def size(self) -> int:
        return int(trim(self.storage).size)


def integer_mpz(n: int) -> MPZ:
    """integer-mpz: a Python int as an MPZ."""
    s = (n > 0) - (n < 0)
    m = abs(n)
    nl = max(1, (m.bit_length() + 63) // 64)
    limbs = np.frombuffer(m.to_bytes(8 * nl, "little"), dtype="<u8").astype(np.uint64)
    return MPZ(s, limbs)


def mpz_integer(z: MPZ) -> int:
    """mpz-integer: an MPZ as a Python int."""
    v = int.from_bytes(_u64(z.storage).astype("<u8").tobytes(), "little")
    return z.sign * v if v else 0


def mpz_size(z: MPZ) -> int:
    return z.size()


# ------------------------------------------------------------------ limbs

def mul_limbs(a, b) -> np.ndarray:
    """Exact product of two limb vectors: NA + NB limbs (hgx_mul)."""
    a, b = _u64(a), _u64(b)
    if a.size == 0 or b.size == 0:
        return np.zeros(a.size + b.size, dtype=np.uint64)
    out = np.empty(a.size + b.size, dtype=np.uint64)
    L = _lib.lib()
    if a is b:
        check(L.hgx_sqr(_ptr(a), a.size, _ptr(out)))
    else:
        check(L.hgx_mul(_ptr(a), a.size, _ptr(b), b.size, _ptr(out)))
    return out


def sqr_limbs(a) -> np.ndarray:
    """Exact square of a limb vector: 2 NA limbs (hgx_sqr)."""
    a = _u64(a)
    if a.size == 0:
        return np.zeros(0, dtype=np.uint64)
    out = np.empty(2 * a.size, dtype=np.uint64)
    check(_lib.lib().hgx_sqr(_ptr(a), a.size, _ptr(out)))
    return out


def mul_batch(pairs) -> list[np.ndarray]:
    """Many independent products in one call (hgx_mul_batch).  PAIRS holds
    (a, b) limb vectors; b may be the same object as a for a square."""
    keep = []
    descs = (MulDesc * len(pairs))()
    outs = []
    for i, (a, b) in enumerate(pairs):
        a = _u64(a)
        b = a if b is pairs[i][0] else _u64(b)
        out = np.empty(a.size + b.size, dtype=np.uint64)
        keep += [a, b]
        outs.append(out)
        descs[i] = MulDesc(_ptr(a), a.size, _ptr(b), b.size, _ptr(out))
    check(_lib.lib().hgx_mul_batch(descs, len(pairs)))
    return outs


class Ticket:
    """A queued product on the library's request server (hgx_submit).
    Holds the operand and output arrays until the product completes: a
    ticket dropped before completion waits for it first, so that the server
    never writes into freed memory."""

    def __init__(self, a: np.ndarray, b: np.ndarray, sqr: bool = False):
        self._done = True  # nothing to wait for unless the submit succeeds
        self.a = _u64(a)
        self.b = self.a if sqr else _u64(b)
        self.out = np.empty(self.a.size + self.b.size, dtype=np.uint64)
        d = MulDesc(_ptr(self.a), self.a.size, _ptr(self.b), self.b.size, _ptr(self.out))
        t = ctypes.c_uint64(0)
        check(_lib.lib().hgx_submit(ctypes.byref(d), ctypes.byref(t)))
        self.ticket = t.value
        self._done = False

    def done(self) -> bool:
        """True once the product is in OUT; raises the request's error.  The
        poll that reports completion releases the ticket (hgx_poll)."""
        if self._done:
            return True
        try:
            rc = check(_lib.lib().hgx_poll(self.ticket))
        except HgxError:
            self._done = True
            raise
        if rc == 1:
            self._done = True
        return self._done

    def wait(self) -> np.ndarray:
        if not self._done:
            self._done = True
            check(_lib.lib().hgx_wait(self.ticket))
        return self.out

    def __del__(self):
        if not getattr(self, "_done", True) and _lib._lib is not None:
            self._done = True
            _lib._lib.hgx_wait(self.ticket)


def submit(a, b=None) -> Ticket:
    """Queue a*b (or a^2 when b is None) on the request server."""
    if b is None:
        return Ticket(a, a, sqr=True)
    return Ticket(a, b)


# ------------------------------------------------------------------ MPZ

def mpz_mul_ntt(x: MPZ, y: MPZ) -> MPZ:
    """MPZ-*/NTT (src/ntt-multiply.lisp:174-254): product with sign
    sign(x)*sign(y) (:254).  Storage holds NX + NY limbs (the reference's
    L-limb storage holds the same value)."""
    a, b = trim(x.storage), trim(y.storage)
    return MPZ(x.sign * y.sign, mul_limbs(a, b))


def mpz_square(x: MPZ) -> MPZ:
    """MPZ-SQUARE (src/ntt-multiply.lisp:101-172): sign +1 (:172)."""
    a = trim(x.storage)
    return MPZ(1, sqr_limbs(a))


def mpz_mul(x: MPZ, y: MPZ) -> MPZ:
    """The mpz/ram MPZ-* method (src/multiply.lisp:12-49): trim, order so
    |x| >= |y|, zero short-cut, square when x is y, else the NTT product;
    the result is trimmed (optimize-mpz, :19)."""
    xs, ys = trim(x.storage), trim(y.storage)
    if xs.size < ys.size:
        x, y, xs, ys = y, x, ys, xs
    if ys.size == 0 or x.sign == 0 or y.sign == 0:
        return MPZ(0, np.zeros(1, dtype=np.uint64))
    if x is y:
        r = mpz_square(x)
    else:
        r = mpz_mul_ntt(x, y)
    return MPZ(r.sign, trim(r.storage))


def mpz_expt(a: MPZ, n: int) -> MPZ:
    """MPZ-EXPT via F-EXPT (src/multiply.lisp:51-79): left-to-right square
    and multiply; every step goes through mpz_mul."""
    if n == 0:
        return integer_mpz(1)
    x = a
    for bit in range(n.bit_length() - 2, -1, -1):
        x = mpz_mul(x, x)
        if (n >> bit) & 1:
            x = mpz_mul(x, a)
    return x


# ------------------------------------------------------------------ transforms

def ntt_forward(a, mod_num: int) -> np.ndarray:
    """NTT-FORWARD (src/number-theoretic-transform.lisp:40-90) of canonical
    residues mod the MOD_NUM-th modulus; returns the bit-reversed spectrum."""
    x = np.array(a, dtype=np.uint64, copy=True)
    check(_lib.lib().hgx_ntt_forward(_ptr(x), x.size, mod_num))
    return x


def ntt_reverse(a, mod_num: int) -> np.ndarray:
    """NTT-REVERSE (src/number-theoretic-transform.lisp:93-143): bit-reversed
    input, natural-order output scaled by 1/N."""
    x = np.array(a, dtype=np.uint64, copy=True)
    check(_lib.lib().hgx_ntt_reverse(_ptr(x), x.size, mod_num))
    return x


def moduli() -> list[int]:
    """scheme-moduli of **scheme** (src/ntt-multiply.lisp:8)."""
    L = _lib.load()
    return [int(L.hgx_modulus(k)) for k in range(L.hgx_num_moduli())]


def product_length(na: int, nb: int) -> int:
    return int(_lib.load().hgx_product_length(na, nb))


# ------------------------------------------------------------------ device

def mul_device(a_ptr: int, na: int, b_ptr: int, nb: int, out_ptr: int, stream: int = 0) -> None:
    """hgx_mul_device on device pointers (e.g. torch tensor data_ptr())."""
    check(_lib.lib().hgx_mul_device(a_ptr, na, b_ptr, nb, out_ptr, stream or None))


def fill_splitmix64(out_ptr: int, n: int, seed: int, force_top: bool = True, stream: int = 0) -> None:
    check(_lib.lib().hgx_fill_splitmix64(out_ptr, n, seed & (2**64 - 1), int(force_top), stream or None))


__all__ = [
    "MPZ", "HgxError", "NTT_MULTIPLY_THRESHOLD", "integer_mpz", "mpz_integer", "mpz_size",
    "mpz_mul", "mpz_mul_ntt", "mpz_square", "mpz_expt", "mul_limbs", "sqr_limbs", "mul_batch",
    "ntt_forward", "ntt_reverse", "moduli", "product_length", "mul_device", "fill_splitmix64",
    "trim", "submit", "Ticket",
]
