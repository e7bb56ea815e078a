"""Dyadic floats, Newton reciprocal / square root and pi, with every product
on the GPU.

Mirrors src/mpd.lisp, src/mpd-reciprocal.lisp, src/mpd-sqrt.lisp and
src/pi.lisp.  An MPD is mantissa * B^exponent with B = 2^64 (src/mpd.lisp:7-13);
the mantissa is an exact integer.  Every MPZ product goes through the
library (mpz_mul, i.e. hgx_mul / hgx_sqr), so MPD-* here computes the same
numbers as the reference's MPD-* (src/mpd.lisp:83-88).  Additions, shifts
and the scalar seeds are host work, as in the reference.

The reference's Newton iterations never truncate (src/pi.lisp:72 is
commented out; "TODO: Should we truncate?" at src/mpd-reciprocal.lisp:69 and
src/mpd-sqrt.lisp:80), so their mantissas grow geometrically.  TRUNCATE=True
applies the reference's own (unused) MPD-TRUNCATE! (src/mpd.lisp:56-70) to
the iterate after every step: a declared deviation that keeps the sizes
bounded and still converges to the requested precision.
"""
from __future__ import annotations

import math
import struct
from dataclasses import dataclass
from typing import Callable

import numpy as np

from . import mpz as _mpz
from . import series as _series

DIGIT_BITS = 64
BASE = 1 << DIGIT_BITS


def _gpu_mul(x: int, y: int) -> int:
    """Exact product as MPZ-* computes it (src/multiply.lisp:12-49): operands
    of at most *ntt-multiply-threshold* limbs stay on the host, every other
    product goes through the library's NTT path."""
    if x == 0 or y == 0:
        return 0
    if min(abs(x).bit_length(), abs(y).bit_length()) <= 64 * _mpz.NTT_MULTIPLY_THRESHOLD:
        return x * y
    s = -1 if (x < 0) != (y < 0) else 1
    if x == y:
        return s * _series._int(_mpz.sqr_limbs(_series._limbs(x)))
    return s * _series._int(_mpz.mul_limbs(_series._limbs(x), _series._limbs(y)))


def _size(m: int) -> int:
    """mpz-size: limbs of |m| (0 for zero)."""
    return (abs(m).bit_length() + 63) // 64


@dataclass
class MPD:
    """mantissa * 2^(64 * exponent) (src/mpd.lisp:7-13)."""
    mantissa: int
    exponent: int

    def rational(self):
        from fractions import Fraction
        return Fraction(self.mantissa) * Fraction(BASE) ** self.exponent


class Context:
    """The MPD operations with a chosen exact multiplier (GPU by default)."""

    def __init__(self, mul: Callable[[int, int], int] | None = None, truncate: bool = False):
        self.mul = mul or _gpu_mul
        self.truncate = truncate
        self.products = 0
        self.max_product_limbs = 0

    # -- src/mpd.lisp
    def mpd_mul(self, a: MPD, b: MPD) -> MPD:
        """mpd-* (src/mpd.lisp:83-88)."""
        self.products += 1
        self.max_product_limbs = max(self.max_product_limbs, _size(a.mantissa) + _size(b.mantissa))
        return MPD(self.mul(a.mantissa, b.mantissa), a.exponent + b.exponent)

    @staticmethod
    def mpd_add(a: MPD, b: MPD) -> MPD:
        """mpd-+ (src/mpd.lisp:90-99): align to the smaller exponent."""
        m = min(a.exponent, b.exponent)
        return MPD((a.mantissa << (64 * (a.exponent - m))) + (b.mantissa << (64 * (b.exponent - m))), m)

    @staticmethod
    def mpd_sub(a: MPD, b: MPD) -> MPD:
        return Context.mpd_add(a, MPD(-b.mantissa, b.exponent))

    @staticmethod
    def mpd_oom(x: MPD) -> tuple[int, bool]:
        """mpd-oom (src/mpd.lisp:23-41)."""
        if x.mantissa == 0:
            return 0, True
        return min(0, _size(x.mantissa) + x.exponent), False

    @staticmethod
    def mpd_half(a: MPD) -> MPD:
        """mpd-half! (src/mpd.lisp:77-81): mantissa * B/2, exponent - 1."""
        return MPD(a.mantissa * (BASE // 2), a.exponent - 1)

    @staticmethod
    def mpd_truncate(a: MPD, digits: int) -> MPD:
        """mpd-truncate! (src/mpd.lisp:56-70) to DIGITS limbs: drop the low
        limbs of the magnitude, keep the sign."""
        s = _size(a.mantissa)
        if digits >= s:
            return a
        d = s - digits
        mag = abs(a.mantissa) >> (64 * d)
        return MPD(mag if a.mantissa >= 0 else -mag, a.exponent + d)

    # -- src/mpd-reciprocal.lisp
    @staticmethod
    def estimate_reciprocal(a: MPD) -> MPD:
        """estimate-reciprocal (src/mpd-reciprocal.lisp:7-25)."""
        assert a.mantissa != 0
        size = _size(a.mantissa)
        msb = abs(a.mantissa) >> (64 * (size - 1))
        r = _round_div(BASE, msb)
        return MPD(-r if a.mantissa < 0 else r, -a.exponent - size)

    def mpd_reciprocal(self, a: MPD, bits: int) -> MPD:
        """mpd-reciprocal (src/mpd-reciprocal.lisp:27-73): x <- x + x h,
        h = 1 - a x, until h has BITS/64 leading zero limbs."""
        digits = max(1, -(-bits // 64))
        x = self.estimate_reciprocal(a)
        while True:
            h = self.mpd_sub(MPD(1, 0), self.mpd_mul(a, x))
            zeros, zero = self.mpd_oom(h)
            if zero or -zeros >= digits:
                return x
            x = self.mpd_add(x, self.mpd_mul(x, h))
            if self.truncate:
                x = self.mpd_truncate(x, digits + 2)

    # -- src/mpd-sqrt.lisp
    @staticmethod
    def estimate_inv_sqrt(a: MPD) -> MPD:
        """estimate-inv-sqrt (src/mpd-sqrt.lisp:7-38).  The reference computes
        (round $base (sqrt k)) in single-float: (sqrt integer) returns a
        single-float and $base is converted to single-float before the
        division; restated here with IEEE binary32 rounding."""
        assert a.mantissa >= 0
        size = _size(a.mantissa)
        msb = a.mantissa >> (64 * (size - 1))
        dne = -1 - size - a.exponent
        k = BASE * msb if dne % 2 != 0 else msb
        est = _round_single(_f32_div(_int_to_f32(BASE), _f32_sqrt(_int_to_f32(k))))
        if dne % 2 != 0:
            dne += 1
        return MPD(est, dne // 2)

    def mpd_inv_sqrt(self, a: MPD, bits: int) -> MPD:
        """mpd-inv-sqrt (src/mpd-sqrt.lisp:40-87): x <- (2x + x h) / 2,
        h = 1 - a x^2."""
        digits = max(1, -(-bits // 64))
        x = self.estimate_inv_sqrt(a)
        while True:
            h = self.mpd_sub(MPD(1, 0), self.mpd_mul(a, self.mpd_mul(x, x)))
            zeros, zero = self.mpd_oom(h)
            if zero or -zeros >= digits:
                return x
            x = self.mpd_half(self.mpd_add(self.mpd_add(x, x), self.mpd_mul(x, h)))
            if self.truncate:
                x = self.mpd_truncate(x, digits + 2)

    def mpd_sqrt(self, a: MPD, bits: int) -> MPD:
        """mpd-sqrt (src/mpd-sqrt.lisp:89-90)."""
        return self.mpd_mul(a, self.mpd_inv_sqrt(a, bits))

    # -- src/pi.lisp
    def mpd_pi(self, prec_bits: int) -> MPD:
        """mpd-pi (src/pi.lisp:53-74): pi/10 to PREC_BITS bits.  Binary
        splitting runs level by level with batched products."""
        guard = prec_bits + DIGIT_BITS
        num_terms = 2 + math.floor(guard / CHUD_BITS_PER_TERM)
        comp = _series.binary_split(_series.chudnovsky_series(), 0, num_terms,
                                    mul_batch=None if self.mul is _gpu_mul else
                                    (lambda pairs: [self.mul(x, y) for x, y in pairs]))
        sqrt = self.mpd_sqrt(MPD(10005, 0), guard)
        recip = self.mpd_reciprocal(MPD(comp.numerator(), 0), guard)
        final = self.mpd_mul(sqrt, recip)
        den = comp.denominator(self.mul if self.mul is not _gpu_mul else None)
        return self.mpd_mul(final, MPD(den, 0))


CHUD_BITS_PER_TERM = math.log2(151931373056000)  # +chud-bits-per-term+, src/pi.lisp:11


def _round_div(n: int, d: int) -> int:
    """CL (round n d) for integers: nearest, ties to even."""
    q, r = divmod(n, d)
    if 2 * r > d or (2 * r == d and q % 2 == 1):
        q += 1
    return q


# ---- IEEE binary32 arithmetic (SBCL single-float), exact restatement
def _int_to_f32(n: int) -> float:
    """Integer to the nearest binary32 (ties to even), as CL COERCE / float
    contagion does; avoids the double rounding of float(n) then float32."""
    if n == 0:
        return 0.0
    e = n.bit_length() - 24
    if e <= 0:
        return float(np.float32(n))
    q, r = divmod(n, 1 << e)
    half = 1 << (e - 1)
    if r > half or (r == half and q & 1):
        q += 1
    return float(np.float32(q * 2.0 ** e))


def _f32_sqrt(x: float) -> float:
    return float(np.sqrt(np.float32(x)))


def _f32_div(a: float, b: float) -> float:
    return float(np.float32(a) / np.float32(b))


def _round_single(x: float) -> int:
    """CL (round single-float): nearest integer, ties to even."""
    from fractions import Fraction
    f = Fraction(struct.unpack("<f", struct.pack("<f", x))[0])
    q, r = divmod(f.numerator, f.denominator)
    if 2 * r > f.denominator or (2 * r == f.denominator and q % 2 == 1):
        q += 1
    return q


def mpd_pi(prec_bits: int, truncate: bool = False) -> MPD:
    return Context(truncate=truncate).mpd_pi(prec_bits)
