"""Checks of a computed pi/10, an MPD (mantissa * 2^(64 exponent), as
mpd-pi returns it, src/pi.lisp:53-74), against values of pi computed by
other methods: BBP hexadecimal digit extraction at chosen positions
(tests/_bbp.c), MPFR's mpfr_const_pi (tests/_mpfr.py), and exact integer
routings through GMP.  Test and benchmark infrastructure only."""
from __future__ import annotations

import numpy as np

from tests import _bbp, _gmp, _mpfr


def pi_scaled(mantissa: int, exponent: int, shift: int) -> int:
    """floor(pi * 2^shift) from the MPD value of pi/10 (pi = 10 * value)."""
    x = 10 * mantissa
    frac_bits = -64 * exponent
    return x >> (frac_bits - shift) if frac_bits >= shift else x << (shift - frac_bits)


def bbp_positions(prec_bits: int) -> list[int]:
    """Hex positions to check for a value good to PREC_BITS bits: near the
    end, the middle and the start of the fraction."""
    last = (prec_bits - 192) // 4
    return [last, last // 2 + 12345 % (last // 4 + 1), min(1000003, last // 3)]


def bbp_check(mantissa: int, exponent: int, positions, threads: int = 8) -> dict:
    """{position d: distance, in units of 2^-64, between frac(16^d pi) from
    the value and from BBP}; each must be within _bbp.error_bound(d)."""
    out = {}
    for d in positions:
        g = pi_scaled(mantissa, exponent, 4 * d + 64) & ((1 << 64) - 1)
        out[d] = _bbp.distance(d, g, threads)
    return out


def bbp_ok(distances: dict) -> bool:
    return all(v <= _bbp.error_bound(d) + 1 for d, v in distances.items())


def mpfr_leading_bits_match(mantissa: int, exponent: int, bits: int) -> bool:
    """The value's first BITS bits of pi equal MPFR's (within 2 units of
    the last hex digit compared, for the truncations on both sides)."""
    n = bits // 4
    want = _mpfr.pi_hex(n)
    got = pi_scaled(mantissa, exponent, 4 * (n - 1))
    return abs(got - want) <= 2


def _limbs(x: int) -> np.ndarray:
    n = max(1, (x.bit_length() + 63) // 64)
    return np.frombuffer(x.to_bytes(8 * n, "little"), dtype="<u8").astype(np.uint64)


def gmp_mul(x: int, y: int) -> int:
    """x * y through GMP's mpz_mul: the exact host routing of products."""
    if x == 0 or y == 0:
        return 0
    s = -1 if (x < 0) != (y < 0) else 1
    r = _gmp.mul(_limbs(abs(x)), _limbs(abs(y)))
    return s * int.from_bytes(r.astype("<u8").tobytes(), "little")
