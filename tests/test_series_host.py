"""Binary splitting (src/binary-splitting.lisp, src/pi.lisp): the oracle's
depth-first restatement pinned by closed forms (mpmath digits of pi and e),
and the level-synchronous driver's host logic (tree, rounds, signs) checked
against it with exact Python products in place of the GPU batch."""
import mpmath
import pytest

from hypergeometrica_amd import series as S
from oracle import series_oracle as SO


def python_batch(pairs):
    return [x * y for x, y in pairs]


def test_oracle_pi_matches_mpmath():
    prec = 300
    s = S.chudnovsky_series()
    n = int(2 + prec / S.CHUD_DECIMALS_PER_TERM)
    lo, up, p, q, b, r = SO.binary_split(s, 0, n)
    from math import isqrt
    val = (isqrt(10005 * 100 ** prec) * (b * q)) // r
    mpmath.mp.dps = prec + 20
    want = int(mpmath.floor(mpmath.pi / 10 * mpmath.mpf(10) ** prec))
    assert abs(val - want) <= 1


def test_oracle_e_matches_mpmath():
    prec = 200
    lo, up, p, q, b, r = SO.binary_split(S.e_series(), 0, prec + 5)
    mpmath.mp.dps = prec + 20
    want = int(mpmath.nint(mpmath.e * mpmath.mpf(10) ** prec))
    got = S.partial_digits(S.Partial(lo, up, p, q, b, r), prec, mul=lambda x, y: x * y)
    assert got == want


@pytest.mark.parametrize("make,lo,up", [(S.chudnovsky_series, 0, 37), (S.ramanujan_series, 0, 23),
                                        (S.e_series, 0, 50), (S.catalan_series, 0, 41), (S.apery_series, 0, 19),
                                        (lambda: S.exp_series(3, 7), 0, 33), (S.chudnovsky_series, 5, 6),
                                        (S.chudnovsky_series, 3, 100)])
def test_level_synchronous_equals_depth_first(make, lo, up):
    s = make()
    want = SO.binary_split(s, lo, up)
    got = S.binary_split(s, lo, up, mul_batch=python_batch)
    assert (got.lower, got.upper, got.p, got.q, got.b, got.r) == want
