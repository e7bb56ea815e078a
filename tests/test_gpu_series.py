"""GPU checks of the native binary-splitting engine (hgx_series_split) on
intervals whose upper tree heights multiply on the GPU: every partial equals
the Python level-synchronous split with GPU products (series.binary_split)
and, where it finishes quickly, the oracle's depth-first split with Python
integers (src/binary-splitting.lisp:100-133)."""
import pytest

from hypergeometrica_amd import series as S
from oracle import series_oracle as SO

pytestmark = pytest.mark.gpu

SERIES = {
    "chudnovsky": S.chudnovsky_series, "ramanujan": S.ramanujan_series, "e": S.e_series,
    "catalan": S.catalan_series, "apery": S.apery_series, "exp(7/2)": lambda: S.exp_series(7, 2),
    "exp(-5/3)": lambda: S.exp_series(-5, 3),
}


def _tuple(x):
    return (x.lower, x.upper, x.p, x.q, x.b, x.r)


@pytest.mark.parametrize("name,lo,up", [
    ("chudnovsky", 0, 3000), ("ramanujan", 0, 3000), ("e", 0, 5000), ("catalan", 0, 4000), ("apery", 0, 2000),
    ("exp(7/2)", 0, 5000), ("exp(-5/3)", 17, 4017), ("chudnovsky", 1000, 2999),
])
def test_native_equals_depth_first(hgx, name, lo, up):
    s = SERIES[name]()
    assert _tuple(S.binary_split_native(s, lo, up)) == SO.binary_split(s, lo, up)


@pytest.mark.parametrize("name,lo,up", [
    ("chudnovsky", 0, 60000), ("ramanujan", 0, 40000), ("e", 0, 100000), ("catalan", 0, 60000),
    ("apery", 0, 30000), ("exp(7/2)", 0, 80000), ("chudnovsky", 777777, 800001),
])
def test_native_equals_python_levels(hgx, name, lo, up):
    s = SERIES[name]()
    assert _tuple(S.binary_split_native(s, lo, up)) == _tuple(S.binary_split(s, lo, up))


def test_native_chudnovsky_pi_digits(hgx):
    # compute-pi/chudnovsky (src/pi.lisp:45-51) with the native partial: the
    # digits of pi/10 against mpmath
    import mpmath
    from math import isqrt
    prec = 20000
    n = int(2 + prec / S.CHUD_DECIMALS_PER_TERM)
    comp = S.binary_split_native(S.chudnovsky_series(), 0, n)
    val = (isqrt(10005 * 100 ** prec) * (comp.b * comp.q)) // comp.r
    mpmath.mp.dps = prec + 30
    want = int(mpmath.floor(mpmath.pi / 10 * mpmath.mpf(10) ** prec))
    assert abs(val - want) <= 1
