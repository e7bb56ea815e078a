"""CPU checks of the native binary-splitting engine (hgx_series_split) on
intervals whose products all have a smaller operand of at most 100 limbs,
which the engine multiplies on the host as the reference's mpz-* does
(src/multiply.lisp:37-42): no device is needed.  Each partial is compared
with the oracle's depth-first binary-split (src/binary-splitting.lisp:100-133),
itself pinned by closed forms in test_series_host.py."""
import pytest

from hypergeometrica_amd import series as S
from hypergeometrica_amd._lib import HGX_EINVAL, HGX_ENOTINIT, HgxError
from oracle import series_oracle as SO

SERIES = {
    "chudnovsky": S.chudnovsky_series, "ramanujan": S.ramanujan_series, "e": S.e_series,
    "catalan": S.catalan_series, "apery": S.apery_series, "exp(-5/3)": lambda: S.exp_series(-5, 3),
    "exp(7/2)": lambda: S.exp_series(7, 2),
}


@pytest.mark.parametrize("name,lo,up", [
    ("chudnovsky", 0, 37), ("chudnovsky", 0, 60), ("chudnovsky", 3, 40), ("chudnovsky", 5, 6),
    ("chudnovsky", 123456789, 123456800), ("ramanujan", 0, 23), ("ramanujan", 0, 70), ("e", 0, 1), ("e", 0, 50),
    ("e", 0, 700), ("catalan", 0, 41), ("catalan", 0, 300), ("apery", 0, 19), ("apery", 0, 60),
    ("apery", 1000000, 1000003), ("exp(-5/3)", 0, 400), ("exp(7/2)", 1, 33),
])
def test_native_equals_depth_first(name, lo, up):
    s = SERIES[name]()
    want = SO.binary_split(s, lo, up)
    got = S.binary_split_native(s, lo, up)
    assert (got.lower, got.upper, got.p, got.q, got.b, got.r) == want


def test_poly_terms_equal_the_reference_forms():
    # the term functions as the reference writes them (src/binary-splitting.lisp:140-231, src/pi.lisp:24-43)
    C = S.CHUD_C
    ref = {
        "chudnovsky": (lambda n: 135914090 + n * 5451401340, lambda n: 8 * C // 12,
                       lambda n: 1 if n == 0 else 5 + n * (-46 + n * (108 + n * -72)),
                       lambda n: 1 if n == 0 else n ** 3 * (C ** 3 // 24)),
        "ramanujan": (lambda n: 1103 + n * 26390, lambda n: 1,
                      lambda n: 1 if n == 0 else -3 + n * (22 + n * (-48 + n * 32)),
                      lambda n: 1 if n == 0 else n ** 3 * (396 ** 4 // 8)),
        "e": (lambda n: 1, lambda n: 1, lambda n: 1, lambda n: 1 if n == 0 else n),
        "catalan": (lambda n: 1, lambda n: 1 + 2 * n, lambda n: 1 if n == 0 else n,
                    lambda n: 1 if n == 0 else 2 + 4 * n),
        "apery": (lambda n: 77 + n * (250 + n * 205), lambda n: 2, lambda n: 1 if n == 0 else -(n ** 5),
                  lambda n: 32 * (1 + 2 * n) ** 5),
        "exp(-5/3)": (lambda n: 1, lambda n: 1, lambda n: 1 if n == 0 else -5, lambda n: 1 if n == 0 else 3 * n),
    }
    for name, fns in ref.items():
        s = SERIES[name]()
        for n in list(range(0, 40)) + [10 ** 6, 2 ** 31 + 7, 2 ** 39]:
            assert (s.a(n), s.b(n), s.p(n), s.q(n)) == tuple(f(n) for f in fns), (name, n)


def test_large_term_values():
    # apery's q(k) = 32 (2k+1)^5 reaches 2^206 at k near 2^40: multi-limb term values
    s = S.apery_series()
    lo = (1 << 40) - 5
    want = SO.binary_split(s, lo, lo + 4)
    got = S.binary_split_native(s, lo, lo + 4)
    assert (got.p, got.q, got.b, got.r) == want[2:]


def test_errors():
    with pytest.raises(HgxError) as e:
        S.binary_split_native(S.e_series(), 5, 5)
    assert e.value.code == HGX_EINVAL
    with pytest.raises(HgxError) as e:
        S.binary_split_native(S.e_series(), 0, (1 << 40) + 1)
    assert e.value.code == HGX_EINVAL
    with pytest.raises(TypeError):
        S.binary_split_native(S.Series(a=lambda n: 1), 0, 3)
    with pytest.raises(ValueError):
        S.binary_split_native(S.Series(a=S.Poly(tuple(range(9)))), 0, 3)


def test_gpu_products_need_init():
    from hypergeometrica_amd import _lib
    if _lib._initialised:
        pytest.skip("the library is initialised in this process")
    with pytest.raises(HgxError) as e:
        S.binary_split_native(S.e_series(), 0, 3000)  # 3000! has more than 100 limbs
    assert e.value.code == HGX_ENOTINIT


@pytest.mark.parametrize("series,lo,up", [
    # p vanishes at k = 5: P and every product with it become zero
    (S.Series(p=S.Poly((-5, 1), at_zero=1), q=S.Poly((0, 3), at_zero=1)), 0, 40),
    # a vanishes everywhere: R = 0
    (S.Series(a=S.Poly((0,)), q=S.Poly((1, 1))), 0, 50),
    # negative b and q: signs of B, Q and R alternate
    (S.Series(a=S.Poly((2, -7)), b=S.Poly((-3,)), p=S.Poly((1, 2, -1)), q=S.Poly((-4, 0, 5), at_zero=-1)), 0, 45),
    # degree-7 terms with the largest coefficients
    (S.Series(a=S.Poly((2**63 - 1,) * 8), p=S.Poly((-(2**63),) + (1,) * 7, at_zero=7), q=S.Poly((1,) * 8)), 0, 9),
])
def test_native_degenerate_series(series, lo, up):
    want = SO.binary_split(series, lo, up)
    got = S.binary_split_native(series, lo, up)
    assert (got.lower, got.upper, got.p, got.q, got.b, got.r) == want
