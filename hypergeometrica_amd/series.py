"""Hypergeometric series by binary splitting, one tree level per batch.

Mirrors src/binary-splitting.lisp: a SERIES is four integer functions
a, b, p, q (:19-34); a PARTIAL over [lower, upper) holds P, Q, B, R (:68-89);
the base case (:100-110) and COMBINE (:112-121) are the reference's.  The
reference recurses depth first (binary-split, :123-133); here the same tree
(split at floor((lower + upper) / 2)) is evaluated bottom-up, one height at a
time, so that every product of a height goes to the GPU in one batched call
(hgx_mul_batch, grouped by transform length).  All arithmetic is exact, so
every partial equals the reference's.

Each COMBINE needs seven products in two dependent rounds:
  round 1: P_l P_r, Q_l Q_r, B_l B_r, Q_r R_l, P_l R_r
  round 2: B_r (Q_r R_l), B_l (P_l R_r)
then R = B_r Q_r R_l + B_l P_l R_r (an addition, done on the host).
"""
from __future__ import annotations

from dataclasses import dataclass
from math import isqrt
from typing import Callable

import numpy as np

from . import mpz as _mpz


@dataclass(frozen=True)
class Poly:
    """A term function that is a polynomial in k with integer coefficients
    (constant term first), or AT_ZERO at k = 0 when given: every series of
    the reference has this form (src/binary-splitting.lisp:140-231,
    src/pi.lisp:24-43), and the native engine (hgx_series_split) takes it."""
    coef: tuple
    at_zero: int | None = None

    def __call__(self, k: int) -> int:
        if k == 0 and self.at_zero is not None:
            return self.at_zero
        v = 0
        for c in reversed(self.coef):
            v = v * k + c
        return v


ONE = Poly((1,))


@dataclass(frozen=True)
class Series:
    """sum_k a(k)/b(k) * p(0)...p(k) / (q(0)...q(k)) (src/binary-splitting.lisp:19-34)."""
    a: Callable[[int], int] = ONE
    b: Callable[[int], int] = ONE
    p: Callable[[int], int] = ONE
    q: Callable[[int], int] = ONE


@dataclass
class Partial:
    """PARTIAL (src/binary-splitting.lisp:68-89) over [lower, upper)."""
    lower: int
    upper: int
    p: int
    q: int
    b: int
    r: int

    def numerator(self) -> int:
        return self.r

    def denominator(self, mul=None) -> int:
        """partial-denominator: B * Q (src/binary-splitting.lisp:87-89)."""
        return (mul or _mul_one)(self.b, self.q)


def _limbs(x: int) -> np.ndarray:
    m = abs(x)
    n = max(1, (m.bit_length() + 63) // 64)
    return np.frombuffer(m.to_bytes(8 * n, "little"), dtype="<u8").astype(np.uint64)


def _int(limbs: np.ndarray) -> int:
    return int.from_bytes(np.ascontiguousarray(limbs, dtype="<u8").tobytes(), "little")


def mul_many(pairs: list[tuple[int, int]], threshold: int | None = None) -> list[int]:
    """Exact products of Python integers.  As the MPZ-* dispatcher does
    (src/multiply.lisp:21-42), a product whose smaller operand has at most
    THRESHOLD limbs (*ntt-multiply-threshold*, 100) stays on the host; all
    others go to the GPU in one batched call."""
    if threshold is None:
        threshold = _mpz.NTT_MULTIPLY_THRESHOLD
    out = [0] * len(pairs)
    idx, batch = [], []
    for i, (x, y) in enumerate(pairs):
        if x == 0 or y == 0:
            continue
        if min(abs(x).bit_length(), abs(y).bit_length()) <= 64 * threshold:
            out[i] = x * y
            continue
        if x is y or x == y:
            a = _limbs(x)
            batch.append((a, a))
        else:
            batch.append((_limbs(x), _limbs(y)))
        idx.append(i)
    if batch:
        prods = _mpz.mul_batch(batch)
        for i, pr in zip(idx, prods):
            x, y = pairs[i]
            s = -1 if (x < 0) != (y < 0) else 1
            out[i] = s * _int(pr)
    return out


def _mul_one(x: int, y: int) -> int:
    return mul_many([(x, y)])[0]


def base_case(series: Series, lower: int) -> Partial:
    """binary-split-base-case=1 (src/binary-splitting.lisp:100-110); its one
    product p * a(lower) is a small integer product."""
    p = series.p(lower)
    return Partial(lower, lower + 1, p, series.q(lower), series.b(lower), p * series.a(lower))


FOLD_TERMS = 32
"""Subtrees of at most this many terms are evaluated by a left fold of their
base cases on the host.  Their partials are the same four integers as the
recursion's (below), and their numbers stay below *ntt-multiply-threshold*
limbs for every series here, so the reference multiplies them on the host
too; every node above them is the reference's own node with its own seven
products."""


def fold_partial(series: Series, lower: int, upper: int) -> Partial:
    """The partial over [lower, upper): base case `lower` combined with the
    base cases `lower + 1, ...` one at a time.  P, Q and B are the products
    of p, q and b over the interval, and R = sum over n of
    a(n) * prod_{m != n} b(m) * prod_{m <= n} p(m) * prod_{m > n} q(m), which
    COMBINE (src/binary-splitting.lisp:112-121) preserves for any split, so
    the fold gives the recursion's integers (tested against it)."""
    a, b, p, q = series.a, series.b, series.p, series.q
    P = p(lower)
    Q, B, R = q(lower), b(lower), P * a(lower)
    BP = B * P  # B * P of the terms so far, kept so that no step multiplies two long numbers
    for k in range(lower + 1, upper):
        pk, qk, bk = p(k), q(k), b(k)
        # COMBINE with the base case (p_k, q_k, b_k, p_k a_k) on the right
        R = (bk * qk) * R + BP * (pk * a(k))
        BP *= bk * pk
        P *= pk
        Q *= qk
        B *= bk
    return Partial(lower, upper, P, Q, B, R)


def _tree(lower: int, upper: int, nodes: list, blocks: list, fold: int) -> int:
    """The reference's split tree down to subtrees of at most FOLD terms
    (recorded in BLOCKS); returns the node's height above them."""
    if upper - lower <= fold:
        blocks.append((lower, upper))
        return 0
    m = (lower + upper) // 2
    h = 1 + max(_tree(lower, m, nodes, blocks, fold), _tree(m, upper, nodes, blocks, fold))
    nodes.append((h, lower, m, upper))
    return h


def binary_split(series: Series, lower: int, upper: int, mul_batch=None, fold: int = FOLD_TERMS) -> Partial:
    """binary-split (src/binary-splitting.lisp:123-133), evaluated one tree
    height at a time with batched products, the subtrees of at most FOLD
    terms by host folds (FOLD = 1: the whole tree, from base cases)."""
    assert upper > lower
    mul = mul_batch or mul_many
    nodes: list = []
    blocks: list = []
    _tree(lower, upper, nodes, blocks, max(1, fold))
    parts: dict[tuple[int, int], Partial] = {
        (lo, up): base_case(series, lo) if up - lo == 1 else fold_partial(series, lo, up) for lo, up in blocks}
    heights: dict[int, list] = {}
    for h, lo, m, up in nodes:
        heights.setdefault(h, []).append((lo, m, up))
    for h in sorted(heights):
        level = heights[h]
        r1 = []
        for lo, m, up in level:
            L, R = parts[(lo, m)], parts[(m, up)]
            r1 += [(L.p, R.p), (L.q, R.q), (L.b, R.b), (R.q, L.r), (L.p, R.r)]
        o1 = mul(r1)
        r2 = []
        for i, (lo, m, up) in enumerate(level):
            L, R = parts[(lo, m)], parts[(m, up)]
            r2 += [(R.b, o1[5 * i + 3]), (L.b, o1[5 * i + 4])]
        o2 = mul(r2)
        for i, (lo, m, up) in enumerate(level):
            parts[(lo, up)] = Partial(lo, up, o1[5 * i], o1[5 * i + 1], o1[5 * i + 2], o2[2 * i] + o2[2 * i + 1])
            del parts[(lo, m)], parts[(m, up)]
    return parts[(lower, upper)]


def binary_split_native(series: Series, lower: int, upper: int) -> Partial:
    """binary-split (src/binary-splitting.lisp:123-133) by the library's
    native engine (hgx_series_split): the same tree evaluated one height at
    a time in C++, with each height's products batched on the GPU (or, for
    small products of a small split, multiplied on host threads).  The
    series' four term functions must be Poly."""
    v = binary_split_native_limbs(series, lower, upper)
    return Partial(lower, upper, *(v[k][0] * _int(v[k][1]) for k in "pqbr"))


def native_desc(series: Series):
    """The hgx_series descriptor of a series whose term functions are Poly."""
    from . import _lib
    fns = (series.a, series.b, series.p, series.q)
    if not all(isinstance(f, Poly) for f in fns):
        raise TypeError("the native engine takes series whose term functions are Poly")
    desc = _lib.SeriesDesc()
    for i, f in enumerate(fns):
        if not 1 <= len(f.coef) <= 8:
            raise ValueError("term polynomials have degree 0..7")
        desc.degree[i] = len(f.coef) - 1
        for j, c in enumerate(f.coef):
            desc.coef[i][j] = c
        desc.at_zero[i] = 0 if f.at_zero is None else 1
        desc.value0[i] = 0 if f.at_zero is None else f.at_zero
    return desc


def binary_split_native_limbs(series: Series, lower: int, upper: int) -> dict[str, tuple[int, np.ndarray]]:
    """binary_split_native's P, Q, B, R as (sign, trimmed little-endian
    limbs), with no conversion to Python integers."""
    import ctypes

    from . import _lib
    desc = native_desc(series)
    L = _lib.load() if not _lib._initialised else _lib.lib()
    res = ctypes.c_void_p(0)
    _lib.check(L.hgx_series_split(ctypes.byref(desc), lower, upper, ctypes.byref(res)))
    out = {}
    try:
        for w, k in enumerate("pqbr"):
            sign, n = ctypes.c_int(0), ctypes.c_size_t(0)
            _lib.check(L.hgx_partial_get(res, w, ctypes.byref(sign), ctypes.byref(n), None))
            buf = np.empty(n.value, dtype=np.uint64)
            if n.value:
                _lib.check(L.hgx_partial_get(res, w, None, None, buf.ctypes.data))
            out[k] = (sign.value, buf)
    finally:
        L.hgx_partial_free(res)
    return out


# ------------------------------------------------------------------ series

def e_series() -> Series:
    """make-e-series (src/binary-splitting.lisp:150-154)."""
    return Series(q=Poly((0, 1), at_zero=1))


def exp_series(num: int, den: int) -> Series:
    """make-exp-series for x = num/den (src/binary-splitting.lisp:140-148)."""
    return Series(p=Poly((num,), at_zero=1), q=Poly((0, den), at_zero=1))


RAMA_A, RAMA_B, RAMA_C = 1103, 26390, 396


def ramanujan_series() -> Series:
    """make-ramanujan-series (src/binary-splitting.lisp:167-184)."""
    return Series(a=Poly((RAMA_A, RAMA_B)),
                  p=Poly((-3, 22, -48, 32), at_zero=1),
                  q=Poly((0, 0, 0, RAMA_C ** 4 // 8), at_zero=1))


CHUD_A, CHUD_B, CHUD_C = 13591409, 545140134, 640320
CHUD_PREFACTOR = 10  # computes pi/10 (src/pi.lisp:14-20)


def chudnovsky_series() -> Series:
    """make-chudnovsky-series (src/pi.lisp:24-43)."""
    return Series(a=Poly((CHUD_A * CHUD_PREFACTOR, CHUD_B * CHUD_PREFACTOR)),
                  b=Poly(((8 * CHUD_C) // 12,)),
                  p=Poly((5, -46, 108, -72), at_zero=1),
                  q=Poly((0, 0, 0, CHUD_C ** 3 // 24), at_zero=1))


def catalan_series() -> Series:
    """make-catalan-series (src/binary-splitting.lisp:203-213)."""
    return Series(b=Poly((1, 2)), p=Poly((0, 1), at_zero=1), q=Poly((2, 4), at_zero=1))


def apery_series() -> Series:
    """make-apery-series (src/binary-splitting.lisp:218-231): q(n) = 32 (1 + 2n)^5 expanded."""
    return Series(a=Poly((77, 250, 205)), b=Poly((2,)),
                  p=Poly((0, 0, 0, 0, 0, -1), at_zero=1),
                  q=Poly((32, 320, 1280, 2560, 2560, 1024)))


CHUD_DECIMALS_PER_TERM = 14.181647462725477  # log10(151931373056000), src/pi.lisp:10


def compute_pi_chudnovsky(prec: int) -> int:
    """compute-pi/chudnovsky (src/pi.lisp:45-51): floor(10^prec * pi / 10)
    with the series summed by batched binary splitting and the final
    division on the host."""
    num_terms = int(2 + prec / CHUD_DECIMALS_PER_TERM)
    sqrt_c = isqrt(10005 * 100 ** prec)
    comp = binary_split(chudnovsky_series(), 0, num_terms)
    den = comp.denominator()
    return (sqrt_c * den) // comp.numerator()


def partial_digits(x: Partial, digits: int, mul=None) -> int:
    """partial-digits (src/binary-splitting.lisp:92-96)."""
    num, den = 10 ** digits * x.numerator(), x.denominator(mul)
    q, r = divmod(num, den)
    if 2 * r > den or (2 * r == den and q % 2 == 1):  # CL ROUND: nearest, ties to even
        q += 1
    return q
