"""Pure-Python restatement of the reference's binary splitting -- TEST
INFRASTRUCTURE ONLY (small cases; the checker of hypergeometrica_amd.series).

Follows src/binary-splitting.lisp step for step: base case (:100-110),
COMBINE (:112-121) and the depth-first recursion split at
floor((lower + upper) / 2) (:123-133), with Python integers for MPZ.
"""
from __future__ import annotations


def base_case(series, lower):
    # binary-split-base-case=1, src/binary-splitting.lisp:100-110
    p = series.p(lower)
    return (lower, lower + 1, p, series.q(lower), series.b(lower), p * series.a(lower))


def combine(left, right):
    # combine, src/binary-splitting.lisp:112-121
    ll, lu, lp, lq, lb, lr = left
    rl, ru, rp, rq, rb, rr = right
    assert lu == rl
    return (ll, ru, lp * rp, lq * rq, lb * rb, rb * (rq * lr) + lb * (lp * rr))


def binary_split(series, lower, upper):
    # binary-split, src/binary-splitting.lisp:123-133
    assert upper > lower
    if upper - lower == 1:
        return base_case(series, lower)
    m = (lower + upper) // 2
    return combine(binary_split(series, lower, m), binary_split(series, m, upper))
