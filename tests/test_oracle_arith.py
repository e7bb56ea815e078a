"""Pins the oracle's modular arithmetic (src/modular-arithmetic.lisp) against
exact integer arithmetic, as tests/multiplication.lisp:9-18 and
tests/arithmetic.lisp:47-64 do."""
import numpy as np
import pytest


def rand_below(rng, m, n):
    hi = rng.integers(0, 2**63, size=n, dtype=np.uint64)
    return np.array([int(x) % m for x in hi], dtype=np.uint64)


@pytest.mark.parametrize("k", [0, 1, 2])
def test_m_mul_fast(oracle, k):
    m = oracle.moduli()[k]
    rng = np.random.default_rng(100 + k)
    a = rand_below(rng, m, 100000)
    b = rand_below(rng, m, 100000)
    edge = np.array([0, 1, m - 1, m - 2, m // 2], dtype=np.uint64)
    a = np.concatenate([a, np.repeat(edge, 5)])
    b = np.concatenate([b, np.tile(edge, 5)])
    got = oracle.m_mul_fast(a, b, k)
    want = [(int(x) * int(y)) % m for x, y in zip(a, b)]
    assert [int(v) for v in got] == want


@pytest.mark.parametrize("k", [0, 1, 2])
def test_m_mul_fast2(oracle, k):
    m = oracle.moduli()[k]
    rng = np.random.default_rng(200 + k)
    b = rand_below(rng, m, 20000)
    a = rand_below(rng, m, 20000)
    got = oracle.m_mul_fast2(b, a, k)
    assert [int(v) for v in got] == [(int(x) * int(y)) % m for x, y in zip(a, b)]


def test_add_sub_mul(oracle):
    L = oracle.lib()
    rng = np.random.default_rng(7)
    for m in oracle.moduli():
        for _ in range(2000):
            a = int(rng.integers(0, 2**62)) % m
            b = int(rng.integers(0, 2**62)) % m
            assert L.hgxo_m_add(a, b, m) == (a + b) % m
            assert L.hgxo_m_sub(a, b, m) == (a - b) % m
            assert L.hgxo_m_mul(a, b, m) == (a * b) % m
        for a, b in [(0, 0), (m - 1, m - 1), (m - 1, 1), (0, m - 1)]:
            assert L.hgxo_m_add(a, b, m) == (a + b) % m
            assert L.hgxo_m_sub(a, b, m) == (a - b) % m


def test_inv_mod_and_expt(oracle):
    L = oracle.lib()
    for m in oracle.moduli():
        for x in (1, 2, 3, 12345, m - 1, 2**40 + 7):
            assert (L.hgxo_inv_mod(x, m) * x) % m == 1
            assert L.hgxo_expt_mod(x, 2**55, m) == pow(x, 2**55, m)


def test_planted_barrett_mistake_is_caught(oracle):
    # a dropped correction step must fail the closed-form check above
    k = 0
    m = oracle.moduli()[k]
    rng = np.random.default_rng(9)
    a = rand_below(rng, m, 5000)
    b = rand_below(rng, m, 5000)
    with oracle.Fault(oracle.Fault.BARRETT_FIX):
        got = oracle.m_mul_fast(a, b, k)
    assert [int(v) for v in got] != [(int(x) * int(y)) % m for x, y in zip(a, b)]
