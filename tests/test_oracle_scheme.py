"""Pins the oracle's scheme constants (src/moduli.lisp) against closed forms
and the reference's own log transcripts."""
import json
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def test_moduli_match_reference_logs(oracle):
    # every "Moduli:" line the reference printed (e.g. logs/24apr21-999-ram.txt:34)
    rows = json.load(open(os.path.join(GOLD, "logs.json")))
    assert len(rows) > 100
    mods = oracle.moduli()
    for r in rows:
        assert r["moduli"] == mods, r


def test_moduli_closed_form(oracle):
    # default-moduli (src/moduli.lisp:257-264): 63-bit primes k*2^55+1, descending
    d = json.load(open(os.path.join(GOLD, "scheme.json")))
    assert oracle.moduli() == d["moduli"]
    assert len(d["moduli"]) == 3
    for p in d["moduli"]:
        assert p.bit_length() == 63 and (p - 1) % 2**55 == 0
    # test-scheme-is-sufficient, tests/moduli.lisp:7-10
    assert sum(p.bit_length() - 1 for p in d["moduli"]) >= 64 + 64 + 50
    assert min(d["two_adicity"]) >= 50


def test_generators(oracle):
    d = json.load(open(os.path.join(GOLD, "scheme.json")))
    assert [oracle.generator(p) for p in d["moduli"]] == d["generators"] == [3, 5, 3]


def test_barrett_inverses(oracle):
    d = json.load(open(os.path.join(GOLD, "scheme.json")))
    assert [oracle.barrett_inverse(p) for p in d["moduli"]] == d["barrett_inverses"]
    # closed form: floor(2^126 / p), computed here with Python integers
    assert d["barrett_inverses"] == [(1 << 126) // p for p in d["moduli"]]
    assert [hex(x) for x in d["barrett_inverses"]] == ["0xa655c4392d7b73a5", "0xbc52640bc52640ba",
                                                       "0xfa232cf252138abb"]


def test_primitive_roots(oracle):
    d = json.load(open(os.path.join(GOLD, "scheme.json")))
    for nbits, ws in d["omega"].items():
        n = 1 << int(nbits)
        for p, w in zip(d["moduli"], ws):
            assert oracle.primitive_root(n, p) == w
            assert pow(w, n, p) == 1 and pow(w, n // 2, p) == p - 1  # order exactly n


def test_convolution_bits_and_lengths_match_logs(oracle):
    # "Transform length" / "Convolution bits" pairs printed by mpz-*/ntt and
    # mpz-square (src/ntt-multiply.lisp:176-178, :103-105)
    rows = json.load(open(os.path.join(GOLD, "logs.json")))
    L = oracle.lib()
    for r in rows:
        assert L.hgxo_convolution_bits(r["length"]) == r["bits"], r
    for n in range(1, 56):
        assert L.hgxo_convolution_bits(1 << n) == 128 + n


def test_least_power_of_two(oracle):
    L = oracle.lib()
    for n, want in [(1, 1), (2, 2), (3, 4), (4, 4), (5, 8), (4097, 8192), (2**26 + 2**26, 2**27)]:
        assert L.hgxo_least_power_of_two_ge(n) == want


def test_primep_matches_closed_form(oracle):
    import sympy
    rng = np.random.default_rng(5)
    for x in list(rng.integers(2, 2**62, size=2000, dtype=np.int64)) + [2**61 - 1, 561, 1105, 2**55 * 197 + 1]:
        assert bool(oracle.lib().hgxo_primep(int(x))) == sympy.isprime(int(x))
