"""Pins the oracle's transforms (src/number-theoretic-transform.lisp:40-143)
against direct DFTs over Z/pZ, as tests/ntt.lisp:121-155 does."""
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def vectors():
    return np.load(os.path.join(GOLD, "ntt_vectors.npz"))


@pytest.mark.parametrize("k", [0, 1, 2])
@pytest.mark.parametrize("nb", [6, 8, 10])
def test_forward_is_bitreversed_direct_dft(oracle, vectors, k, nb):
    x = vectors[f"in_{k}_{nb}"]
    assert np.array_equal(oracle.ntt_forward(x, k), vectors[f"fwd_{k}_{nb}"])


@pytest.mark.parametrize("k", [0, 1, 2])
@pytest.mark.parametrize("nb", [6, 8, 10])
def test_reverse_inverts_forward(oracle, vectors, k, nb):
    # test-inversion-properties (tests/ntt.lisp:121-131) and the reverse half
    # of test-ntt-from-various-definitions (:148-155)
    x = vectors[f"in_{k}_{nb}"]
    assert np.array_equal(oracle.ntt_reverse(vectors[f"fwd_{k}_{nb}"], k), x)


@pytest.mark.parametrize("k", [0, 1, 2])
def test_small_lengths_direct(oracle, k):
    m = oracle.moduli()[k]
    g = oracle.generator(m)
    rng = np.random.default_rng(k)
    for nb in range(0, 6):
        N = 1 << nb
        x = [int(v) % m for v in rng.integers(0, 2**63, size=N, dtype=np.uint64)]
        w = pow(g, (m - 1) >> nb, m)
        X = oracle.direct_dft(x, m, w)
        got = oracle.ntt_forward(np.array(x, dtype=np.uint64), k)
        assert [int(v) for v in got] == [X[oracle.bitrev(i, nb)] for i in range(N)]
        # inverse: natural order, scaled by 1/N
        inv = oracle.ntt_reverse(got, k)
        assert [int(v) for v in inv] == x


def test_planted_twiddle_mistake_is_caught(oracle, vectors):
    x = vectors["in_0_8"]
    with oracle.Fault(oracle.Fault.TWIDDLE_INDEX):
        got = oracle.ntt_forward(x, 0)
    assert not np.array_equal(got, vectors["fwd_0_8"])


def test_planted_scale_mistake_is_caught(oracle, vectors):
    with oracle.Fault(oracle.Fault.INVERSE_SCALE):
        got = oracle.ntt_reverse(vectors["fwd_1_6"], 1)
    assert not np.array_equal(got, vectors["in_1_6"])
