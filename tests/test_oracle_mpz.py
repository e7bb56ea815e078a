"""Pins the oracle's multiply path (mpz-*/ntt and mpz-square,
src/ntt-multiply.lisp:101-254): the product of two integers is unique, so
Python ints, schoolbook and GMP fix every output limb."""
import os

import numpy as np
import pytest

from tests import _gmp

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def exact(a, b):
    from oracle.oracle import limbs_to_int
    return limbs_to_int(a) * limbs_to_int(b)


def test_golden_products(oracle):
    d = np.load(os.path.join(GOLD, "mul_vectors.npz"))
    n = len([k for k in d.files if k.startswith("p_")])
    for i in range(n):
        a, b, p = d[f"a_{i}"], d[f"b_{i}"], d[f"p_{i}"]
        got = oracle.mpz_mul_ntt(a, b)
        assert got.size == oracle.product_length(a.size, b.size)
        assert np.array_equal(got[: a.size + b.size], p), i
        assert not got[a.size + b.size:].any()
        if a is b or np.array_equal(a, b):
            s = oracle.mpz_square(a)
            assert np.array_equal(s[: 2 * a.size], p), i


@pytest.mark.parametrize("na,nb", [(1, 1), (1, 2), (2, 2), (5, 3), (100, 101), (257, 255), (1024, 1024),
                                   (1023, 1025), (4095, 1), (3000, 5000)])
def test_products_vs_python_and_schoolbook(oracle, na, nb):
    a = oracle.splitmix64(na * 1000 + nb, na)
    b = oracle.splitmix64(nb * 1000 + na + 7, nb)
    got = oracle.mpz_mul_ntt(a, b)
    assert oracle.limbs_to_int(got) == exact(a, b)
    assert np.array_equal(got[: na + nb], oracle.schoolbook(a, b))


@pytest.mark.skipif(not _gmp.available(), reason="libgmp.so.10 not present")
def test_products_vs_gmp(oracle):
    for n in (4096, 20000, 65536):
        a = oracle.splitmix64(0x48475800, n)
        b = oracle.splitmix64(0x48475801, n)
        got = oracle.mpz_mul_ntt(a, b)
        assert np.array_equal(got[: 2 * n], _gmp.mul(a, b))


@pytest.mark.parametrize("threads", [1, 2, 6])
def test_parallel_tasks_same_product(oracle, threads):
    """The with-parallel-work form (tasks per operand and modulus,
    src/ntt-multiply.lisp:60-86, :202-234) gives the serial product and the
    exact one."""
    for na, nb in [(1, 1), (101, 7), (3000, 2000), (40000, 25000)]:
        a = oracle.splitmix64(0x5100 + na, na)
        b = oracle.splitmix64(0x5200 + nb, nb)
        got = oracle.mpz_mul_ntt_parallel(a, b, threads)
        assert np.array_equal(got, oracle.mpz_mul_ntt(a, b)), (na, nb)
        assert oracle.limbs_to_int(got[: na + nb]) == exact(a, b)


def test_reference_style_random_mpz(oracle):
    # test-mpz-plus-minus-times-randomly (tests/mpz.lisp:32-47): ~333-bit values
    rng = np.random.default_rng(3)
    for _ in range(10):
        a = int(rng.integers(1, 2**62)) ** 6
        b = int(rng.integers(1, 2**62)) ** 5 + 12345
        la = oracle.int_to_limbs(a)
        lb = oracle.int_to_limbs(b)
        assert oracle.limbs_to_int(oracle.mpz_mul_ntt(la, lb)) == a * b
        assert oracle.limbs_to_int(oracle.mpz_square(la)) == a * a


@pytest.mark.parametrize("n", [1, 2, 3, 64, 127, 128, 1000, 2048])
def test_all_ones_square(oracle, n):
    a = np.full(n, 2**64 - 1, dtype=np.uint64)
    want = oracle.limbs_to_int(a) ** 2
    assert oracle.limbs_to_int(oracle.mpz_square(a)) == want
    assert oracle.limbs_to_int(oracle.mpz_mul_ntt(a, a.copy())) == want


def test_crt_is_exact_reconstruction(oracle):
    mods = oracle.moduli()
    M = mods[0] * mods[1] * mods[2]
    rng = np.random.default_rng(11)
    for _ in range(3000):
        x = int(rng.integers(0, 2**62)) * int(rng.integers(0, 2**62)) * int(rng.integers(0, 2**62))
        x %= M
        assert oracle.crt([x % m for m in mods]) == x
    for x in (0, 1, M - 1, (1 << 155) - 1):
        assert oracle.crt([x % m for m in mods]) == x


def test_add_big_digit(oracle):
    import ctypes
    L = oracle.lib()
    st = np.array([2**64 - 1, 2**64 - 1, 5, 0], dtype=np.uint64)
    d = np.array([1], dtype=np.uint64)
    P = ctypes.POINTER(ctypes.c_uint64)
    assert L.hgxo_add_big_digit(d.ctypes.data_as(P), 1, st.ctypes.data_as(P), 4, 0) == 0
    assert [int(v) for v in st] == [0, 0, 6, 0]
    # a non-zero value past the end is the reference's "grave inconsistency"
    d2 = np.array([1, 1], dtype=np.uint64)
    assert L.hgxo_add_big_digit(d2.ctypes.data_as(P), 2, st.ctypes.data_as(P), 4, 3) == -1


@pytest.mark.parametrize("fault", ["CRT_DROP_TERM", "CARRY", "TWIDDLE_INDEX", "INVERSE_SCALE", "BARRETT_FIX"])
def test_planted_mistakes_fail_the_product_pin(oracle, fault):
    a = oracle.splitmix64(1, 700)
    b = oracle.splitmix64(2, 900)
    with oracle.Fault(getattr(oracle.Fault, fault)):
        got = oracle.mpz_mul_ntt(a, b)
    assert oracle.limbs_to_int(got) != exact(a, b)
