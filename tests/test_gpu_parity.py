"""GPU parity: the MI355X library (through the C ABI) against the oracle,
the golden fixtures, GMP and exact integers.  Integer work: bit-exact."""
import os
import sys

import time

import numpy as np
import pytest

from tests import _gmp
from tests._checksum import Q, mod_q

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def residues(oracle, k, n, seed):
    m = oracle.moduli()[k]
    x = oracle.splitmix64(seed, n)
    return np.array([int(v) % m for v in x], dtype=np.uint64) if n < 4096 else (x % np.uint64(m))


# ---------------------------------------------------------------- transforms

@pytest.mark.parametrize("nb", [1, 2, 3, 4, 5, 8, 11, 12, 13, 16, 17, 20, 21])
@pytest.mark.parametrize("k", [0, 1, 2])
def test_ntt_forward_reverse_match_oracle(hgx, oracle, k, nb):
    say(f"ntt k={k} n=2^{nb}")
    x = residues(oracle, k, 1 << nb, 1000 * k + nb)
    want = oracle.ntt_forward(x, k)
    got = hgx.ntt_forward(x, k)
    assert np.array_equal(got, want)
    back = hgx.ntt_reverse(got, k)
    assert np.array_equal(back, oracle.ntt_reverse(want, k))
    assert np.array_equal(back, x)


@pytest.mark.parametrize("k", [0, 1, 2])
@pytest.mark.parametrize("nb", [6, 8, 10])
def test_ntt_golden_vectors(hgx, k, nb):
    v = np.load(os.path.join(GOLD, "ntt_vectors.npz"))
    assert np.array_equal(hgx.ntt_forward(v[f"in_{k}_{nb}"], k), v[f"fwd_{k}_{nb}"])
    assert np.array_equal(hgx.ntt_reverse(v[f"fwd_{k}_{nb}"], k), v[f"in_{k}_{nb}"])


def test_ntt_rejects_bad_arguments(hgx):
    from hypergeometrica_amd import HgxError
    with pytest.raises(HgxError):
        hgx.ntt_forward(np.zeros(6, dtype=np.uint64), 0)
    with pytest.raises(HgxError):
        hgx.ntt_forward(np.zeros(8, dtype=np.uint64), 3)


# ---------------------------------------------------------------- products

SIZES = [(1, 1), (1, 2), (3, 5), (100, 101), (101, 101), (777, 3), (2048, 2048), (2048, 2049), (1000, 7000),
         (4096, 4096), (8192, 1), (65536, 65536), (65541, 32768), (100000, 31072)]


@pytest.mark.parametrize("na,nb", SIZES)
def test_mul_matches_oracle(hgx, oracle, na, nb):
    say(f"mul {na}x{nb}")
    a = oracle.splitmix64(0x48475800 + na, na)
    b = oracle.splitmix64(0x48475900 + nb, nb)
    got = hgx.mul_limbs(a, b)
    want = oracle.mpz_mul_ntt(a, b)
    assert np.array_equal(got, want[: na + nb])


@pytest.mark.parametrize("bits", range(13, 21))
def test_every_row_pass_length(hgx, oracle, bits):
    # transform lengths 2^13 .. 2^20: one column pass and rows of 2^5 .. 2^12
    # words; up to 2^19 a single product's row pass runs on 2048-word tiles,
    # at 2^20 on 4096-word tiles
    n = 1 << (bits - 1)
    say(f"row pass L=2^{bits}")
    a = oracle.splitmix64(0x524F5700 + bits, n)
    b = oracle.splitmix64(0x524F5780 + bits, n - 5)
    assert np.array_equal(hgx.mul_limbs(a, b), oracle.mpz_mul_ntt(a, b)[: 2 * n - 5])
    assert np.array_equal(hgx.sqr_limbs(a), oracle.mpz_square(a)[: 2 * n])


@pytest.mark.parametrize("n", [1, 2, 50, 101, 2048, 2049, 30000, 65536])
def test_sqr_matches_oracle(hgx, oracle, n):
    say(f"sqr {n}")
    a = oracle.splitmix64(0x53515200 + n, n)
    got = hgx.sqr_limbs(a)
    want = oracle.mpz_square(a)
    assert np.array_equal(got, want[: 2 * n])


def test_golden_products(hgx):
    d = np.load(os.path.join(GOLD, "mul_vectors.npz"))
    n = len([k for k in d.files if k.startswith("p_")])
    for i in range(n):
        a, b, p = d[f"a_{i}"], d[f"b_{i}"], d[f"p_{i}"]
        assert np.array_equal(hgx.mul_limbs(a, b), p), i
        if np.array_equal(a, b):
            assert np.array_equal(hgx.sqr_limbs(a), p), i


@pytest.mark.parametrize("n", [1, 7, 16, 17, 2048, 2049, 65536, 300000, 1 << 20])
def test_all_ones_carry_chains(hgx, oracle, n):
    # (2^(64n) - 1)^2 has ~n all-ones limbs above a zero run: the longest carry chains
    say(f"all-ones {n}")
    a = np.full(n, 2**64 - 1, dtype=np.uint64)
    got = hgx.sqr_limbs(a)
    want = np.zeros(2 * n, dtype=np.uint64)
    want[0] = 1
    want[n] = 2**64 - 2
    want[n + 1:] = 2**64 - 1
    assert np.array_equal(got, want)
    got2 = hgx.mul_limbs(a, a.copy())
    assert np.array_equal(got2, want)


@pytest.mark.parametrize("na,nb", [(3, 5), (5000, 3), (70000, 70000)])
def test_all_ones_times_mixed(hgx, oracle, na, nb):
    a = np.full(na, 2**64 - 1, dtype=np.uint64)
    b = oracle.splitmix64(na + nb, nb)
    got = hgx.mul_limbs(a, b)
    if na + nb <= 200000:
        assert np.array_equal(got, oracle.mpz_mul_ntt(a, b)[: na + nb])


@pytest.mark.skipif(not _gmp.available(), reason="libgmp.so.10 not present")
@pytest.mark.parametrize("na,nb", [(1 << 18, 1 << 18), ((1 << 20) + 3, 12345), (1 << 21, 1 << 20)])
def test_large_products_match_gmp(hgx, oracle, na, nb):
    say(f"gmp {na}x{nb}")
    a = oracle.splitmix64(na, na)
    b = oracle.splitmix64(nb + 1, nb)
    assert np.array_equal(hgx.mul_limbs(a, b), _gmp.mul(a, b))


def test_headline_size_checksum(hgx, oracle):
    # 2^26 x 2^26 limbs, L = 2^27: product residue mod 2^61-1 equals the
    # product of the operands' residues; top and bottom limbs exact.
    n = 1 << 26
    say("checksum 2^26 x 2^26")
    a = oracle.splitmix64(0x48475801, n)
    b = oracle.splitmix64(0x48475802, n)
    got = hgx.mul_limbs(a, b)
    assert mod_q(got) == (mod_q(a) * mod_q(b)) % Q
    # lowest limbs: a*b mod 2^(64*4) from the low limbs alone
    lo = (oracle.limbs_to_int(a[:4]) * oracle.limbs_to_int(b[:4])) % (1 << 256)
    assert oracle.limbs_to_int(got[:4]) == lo
    assert got[-1] != 0 or got[-2] != 0


def test_host_transfers_through_staging(hgx, oracle):
    # operands, product and transform above the 32 MiB staging chunk: uploads
    # overlap their DMA with the next chunk's copy, downloads are copied out
    # when a buffer is reused or at the end of the call
    na, nb = (1 << 22) + 12345, (1 << 22) - 777
    say(f"staged transfers {na}x{nb}")
    a = oracle.splitmix64(0x5170, na)
    b = oracle.splitmix64(0x5171, nb)
    assert np.array_equal(hgx.mul_limbs(a, b), _gmp.mul(a, b))
    assert np.array_equal(hgx.sqr_limbs(a), _gmp.mul(a, a.copy()))
    x = residues(oracle, 1, 1 << 23, 0x5172)
    assert np.array_equal(hgx.ntt_reverse(hgx.ntt_forward(x, 1), 1), x)


def test_concurrent_host_calls_from_threads(hgx, oracle):
    # the reference's :hypergeometrica-parallel build calls mpz-* from
    # lparallel workers: eight threads call hgx_mul / hgx_sqr at once, with
    # sizes on both sides of the 4 MiB staging threshold
    import threading
    shapes = [(3000, 2000), ((1 << 19) + 5, 1 << 19), (100000, 31072), (1 << 16, 1 << 16)]
    say("concurrent host calls, 8 threads")
    results, errors = {}, []

    def worker(w):
        try:
            for r, (na, nb) in enumerate(shapes):
                a = oracle.splitmix64(0x7000 + 16 * w + r, na)
                b = oracle.splitmix64(0x7800 + 16 * w + r, nb)
                if (w + r) % 3 == 0:
                    results[(w, r)] = (hgx.sqr_limbs(a), _gmp.mul(a, a.copy()))
                else:
                    results[(w, r)] = (hgx.mul_limbs(a, b), _gmp.mul(a, b))
        except Exception as e:  # reported below with the worker's index
            errors.append((w, repr(e)))

    threads = [threading.Thread(target=worker, args=(w,)) for w in range(8)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors
    assert len(results) == 8 * len(shapes)
    for key, (got, want) in results.items():
        assert np.array_equal(got, want), key


def test_shutdown_during_concurrent_host_calls(hgx, oracle):
    # hgx_mul calls on call slots from four threads while the library shuts
    # down and starts again: every call returns either the exact product or
    # HGX_ENOTINIT, and none hangs (hgx_shutdown waits for calls holding a
    # slot; later calls fail until hgx_init)
    import threading
    from hypergeometrica_amd import _lib
    L = _lib.lib()
    a = oracle.splitmix64(0x5301, 1 << 15)
    b = oracle.splitmix64(0x5302, 3 << 13)
    want = _gmp.mul(a, b)
    stop = threading.Event()
    seen = {"ok": 0, "notinit": 0, "bad": []}

    def worker():
        out = np.empty(a.size + b.size, dtype=np.uint64)
        while not stop.is_set():
            out[:] = 0
            rc = L.hgx_mul(a.ctypes.data, a.size, b.ctypes.data, b.size, out.ctypes.data)
            if rc == 0:
                if np.array_equal(out, want):
                    seen["ok"] += 1
                else:
                    seen["bad"].append("wrong product")
            elif rc == _lib.HGX_ENOTINIT:
                seen["notinit"] += 1
            else:
                seen["bad"].append(rc)

    import gc
    gc.collect()  # no device block of an earlier test is released after the restarts
    threads = [threading.Thread(target=worker) for _ in range(4)]
    for t in threads:
        t.start()
    try:
        for _ in range(3):
            time.sleep(0.2)
            L.hgx_shutdown()
            _lib.generation += 1  # blocks handed out before are gone
            time.sleep(0.05)
            _lib.check(L.hgx_init(0, 0))
        time.sleep(0.2)
    finally:
        stop.set()
        for t in threads:
            t.join(timeout=60)
    assert not any(t.is_alive() for t in threads)
    assert not seen["bad"], seen["bad"][:5]
    assert seen["ok"] > 0


def test_overlapping_product_storage_rejected(hgx):
    from hypergeometrica_amd import _lib
    L = _lib.lib()
    buf = np.zeros(64, dtype=np.uint64)
    buf[:8] = 3
    p = buf.ctypes.data
    assert L.hgx_mul(p, 4, p + 32, 4, p + 16) == _lib.HGX_EINVAL
    assert L.hgx_mul(p, 4, p + 32, 4, p + 40) == _lib.HGX_EINVAL
    assert L.hgx_sqr(p, 4, p + 8) == _lib.HGX_EINVAL
    assert L.hgx_mul(p, 4, p + 32, 4, p + 64) == 0
    # (3 (1 + X + X^2 + X^3))^2 with X = 2^64
    assert buf[8:16].tolist() == [9, 18, 27, 36, 27, 18, 9, 0]


# ---------------------------------------------------------------- batch / serving

def test_batch_equals_per_call(hgx, oracle):
    rng = np.random.default_rng(0x48475803)
    pairs = []
    for i in range(40):
        na = int(2 ** rng.uniform(0, 14))
        nb = int(2 ** rng.uniform(0, 14))
        a = oracle.splitmix64(3 * i + 1, max(na, 1))
        b = a if i % 7 == 0 else oracle.splitmix64(3 * i + 2, max(nb, 1))
        pairs.append((a, b))
    outs = hgx.mul_batch(pairs)
    for (a, b), o in zip(pairs, outs):
        assert np.array_equal(o, hgx.mul_limbs(a, b))
        assert oracle.limbs_to_int(o) == oracle.limbs_to_int(a) * oracle.limbs_to_int(b)


@pytest.mark.parametrize("n", [4, 5, 6, 7, 8, 9, 10])
def test_small_products_packed_per_tile(hgx, oracle, n):
    # batches of one small transform length 2^n (several products per tile,
    # a partial last tile): random, all-ones and square operands against
    # exact integers, for every count from one product to several tiles
    L = 1 << n
    rng = np.random.default_rng(0x48475806 + n)
    for count in (1, 2, 3, 31, 33, 4096 // L, 4096 // L + 1, 257):
        pairs = []
        for i in range(count):
            na = int(rng.integers(L // 4 + 1, L // 2 + 1))
            nb = int(rng.integers(max(1, L // 2 - na + 1), L - na + 1))
            if i % 5 == 1:
                a = np.full(na, np.uint64(2**64 - 1), dtype=np.uint64)
                b = np.full(nb, np.uint64(2**64 - 1), dtype=np.uint64)
            else:
                a = oracle.splitmix64(1000 * n + 3 * i + 1, na)
                b = oracle.splitmix64(1000 * n + 3 * i + 2, nb)
            if i % 4 == 3 and na > L // 4:
                b = a  # a square (same operand)
            pairs.append((a, b))
        outs = hgx.mul_batch(pairs)
        for (a, b), o in zip(pairs, outs):
            assert oracle.limbs_to_int(o) == oracle.limbs_to_int(a) * oracle.limbs_to_int(b), (n, count)


def test_mpz_mirror(hgx):
    import random
    rnd = random.Random(4)
    for _ in range(10):
        a = rnd.randrange(-10**300, 10**300)
        b = rnd.randrange(-10**250, 10**250)
        za, zb = hgx.integer_mpz(a), hgx.integer_mpz(b)
        assert hgx.mpz_integer(hgx.mpz_mul(za, zb)) == a * b
        assert hgx.mpz_integer(hgx.mpz_mul(za, za)) == a * a
        assert hgx.mpz_integer(hgx.mpz_square(za)) == a * a
    assert hgx.mpz_integer(hgx.mpz_expt(hgx.integer_mpz(10), 9001)) == 10**9001


def test_batch_device_grouping_many_lengths(hgx, oracle):
    # 300 products spanning every pass structure, batched in one call
    rng = np.random.default_rng(0x48475804)
    pairs = []
    for i in range(300):
        na = int(2 ** rng.uniform(0, 15.5))
        nb = int(2 ** rng.uniform(0, 15.5))
        a = oracle.splitmix64(7 * i + 1, max(na, 1))
        pairs.append((a, a if i % 5 == 0 else oracle.splitmix64(7 * i + 2, max(nb, 1))))
    outs = hgx.mul_batch(pairs)
    for (a, b), o in zip(pairs, outs):
        assert oracle.limbs_to_int(o) == oracle.limbs_to_int(a) * oracle.limbs_to_int(b)


def test_server_results_alone_and_in_the_mix(hgx, oracle):
    rng = np.random.default_rng(0x48475805)
    reqs = []
    for i in range(60):
        na = int(2 ** rng.uniform(4, 17))
        nb = int(2 ** rng.uniform(4, 17))
        a = oracle.splitmix64(11 * i + 1, na)
        b = None if i % 6 == 0 else oracle.splitmix64(11 * i + 2, nb)
        reqs.append((a, b))
    # alone: one request in flight at a time
    alone = [hgx.submit(a, b).wait().copy() for a, b in reqs]
    # in the mix: all submitted before any is waited on
    tickets = [hgx.submit(a, b) for a, b in reqs]
    mixed = [t.wait() for t in tickets]
    for (a, b), x, y in zip(reqs, alone, mixed):
        assert np.array_equal(x, y)
        bb = a if b is None else b
        assert oracle.limbs_to_int(x) == oracle.limbs_to_int(a) * oracle.limbs_to_int(bb)


def test_server_rejects_empty_operand(hgx):
    from hypergeometrica_amd import HgxError
    with pytest.raises(HgxError):
        hgx.submit(np.zeros(0, dtype=np.uint64), np.ones(3, dtype=np.uint64))


# ---------------------------------------------------------------- callers

def test_binary_splitting_gpu_equals_depth_first(hgx):
    from hypergeometrica_amd import series as S
    from oracle import series_oracle as SO
    s = S.chudnovsky_series()
    got = S.binary_split(s, 0, 300)
    assert (got.lower, got.upper, got.p, got.q, got.b, got.r) == SO.binary_split(s, 0, 300)


@pytest.mark.parametrize("bits,truncate", [(1000, False), (10000, False), (1 << 16, True)])
def test_mpd_pi_gpu_bit_identical_to_cpu_routing(hgx, bits, truncate):
    from hypergeometrica_amd import mpd as M
    gpu = M.Context(truncate=truncate).mpd_pi(bits)
    cpu = M.Context(mul=lambda x, y: x * y, truncate=truncate).mpd_pi(bits)
    assert gpu == cpu


@pytest.mark.parametrize("bits,truncate,device_heights", [
    (1000, False, "0"), (10000, False, "1"), (1 << 16, True, "1"), (1 << 18, True, "0"), (1 << 18, True, "1"),
])
def test_mpd_pi_device_chain_equals_host_routing(hgx, monkeypatch, bits, truncate, device_heights):
    # DeviceContext.mpd_pi: the native split's limbs go to device memory as
    # they are, B*Q and the Newton steps run there; the same integers as the
    # host iteration with exact Python products
    from hypergeometrica_amd import mpd as M
    from hypergeometrica_amd import mpd_device as D
    monkeypatch.setenv("HGX_SERIES_DEVICE", device_heights)
    dev = D.DeviceContext(truncate=truncate).mpd_pi(bits)
    cpu = M.Context(mul=lambda x, y: x * y, truncate=truncate).mpd_pi(bits)
    assert dev == cpu


@pytest.mark.parametrize("k", [4096, 6000, 1 << 17, 1 << 21])
def test_product_that_is_all_ones(hgx, k):
    # (2^(32k) - 1)(2^(32k) + 1) = 2^(64k) - 1: every limb of the product is
    # all ones, so carries cross every run and every 1024-limb CRT block
    half = k // 2
    a = np.full(half, 2**64 - 1, dtype=np.uint64)
    b = np.zeros(half + 1, dtype=np.uint64)
    b[0] = 1
    b[half] = 1
    got = hgx.mul_limbs(a, b)
    assert got.size == k + 1
    assert np.all(got[:k] == np.uint64(2**64 - 1)) and got[k] == 0


@pytest.mark.parametrize("na,nb", [(3000, 5000), (2047, 2049), (1 << 19, (1 << 19) + 3)])
def test_all_ones_uneven(hgx, oracle, na, nb):
    a = np.full(na, 2**64 - 1, dtype=np.uint64)
    b = np.full(nb, 2**64 - 1, dtype=np.uint64)
    got = hgx.mul_limbs(a, b)
    # (2^A - 1)(2^B - 1) = 2^(A+B) - 2^A - 2^B + 1
    A, B = 64 * na, 64 * nb
    assert oracle.limbs_to_int(got) == (1 << (A + B)) - (1 << A) - (1 << B) + 1


# ---------------------------------------------------------------- four passes

def test_four_pass_transform_impulse(hgx, oracle):
    # L = 2^29: column passes of 8 + 8 + 8 bits and a 5-bit row pass, both
    # twiddle forms.  The inverse transform of the impulse at bit-reversed
    # slot k0 is x_j = L^-1 w^(-j k0) (checked at sampled j against Python's
    # pow); the forward transform must give the impulse back.
    n, k = 29, 1
    L = 1 << n
    p = oracle.moduli()[k]
    w = pow(oracle.generator(p), (p - 1) >> n, p)
    winv, Linv = pow(w, -1, p), pow(L, -1, p)
    k0 = 0x12345679 % L
    say("impulse transform 2^29")
    e = np.zeros(L, dtype=np.uint64)
    e[oracle.bitrev(k0, n)] = 1
    x = hgx.ntt_reverse(e, k)
    rng = np.random.default_rng(0x48475829)
    for j in [0, 1, 2, L - 1] + [int(v) for v in rng.integers(0, L, 300)]:
        assert int(x[j]) == Linv * pow(winv, (j * k0) % (p - 1), p) % p, j
    del e
    y = hgx.ntt_forward(x, k)
    assert int(y[oracle.bitrev(k0, n)]) == 1 and np.count_nonzero(y) == 1


def test_four_pass_product_checksum(hgx, oracle):
    # 2^28 x 2^28 limbs, L = 2^29 (four passes): residue mod 2^61-1 and the
    # exact low limbs, as for the 2^26 product
    n = 1 << 28
    say("checksum 2^28 x 2^28")
    a = oracle.splitmix64(0x48475811, n)
    b = oracle.splitmix64(0x48475812, n)
    got = hgx.mul_limbs(a, b)
    assert mod_q(got) == (mod_q(a) * mod_q(b)) % Q
    lo = (oracle.limbs_to_int(a[:4]) * oracle.limbs_to_int(b[:4])) % (1 << 256)
    assert oracle.limbs_to_int(got[:4]) == lo
    assert got[-1] != 0 or got[-2] != 0


# ---------------------------------------------------------------- Newton in HBM

@pytest.mark.parametrize("bits,truncate", [(1000, False), (8000, False), (20000, True), (1 << 18, True)])
def test_device_newton_equals_host(hgx, bits, truncate):
    # mpd-reciprocal and mpd-inv-sqrt with every iterate in device memory give
    # the integers of the host iteration (same steps, exact arithmetic)
    import random

    from hypergeometrica_amd import mpd as M
    from hypergeometrica_amd import mpd_device as D
    rnd = random.Random(bits)
    host = M.Context(mul=lambda x, y: x * y, truncate=truncate)
    dev = D.DeviceContext(truncate=truncate)
    for mant, exp in [(rnd.getrandbits(bits // 2) | 1, -3), (-(rnd.getrandbits(bits // 3) | 1), 2), (10005, 0)]:
        a = M.MPD(mant, exp)
        say("device newton", bits, truncate, a.mantissa.bit_length())
        assert dev.mpd_reciprocal(a, bits) == host.mpd_reciprocal(a, bits)
        if mant > 0:
            assert dev.mpd_inv_sqrt(a, bits) == host.mpd_inv_sqrt(a, bits)
            assert dev.mpd_sqrt(a, bits) == host.mpd_sqrt(a, bits)


def test_device_add_sub_carry_chains(hgx):
    # magnitude add / subtract across blocks of 1024 limbs: all-ones runs,
    # borrows through zero runs, uneven lengths, in place
    import torch
    from hypergeometrica_amd import _lib
    rnd = np.random.default_rng(5)
    L = _lib.lib()
    s = None  # the library's stream; torch.cuda.synchronize() orders the reads after it

    def val(t):
        return int.from_bytes(t.cpu().numpy().view(np.uint64).astype("<u8").tobytes(), "little")

    for na, nb in [(1, 1), (5000, 3), (4096, 4096), (70000, 65536), (1 << 20, 1 << 20)]:
        for kind in ["random", "ones"]:
            if kind == "ones":
                a = np.full(na, 2**64 - 1, dtype=np.uint64)
                b = np.zeros(nb, dtype=np.uint64)
                b[0] = 1
            else:
                a = rnd.integers(0, 2**64, size=na, dtype=np.uint64)
                b = rnd.integers(0, 2**64, size=nb, dtype=np.uint64)
            A = torch.from_numpy(a.view(np.int64)).cuda()
            B = torch.from_numpy(b.view(np.int64)).cuda()
            n = max(na, nb)
            out = torch.empty(n + 1, dtype=torch.int64, device="cuda")
            torch.cuda.synchronize()
            _lib.check(L.hgx_add_device(A.data_ptr(), na, B.data_ptr(), nb, out.data_ptr(), s))
            torch.cuda.synchronize()
            ia, ib = val(A), val(B)
            assert val(out) == ia + ib
            big, small, nbig, nsmall = (A, B, na, nb) if ia >= ib else (B, A, nb, na)
            out2 = torch.empty(n, dtype=torch.int64, device="cuda")
            _lib.check(L.hgx_sub_device(big.data_ptr(), nbig, small.data_ptr(), nsmall, out2.data_ptr(), s))
            torch.cuda.synchronize()
            assert val(out2) == abs(ia - ib)
    # in place: out aliases the first operand (which has room for the carry limb)
    A = torch.full((4097,), -1, dtype=torch.int64, device="cuda")
    A[4096] = 0
    one = torch.ones(1, dtype=torch.int64, device="cuda")
    torch.cuda.synchronize()  # torch's fills first: the library works on its own stream
    _lib.check(L.hgx_add_device(A.data_ptr(), 4096, one.data_ptr(), 1, A.data_ptr(), s))
    torch.cuda.synchronize()
    assert val(A) == 1 << (64 * 4096)


# ---------------------------------------------------------------- graphs

def test_batch_graph_replays_equal_eager(hgx, oracle):
    # a batched level is captured once and replayed; every replay, with other
    # eager work between, equals the eager launches
    from hypergeometrica_amd import _lib
    L = _lib.lib()
    rng = np.random.default_rng(0x4847580A)
    pairs = []
    for i in range(48):
        na, nb = int(2 ** rng.uniform(0, 15)), int(2 ** rng.uniform(0, 15))
        a = oracle.splitmix64(11 * i + 1, max(na, 1))
        b = a if i % 9 == 0 else oracle.splitmix64(11 * i + 2, max(nb, 1))
        pairs.append((a, b))
    for i in range(40):  # small products: groups of several per tile (packed transforms)
        na, nb = int(rng.integers(20, 200)), int(rng.integers(20, 200))
        pairs.append((oracle.splitmix64(3000 + i, na), oracle.splitmix64(4000 + i, nb)))
    _lib.check(L.hgx_graphs_enable(0))
    eager = hgx.mul_batch(pairs)
    _lib.check(L.hgx_graphs_enable(1))
    try:
        for rep in range(4):
            say("graph replay", rep)
            got = hgx.mul_batch(pairs)
            assert all(np.array_equal(g, e) for g, e in zip(got, eager)), rep
            # eager work between replays: another shape and a single product
            other = hgx.mul_batch(pairs[: 5 + rep])
            assert all(np.array_equal(g, e) for g, e in zip(other, eager[: 5 + rep]))
            x, y = oracle.splitmix64(rep, 3000), oracle.splitmix64(rep + 9, 5000)
            assert oracle.limbs_to_int(hgx.mul_limbs(x, y)) == oracle.limbs_to_int(x) * oracle.limbs_to_int(y)
    finally:
        _lib.check(L.hgx_graphs_enable(1))
    for (a, b), e in zip(pairs, eager):
        assert oracle.limbs_to_int(e) == oracle.limbs_to_int(a) * oracle.limbs_to_int(b)


# ---------------------------------------------------------------- pass structures

def _plan(L, bits):
    import ctypes
    n, d, m = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    from hypergeometrica_amd import _lib
    _lib.check(L.hgx_plan_get(bits, ctypes.byref(n), ctypes.byref(d), ctypes.byref(m)))
    return n.value, d.value, m.value


@pytest.mark.parametrize("bits", [16, 20, 21, 28])
def test_every_pass_structure_gives_the_same_results(hgx, oracle, bits):
    # each column-pass count, direct-table limit and direct Montgomery flag:
    # products against GMP (or the mod 2^61-1 residue at 2^28), squares, and
    # transforms against the oracle; the default is restored after.  At 2^28
    # the direct-table limits change only the 2^20 pass: one is enough.  At
    # 2^16 every launch is small: the 512-thread column kernels run with each
    # twiddle form.
    from hypergeometrica_amd import _lib
    L = _lib.lib()
    default = _plan(L, bits)
    lo, hi = (bits - 12 + 7) // 8, (bits - 4) // 8
    limits = [12, 16, 20, 22] if bits <= 21 else [12]
    try:
        for ncol, dm, md in [(c, d, m) for c in range(lo, hi + 1) for d in limits for m in [0, 1]]:
            say("plan", bits, ncol, dm, md)
            _lib.check(L.hgx_plan_set(bits, ncol, dm, md))
            assert _plan(L, bits) == (ncol, dm, md)
            a = oracle.splitmix64(bits * 100 + ncol, (1 << (bits - 1)) - 7)
            b = oracle.splitmix64(dm + md, (1 << (bits - 1)) - 3)
            got = hgx.mul_limbs(a, b)
            if bits <= 21:
                assert np.array_equal(got, _gmp.mul(a, b))
                assert np.array_equal(hgx.sqr_limbs(a), _gmp.mul(a, a))
                x = residues(oracle, 1, 1 << bits, 0x485 + dm)
                assert np.array_equal(hgx.ntt_forward(x, 1), oracle.ntt_forward(x, 1))
                assert np.array_equal(hgx.ntt_reverse(hgx.ntt_forward(x, 1), 1), x)
            else:
                assert mod_q(got) == (mod_q(a) * mod_q(b)) % Q
    finally:
        _lib.check(L.hgx_plan_set(bits, *default))


def test_autotune_keeps_valid_choices(hgx, oracle, tmp_path, monkeypatch):
    import ctypes
    from hypergeometrica_amd import _lib
    L = _lib.lib()
    cache = tmp_path / "tune.txt"
    monkeypatch.setenv("HGX_TUNE_CACHE", str(cache))
    saved = {b: _plan(L, b) for b in range(18, 22)}
    saved_k = {b: _kernels(L, b) for b in range(18, 22)}
    try:
        _lib.check(L.hgx_autotune(18, 21, 1))
        lines = cache.read_text().split("\n")
        rows = [ln.split() for ln in lines if ln.strip()]
        assert rows and all(len(r) == 7 for r in rows)
        for b in range(18, 22):
            ncol, dm, md = _plan(L, b)
            assert (b - 12 + 7) // 8 <= ncol <= (b - 4) // 8 and 12 <= dm <= 24 and md in (0, 1)
            colw, rowt = _kernels(L, b)
            assert colw in (0, 256, 512) and rowt in (0, 10, 11, 12)
            assert b >= 20 or (colw, rowt) == (0, 0)
            na = (1 << (b - 1)) - 5
            a, bb = oracle.splitmix64(b, na), oracle.splitmix64(b + 50, na + 1)
            assert np.array_equal(hgx.mul_limbs(a, bb), _gmp.mul(a, bb))
    finally:
        for b, c in saved.items():
            _lib.check(L.hgx_plan_set(b, *c))
            _lib.check(L.hgx_kernel_set(b, *saved_k[b]))


def _kernels(L, bits):
    import ctypes
    from hypergeometrica_amd import _lib
    c, r = ctypes.c_int(), ctypes.c_int()
    _lib.check(L.hgx_kernel_get(bits, ctypes.byref(c), ctypes.byref(r)))
    return c.value, r.value


@pytest.mark.parametrize("bits", [17, 20, 22, 27])
def test_every_kernel_choice_gives_the_same_products(hgx, oracle, bits):
    # column tiles of 256 and 512 threads and row tiles of 2^12, 2^11 and
    # 2^10 words, on each column-pass count: products against GMP (the mod
    # 2^61-1 residue and exact low limbs at 2^27), squares, and the default
    # kernels restored after
    from hypergeometrica_amd import _lib
    L = _lib.lib()
    default, default_k = _plan(L, bits), _kernels(L, bits)
    lo, hi = (bits - 12 + 7) // 8, (bits - 4) // 8
    try:
        for ncol in range(lo, hi + 1):
            srow = bits - 8 * ncol
            _lib.check(L.hgx_plan_set(bits, ncol, default[1], default[2]))
            for colw in (256, 512):
                for rowt in (12, 11, 10):
                    if srow > rowt:
                        assert L.hgx_kernel_set(bits, colw, rowt) == 0  # accepted, falls back to 4096-word tiles
                    _lib.check(L.hgx_kernel_set(bits, colw, rowt))
                    assert _kernels(L, bits) == (colw, rowt)
                    say("kernels", bits, ncol, colw, rowt)
                    a = oracle.splitmix64(bits * 1000 + colw + rowt, (1 << (bits - 1)) - 11)
                    b = oracle.splitmix64(ncol * 7 + rowt, (1 << (bits - 1)) - 2)
                    got = hgx.mul_limbs(a, b)
                    if bits <= 22:
                        assert np.array_equal(got, _gmp.mul(a, b)), (ncol, colw, rowt)
                        assert np.array_equal(hgx.sqr_limbs(b), _gmp.mul(b, b.copy())), (ncol, colw, rowt)
                    else:
                        from tests import _residue as RES
                        assert RES.residue(got) == RES.mulmod(RES.residue(a), RES.residue(b)), (ncol, colw, rowt)
                        lo4 = (oracle.limbs_to_int(a[:4]) * oracle.limbs_to_int(b[:4])) % (1 << 256)
                        assert oracle.limbs_to_int(got[:4]) == lo4
        assert L.hgx_kernel_set(bits, 384, 0) == _lib.HGX_EINVAL
        assert L.hgx_kernel_set(bits, 0, 9) == _lib.HGX_EINVAL
    finally:
        _lib.check(L.hgx_plan_set(bits, *default))
        _lib.check(L.hgx_kernel_set(bits, *default_k))


# ---------------------------------------------------------------- unbalanced products

UNBALANCED = [(300000, 3), (300000, 1500), (1 << 20, 1000), (3000000, 5000), (1 << 22, 4096), (1 << 23, 65536)]


def test_unbalanced_products_split_equal_gmp_and_unsplit(hgx, oracle):
    # products split into chunk products (and an image sum) against GMP and
    # against the one-transform path, both operand orders
    from hypergeometrica_amd import _lib
    L = _lib.lib()
    for na, nb in UNBALANCED:
        say("unbalanced", na, nb)
        a = oracle.splitmix64(na + 17, na)
        b = oracle.splitmix64(nb + 23, nb)
        got = hgx.mul_limbs(a, b)
        assert np.array_equal(got, _gmp.mul(a, b)), (na, nb)
        assert np.array_equal(hgx.mul_limbs(b, a), got)
        _lib.check(L.hgx_split_enable(0))
        try:
            assert np.array_equal(hgx.mul_limbs(a, b), got)
        finally:
            _lib.check(L.hgx_split_enable(1))
    # all-ones operands: carries between the two images cross every chunk
    a = np.full(1 << 20, 2**64 - 1, dtype=np.uint64)
    b = np.full(3000, 2**64 - 1, dtype=np.uint64)
    A, B = 64 * a.size, 64 * b.size
    assert oracle.limbs_to_int(hgx.mul_limbs(a, b)) == (1 << (A + B)) - (1 << A) - (1 << B) + 1


def test_unbalanced_in_batches_and_server(hgx, oracle):
    pairs = [(oracle.splitmix64(i, na), oracle.splitmix64(i + 50, nb))
             for i, (na, nb) in enumerate(UNBALANCED[:4] + [(3000, 2000), (40000, 41000)])]
    outs = hgx.mul_batch(pairs)
    for (a, b), o in zip(pairs, outs):
        assert np.array_equal(o, _gmp.mul(a, b))
    tickets = [hgx.submit(a, b) for a, b in pairs]
    for (a, b), t in zip(pairs, tickets):
        assert np.array_equal(t.wait(), _gmp.mul(a, b))


# ---------------------------------------------------------------- host-staged products
# A limit on the device memory of one product forces the chunked path that
# products beyond HBM take; the kernel timing shows its window additions.

STAGED = [(300001, 200003), (1 << 20, 1 << 20), (1 << 21, 5), (1 << 21, 70000), (777777, 777776)]


@pytest.fixture
def staging_limit(hgx):
    from hypergeometrica_amd import _lib
    L = _lib.lib()
    yield lambda nbytes: _lib.check(L.hgx_staging_limit(nbytes))
    _lib.check(L.hgx_staging_limit(0))


def _mul_staged(hgx, a, b):
    from hypergeometrica_amd import _lib
    _lib.timing_enable(True)
    try:
        got = hgx.mul_limbs(a, b)
        assert "stage_add" in _lib.timing_read(), "the product was not staged"
    finally:
        _lib.timing_enable(False)
    return got


def test_staged_products_equal_gmp(hgx, oracle, staging_limit):
    staging_limit(16 << 20)
    for na, nb in STAGED:
        say("staged", na, nb)
        a = oracle.splitmix64(na + 31, na)
        b = oracle.splitmix64(nb + 37, nb)
        got = _mul_staged(hgx, a, b)
        assert np.array_equal(got, _gmp.mul(a, b)), (na, nb)
        assert np.array_equal(_mul_staged(hgx, b, a), got)
    say("staged square")
    a = oracle.splitmix64(99, 600001)
    assert np.array_equal(_mul_staged(hgx, a, a), _gmp.mul(a, a))
    # all-ones operands: carries run through every window
    say("staged all-ones")
    a = np.full(700000, 2**64 - 1, dtype=np.uint64)
    b = np.full(650001, 2**64 - 1, dtype=np.uint64)
    A, B = 64 * a.size, 64 * b.size
    assert oracle.limbs_to_int(_mul_staged(hgx, a, b)) == (1 << (A + B)) - (1 << A) - (1 << B) + 1
    assert oracle.limbs_to_int(_mul_staged(hgx, a, a)) == (1 << (2 * A)) - (1 << (A + 1)) + 1


def test_staged_in_batches_and_server(hgx, oracle, staging_limit):
    staging_limit(16 << 20)
    pairs = [(oracle.splitmix64(i, na), oracle.splitmix64(i + 70, nb))
             for i, (na, nb) in enumerate([(400000, 300000), (3000, 2000), (1 << 20, 4000), (40000, 41000)])]
    outs = hgx.mul_batch(pairs)
    for (a, b), o in zip(pairs, outs):
        assert np.array_equal(o, _gmp.mul(a, b))
    tickets = [hgx.submit(a, b) for a, b in pairs]
    for (a, b), t in zip(pairs, tickets):
        assert np.array_equal(t.wait(), _gmp.mul(a, b))


def test_staged_headline_size_equals_in_hbm(hgx, oracle, staging_limit):
    # 2^26 x 2^26 limbs within 1 GiB of device memory: identical to the
    # in-HBM product, which needs about 8 GiB
    n = 1 << 26
    say("staged 2^26 x 2^26")
    a = oracle.splitmix64(0x48475801, n)
    b = oracle.splitmix64(0x48475802, n)
    whole = hgx.mul_limbs(a, b)
    staging_limit(1 << 30)
    got = _mul_staged(hgx, a, b)
    assert np.array_equal(got, whole)
    assert mod_q(got) == (mod_q(a) * mod_q(b)) % Q
