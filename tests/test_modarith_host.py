"""The device's modular arithmetic (hypergeometrica_amd/csrc/modarith.hpp),
compiled for the host with g++ and checked against exact integers.

The header's functions are host-and-device: the host build computes the same
32-bit formulas (v_mul_u32_u24 is emulated as the low 32 bits of a 24 x 24-bit
product), so these tests pin the exact instruction sequences the kernels run:
Shoup's product with the split middle column and the q*c term of the Proth
primes, the special Montgomery REDC, the select-based conditional subtraction
and the additions the butterflies use.  Expected values: (a*b) mod p and
a*b*2^-64 mod p with Python integers (the reference's m*/fast and m*/fast2
compute the same residues, src/modular-arithmetic.lisp:167-231)."""
import ctypes
import os
import random
import shutil
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HEADER_DIR = os.path.join(os.path.dirname(HERE), "hypergeometrica_amd", "csrc")
PRIMES = [0x6280000000000001, 0x5700000000000001, 0x4180000000000001]
M64 = (1 << 64) - 1

HARNESS = r"""
#include <stddef.h>
#include "modarith.hpp"
static Mod mk(u64 p) { return Mod{p, 0 - p, (u32)((p - 1) >> 55)}; }
extern "C" {
void v_shoup(const u64 *a, const u64 *w, const u64 *wq, u64 *out, size_t n, u64 p) {
  const Mod M = mk(p);
  for (size_t i = 0; i < n; i++) out[i] = shoup_mul(a[i], w[i], wq[i], M);
}
void v_mont(const u64 *a, const u64 *b, u64 *out, size_t n, u64 p) {
  const Mod M = mk(p);
  for (size_t i = 0; i < n; i++) out[i] = mont_mul(a[i], b[i], M);
}
void v_add(const u64 *a, const u64 *b, u64 *out, size_t n, u64 p) {
  const Mod M = mk(p);
  for (size_t i = 0; i < n; i++) out[i] = mod_add(a[i], b[i], M);
}
void v_sub(const u64 *a, const u64 *b, u64 *out, size_t n, u64 p) {
  const Mod M = mk(p);
  for (size_t i = 0; i < n; i++) out[i] = mod_sub(a[i], b[i], M);
}
void v_csub(const u64 *a, u64 *out, size_t n, u64 p) {
  const Mod M = mk(p);
  for (size_t i = 0; i < n; i++) out[i] = csub(a[i], M);
}
void v_reduce(const u64 *a, u64 *out, size_t n, u64 p) {
  const Mod M = mk(p);
  for (size_t i = 0; i < n; i++) out[i] = mod_reduce64(a[i], M);
}
void v_mul128(const u64 *a, const u64 *b, u64 *lo, u64 *hi, size_t n) {
  for (size_t i = 0; i < n; i++) mul64x64(a[i], b[i], lo[i], hi[i]);
}
void v_table(const u64 *e, u64 *value, u64 *comp, size_t n, u64 rm, u64 sm, u64 one, u64 pinv, u64 p) {
  const Mod M = mk(p);
  for (size_t i = 0; i < n; i++) pow_table_entry(e[i], rm, sm, one, pinv, M, value[i], comp[i]);
}
}
"""


@pytest.fixture(scope="module")
def ma(tmp_path_factory):
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("modarith")
    src = d / "h.cpp"
    src.write_text(HARNESS)
    so = d / "libmodarith_host.so"
    subprocess.run([cxx, "-O2", "-std=c++17", "-fPIC", "-shared", "-I", HEADER_DIR, "-o", str(so), str(src)],
                   check=True)
    L = ctypes.CDLL(str(so))
    u = ctypes.c_uint64
    P = ctypes.c_void_p
    L.v_shoup.argtypes = [P, P, P, P, ctypes.c_size_t, u]
    for f in ("v_mont", "v_add", "v_sub"):
        getattr(L, f).argtypes = [P, P, P, ctypes.c_size_t, u]
    for f in ("v_csub", "v_reduce"):
        getattr(L, f).argtypes = [P, P, ctypes.c_size_t, u]
    L.v_mul128.argtypes = [P, P, P, P, ctypes.c_size_t]
    L.v_table.argtypes = [P, P, P, ctypes.c_size_t, u, u, u, u, u]
    return L


def arr(xs):
    return np.array(xs, dtype=np.uint64)


def ptr(a):
    return a.ctypes.data


def call(fn, *args):
    """FN on uint64 arrays made from lists (kept alive across the call) and
    the other arguments as given"""
    keep = [arr(x) if isinstance(x, list) else x for x in args]
    fn(*[ptr(x) if isinstance(x, np.ndarray) else x for x in keep])


def edge_words(p):
    return [0, 1, 2, p - 2, p - 1, p, p + 1, 2 * p - 1, 2 * p, (1 << 63) - 1, 1 << 63, M64 - 1, M64,
            (1 << 32) - 1, 1 << 32, 0xFFFFFFFF00000000, 0x00000000FFFFFFFF | (1 << 62)]


def cases(p, n, seed, a_max, b_max):
    """every pair of edge words in range, then N random pairs"""
    rnd = random.Random(seed)
    ea = [x for x in edge_words(p) if x <= a_max]
    eb = [x for x in edge_words(p) if x <= b_max]
    a = [x for x in ea for _ in eb] + [rnd.randrange(a_max + 1) for _ in range(n)]
    b = [y for _ in ea for y in eb] + [rnd.randrange(b_max + 1) for _ in range(n)]
    return a, b


@pytest.mark.parametrize("k", [0, 1, 2])
def test_shoup_product_any_64_bit_multiplicand(ma, k):
    # Shoup's product takes any 64-bit a (the DIF butterflies pass
    # u + p - v < 2p, lazy inverse outputs are below 2p) and w < p
    p = PRIMES[k]
    a, w = cases(p, 20000, 100 + k, M64, p - 1)
    wq = [(x << 64) // p for x in w]
    out = np.zeros(len(a), dtype=np.uint64)
    call(ma.v_shoup, a, w, wq, out, len(a), p)
    want = [(x * y) % p for x, y in zip(a, w)]
    assert out.tolist() == want


@pytest.mark.parametrize("k", [0, 1, 2])
def test_montgomery_product(ma, k):
    # a < 2^64, b < p: a b 2^-64 mod p, canonical
    p = PRIMES[k]
    a, b = cases(p, 20000, 200 + k, M64, p - 1)
    out = np.zeros(len(a), dtype=np.uint64)
    call(ma.v_mont, a, b, out, len(a), p)
    rinv = pow(1 << 64, -1, p)
    assert out.tolist() == [(x * y * rinv) % p for x, y in zip(a, b)]


@pytest.mark.parametrize("k", [0, 1, 2])
def test_additions_and_reductions(ma, k):
    p = PRIMES[k]
    a, b = cases(p, 20000, 300 + k, p - 1, p - 1)
    out = np.zeros(len(a), dtype=np.uint64)
    call(ma.v_add, a, b, out, len(a), p)
    assert out.tolist() == [(x + y) % p for x, y in zip(a, b)]
    call(ma.v_sub, a, b, out, len(a), p)
    assert out.tolist() == [(x - y) % p for x, y in zip(a, b)]
    rnd = random.Random(400 + k)
    r = [x for x in edge_words(p) if x < 2 * p] + [rnd.randrange(2 * p) for _ in range(20000)]
    out = np.zeros(len(r), dtype=np.uint64)
    call(ma.v_csub, r, out, len(r), p)
    assert out.tolist() == [x % p for x in r]
    x = edge_words(p) + [random.Random(500 + k).getrandbits(64) for _ in range(20000)]
    out = np.zeros(len(x), dtype=np.uint64)
    call(ma.v_reduce, x, out, len(x), p)
    assert out.tolist() == [v % p for v in x]


def test_full_64x64_product(ma):
    rnd = random.Random(600)
    a = [0, 1, M64, M64, 1 << 63] + [rnd.getrandbits(64) for _ in range(20000)]
    b = [0, M64, 1, M64, 1 << 63] + [rnd.getrandbits(64) for _ in range(20000)]
    lo = np.zeros(len(a), dtype=np.uint64)
    hi = np.zeros(len(a), dtype=np.uint64)
    call(ma.v_mul128, a, b, lo, hi, len(a))
    assert [int(h) << 64 | int(l) for h, l in zip(hi, lo)] == [x * y for x, y in zip(a, b)]


def test_planted_mistake_in_shoup_companion_is_caught(ma):
    # a companion one too small (floor(w 2^64 / p) - 1) leaves the quotient
    # short by one for some a: the result is then off by p, which this pin sees
    p = PRIMES[0]
    rnd = random.Random(700)
    a = [rnd.getrandbits(64) for _ in range(5000)]
    w = [rnd.randrange(1, p) for _ in range(5000)]
    wq = [max(0, ((x << 64) // p) - 1) for x in w]
    out = np.zeros(len(a), dtype=np.uint64)
    call(ma.v_shoup, a, w, wq, out, len(a), p)
    assert out.tolist() != [(x * y) % p for x, y in zip(a, w)]


@pytest.mark.parametrize("k", [0, 1, 2])
def test_twiddle_table_entries(ma, k):
    # the root tables' entries as the device fills them (k_pow_table):
    # SCALE * R^E mod p and its Shoup companion floor(value 2^64 / p), for
    # exponents up to 2^27 (the largest i*j of a table) and the scales the
    # plans use (1, the inverse of L, R/L)
    p = PRIMES[k]
    rnd = random.Random(800 + k)
    g = [3, 5, 3][k]
    mont = lambda x: (x << 64) % p
    pinv = pow(p, -1, 1 << 64)
    for bits, scale in [(27, 1), (20, pow(1 << 20, -1, p)), (13, ((1 << 64) * pow(1 << 13, -1, p)) % p),
                        (8, rnd.randrange(1, p))]:
        r = pow(g, (p - 1) >> bits, p)
        e = [0, 1, 2, (1 << bits) - 1, 1 << bits, (1 << 27) - 1] + [rnd.randrange(1 << 27) for _ in range(3000)]
        val = np.zeros(len(e), dtype=np.uint64)
        comp = np.zeros(len(e), dtype=np.uint64)
        call(ma.v_table, e, val, comp, len(e), mont(r), mont(scale), mont(1), pinv, p)
        want = [(scale * pow(r, x, p)) % p for x in e]
        assert val.tolist() == want
        assert comp.tolist() == [(v << 64) // p for v in want]
