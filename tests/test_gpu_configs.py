"""GPU parity at the measured configurations' own shapes and at the largest
transform lengths the library takes:

* the single large product (2^26 x 2^26 limbs, L = 2^27) limb for limb
  against GMP, on the benchmark's own operands;
* the binary-splitting level (4096 products, operands log-uniform in
  [2^8, 2^16] limbs, seed 0x48475803) batched on the device, every product
  against GMP;
* transforms and products at L = 2^30, 2^31 and 2^32, beyond the 32-bit
  offsets of the smaller lengths: an impulse through the inverse and forward
  transforms (closed form), and products by their residues mod 2^61 - 1.

Integer work: bit-exact."""
import sys

import numpy as np
import pytest

from tests import _gmp
from tests import _residue as RES

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.mark.skipif(not _gmp.available(), reason="libgmp.so.10 not present")
def test_headline_product_equals_gmp(hgx):
    import torch

    from hypergeometrica_amd import _lib
    from hypergeometrica_amd import workload as W
    L = _lib.lib()
    n = 1 << 26
    dev = torch.device("cuda:0")
    A = torch.empty(n, dtype=torch.int64, device=dev)
    B = torch.empty(n, dtype=torch.int64, device=dev)
    O = torch.empty(2 * n, dtype=torch.int64, device=dev)
    hgx.fill_splitmix64(A.data_ptr(), n, W.SINGLE_SEED)
    hgx.fill_splitmix64(B.data_ptr(), n, W.SINGLE_SEED + 1)
    _lib.check(L.hgx_mul_device(A.data_ptr(), n, B.data_ptr(), n, O.data_ptr(), None))
    torch.cuda.synchronize()
    a = A.cpu().numpy().view(np.uint64)
    b = B.cpu().numpy().view(np.uint64)
    got = O.cpu().numpy().view(np.uint64)
    del A, B, O
    torch.cuda.empty_cache()
    assert np.array_equal(a, W.splitmix64(W.SINGLE_SEED, n, force_top=True))
    say("GMP 2^26 x 2^26")
    want = _gmp.mul(a, b)
    assert np.array_equal(got, want)


@pytest.mark.skipif(not _gmp.available(), reason="libgmp.so.10 not present")
def test_batched_level_workload_equals_gmp(hgx):
    import torch

    from hypergeometrica_amd import _lib
    from hypergeometrica_amd import workload as W
    L = _lib.lib()
    count = 4096
    na, nb = W.batch_shapes(count, W.BATCH_SEED)
    tot = sum(na) + sum(nb)
    dev = torch.device("cuda:0")
    IN = torch.empty(tot, dtype=torch.int64, device=dev)
    OUT = torch.zeros(tot, dtype=torch.int64, device=dev)
    descs = (_lib.MulDesc * count)()
    at, spans = 0, []
    for i in range(count):
        a = IN.data_ptr() + 8 * at
        hgx.fill_splitmix64(a, na[i], W.BATCH_SEED + 2 * i + 1)
        b = a + 8 * na[i]
        hgx.fill_splitmix64(b, nb[i], W.BATCH_SEED + 2 * i + 2)
        descs[i] = _lib.MulDesc(a, na[i], b, nb[i], OUT.data_ptr() + 8 * at)
        spans.append(at)
        at += na[i] + nb[i]
    _lib.check(L.hgx_mul_batch_device(descs, count, None))
    torch.cuda.synchronize()
    inp = IN.cpu().numpy().view(np.uint64)
    out = OUT.cpu().numpy().view(np.uint64)
    say(f"GMP on {count} batched products, {tot} limbs")
    for i in range(count):
        s = spans[i]
        a, b = inp[s:s + na[i]], inp[s + na[i]:s + na[i] + nb[i]]
        assert np.array_equal(out[s:s + na[i] + nb[i]], _gmp.mul(a, b)), (i, na[i], nb[i])


@pytest.mark.parametrize("bits,k", [(30, 0), (32, 2)])
def test_largest_transforms_impulse(hgx, oracle, bits, k):
    # L = 2^30 and 2^32 (three 8-bit column passes and a row pass of 6 or 8
    # bits): the inverse transform of the impulse at bit-reversed slot k0 is
    # x_j = L^-1 w^(-j k0), checked at sampled j; the forward transform gives
    # the impulse back
    from hypergeometrica_amd import _lib
    assert bits <= _lib.lib().hgx_max_transform_bits()
    L = 1 << bits
    p = oracle.moduli()[k]
    w = pow(oracle.generator(p), (p - 1) >> bits, p)
    winv, Linv = pow(w, -1, p), pow(L, -1, p)
    k0 = 0x9E3779B97 % L
    say(f"impulse transform 2^{bits}")
    e = np.zeros(L, dtype=np.uint64)
    e[oracle.bitrev(k0, bits)] = 1
    x = hgx.ntt_reverse(e, k)
    del e
    rng = np.random.default_rng(0x48475830 + bits)
    for j in [0, 1, 2, L // 2 + 1, L - 1] + [int(v) for v in rng.integers(0, L, 300)]:
        assert int(x[j]) == Linv * pow(winv, (j * k0) % (p - 1), p) % p, j
    y = hgx.ntt_forward(x, k)
    del x
    assert int(y[oracle.bitrev(k0, bits)]) == 1 and np.count_nonzero(y) == 1


@pytest.mark.parametrize("n,square", [((1 << 29) - 12345, False), (1 << 30, True)])
def test_largest_products_checksum(hgx, oracle, n, square):
    # a 2^29-limb product (L = 2^30) and the square of a 2^30-limb operand
    # (L = 2^31), checked by residues mod 2^61 - 1 and their exact low limbs
    say(f"{'square' if square else 'product'} of {n} limbs")
    a = oracle.splitmix64(0x48475840 + n, n)
    if square:
        got = hgx.sqr_limbs(a)
        b = a
    else:
        b = oracle.splitmix64(0x48475841 + n, n + 777)
        got = hgx.mul_limbs(a, b)
    assert got.size == a.size + b.size
    assert RES.residue(got) == RES.mulmod(RES.residue(a), RES.residue(b))
    lo = (oracle.limbs_to_int(a[:4]) * oracle.limbs_to_int(b[:4])) % (1 << 256)
    assert oracle.limbs_to_int(got[:4]) == lo
    hi = oracle.limbs_to_int(a[-2:]) * oracle.limbs_to_int(b[-2:])
    top = oracle.limbs_to_int(got[-4:])
    # the top four limbs hold the product of the top two limbs of each
    # operand, plus at most the carry of everything below
    assert hi <= top <= hi + (1 << 130)


def test_mpd_pi_2e24_bits_three_ways(hgx):
    # mpd-pi to 2^24 bits with truncating Newton: GPU routings with host and
    # with HBM iterates bit-identical to GMP-routed host products; pi's hex
    # digits at the end, middle and start equal BBP's; the leading 2^24 bits
    # equal MPFR's
    from hypergeometrica_amd import mpd as M
    from hypergeometrica_amd import mpd_device as D
    from tests import _mpfr
    from tests import _pi_check as PC
    bits = 1 << 24
    say("mpd-pi 2^24 bits")
    r = M.Context(truncate=True).mpd_pi(bits)
    assert D.DeviceContext(truncate=True).mpd_pi(bits) == r
    if _gmp.available():
        assert M.Context(mul=PC.gmp_mul, truncate=True).mpd_pi(bits) == r
    dist = PC.bbp_check(r.mantissa, r.exponent, PC.bbp_positions(bits))
    assert PC.bbp_ok(dist), dist
    if _mpfr.available():
        assert PC.mpfr_leading_bits_match(r.mantissa, r.exponent, bits)
