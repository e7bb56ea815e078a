"""CPU checks of the workload generators (hypergeometrica_amd.workload) and
of the mod 2^61 - 1 residue checker (tests/_residue.c) that the GPU tests and
the benchmark's serving leg use."""
import numpy as np

from tests import _residue as R
from tests._checksum import Q, mod_q


def test_splitmix64_matches_the_oracle_streams(oracle):
    from hypergeometrica_amd import workload as W
    for seed in (0, 1, 0x48475801, 2**64 - 1):
        assert W.splitmix64(seed, 1000).tolist() == oracle.splitmix64_py(seed, 1000)
        assert np.array_equal(W.splitmix64(seed, 5000), oracle.splitmix64(seed, 5000))
    assert W.splitmix64(7, 0).size == 0


def test_log_uniform_sizes_and_gaps():
    from hypergeometrica_amd import workload as W
    na, nb = W.batch_shapes(4096)
    assert len(na) == len(nb) == 4096
    assert min(na + nb) >= 2**8 and max(na + nb) < 2**16
    # log-uniform: about half of the operands below 2^12
    assert 0.45 < sum(n < 2**12 for n in na + nb) / 8192 < 0.55
    gaps = W.poisson_gaps(0x48475805, 20000, 2000.0)
    assert min(gaps) > 0
    assert abs(np.mean(gaps) - 1 / 2000.0) < 0.03 / 2000.0
    sa, sb = W.serve_shapes(2000)
    assert min(sa + sb) >= 2**10 and max(sa + sb) < 2**24


def test_residue_checker_matches_python_integers():
    rng = np.random.default_rng(0x52)
    for n in (0, 1, 2, 60, 61, 62, 122, 123, 1000, 65537):
        a = rng.integers(0, 2**64, size=n, dtype=np.uint64)
        v = int.from_bytes(a.astype("<u8").tobytes(), "little")
        assert R.residue(a) == v % Q
        if n:
            assert R.residue(a) == mod_q(a)
        P = R.prefix_residues(a)
        for off, m in ((0, n), (n // 3, n // 2), (n // 2, n - n // 2), (n, 0)):
            w = int.from_bytes(a[off:off + m].astype("<u8").tobytes(), "little") if m else 0
            assert R.slice_residue(P, off, m) == w % Q
    ones = np.full(777, 2**64 - 1, dtype=np.uint64)
    assert R.residue(ones) == (2**(64 * 777) - 1) % Q
    # a product: (a b) mod q = (a mod q)(b mod q) mod q, and a wrong limb is caught
    a = rng.integers(0, 2**64, size=300, dtype=np.uint64)
    b = rng.integers(0, 2**64, size=200, dtype=np.uint64)
    p = int.from_bytes(a.astype("<u8").tobytes(), "little") * int.from_bytes(b.astype("<u8").tobytes(), "little")
    limbs = np.frombuffer(p.to_bytes(8 * 500, "little"), dtype="<u8").astype(np.uint64)
    assert R.residue(limbs) == R.mulmod(R.residue(a), R.residue(b))
    limbs[137] ^= np.uint64(1 << 20)
    assert R.residue(limbs) != R.mulmod(R.residue(a), R.residue(b))
    # swapped limbs 61 apart have the same weight; 1 apart, not
    x = rng.integers(0, 2**64, size=200, dtype=np.uint64)
    y = x.copy()
    y[[5, 6]] = y[[6, 5]]
    assert x[5] == x[6] or R.residue(x) != R.residue(y)


def test_pi_checkers_pinned_by_mpmath():
    # BBP digit extraction, MPFR's pi and the MPD scaling agree with mpmath's
    # pi; a value off in one bit fails them
    import mpmath

    from tests import _bbp, _mpfr
    from tests import _pi_check as PC
    prec = 20000
    mpmath.mp.prec = prec + 256
    # pi/10 as an MPD of 64-bit digits, truncated to prec bits
    e = -(prec // 64)
    mant = int(mpmath.floor(mpmath.pi / 10 * mpmath.mpf(2) ** (-64 * e)))
    for d in PC.bbp_positions(prec) + [0, 1, 100]:
        g = int(mpmath.floor(mpmath.frac(mpmath.pi * mpmath.mpf(16) ** d) * mpmath.mpf(2) ** 64))
        assert _bbp.matches(d, g), d
    dist = PC.bbp_check(mant, e, PC.bbp_positions(prec))
    assert PC.bbp_ok(dist), dist
    bad = mant ^ (1 << (prec - 4 * PC.bbp_positions(prec)[0] - 40))
    assert not PC.bbp_ok(PC.bbp_check(bad, e, PC.bbp_positions(prec)))
    if _mpfr.available():
        assert PC.mpfr_leading_bits_match(mant, e, 16000)
        assert not PC.mpfr_leading_bits_match(mant ^ (1 << (prec - 12000)), e, 16000)
    from tests import _gmp
    if _gmp.available():
        assert PC.gmp_mul(-(3 ** 5000), 7 ** 3000) == -(3 ** 5000) * 7 ** 3000
