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
