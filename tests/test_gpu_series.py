"""GPU checks of the native binary-splitting engine (hgx_series_split) on
intervals whose upper tree heights multiply on the GPU: every partial equals
the Python level-synchronous split with GPU products (series.binary_split)
and, where it finishes quickly, the oracle's depth-first split with Python
integers (src/binary-splitting.lisp:100-133)."""
import pytest

from hypergeometrica_amd import series as S
from oracle import series_oracle as SO

pytestmark = pytest.mark.gpu

SERIES = {
    "chudnovsky": S.chudnovsky_series, "ramanujan": S.ramanujan_series, "e": S.e_series,
    "catalan": S.catalan_series, "apery": S.apery_series, "exp(7/2)": lambda: S.exp_series(7, 2),
    "exp(-5/3)": lambda: S.exp_series(-5, 3),
}


def _tuple(x):
    return (x.lower, x.upper, x.p, x.q, x.b, x.r)


@pytest.mark.parametrize("name,lo,up", [
    ("chudnovsky", 0, 3000), ("ramanujan", 0, 3000), ("e", 0, 5000), ("catalan", 0, 4000), ("apery", 0, 2000),
    ("exp(7/2)", 0, 5000), ("exp(-5/3)", 17, 4017), ("chudnovsky", 1000, 2999),
])
def test_native_equals_depth_first(hgx, name, lo, up):
    s = SERIES[name]()
    assert _tuple(S.binary_split_native(s, lo, up)) == SO.binary_split(s, lo, up)


@pytest.mark.parametrize("name,lo,up", [
    ("chudnovsky", 0, 60000), ("ramanujan", 0, 40000), ("e", 0, 100000), ("catalan", 0, 60000),
    ("apery", 0, 30000), ("exp(7/2)", 0, 80000), ("chudnovsky", 777777, 800001),
])
def test_native_equals_python_levels(hgx, name, lo, up):
    s = SERIES[name]()
    assert _tuple(S.binary_split_native(s, lo, up)) == _tuple(S.binary_split(s, lo, up))


def test_native_chudnovsky_pi_digits(hgx):
    # compute-pi/chudnovsky (src/pi.lisp:45-51) with the native partial: the
    # digits of pi/10 against mpmath
    import mpmath
    from math import isqrt
    prec = 20000
    n = int(2 + prec / S.CHUD_DECIMALS_PER_TERM)
    comp = S.binary_split_native(S.chudnovsky_series(), 0, n)
    val = (isqrt(10005 * 100 ** prec) * (comp.b * comp.q)) // comp.r
    mpmath.mp.dps = prec + 30
    want = int(mpmath.floor(mpmath.pi / 10 * mpmath.mpf(10) ** prec))
    assert abs(val - want) <= 1


@pytest.mark.parametrize("mode", ["0", "1"])
@pytest.mark.parametrize("name,lo,up", [
    ("chudnovsky", 0, 3000), ("ramanujan", 0, 3000), ("e", 0, 5000), ("catalan", 0, 4000), ("apery", 0, 2000),
    ("exp(7/2)", 0, 5000), ("exp(-5/3)", 17, 4017), ("chudnovsky", 1000, 2999),
])
def test_native_host_and_device_heights_equal_depth_first(hgx, monkeypatch, mode, name, lo, up):
    # HGX_SERIES_DEVICE=1: every height above the folded leaves in device
    # memory (products batched on the device, signed sums by the device);
    # =0: every height on the host
    monkeypatch.setenv("HGX_SERIES_DEVICE", mode)
    s = SERIES[name]()
    assert _tuple(S.binary_split_native(s, lo, up)) == SO.binary_split(s, lo, up)


# a = 1, p = -1 (1 at k = 0), q = b = 1: R over [lo, up) is a sum of +-1 that
# cancels to zero on many intervals, so device sums meet equal magnitudes
# of opposite signs (a zero result) and zero operands
CANCEL = S.Series(a=S.Poly((1,)), p=S.Poly((-1,), at_zero=1))


@pytest.mark.parametrize("mode", ["0", "1"])
@pytest.mark.parametrize("series,lo,up", [
    (CANCEL, 0, 66), (CANCEL, 0, 130), (CANCEL, 0, 1000), (CANCEL, 5, 1000), (CANCEL, 1, 4097),
    (S.Series(p=S.Poly((-5, 1), at_zero=1), q=S.Poly((0, 3), at_zero=1)), 0, 400),
    (S.Series(a=S.Poly((2, -7)), b=S.Poly((-3,)), p=S.Poly((1, 2, -1)), q=S.Poly((-4, 0, 5), at_zero=-1)), 0, 900),
])
def test_device_sums_signs_and_cancellation(hgx, monkeypatch, mode, series, lo, up):
    monkeypatch.setenv("HGX_SERIES_DEVICE", mode)
    assert _tuple(S.binary_split_native(series, lo, up)) == SO.binary_split(series, lo, up)


@pytest.mark.parametrize("name,lo,up", [("chudnovsky", 0, 60000), ("apery", 0, 30000), ("exp(7/2)", 0, 80000)])
def test_device_heights_equal_python_levels(hgx, monkeypatch, name, lo, up):
    monkeypatch.setenv("HGX_SERIES_DEVICE", "1")
    s = SERIES[name]()
    assert _tuple(S.binary_split_native(s, lo, up)) == _tuple(S.binary_split(s, lo, up))


def test_default_move_to_device_equals_host_heights(hgx, monkeypatch):
    # 400 000 Chudnovsky terms: the upper heights' products exceed 2^20 limbs,
    # so the default rule moves them to device memory
    s = S.chudnovsky_series()
    monkeypatch.delenv("HGX_SERIES_DEVICE", raising=False)
    dev = _tuple(S.binary_split_native(s, 0, 400000))
    monkeypatch.setenv("HGX_SERIES_DEVICE", "0")
    host = _tuple(S.binary_split_native(s, 0, 400000))
    assert dev == host


def test_concurrent_splits_from_threads(hgx, monkeypatch):
    # four threads split at once (device heights forced on, so their rounds,
    # sums and regions interleave on the device); every partial equals the
    # oracle's
    import threading
    monkeypatch.setenv("HGX_SERIES_DEVICE", "1")
    jobs = [("chudnovsky", 0, 3000), ("apery", 0, 2000), ("exp(-5/3)", 17, 4017), ("catalan", 0, 4000)]
    got, errors = {}, []

    def worker(i):
        try:
            name, lo, up = jobs[i]
            for _ in range(3):
                got[i] = _tuple(S.binary_split_native(SERIES[name](), lo, up))
        except Exception as e:  # reported below with the job
            errors.append((jobs[i], repr(e)))

    threads = [threading.Thread(target=worker, args=(i,)) for i in range(len(jobs))]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors
    for i, (name, lo, up) in enumerate(jobs):
        assert got[i] == SO.binary_split(SERIES[name](), lo, up), jobs[i]


@pytest.mark.parametrize("fail_height", [1, 2, 5])
@pytest.mark.parametrize("name,lo,up", [("chudnovsky", 0, 3000), ("exp(-5/3)", 17, 4017)])
def test_device_memory_exhausted_goes_on_on_the_host(hgx, monkeypatch, fail_height, name, lo, up):
    # the device rounds of one height report HGX_ENOMEM (as when device
    # memory runs out): every live value returns to host memory and the split
    # goes on there, with the same partial
    monkeypatch.setenv("HGX_SERIES_DEVICE", "1")
    monkeypatch.setenv("HGX_SERIES_FAIL_HEIGHT", str(fail_height))
    s = SERIES[name]()
    assert _tuple(S.binary_split_native(s, lo, up)) == SO.binary_split(s, lo, up)


@pytest.mark.parametrize("name,lo,up", [("chudnovsky", 0, 3000), ("catalan", 0, 4000), ("exp(-5/3)", 17, 4017)])
def test_leaves_too_large_for_the_device_fold_fold_on_the_host(hgx, monkeypatch, name, lo, up):
    # HGX_SERIES_FAIL_HEIGHT=0: the device fold gets buffers of its smallest
    # capacity, which leaves of more than a few limbs outgrow; the kernel
    # reports them and the split folds every leaf on the host instead
    monkeypatch.setenv("HGX_SERIES_DEVICE", "1")
    monkeypatch.setenv("HGX_SERIES_FAIL_HEIGHT", "0")
    s = SERIES[name]()
    assert _tuple(S.binary_split_native(s, lo, up)) == SO.binary_split(s, lo, up)


@pytest.mark.parametrize("name,lo,up", [("chudnovsky", 0, 1), ("chudnovsky", 0, 32), ("apery", 5, 20),
                                        ("exp(7/2)", 3, 36), ("e", 0, 65)])
def test_single_and_few_leaves_folded_on_the_device(hgx, monkeypatch, name, lo, up):
    # one or a few leaves (at most 32 terms each) folded on the device, the
    # result read back from the leaves' own region or after a height or two
    monkeypatch.setenv("HGX_SERIES_DEVICE", "1")
    s = SERIES[name]()
    assert _tuple(S.binary_split_native(s, lo, up)) == SO.binary_split(s, lo, up)


@pytest.mark.parametrize("series,lo,up", [
    (S.chudnovsky_series(), 1 << 39, (1 << 39) + 300),  # q(k) ~ 2^170: 3-limb term values
    (S.Series(p=S.Poly((0, 0, 0, 0, 0, 0, 0, 1 << 20)), q=S.Poly((3, 1))), 1 << 33, (1 << 33) + 300),  # 4 limbs
    (S.Series(p=S.Poly((0, 0, 0, 0, 0, 0, 0, 1 << 40)), q=S.Poly((3, 1))), 1 << 39, (1 << 39) + 300),  # > 256 bits
])
def test_device_fold_of_wide_term_values(hgx, monkeypatch, series, lo, up):
    # term values of 3 and 4 limbs fold on the device; wider ones (the
    # last series) are folded on the host
    monkeypatch.setenv("HGX_SERIES_DEVICE", "1")
    assert _tuple(S.binary_split_native(series, lo, up)) == SO.binary_split(series, lo, up)


@pytest.mark.parametrize("mode", ["0", "1"])
def test_partial_read_into_device_memory(hgx, monkeypatch, mode):
    # hgx_partial_get_device: a split that ran in device memory (=1) hands
    # its partial over device to device, one that ran on the host (=0)
    # uploads it; both equal the oracle's
    from hypergeometrica_amd import mpd_device as MD
    monkeypatch.setenv("HGX_SERIES_DEVICE", mode)
    s = S.chudnovsky_series()
    got = MD.split_to_device(s, 0, 3000)
    want = SO.binary_split(s, 0, 3000)[2:]
    for k, w in zip("pqbr", want):
        assert MD.to_host(got[k]).mantissa == w, k
