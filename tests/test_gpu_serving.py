"""GPU tests of the request server (hgx_submit / hgx_poll / hgx_wait).

The served products of a seeded Poisson stream over the serving
configuration's operand sizes (log-uniform in [2^10, 2^24] limbs) against
their solo runs and GMP; failures forced from the host (a copy the server
reports failed without issuing it); shutdown with requests queued; ticket
release; and host-staged requests beside small ones.  Integer work:
bit-exact."""
import ctypes
import os
import sys
import threading
import time

import numpy as np
import pytest

from tests import _gmp

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def _restart(env: dict):
    """Stop the library (its request server included) and start it again
    with ENV applied (None removes a variable): the server reads its
    settings when the next request starts it."""
    from hypergeometrica_amd import _lib
    L = _lib.lib()
    L.hgx_shutdown()
    for k, v in env.items():
        if v is None:
            os.environ.pop(k, None)
        else:
            os.environ[k] = str(v)
    _lib.check(L.hgx_init(0, 0))


@pytest.fixture
def server_env():
    """Restart the server with given settings; restore the defaults after."""
    used = {}

    def apply(**env):
        used.update(env)
        _restart(env)

    yield apply
    if used:
        _restart({k: None for k in used})


def _stream_operands(oracle, count, seed):
    from hypergeometrica_amd import workload as W
    na, nb = W.serve_shapes(count, seed)
    reqs = []
    for i in range(count):
        a = W.splitmix64(seed + 16 + 2 * i, na[i], force_top=True)
        b = a if i % 8 == 3 else W.splitmix64(seed + 17 + 2 * i, nb[i], force_top=True)
        reqs.append((a, b))
    return reqs


# the server's two stream layouts: uploads and downloads on streams shared
# by the compute lanes (the default), and every copy on its round's lane
MODES = {"lane_streams": {"HGX_SERVE_COPY_STREAMS": 0}, "copy_streams": {"HGX_SERVE_COPY_STREAMS": 1}}


@pytest.mark.parametrize("mode", list(MODES))
def test_poisson_stream_equals_solo_and_gmp(hgx, oracle, server_env, mode):
    # 24 requests of the serving mix, submitted at Poisson arrival times
    # (2000/s) so that they overlap on the device; every product equals its
    # solo run and GMP's
    from hypergeometrica_amd import _lib
    from hypergeometrica_amd import workload as W
    server_env(**MODES[mode])
    count, seed = 24, 0x48475804
    reqs = _stream_operands(oracle, count, seed)
    say("serving mix:", [(a.size, b.size) for a, b in reqs])
    solo = []
    for a, b in reqs:
        solo.append(hgx.submit(a, None if b is a else b).wait().copy())
    gaps = W.poisson_gaps(seed + 1, count, 2000.0)
    tickets = []
    t = time.perf_counter()
    for (a, b), g in zip(reqs, gaps):
        t += g
        while time.perf_counter() < t:
            pass
        tickets.append(hgx.submit(a, None if b is a else b))
    mixed = [tk.wait() for tk in tickets]
    for i, ((a, b), x, y) in enumerate(zip(reqs, solo, mixed)):
        assert np.array_equal(x, y), i
        want = _gmp.mul(a, b if b is not a else a.copy())
        assert np.array_equal(y, want), (i, a.size, b.size)
    st = _lib.server_stats()
    assert (st["uploads"], st["downloads"]) == ((1, 2) if mode == "copy_streams" else (0, 0))


@pytest.mark.parametrize("mode", list(MODES))
def test_failed_copy_fails_only_its_request(hgx, oracle, server_env, mode):
    # the second copy the server issues (the first request's second operand)
    # is reported failed: that request fails with the upload error, is not
    # downloaded, and the next one is served correctly
    from hypergeometrica_amd._lib import HGX_EDEVICE
    server_env(HGX_SERVE_FAIL_COPY=2, **MODES[mode])
    from hypergeometrica_amd import _lib
    L = _lib.lib()
    a = oracle.splitmix64(0x6001, 5000)
    b = oracle.splitmix64(0x6002, 7000)
    out = np.zeros(a.size + b.size, dtype=np.uint64)
    d = _lib.MulDesc(a.ctypes.data, a.size, b.ctypes.data, b.size, out.ctypes.data)
    t = ctypes.c_uint64(0)
    _lib.check(L.hgx_submit(ctypes.byref(d), ctypes.byref(t)))
    assert L.hgx_wait(t) == HGX_EDEVICE
    assert "upload" in L.hgx_last_error().decode()
    assert not out.any(), "a failed request's product was downloaded"
    assert np.array_equal(hgx.submit(a, b).wait(), _gmp.mul(a, b))
    assert np.array_equal(hgx.submit(b, a).wait(), _gmp.mul(b, a))
    assert _lib.server_stats()["failed"] == 1


def test_poll_releases_the_ticket(hgx, oracle):
    from hypergeometrica_amd import _lib
    L = _lib.lib()
    a = oracle.splitmix64(0x6003, 3000)
    out = np.zeros(6000, dtype=np.uint64)
    d = _lib.MulDesc(a.ctypes.data, a.size, a.ctypes.data, a.size, out.ctypes.data)
    t = ctypes.c_uint64(0)
    _lib.check(L.hgx_submit(ctypes.byref(d), ctypes.byref(t)))
    deadline = time.time() + 30
    while _lib.check(L.hgx_poll(t)) == 0:
        assert time.time() < deadline
    assert L.hgx_poll(t) == _lib.HGX_EINVAL
    assert L.hgx_wait(t) == _lib.HGX_EINVAL
    assert np.array_equal(out, _gmp.mul(a, a.copy()))


def test_shutdown_completes_every_ticket(hgx, oracle):
    # requests queued when the library shuts down complete: those already on
    # the device with their products, the rest with HGX_ENOTINIT; no wait
    # is left hanging
    from hypergeometrica_amd import _lib
    L = _lib.lib()
    rng = np.random.default_rng(0x6004)
    keep, tickets = [], []
    for i in range(200):
        na, nb = int(2 ** rng.uniform(8, 17)), int(2 ** rng.uniform(8, 17))
        a = oracle.splitmix64(0x6100 + 2 * i, na)
        b = oracle.splitmix64(0x6101 + 2 * i, nb)
        out = np.zeros(na + nb, dtype=np.uint64)
        d = _lib.MulDesc(a.ctypes.data, na, b.ctypes.data, nb, out.ctypes.data)
        t = ctypes.c_uint64(0)
        _lib.check(L.hgx_submit(ctypes.byref(d), ctypes.byref(t)))
        keep.append((a, b, out))
        tickets.append(t.value)
    codes = {}

    def waiter():
        for t in tickets:
            codes[t] = L.hgx_wait(ctypes.c_uint64(t))

    th = threading.Thread(target=waiter)
    th.start()
    L.hgx_shutdown()
    th.join(timeout=120)
    assert not th.is_alive(), "a wait did not return after hgx_shutdown"
    _lib.check(L.hgx_init(0, 0))
    done = [t for t in tickets if codes[t] == 0]
    say(f"shutdown: {len(done)} of {len(tickets)} requests completed")
    assert set(codes.values()) <= {0, _lib.HGX_ENOTINIT}
    for t, (a, b, out) in zip(tickets, keep):
        if codes[t] == 0:
            assert np.array_equal(out, _gmp.mul(a, b))
    a, b, _ = keep[0]
    assert np.array_equal(hgx.submit(a, b).wait(), _gmp.mul(a, b))  # served again after the restart


def test_staged_request_does_not_stall_small_ones(hgx, oracle):
    # a request beyond the staging limit is computed host-staged by the
    # server's stager thread; small requests submitted after it complete
    # while it runs
    from hypergeometrica_amd import _lib
    L = _lib.lib()
    _lib.check(L.hgx_staging_limit(48 << 20))
    try:
        n = 1 << 23
        A = oracle.splitmix64(0x6201, n)
        B = oracle.splitmix64(0x6202, n)
        big = hgx.submit(A, B)
        small = []
        for i in range(40):
            a = oracle.splitmix64(0x6300 + 2 * i, 2000 + 97 * i)
            b = oracle.splitmix64(0x6301 + 2 * i, 3000 + 31 * i)
            small.append((a, b, hgx.submit(a, b)))
        for a, b, t in small:
            assert np.array_equal(t.wait(), _gmp.mul(a, b))
        big_pending = not big.done()
        say("staged request still running after the small ones:", big_pending)
        assert big_pending, "the small requests waited for the staged one"
        assert np.array_equal(big.wait(), _gmp.mul(A, B))
        assert _lib.server_stats()["staged"] >= 1
    finally:
        _lib.check(L.hgx_staging_limit(0))
