"""GPU tests of the device limb-vector entry points (hgx_device_alloc /
free / upload / download / copy / fill / significant / compare,
hgx_shl_device, hgx_host_alloc / free) through ctypes alone, and of the
Newton chain that keeps its iterates in HBM with them (mpd_device, which
uses no other device API): results equal exact integers and the host
iteration."""
import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def _alloc(L, n):
    from hypergeometrica_amd import _lib
    p = ctypes.c_void_p(0)
    _lib.check(L.hgx_device_alloc(n, ctypes.byref(p)))
    return int(p.value)


def _val(a):
    return int.from_bytes(np.ascontiguousarray(a, dtype=np.uint64).astype("<u8").tobytes(), "little")


def test_device_primitives(hgx):
    from hypergeometrica_amd import _lib
    L = _lib.lib()
    rng = np.random.default_rng(0x44)
    for n in (1, 3, 1000, 4099, 300001, 1 << 22):
        a = rng.integers(0, 2**64, size=n, dtype=np.uint64)
        a[-1] |= np.uint64(1)
        d = _alloc(L, n)
        _lib.check(L.hgx_upload(d, a.ctypes.data, n, None))
        back = np.zeros(n, dtype=np.uint64)
        _lib.check(L.hgx_download(back.ctypes.data, d, n, None))
        assert np.array_equal(back, a), n
        # significant limbs: the top non-zero one, after clearing the top part
        out = ctypes.c_size_t(0)
        _lib.check(L.hgx_device_significant(d, n, ctypes.byref(out), None))
        assert out.value == n
        k = n // 2
        _lib.check(L.hgx_device_fill(d + 8 * k, n - k, 0, None))
        _lib.check(L.hgx_device_significant(d, n, ctypes.byref(out), None))
        want = k
        while want and a[want - 1] == 0:
            want -= 1
        assert out.value == want, (n, k)
        # compare: equal, longer, one limb higher, one limb lower
        e = _alloc(L, n)
        _lib.check(L.hgx_upload(e, a.ctypes.data, n, None))
        _lib.check(L.hgx_upload(d, a.ctypes.data, n, None))
        c = ctypes.c_int(9)
        _lib.check(L.hgx_device_compare(d, n, e, n, ctypes.byref(c), None))
        assert c.value == 0
        j = int(rng.integers(0, n))
        bumped = a.copy()
        bumped[j] = bumped[j] + np.uint64(1) if bumped[j] != np.uint64(2**64 - 1) else np.uint64(0)
        _lib.check(L.hgx_upload(e, bumped.ctypes.data, n, None))
        _lib.check(L.hgx_device_compare(d, n, e, n, ctypes.byref(c), None))
        assert c.value == (1 if _val(a) > _val(bumped) else -1)
        _lib.check(L.hgx_device_compare(d, n, e, max(1, n // 3), ctypes.byref(c), None))
        assert c.value == (1 if _val(a) > _val(bumped[:max(1, n // 3)]) else (0 if n == 1 and _val(a) == _val(bumped) else -1))
        # shifts: a * 2^s in n + 1 limbs, and copies
        o = _alloc(L, n + 1)
        for s in (0, 1, 63):
            _lib.check(L.hgx_shl_device(d, n, s, o, None))
            got = np.zeros(n + 1, dtype=np.uint64)
            _lib.check(L.hgx_download(got.ctypes.data, o, n + 1, None))
            assert _val(got) == _val(a) << s, (n, s)
        _lib.check(L.hgx_device_copy(o, d, n, None))
        got = np.zeros(n, dtype=np.uint64)
        _lib.check(L.hgx_download(got.ctypes.data, o, n, None))
        assert np.array_equal(got, a)
        for p in (d, e, o):
            _lib.check(L.hgx_device_free(p, None))
    # rejections
    d = _alloc(L, 8)
    assert L.hgx_shl_device(d, 4, 64, d + 64, None) == _lib.HGX_EINVAL
    assert L.hgx_shl_device(d, 4, 3, d + 8, None) == _lib.HGX_EINVAL      # overlaps its input
    assert L.hgx_device_copy(d + 8, d, 4, None) == _lib.HGX_EINVAL        # overlapping copy
    assert L.hgx_device_free(d + 8, None) == _lib.HGX_EINVAL              # not a block
    _lib.check(L.hgx_device_free(d, None))
    # blocks are reused after release
    ptrs = set()
    for _ in range(20):
        p = _alloc(L, 12345)
        ptrs.add(p)
        _lib.check(L.hgx_device_free(p, None))
        _lib.check(L.hgx_synchronize(None))
    assert len(ptrs) < 20
    # pinned host memory: a product read from and written into it
    h = ctypes.c_void_p(0)
    _lib.check(L.hgx_host_alloc(8 * 4000, ctypes.byref(h)))
    buf = np.ctypeslib.as_array((ctypes.c_uint64 * 4000).from_address(h.value))
    buf[:2000] = rng.integers(0, 2**64, size=2000, dtype=np.uint64)
    _lib.check(L.hgx_mul(h.value, 1000, h.value + 8 * 1000, 1000, h.value + 8 * 2000))
    assert _val(buf[2000:4000]) == _val(buf[:1000]) * _val(buf[1000:2000])
    del buf
    _lib.check(L.hgx_host_free(h.value))
