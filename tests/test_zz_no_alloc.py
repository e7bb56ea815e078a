"""GPU test of hgx_mul_device_noalloc, the entry point that allocates
nothing: through the C ABI, against hgx_mul_device on the same device
operands, bit for bit, on two calls in a row (the second with a smaller
input), plus its size limit and the resident-tables requirement."""
import ctypes
import sys

import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def test_noalloc_equals_mul_device(hgx):
    import torch

    from hypergeometrica_amd import _lib
    L = _lib.lib()
    dev = torch.device("cuda:0")
    # the largest product it takes (L = 2^20), then smaller ones: a 2-pass
    # product, a single-tile product and a square
    shapes = [((1 << 19) + 4321, (1 << 19) - 4321), (40000, 25000), (1500, 900), (70000, None)]
    ops, refs = [], []
    for i, (na, nb) in enumerate(shapes):
        A = torch.empty(na, dtype=torch.int64, device=dev)
        hgx.fill_splitmix64(A.data_ptr(), na, 0x4E41 + 2 * i)
        if nb is None:
            B, nb = A, na
        else:
            B = torch.empty(nb, dtype=torch.int64, device=dev)
            hgx.fill_splitmix64(B.data_ptr(), nb, 0x4E42 + 2 * i)
        R = torch.zeros(na + nb, dtype=torch.int64, device=dev)
        # the existing entry point; it also builds the root tables of the length
        _lib.check(L.hgx_mul_device(A.data_ptr(), na, B.data_ptr(), nb, R.data_ptr(), None))
        ops.append((A, na, B, nb))
        refs.append(R)
    torch.cuda.synchronize()
    outs = [torch.full_like(R, -1) for R in refs]
    say("noalloc: first call (2^20), second smaller, then the rest")
    free0 = torch.cuda.mem_get_info()[0]
    for (A, na, B, nb), O in zip(ops, outs):
        _lib.check(L.hgx_mul_device_noalloc(A.data_ptr(), na, B.data_ptr(), nb, O.data_ptr(), None))
    torch.cuda.synchronize()
    assert torch.cuda.mem_get_info()[0] == free0, "device memory changed across the calls"
    for i, (O, R) in enumerate(zip(outs, refs)):
        assert torch.equal(O, R), shapes[i]
    # again in reverse order (small after large, large after small): same
    outs2 = [torch.full_like(R, -1) for R in refs]
    for (A, na, B, nb), O in reversed(list(zip(ops, outs2))):
        _lib.check(L.hgx_mul_device_noalloc(A.data_ptr(), na, B.data_ptr(), nb, O.data_ptr(), None))
    torch.cuda.synchronize()
    for O, R in zip(outs2, refs):
        assert torch.equal(O, R)


def test_noalloc_limits(hgx):
    import torch

    from hypergeometrica_amd import _lib
    L = _lib.lib()
    dev = torch.device("cuda:0")
    A = torch.empty((1 << 19) + 1, dtype=torch.int64, device=dev)
    O = torch.empty(1 << 21, dtype=torch.int64, device=dev)
    hgx.fill_splitmix64(A.data_ptr(), A.numel(), 0x4E50)
    # above the scratch array's transform length
    assert L.hgx_mul_device_noalloc(A.data_ptr(), A.numel(), A.data_ptr(), (1 << 19), O.data_ptr(), None) \
        == _lib.HGX_ETOOBIG
    assert b"2^20" in L.hgx_last_error()
    assert L.hgx_mul_device_noalloc(None, 4, A.data_ptr(), 4, O.data_ptr(), None) == _lib.HGX_EINVAL
    # root tables of a length not resident: a fresh pass structure for 2^16
    c, dm, md = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    _lib.check(L.hgx_plan_get(16, ctypes.byref(c), ctypes.byref(dm), ctypes.byref(md)))
    try:
        _lib.check(L.hgx_plan_set(16, c.value, 12 if dm.value != 12 else 16, md.value))
        n = 1 << 15
        assert L.hgx_mul_device_noalloc(A.data_ptr(), n, A.data_ptr() + 8, n, O.data_ptr(), None) \
            == _lib.HGX_ENOTINIT
        R = torch.empty(2 * n, dtype=torch.int64, device=dev)
        _lib.check(L.hgx_mul_device(A.data_ptr(), n, A.data_ptr() + 8, n, R.data_ptr(), None))
        _lib.check(L.hgx_mul_device_noalloc(A.data_ptr(), n, A.data_ptr() + 8, n, O.data_ptr(), None))
        torch.cuda.synchronize()
        assert torch.equal(O[:2 * n], R)
    finally:
        _lib.check(L.hgx_plan_set(16, c.value, dm.value, md.value))
