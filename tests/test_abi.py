"""CPU checks of the C ABI: the library loads and exports every symbol
include/hgx.h declares (no device calls)."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "hgx.h")).read()
    src = re.sub(r"//.*", "", src)
    return sorted(set(re.findall(r"\b(hgx_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_the_boundary():
    syms = declared_symbols()
    for s in ("hgx_init", "hgx_mul", "hgx_sqr", "hgx_mul_batch", "hgx_ntt_forward", "hgx_ntt_reverse",
              "hgx_submit", "hgx_poll", "hgx_wait", "hgx_last_error"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    import ctypes
    from hypergeometrica_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libhgx.so is not built: run __graft_entry__.build()")
    L = ctypes.CDLL(_lib.LIB_PATH)
    for s in declared_symbols():
        assert hasattr(L, s), s
    assert sorted(_lib.EXPORTS) == declared_symbols()


def test_host_only_queries():
    from hypergeometrica_amd import _lib
    L = _lib.load()
    assert L.hgx_num_moduli() == 3
    assert [L.hgx_modulus(k) for k in range(3)] == [0x6280000000000001, 0x5700000000000001, 0x4180000000000001]
    assert L.hgx_product_length(2**16, 2**16) == 2**17
    assert L.hgx_product_length(3, 2) == 8
    assert L.hgx_product_length(1, 1) == 2


def test_calls_before_init_fail_loudly():
    # no GPU in this process: hgx_mul must fail, never compute on the CPU
    import numpy as np
    from hypergeometrica_amd import _lib
    L = _lib.load()
    if _lib._initialised:
        pytest.skip("device already initialised in this process")
    a = np.ones(4, dtype=np.uint64)
    out = np.zeros(8, dtype=np.uint64)
    rc = L.hgx_mul(a.ctypes.data, 4, a.ctypes.data, 4, out.ctypes.data)
    assert rc == _lib.HGX_ENOTINIT
    assert not out.any()
