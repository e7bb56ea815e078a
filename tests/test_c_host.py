"""The C ABI from a C program (tests/c/abi_host.c, built with gcc against
include/hgx.h and libhgx.so): scheme queries, the error code of a product
before hgx_init, and a host-only binary split of the Chudnovsky series
against the oracle's depth-first split (src/binary-splitting.lisp:100-133).
No device is touched."""
import os
import subprocess

import pytest

from hypergeometrica_amd import series as S
from oracle import series_oracle as SO

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def c_host(tmp_path_factory):
    lib = os.path.join(ROOT, "hypergeometrica_amd", "libhgx.so")
    if not os.path.exists(lib):
        pytest.fail("libhgx.so is not built: run __graft_entry__.build()")
    exe = str(tmp_path_factory.mktemp("c") / "abi_host")
    subprocess.run(["gcc", "-O1", "-std=c11", "-I", os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "c",
                    "abi_host.c"), "-L", os.path.dirname(lib), "-lhgx", "-Wl,-rpath," + os.path.dirname(lib),
                    "-o", exe], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout.splitlines()


def test_queries_and_uninitialised_product(c_host):
    lines = dict(line.split(" ", 1) for line in c_host if not line.startswith("series"))
    assert lines["moduli"].split() == [str(0x6280000000000001), str(0x5700000000000001), str(0x4180000000000001)]
    assert lines["length"] == str(1 << 17)
    assert int(lines["mul_uninit"]) == -5  # HGX_ENOTINIT


def test_series_split_from_c(c_host):
    got = {}
    for line in c_host:
        if line.startswith("series "):
            f = line.split()
            w, sign, n = int(f[1]), int(f[2]), int(f[3])
            mag = sum(int(x) << (64 * i) for i, x in enumerate(f[4:4 + n]))
            got[w] = sign * mag
    _, _, p, q, b, r = SO.binary_split(S.chudnovsky_series(), 0, 30)
    assert (got[0], got[1], got[2], got[3]) == (p, q, b, r)
