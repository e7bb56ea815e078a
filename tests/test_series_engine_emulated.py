"""The native binary-splitting engine's host logic with its device side
emulated on the host (tests/c/series_emulated.cpp, built with g++ and the
address and undefined-behaviour sanitizers together with
hypergeometrica_amd/csrc/hgx_series.cpp): heights kept in "device" memory
(HGX_SERIES_DEVICE=1), heights on the host (=0), and a move back to the host
when a device height reports exhausted memory (HGX_SERIES_FAIL_HEIGHT).
Every partial equals the oracle's depth-first split
(src/binary-splitting.lisp:100-133).  No device is touched; the kernels are
tested on the GPU (tests/test_gpu_series.py)."""
import os
import subprocess

import pytest

from hypergeometrica_amd import series as S
from oracle import series_oracle as SO

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SERIES = {
    0: S.chudnovsky_series,
    1: lambda: S.Series(a=S.Poly((2, -7)), b=S.Poly((-3,)), p=S.Poly((1, 2, -1)), q=S.Poly((-4, 0, 5), at_zero=-1)),
    2: lambda: S.Series(a=S.Poly((1,)), p=S.Poly((-1,), at_zero=1)),
    3: lambda: S.Series(p=S.Poly((0, 0, 0, 0, 0, 0, 0, 1 << 40)), q=S.Poly((3, 1))),
    4: lambda: S.Series(p=S.Poly((0, 0, 0, 0, 0, 0, 0, 1 << 20)), q=S.Poly((3, 1))),
}
K39, K33 = 1 << 39, 1 << 33


@pytest.fixture(scope="module")
def engine(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("emu") / "series_emulated")
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                    "-fno-omit-frame-pointer", "-o", exe,
                    os.path.join(ROOT, "hypergeometrica_amd", "csrc", "hgx_series.cpp"),
                    os.path.join(ROOT, "tests", "c", "series_emulated.cpp"), "-lpthread"], check=True)
    return exe


def _run(exe, which, lo, up, env):
    e = dict(os.environ)
    e.pop("HGX_SERIES_DEVICE", None)
    e.pop("HGX_SERIES_FAIL_HEIGHT", None)
    e.update(env)
    r = subprocess.run([exe, str(which), str(lo), str(up)], capture_output=True, text=True, env=e, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr[-3000:]
    vals = []
    for line in r.stdout.split():
        sign, hx = line.split(":")
        vals.append(int(sign) * (int(hx, 16) if hx else 0))
    return tuple(vals), r.stderr


@pytest.mark.parametrize("env", [
    {"HGX_SERIES_DEVICE": "0"}, {"HGX_SERIES_DEVICE": "1"},
    {"HGX_SERIES_DEVICE": "1", "HGX_SERIES_FAIL_HEIGHT": "0"},
    {"HGX_SERIES_DEVICE": "1", "HGX_SERIES_FAIL_HEIGHT": "1"},
    {"HGX_SERIES_DEVICE": "1", "HGX_SERIES_FAIL_HEIGHT": "3"},
    {"HGX_SERIES_DEVICE": "1", "HGX_SERIES_FAIL_HEIGHT": "6"},
], ids=["host", "device", "fail-h0", "fail-h1", "fail-h3", "fail-h6"])
@pytest.mark.parametrize("which,lo,up", [(0, 0, 3000), (0, 1000, 2999), (1, 3, 5000), (2, 0, 1000), (2, 5, 4097),
                                         (3, K39, K39 + 300), (0, K39, K39 + 300), (4, K33, K33 + 300)])
def test_emulated_device_heights_equal_depth_first(engine, env, which, lo, up):
    """Leaves fold on the device whenever the split goes there and their
    term values fit 256 bits (series 3's do not: the host folds them; the
    Chudnovsky terms near k = 2^39 take 3 limbs, series 4's 4); with buffers
    made too small (fail-h0) a leaf fails and the host folds."""
    vals, err = _run(engine, which, lo, up, env)
    assert vals == SO.binary_split(SERIES[which](), lo, up)[2:]
    if env["HGX_SERIES_DEVICE"] == "0" or which == 3:
        assert "emulated device fold" not in err
    elif env.get("HGX_SERIES_FAIL_HEIGHT") == "0" and which in (0, 1, 4):
        assert "a leaf failed" in err
    else:
        assert "emulated device fold" in err and "done" in err
