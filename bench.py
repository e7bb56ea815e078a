"""Benchmark of the exact NTT multiply on one MI355X.

Workload (one "step"): one exact product of two 2^26-limb integers
(uniform random 64-bit limbs, SplitMix64 seeds 0x48475801/0x48475802,
top limb non-zero), transform length L = 2^27 over the three 63-bit primes,
with operands and product resident in HBM.  value = output limbs (2^27) per
second over the timed steps.

Also reported:
  roofline      the dominant kernel's algorithmic HBM bytes per launch over
                its mean duration (HIP events on its launch stream, recorded
                by the library around every kernel during the timed steps),
                against 8 TB/s;
  cpu_baseline  the oracle (C restatement of the reference's NTT multiply,
                1 thread) timed on this host on a bounded sample;
  latency       p50 / p99 of per-step wall time;
  legs          the 2^bits-limb square, 4096 batched products against the
                per-call path, 2^16-limb per-call latency, the request server
                under Poisson load, host-staged products and mpd-pi.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-<leg> ...]
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# hardware queues of this process's HIP runtime (read when it starts, so
# before the library loads): the request server runs six streams beside the
# library's, and with HIP's default of four queues streams share queues and
# wait behind each other's copies (served rate 3.8 -> 4.0e9 limbs/s with 8)
os.environ.setdefault("GPU_MAX_HW_QUEUES", "8")
# the request server admits requests under a fixed HBM budget: an arena of
# 64 GiB for operands, products and work arrays (read when the server starts)
SERVE_ARENA_BYTES = 64 << 30
os.environ.setdefault("HGX_SERVE_ARENA_BYTES", str(SERVE_ARENA_BYTES))

from hypergeometrica_amd import workload as W  # noqa: E402

HBM_PEAK = 8.0e12
PROBE_OPS = {"shoup": 0, "mont": 1, "addsub": 2, "fwd_butterfly": 3, "inv_butterfly": 4, "lds_exchange": 5,
             "wave_exchange": 6}
METRIC = "NTT-multiply limbs/s batched vs per-call on GPU; %HBM roofline; p99 latency"


def kernel_bytes(name: str, L: int, na: int, nb: int, nops: int) -> float | None:
    """Algorithmic HBM bytes of one launch: every word the kernel must read
    or write once (3 primes, 8-byte words)."""
    K = 3
    runs = L // 16
    if name == "fwd_col_first":
        return 8 * (na + nb if nops == 2 else na) + 8 * K * L * nops
    if name == "fwd_col":
        return 16 * K * L * nops
    if name in ("row_mul",):
        return 8 * K * L * 3          # read both spectra, write the product spectrum
    if name in ("row_sqr",):
        return 8 * K * L * 2
    if name == "inv_col":
        return 16 * K * L
    if name == "crt":
        return 8 * K * L + 8 * (na + nb) + 20 * runs
    return None


def alu_probe(_lib, iters: int = 4096) -> dict:
    """Rates of the library's own modular arithmetic on this GPU
    (hgx_modmul_probe: register-only, every CU busy) and the shader clock
    under that load: the integer-ALU ceiling of the transform kernels."""
    import ctypes
    out = (ctypes.c_double * 3)()
    res = {}
    for name, op in PROBE_OPS.items():
        _lib.check(_lib.lib().hgx_modmul_probe(op, iters, out))
        res[name] = {"ops_per_s": out[0], "clock_ghz": round(out[1], 3)}
    return res


def kernel_alu_ops(name: str, L: int, srow: int) -> dict | None:
    """Modular operations of one launch of a transform kernel of the
    product of two operands at transform length L (3 primes; column passes
    of 8 stages, a row pass of SROW stages; the outermost column pass
    applies tile-factored twiddles, two Shoup products per word, the inner
    one direct twiddles, one): radix-2 butterflies, twiddle products and
    the pointwise Montgomery products."""
    K, half = 3, L // 2
    if name == "fwd_col_first":
        return {"fwd_butterfly": K * 2 * 8 * half, "shoup": K * 2 * L * 2}
    if name == "fwd_col":
        return {"fwd_butterfly": K * 2 * 8 * half, "shoup": K * 2 * L}
    if name == "row_mul":
        return {"fwd_butterfly": K * 2 * srow * half, "inv_butterfly": K * srow * half, "mont": K * L}
    if name == "inv_col":  # the inner (direct) and outermost (factored) passes, averaged
        return {"inv_butterfly": K * 8 * half, "shoup": K * L * 3 // 2}
    return None


def alu_floor_s(ops: dict, probe: dict) -> float:
    """Seconds the operations take at the probe's rates."""
    return sum(n / probe[op]["ops_per_s"] for op, n in ops.items())


def cpu_baseline(sample_bits: int = 22, headline_bits: int = 26) -> dict:
    """Time the oracle's mpz-*/ntt restatement on one 2^sample_bits x
    2^sample_bits multiply (L = 2^(sample_bits+1), about 11 s): single-threaded
    (the reference's default build, the reported value), and with the
    reference's parallel tasks (one thread per operand and modulus, at most
    6); GMP's mpz_mul on the same operands is a sanity line, not the
    baseline.  All three products are compared.  The reported value is the
    rate at the headline shape (L = 2^(headline_bits+1)) at the sample's time
    per butterfly: the radix-2 transforms dominate and cost L log2 L, so the
    sample's rate is scaled by log2 L_sample / log2 L_headline (CRT and
    pointwise work, linear in L, are charged the same factor: the CPU rate is
    if anything understated)."""
    import numpy as np

    from oracle import oracle as O
    O.build()
    n = 1 << sample_bits
    a = O.splitmix64(0x48475801, n)
    b = O.splitmix64(0x48475802, n)
    t0 = time.perf_counter()
    p1 = O.mpz_mul_ntt(a, b)
    dt = time.perf_counter() - t0
    t0 = time.perf_counter()
    p6 = O.mpz_mul_ntt_parallel(a, b, 6)
    dt6 = time.perf_counter() - t0
    gmp = None
    try:
        from tests import _gmp
        if _gmp.available():
            t0 = time.perf_counter()
            pg = _gmp.mul(a, b)
            dtg = time.perf_counter() - t0
            gmp = {"value": (2 * n) / dtg, "unit": "limbs/s", "cores": 1, "seconds": round(dtg, 3),
                   "same_product": bool(np.array_equal(pg, p1[: 2 * n]))}
    except (ImportError, OSError):
        gmp = None
    host = f"host {platform.processor() or platform.machine()}, {os.cpu_count()} cpus"
    scale = (sample_bits + 1) / (headline_bits + 1)
    return {
        "value": (2 * n) / dt * scale,
        "unit": "limbs/s",
        "cores": 1,
        "kind": "port",
        "sample": f"one {n}x{n}-limb multiply (L=2^{sample_bits + 1}) through the oracle's restatement of "
                  f"mpz-*/ntt, 1 thread, {dt:.2f} s ({(2 * n) / dt:.4g} limbs/s), scaled to L=2^{headline_bits + 1} "
                  f"by {sample_bits + 1}/{headline_bits + 1} (time per radix-2 butterfly held); {host}",
        "sample_value": (2 * n) / dt,
        "parallel_tasks": {"value": (2 * n) / dt6 * scale, "unit": "limbs/s", "cores": 6, "seconds": round(dt6, 3),
                           "same_product": bool(np.array_equal(p6, p1)),
                           "sample": "same multiply, one thread per (operand, prime) forward task and per prime "
                                     "pointwise and inverse task (with-parallel-work), CRT serial; scaled alike"},
        "gmp_sanity": gmp,
    }


class Dev:
    """LIMBS words of device memory from the library (hgx_device_alloc)."""

    def __init__(self, _lib, limbs: int):
        import ctypes
        p = ctypes.c_void_p(0)
        _lib.check(_lib.lib().hgx_device_alloc(limbs, ctypes.byref(p)))
        self._lib, self.ptr, self.n = _lib, int(p.value), limbs

    def free(self):
        if self.ptr:
            self._lib.check(self._lib.lib().hgx_device_free(self.ptr, None))
            self.ptr = 0


class Pinned:
    """LIMBS words of pinned host memory from the library (hgx_host_alloc),
    with a numpy view."""

    def __init__(self, _lib, limbs: int):
        import ctypes

        import numpy as np
        p = ctypes.c_void_p(0)
        _lib.check(_lib.lib().hgx_host_alloc(8 * limbs, ctypes.byref(p)))
        self._lib, self.ptr, self.n = _lib, int(p.value), limbs
        self.arr = np.ctypeslib.as_array((ctypes.c_uint64 * limbs).from_address(self.ptr))

    def free(self):
        if self.ptr:
            self.arr = None
            self._lib.check(self._lib.lib().hgx_host_free(self.ptr))
            self.ptr = 0


def sync(_lib) -> None:
    """Wait for the work queued on the library's stream."""
    _lib.check(_lib.lib().hgx_synchronize(None))


def dev_equal(_lib, a: int, b: int, n: int) -> bool:
    """Whether N limbs at device addresses A and B are equal."""
    import ctypes
    c = ctypes.c_int(9)
    _lib.check(_lib.lib().hgx_device_compare(a, n, b, n, ctypes.byref(c), None))
    return c.value == 0


def download(_lib, ptr: int, n: int):
    import numpy as np
    out = np.empty(n, dtype=np.uint64)
    _lib.check(_lib.lib().hgx_download(out.ctypes.data, ptr, n, None))
    return out


def batched_leg(h, _lib, count: int = 4096, reps: int = 3, probe: dict | None = None) -> dict:
    """Binary-splitting level: COUNT independent products, operand sizes
    log-uniform in [2^8, 2^16] limbs (seed 0x48475803), operands in HBM.
    Batched (hgx_mul_batch_device, grouped by transform length) against the
    per-call path (hgx_mul_device per product); outputs compared.  With the
    ALU PROBE's rates, the time of the level's modular arithmetic (3 primes:
    radix-2 butterflies, pointwise products, and for transforms above one
    4096-word tile the twiddle products of the one column pass each such
    length has up to 2^20, one per word of each forward and of the inverse
    transform) and the batched time's fraction of it."""
    na, nb = W.batch_shapes(count, W.BATCH_SEED)
    tot = sum(na) + sum(nb)
    IN, OUT1, OUT2 = Dev(_lib, tot), Dev(_lib, tot), Dev(_lib, tot)
    descs1 = (_lib.MulDesc * count)()
    descs2 = (_lib.MulDesc * count)()
    at = 0
    for i in range(count):
        a = IN.ptr + 8 * at
        h.fill_splitmix64(a, na[i], W.BATCH_SEED + 2 * i + 1)
        b = a + 8 * na[i]
        h.fill_splitmix64(b, nb[i], W.BATCH_SEED + 2 * i + 2)
        descs1[i] = _lib.MulDesc(a, na[i], b, nb[i], OUT1.ptr + 8 * at)
        descs2[i] = _lib.MulDesc(a, na[i], b, nb[i], OUT2.ptr + 8 * at)
        at += na[i] + nb[i]
    L_ = _lib.lib()
    # warm up both paths (plans, work memory)
    _lib.check(L_.hgx_mul_batch_device(descs1, count, None))
    for i in range(count):
        _lib.check(L_.hgx_mul_device(descs2[i].a, descs2[i].na, descs2[i].b, descs2[i].nb, descs2[i].out, None))
    sync(_lib)
    tb = []
    for _ in range(reps):
        t0 = time.perf_counter()
        _lib.check(L_.hgx_mul_batch_device(descs1, count, None))
        sync(_lib)
        tb.append(time.perf_counter() - t0)
    tp = []
    for _ in range(reps):
        t0 = time.perf_counter()
        for i in range(count):
            _lib.check(L_.hgx_mul_device(descs2[i].a, descs2[i].na, descs2[i].b, descs2[i].nb, descs2[i].out, None))
        sync(_lib)
        tp.append(time.perf_counter() - t0)
    same = dev_equal(_lib, OUT1.ptr, OUT2.ptr, tot)
    for d in (IN, OUT1, OUT2):
        d.free()
    b, p = min(tb), min(tp)
    res = {"workload": f"{count} products, operand limbs log-uniform in [2^8, 2^16], seed 0x48475803",
           "limbs": tot, "batched_s": round(b, 5), "per_call_s": round(p, 5),
           "batched_limbs_per_s": tot / b, "per_call_limbs_per_s": tot / p,
           "speedup": round(p / b, 2), "outputs_identical": same}
    if probe:
        ops = {"fwd_butterfly": 0, "inv_butterfly": 0, "mont": 0, "shoup": 0}
        for x, y in zip(na, nb):
            n = (x + y - 1).bit_length()
            L = 1 << n
            ops["fwd_butterfly"] += 3 * 2 * (L // 2) * n
            ops["inv_butterfly"] += 3 * (L // 2) * n
            ops["mont"] += 3 * L
            if n > 12:
                ops["shoup"] += 3 * 3 * L
        fl = alu_floor_s(ops, probe)
        res["alu_floor_ms"] = round(1e3 * fl, 3)
        res["alu_frac"] = round(fl / b, 4)
    return res


def square_leg(h, _lib, bits: int = 26, steps: int = 10) -> dict:
    """mpz-square (src/ntt-multiply.lisp:101-172) of one 2^BITS-limb operand
    (seed 0x48475801, L = 2^(BITS+1)), operand and square in HBM: STEPS timed
    calls after two warm-up calls, output limbs (2n) per second and the
    per-kernel HIP-event durations of the square's launches.  The result is
    checked against the general product of the operand with a copy of itself."""
    L_ = _lib.lib()
    n = 1 << bits
    A, A2, O, S = Dev(_lib, n), Dev(_lib, n), Dev(_lib, 2 * n), Dev(_lib, 2 * n)
    h.fill_splitmix64(A.ptr, n, 0x48475801)
    _lib.check(L_.hgx_device_copy(A2.ptr, A.ptr, n, None))

    def sq():
        _lib.check(L_.hgx_sqr_device(A.ptr, n, O.ptr, None))

    for _ in range(2):
        sq()
    sync(_lib)
    t0 = time.perf_counter()
    for _ in range(steps):
        sq()
    sync(_lib)
    dt = (time.perf_counter() - t0) / steps
    _lib.timing_enable(True)
    for _ in range(steps):
        sq()
    kt = _lib.timing_read()
    _lib.timing_enable(False)
    _lib.check(L_.hgx_mul_device(A.ptr, n, A2.ptr, n, S.ptr, None))
    same = dev_equal(_lib, S.ptr, O.ptr, 2 * n)
    for d in (A, A2, O, S):
        d.free()
    return {"workload": f"square of one 2^{bits}-limb operand, L=2^{bits + 1}, operand in HBM",
            "ms": round(1e3 * dt, 3), "limbs_per_s": 2 * n / dt, "equals_general_product": same,
            "kernels_ms": {k: round(t / c, 4) for k, (c, t) in sorted(kt.items())}}


def latency_leg(h, _lib, n: int = 1 << 16, reps: int = 200) -> dict:
    """One 2^16 x 2^16-limb product (L = 2^17, seeds 0x48475801/2), REPS
    calls: host-to-host through hgx_mul (pinned host buffers: limbs in, the
    product out) and device-resident through hgx_mul_device.  Per-call wall
    time p50 / p99 and limbs/s."""
    import numpy as np

    L_ = _lib.lib()
    A, B, O = Dev(_lib, n), Dev(_lib, n), Dev(_lib, 2 * n)
    h.fill_splitmix64(A.ptr, n, 0x48475801)
    h.fill_splitmix64(B.ptr, n, 0x48475802)
    hA, hB, hO = Pinned(_lib, n), Pinned(_lib, n), Pinned(_lib, 2 * n)
    _lib.check(L_.hgx_download(hA.ptr, A.ptr, n, None))
    _lib.check(L_.hgx_download(hB.ptr, B.ptr, n, None))

    def timed(fn):
        for _ in range(10):
            fn()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        ts.sort()
        return ts

    def host_call():
        _lib.check(L_.hgx_mul(hA.ptr, n, hB.ptr, n, hO.ptr))

    def dev_call():
        _lib.check(L_.hgx_mul_device(A.ptr, n, B.ptr, n, O.ptr, None))
        sync(_lib)

    th = timed(host_call)
    td = timed(dev_call)
    same = bool(np.array_equal(hO.arr, download(_lib, O.ptr, 2 * n)))
    want = hO.arr.copy()

    def threaded(nthreads: int, calls: int = 100) -> dict:
        """NTHREADS threads each making CALLS hgx_mul calls of this product
        from and to pageable memory of their own (numpy arrays), as the
        reference's parallel build calls mpz-* from lparallel workers."""
        import threading
        bufs = [(hA.arr.copy(), hB.arr.copy(), np.empty(2 * n, dtype=np.uint64)) for _ in range(nthreads)]
        go = threading.Barrier(nthreads + 1)
        errs = []

        def work(k):
            a, b, o = bufs[k]
            go.wait()
            try:
                for _ in range(calls):
                    _lib.check(L_.hgx_mul(a.ctypes.data, n, b.ctypes.data, n, o.ctypes.data))
            except Exception as e:  # reported with the leg
                errs.append(repr(e))

        ths = [threading.Thread(target=work, args=(k,)) for k in range(nthreads)]
        for t in ths:
            t.start()
        go.wait()
        t0 = time.perf_counter()
        for t in ths:
            t.join()
        dt = time.perf_counter() - t0
        ok = not errs and all(np.array_equal(o, want) for _, _, o in bufs)
        return {"threads": nthreads, "calls": nthreads * calls, "seconds": round(dt, 4),
                "limbs_per_s": 2 * n * nthreads * calls / dt, "outputs_identical": ok}

    conc = [threaded(k) for k in (1, 2, 4, 8)]
    for d in (A, B, O, hA, hB, hO):
        d.free()

    def q(ts, f):
        return round(1e3 * ts[min(len(ts) - 1, int(f * len(ts)))], 4)

    return {"workload": f"one {n} x {n}-limb product (L=2^{(2 * n).bit_length() - 1}), {reps} calls",
            "host_p50_ms": q(th, 0.5), "host_p99_ms": q(th, 0.99), "host_limbs_per_s": 2 * n / th[len(th) // 2],
            "device_p50_ms": q(td, 0.5), "device_p99_ms": q(td, 0.99), "device_limbs_per_s": 2 * n / td[len(td) // 2],
            "outputs_identical": same, "concurrent_host_calls": conc}


def pi_leg(bits: int = 1 << 28, identical_bits: int = 1 << 22, mpfr_bits: int = 1 << 24, host_iterates: bool = False,
           threads: int = 16) -> dict:
    """mpd-pi (Chudnovsky binary splitting, Newton reciprocal and inverse
    square root, src/pi.lisp:53-74) to BITS bits with the reference's
    mpd-truncate! applied after each Newton step (src/mpd.lisp:56-70; a
    declared deviation: untruncated iterates grow without bound), every
    product above the 100-limb threshold on the GPU, binary-splitting levels
    batched and the Newton iterates in HBM (mpd_device); optionally also with
    host-resident iterates.  Checked three ways: hexadecimal digits of pi at
    positions near the end, the middle and the start by BBP digit extraction;
    the leading MPFR_BITS bits against MPFR's pi; and, at IDENTICAL_BITS, the
    GPU routings (host and HBM iterates) bit-identical to GMP-routed host
    products."""
    from hypergeometrica_amd import mpd as M
    from hypergeometrica_amd import mpd_device as D
    from tests import _mpfr
    from tests import _pi_check as PC
    res = {"workload": f"mpd-pi to 2^{bits.bit_length() - 1} bits, truncating Newton, GPU products "
                       f"(binary-splitting levels batched, Newton iterates in HBM)"}
    t0 = time.perf_counter()
    ctx = D.DeviceContext(truncate=True)
    r = ctx.mpd_pi(bits)
    res["gpu_newton_in_hbm_s"] = round(time.perf_counter() - t0, 3)
    res["phases_s"] = ctx.phase_s
    if host_iterates:
        t0 = time.perf_counter()
        g = M.Context(truncate=True)
        rh = g.mpd_pi(bits)
        res["gpu_host_iterates_s"] = round(time.perf_counter() - t0, 3)
        res["host_and_hbm_iterates_identical"] = rh == r
        res["products"], res["max_product_limbs"] = g.products, g.max_product_limbs
        del rh
    pos = PC.bbp_positions(bits)
    t0 = time.perf_counter()
    dist = PC.bbp_check(r.mantissa, r.exponent, pos, threads=threads)
    res["bbp"] = {"hex_positions": pos, "distance_2^-64": [dist[d] for d in pos],
                  "bound_2^-64": [8 * (d + 17) for d in pos], "match": PC.bbp_ok(dist),
                  "seconds": round(time.perf_counter() - t0, 2)}
    if _mpfr.available():
        t0 = time.perf_counter()
        res["mpfr"] = {"leading_bits": mpfr_bits,
                       "match": PC.mpfr_leading_bits_match(r.mantissa, r.exponent, mpfr_bits),
                       "seconds": round(time.perf_counter() - t0, 2)}
    del r
    t0 = time.perf_counter()
    g = M.Context(truncate=True)
    a = g.mpd_pi(identical_bits)
    b = D.DeviceContext(truncate=True).mpd_pi(identical_bits)
    c = M.Context(mul=PC.gmp_mul, truncate=True).mpd_pi(identical_bits)
    res["routing"] = {"bits": identical_bits, "gpu_host_iterates_eq_gpu_hbm_eq_gmp": a == b == c,
                      "products": g.products, "seconds": round(time.perf_counter() - t0, 2)}
    return res


def pcie_probe(_lib, mib: int = 1024, reps: int = 3) -> dict:
    """Host<->device copy rates of pinned memory on this box: each direction
    alone and both at once (two threads, one stream each), best of REPS."""
    import ctypes
    import threading
    L_ = _lib.lib()
    n = mib << 17
    H1, H2, D1, D2 = Pinned(_lib, n), Pinned(_lib, n), Dev(_lib, n), Dev(_lib, n)
    s1, s2 = ctypes.c_void_p(0), ctypes.c_void_p(0)
    _lib.check(L_.hgx_stream_create(ctypes.byref(s1)))
    _lib.check(L_.hgx_stream_create(ctypes.byref(s2)))

    def up():
        _lib.check(L_.hgx_upload(D1.ptr, H1.ptr, n, s1))

    def down():
        _lib.check(L_.hgx_download(H2.ptr, D2.ptr, n, s2))

    def both():
        th = [threading.Thread(target=f) for f in (up, down)]
        for t in th:
            t.start()
        for t in th:
            t.join()

    def timed(fn):
        best = float("inf")
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            best = min(best, time.perf_counter() - t0)
        return 8 * n / best / 1e9

    r_up, r_down, r_bi = timed(up), timed(down), timed(both)
    _lib.check(L_.hgx_stream_destroy(s1))
    _lib.check(L_.hgx_stream_destroy(s2))
    for d in (H1, H2, D1, D2):
        d.free()
    # a served request moves 8 bytes per output limb up (its operands) and 8
    # down (its product), both directions at once
    return {"h2d_GBps": round(r_up, 1), "d2h_GBps": round(r_down, 1), "both_ways_GBps_each": round(r_bi, 1),
            "bound_limbs_per_s": r_bi * 1e9 / 8}


def copy_replay(_lib, count: int = 3000, seed: int = W.SERVE_SEED, lanes: int = 6) -> dict:
    """The copies of the serving stream without its kernels: the first COUNT
    requests' operand uploads and product downloads (sizes of serve_shapes
    SEED, pinned host memory), LANES threads each on a stream of its own
    taking the next request in turn (upload a, upload b, download the
    product), all requests available at once.  What the copy engines give
    for this size mix at this concurrency: the served rate's ceiling."""
    import ctypes
    import threading
    L_ = _lib.lib()
    na_l, nb_l = W.serve_shapes(count, seed)
    nmax = 1 << 24
    HA, HB = Pinned(_lib, nmax), Pinned(_lib, nmax)
    streams, bufs = [], []
    try:
        for _ in range(lanes):
            s = ctypes.c_void_p(0)
            _lib.check(L_.hgx_stream_create(ctypes.byref(s)))
            streams.append(s)
            bufs.append((Dev(_lib, nmax), Dev(_lib, nmax), Dev(_lib, 2 * nmax), Pinned(_lib, 2 * nmax)))
        nxt = [0]
        lock = threading.Lock()
        errs = []

        def worker(k):
            A, B, O, H = bufs[k]
            s = streams[k]
            try:
                while True:
                    with lock:
                        i = nxt[0]
                        nxt[0] += 1
                    if i >= count:
                        return
                    na, nb = na_l[i], nb_l[i]
                    _lib.check(L_.hgx_upload(A.ptr, HA.ptr, na, s))
                    _lib.check(L_.hgx_upload(B.ptr, HB.ptr, nb, s))
                    _lib.check(L_.hgx_download(H.ptr, O.ptr, na + nb, s))
            except Exception as e:  # noqa: BLE001 -- re-raised below
                errs.append(e)

        th = [threading.Thread(target=worker, args=(k,)) for k in range(lanes)]
        t0 = time.perf_counter()
        for t in th:
            t.start()
        for t in th:
            t.join()
        dt = time.perf_counter() - t0
        if errs:
            raise errs[0]
    finally:
        for s in streams:
            _lib.check(L_.hgx_stream_destroy(s))
        for group in bufs:
            for b in group:
                b.free()
        HA.free()
        HB.free()
    limbs = sum(na_l) + sum(nb_l)
    return {"requests": count, "lanes": lanes, "seconds": round(dt, 4), "limbs_per_s": limbs / dt,
            "GBps_each_way": round(8 * limbs / dt / 1e9, 1)}


class _Stream:
    """Pinned host buffers of the serving legs: two random 2^24-limb operand
    buffers (requests take slices of them) with their prefix residues mod
    2^61 - 1, and SLOTS product buffers."""

    def __init__(self, _lib, slots: int, seed: int):
        from tests import _residue as RES
        self.nmax = 1 << 24
        self.A, self.B = Pinned(_lib, self.nmax), Pinned(_lib, self.nmax)
        self.a, self.b = self.A.arr, self.B.arr
        self.a[:] = W.splitmix64(seed + 3, self.nmax)
        self.b[:] = W.splitmix64(seed + 4, self.nmax)
        self.PA = RES.prefix_residues(self.a)
        self.PB = RES.prefix_residues(self.b)
        self.outs = [Pinned(_lib, 2 * self.nmax) for _ in range(slots)]

    def release(self):
        self.a = self.b = None
        for d in [self.A, self.B] + self.outs:
            d.free()


def serve_stream(_lib, st: _Stream, rate: float, seconds: float, seed: int, checkers: int = 12,
                 gmp_sample: int = 0, check: bool = True) -> dict:
    """Poisson arrivals at RATE/s for SECONDS through hgx_submit, operand
    sizes log-uniform in [2^10, 2^24] limbs (SEED), operands slices of the
    pinned buffers of ST, products into its slots.  Latency = completion seen
    by the client - scheduled arrival.  Every product is checked: its residue
    mod 2^61 - 1 (a checker thread pool, which frees the slot afterwards)
    against the product of the operands' residues; GMP_SAMPLE of them are
    also compared limb for limb with GMP after the stream."""
    import ctypes
    import queue
    from concurrent.futures import ThreadPoolExecutor

    import numpy as np

    from tests import _gmp
    from tests import _residue as RES
    L_ = _lib.lib()
    count = max(1, int(rate * seconds))
    na_l, nb_l = W.serve_shapes(count, seed)
    gaps = W.poisson_gaps(seed + 1, count, rate)
    offs = [int(v) for v in W.splitmix64(seed + 2, count)]
    free = queue.SimpleQueue()
    for k in range(len(st.outs)):
        free.put(k)
    sample_every = max(1, count // gmp_sample) if gmp_sample else 0
    kept = {}
    mism = []
    checked = [0]
    pool = ThreadPoolExecutor(max_workers=checkers)

    def check_one(i, slot, n, want):
        try:
            if not check:
                return
            got = RES.residue_ptr(st.outs[slot].ptr, n)
            if sample_every and i % sample_every == 0:
                kept[i] = st.outs[slot].arr[:n].copy()
            if got != want:
                mism.append(i)
            checked[0] += 1
        finally:
            free.put(slot)

    inflight = {}
    lat, arr_t, limbs = [], [], 0

    def reap():
        nonlocal limbs
        # hgx_poll releases a ticket when it reports completion; a failed
        # request raises here
        done = [t for t in inflight if _lib.check(L_.hgx_poll(ctypes.c_uint64(t))) == 1]
        now = time.perf_counter()
        for t in done:
            i, arr, slot, n, want = inflight.pop(t)
            lat.append(now - arr)
            arr_t.append(arr)
            limbs += n
            pool.submit(check_one, i, slot, n, want)
        return bool(done)

    stats0 = _lib.server_stats()
    t0 = time.perf_counter()
    sched = t0
    for i in range(count):
        sched += gaps[i]
        na, nb = na_l[i], nb_l[i]
        while time.perf_counter() < sched:
            reap()
        while True:
            try:
                slot = free.get_nowait()
                break
            except queue.Empty:
                reap()
        oa = (offs[i] >> 8) % (st.nmax - na + 1)
        ob = (offs[i] >> 32) % (st.nmax - nb + 1)
        want = RES.mulmod(RES.slice_residue(st.PA, oa, na), RES.slice_residue(st.PB, ob, nb))
        d = _lib.MulDesc(st.A.ptr + 8 * oa, na, st.B.ptr + 8 * ob, nb, st.outs[slot].ptr)
        t = ctypes.c_uint64(0)
        _lib.check(L_.hgx_submit(ctypes.byref(d), ctypes.byref(t)))
        inflight[t.value] = (i, sched, slot, na + nb, want)
    while inflight:
        reap()
    wall = time.perf_counter() - t0
    pool.shutdown(wait=True)
    stats1 = _lib.server_stats()
    gmp_bad = 0
    for i, got in kept.items():
        na, nb = na_l[i], nb_l[i]
        oa = (offs[i] >> 8) % (st.nmax - na + 1)
        ob = (offs[i] >> 32) % (st.nmax - nb + 1)
        if not np.array_equal(got, _gmp.mul(st.a[oa:oa + na], st.b[ob:ob + nb])):
            gmp_bad += 1
    order = sorted(range(len(lat)), key=lambda k: arr_t[k])
    half = len(order) // 2

    def q(xs, f):
        xs = sorted(xs)
        return round(1e3 * xs[min(len(xs) - 1, int(f * len(xs)))], 3) if xs else None

    first = [lat[k] for k in order[:half]]
    second = [lat[k] for k in order[half:]]
    p99a, p99b = q(first, 0.99), q(second, 0.99)
    launched = stats1["launch_s"] - stats0["launch_s"]
    return {"rate_per_s": rate, "seconds": seconds, "requests": count, "limbs": limbs,
            "offered_limbs_per_s": (sum(na_l) + sum(nb_l)) / seconds, "wall_s": round(wall, 4),
            "limbs_per_s": limbs / wall, "p50_ms": q(lat, 0.5), "p99_ms": q(lat, 0.99),
            "p99_first_half_ms": p99a, "p99_second_half_ms": p99b,
            "bounded": bool(p99a and p99b and p99b <= 1.5 * p99a and wall <= seconds + 1.0),
            "outputs_checked": checked[0], "mismatches": len(mism), "gmp_checked": len(kept),
            "gmp_mismatches": gmp_bad,
            "server_rounds": stats1["rounds"] - stats0["rounds"],
            "server_thread_launch_frac": round(launched / wall, 4),
            "server_thread_retire_frac": round((stats1["retire_s"] - stats0["retire_s"]) / wall, 4),
            "server_thread_wait_frac": round((stats1["wait_s"] - stats0["wait_s"]) / wall, 4)}


def serving_leg(_lib, rate: float = 2000.0, seconds: float = 10.0, slots: int = 32,
                seed: int = W.SERVE_SEED, search=(1500.0, 1200.0, 1000.0, 800.0, 600.0, 400.0),
                search_seconds: float = 4.0) -> dict:
    """Request server under load (the serving configuration): Poisson
    arrivals at RATE/s for SECONDS, then streams of SEARCH_SECONDS at
    decreasing rates until one keeps its p99 bounded (the second half's p99
    at most 1.5x the first half's, and the stream drained within a second of
    its end).  The root tables of every transform length in the range are
    built first (one request per length), as a server keeps them resident.
    Every served product is checked (serve_stream)."""
    import ctypes
    L_ = _lib.lib()
    st = _Stream(_lib, slots, seed)
    try:
        for bits in range(10, 25):
            d = _lib.MulDesc(st.A.ptr, 1 << bits, st.B.ptr, 1 << bits, st.outs[0].ptr)
            t = ctypes.c_uint64(0)
            _lib.check(L_.hgx_submit(ctypes.byref(d), ctypes.byref(t)))
            _lib.check(L_.hgx_wait(t))
        main = serve_stream(_lib, st, rate, seconds, seed, gmp_sample=6)
        sustained = None
        tried = []
        for r in search:
            res = serve_stream(_lib, st, r, search_seconds, seed + 0x100 + int(r))
            tried.append({k: res[k] for k in ("rate_per_s", "limbs_per_s", "p50_ms", "p99_ms", "p99_first_half_ms",
                                               "p99_second_half_ms", "bounded", "outputs_checked", "mismatches")})
            if res["bounded"]:
                sustained = res
                break
        stats = _lib.server_stats()
    finally:
        st.release()
    main["arena_bytes"] = int(os.environ.get("HGX_SERVE_ARENA_BYTES", "0"))
    main["workload"] = (f"Poisson arrivals {rate:.0f}/s for {seconds} s (seed {seed:#x}), operand limbs "
                        f"log-uniform in [2^10, 2^24], pinned host buffers, {slots} product slots, server arena "
                        f"{main['arena_bytes'] >> 30} GiB, root tables built before the stream, every product "
                        f"checked mod 2^61-1 and {main['gmp_checked']} against GMP")
    main["sustained"] = sustained
    main["rate_search"] = tried
    main["server"] = {"lanes": stats["lanes"], "split_downloads": stats["split_downloads"],
                      "uploads": stats["uploads"], "downloads": stats["downloads"],
                      "staged": stats["staged"], "failed": stats["failed"]}
    return main


def staged_leg(h, _lib, bits: int = 26, limits=(4 << 30, 1 << 30), reps: int = 2) -> dict:
    """Host-memory product of two 2^bits-limb operands: whole in HBM (operands
    up and the product down over PCIe), then host-staged under device-memory
    limits (chunk products summed diagonal by diagonal); every result equal."""
    import numpy as np
    L_ = _lib.lib()
    n = 1 << bits
    A, B = Dev(_lib, n), Dev(_lib, n)
    h.fill_splitmix64(A.ptr, n, 0x48475801)
    h.fill_splitmix64(B.ptr, n, 0x48475802)
    a = download(_lib, A.ptr, n)
    b = download(_lib, B.ptr, n)
    A.free()
    B.free()

    def run():
        t = []
        out = None
        for _ in range(reps):
            t0 = time.perf_counter()
            out = h.mul_limbs(a, b)
            t.append(time.perf_counter() - t0)
        return min(t), out

    res = {"workload": f"2^{bits} x 2^{bits} limbs from host memory", "limbs": 2 * n}
    t, whole = run()
    res["whole_in_hbm"] = {"ms": round(1e3 * t, 1), "limbs_per_s": 2 * n / t}
    res["staged"] = []
    try:
        for lim in limits:
            _lib.check(L_.hgx_staging_limit(lim))
            t, got = run()
            res["staged"].append({"limit_bytes": lim, "ms": round(1e3 * t, 1), "limbs_per_s": 2 * n / t,
                                  "identical": bool(np.array_equal(got, whole))})
    finally:
        _lib.check(L_.hgx_staging_limit(0))
    return res


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--bits", type=int, default=26, help="operand size 2^bits limbs each")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-bits", type=int, default=22)
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-vs-per-call leg")
    ap.add_argument("--no-square", action="store_true", help="skip the 2^bits-limb square leg")
    ap.add_argument("--no-serve", action="store_true", help="skip the request-server leg")
    ap.add_argument("--serve-seconds", type=float, default=10.0, help="length of the 2000/s serving stream")
    ap.add_argument("--serve-slots", type=int, default=32, help="pinned product buffers of the serving client")
    ap.add_argument("--no-serve-search", action="store_true", help="skip the sustainable-rate search")
    ap.add_argument("--no-latency", action="store_true", help="skip the 2^16 x 2^16 per-call latency leg")
    ap.add_argument("--no-pi", action="store_true", help="skip the mpd-pi leg")
    ap.add_argument("--no-staged", action="store_true", help="skip the host-staged product leg")
    ap.add_argument("--pi-bits", type=int, default=28, help="mpd-pi precision 2^bits")
    ap.add_argument("--pi-host-iterates", action="store_true", help="also time mpd-pi with host-resident iterates")
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner (default pass structures)")
    ap.add_argument("--calib-copy", type=int, default=0, metavar="MIB",
                    help="first run one plain device copy of MIB MiB (reference kernel for HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); products are independent per GPU "
              f"and no multi-GPU path is part of this build", file=sys.stderr)
        return 2

    import hypergeometrica_amd as h
    from hypergeometrica_amd import _lib

    # every device buffer, stream and pinned host buffer comes from the
    # library: the process runs on the one HIP runtime the library is built
    # against
    L_ = _lib.lib(0)
    n = 1 << args.bits
    na = nb = n
    A, B, OUT = Dev(_lib, na), Dev(_lib, nb), Dev(_lib, na + nb)
    L_bits = (2 * n - 1).bit_length()
    h.fill_splitmix64(A.ptr, na, 0x48475801)
    h.fill_splitmix64(B.ptr, nb, 0x48475802)
    sync(_lib)
    if args.calib_copy:
        words = args.calib_copy << 17
        src, dst = Dev(_lib, words), Dev(_lib, words)
        _lib.check(L_.hgx_device_fill(src.ptr, words, 1, None))
        _lib.check(L_.hgx_device_copy(dst.ptr, src.ptr, words, None))
        sync(_lib)
        print(f"# calibration copy: {words * 8} bytes read, {words * 8} bytes written", file=sys.stderr, flush=True)
        src.free()
        dst.free()

    # the autotuner picks the pass structure of every length the legs use
    tuned = {}
    if not args.no_tune:
        import ctypes
        t0 = time.perf_counter()
        _lib.check(L_.hgx_autotune(13, L_bits, 2))
        for b in range(13, L_bits + 1):
            c, dm, md = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
            _lib.check(L_.hgx_plan_get(b, ctypes.byref(c), ctypes.byref(dm), ctypes.byref(md)))
            tuned[b] = (c.value, dm.value, md.value)
        print(f"# autotuned 2^13..2^{L_bits} in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)

    def step():
        _lib.check(L_.hgx_mul_device(A.ptr, na, B.ptr, nb, OUT.ptr, None))

    for _ in range(args.warmup):
        step()
    sync(_lib)
    print(f"# warmup done ({args.warmup} steps)", file=sys.stderr, flush=True)

    # timed region: K steps, wall clock, synchronised on both sides; the
    # library records a HIP event pair around every kernel on its launch
    # stream during these steps (pooled events, no measurable overhead)
    per_step = []
    _lib.timing_enable(True)
    sync(_lib)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s0 = time.perf_counter()
        step()
        sync(_lib)
        per_step.append(time.perf_counter() - s0)
    sync(_lib)
    wall = time.perf_counter() - t0
    kt = _lib.timing_read()
    _lib.timing_enable(False)
    ms_per_step = 1e3 * wall / args.steps
    limbs = na + nb
    value = limbs * args.steps / wall
    print(f"# timed {args.steps} steps: {ms_per_step:.3f} ms/step", file=sys.stderr, flush=True)
    L = h.product_length(na, nb)
    kernels = {}
    for name, (cnt, tot) in kt.items():
        avg = tot / cnt
        byts = kernel_bytes(name, L, na, nb, 2)
        kernels[name] = {"launches_per_step": cnt / args.steps, "avg_ms": round(avg, 4),
                         "ms_per_step": round(tot / args.steps, 4),
                         "alg_bytes": byts,
                         "GBps": round(byts / (avg * 1e-3) / 1e9, 1) if byts else None}
    dom = max(kernels, key=lambda k: kernels[k]["ms_per_step"])
    d = kernels[dom]
    achieved = d["alg_bytes"] / (d["avg_ms"] * 1e-3) if d["alg_bytes"] else None
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"traffic_{dom}.json")
    if os.path.exists(tpath):
        try:
            traffic = json.load(open(tpath)).get("bytes_per_launch")
        except Exception:
            traffic = None
    # integer-ALU roofline of the same kernel: the time its modular
    # arithmetic takes at the rates of the register-only probe, over its
    # measured time
    probe = alu_probe(_lib)
    srow = (tuned.get(L_bits, (None,))[0] if tuned else None)
    srow = (L_bits - 8 * srow) if srow is not None else (L_bits - 16)
    for name, k in kernels.items():
        ops = kernel_alu_ops(name, L, srow)
        if ops:
            k["alu_floor_ms"] = round(1e3 * alu_floor_s(ops, probe), 4)
            k["alu_frac"] = round(k["alu_floor_ms"] / k["avg_ms"], 4)
    valu = None
    ops = kernel_alu_ops(dom, L, srow)
    if ops:
        fl = alu_floor_s(ops, probe)
        valu = {"kernel": dom, "bound": "alu", "ops_per_launch": ops, "floor_ms": round(1e3 * fl, 4),
                "measured_ms": d["avg_ms"], "frac": round(1e3 * fl / d["avg_ms"], 4),
                "probe": probe, "probe_clock_ghz": probe["fwd_butterfly"]["clock_ghz"],
                "note": "frac = time of the launch's butterflies, twiddle and pointwise products at the register-only "
                        "probe's rates (hgx_modmul_probe) over the launch's measured time"}
    roofline = {"bound": "hbm", "kernel": dom,
                "achieved": round(achieved / 1e9, 1) if achieved else None,
                "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK, 4) if achieved else None,
                "traffic": traffic}
    # whole-multiply fraction against the fused-minimum bytes B_mul
    P = -(-(L.bit_length() - 1) // 12)
    b_mul = 16 * (na + nb) + 3 * L * (48 * P - 16)
    per_step_sorted = sorted(per_step)
    p50 = per_step_sorted[len(per_step_sorted) // 2]
    p99 = per_step_sorted[min(len(per_step_sorted) - 1, int(0.99 * len(per_step_sorted)))]

    for d in (A, B, OUT):
        d.free()
    square = None
    if not args.no_square:
        print("# square ...", file=sys.stderr, flush=True)
        square = square_leg(h, _lib, bits=args.bits)

    batched = None
    if not args.no_batch:
        print("# batched vs per-call ...", file=sys.stderr, flush=True)
        batched = batched_leg(h, _lib, probe=probe)

    latency = None
    if not args.no_latency:
        print("# per-call latency ...", file=sys.stderr, flush=True)
        latency = latency_leg(h, _lib)

    serving = None
    if not args.no_serve:
        print("# serving ...", file=sys.stderr, flush=True)
        serving = serving_leg(_lib, seconds=args.serve_seconds, slots=args.serve_slots,
                              search=() if args.no_serve_search else (1500.0, 1200.0, 1000.0, 800.0, 600.0, 400.0))
        serving["pcie"] = pcie_probe(_lib)
        serving["frac_of_pcie_bound"] = round(serving["limbs_per_s"] / serving["pcie"]["bound_limbs_per_s"], 4)
        serving["copy_replay"] = copy_replay(_lib, lanes=serving["server"]["lanes"])
        serving["frac_of_copy_replay"] = round(serving["limbs_per_s"] / serving["copy_replay"]["limbs_per_s"], 4)

    staged = None
    if not args.no_staged:
        print("# host-staged products ...", file=sys.stderr, flush=True)
        staged = staged_leg(h, _lib, bits=args.bits)

    pi = None
    if not args.no_pi:
        print("# mpd-pi ...", file=sys.stderr, flush=True)
        pi = pi_leg(1 << args.pi_bits, host_iterates=args.pi_host_iterates)

    cpu = None
    if not args.no_cpu:
        print("# cpu baseline ...", file=sys.stderr, flush=True)
        cpu = cpu_baseline(args.cpu_bits)

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "limbs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {"workload": f"single exact multiply, 2^{args.bits} x 2^{args.bits} uniform 64-bit limbs, "
                               f"L=2^{L.bit_length() - 1}, 3 primes, operands resident in HBM",
                   "limbs_per_step": limbs, "transform_length": L,
                   "pass_structure": {"column_passes": tuned.get(L_bits, (None,) * 3)[0],
                                      "direct_table_max_bits": tuned.get(L_bits, (None,) * 3)[1],
                                      "direct_montgomery_tables": tuned.get(L_bits, (None,) * 3)[2],
                                      "autotuned": bool(tuned)}},
        "roofline": roofline,
        "valu_roofline": valu,
        "multiply_roofline": {"alg_bytes": b_mul, "frac": round(b_mul / (ms_per_step * 1e-3) / HBM_PEAK, 4)},
        "latency_ms": {"p50": round(1e3 * p50, 3), "p99": round(1e3 * p99, 3)},
        "kernels": kernels,
        "square": square,
        "batched": batched,
        "latency_2e16": latency,
        "serving": serving,
        "host_staged": staged,
        "mpd_pi": pi,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    # release the library while the HIP runtime is fully up, before the
    # interpreter's and the runtime's own teardown
    _lib.shutdown()
    return 0


if __name__ == "__main__":
    sys.exit(main())
