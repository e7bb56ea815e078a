"""Newton iterations with the mantissas resident in device memory, through
the C ABI alone.

The reference's mpd-reciprocal (src/mpd-reciprocal.lisp:27-73) and
mpd-inv-sqrt (src/mpd-sqrt.lisp:40-87) chain dependent mpd-* products and
mpd-+ additions.  Here every mantissa of the iteration stays in HBM, in
blocks of hgx_device_alloc: products through hgx_mul_device / hgx_sqr_device
(the square for x*x, as the reference's (eq x y) dispatch), sums and
differences through hgx_add_device / hgx_sub_device, the alignment of
mpd-+ by hgx_device_fill and hgx_device_copy, mpd-half! by hgx_shl_device,
mpd-truncate! as a view.  Only the scalars the iteration branches on are
read back (hgx_device_significant for sizes, hgx_device_compare for the sign
of a difference, one limb for the seed).  Every call is ordered on the
library's stream.  The steps are those of Context in mpd.py, on the same
integers, so the results are bit-identical to it; an SBCL host would make
the same calls through sb-alien (lisp/hgx.lisp).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib
from . import mpd as _mpd
from .mpz import check


class DBuf:
    """A block of device limbs (hgx_device_alloc), released (stream-ordered,
    hgx_device_free) when its last view goes."""

    __slots__ = ("ptr", "n", "gen")

    def __init__(self, n: int):
        p = ctypes.c_void_p(0)
        check(_lib.lib().hgx_device_alloc(max(n, 1), ctypes.byref(p)))
        self.ptr, self.n, self.gen = int(p.value), n, _lib.generation

    def __del__(self):
        if getattr(self, "gen", -1) == _lib.generation and _lib._initialised and _lib._lib is not None:
            _lib._lib.hgx_device_free(self.ptr, None)


@dataclass
class DMPD:
    """sign * magnitude * 2^(64 * exponent); the magnitude is SIZE limbs at
    limb OFF of BUF (trimmed: its top limb is non-zero), none for zero."""
    sign: int
    buf: DBuf | None
    off: int
    size: int
    exponent: int

    @property
    def ptr(self) -> int:
        return self.buf.ptr + 8 * self.off


def _L():
    return _lib.lib()


def _significant(ptr: int, n: int) -> int:
    out = ctypes.c_size_t(0)
    check(_L().hgx_device_significant(ptr, n, ctypes.byref(out), None))
    return int(out.value)


def _make(sign: int, buf: DBuf, n: int, exponent: int) -> DMPD:
    """A result of up to N limbs in BUF, trimmed (sign 0 when it is zero)."""
    k = _significant(buf.ptr, n)
    return DMPD(sign if k else 0, buf if k else None, 0, k, exponent)


def to_device(x: _mpd.MPD) -> DMPD:
    if x.mantissa == 0:
        return DMPD(0, None, 0, 0, x.exponent)
    m = abs(x.mantissa)
    n = (m.bit_length() + 63) // 64
    host = np.frombuffer(m.to_bytes(8 * n, "little"), dtype="<u8")  # read-only view, uploaded before it goes
    buf = DBuf(n)
    check(_L().hgx_upload(buf.ptr, host.ctypes.data, n, None))
    return DMPD(1 if x.mantissa > 0 else -1, buf, 0, n, x.exponent)


def from_limbs(sign: int, limbs: np.ndarray, exponent: int = 0) -> DMPD:
    """A device value from trimmed host limbs (no Python integer)."""
    n = int(limbs.size)
    if sign == 0 or n == 0:
        return DMPD(0, None, 0, 0, exponent)
    host = np.ascontiguousarray(limbs, dtype=np.uint64)
    buf = DBuf(n)
    check(_L().hgx_upload(buf.ptr, host.ctypes.data, n, None))
    return DMPD(sign, buf, 0, n, exponent)


def split_to_device(series, lower: int, upper: int, which: str = "pqbr") -> dict[str, DMPD]:
    """The native split's (hgx_series_split) P, Q, B, R named in WHICH as
    device values: copied device to device when the split ran in device
    memory (hgx_partial_get_device), uploaded when it ran on the host."""
    from . import series as _series
    L = _L()
    desc = _series.native_desc(series)
    res = ctypes.c_void_p(0)
    check(L.hgx_series_split(ctypes.byref(desc), lower, upper, ctypes.byref(res)))
    out = {}
    try:
        for w, k in enumerate("pqbr"):
            if k not in which:
                continue
            sign, n = ctypes.c_int(0), ctypes.c_size_t(0)
            check(L.hgx_partial_get_device(res, w, ctypes.byref(sign), ctypes.byref(n), None, None))
            if sign.value == 0 or n.value == 0:
                out[k] = DMPD(0, None, 0, 0, 0)
                continue
            buf = DBuf(n.value)
            check(L.hgx_partial_get_device(res, w, None, None, buf.ptr, None))
            out[k] = DMPD(sign.value, buf, 0, n.value, 0)
    finally:
        L.hgx_partial_free(res)
    return out


def to_host(x: DMPD) -> _mpd.MPD:
    if x.sign == 0:
        return _mpd.MPD(0, x.exponent)
    host = np.empty(x.size, dtype=np.uint64)
    check(_L().hgx_download(host.ctypes.data, x.ptr, x.size, None))
    return _mpd.MPD(x.sign * int.from_bytes(np.ascontiguousarray(host, dtype="<u8").tobytes(), "little"), x.exponent)


def _mag_cmp(a: DMPD, b: DMPD) -> int:
    """Sign of |a| - |b|."""
    out = ctypes.c_int(0)
    check(_L().hgx_device_compare(a.ptr, a.size, b.ptr, b.size, ctypes.byref(out), None))
    return out.value


class DeviceContext:
    """mpd-*, mpd-+ and the Newton iterations on device-resident mantissas."""

    def __init__(self, truncate: bool = False):
        self.truncate = truncate
        self.products = 0
        self.max_product_limbs = 0
        self.phase_s: dict[str, float] = {}  # mpd_pi's steps

    # -- src/mpd.lisp
    def mpd_mul(self, a: DMPD, b: DMPD) -> DMPD:
        """mpd-* (src/mpd.lisp:83-88): mpz-* of the mantissas on the device."""
        self.products += 1
        self.max_product_limbs = max(self.max_product_limbs, a.size + b.size)
        e = a.exponent + b.exponent
        if a.sign == 0 or b.sign == 0:
            return DMPD(0, None, 0, 0, e)
        out = DBuf(a.size + b.size)
        if a is b:
            check(_L().hgx_sqr_device(a.ptr, a.size, out.ptr, None))
        else:
            check(_L().hgx_mul_device(a.ptr, a.size, b.ptr, b.size, out.ptr, None))
        return _make(a.sign * b.sign, out, a.size + b.size, e)

    @staticmethod
    def _shifted(x: DMPD, k: int) -> DMPD:
        """The magnitude times 2^(64 k): K zero limbs below it."""
        if k == 0:
            return x
        out = DBuf(x.size + k)
        check(_L().hgx_device_fill(out.ptr, k, 0, None))
        check(_L().hgx_device_copy(out.ptr + 8 * k, x.ptr, x.size, None))
        return DMPD(x.sign, out, 0, x.size + k, x.exponent - k)

    def mpd_add(self, a: DMPD, b: DMPD) -> DMPD:
        """mpd-+ (src/mpd.lisp:90-99): align to the smaller exponent, add."""
        m = min(a.exponent, b.exponent)
        if a.sign == 0:
            return self._shifted(b, b.exponent - m) if b.sign else DMPD(0, None, 0, 0, m)
        if b.sign == 0:
            return self._shifted(a, a.exponent - m)
        x, y = self._shifted(a, a.exponent - m), self._shifted(b, b.exponent - m)
        n = max(x.size, y.size)
        if a.sign == b.sign:
            out = DBuf(n + 1)
            check(_L().hgx_add_device(x.ptr, x.size, y.ptr, y.size, out.ptr, None))
            return _make(a.sign, out, n + 1, m)
        c = _mag_cmp(x, y)
        if c == 0:
            return DMPD(0, None, 0, 0, m)
        if c < 0:
            x, y, s = y, x, b.sign
        else:
            s = a.sign
        out = DBuf(n)
        check(_L().hgx_sub_device(x.ptr, x.size, y.ptr, y.size, out.ptr, None))
        return _make(s, out, n, m)

    def mpd_sub(self, a: DMPD, b: DMPD) -> DMPD:
        return self.mpd_add(a, DMPD(-b.sign, b.buf, b.off, b.size, b.exponent))

    @staticmethod
    def mpd_oom(x: DMPD) -> tuple[int, bool]:
        """mpd-oom (src/mpd.lisp:23-41)."""
        if x.sign == 0:
            return 0, True
        return min(0, x.size + x.exponent), False

    @staticmethod
    def mpd_half(a: DMPD) -> DMPD:
        """mpd-half! (src/mpd.lisp:77-81): mantissa * 2^63, exponent - 1."""
        if a.sign == 0:
            return DMPD(0, None, 0, 0, a.exponent - 1)
        out = DBuf(a.size + 1)
        check(_L().hgx_shl_device(a.ptr, a.size, 63, out.ptr, None))
        return _make(a.sign, out, a.size + 1, a.exponent - 1)

    @staticmethod
    def mpd_truncate(a: DMPD, digits: int) -> DMPD:
        """mpd-truncate! (src/mpd.lisp:56-70): keep the DIGITS high limbs (a
        view; the top limb stays non-zero)."""
        if digits >= a.size:
            return a
        d = a.size - digits
        return DMPD(a.sign, a.buf, a.off + d, digits, a.exponent + d)

    def _one(self) -> DMPD:
        return to_device(_mpd.MPD(1, 0))

    @staticmethod
    def _top(a: DMPD) -> int:
        v = np.zeros(1, dtype=np.uint64)
        check(_L().hgx_download(v.ctypes.data, a.ptr + 8 * (a.size - 1), 1, None))
        return int(v[0])

    # -- src/mpd-reciprocal.lisp
    def mpd_reciprocal(self, a: _mpd.MPD, bits: int) -> _mpd.MPD:
        """mpd-reciprocal (src/mpd-reciprocal.lisp:27-73) with A and every
        iterate in device memory."""
        return to_host(self.reciprocal(to_device(a), bits))

    def reciprocal(self, da: DMPD, bits: int) -> DMPD:
        """mpd-reciprocal of a device value, left in device memory."""
        digits = max(1, -(-bits // 64))
        r = _mpd._round_div(_mpd.BASE, self._top(da))
        x = to_device(_mpd.MPD(-r if da.sign < 0 else r, -da.exponent - da.size))
        one = self._one()
        while True:
            h = self.mpd_sub(one, self.mpd_mul(da, x))
            zeros, zero = self.mpd_oom(h)
            if zero or -zeros >= digits:
                return x
            x = self.mpd_add(x, self.mpd_mul(x, h))
            if self.truncate:
                x = self.mpd_truncate(x, digits + 2)

    # -- src/mpd-sqrt.lisp
    def mpd_inv_sqrt(self, a: _mpd.MPD, bits: int) -> _mpd.MPD:
        """mpd-inv-sqrt (src/mpd-sqrt.lisp:40-87) with A and every iterate in
        device memory; x*x goes to the squaring path."""
        return to_host(self.inv_sqrt(a, bits))

    def inv_sqrt(self, a: _mpd.MPD, bits: int) -> DMPD:
        """mpd-inv-sqrt, the result left in device memory."""
        digits = max(1, -(-bits // 64))
        da = to_device(a)
        x = to_device(_mpd.Context.estimate_inv_sqrt(a))
        one = self._one()
        while True:
            h = self.mpd_sub(one, self.mpd_mul(da, self.mpd_mul(x, x)))
            zeros, zero = self.mpd_oom(h)
            if zero or -zeros >= digits:
                return x
            x = self.mpd_half(self.mpd_add(self.mpd_add(x, x), self.mpd_mul(x, h)))
            if self.truncate:
                x = self.mpd_truncate(x, digits + 2)

    def mpd_sqrt(self, a: _mpd.MPD, bits: int) -> _mpd.MPD:
        """mpd-sqrt (src/mpd-sqrt.lisp:89-90)."""
        return to_host(self.sqrt(a, bits))

    def sqrt(self, a: _mpd.MPD, bits: int) -> DMPD:
        """mpd-sqrt, the result left in device memory."""
        return self.mpd_mul(to_device(a), self.inv_sqrt(a, bits))

    # -- src/pi.lisp
    def mpd_pi(self, prec_bits: int) -> _mpd.MPD:
        """mpd-pi (src/pi.lisp:53-74): binary splitting by the library's
        native engine (hgx_series_split), then the Newton steps and the
        final products on the device.  The partial's limbs go to device
        memory device to device when the split ran there, and are uploaded
        otherwise (partial-denominator B*Q as a device product); only the
        result comes back as an integer.  The square root of 10005
        does not depend on the series: it runs on a second thread while the
        split runs (the same integers either way)."""
        import threading
        import time

        from . import series as _series
        guard = prec_bits + _mpd.DIGIT_BITS
        num_terms = 2 + int(guard // _mpd.CHUD_BITS_PER_TERM)
        box: dict = {}

        def root():
            t0 = time.perf_counter()
            try:
                box["sqrt"] = self.sqrt(_mpd.MPD(10005, 0), guard)
            except BaseException as e:  # raised again on the caller's thread
                box["error"] = e
            box["s"] = time.perf_counter() - t0

        side = threading.Thread(target=root)
        t = [time.perf_counter()]
        side.start()
        try:
            comp = split_to_device(_series.chudnovsky_series(), 0, num_terms, "bqr")
        finally:
            side.join()
        if "error" in box:
            raise box["error"]
        t.append(time.perf_counter())
        recip = self.reciprocal(comp["r"], guard)
        t.append(time.perf_counter())
        final = self.mpd_mul(box["sqrt"], recip)
        den = self.mpd_mul(comp["b"], comp["q"])
        out = to_host(self.mpd_mul(final, den))
        t.append(time.perf_counter())
        # seconds of each step (each ends with a read-back, so its device work
        # is done); the square root overlaps the split
        self.phase_s = {"binary_split_and_sqrt_10005": round(t[1] - t[0], 3), "sqrt_10005": round(box["s"], 3),
                        "reciprocal": round(t[2] - t[1], 3), "products_and_result": round(t[3] - t[2], 3)}
        return out
