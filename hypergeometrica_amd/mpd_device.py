"""Newton iterations with the mantissas resident in device memory.

The reference's mpd-reciprocal (src/mpd-reciprocal.lisp:27-73) and
mpd-inv-sqrt (src/mpd-sqrt.lisp:40-87) chain dependent mpd-* products and
mpd-+ additions.  Here every mantissa of the iteration stays in HBM: products
through hgx_mul_device / hgx_sqr_device (the square for x*x, as the
reference's (eq x y) dispatch), sums and differences through hgx_add_device /
hgx_sub_device, limb shifts and truncation as device copies and views.  Only
the scalars the iteration branches on (sizes, the leading limb for the seed,
magnitude comparisons) are read back.  Every step computes the same integers
as Context in mpd.py, so the results are bit-identical to it.

Limb vectors are torch int64 tensors on the device holding uint64 bit
patterns, little endian, trimmed (the top limb non-zero); torch is used for
device memory and for the stream every call is ordered on.  Import torch
before libhgx.so is first loaded (this module does): torch ships its own HIP
runtime, and the process must bind both to it.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch  # noqa: F401  (before the library is loaded: see above)

from . import _lib
from . import mpd as _mpd
from .mpz import check

_SIGN_BIT = -(1 << 63)


def _torch():
    import torch
    return torch


def _stream() -> int:
    """The current torch stream, on which the library calls are queued so
    that they are ordered with torch's own work.  The null stream would send
    them to the library's stream instead, unordered with torch's."""
    h = _torch().cuda.current_stream().cuda_stream
    if not h:
        raise RuntimeError("mpd_device runs on a non-default torch stream (DeviceContext sets one)")
    return h


@dataclass
class DMPD:
    """sign * magnitude * 2^(64 * exponent); magnitude = LIMBS (device), empty for zero."""
    sign: int
    limbs: "object"
    exponent: int

    @property
    def size(self) -> int:
        return int(self.limbs.numel())


def _trim(t):
    """Drop high zero limbs (reads the index of the top non-zero limb)."""
    torch = _torch()
    nz = torch.nonzero(t)
    if nz.numel() == 0:
        return t[:0]
    return t[: int(nz[-1, 0]) + 1]


def to_device(x: _mpd.MPD) -> DMPD:
    torch = _torch()
    if x.mantissa == 0:
        return DMPD(0, torch.empty(0, dtype=torch.int64, device="cuda"), x.exponent)
    m = abs(x.mantissa)
    n = (m.bit_length() + 63) // 64
    host = np.frombuffer(m.to_bytes(8 * n, "little"), dtype="<i8").copy()
    return DMPD(1 if x.mantissa > 0 else -1, torch.from_numpy(host).to("cuda"), x.exponent)


def to_host(x: DMPD) -> _mpd.MPD:
    if x.sign == 0:
        return _mpd.MPD(0, x.exponent)
    b = x.limbs.cpu().numpy().astype("<i8").tobytes()
    return _mpd.MPD(x.sign * int.from_bytes(b, "little"), x.exponent)


def _mag_cmp(a, b) -> int:
    """Sign of |a| - |b| for trimmed magnitudes."""
    torch = _torch()
    if a.numel() != b.numel():
        return 1 if a.numel() > b.numel() else -1
    nz = torch.nonzero(a != b)
    if nz.numel() == 0:
        return 0
    i = int(nz[-1, 0])
    # unsigned order of the two limbs: flip the sign bits, compare as signed
    x, y = int(a[i]) ^ _SIGN_BIT, int(b[i]) ^ _SIGN_BIT
    return 1 if x > y else -1


class DeviceContext:
    """mpd-*, mpd-+ and the Newton iterations on device-resident mantissas."""

    def __init__(self, truncate: bool = False):
        self.truncate = truncate
        self.products = 0
        self.max_product_limbs = 0
        self.stream = _torch().cuda.Stream()  # every device step of this context runs here

    # -- src/mpd.lisp
    def mpd_mul(self, a: DMPD, b: DMPD) -> DMPD:
        """mpd-* (src/mpd.lisp:83-88): mpz-* of the mantissas on the device."""
        torch = _torch()
        self.products += 1
        self.max_product_limbs = max(self.max_product_limbs, a.size + b.size)
        if a.sign == 0 or b.sign == 0:
            return DMPD(0, a.limbs[:0], a.exponent + b.exponent)
        out = torch.empty(a.size + b.size, dtype=torch.int64, device="cuda")
        L = _lib.lib()
        if a is b:
            check(L.hgx_sqr_device(a.limbs.data_ptr(), a.size, out.data_ptr(), _stream()))
        else:
            check(L.hgx_mul_device(a.limbs.data_ptr(), a.size, b.limbs.data_ptr(), b.size, out.data_ptr(), _stream()))
        return DMPD(a.sign * b.sign, _trim(out), a.exponent + b.exponent)

    @staticmethod
    def _shifted(x: DMPD, k: int):
        """The magnitude times 2^(64 k): K zero limbs below it."""
        torch = _torch()
        if k == 0:
            return x.limbs
        out = torch.zeros(x.size + k, dtype=torch.int64, device="cuda")
        out[k:] = x.limbs
        return out

    def mpd_add(self, a: DMPD, b: DMPD) -> DMPD:
        """mpd-+ (src/mpd.lisp:90-99): align to the smaller exponent, add."""
        torch = _torch()
        m = min(a.exponent, b.exponent)
        if a.sign == 0:
            return DMPD(b.sign, self._shifted(b, b.exponent - m), m) if b.sign else DMPD(0, a.limbs, m)
        if b.sign == 0:
            return DMPD(a.sign, self._shifted(a, a.exponent - m), m)
        x, y = self._shifted(a, a.exponent - m), self._shifted(b, b.exponent - m)
        L = _lib.lib()
        n = max(x.numel(), y.numel())
        if a.sign == b.sign:
            out = torch.empty(n + 1, dtype=torch.int64, device="cuda")
            check(L.hgx_add_device(x.data_ptr(), x.numel(), y.data_ptr(), y.numel(), out.data_ptr(), _stream()))
            return DMPD(a.sign, _trim(out), m)
        c = _mag_cmp(x, y)
        if c == 0:
            return DMPD(0, x[:0], m)
        if c < 0:
            x, y, s = y, x, b.sign
        else:
            s = a.sign
        out = torch.empty(n, dtype=torch.int64, device="cuda")
        check(L.hgx_sub_device(x.data_ptr(), x.numel(), y.data_ptr(), y.numel(), out.data_ptr(), _stream()))
        return DMPD(s, _trim(out), m)

    def mpd_sub(self, a: DMPD, b: DMPD) -> DMPD:
        return self.mpd_add(a, DMPD(-b.sign, b.limbs, b.exponent))

    @staticmethod
    def mpd_oom(x: DMPD) -> tuple[int, bool]:
        """mpd-oom (src/mpd.lisp:23-41)."""
        if x.sign == 0:
            return 0, True
        return min(0, x.size + x.exponent), False

    @staticmethod
    def mpd_half(a: DMPD) -> DMPD:
        """mpd-half! (src/mpd.lisp:77-81): mantissa * 2^63, exponent - 1."""
        torch = _torch()
        if a.sign == 0:
            return DMPD(0, a.limbs, a.exponent - 1)
        n = a.size
        out = torch.empty(n + 1, dtype=torch.int64, device="cuda")
        # limb i of the result: bit 0 of limb i at bit 63, bits 1..63 of limb i - 1 below it
        low = torch.where(torch.bitwise_and(a.limbs, 1) == 1, torch.full_like(a.limbs, _SIGN_BIT),
                          torch.zeros_like(a.limbs))
        high = torch.bitwise_and(torch.bitwise_right_shift(a.limbs, 1), (1 << 63) - 1)
        out[:n] = low
        out[n] = 0
        out[1:] |= high
        return DMPD(a.sign, _trim(out), a.exponent - 1)

    @staticmethod
    def mpd_truncate(a: DMPD, digits: int) -> DMPD:
        """mpd-truncate! (src/mpd.lisp:56-70): keep the DIGITS high limbs."""
        s = a.size
        if digits >= s:
            return a
        d = s - digits
        return DMPD(a.sign, _trim(a.limbs[d:]), a.exponent + d)

    def _one(self) -> DMPD:
        return to_device(_mpd.MPD(1, 0))

    def _top(self, a: DMPD) -> int:
        return int(a.limbs[-1]) & ((1 << 64) - 1)

    # -- src/mpd-reciprocal.lisp
    def _mpd_reciprocal(self, a: _mpd.MPD, bits: int) -> _mpd.MPD:
        """mpd-reciprocal (src/mpd-reciprocal.lisp:27-73) with A and every
        iterate in device memory."""
        digits = max(1, -(-bits // 64))
        da = to_device(a)
        size = da.size
        r = _mpd._round_div(_mpd.BASE, self._top(da))
        x = to_device(_mpd.MPD(-r if a.mantissa < 0 else r, -a.exponent - size))
        one = self._one()
        while True:
            h = self.mpd_sub(one, self.mpd_mul(da, x))
            zeros, zero = self.mpd_oom(h)
            if zero or -zeros >= digits:
                return to_host(x)
            x = self.mpd_add(x, self.mpd_mul(x, h))
            if self.truncate:
                x = self.mpd_truncate(x, digits + 2)

    # -- src/mpd-sqrt.lisp
    def _mpd_inv_sqrt(self, a: _mpd.MPD, bits: int) -> _mpd.MPD:
        """mpd-inv-sqrt (src/mpd-sqrt.lisp:40-87) with A and every iterate in
        device memory; x*x goes to the squaring path."""
        digits = max(1, -(-bits // 64))
        da = to_device(a)
        x = to_device(_mpd.Context.estimate_inv_sqrt(a))
        one = self._one()
        while True:
            h = self.mpd_sub(one, self.mpd_mul(da, self.mpd_mul(x, x)))
            zeros, zero = self.mpd_oom(h)
            if zero or -zeros >= digits:
                return to_host(x)
            x = self.mpd_half(self.mpd_add(self.mpd_add(x, x), self.mpd_mul(x, h)))
            if self.truncate:
                x = self.mpd_truncate(x, digits + 2)

    def _mpd_sqrt(self, a: _mpd.MPD, bits: int) -> _mpd.MPD:
        """mpd-sqrt (src/mpd-sqrt.lisp:89-90)."""
        inv = to_device(self._mpd_inv_sqrt(a, bits))
        return to_host(self.mpd_mul(to_device(a), inv))

    # -- src/pi.lisp
    def _mpd_pi(self, prec_bits: int) -> _mpd.MPD:
        """mpd-pi (src/pi.lisp:53-74): binary splitting with batched GPU
        products, then the Newton steps and the final products on the device."""
        from . import series as _series
        guard = prec_bits + _mpd.DIGIT_BITS
        num_terms = 2 + int(guard // _mpd.CHUD_BITS_PER_TERM)
        comp = _series.binary_split(_series.chudnovsky_series(), 0, num_terms)
        sqrt = to_device(self._mpd_sqrt(_mpd.MPD(10005, 0), guard))
        recip = to_device(self._mpd_reciprocal(_mpd.MPD(comp.numerator(), 0), guard))
        final = self.mpd_mul(sqrt, recip)
        den = to_device(_mpd.MPD(comp.denominator(), 0))
        return to_host(self.mpd_mul(final, den))

    def _on_stream(self, fn, *args):
        torch = _torch()
        with torch.cuda.stream(self.stream):
            out = fn(*args)
        self.stream.synchronize()
        return out

    def mpd_reciprocal(self, a: _mpd.MPD, bits: int) -> _mpd.MPD:
        """mpd-reciprocal (src/mpd-reciprocal.lisp:27-73) with A and every
        iterate in device memory."""
        return self._on_stream(self._mpd_reciprocal, a, bits)

    def mpd_inv_sqrt(self, a: _mpd.MPD, bits: int) -> _mpd.MPD:
        """mpd-inv-sqrt (src/mpd-sqrt.lisp:40-87) with A and every iterate in
        device memory; x*x goes to the squaring path."""
        return self._on_stream(self._mpd_inv_sqrt, a, bits)

    def mpd_sqrt(self, a: _mpd.MPD, bits: int) -> _mpd.MPD:
        """mpd-sqrt (src/mpd-sqrt.lisp:89-90)."""
        return self._on_stream(self._mpd_sqrt, a, bits)

    def mpd_pi(self, prec_bits: int) -> _mpd.MPD:
        """mpd-pi (src/pi.lisp:53-74): binary splitting with batched GPU
        products, then the Newton steps and the final products on the device."""
        return self._on_stream(self._mpd_pi, prec_bits)
