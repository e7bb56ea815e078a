"""Size-independent check of huge products: residues modulo the Mersenne
prime q = 2^61 - 1.  Since 2^64 = 8 (mod q), limb i contributes
limb * 2^(3i mod 61), a 61-bit rotation, so the residue of a 2^27-limb
number takes a few vectorised numpy passes."""
from __future__ import annotations

import numpy as np

Q = (1 << 61) - 1


def mod_q(limbs: np.ndarray, chunk: int = 1 << 24) -> int:
    a = np.ascontiguousarray(limbs, dtype=np.uint64)
    q = np.uint64(Q)
    total = 0
    for i0 in range(0, a.size, chunk):
        c = a[i0:i0 + chunk]
        v = (c & q) + (c >> np.uint64(61))          # < 2^61 + 8
        v = np.where(v >= q, v - q, v)
        s = ((np.arange(c.size, dtype=np.uint64) + np.uint64(i0)) * np.uint64(3)) % np.uint64(61)
        rot = ((v << s) & q) | (v >> (np.uint64(61) - s))
        rot = np.where(s == 0, v, rot)
        lo = int(np.sum(rot & np.uint64((1 << 31) - 1), dtype=np.uint64))
        hi = int(np.sum(rot >> np.uint64(31), dtype=np.uint64))
        total += lo + (hi << 31)
    return total % Q
