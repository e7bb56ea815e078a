"""Writes the golden fixtures under tests/golden/.

Run from the repository root:  python tests/golden/gen_golden.py

Sources:
  * logs.json -- every ("Transform length", "Convolution bits", "Moduli")
    triple printed by the reference's verbose multiply (logs/*.txt, e.g.
    logs/24apr21-999-ram.txt:32-34).  Needs /root/reference (read as text).
  * scheme.json -- closed-form scheme constants computed here with exact
    Python integers: primes i*2^55+1 of 63 bits, smallest prime
    generators, Barrett inverses floor(2^126/p), primitive roots.
  * ntt_vectors.npz -- direct DFTs over Z/pZ of seeded vectors (N = 2^6,
    2^8, 2^10) for each prime: forward output in bit-reversed order.
  * mul_vectors.npz -- seeded operands and their exact products (Python int).
Nothing here imports or runs the reference.
"""
from __future__ import annotations

import glob
import json
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference"


def is_prime(n: int) -> bool:
    if n < 2:
        return False
    small = [2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41]
    for p in small:
        if n % p == 0:
            return n == p
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    for a in small:
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(s - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


def prime_factors(n: int) -> list[int]:
    f, d = [], 2
    while d * d <= n:
        if n % d == 0:
            f.append(d)
            while n % d == 0:
                n //= d
        d += 1
    if n > 1:
        f.append(n)
    return f


def smallest_prime_generator(p: int) -> int:
    fs = prime_factors(p - 1)
    g = 2
    while True:
        if is_prime(g) and all(pow(g, (p - 1) // q, p) != 1 for q in fs):
            return g
        g += 1


def splitmix64(seed: int, n: int) -> np.ndarray:
    M = 2**64 - 1
    out = np.empty(n, dtype=np.uint64)
    for i in range(n):
        z = (seed + (i + 1) * 0x9E3779B97F4A7C15) & M
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        out[i] = z ^ (z >> 31)
    return out


def to_int(a) -> int:
    return int.from_bytes(np.asarray(a, dtype="<u8").tobytes(), "little")


def to_limbs(x: int, n: int) -> np.ndarray:
    return np.frombuffer(x.to_bytes(8 * n, "little"), dtype="<u8").astype(np.uint64)


def bitrev(i: int, bits: int) -> int:
    return int(format(i, f"0{bits}b")[::-1], 2) if bits else 0


def gen_logs():
    rows = []
    for path in sorted(glob.glob(os.path.join(REF, "logs", "*.txt"))):
        lines = open(path, errors="replace").read().splitlines()
        for i, ln in enumerate(lines):
            m = re.match(r"Transform length: (\d+)", ln)
            if not m or i + 2 >= len(lines):
                continue
            b = re.match(r"Convolution bits: (\d+)", lines[i + 1])
            mo = re.match(r"Moduli: (.*)", lines[i + 2])
            if b and mo:
                mods = [int(x.strip().replace("#x", ""), 16) for x in mo.group(1).split(",")]
                rows.append({"file": "logs/" + os.path.basename(path), "line": i + 1,
                             "length": int(m.group(1)), "bits": int(b.group(1)), "moduli": mods})
    json.dump(rows, open(os.path.join(HERE, "logs.json"), "w"), indent=0)
    print("logs.json:", len(rows), "records")


def gen_scheme():
    primes = [i * 2**55 + 1 for i in range(255, 127, -1) if is_prime(i * 2**55 + 1)]
    primes = [p for p in primes if p.bit_length() == 63]
    gens = [smallest_prime_generator(p) for p in primes]
    d = {
        "moduli": primes,
        "generators": gens,
        "barrett_inverses": [(1 << 126) // p for p in primes],
        "two_adicity": [((p - 1) & -(p - 1)).bit_length() - 1 for p in primes],
        "omega": {str(n): [pow(g, (p - 1) >> n, p) for p, g in zip(primes, gens)] for n in (1, 2, 8, 12, 17, 27)},
    }
    json.dump(d, open(os.path.join(HERE, "scheme.json"), "w"), indent=1)
    print("scheme.json:", [hex(p) for p in primes], gens)
    return primes, gens


def gen_ntt(primes, gens):
    out = {}
    for k, (p, g) in enumerate(zip(primes, gens)):
        for nb in (6, 8, 10):
            N = 1 << nb
            x = [int(v) % p for v in splitmix64(0x4E545400 + 16 * k + nb, N)]
            w = pow(g, (p - 1) >> nb, p)
            X = [sum(x[j] * pow(w, j * kk, p) for j in range(N)) % p for kk in range(N)]
            fwd = [X[bitrev(i, nb)] for i in range(N)]
            out[f"in_{k}_{nb}"] = np.array(x, dtype=np.uint64)
            out[f"fwd_{k}_{nb}"] = np.array(fwd, dtype=np.uint64)
    np.savez_compressed(os.path.join(HERE, "ntt_vectors.npz"), **out)
    print("ntt_vectors.npz:", len(out), "arrays")


def gen_mul():
    cases = [(1, 1), (2, 1), (3, 3), (101, 101), (150, 120), (1000, 37), (2048, 2048), (2049, 2047),
             (4096, 1), (5000, 3000)]
    out = {}
    for i, (na, nb) in enumerate(cases):
        a = splitmix64(0x48475800 + 2 * i, na)
        b = splitmix64(0x48475801 + 2 * i, nb)
        prod = to_int(a) * to_int(b)
        out[f"a_{i}"] = a
        out[f"b_{i}"] = b
        out[f"p_{i}"] = to_limbs(prod, na + nb)
    # all-ones operands (longest carry chains)
    for j, n in enumerate((128, 4096, 5000)):
        a = np.full(n, 2**64 - 1, dtype=np.uint64)
        out[f"a_{len(cases) + j}"] = a
        out[f"b_{len(cases) + j}"] = a
        out[f"p_{len(cases) + j}"] = to_limbs(to_int(a) ** 2, 2 * n)
    np.savez_compressed(os.path.join(HERE, "mul_vectors.npz"), **out)
    print("mul_vectors.npz:", len(out) // 3, "products")


if __name__ == "__main__":
    if os.path.isdir(REF):
        gen_logs()
    else:
        print("reference not present; logs.json left as is", file=sys.stderr)
    primes, gens = gen_scheme()
    gen_ntt(primes, gens)
    gen_mul()
