"""Synthetic inputs of the measured configurations.

Limbs come from the SplitMix64 stream out[i] = mix(seed + (i+1) * 0x9E3779B97F4A7C15),
the same stream hgx_fill_splitmix64 writes on the device (with FORCE_TOP the
last limb is made non-zero, so that the limb count is exact).  Operand sizes
are log-uniform, n = floor(2^(lo + (hi - lo) u)) with u = next / 2^64 drawn
independently per operand, and request arrivals are Poisson,
gap = -ln(1 - u) / rate.  Seeds per configuration are 0x48475800 + index.

The reference ships no benchmark inputs (its tests randomize against CL
bignums, tests/mpz.lisp:32-47); these generators define the workloads the
benchmark and the tests share.
"""
from __future__ import annotations

import math

import numpy as np

GAMMA = 0x9E3779B97F4A7C15
M64 = (1 << 64) - 1

SINGLE_SEED = 0x48475801   # operands of the single large product: seeds +0 and +1
BATCH_SEED = 0x48475803    # the binary-splitting level of 4096 products
SERVE_SEED = 0x48475804    # the serving stream


def splitmix64(seed: int, n: int, force_top: bool = False) -> np.ndarray:
    """N words of the SplitMix64 stream of SEED (vectorised, wrapping uint64)."""
    i = np.arange(1, n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed & M64) + i * np.uint64(GAMMA)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    if force_top and n and z[-1] == 0:
        z[-1] = 1
    return z


def uniforms(seed: int, count: int) -> list[float]:
    """COUNT values next / 2^64 of the stream of SEED, in [0, 1)."""
    return [int(v) / 2.0**64 for v in splitmix64(seed, count)]


def log_uniform_sizes(seed: int, count: int, lo_bits: float, hi_bits: float) -> list[int]:
    """n = floor(2^(lo + (hi - lo) u)) for COUNT draws of the stream of SEED."""
    return [max(1, int(2 ** (lo_bits + (hi_bits - lo_bits) * u))) for u in uniforms(seed, count)]


def poisson_gaps(seed: int, count: int, rate: float) -> list[float]:
    """Inter-arrival times of a Poisson stream of RATE per second."""
    return [-math.log(1.0 - (u + 2.0**-65)) / rate for u in uniforms(seed, count)]


def batch_shapes(count: int = 4096, seed: int = BATCH_SEED) -> tuple[list[int], list[int]]:
    """Operand sizes of the binary-splitting level: COUNT products with both
    operands log-uniform in [2^8, 2^16] limbs.  Product i's operands use the
    streams SEED + 2i + 1 and SEED + 2i + 2."""
    sizes = log_uniform_sizes(seed, 2 * count, 8, 16)
    return sizes[0::2], sizes[1::2]


def serve_shapes(count: int, seed: int = SERVE_SEED) -> tuple[list[int], list[int]]:
    """Operand sizes of the serving stream: log-uniform in [2^10, 2^24] limbs."""
    sizes = log_uniform_sizes(seed, 2 * count, 10, 24)
    return sizes[0::2], sizes[1::2]
