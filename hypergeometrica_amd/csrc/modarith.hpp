// modarith.hpp -- 63-bit modular arithmetic for the three Proth primes of
// Hypergeometrica's default scheme, built from 32-bit integer VALU ops.
//
// The moduli are p = c * 2^55 + 1 with c in {197, 174, 131}
// (default-moduli, src/moduli.lisp:257-264).  All three exceed 2^62, so
// 2p < 2^64 < 4p: residues are kept canonical in [0, p) between butterflies
// (as m+ / m- do, src/modular-arithmetic.lisp:25-48) and only transient
// values reach [0, 2p).
//
// Multiplication is Montgomery's REDC with R = 2^64.  For p = c*2^55 + 1,
// p^-1 = 1 - c*2^55 (mod 2^64), so m = -T_lo * p^-1 = ((T_lo*c) << 55) - T_lo
// needs only the low 9 bits of T_lo*c, and m*p = m + (m*c) << 55 needs only
// the 72-bit product m*c.  Twiddles and constants are stored in Montgomery
// form (x*R mod p) so that mont_mul(a, w~) = a*w mod p for plain a.
// This replaces the reference's Barrett m*/fast (:167-202) and Shoup
// m*/fast2 (:206-231): the products are the same residues.
#pragma once
#include <stdint.h>
#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define HGX_HD __host__ __device__ __forceinline__
#else
#define HGX_HD static inline
#endif

typedef uint64_t u64;
typedef uint32_t u32;

// 64 x 64 -> 128 product from four 32 x 32 -> 64 multiply-adds
// (v_mad_u64_u32 on gfx950).
HGX_HD void mul64x64(u64 a, u64 b, u64 &lo, u64 &hi) {
  u32 a0 = (u32)a, a1 = (u32)(a >> 32), b0 = (u32)b, b1 = (u32)(b >> 32);
  u64 t00 = (u64)a0 * b0;
  u64 t10 = (u64)a1 * b0 + (t00 >> 32);
  u64 t01 = (u64)a0 * b1 + (u32)t10;
  hi = (u64)a1 * b1 + (t10 >> 32) + (t01 >> 32);
  lo = (u64)(u32)t00 | (t01 << 32);
}

// a*b*2^-64 mod p, canonical, for a < 2^64 and b < p (so a*b < p*2^64).
HGX_HD u64 mont_mul(u64 a, u64 b, u64 p, u32 c) {
  u64 lo, hi;
  mul64x64(a, b, lo, hi);
  // m = lo * (c*2^55 - 1) mod 2^64
  u64 m = ((u64)((u32)lo * c) << 55) - lo;
  // m*c as a 72-bit number: mc_hi * 2^32 + mc_lo32
  u64 mc_lo = (u64)(u32)m * c;
  u64 mc_hi = (u64)(u32)(m >> 32) * c + (mc_lo >> 32);
  // floor(m*c / 2^9), and (m*c mod 2^9) * 2^55 + m with its carry
  u64 q = (mc_hi << 23) | ((u32)mc_lo >> 9);
  u64 low = ((u64)((u32)mc_lo & 511u) << 55) + m;
  u64 carry = (low < m) ? 1 : 0;
  // t = (T + m*p) / 2^64;  T_lo + lo64(m*p) == 0 mod 2^64, carrying 1 iff T_lo != 0
  u64 t = hi + q + carry + (lo != 0 ? 1 : 0);
  return (t >= p) ? t - p : t;
}

HGX_HD u64 mod_add(u64 a, u64 b, u64 p) {
  u64 s = a + b;
  return (s >= p) ? s - p : s;
}

HGX_HD u64 mod_sub(u64 a, u64 b, u64 p) {
  u64 d = a - b;
  return (a < b) ? d + p : d;
}

// x mod p for any 64-bit x (x < 4p because p > 2^62).
HGX_HD u64 mod_reduce64(u64 x, u64 p) {
  u64 p2 = p << 1;
  x = (x >= p2) ? x - p2 : x;
  return (x >= p) ? x - p : x;
}
