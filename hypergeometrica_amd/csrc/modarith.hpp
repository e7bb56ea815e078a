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

// low 32 bits of a 24 x 24-bit product on the full-rate multiplier
// (v_mul_u32_u24); the compiler would otherwise pick v_mul_lo_u32.
HGX_HD u32 mul_u24(u32 a, u32 b) {
#if defined(__HIP_DEVICE_COMPILE__)
  u32 r;
  asm("v_mul_u32_u24 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
  return r;
#else
  return (a & 0xFFFFFFu) * (b & 0xFFFFFFu);
#endif
}

// high 32 bits of a 32 x 32-bit product (v_mul_hi_u32)
HGX_HD u32 mulhi32(u32 a, u32 b) { return (u32)(((u64)a * b) >> 32); }

// 32 x 32 + 64 -> 65-bit multiply-add: the low 64 bits, and the carry out in
// CARRY.  v_mad_u64_u32 writes that carry to an SGPR lane mask, a form the
// compiler does not select from C (it adds the carry with a compare and two
// selects, or builds a zero-extended addend with a move); hence the asm,
// whose only inputs and outputs are registers.
HGX_HD u64 mad64_carry(u32 a, u32 b, u64 c, u32 &carry) {
#if defined(__HIP_DEVICE_COMPILE__)
  u64 d;
  unsigned long long m;
  asm("v_mad_u64_u32 %0, %1, %2, %3, %4" : "=v"(d), "=s"(m) : "v"(a), "v"(b), "v"(c));
  asm("v_cndmask_b32 %0, 0, 1, %1" : "=v"(carry) : "s"(m));
  return d;
#else
  const unsigned __int128 s = (unsigned __int128)((u64)a * b) + c;
  carry = (u32)(s >> 64);
  return (u64)s;
#endif
}

// One modulus as the kernels use it: p, its negation 2^64 - p (kept as its
// own value so that r - p compiles to a single 64-bit add, with no carry
// through an SGPR), and the Proth factor c of p = c*2^55 + 1.
struct Mod {
  u64 p;
  u64 np;
  u32 c;
};

// r mod p for r in [0, 2p): t = r - p lies in (-p, p), and p < 2^63, so the
// sign of t (as a 64-bit two's-complement value) says whether r was already
// reduced; a bitwise select (v_bfi_b32) picks r or t.
HGX_HD u64 csub(u64 r, const Mod &M) {
  const u64 t = r + M.np;
  const u32 m = (u32)((int32_t)(u32)(t >> 32) >> 31);
  const u64 mm = ((u64)m << 32) | m;
  return (r & mm) | (t & ~mm);
}

// a*b*2^-64 mod p, canonical, for a < 2^64 and b < p (so a*b < p*2^64).
HGX_HD u64 mont_mul(u64 a, u64 b, const Mod &M) {
  const u32 c = M.c;
  u64 t00, hi;
  {
    const u32 a0 = (u32)a, a1 = (u32)(a >> 32), b0 = (u32)b, b1 = (u32)(b >> 32);
    t00 = (u64)a0 * b0;
    const u64 t10 = (u64)a1 * b0 + (t00 >> 32);
    const u64 t01 = (u64)a0 * b1 + (u32)t10;
    hi = (u64)a1 * b1 + (t10 >> 32) + (t01 >> 32);
    t00 = (u64)(u32)t00 | (t01 << 32);
  }
  const u32 lo0 = (u32)t00, lo1 = (u32)(t00 >> 32);
  // m = lo * (c*2^55 - 1) mod 2^64 = ((lo*c) << 55) - lo: the low word is
  // -lo0, the high word (lo0*c mod 2^9) << 23 - lo1 - borrow
  const u32 mlo = 0u - lo0;
  const u32 mhi = (mul_u24(lo0 & 511u, c) << 23) - lo1 - (lo0 != 0u ? 1u : 0u);
  // m*c as a 72-bit number mc_hi*2^32 + (u32)mc_lo
  const u64 mc_lo = (u64)mlo * c;
  const u64 mc_hi = (u64)mhi * c + (mc_lo >> 32);
  // lo64(m*p) = ((m*c mod 2^9) << 55) + m; its carry shows in the high word
  const u32 x = mhi + (((u32)mc_lo & 511u) << 23);
  const u32 carry2 = (x < mhi) ? 1u : 0u;
  // T_lo + lo64(m*p) == 0 mod 2^64, carrying 1 iff T_lo != 0
  const u32 nz = ((lo0 | lo1) != 0u) ? 1u : 0u;
  // t = hi + floor(m*c / 2^9) + carry2 + nz; the small terms join the low
  // 23 bits of floor(m*c / 2^9) without a carry out of 32 bits
  const u32 qlo = ((u32)mc_lo >> 9) + carry2 + nz;
  const u64 t = hi + (mc_hi << 23) + (u64)qlo;
  return csub(t, M);
}

// base^e in Montgomery form (BASE and ONE = 2^64 mod p in Montgomery form,
// canonical), by squaring and multiplying
HGX_HD u64 mont_pow(u64 base, u64 e, u64 one, const Mod &M) {
  u64 r = one;
  while (e) {
    if (e & 1) r = mont_mul(r, base, M);
    base = mont_mul(base, base, M);
    e >>= 1;
  }
  return r;
}

// Entry E of a twiddle table: the value SCALE * R^E mod p and its Shoup
// companion floor(value * 2^64 / p), from RM, SM and ONE (R, SCALE and 1 in
// Montgomery form, x 2^64 mod p) and PINV = p^-1 mod 2^64.  With x~ the
// Montgomery form of the value, value = REDC(x~), and since value * 2^64 =
// companion * p + x~, the companion is the exact quotient -x~ p^-1 mod 2^64:
// no division.
HGX_HD void pow_table_entry(u64 e, u64 rm, u64 sm, u64 one, u64 pinv, const Mod &M, u64 &value, u64 &companion) {
  const u64 x = mont_mul(sm, mont_pow(rm, e, one, M), M);
  value = mont_mul(x, 1, M);
  companion = (0 - x) * pinv;
}

// a*w mod p, canonical, by Shoup's method with the precomputed companion
// wq = floor(w * 2^64 / p) (calc-b-inv and m*/fast2,
// src/modular-arithmetic.lisp:206-231): q = hi64(a*wq) is exact, so
// r = a*w - q*p lies in [0, 2p) for any a < 2^64 and one subtraction makes it
// canonical.  q*p mod 2^64 = q + ((q*c mod 2^9) << 55) for p = c*2^55 + 1.
HGX_HD u64 shoup_mul(u64 a, u64 w, u64 wq, const Mod &M) {
  const u32 a0 = (u32)a, a1 = (u32)(a >> 32), w0 = (u32)w, w1 = (u32)(w >> 32);
  const u32 q0 = (u32)wq, q1 = (u32)(wq >> 32);
  // q = hi64(a*wq) = a1 q1 + floor((a1 q0 + a0 q1 + hi32(a0 q0)) / 2^32):
  // y = a0 q1 + hi32(a0 q0) stays below 2^64, a1 q0 + y may carry into bit
  // 64, and that carry is bit 32 of the addend of a1 q1
  const u64 y = (u64)a0 * q1 + mulhi32(a0, q0);
  u32 cx;
  const u64 x = mad64_carry(a1, q0, y, cx);
  const u64 q = (u64)a1 * q1 + ((x >> 32) | ((u64)cx << 32));
  // r = a*w - q*p mod 2^64 with q*p = q + ((q*c mod 2^9) << 55).  The terms
  // that reach only the high word are summed in 32 bits; (q mod 2^24) *
  // (2^24 - c) << 23 is -(q*c mod 2^9) << 23 mod 2^32, so q needs no mask.
  const u32 h = a0 * w1 + a1 * w0 + (mul_u24((u32)q, (1u << 24) - M.c) << 23);
  const u64 r = (u64)a0 * w0 + (((u64)h << 32) - q);
  return csub(r, M);
}

HGX_HD u64 mod_add(u64 a, u64 b, const Mod &M) { return csub(a + b, M); }

// a - b mod p for canonical a, b: a - b lies in (-p, p); add p back when
// it is negative.
HGX_HD u64 mod_sub(u64 a, u64 b, const Mod &M) {
  const u64 d = a - b;
  const u32 m = (u32)((int32_t)(u32)(d >> 32) >> 31);
  return d + ((((u64)m << 32) | m) & M.p);
}

// x mod p for any 64-bit x (x < 4p because p > 2^62).
HGX_HD u64 mod_reduce64(u64 x, const Mod &M) {
  const u64 p2 = M.p << 1;
  x = (x >= p2) ? x - p2 : x;
  return (x >= M.p) ? x - M.p : x;
}
