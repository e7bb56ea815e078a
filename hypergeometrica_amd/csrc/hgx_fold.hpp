// hgx_fold.hpp -- the leaf folds of binary splitting (hgx_series.cpp) as
// host-and-device code: one thread folds one leaf, an interval of at most 32
// terms, into its partial P, Q, B, R.  The fold is the one the engine runs on
// the host: the base case at LO (binary-split-base-case=1,
// src/binary-splitting.lisp:100-110) combined with the terms LO+1, ... one at
// a time on the right, COMBINE (:112-121) with a one-term partial:
//   R <- R b_k q_k + BP p_k a_k,  BP <- BP b_k p_k,
//   P <- P p_k,  Q <- Q q_k,  B <- B b_k      (BP = B P of the terms so far).
//
// Term values are signed integers of at most kFoldTV limbs, kept at full
// width (zero limbs above their length) so that every loop over them has a
// fixed trip count and they stay in registers.  The growing values of a leaf
// live in strided buffers -- limb i of leaf t at buf[b][t + i * stride] --
// so that the threads of a wave, which fold consecutive leaves, read and
// write consecutive words.  Each value has two buffers and moves to the
// other one at every term; a leaf starts in the pair member that makes its
// last term land in the even one (kFoldP, kFoldQ, kFoldB, kFoldR).  Every
// write is checked against the buffer's capacity: a leaf that would exceed
// one reports kFoldFailed instead, and the engine folds on the host.
#pragma once
#include <math.h>

#include "modarith.hpp"

constexpr int kFoldTV = 4;             // limbs of a term value
constexpr int kFoldTV2 = 2 * kFoldTV;  // of a product of two
constexpr int kFoldBufs = 11;
enum { kFoldP = 0, kFoldQ = 2, kFoldB = 4, kFoldBP = 6, kFoldR = 8, kFoldY = 10 };
constexpr int kFoldFailed = -2147483647 - 1;  // a leaf's first code when its fold failed

// a(k), b(k), p(k), q(k) (F[0..3]): polynomials in k, constant term first,
// or V0 at k = 0 when AT_ZERO is set
struct FoldPoly {
  int deg;
  int at_zero;
  long long v0;
  long long c[8];
};
struct FoldSeries {
  FoldPoly f[4];
};

// buffer b of leaf t starts at buf[b] + t and holds cap[b] limbs (the two
// buffers of a pair have the same capacity)
struct FoldSlots {
  u64 *buf[kFoldBufs];
  u32 cap[kFoldBufs];
  u64 stride;
};

template <int N>
struct FoldSmall {
  int s;
  u64 m[N];
};

template <int N>
HGX_HD int fold_len(const FoldSmall<N> &v) {
  int n = 0;
#pragma unroll
  for (int i = 0; i < N; i++)
    if (v.m[i]) n = i + 1;
  return n;
}

template <int N>
HGX_HD void fold_zero(FoldSmall<N> &v) {
  v.s = 0;
#pragma unroll
  for (int i = 0; i < N; i++) v.m[i] = 0;
}

// v = v * k + c for k >= 0 (one Horner step); false when v needs more than
// N limbs
template <int N>
HGX_HD bool fold_horner(FoldSmall<N> &v, u64 k, long long c) {
  u64 carry = 0;
#pragma unroll
  for (int i = 0; i < N; i++) {
    u64 lo, hi;
    mul64x64(v.m[i], k, lo, hi);
    lo += carry;
    hi += (lo < carry) ? 1 : 0;
    v.m[i] = lo;
    carry = hi;
  }
  if (carry) return false;
  if (fold_len(v) == 0) v.s = 0;
  if (c == 0) return true;
  const int cs = c < 0 ? -1 : 1;
  const u64 cm = c < 0 ? (u64)0 - (u64)c : (u64)c;
  if (v.s == 0) {  // every limb is zero
    v.s = cs;
    v.m[0] = cm;
    return true;
  }
  if (v.s == cs) {  // magnitudes add
    u64 cy = cm;
#pragma unroll
    for (int i = 0; i < N; i++) {
      v.m[i] += cy;
      cy = (v.m[i] < cy) ? 1 : 0;
    }
    return cy == 0;
  }
  bool one = true;  // |v| fits one limb
#pragma unroll
  for (int i = 1; i < N; i++) one = one && v.m[i] == 0;
  if (one && v.m[0] < cm) {  // |v| < |c|: the sign turns
    v.m[0] = cm - v.m[0];
    v.s = cs;
    return true;
  }
  u64 b = cm;
#pragma unroll
  for (int i = 0; i < N; i++) {
    const u64 old = v.m[i];
    v.m[i] = old - b;
    b = (old < b) ? 1 : 0;
  }
  if (fold_len(v) == 0) v.s = 0;
  return true;
}

template <int N>
HGX_HD bool fold_eval(const FoldPoly &f, u64 k, FoldSmall<N> &out) {
  fold_zero(out);
  if (k == 0 && f.at_zero) return fold_horner(out, 0, f.v0);
  for (int i = f.deg; i >= 0; i--)
    if (!fold_horner(out, k, f.c[i])) return false;
  return true;
}

// z = x * y of two term values.  Rows run over all of x's limbs (the zero
// ones too): a row that stopped at x's length would store its carry at a
// computed index, which sends z to scratch memory.
HGX_HD void fold_small_mul(const FoldSmall<kFoldTV> &x, const FoldSmall<kFoldTV> &y, FoldSmall<kFoldTV2> &z) {
  fold_zero(z);
  if (!x.s || !y.s) return;
  const int yn = fold_len(y);
#pragma unroll
  for (int j = 0; j < kFoldTV; j++) {
    if (j < yn) {
      u64 carry = 0;
#pragma unroll
      for (int i = 0; i < kFoldTV; i++) {
        u64 lo, hi;
        mul64x64(x.m[i], y.m[j], lo, hi);
        lo += carry;
        hi += (lo < carry) ? 1 : 0;
        lo += z.m[i + j];
        hi += (lo < z.m[i + j]) ? 1 : 0;
        z.m[i + j] = lo;
        carry = hi;
      }
      z.m[kFoldTV + j] = carry;
    }
  }
  z.s = x.s * y.s;
}

// a signed value of a leaf: limb i at p[i * stride], N limbs (trimmed)
struct FoldBig {
  u64 *p;
  int s;
  u32 n;
};

HGX_HD void fold_swap(FoldBig &a, FoldBig &b) {
  const FoldBig t = a;
  a = b;
  b = t;
}

// O = X * V (O's buffer is not X's); false when O needs more than CAP limbs
template <int N>
HGX_HD bool fold_mul(const FoldBig &x, const FoldSmall<N> &v, FoldBig &o, u32 cap, u64 st) {
  o.s = 0;
  o.n = 0;
  if (!x.s || !v.s) return true;
  const int vn = fold_len(v);
  const u32 n = x.n + (u32)vn;
  if (n > cap) return false;
  u64 carry = 0;
  // (unrolled: a thread issues several strided loads ahead of its carry chain)
#pragma unroll 4
  for (u32 i = 0; i < x.n; i++) {  // the first row writes
    u64 lo, hi;
    mul64x64(x.p[i * st], v.m[0], lo, hi);
    lo += carry;
    hi += (lo < carry) ? 1 : 0;
    o.p[i * st] = lo;
    carry = hi;
  }
  o.p[x.n * st] = carry;
#pragma unroll
  for (int j = 1; j < N; j++) {  // later rows accumulate
    if (j < vn) {
      const u64 vj = v.m[j];
      carry = 0;
#pragma unroll 4
      for (u32 i = 0; i < x.n; i++) {
        u64 lo, hi;
        mul64x64(x.p[i * st], vj, lo, hi);
        lo += carry;
        hi += (lo < carry) ? 1 : 0;
        const u64 r = o.p[(i + j) * st];
        lo += r;
        hi += (lo < r) ? 1 : 0;
        o.p[(i + j) * st] = lo;
        carry = hi;
      }
      o.p[(x.n + j) * st] = carry;
    }
  }
  o.n = n;
  while (o.n > 0 && o.p[(o.n - 1) * st] == 0) o.n--;
  o.s = o.n ? x.s * v.s : 0;
  return true;
}

// O = V, all N limbs written (fixed indices); false when CAP < N
template <int N>
HGX_HD bool fold_set(FoldBig &o, const FoldSmall<N> &v, u32 cap, u64 st) {
  if ((u32)N > cap) return false;
  const int vn = fold_len(v);
#pragma unroll
  for (int i = 0; i < N; i++) o.p[i * st] = v.m[i];
  o.n = (u32)vn;
  o.s = vn ? v.s : 0;
  return true;
}

// X += Y for signed X, Y (mpz-+), in X's buffer of CAP limbs; false when
// the sum may need more
HGX_HD bool fold_add(FoldBig &x, const FoldBig &y, u32 cap, u64 st) {
  if (!y.s) return true;
  if (!x.s) {
    if (y.n > cap) return false;
    for (u32 i = 0; i < y.n; i++) x.p[i * st] = y.p[i * st];
    x.n = y.n;
    x.s = y.s;
    return true;
  }
  if (x.s == y.s) {  // |x| += |y|
    const u32 n = x.n > y.n ? x.n : y.n;
    if (n + 1 > cap) return false;
    u64 c = 0;
    for (u32 i = 0; i < n; i++) {
      const u64 a = i < x.n ? x.p[i * st] : 0, b = i < y.n ? y.p[i * st] : 0;
      const u64 s = a + b;
      const u64 c1 = (s < a) ? 1 : 0;
      const u64 s2 = s + c;
      const u64 c2 = (s2 < s) ? 1 : 0;
      x.p[i * st] = s2;
      c = c1 + c2;
    }
    x.n = n;
    if (c) {
      x.p[n * st] = c;
      x.n = n + 1;
    }
    return true;
  }
  // opposite signs: the smaller magnitude comes off the larger
  int cmp = 0;
  if (x.n != y.n) {
    cmp = x.n < y.n ? -1 : 1;
  } else {
    for (u32 i = x.n; i-- > 0;) {
      const u64 a = x.p[i * st], b = y.p[i * st];
      if (a != b) {
        cmp = a < b ? -1 : 1;
        break;
      }
    }
  }
  if (cmp == 0) {
    x.s = 0;
    x.n = 0;
    return true;
  }
  const bool x_big = cmp > 0;
  const u32 n = x_big ? x.n : y.n;
  if (n > cap) return false;
  u64 borrow = 0;
  for (u32 i = 0; i < n; i++) {
    const u64 xv = i < x.n ? x.p[i * st] : 0, yv = i < y.n ? y.p[i * st] : 0;
    const u64 a = x_big ? xv : yv, b = x_big ? yv : xv;
    const u64 d = a - b;
    const u64 b1 = (a < b) ? 1 : 0;
    const u64 d2 = d - borrow;
    const u64 b2 = (d < borrow) ? 1 : 0;
    x.p[i * st] = d2;
    borrow = b1 + b2;
  }
  x.n = n;
  while (x.n > 0 && x.p[(x.n - 1) * st] == 0) x.n--;
  x.s = x.n ? (x_big ? x.s : y.s) : 0;
  return true;
}

// The partial of [LO, UP) for leaf T: CODE[w] = sign * limbs of P, Q, B, R
// (w = 0..3), whose limbs end in buffers kFoldP, kFoldQ, kFoldB, kFoldR.
// False when a term value or a value of the leaf outgrows its capacity
// (the buffers then hold garbage).
HGX_HD bool fold_leaf(const FoldSeries &S, u64 lo, u64 up, u64 t, const FoldSlots &F, int *code) {
  const u64 st = F.stride;
  // the first buffer of each pair (selected, not indexed: the slots stay
  // where the kernel's arguments are)
  const bool odd = ((up - lo - 1) & 1) != 0;
  FoldBig P{(odd ? F.buf[kFoldP + 1] : F.buf[kFoldP]) + t, 0, 0};
  FoldBig P2{(odd ? F.buf[kFoldP] : F.buf[kFoldP + 1]) + t, 0, 0};
  FoldBig Q{(odd ? F.buf[kFoldQ + 1] : F.buf[kFoldQ]) + t, 0, 0};
  FoldBig Q2{(odd ? F.buf[kFoldQ] : F.buf[kFoldQ + 1]) + t, 0, 0};
  FoldBig B{(odd ? F.buf[kFoldB + 1] : F.buf[kFoldB]) + t, 0, 0};
  FoldBig B2{(odd ? F.buf[kFoldB] : F.buf[kFoldB + 1]) + t, 0, 0};
  FoldBig BP{(odd ? F.buf[kFoldBP + 1] : F.buf[kFoldBP]) + t, 0, 0};
  FoldBig BP2{(odd ? F.buf[kFoldBP] : F.buf[kFoldBP + 1]) + t, 0, 0};
  FoldBig R{(odd ? F.buf[kFoldR + 1] : F.buf[kFoldR]) + t, 0, 0};
  FoldBig R2{(odd ? F.buf[kFoldR] : F.buf[kFoldR + 1]) + t, 0, 0};
  FoldBig Y{F.buf[kFoldY] + t, 0, 0};
  const u32 cP = F.cap[kFoldP], cQ = F.cap[kFoldQ], cB = F.cap[kFoldB], cBP = F.cap[kFoldBP], cR = F.cap[kFoldR],
            cY = F.cap[kFoldY];
  FoldSmall<kFoldTV> pk, qk, bk, ak;
  FoldSmall<kFoldTV2> t1, t2;
  if (!fold_eval(S.f[2], lo, pk) || !fold_eval(S.f[3], lo, qk) || !fold_eval(S.f[1], lo, bk) ||
      !fold_eval(S.f[0], lo, ak))
    return false;
  fold_small_mul(pk, ak, t1);
  fold_small_mul(bk, pk, t2);
  if (!fold_set(P, pk, cP, st) || !fold_set(Q, qk, cQ, st) || !fold_set(B, bk, cB, st) || !fold_set(R, t1, cR, st) ||
      !fold_set(BP, t2, cBP, st))
    return false;
  for (u64 k = lo + 1; k < up; k++) {
    if (!fold_eval(S.f[2], k, pk) || !fold_eval(S.f[3], k, qk) || !fold_eval(S.f[1], k, bk) ||
        !fold_eval(S.f[0], k, ak))
      return false;
    fold_small_mul(bk, qk, t1);
    fold_small_mul(pk, ak, t2);
    if (!fold_mul(R, t1, R2, cR, st) || !fold_mul(BP, t2, Y, cY, st) || !fold_add(R2, Y, cR, st)) return false;
    fold_small_mul(bk, pk, t1);
    if (!fold_mul(BP, t1, BP2, cBP, st) || !fold_mul(P, pk, P2, cP, st) || !fold_mul(Q, qk, Q2, cQ, st) ||
        !fold_mul(B, bk, B2, cB, st))
      return false;
    fold_swap(R, R2);
    fold_swap(BP, BP2);
    fold_swap(P, P2);
    fold_swap(Q, Q2);
    fold_swap(B, B2);
  }
  code[0] = P.s * (int)P.n;
  code[1] = Q.s * (int)Q.n;
  code[2] = B.s * (int)B.n;
  code[3] = R.s * (int)R.n;
  return true;
}

// Leaf T's P, Q, B, R one after the other at DST (|CODE[w]| limbs each)
HGX_HD void fold_pack_leaf(const FoldSlots &F, u64 t, const int *code, u64 *dst) {
  const int bufs[4] = {kFoldP, kFoldQ, kFoldB, kFoldR};
#pragma unroll
  for (int w = 0; w < 4; w++) {
    const u32 n = (u32)(code[w] < 0 ? -code[w] : code[w]);
    const u64 *src = F.buf[bufs[w]] + t;
    for (u32 i = 0; i < n; i++) dst[i] = src[i * F.stride];
    dst += n;
  }
}

// ---- host side: capacities ----
// An upper bound on the bits of |f(k)| for LO <= k < UP (0 when f is zero
// there): sum |c_i| (UP-1)^i, or |v0| at k = 0, with a bit to spare for the
// rounding of the long double sum.
inline int fold_term_bits(const FoldPoly &f, u64 lo, u64 up) {
  long double m = 0, kk = 1;
  const long double k = (long double)(up - 1);
  for (int i = 0; i <= f.deg; i++) {
    const long double c = f.c[i] < 0 ? -(long double)f.c[i] : (long double)f.c[i];
    m += c * kk;
    kk *= k;
  }
  if (lo == 0 && f.at_zero) {
    const long double v = f.v0 < 0 ? -(long double)f.v0 : (long double)f.v0;
    if (v > m) m = v;
  }
  if (m < 1) return 0;
  return (int)floorl(log2l(m)) + 2;
}

// Capacities of the buffers for leaves of at most TERMS terms over
// [LO, UP): false when a term value may exceed kFoldTV limbs (the engine
// then folds on the host).  LIMBS receives an upper bound on a leaf's
// P, Q, B, R limbs together.
inline bool fold_caps(const FoldSeries &S, u64 lo, u64 up, u64 terms, u32 cap[kFoldBufs], u64 &limbs) {
  int bits[4];
  for (int i = 0; i < 4; i++) {
    bits[i] = fold_term_bits(S.f[i], lo, up);
    if (bits[i] > 64 * kFoldTV) return false;
  }
  const u64 ba = (u64)bits[0], bb = (u64)bits[1], bp = (u64)bits[2], bq = (u64)bits[3];
  const u64 T = terms;
  // bits after a leaf's last term: P, Q, B and BP are products of T term
  // values each, R a sum of T products of 2T - 1 term values and one a
  // (the sum adds 5 bits for T <= 32)
  auto limbs_of = [](u64 b) { return (b + 63) / 64; };
  const u64 lp = limbs_of(T * bp), lq = limbs_of(T * bq), lb = limbs_of(T * bb), lbp = limbs_of(T * (bb + bp));
  const u64 lr = limbs_of(T * (bb + (bp > bq ? bp : bq)) + ba + 6);
  // a product is written at its operands' full length before it is
  // trimmed, and a sum may need one more limb
  const u64 pad = kFoldTV2 + 1;
  const u64 c[kFoldBufs] = {lp + pad, lp + pad, lq + pad, lq + pad, lb + pad, lb + pad,
                            lbp + pad, lbp + pad, lr + pad, lr + pad, lr + pad};
  for (int b = 0; b < kFoldBufs; b++) {
    if (c[b] >= ((u64)1 << 30)) return false;
    cap[b] = (u32)c[b];
  }
  limbs = lp + lq + lb + lr;
  return true;
}
