//
// hgx_oracle.c -- CPU restatement of Hypergeometrica's exact NTT multiply.
//
// TEST INFRASTRUCTURE ONLY: this file is the oracle the MI355X library is
// checked against.  It is loaded by tests/, by __graft_entry__.smoke() and by
// the cpu_baseline leg of bench.py, and by nothing on the product path.
//
// The reference is Common Lisp (SBCL).  Each function below follows one
// reference function, cited as file:line, step for step: the same loop
// order, the same Barrett (m*/fast) and Shoup (m*/fast2) reductions, the same
// direct CRT and the same carry ripple.  Where SBCL bignums are used by the
// reference, fixed-width multiword arithmetic on uint64_t words is used here;
// the values are the same.
//
// Pinning (see tests/test_oracle_*.py):
//   - moduli, generators, roots, Barrett inverses: closed forms (sympy) and
//     the "Moduli:" / "Convolution bits:" lines of the reference's logs;
//   - m+, m-, m*, m*/fast, m*/fast2: (a*b) mod m with Python integers;
//   - ntt-forward / ntt-reverse: direct DFT over Z/pZ (bit-reversed), the
//     same check as tests/ntt.lisp:133-155, and inverse(forward) = id;
//   - CRT: exact reconstruction (unique below M) and a Garner round trip;
//   - mpz-*/ntt, mpz-square: the exact integer product (Python int, GMP,
//     schoolbook), as tests/mpz.lisp:32-47 does against CL `*`.
// Planted-mistake switches (hgxo_fault) let the tests show each pin fails
// when the restatement is wrong.
#include "hgx_oracle.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

int hgxo_fault = HGXO_FAULT_NONE;

/* ------------------------------------------------------------------ */
/* L0 / L1: modular primitives                                          */
/* ------------------------------------------------------------------ */

/* expt-mod, src/modular-arithmetic.lisp:118-131 (square and multiply). */
uint64_t hgxo_expt_mod(uint64_t a, uint64_t n, uint64_t m) {
  uint64_t result = 1 % m;
  a %= m;
  for (;;) {
    if (n & 1) result = (uint64_t)(((u128)result * a) % m);
    n >>= 1;
    if (n == 0) return result;
    a = (uint64_t)(((u128)a * a) % m);
  }
}

/* inv-mod, src/modular-arithmetic.lisp:90-101 (extended Euclid). */
uint64_t hgxo_inv_mod(uint64_t x, uint64_t m) {
  __int128 b = m, xx = x % m, a = 0, u = 1;
  while (xx != 0) {
    __int128 q = b / xx, r = b % xx;
    __int128 na = u, nu = a - u * q;
    b = xx;
    xx = r;
    a = na;
    u = nu;
  }
  if (b != 1) return 0; /* not invertible: the reference signals an error */
  a %= (__int128)m;
  if (a < 0) a += m;
  return (uint64_t)a;
}

/* m-, src/modular-arithmetic.lisp:25-35 */
uint64_t hgxo_m_sub(uint64_t a, uint64_t b, uint64_t m) {
  if (a < b) return (m - b) + a;
  return a - b;
}

/* m+, src/modular-arithmetic.lisp:37-48: a + b = a - (m - b) */
uint64_t hgxo_m_add(uint64_t a, uint64_t b, uint64_t m) {
  if (b == 0) return a;
  return hgxo_m_sub(a, m - b, m);
}

/* m*, src/modular-arithmetic.lisp:81-88 (mul128 then div128 remainder). */
uint64_t hgxo_m_mul(uint64_t a, uint64_t b, uint64_t m) {
  return (uint64_t)(((u128)a * b) % m);
}

// mod128/fast, src/modular-arithmetic.lisp:167-190.  Barrett reduction of
// lo + hi*2^64 (< 2^126) with m-inv = floor(2^126/m): q = hi64(x>>62 * minv),
// r = x - q*m - 2m, then up to two +m corrections selected by the sign word
// (ub64/2 of the high word is 0 or a mask covering m, src/digit.lisp:75-81).
uint64_t hgxo_mod128_fast(uint64_t lo, uint64_t hi, uint64_t m, uint64_t minv) {
  const int lg = 62; /* +lg-modulus+, src/modular-arithmetic.lisp:12-13 */
  uint64_t a1 = (lo >> lg) | ((hi & ((1ULL << lg) - 1)) << (64 - lg));
  uint64_t q = (uint64_t)(((u128)a1 * minv) >> 64);
  u128 x = ((u128)hi << 64) | lo;
  x -= (u128)q * m;
  x -= (u128)(2 * m);
  uint64_t xhi = (uint64_t)(x >> 64);
  x += (u128)(m & (xhi >> 1));
  xhi = (uint64_t)(x >> 64);
  uint64_t r = (uint64_t)x;
  if (hgxo_fault == HGXO_FAULT_BARRETT_FIX) return r;
  return r + (m & xhi);
}

// m*/fast, src/modular-arithmetic.lisp:192-202
uint64_t hgxo_m_mul_fast(uint64_t a, uint64_t b, uint64_t m, uint64_t minv) {
  u128 p = (u128)a * b;
  return hgxo_mod128_fast((uint64_t)p, (uint64_t)(p >> 64), m, minv);
}

/* calc-b-inv, src/modular-arithmetic.lisp:206-209: floor(b*2^64 / m). */
uint64_t hgxo_calc_b_inv(uint64_t b, uint64_t m) {
  return (uint64_t)((((u128)b) << 64) / m);
}

// m*/fast2, src/modular-arithmetic.lisp:211-231 (Shoup): r = a*b - q*m mod
// 2^64 with q = hi64(a*b_inv); one conditional subtraction.
uint64_t hgxo_m_mul_fast2(uint64_t b, uint64_t binv, uint64_t a, uint64_t m) {
  uint64_t q = (uint64_t)(((u128)a * binv) >> 64);
  uint64_t r = a * b - q * m;
  if (r >= m) r -= m;
  return r;
}

void hgxo_m_mul_fast_vec(const uint64_t *a, const uint64_t *b, uint64_t *out,
                         size_t n, uint64_t m, uint64_t minv) {
  for (size_t i = 0; i < n; i++) out[i] = hgxo_m_mul_fast(a[i], b[i], m, minv);
}

void hgxo_m_mul_fast2_vec(const uint64_t *b, const uint64_t *a, uint64_t *out,
                          size_t n, uint64_t m) {
  for (size_t i = 0; i < n; i++)
    out[i] = hgxo_m_mul_fast2(b[i], hgxo_calc_b_inv(b[i], m), a[i], m);
}

/* ------------------------------------------------------------------ */
/* Moduli and scheme tables                                            */
/* ------------------------------------------------------------------ */

// primep, src/moduli.lisp:23-45.  The reference draws 10 random Miller-Rabin
// witnesses; this restatement uses the first twelve primes as witnesses,
// which is deterministic for every n < 3.3e24 and therefore agrees with the
// reference whenever the reference's random test is right.
int hgxo_primep(uint64_t n) {
  static const uint64_t w[12] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
  if (n < 2) return 0;
  for (int i = 0; i < 12; i++) {
    if (n == w[i]) return 1;
    if (n % w[i] == 0) return 0;
  }
  uint64_t d = n - 1;
  int s = 0;
  while ((d & 1) == 0) { d >>= 1; s++; }
  for (int i = 0; i < 12; i++) {
    uint64_t x = hgxo_expt_mod(w[i], d, n);
    if (x == 1 || x == n - 1) continue;
    int liar = 0;
    for (int r = 1; r < s; r++) {
      x = (uint64_t)(((u128)x * x) % n);
      if (x == n - 1) { liar = 1; break; }
    }
    if (!liar) return 0;
  }
  return 1;
}

static int integer_length_u64(uint64_t x) {
  return x ? 64 - __builtin_clzll(x) : 0;
}

// default-moduli, src/moduli.lisp:257-264, with find-suitable-moduli
// (src/moduli.lisp:109-124) for transform length 2^55: primes p = i*2^55+1,
// i = 1, 2, ...  The reference collects 1000 of them and keeps the ones of
// exactly +modulus-bits+ = 63 bits, sorted descending.  Only i in [128, 256)
// gives 63-bit p, and 1000 primes cannot all lie below i = 256, so the kept
// set is exactly the primes with i in [128, 256); that range is what is
// enumerated here.
int hgxo_default_moduli(uint64_t *out, int max_out) {
  int count = 0;
  for (uint64_t i = 255; i >= 128; i--) {
    uint64_t p = (i << 55) + 1;
    if (integer_length_u64(p) != 63) continue;
    if (hgxo_primep(p)) {
      if (count < max_out) out[count] = p;
      count++;
    }
  }
  return count;
}

// factorization of m-1 (trial division) for generator-test-function,
// src/moduli.lisp:164-171.
static int prime_factors(uint64_t n, uint64_t *f, int maxf) {
  int k = 0;
  for (uint64_t d = 2; d * d <= n; d += (d == 2 ? 1 : 2)) {
    if (n % d == 0) {
      if (k < maxf) f[k++] = d;
      while (n % d == 0) n /= d;
    }
  }
  if (n > 1 && k < maxf) f[k++] = n;
  return k;
}

static uint64_t next_prime(uint64_t n) {
  /* next-prime, src/moduli.lisp:47-53 */
  if (n < 2) return 2;
  uint64_t i = (n & 1) ? n : n - 1;
  for (i += 2;; i += 2)
    if (hgxo_primep(i)) return i;
}

// find-finite-field-generator, src/moduli.lisp:173-180: smallest prime g with
// g^((m-1)/f) != 1 for every prime factor f of m-1.
uint64_t hgxo_find_generator(uint64_t m) {
  uint64_t f[64];
  int nf = prime_factors(m - 1, f, 64);
  for (uint64_t g = 2;; g = next_prime(g)) {
    int ok = 1;
    for (int i = 0; i < nf; i++)
      if (hgxo_expt_mod(g, (m - 1) / f[i], m) == 1) { ok = 0; break; }
    if (ok) return g;
  }
}

/* primitive-root-from-generator, src/moduli.lisp:182-187 */
uint64_t hgxo_primitive_root(uint64_t g, uint64_t n, uint64_t m) {
  return hgxo_expt_mod(g, (m - 1) / n, m);
}

static int count_trailing_zeroes(uint64_t n) { return __builtin_ctzll(n); }

static hgxo_scheme g_scheme;
static int g_scheme_ready = 0;

#define ROOT_IDX(s, n, m, k, i) \
  ((((size_t)(n) * (s)->num_moduli + (m)) * (size_t)((s)->max_transform_length + 1) + (k)) * 2 + (i))
#define INVLEN_IDX(s, k, m, i) ((((size_t)(k) * (s)->num_moduli) + (m)) * 2 + (i))

/* make-modular-scheme, src/moduli.lisp:298-353. */
static void make_modular_scheme(hgxo_scheme *s, const uint64_t *moduli, int nm) {
  memset(s, 0, sizeof(*s));
  s->num_moduli = nm;
  /* sorted descending (:300-301) */
  for (int i = 0; i < nm; i++) s->moduli[i] = moduli[i];
  for (int i = 0; i < nm; i++)
    for (int j = i + 1; j < nm; j++)
      if (s->moduli[j] > s->moduli[i]) {
        uint64_t t = s->moduli[i]; s->moduli[i] = s->moduli[j]; s->moduli[j] = t;
      }
  int max_trans = 64, lg_mod = 64;
  for (int i = 0; i < nm; i++) {
    int tz = count_trailing_zeroes(s->moduli[i] - 1); /* max-transform-length-bits :266-267 */
    if (tz < max_trans) max_trans = tz;
    int lg = integer_length_u64(s->moduli[i]) - 1;
    if (lg < lg_mod) lg_mod = lg;
  }
  s->max_transform_length = max_trans;
  size_t nroots = (size_t)(max_trans + 1) * nm * (max_trans + 1) * 2;
  s->roots = (uint64_t *)calloc(nroots, sizeof(uint64_t));
  s->inverse_roots = (uint64_t *)calloc(nroots, sizeof(uint64_t));
  s->inverse_lengths = (uint64_t *)calloc((size_t)(max_trans + 1) * nm * 2, sizeof(uint64_t));
  for (int i = 0; i < nm; i++) s->generators[i] = hgxo_find_generator(s->moduli[i]);
  /* roots (:315-332) */
  for (int power = 0; power <= max_trans; power++) {
    for (int i = 0; i < nm; i++) {
      uint64_t m = s->moduli[i];
      uint64_t w = hgxo_primitive_root(s->generators[i], 1ULL << power, m);
      uint64_t wi = hgxo_inv_mod(w, m);
      uint64_t wk = w, wik = wi;
      for (int k = 0; k <= max_trans; k++) {
        /* expt-mod/2^n w k m = w^(2^k) */
        if (k > 0) {
          wk = hgxo_m_mul(wk, wk, m);
          wik = hgxo_m_mul(wik, wik, m);
        }
        s->roots[ROOT_IDX(s, power, i, k, 0)] = wk;
        s->roots[ROOT_IDX(s, power, i, k, 1)] = hgxo_calc_b_inv(wk, m);
        s->inverse_roots[ROOT_IDX(s, power, i, k, 0)] = wik;
        s->inverse_roots[ROOT_IDX(s, power, i, k, 1)] = hgxo_calc_b_inv(wik, m);
      }
    }
  }
  /* reciprocal lengths (:334-340) */
  for (int i = 0; i < nm; i++)
    for (int k = 0; k <= max_trans; k++) {
      uint64_t m = s->moduli[i];
      uint64_t inv = hgxo_inv_mod((uint64_t)(((u128)1 << k) % m), m);
      s->inverse_lengths[INVLEN_IDX(s, k, i, 0)] = inv;
      s->inverse_lengths[INVLEN_IDX(s, k, i, 1)] = hgxo_calc_b_inv(inv, m);
    }
  /* Barrett inverses floor(2^(64+lg)/m) (:342-345) */
  for (int i = 0; i < nm; i++)
    s->inverses[i] = (uint64_t)((((u128)1) << (64 + lg_mod)) / s->moduli[i]);
}

/* **scheme**, src/ntt-multiply.lisp:8 */
const hgxo_scheme *hgxo_default_scheme(void) {
  if (!g_scheme_ready) {
    uint64_t mods[HGXO_MAX_MODULI];
    int n = hgxo_default_moduli(mods, HGXO_MAX_MODULI);
    if (n > HGXO_MAX_MODULI) n = HGXO_MAX_MODULI;
    make_modular_scheme(&g_scheme, mods, n);
    g_scheme_ready = 1;
  }
  return &g_scheme;
}

/* least-power-of-two->=, src/math-utilities.lisp:24-28 */
uint64_t hgxo_least_power_of_two_ge(uint64_t n) {
  if (n != 0 && (n & (n - 1)) == 0) return n;
  return 1ULL << integer_length_u64(n);
}

/* integer-length (* length (expt (1- $base) 2)), src/ntt-multiply.lisp:176-177 */
int hgxo_convolution_bits(uint64_t length) {
  /* (2^64-1)^2 = 2^128 - 2^65 + 1, as words {1, 2^64-2}; times length. */
  u128 lo = (u128)1 * length;               /* word 0 .. */
  u128 mid = (u128)(~0ULL - 1) * length;    /* word 1 .. */
  uint64_t w0 = (uint64_t)lo;
  u128 t = (lo >> 64) + (uint64_t)mid;
  uint64_t w1 = (uint64_t)t;
  u128 t2 = (t >> 64) + (mid >> 64);
  uint64_t w2 = (uint64_t)t2;
  (void)w0;
  if (w2) return 128 + integer_length_u64(w2);
  if (w1) return 64 + integer_length_u64(w1);
  return integer_length_u64(w0);
}

/* num-moduli-needed-for-bits, src/moduli.lisp:356-372 */
int hgxo_num_moduli_needed_for_bits(const hgxo_scheme *s, int bits) {
  int needed = 0;
  while (bits > 0) {
    if (needed == s->num_moduli) return -1; /* the reference asserts */
    bits -= integer_length_u64(s->moduli[needed]);
    needed++;
  }
  return needed;
}

/* ------------------------------------------------------------------ */
/* Transforms                                                          */
/* ------------------------------------------------------------------ */

static int lg_size(size_t n) { return integer_length_u64((uint64_t)n) - 1; }

// ntt-forward, src/number-theoretic-transform.lisp:40-90.  Decimation in
// frequency; stages lsubn = lg N .. 2 with twiddle w^j advanced by Shoup
// multiplication by dw = roots[lgN, mod, lgN - lsubn]; final radix-2 stage
// without multiplication.  Output in bit-reversed order.
void hgxo_ntt_forward(uint64_t *a, size_t n, int mod_num) {
  const hgxo_scheme *s = hgxo_default_scheme();
  uint64_t m = s->moduli[mod_num], minv = s->inverses[mod_num];
  int ln = lg_size(n);
  for (int lsubn = ln; lsubn >= 2; lsubn--) {
    size_t subn = (size_t)1 << lsubn, half = subn / 2;
    int kidx = ln - lsubn;
    if (hgxo_fault == HGXO_FAULT_TWIDDLE_INDEX && lsubn == 2) kidx += 1;
    uint64_t dw = s->roots[ROOT_IDX(s, ln, mod_num, kidx, 0)];
    uint64_t dwi = s->roots[ROOT_IDX(s, ln, mod_num, kidx, 1)];
    uint64_t wj = 1;
    for (size_t j = 0; j < half; j++) {
      for (size_t r = 0; r + subn <= n; r += subn) {
        uint64_t u = a[r + j], v = a[r + j + half];
        a[r + j] = hgxo_m_add(u, v, m);
        a[r + j + half] = hgxo_m_mul_fast(wj, hgxo_m_sub(u, v, m), m, minv);
      }
      wj = hgxo_m_mul_fast2(dw, dwi, wj, m);
    }
  }
  if (ln > 0) {
    for (size_t r = 0; r < n; r += 2) {
      uint64_t u = a[r], v = a[r + 1];
      a[r] = hgxo_m_add(u, v, m);
      a[r + 1] = hgxo_m_sub(u, v, m);
    }
  }
}

// ntt-reverse, src/number-theoretic-transform.lisp:93-143.  Decimation in
// time on bit-reversed input; the first radix-2 stage folds in 1/N by Shoup
// multiplication; stages ldm = 2 .. lg N use the inverse roots.
void hgxo_ntt_reverse(uint64_t *a, size_t n, int mod_num) {
  const hgxo_scheme *s = hgxo_default_scheme();
  uint64_t m = s->moduli[mod_num], minv = s->inverses[mod_num];
  int ldn = lg_size(n);
  uint64_t invn = s->inverse_lengths[INVLEN_IDX(s, ldn, mod_num, 0)];
  uint64_t invni = s->inverse_lengths[INVLEN_IDX(s, ldn, mod_num, 1)];
  if (hgxo_fault == HGXO_FAULT_INVERSE_SCALE) { invn = 1; invni = hgxo_calc_b_inv(1, m); }
  if (ldn > 0) {
    for (size_t r = 0; r < n; r += 2) {
      uint64_t u = a[r], v = a[r + 1];
      a[r] = hgxo_m_mul_fast2(invn, invni, hgxo_m_add(u, v, m), m);
      a[r + 1] = hgxo_m_mul_fast2(invn, invni, hgxo_m_sub(u, v, m), m);
    }
  }
  for (int ldm = 2; ldm <= ldn; ldm++) {
    size_t subn = (size_t)1 << ldm, half = subn / 2;
    uint64_t dw = s->inverse_roots[ROOT_IDX(s, ldn, mod_num, ldn - ldm, 0)];
    uint64_t dwi = s->inverse_roots[ROOT_IDX(s, ldn, mod_num, ldn - ldm, 1)];
    uint64_t wj = 1;
    for (size_t j = 0; j < half; j++) {
      for (size_t r = 0; r + subn <= n; r += subn) {
        uint64_t u = a[r + j];
        uint64_t v = hgxo_m_mul_fast(wj, a[r + j + half], m, minv);
        a[r + j] = hgxo_m_add(u, v, m);
        a[r + j + half] = hgxo_m_sub(u, v, m);
      }
      wj = hgxo_m_mul_fast2(dw, dwi, wj, m);
    }
  }
}

/* ------------------------------------------------------------------ */
/* CRT and carry                                                       */
/* ------------------------------------------------------------------ */

/* little-endian multiword helpers, NW words */
#define NW 5
static void mw_zero(uint64_t *a) { memset(a, 0, NW * sizeof(uint64_t)); }
static void mw_add(uint64_t *a, const uint64_t *b) {
  u128 c = 0;
  for (int i = 0; i < NW; i++) { c += (u128)a[i] + b[i]; a[i] = (uint64_t)c; c >>= 64; }
}
static void mw_sub(uint64_t *a, const uint64_t *b) {
  uint64_t borrow = 0;
  for (int i = 0; i < NW; i++) {
    uint64_t bi = b[i] + borrow;
    uint64_t nb = (bi < borrow) || (a[i] < bi);
    a[i] -= bi;
    borrow = nb;
  }
}
static int mw_cmp(const uint64_t *a, const uint64_t *b) {
  for (int i = NW - 1; i >= 0; i--) {
    if (a[i] != b[i]) return a[i] < b[i] ? -1 : 1;
  }
  return 0;
}
static void mw_mul_word(uint64_t *out, const uint64_t *a, uint64_t w) {
  u128 c = 0;
  for (int i = 0; i < NW; i++) { c += (u128)a[i] * w; out[i] = (uint64_t)c; c >>= 64; }
}
static void mw_shl1(uint64_t *a) {
  for (int i = NW - 1; i > 0; i--) a[i] = (a[i] << 1) | (a[i - 1] >> 63);
  a[0] <<= 1;
}

static int g_crt_ready = 0;
static uint64_t g_composite[NW];
static uint64_t g_factors[3][NW];
static uint64_t g_composite_shifted[80][NW]; /* composite * 2^j */

// The constants of the CRT block, src/ntt-multiply.lisp:243-246 (the same
// block appears for squares at :160-164): composite M = prod m_k, complements
// M/m_k, inverses (M/m_k)^-1 mod m_k, factors f_k = complement * inverse.
static void crt_init(void) {
  if (g_crt_ready) return;
  const hgxo_scheme *s = hgxo_default_scheme();
  uint64_t one[NW] = {1, 0, 0, 0, 0};
  uint64_t acc[NW];
  memcpy(acc, one, sizeof acc);
  for (int k = 0; k < 3; k++) { uint64_t t[NW]; mw_mul_word(t, acc, s->moduli[k]); memcpy(acc, t, sizeof acc); }
  memcpy(g_composite, acc, sizeof acc);
  for (int k = 0; k < 3; k++) {
    uint64_t comp[NW];
    memcpy(comp, one, sizeof comp);
    for (int j = 0; j < 3; j++)
      if (j != k) { uint64_t t[NW]; mw_mul_word(t, comp, s->moduli[j]); memcpy(comp, t, sizeof comp); }
    /* complement mod m_k */
    u128 r = 0;
    for (int i = NW - 1; i >= 0; i--) r = ((r << 64) | comp[i]) % s->moduli[k];
    uint64_t inv = hgxo_inv_mod((uint64_t)r, s->moduli[k]);
    mw_mul_word(g_factors[k], comp, inv);
  }
  memcpy(g_composite_shifted[0], g_composite, sizeof g_composite);
  for (int j = 1; j < 80; j++) {
    memcpy(g_composite_shifted[j], g_composite_shifted[j - 1], sizeof g_composite);
    mw_shl1(g_composite_shifted[j]);
  }
  g_crt_ready = 1;
}

// The per-coefficient CRT, src/ntt-multiply.lisp:247-252:
// (mod (sum f_k * r_k) composite).  The remainder is taken by shift and
// subtract; any exact remainder is the same number.
void hgxo_crt(const uint64_t *residues, uint64_t out[4]) {
  crt_init();
  uint64_t sum[NW];
  mw_zero(sum);
  int nk = (hgxo_fault == HGXO_FAULT_CRT_DROP_TERM) ? 2 : 3;
  for (int k = 0; k < nk; k++) {
    uint64_t t[NW];
    mw_mul_word(t, g_factors[k], residues[k]);
    mw_add(sum, t);
  }
  for (int j = 79; j >= 0; j--) {
    if (g_composite_shifted[j][NW - 1] >> 63) continue; /* would overflow NW words */
    if (mw_cmp(sum, g_composite_shifted[j]) >= 0) mw_sub(sum, g_composite_shifted[j]);
  }
  for (int i = 0; i < 4; i++) out[i] = sum[i];
}

// add-big-digit, src/ntt-multiply.lisp:10-51: add D into STORAGE at digit I
// with a carry ripple; a non-zero value landing at or beyond LENGTH is the
// reference's "grave inconsistency" error.  (The reference's bignum branch
// also touches a zero sign word past the end when the digit's top bit is
// set; adding zero there changes no value, so zero words past the end are
// skipped here.)
int hgxo_add_big_digit(const uint64_t *d, int nd, uint64_t *storage,
                       size_t length, size_t i) {
  uint64_t carry = 0;
  size_t k = 0;
  for (;; k++) {
    uint64_t w = (k < (size_t)nd) ? d[k] : 0;
    if (w == 0 && carry == 0 && k >= (size_t)nd) break;
    if (i + k >= length) {
      if (w == 0 && carry == 0) continue;
      return -1;
    }
    u128 sum = (u128)storage[i + k] + w + carry;
    storage[i + k] = (uint64_t)sum;
    carry = (uint64_t)(sum >> 64);
    if (hgxo_fault == HGXO_FAULT_CARRY) carry = 0;
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* The multiply path                                                   */
/* ------------------------------------------------------------------ */

// %make-ntt-work, src/number-theoretic-transform.lisp:7-22: a[i] = limb mod m
// for i below the operand length, zero above.
static uint64_t *make_ntt_work(const uint64_t *x, size_t nx, size_t length, uint64_t m) {
  uint64_t *a = (uint64_t *)calloc(length, sizeof(uint64_t));
  for (size_t i = 0; i < nx; i++) a[i] = x[i] % m;
  return a;
}

size_t hgxo_product_length(size_t nx, size_t ny) {
  return (size_t)hgxo_least_power_of_two_ge((uint64_t)(nx + ny));
}

static void crt_and_carry(uint64_t **ntts, size_t length, uint64_t *result) {
  memset(result, 0, length * sizeof(uint64_t));
  for (size_t i = 0; i < length; i++) {
    uint64_t r[3] = {ntts[0][i], ntts[1][i], ntts[2][i]};
    uint64_t c[4];
    hgxo_crt(r, c);
    int nd = 4;
    while (nd > 0 && c[nd - 1] == 0) nd--;
    if (nd) hgxo_add_big_digit(c, nd, result, length, i);
  }
}

// mpz-*/ntt, src/ntt-multiply.lisp:174-254.  Length = least power of two
// >= nx + ny; all scheme moduli are prepared; forward transforms of both
// operands; Barrett pointwise product into X's work (multiply-pointwise!,
// :88-99); inverse transforms; CRT and carry into an L-limb result.
size_t hgxo_mpz_mul_ntt(const uint64_t *x, size_t nx, const uint64_t *y,
                        size_t ny, uint64_t *out) {
  const hgxo_scheme *s = hgxo_default_scheme();
  size_t length = hgxo_product_length(nx, ny);
  int bits = hgxo_convolution_bits(length);
  int num_moduli = hgxo_num_moduli_needed_for_bits(s, bits);
  if (num_moduli != 3) return 0;
  uint64_t *nx_work[3], *ny_work[3];
  for (int k = 0; k < 3; k++) {
    nx_work[k] = make_ntt_work(x, nx, length, s->moduli[k]);
    ny_work[k] = make_ntt_work(y, ny, length, s->moduli[k]);
  }
  for (int k = 0; k < num_moduli; k++) {
    hgxo_ntt_forward(nx_work[k], length, k);
    hgxo_ntt_forward(ny_work[k], length, k);
  }
  for (int k = 0; k < num_moduli; k++)
    for (size_t i = 0; i < length; i++)
      nx_work[k][i] = hgxo_m_mul_fast(nx_work[k][i], ny_work[k][i], s->moduli[k], s->inverses[k]);
  for (int k = 0; k < num_moduli; k++) hgxo_ntt_reverse(nx_work[k], length, k);
  crt_and_carry(nx_work, length, out);
  for (int k = 0; k < 3; k++) { free(nx_work[k]); free(ny_work[k]); }
  return length;
}

// mpz-*/ntt with the reference's :hypergeometrica-parallel feature:
// with-parallel-work / with-task (src/ntt-multiply.lisp:60-86) submit one task
// per (operand, modulus) for the forward transforms (:202-209), one per modulus
// for the pointwise products (:214-220) and one per modulus for the inverse
// transforms (:229-234), joining after each group; residue preparation and the
// CRT stay serial (:178-179, :242-252).  At most NTHREADS tasks run at once
// (lparallel's pool, src/config.lisp:18).  Same product as hgxo_mpz_mul_ntt.
typedef struct {
  uint64_t *a;
  const uint64_t *b;
  size_t n;
  int k;
} par_task;

static void *par_forward(void *p) {
  par_task *t = (par_task *)p;
  hgxo_ntt_forward(t->a, t->n, t->k);
  return NULL;
}

static void *par_pointwise(void *p) {
  par_task *t = (par_task *)p;
  const hgxo_scheme *s = hgxo_default_scheme();
  for (size_t i = 0; i < t->n; i++)
    t->a[i] = hgxo_m_mul_fast(t->a[i], t->b[i], s->moduli[t->k], s->inverses[t->k]);
  return NULL;
}

static void *par_reverse(void *p) {
  par_task *t = (par_task *)p;
  hgxo_ntt_reverse(t->a, t->n, t->k);
  return NULL;
}

// run NTASKS tasks, at most NTHREADS at a time; 0 on success
static int run_tasks(void *(*fn)(void *), par_task *tasks, int ntasks, int nthreads) {
  pthread_t th[6];
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 6) nthreads = 6;
  for (int i = 0; i < ntasks; i += nthreads) {
    const int m = (ntasks - i < nthreads) ? ntasks - i : nthreads;
    int started = 0, rc = 0;
    for (int j = 0; j < m; j++) {
      if (pthread_create(&th[j], NULL, fn, &tasks[i + j]) != 0) {
        rc = -1;
        break;
      }
      started++;
    }
    for (int j = 0; j < started; j++) pthread_join(th[j], NULL);
    if (rc) return rc;
  }
  return 0;
}

size_t hgxo_mpz_mul_ntt_parallel(const uint64_t *x, size_t nx, const uint64_t *y,
                                 size_t ny, uint64_t *out, int nthreads) {
  const hgxo_scheme *s = hgxo_default_scheme();
  size_t length = hgxo_product_length(nx, ny);
  int num_moduli = hgxo_num_moduli_needed_for_bits(s, hgxo_convolution_bits(length));
  if (num_moduli != 3) return 0;
  uint64_t *nx_work[3], *ny_work[3];
  for (int k = 0; k < 3; k++) {
    nx_work[k] = make_ntt_work(x, nx, length, s->moduli[k]);
    ny_work[k] = make_ntt_work(y, ny, length, s->moduli[k]);
  }
  par_task fwd[6], pw[3], inv[3];
  for (int k = 0; k < 3; k++) {
    fwd[2 * k] = (par_task){nx_work[k], NULL, length, k};
    fwd[2 * k + 1] = (par_task){ny_work[k], NULL, length, k};
    pw[k] = (par_task){nx_work[k], ny_work[k], length, k};
    inv[k] = (par_task){nx_work[k], NULL, length, k};
  }
  size_t rc = length;
  if (run_tasks(par_forward, fwd, 6, nthreads) || run_tasks(par_pointwise, pw, 3, nthreads) ||
      run_tasks(par_reverse, inv, 3, nthreads))
    rc = 0;
  else
    crt_and_carry(nx_work, length, out);
  for (int k = 0; k < 3; k++) { free(nx_work[k]); free(ny_work[k]); }
  return rc;
}

// mpz-square, src/ntt-multiply.lisp:101-172: length = least power of two
// >= 2n, forward transforms of the one operand, in-place squaring, inverse,
// CRT and carry into a fresh length-limb result.
size_t hgxo_mpz_square(const uint64_t *x, size_t nx, uint64_t *out) {
  const hgxo_scheme *s = hgxo_default_scheme();
  size_t length = (size_t)hgxo_least_power_of_two_ge((uint64_t)(2 * nx));
  int num_moduli = hgxo_num_moduli_needed_for_bits(s, hgxo_convolution_bits(length));
  if (num_moduli != 3) return 0;
  uint64_t *w[3];
  for (int k = 0; k < 3; k++) w[k] = make_ntt_work(x, nx, length, s->moduli[k]);
  for (int k = 0; k < num_moduli; k++) hgxo_ntt_forward(w[k], length, k);
  for (int k = 0; k < num_moduli; k++)
    for (size_t i = 0; i < length; i++)
      w[k][i] = hgxo_m_mul_fast(w[k][i], w[k][i], s->moduli[k], s->inverses[k]);
  for (int k = 0; k < num_moduli; k++) hgxo_ntt_reverse(w[k], length, k);
  crt_and_carry(w, length, out);
  for (int k = 0; k < 3; k++) free(w[k]);
  return length;
}

// %multiply-storage/schoolboy-style product (src/mpz-ram.lisp:404-431),
// used here only as an independent pin of the NTT product.
void hgxo_schoolbook(const uint64_t *x, size_t nx, const uint64_t *y,
                     size_t ny, uint64_t *out) {
  memset(out, 0, (nx + ny) * sizeof(uint64_t));
  for (size_t i = 0; i < nx; i++) {
    u128 carry = 0;
    for (size_t j = 0; j < ny; j++) {
      u128 t = (u128)x[i] * y[j] + out[i + j] + (uint64_t)carry;
      out[i + j] = (uint64_t)t;
      carry = t >> 64;
    }
    size_t k = i + ny;
    while (carry) {
      u128 t = (u128)out[k] + (uint64_t)carry;
      out[k] = (uint64_t)t;
      carry = t >> 64;
      k++;
    }
  }
}

void hgxo_splitmix64(uint64_t seed, uint64_t *out, size_t n) {
  const uint64_t golden = 0x9E3779B97F4A7C15ULL;
  for (size_t i = 0; i < n; i++) {
    uint64_t z = seed + (uint64_t)(i + 1) * golden;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    out[i] = z ^ (z >> 31);
  }
}
