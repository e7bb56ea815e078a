/*
 * hgx_oracle.h -- CPU restatement of Hypergeometrica's exact NTT multiply.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker that the MI355X library
 * (hypergeometrica_amd/libhgx.so) is compared against.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.  The
 * product path never calls it.
 *
 * Every function restates one reference function; the citation is on the
 * definition in hgx_oracle.c.  All arithmetic is exact integer arithmetic.
 */
#ifndef HGX_ORACLE_H
#define HGX_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HGXO_MAX_MODULI 16
#define HGXO_MAX_TRANS 56 /* max-transform-length + 1 (2-adicity 55 -> 0..55) */

/* Scheme tables (src/moduli.lisp:272-296). */
typedef struct {
  int num_moduli;
  int max_transform_length; /* exponent of two */
  uint64_t moduli[HGXO_MAX_MODULI];
  uint64_t generators[HGXO_MAX_MODULI];
  /* [n][m][k][i]: the 2^k-th power of the primitive 2^n-th root of unity of
     MODULI[m]; i = 0 the value, i = 1 its Shoup companion floor(w*2^64/m). */
  uint64_t *roots;
  uint64_t *inverse_roots;
  /* [k][m][i]: 1/2^k mod MODULI[m] and its Shoup companion. */
  uint64_t *inverse_lengths;
  /* Barrett inverses floor(2^(64+lg) / m), lg = min over moduli of lg(m). */
  uint64_t inverses[HGXO_MAX_MODULI];
} hgxo_scheme;

/* Fault-injection switch used only by the oracle's own mutation tests: a
   non-zero value plants one known mistake so the tests can show that the
   pins catch it.  0 = faithful restatement. */
extern int hgxo_fault;
enum {
  HGXO_FAULT_NONE = 0,
  HGXO_FAULT_CRT_DROP_TERM = 1,   /* drop the last prime's CRT term */
  HGXO_FAULT_TWIDDLE_INDEX = 2,   /* use the next root-table slot in stage 2 */
  HGXO_FAULT_INVERSE_SCALE = 3,   /* forget the 1/N factor of ntt-reverse */
  HGXO_FAULT_BARRETT_FIX = 4,     /* skip the second Barrett correction */
  HGXO_FAULT_CARRY = 5            /* drop the carry out of add-big-digit */
};

/* L0/L1 primitives */
int hgxo_primep(uint64_t n);
uint64_t hgxo_expt_mod(uint64_t a, uint64_t n, uint64_t m);
uint64_t hgxo_inv_mod(uint64_t x, uint64_t m);
uint64_t hgxo_m_add(uint64_t a, uint64_t b, uint64_t m);
uint64_t hgxo_m_sub(uint64_t a, uint64_t b, uint64_t m);
uint64_t hgxo_m_mul(uint64_t a, uint64_t b, uint64_t m);
uint64_t hgxo_mod128_fast(uint64_t lo, uint64_t hi, uint64_t m, uint64_t minv);
uint64_t hgxo_m_mul_fast(uint64_t a, uint64_t b, uint64_t m, uint64_t minv);
uint64_t hgxo_calc_b_inv(uint64_t b, uint64_t m);
uint64_t hgxo_m_mul_fast2(uint64_t b, uint64_t binv, uint64_t a, uint64_t m);

/* vectorised helpers for tests: out[i] = op(a[i], b[i]) */
void hgxo_m_mul_fast_vec(const uint64_t *a, const uint64_t *b, uint64_t *out,
                         size_t n, uint64_t m, uint64_t minv);
void hgxo_m_mul_fast2_vec(const uint64_t *b, const uint64_t *a, uint64_t *out,
                          size_t n, uint64_t m);

/* moduli and scheme */
int hgxo_default_moduli(uint64_t *out, int max_out);
uint64_t hgxo_find_generator(uint64_t m);
uint64_t hgxo_primitive_root(uint64_t g, uint64_t n, uint64_t m);
const hgxo_scheme *hgxo_default_scheme(void);
int hgxo_num_moduli_needed_for_bits(const hgxo_scheme *s, int bits);
int hgxo_convolution_bits(uint64_t length);
uint64_t hgxo_least_power_of_two_ge(uint64_t n);

/* transforms (in place; n must be a power of two) */
void hgxo_ntt_forward(uint64_t *a, size_t n, int mod_num);
void hgxo_ntt_reverse(uint64_t *a, size_t n, int mod_num);

/* the multiply path.  OUT receives LENGTH limbs (the reference's result
   storage is the full transform length); returns LENGTH. */
size_t hgxo_mpz_mul_ntt(const uint64_t *x, size_t nx, const uint64_t *y,
                        size_t ny, uint64_t *out);
/* The NTT multiply with :hypergeometrica-parallel: one thread per (operand, modulus)
   forward task, per modulus pointwise and inverse task, at most NTHREADS (<= 6)
   at once (src/ntt-multiply.lisp:60-86, :202-234).  0 if a thread fails. */
size_t hgxo_mpz_mul_ntt_parallel(const uint64_t *x, size_t nx, const uint64_t *y,
                                 size_t ny, uint64_t *out, int nthreads);
size_t hgxo_mpz_square(const uint64_t *x, size_t nx, uint64_t *out);
size_t hgxo_product_length(size_t nx, size_t ny);

/* CRT of one coefficient: out[0..3] = (sum_k f_k r_k) mod M. */
void hgxo_crt(const uint64_t *residues, uint64_t out[4]);

/* add-big-digit: add the little-endian big digit D (nd words) into
   STORAGE[i..]; returns -1 on the reference's "grave inconsistency". */
int hgxo_add_big_digit(const uint64_t *d, int nd, uint64_t *storage,
                       size_t length, size_t i);

/* schoolbook product (pins the NTT product independently). */
void hgxo_schoolbook(const uint64_t *x, size_t nx, const uint64_t *y,
                     size_t ny, uint64_t *out);

/* SplitMix64 stream: out[i] = mix(seed + (i + 1) * golden). */
void hgxo_splitmix64(uint64_t seed, uint64_t *out, size_t n);

#ifdef __cplusplus
}
#endif
#endif
