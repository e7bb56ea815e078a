/* Hexadecimal digits of pi at a position, by the Bailey-Borwein-Plouffe
 * digit extraction (pi = sum_k 16^-k (4/(8k+1) - 2/(8k+4) - 1/(8k+5) -
 * 1/(8k+6)); the reference states the same formula in src/bbp.lisp:8-42).
 * An independent check of the high-precision pi; test and benchmark
 * infrastructure only, built with gcc into tests/libhgx_bbp.so.
 *
 * hgxc_bbp_frac(d, threads) returns frac(16^d pi) as a 64-bit fixed-point
 * fraction.  Every term frac(16^(d-k) / (8k+j)) is taken as
 * floor((16^(d-k) mod m) 2^64 / m) (exact, 128-bit division), and the sums
 * run mod 2^64, so the result is below the true value by less than
 * 8 (d + 17) units of 2^-64. */
#include <pthread.h>
#include <stddef.h>
#include <stdint.h>

typedef unsigned __int128 u128;

static uint64_t powmod16(uint64_t e, uint64_t m) { /* 16^e mod m, m < 2^32 */
  uint64_t r = 1 % m, b = 16 % m;
  while (e) {
    if (e & 1) r = (r * b) % m;
    b = (b * b) % m;
    e >>= 1;
  }
  return r;
}

/* sum over k in [k0, k1) of frac(16^(d-k) / (8k + j)), times coefficient c */
static uint64_t head(uint64_t d, uint64_t k0, uint64_t k1) {
  uint64_t s = 0;
  for (uint64_t k = k0; k < k1; k++) {
    const uint64_t e = d - k;
    const uint64_t m1 = 8 * k + 1, m4 = 8 * k + 4, m5 = 8 * k + 5, m6 = 8 * k + 6;
    const uint64_t t1 = (uint64_t)(((u128)powmod16(e, m1) << 64) / m1);
    const uint64_t t4 = (uint64_t)(((u128)powmod16(e, m4) << 64) / m4);
    const uint64_t t5 = (uint64_t)(((u128)powmod16(e, m5) << 64) / m5);
    const uint64_t t6 = (uint64_t)(((u128)powmod16(e, m6) << 64) / m6);
    s += 4 * t1 - 2 * t4 - t5 - t6;
  }
  return s;
}

typedef struct {
  uint64_t d, k0, k1, sum;
} part_t;

static void *run(void *arg) {
  part_t *p = (part_t *)arg;
  p->sum = head(p->d, p->k0, p->k1);
  return NULL;
}

uint64_t hgxc_bbp_frac(uint64_t d, int threads) {
  if (threads < 1) threads = 1;
  if (threads > 64) threads = 64;
  pthread_t th[64];
  part_t parts[64];
  const uint64_t n = d + 1, per = (n + threads - 1) / threads;
  int started = 0;
  for (int i = 0; i < threads; i++) {
    parts[i].d = d;
    parts[i].k0 = (uint64_t)i * per < n ? (uint64_t)i * per : n;
    parts[i].k1 = parts[i].k0 + per < n ? parts[i].k0 + per : n;
    parts[i].sum = 0;
    if (pthread_create(&th[i], NULL, run, &parts[i]) == 0) {
      started = i + 1;
    } else {
      run(&parts[i]); /* no thread: this part here */
      th[i] = 0;
    }
  }
  uint64_t s = 0;
  for (int i = 0; i < threads; i++) {
    if (i < started && th[i]) pthread_join(th[i], NULL);
    s += parts[i].sum;
  }
  /* tail k > d: 16^(d-k) / (8k + j) = 2^(-4i) / m, below 2^-64 for i >= 16 */
  for (uint64_t i = 1; i < 17; i++) {
    const uint64_t k = d + i;
    const u128 num = (u128)1 << (64 - 4 * i);
    s += 4 * (uint64_t)(num / (8 * k + 1)) - 2 * (uint64_t)(num / (8 * k + 4)) - (uint64_t)(num / (8 * k + 5)) -
         (uint64_t)(num / (8 * k + 6));
  }
  return s;
}
