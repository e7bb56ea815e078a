/* Residues of little-endian uint64 limb vectors modulo the Mersenne prime
 * q = 2^61 - 1: the size-independent check of products used by the tests and
 * by the benchmark's serving leg ((a b) mod q = (a mod q)(b mod q) mod q).
 * Since 2^61 = 1 (mod q), 2^64 = 8 and limb i weighs 2^(3i mod 61): a 61-bit
 * rotation of the limb's own residue.  Test and benchmark infrastructure
 * only; built with gcc by the build step into tests/libhgx_residue.so. */
#include <stddef.h>
#include <stdint.h>

#define Q ((((uint64_t)1) << 61) - 1)

static inline uint64_t fold(uint64_t x) { /* x mod q, canonical */
  uint64_t r = (x & Q) + (x >> 61);
  return r >= Q ? r - Q : r;
}

static inline uint64_t rot(uint64_t v, unsigned s) { /* v * 2^s mod q, v < q, s < 61 */
  return s ? (((v << s) & Q) | (v >> (61 - s))) : v;
}

/* the terms of limbs i0 .. i0 + m (m <= 61) with shifts T[0..m), summed as
 * 31-bit and 30-bit halves */
__attribute__((target_clones("avx2", "default")))
static void block_sums(const uint64_t *a, size_t m, const unsigned *T, uint64_t *l, uint64_t *h) {
  uint64_t sl = 0, sh = 0;
  for (size_t k = 0; k < m; k++) {
    uint64_t x = a[k];
    uint64_t v = (x & Q) + (x >> 61);
    v = v >= Q ? v - Q : v;
    const uint64_t t = ((v << T[k]) & Q) | (v >> (61 - T[k]));  /* v * 2^T[k] mod q */
    sl += t & 0x7fffffffu;
    sh += t >> 31;
  }
  *l += sl;
  *h += sh;
}

/* sum over i < n of a[i] * 2^(64 (i + first)) mod q: limb i weighs
 * 2^(3 (i + first) mod 61), a shift pattern of period 61 */
uint64_t hgxc_residue_at(const uint64_t *a, size_t n, uint64_t first) {
  unsigned T[61];
  const unsigned s0 = (unsigned)((3 * (first % 61)) % 61);
  for (unsigned k = 0; k < 61; k++) T[k] = (s0 + 3 * k) % 61;
  uint64_t lo = 0, hi = 0;
  size_t i = 0;
  while (i < n) {
    /* fold at least every 2^31 terms: the halves are below 2^31 */
    const size_t end = n - i > ((size_t)61 << 26) ? i + ((size_t)61 << 26) : n;
    uint64_t l = 0, h = 0;
    for (; i < end; i += 61) block_sums(a + i, end - i < 61 ? end - i : 61, T, &l, &h);
    lo = fold(lo + fold(l));
    hi = fold(hi + fold(h));
  }
  return fold(lo + rot(hi, 31));
}

uint64_t hgxc_residue(const uint64_t *a, size_t n) { return hgxc_residue_at(a, n, 0); }

/* P[k] = residue of a[0..k) for k = 0..n (n + 1 words) */
void hgxc_prefix_residues(const uint64_t *a, size_t n, uint64_t *P) {
  uint64_t acc = 0;
  unsigned s = 0;
  P[0] = 0;
  for (size_t i = 0; i < n; i++) {
    acc += rot(fold(a[i]), s);
    acc = acc >= Q ? acc - Q : acc;
    P[i + 1] = acc;
    s += 3;
    if (s >= 61) s -= 61;
  }
}

/* (x * y) mod q */
uint64_t hgxc_mulmod(uint64_t x, uint64_t y) {
  const unsigned __int128 p = (unsigned __int128)fold(x) * fold(y);
  return fold(fold((uint64_t)p & Q) + fold((uint64_t)(p >> 61)));
}

/* residue of the slice a[off .. off + n) of the vector whose prefix
 * residues are P: (P[off + n] - P[off]) * 2^(-64 off) mod q */
uint64_t hgxc_slice_residue(const uint64_t *P, size_t off, size_t n) {
  uint64_t d = P[off + n] >= P[off] ? P[off + n] - P[off] : P[off + n] + Q - P[off];
  const unsigned s = (unsigned)((3 * (off % 61)) % 61);
  return rot(d, (61 - s) % 61);
}
