// series_emulated.cpp -- the binary-splitting engine (hgx_series.cpp)
// linked against host stand-ins for the library's device side, so that its
// host logic runs on a machine without a GPU: "device" memory is host
// memory, products are schoolbook, and the device signed sums are restated
// one limb at a time.  The engine's own code -- tree, folds, regions, the
// move of the values to the device and back when device memory runs out --
// is what runs.
//
// Usage: series_emulated WHICH LOWER UPPER.  WHICH 0: the Chudnovsky series;
// 1: a series with negative a, b, p and q; 2: a = 1, p = -1 (1 at k = 0),
// whose partial sums cancel; 3: p = 2^40 k^7, q = k + 3, whose term values
// near k = 2^39 exceed what the device fold takes; 4: p = 2^20 k^7, q = k + 3,
// whose term values near k = 2^33 take 4 limbs.  Prints P, Q, B, R as
// SIGN:HEX, one per line.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../hypergeometrica_amd/csrc/hgx_fold.hpp"
#include "../../hypergeometrica_amd/csrc/hgx_sum.h"
#include "../../include/hgx.h"

typedef unsigned __int128 u128;

int hgx_internal_fail(int code, const char *) { return code; }
bool hgx_internal_ready() { return true; }
extern "C" const char *hgx_last_error(void) { return "emulated"; }

extern "C" int hgx_mul_batch(const hgx_mul_desc *d, size_t count) {
  for (size_t k = 0; k < count; k++) {
    const uint64_t *a = d[k].a, *b = d[k].b;
    const size_t na = d[k].na, nb = d[k].nb;
    uint64_t *r = d[k].out;
    memset(r, 0, 8 * (na + nb));
    for (size_t j = 0; j < nb; j++) {
      uint64_t c = 0;
      for (size_t i = 0; i < na; i++) {
        const u128 t = (u128)a[i] * b[j] + r[i + j] + c;
        r[i + j] = (uint64_t)t;
        c = (uint64_t)(t >> 64);
      }
      r[j + na] = c;
    }
  }
  return HGX_OK;
}

// device memory and copies
extern "C" int hgx_device_alloc(size_t n, uint64_t **d) {
  *d = (uint64_t *)malloc(8 * n);
  return *d ? HGX_OK : HGX_ENOMEM;
}
extern "C" int hgx_device_free(uint64_t *d, void *) {
  free(d);
  return HGX_OK;
}
extern "C" int hgx_upload(uint64_t *d, const uint64_t *s, size_t n, void *) {
  memcpy(d, s, 8 * n);
  return HGX_OK;
}
extern "C" int hgx_download(uint64_t *d, const uint64_t *s, size_t n, void *) {
  memcpy(d, s, 8 * n);
  return HGX_OK;
}
int hgx_internal_upload_gather(uint64_t *d_dst, const uint64_t *const *src, const size_t *nlimbs, size_t count) {
  for (size_t k = 0; k < count; k++) {
    memcpy(d_dst, src[k], 8 * nlimbs[k]);
    d_dst += nlimbs[k];
  }
  return HGX_OK;
}
extern "C" int hgx_mul_batch_device(const hgx_mul_desc *d, size_t count, void *) { return hgx_mul_batch(d, count); }
extern "C" int hgx_device_significant(const uint64_t *d, size_t n, size_t *out, void *) {
  while (n && !d[n - 1]) n--;
  *out = n;
  return HGX_OK;
}
int hgx_internal_copy_device(uint64_t *d, const uint64_t *s, size_t n) {
  memcpy(d, s, 8 * n);
  return HGX_OK;
}
size_t hgx_internal_device_room() { return (size_t)1 << 40; }

// the device leaf folds (hgx_sum.h): the kernels' own fold_leaf and
// fold_pack_leaf, one leaf after another, each fold buffer a separate
// allocation of exactly its capacity (so that the sanitizer sees a write
// past it); reports on stderr that it ran
int hgx_internal_fold_device(const FoldSeries *S, const uint64_t *bounds, size_t count, const uint32_t *cap,
                             int32_t *codes, uint64_t **d_vals, size_t *nlimbs) {
  const size_t st = (count + 63) & ~(size_t)63;
  FoldSlots F;
  for (int b = 0; b < kFoldBufs; b++) {
    F.buf[b] = (uint64_t *)malloc(8 * st * cap[b]);
    F.cap[b] = cap[b];
  }
  F.stride = st;
  bool ok = true;
  for (size_t t = 0; t < count && ok; t++) ok = fold_leaf(*S, bounds[2 * t], bounds[2 * t + 1], t, F, codes + 4 * t);
  size_t total = 0;
  for (size_t k = 0; ok && k < 4 * count; k++) total += (size_t)llabs(codes[k]);
  uint64_t *out = ok ? (uint64_t *)malloc(8 * (total + 1)) : nullptr;
  for (size_t t = 0, at = 0; ok && t < count; t++) {
    fold_pack_leaf(F, t, codes + 4 * t, out + at);
    for (int w = 0; w < 4; w++) at += (size_t)llabs(codes[4 * t + w]);
  }
  for (int b = 0; b < kFoldBufs; b++) free(F.buf[b]);
  fprintf(stderr, "emulated device fold: %zu leaves, %s\n", count, ok ? "done" : "a leaf failed");
  if (!ok) return HGX_ETOOBIG;
  *d_vals = out;
  *nlimbs = total;
  return HGX_OK;
}

// the device signed sums (hgx_sum.h), one limb at a time
static int mag_cmp(const std::vector<uint64_t> &a, const std::vector<uint64_t> &b) {
  for (size_t i = a.size(); i-- > 0;)
    if (a[i] != b[i]) return a[i] < b[i] ? -1 : 1;
  return 0;
}
int hgx_internal_sums_device(HgxSumJob *jobs, size_t count, int *signs, uint64_t *sizes) {
  for (size_t j = 0; j < count; j++) {
    HgxSumJob &J = jobs[j];
    if (J.sx == 0) J.nx = 0;
    J.n = (J.nx > J.ny ? J.nx : J.ny) + 1;
    std::vector<uint64_t> x(J.n, 0), y(J.n, 0), r(J.n, 0);
    for (size_t i = 0; i < J.nx; i++) x[i] = J.x[i];
    for (size_t i = 0; i < J.ny; i++) y[i] = J.y[i];
    int sg;
    if (J.sx == 0 || J.sx == J.sy) {
      u128 c = 0;
      for (size_t i = 0; i < J.n; i++) {
        c += (u128)x[i] + y[i];
        r[i] = (uint64_t)c;
        c >>= 64;
      }
      sg = J.sy;
    } else {
      const int c = mag_cmp(x, y);
      const std::vector<uint64_t> &a = c >= 0 ? x : y, &b = c >= 0 ? y : x;
      uint64_t borrow = 0;
      for (size_t i = 0; i < J.n; i++) {
        const uint64_t d = a[i] - b[i];
        const uint64_t nb = (a[i] < b[i]) || (d < borrow);
        r[i] = d - borrow;
        borrow = nb;
      }
      sg = c == 0 ? 0 : (c > 0 ? J.sx : J.sy);
    }
    for (size_t i = 0; i < J.n; i++) J.x[i] = r[i];
    size_t top = J.n;
    while (top && !r[top - 1]) top--;
    sizes[j] = top;
    signs[j] = top ? sg : 0;
  }
  return HGX_OK;
}

int main(int argc, char **argv) {
  if (argc != 4) return 2;
  hgx_series s;
  memset(&s, 0, sizeof s);
  const int which = atoi(argv[1]);
  const uint64_t lo = strtoull(argv[2], 0, 10), up = strtoull(argv[3], 0, 10);
  if (which == 0) {  // make-chudnovsky-series (src/pi.lisp:24-43)
    s.degree[0] = 1;
    s.coef[0][0] = 135914090;
    s.coef[0][1] = 5451401340LL;
    s.coef[1][0] = 426880;
    s.degree[2] = 3;
    s.coef[2][0] = 5;
    s.coef[2][1] = -46;
    s.coef[2][2] = 108;
    s.coef[2][3] = -72;
    s.at_zero[2] = 1;
    s.value0[2] = 1;
    s.degree[3] = 3;
    s.coef[3][3] = 10939058860032000LL;
    s.at_zero[3] = 1;
    s.value0[3] = 1;
  } else if (which == 1) {
    s.degree[0] = 1;
    s.coef[0][0] = 2;
    s.coef[0][1] = -7;
    s.coef[1][0] = -3;
    s.degree[2] = 2;
    s.coef[2][0] = 1;
    s.coef[2][1] = 2;
    s.coef[2][2] = -1;
    s.degree[3] = 2;
    s.coef[3][0] = -4;
    s.coef[3][2] = 5;
    s.at_zero[3] = 1;
    s.value0[3] = -1;
  } else if (which == 2) {
    s.coef[0][0] = 1;
    s.coef[1][0] = 1;
    s.coef[2][0] = -1;
    s.at_zero[2] = 1;
    s.value0[2] = 1;
    s.coef[3][0] = 1;
  } else {  // 3: p(k) = 2^40 k^7, beyond 256 bits near k = 2^39; 4: p(k) = 2^20 k^7, under 256 bits near 2^33
    s.coef[0][0] = 1;
    s.coef[1][0] = 1;
    s.degree[2] = 7;
    s.coef[2][7] = which == 3 ? 1LL << 40 : 1LL << 20;
    s.degree[3] = 1;
    s.coef[3][0] = 3;
    s.coef[3][1] = 1;
  }
  hgx_partial *p = nullptr;
  const int rc = hgx_series_split(&s, lo, up, &p);
  if (rc) {
    printf("rc %d\n", rc);
    return 1;
  }
  for (int k = 0; k < 4; k++) {
    int sg = 0;
    size_t n = 0;
    hgx_partial_get(p, k, &sg, &n, nullptr);
    std::vector<uint64_t> v(n);
    hgx_partial_get(p, k, &sg, &n, v.data());
    printf("%d:", sg);
    for (size_t i = n; i-- > 0;) printf("%016llx", (unsigned long long)v[i]);
    printf("\n");
  }
  hgx_partial_free(p);
  return 0;
}
