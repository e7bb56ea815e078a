/* A C host of the library: the calls an SBCL or C caller makes through
 * include/hgx.h that need no device.  Prints, one per line:
 *   moduli <p0> <p1> <p2>
 *   length <hgx_product_length(65536, 65536)>
 *   mul_uninit <return code of hgx_mul before hgx_init>
 *   series <which> <sign> <nlimbs> <limb0> ... (hgx_series_split of the
 *          Chudnovsky series over [0, 30), all its products on the host)
 */
#include <inttypes.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "hgx.h"

int main(void) {
  printf("moduli %" PRIu64 " %" PRIu64 " %" PRIu64 "\n", hgx_modulus(0), hgx_modulus(1), hgx_modulus(2));
  printf("length %zu\n", hgx_product_length(65536, 65536));
  uint64_t a[2] = {3, 4}, b[2] = {5, 6}, out[4];
  printf("mul_uninit %d\n", hgx_mul(a, 2, b, 2, out));

  /* make-chudnovsky-series (src/pi.lisp:24-43) */
  hgx_series s;
  memset(&s, 0, sizeof s);
  s.degree[0] = 1; s.coef[0][0] = 135914090; s.coef[0][1] = 5451401340;      /* a */
  s.degree[1] = 0; s.coef[1][0] = 426880;                                     /* b */
  s.degree[2] = 3; s.coef[2][0] = 5; s.coef[2][1] = -46; s.coef[2][2] = 108;  /* p */
  s.coef[2][3] = -72; s.at_zero[2] = 1; s.value0[2] = 1;
  s.degree[3] = 3; s.coef[3][3] = 10939058860032000; s.at_zero[3] = 1; s.value0[3] = 1;  /* q */
  hgx_partial *res = NULL;
  int rc = hgx_series_split(&s, 0, 30, &res);
  if (rc) {
    printf("series_error %d %s\n", rc, hgx_last_error());
    return 1;
  }
  for (int w = 0; w < 4; w++) {
    int sign = 0;
    size_t n = 0;
    if (hgx_partial_get(res, w, &sign, &n, NULL)) return 1;
    uint64_t *limbs = (uint64_t *)calloc(n ? n : 1, sizeof(uint64_t));
    if (!limbs || hgx_partial_get(res, w, NULL, NULL, limbs)) return 1;
    printf("series %d %d %zu", w, sign, n);
    for (size_t i = 0; i < n; i++) printf(" %" PRIu64, limbs[i]);
    printf("\n");
    free(limbs);
  }
  hgx_partial_free(res);
  return 0;
}
