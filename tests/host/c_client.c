/* A C99 client of the C ABI (include/tbh.h): the MIS network of a triangle
 * written by hand as one tbh_slice, its complexity and peak memory (host-only
 * calls), its BATCHED step count, and -- with --gpu -- its contraction in both
 * modes on device 0.  Prints one line of key=value pairs. */
#include <stdio.h>
#include <string.h>

#include "tbh.h"

int main(int argc, char** argv) {
  /* vertices 0, 1, 2; leaves: the three vertex tensors, then the edges
   * (0,1), (1,2), (0,2) */
  static const int32_t ix_ptr[] = {0, 1, 2, 3, 5, 7, 9};
  static const int32_t ix[] = {0, 1, 2, 0, 1, 1, 2, 0, 2};
  /* post-order tree: node 6 = (0, 3), 7 = (1, 6), 8 = (4, 7), 9 = (2, 8),
   * 10 = (5, 9), the root */
  static const int32_t tree[] = {0, 3, 1, 6, 4, 7, 2, 8, 5, 9};
  tbh_slice s;
  memset(&s, 0, sizeof s);
  s.nv = 3;
  s.weights = NULL;
  s.nleaves = 6;
  s.ix_ptr = ix_ptr;
  s.ix = ix;
  s.tree = tree;
  s.r = 0.0;

  double tc = 0, sc = 0, rwc = 0, peak = 0;
  if (tbh_complexity(&s, &tc, &sc, &rwc) != TBH_OK || tbh_peak_memory(&s, &peak) != TBH_OK) {
    fprintf(stderr, "complexity: %s\n", tbh_last_error());
    return 1;
  }
  const int steps = tbh_plan_steps(&s, TBH_F32, TBH_BATCHED, NULL, 0);
  if (steps < 0) {
    fprintf(stderr, "plan: %s\n", tbh_last_error());
    return 1;
  }
  printf("tc=%.6f sc=%.6f rwc=%.6f peak=%.6f steps=%d", tc, sc, rwc, peak, steps);

  /* a malformed tree (leaf 0 used twice) is refused with TBH_EINVAL */
  static const int32_t bad_tree[] = {0, 0, 1, 6, 4, 7, 2, 8, 5, 9};
  tbh_slice b = s;
  b.tree = bad_tree;
  printf(" malformed=%d", tbh_complexity(&b, &tc, &sc, &rwc));

  if (argc > 1 && strcmp(argv[1], "--gpu") == 0) {
    tbh_ctx* ctx = NULL;
    if (tbh_create(0, (size_t)1 << 28, &ctx) != TBH_OK) {
      fprintf(stderr, "create: %s\n", tbh_last_error());
      return 1;
    }
    float batched = 0, plain = 0;
    int32_t as_int = 0;
    int rc = tbh_contract_slices(ctx, TBH_F32, &s, 1, &batched, TBH_BATCHED);
    if (rc == TBH_OK) rc = tbh_contract_slices(ctx, TBH_F32, &s, 1, &plain, TBH_PLAIN);
    if (rc == TBH_OK) rc = tbh_contract_slices(ctx, TBH_I32, &s, 1, &as_int, TBH_BATCHED);
    if (rc != TBH_OK) {
      fprintf(stderr, "contract: %s\n", tbh_last_error());
      tbh_destroy(ctx);
      return 1;
    }
    printf(" batched=%g plain=%g int32=%d", batched, plain, (int)as_int);
    tbh_destroy(ctx);
  }
  printf("\n");
  fflush(stdout);
  return 0;
}
