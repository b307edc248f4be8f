/*
 * lm_median_oracle.c - CPU oracle of the median-filter background.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the library's device median,
 * never the product.  Only tests/ and bench.py's check legs may load it.
 *
 * compute_background (laue_index/preprocess.py:43-64) without diplib calls
 * scipy.ndimage.median_filter(bg, size=filter_radius) once per pass
 * (:59-63).  That filter (mode 'reflect', origin 0) sets each pixel to the
 * element of 0-based rank s*s/2 among the s x s window whose first row and
 * column lie s/2 before the pixel; coordinates outside the frame are
 * mirrored with period 2n (d c b a | a b c d | d c b a), repeatedly for
 * windows wider than the frame.  (scipy 1.15 returns exactly that while
 * s < 8n on each axis; from s = 8n on it reads past its line buffers and
 * returns values no window holds.  This oracle, like the device, keeps the
 * periodic extension for every s up to 255.)
 *
 * This file restates that definition on dense ranks (the caller maps the
 * frame's values to 0..D-1 in increasing order and back; a rank order
 * statistic commutes with that map): along each output row, a running
 * histogram of the window's ranks (Huang's method: s ranks leave and s enter
 * per step) with 64-rank coarse counts, and the rank-s*s/2 element found by
 * moving from the previous one.  Rows are independent (OpenMP over rows).
 * It is pinned in the tests against scipy's median_filter and against a
 * sort of every window.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "lm_oracle.h"

/* scipy 'reflect' index: period 2n, i in [-inf, inf) -> [0, n) */
static int mirror(long i, int n) {
  const long p = 2L * n;
  long r = i % p;
  if (r < 0) r += p;
  return (int)(r < n ? r : p - 1 - r);
}

int lmo_median_ranks(const uint32_t *rk, int W, int H, int s, uint32_t D, uint32_t *out) {
  if (W < 1 || H < 1 || s < 1 || s > 255 || D < 1) return -1;
  const int lo = s / 2;
  const long k = ((long)s * s) / 2;
  const uint32_t ncoarse = (D + 63u) / 64u;
  int fail = 0;
#pragma omp parallel
  {
    uint16_t *fine = (uint16_t *)calloc(D, sizeof(uint16_t));
    uint32_t *coarse = (uint32_t *)calloc(ncoarse, sizeof(uint32_t));
    int *rows = (int *)malloc((size_t)s * sizeof(int));
    if (!fine || !coarse || !rows) {
#pragma omp atomic write
      fail = 1;
    }
#pragma omp for schedule(dynamic, 4)
    for (int y = 0; y < H; ++y) {
      if (fail) continue;
      for (int t = 0; t < s; ++t) rows[t] = mirror((long)y - lo + t, H);
      uint32_t cur = 0; /* coarse bin of the median */
      long below = 0;   /* window elements in coarse bins < cur */
#define LMO_COLUMN(cx, d)                                      \
  do {                                                         \
    for (int t = 0; t < s; ++t) {                              \
      const uint32_t e = rk[(size_t)rows[t] * W + (cx)];       \
      fine[e] = (uint16_t)(fine[e] + (d));                     \
      coarse[e >> 6] += (uint32_t)(d);                         \
      if ((e >> 6) < cur) below += (d);                        \
    }                                                          \
  } while (0)
      for (int j = 0; j < s; ++j) LMO_COLUMN(mirror((long)j - lo, W), 1);
      for (int x = 0; x < W; ++x) {
        if (x > 0) {
          LMO_COLUMN(mirror((long)x - 1 - lo, W), -1);
          LMO_COLUMN(mirror((long)x - lo + s - 1, W), 1);
        }
        while (below > k) {
          --cur;
          below -= coarse[cur];
        }
        while (below + (long)coarse[cur] <= k) {
          below += coarse[cur];
          ++cur;
        }
        uint32_t r = cur * 64u;
        long acc = below;
        while (acc + (long)fine[r] <= k) acc += fine[r++];
        out[(size_t)y * W + x] = r;
      }
      /* the last window leaves: the histograms are zero again */
      for (int j = 0; j < s; ++j) LMO_COLUMN(mirror((long)W - 1 - lo + j, W), -1);
#undef LMO_COLUMN
    }
    free(fine);
    free(coarse);
    free(rows);
  }
  return fail ? -2 : 0;
}
