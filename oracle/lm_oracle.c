/*
 * lm_oracle.c - CPU oracle for the LaueMatching indexing hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path, never the product.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it.  It restates the reference LaueMatchingCPU behaviour function by
 * function; each function cites the reference file:line it follows
 * (paths relative to the AdvancedPhotonSource/LaueMatching repository).
 *
 * Build: gcc -O3 -fPIC -fopenmp -ffp-contract=off (see oracle/Makefile).  The
 * reference is compiled by gcc for baseline x86-64 (SSE2, no FMA), so every
 * double operation here is a single IEEE-rounded op in source order.
 */
#define _GNU_SOURCE
#include "lm_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* src/LaueMatchingHeaders.h:31-42 */
#define LMO_DEG2RAD 0.0174532925199433
#define LMO_HC_KEVNM 1.2398419739
#define LMO_EPS 1E-12

/* src/LaueMatchingCPU.c:224-239.  Rodrigues formula; each entry keeps the
 * reference's association: ((1-c)*a)*b +/- s*v. */
void lmo_detector_rotT(const double r[3], double rotT[3][3]) {
  double ang = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
  double v0 = r[0] / ang, v1 = r[1] / ang, v2 = r[2] / ang;
  double c = cos(ang), s = sin(ang), omc = 1 - cos(ang);
  double R[3][3];
  R[0][0] = c + omc * (v0 * v0);
  R[0][1] = omc * v0 * v1 - s * v2;
  R[0][2] = omc * v0 * v2 + s * v1;
  R[1][0] = omc * v1 * v0 + s * v2;
  R[1][1] = c + omc * (v1 * v1);
  R[1][2] = omc * v1 * v2 - s * v0;
  R[2][0] = omc * v2 * v0 - s * v1;
  R[2][1] = omc * v2 * v1 + s * v0;
  R[2][2] = c + omc * (v2 * v2);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) rotT[i][j] = R[j][i];
}

static double zero_small(double v) { return fabs(v) < LMO_EPS ? 0 : v; }

/* src/LaueMatchingHeaders.h:379-436 (real-space basis, then reciprocal basis
 * as cross products scaled by 2*pi/V; rhombohedral settings for the listed
 * space groups). */
void lmo_calc_recip(const double lat[6], int sgnum, double recip[3][3]) {
  const double a = lat[0], b = lat[1], cc = lat[2];
  int rh = (sgnum == 146 || sgnum == 148 || sgnum == 155 || sgnum == 160 || sgnum == 161 ||
            sgnum == 166 || sgnum == 167);
  double ca = cos(lat[3] * LMO_DEG2RAD), cb = cos(lat[4] * LMO_DEG2RAD);
  double cg = cos(lat[5] * LMO_DEG2RAD), sg = sin(lat[5] * LMO_DEG2RAD);
  double phi = sqrt(1.0 - ca * ca - cb * cb - cg * cg + 2 * ca * cb * cg);
  double vol = a * b * cc * phi;
  double pv = (2 * M_PI) / vol;
  double A[3], B[3], C[3];
  if (!rh) {
    A[0] = zero_small(a); A[1] = 0.0; A[2] = 0.0;
    B[0] = b * cg; B[1] = zero_small(b * sg); B[2] = 0;
    C[0] = cc * cb; C[1] = cc * (ca - cb * cg) / sg; C[2] = zero_small(cc * phi / sg);
  } else {
    double p = sqrt(1.0 + 2 * ca), q = sqrt(1.0 - ca);
    double u = (a / 3.0) * (p - q), w = (a / 3.0) * (p + 2 * q);
    A[0] = w; A[1] = u; A[2] = u;
    B[0] = u; B[1] = w; B[2] = u;
    C[0] = u; C[1] = u; C[2] = w;
  }
  recip[0][0] = zero_small((B[1] * C[2] - B[2] * C[1]) * pv);
  recip[1][0] = zero_small((B[2] * C[0] - B[0] * C[2]) * pv);
  recip[2][0] = zero_small((B[0] * C[1] - B[1] * C[0]) * pv);
  recip[0][1] = zero_small((C[1] * A[2] - C[2] * A[1]) * pv);
  recip[1][1] = zero_small((C[2] * A[0] - C[0] * A[2]) * pv);
  recip[2][1] = zero_small((C[0] * A[1] - C[1] * A[0]) * pv);
  recip[0][2] = zero_small((A[1] * B[2] - A[2] * B[1]) * pv);
  recip[1][2] = zero_small((A[2] * B[0] - A[0] * B[2]) * pv);
  recip[2][2] = zero_small((A[0] * B[1] - A[1] * B[0]) * pv);
}

/* src/LaueMatchingHeaders.h:349-357 */
void lmo_matmul33(const double m[3][3], const double n[3][3], double res[3][3]) {
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++) res[r][c] = m[r][0] * n[0][c] + m[r][1] * n[1][c] + m[r][2] * n[2][c];
}

/* src/LaueMatchingHeaders.h:361-369: accumulate from zero, column by column. */
static void matvec(const double m[3][3], const double v[3], double r[3]) {
  for (int i = 0; i < 3; i++) {
    double acc = 0;
    acc += m[i][0] * v[0];
    acc += m[i][1] * v[1];
    acc += m[i][2] * v[2];
    r[i] = acc;
  }
}

/* Pixel index from a detector coordinate: the reference truncates
 * (int)(t) and keeps 0 <= ipx <= N-1 (src/LaueMatchingCPU.c:627-632), which
 * for every double t (x86 cvttsd2si gives INT_MIN out of range / NaN) is
 * exactly -1 < t < N. */
static int to_pixel(double t, int npx, int *ip) {
  if (!(t > -1.0 && t < (double)npx)) return 0;
  *ip = (int)t;
  return 1;
}

/* src/LaueMatchingCPU.c:604-636 (one hkl of one orientation). */
int lmo_project(const double M[3][3], const int hkli[3], const lmo_geom *g, int *ipx,
                int *ipy, double qhat_out[3], double *E_out) {
  double hkl[3] = {(double)hkli[0], (double)hkli[1], (double)hkli[2]};
  double q[3];
  matvec(M, hkl, q);
  double qlen = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
  if (qlen == 0) return 0;
  double qh[3] = {q[0] / qlen, q[1] / qlen, q[2] / qlen};
  double twodot = 2 * qh[2];
  double kf[3] = {0 - twodot * qh[0], 0 - twodot * qh[1], 1.0 - twodot * qh[2]};
  double xyz[3];
  matvec(g->rotT, kf, xyz);
  if (qhat_out) {
    qhat_out[0] = qh[0];
    qhat_out[1] = qh[1];
    qhat_out[2] = qh[2];
  }
  if (xyz[2] <= 0) return 0;
  double X = xyz[0] * g->P[2] / xyz[2];
  double Y = xyz[1] * g->P[2] / xyz[2];
  double xp = X - g->P[0], yp = Y - g->P[1];
  if (!to_pixel((xp / g->pxX) + (0.5 * (g->nrPxX - 1)), g->nrPxX, ipx)) return 0;
  if (!to_pixel((yp / g->pxY) + (0.5 * (g->nrPxY - 1)), g->nrPxY, ipy)) return 0;
  double sinTheta = -qh[2];
  double E = LMO_HC_KEVNM * qlen / (4 * M_PI * sinTheta);
  if (E_out) *E_out = E;
  if (E < g->Elo || E > g->Ehi) return 0;
  return 1;
}

/* src/LaueMatchingCPU.c:591-667 (DoFwd=1 spot-list assembly) plus the
 * float64-image scoring of that branch (:647-653) and the hit test
 * (:683-689).  Dedup: a pixel already taken by an earlier hkl of the same
 * orientation is skipped (the reference uses a per-thread W*H mask, :637-641;
 * scanning the orientation's own list is the same predicate). */
long lmo_forward(const double *orients, size_t n, const int *hkls, int nhkls,
                 const lmo_geom *g, int M, uint16_t *table, const double *img,
                 int minNrSpots, double minIntensity, double *matched, int nthreads) {
  const size_t rec = (size_t)(1 + 2 * M);
  long hits = 0;
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(static) num_threads(nthreads) reduction(+ : hits)
  for (size_t o = 0; o < n; o++) {
    uint16_t *r = table + o * rec;
    memset(r, 0, rec * sizeof(uint16_t));
    double tO[3][3], Mx[3][3];
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) tO[i][j] = orients[o * 9 + (size_t)i * 3 + (size_t)j];
    lmo_matmul33(tO, g->recip, Mx);
    int ns = 0, nSpots = 0;
    double totInt = 0;
    for (int h = 0; h < nhkls && ns < M; h++) {
      int ipx, ipy;
      if (!lmo_project(Mx, hkls + 3 * h, g, &ipx, &ipy, NULL, NULL)) continue;
      int dup = 0;
      for (int s = 0; s < ns; s++)
        if (r[1 + 2 * s] == (uint16_t)ipx && r[2 + 2 * s] == (uint16_t)ipy) {
          dup = 1;
          break;
        }
      if (dup) continue;
      r[1 + 2 * ns] = (uint16_t)ipx;
      r[2 + 2 * ns] = (uint16_t)ipy;
      if (img) {
        double v = img[(size_t)ipy * g->nrPxX + ipx];
        if (v > 0) {
          totInt += v;
          nSpots++;
        }
      }
      ns++;
    }
    r[0] = (uint16_t)ns;
    if (img && matched) {
      if (nSpots >= minNrSpots && totInt >= minIntensity) {
        matched[o] = totInt * sqrt((double)nSpots);
        hits++;
      } else {
        matched[o] = 0;
      }
    }
  }
  return img ? hits : (long)n;
}

/* src/LaueMatchingCPU.c:355-377 */
double lmo_quantize_u8(const double *img, size_t npx, uint8_t *out) {
  double mx = 0;
  for (size_t i = 0; i < npx; i++)
    if (img[i] > mx) mx = img[i];
  double scale = (mx > 0) ? mx / 255.0 : 1.0;
  double inv = (mx > 0) ? 255.0 / mx : 0.0;
  for (size_t i = 0; i < npx; i++) {
    if (img[i] > 0) {
      double v = img[i] * inv;
      out[i] = (v >= 255.0) ? 255 : (v < 1.0) ? 1 : (uint8_t)v;
    } else {
      out[i] = 0;
    }
  }
  return scale;
}

/* src/LaueMatchingCPU.c:668-689 (DoFwd=0 branch and hit test). */
long lmo_compare_u8(const uint16_t *table, size_t n, int M, const uint8_t *img, int nrPxX,
                    double imScale, int minNrSpots, double minIntensity, double *matched,
                    int32_t *nspots_out, double *totint_out, int nthreads) {
  const size_t rec = (size_t)(1 + 2 * M);
  long hits = 0;
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(static) num_threads(nthreads) reduction(+ : hits)
  for (size_t o = 0; o < n; o++) {
    const uint16_t *r = table + o * rec;
    int ns = (int)r[0];
    int nSpots = 0;
    double totInt = 0;
    for (int s = 0; s < ns; s++) {
      int ipx = r[1 + 2 * s], ipy = r[2 + 2 * s];
      uint8_t raw = img[(size_t)ipy * nrPxX + ipx];
      if (raw > 0) {
        totInt += (double)raw * imScale;
        nSpots++;
      }
    }
    if (nspots_out) nspots_out[o] = nSpots;
    if (totint_out) totint_out[o] = totInt;
    if (nSpots >= minNrSpots && totInt >= minIntensity) {
      matched[o] = totInt * sqrt((double)nSpots);
      hits++;
    } else {
      matched[o] = 0;
    }
  }
  return hits;
}

/* src/LaueMatchingCPU.c:772-788 */
long lmo_collect(const double *matched, size_t n, uint64_t *rowNrs, double *mA, long cap) {
  long k = 0;
  for (size_t o = 0; o < n; o++) {
    if (matched[o] != 0) {
      if (k < cap) {
        if (rowNrs) rowNrs[k] = o;
        if (mA) mA[k] = matched[o];
      }
      k++;
    }
  }
  return k;
}
