/*
 * lm_merge_oracle.c - CPU oracle of the duplicate-orientation merge, the
 * consumer of the hit list.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the library's host and device
 * merges, never the product.  Only tests/ may load it.
 *
 * Restates mergeDuplicateOrientations (src/LaueMatchingHeaders.h:949-1065)
 * with its quaternion helpers (normalizeQuat :157-165, QuaternionProduct
 * :167-180, MakeSymmetries :182-220, BringDownToFundamentalRegion :222-246,
 * GetMisOrientation :248-260, OrientMat2Quat :262-296) and the point-group
 * tables (:83-148), in the reference's operation order, float64 throughout
 * except the float32 misorientation matrix (:1020-1030):
 *
 *   1. more than 50000 candidates: qsort by score, descending, with the
 *      reference comparator (:958-962, :976-997), keep the first 50000;
 *   2. unit quaternion of every candidate's orientation matrix (:1005-1011);
 *   3. misoDist(i, j) = (float)GetMisOrientation(q_i, q_j) for j < i;
 *   4. greedy pass in list order: an unassigned candidate opens a cluster and
 *      takes every later unassigned candidate with misoDist <= maxAngle; the
 *      best member is replaced only on a strictly greater score (:1033-1059).
 *
 * One change of evaluation order, not of values: step 3 evaluates an entry of
 * the matrix when step 4 reads it (the entries a cluster head compares
 * against), instead of filling the whole lower triangle first.  Every entry
 * is a pure function of two quaternions, so each one that is read has the
 * value the reference stores; entries step 4 never reads are not computed.
 * This keeps the oracle usable at the 50000 cap in tests.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "lm_oracle.h"

/* src/LaueMatchingHeaders.h:32 */
#define LMO_RAD2DEG 57.2957795130823
/* src/LaueMatchingHeaders.h:953 */
#define LMO_MERGE_CAP 50000

/* Point-group rotations (w, x, y, z), src/LaueMatchingHeaders.h:83-148. */
static const double T_TRIC[1][4] = {{1.0, 0.0, 0.0, 0.0}};
static const double T_MONO[2][4] = {{1.0, 0.0, 0.0, 0.0}, {0.0, 1.0, 0.0, 0.0}};
static const double T_ORT[4][4] = {
    {1.0, 0.0, 0.0, 0.0}, {1.0, 1.0, 0.0, 0.0}, {0.0, 0.0, 1.0, 0.0}, {0.0, 0.0, 0.0, 1.0}};
static const double T_TET[8][4] = {
    {1.0, 0.0, 0.0, 0.0},          {0.70711, 0.0, 0.0, 0.70711},
    {0.0, 0.0, 0.0, 1.0},          {0.70711, -0.0, -0.0, -0.70711},
    {0.0, 1.0, 0.0, 0.0},          {0.0, 0.0, 1.0, 0.0},
    {0.0, 0.70711, 0.70711, 0.0},  {0.0, -0.70711, 0.70711, 0.0}};
static const double T_TRIG[6][4] = {
    {1.0, 0.0, 0.0, 0.0},         {0.0, 0.86603, -0.5, 0.0},
    {0.5, 0.0, 0.0, 0.86603},     {0.0, 0.0, 1.0, 0.0},
    {0.5, -0.0, -0.0, -0.86603},  {0.0, 0.86603, 0.5, 0.0}};
static const double T_HEX[12][4] = {
    {1.0, 0.0, 0.0, 0.0},         {0.86603, 0.0, 0.0, 0.5},
    {0.5, 0.0, 0.0, 0.86603},     {0.0, 0.0, 0.0, 1.0},
    {0.5, -0.0, -0.0, -0.86603},  {0.86603, -0.0, -0.0, -0.5},
    {0.0, 1.0, 0.0, 0.0},         {0.0, 0.86603, 0.5, 0.0},
    {0.0, 0.5, 0.86603, 0.0},     {0.0, 0.0, 1.0, 0.0},
    {0.0, -0.5, 0.86603, 0.0},    {0.0, -0.86603, 0.5, 0.0}};
static const double T_CUB[24][4] = {
    {1.0, 0.0, 0.0, 0.0},          {0.70711, 0.70711, 0.0, 0.0},
    {0.0, 1.0, 0.0, 0.0},          {0.70711, -0.70711, 0.0, 0.0},
    {0.70711, 0.0, 0.70711, 0.0},  {0.0, 0.0, 1.0, 0.0},
    {0.70711, 0.0, -0.70711, 0.0}, {0.70711, 0.0, 0.0, 0.70711},
    {0.0, 0.0, 0.0, 1.0},          {0.70711, 0.0, 0.0, -0.70711},
    {0.5, 0.5, 0.5, 0.5},          {0.5, -0.5, -0.5, -0.5},
    {0.5, -0.5, 0.5, 0.5},         {0.5, 0.5, -0.5, -0.5},
    {0.5, 0.5, -0.5, 0.5},         {0.5, -0.5, 0.5, -0.5},
    {0.5, -0.5, -0.5, 0.5},        {0.5, 0.5, 0.5, -0.5},
    {0.0, 0.70711, 0.70711, 0.0},  {0.0, -0.70711, 0.70711, 0.0},
    {0.0, 0.70711, 0.0, 0.70711},  {0.0, 0.70711, 0.0, -0.70711},
    {0.0, 0.0, 0.70711, 0.70711},  {0.0, 0.0, 0.70711, -0.70711}};

typedef struct {
  int n;
  const double (*q)[4];
} sym_t;

/* MakeSymmetries, src/LaueMatchingHeaders.h:182-220: the table by space-group
 * range. */
static sym_t sym_for(int sg) {
  sym_t s;
  if (sg <= 2) {
    s.n = 1; s.q = T_TRIC;
  } else if (sg <= 15) {
    s.n = 2; s.q = T_MONO;
  } else if (sg <= 74) {
    s.n = 4; s.q = T_ORT;
  } else if (sg <= 142) {
    s.n = 8; s.q = T_TET;
  } else if (sg <= 167) {
    s.n = 6; s.q = T_TRIG;
  } else if (sg <= 194) {
    s.n = 12; s.q = T_HEX;
  } else {
    s.n = 24; s.q = T_CUB;
  }
  return s;
}

/* normalizeQuat, :157-165 */
static void q_norm(double q[4]) {
  double nrm = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
  q[0] /= nrm;
  q[1] /= nrm;
  q[2] /= nrm;
  q[3] /= nrm;
}

/* QuaternionProduct(q, r, Q), :167-180 */
static void q_prod(const double q[4], const double r[4], double Q[4]) {
  Q[0] = r[0] * q[0] - r[1] * q[1] - r[2] * q[2] - r[3] * q[3];
  Q[1] = r[1] * q[0] + r[0] * q[1] + r[3] * q[2] - r[2] * q[3];
  Q[2] = r[2] * q[0] + r[0] * q[2] + r[1] * q[3] - r[3] * q[1];
  Q[3] = r[3] * q[0] + r[0] * q[3] + r[2] * q[1] - r[1] * q[2];
  if (Q[0] < 0) {
    Q[0] = -Q[0];
    Q[1] = -Q[1];
    Q[2] = -Q[2];
    Q[3] = -Q[3];
  }
  q_norm(Q);
}

/* BringDownToFundamentalRegion, :222-246: the product with the largest
 * scalar part, the first one on ties (strict <), normalised again. */
static void q_fundamental(const sym_t *s, const double in[4], double out[4]) {
  double prods[24][4];
  double maxCos = -10000;
  int best = 0;
  for (int i = 0; i < s->n; i++) {
    double sy[4] = {s->q[i][0], s->q[i][1], s->q[i][2], s->q[i][3]};
    q_prod(in, sy, prods[i]);
    if (maxCos < prods[i][0]) {
      maxCos = prods[i][0];
      best = i;
    }
  }
  for (int k = 0; k < 4; k++) out[k] = prods[best][k];
  q_norm(out);
}

/* GetMisOrientation, :248-260 (degrees). */
static double miso_deg(const sym_t *s, const double q1[4], const double q2[4]) {
  double f1[4], f2[4], inv[4], p[4], m[4];
  q_fundamental(s, q1, f1);
  q_fundamental(s, q2, f2);
  inv[0] = -f1[0];
  inv[1] = f1[1];
  inv[2] = f1[2];
  inv[3] = f1[3];
  q_prod(inv, f2, p);
  q_fundamental(s, p, m);
  if (m[0] > 1) m[0] = 1;
  return 2 * (acos(m[0])) * LMO_RAD2DEG;
}

/* OrientMat2Quat, :262-296 (row-major 3x3; w made non-negative, then
 * normalised). */
void lmo_quat_from_matrix(const double m[9], double q[4]) {
  double tr = m[0] + m[4] + m[8];
  if (tr > 0) {
    double s = 0.5 / sqrt(tr + 1.0);
    q[0] = 0.25 / s;
    q[1] = (m[7] - m[5]) * s;
    q[2] = (m[2] - m[6]) * s;
    q[3] = (m[3] - m[1]) * s;
  } else if (m[0] > m[4] && m[0] > m[8]) {
    double s = 2.0 * sqrt(1.0 + m[0] - m[4] - m[8]);
    q[0] = (m[7] - m[5]) / s;
    q[1] = 0.25 * s;
    q[2] = (m[1] + m[3]) / s;
    q[3] = (m[2] + m[6]) / s;
  } else if (m[4] > m[8]) {
    double s = 2.0 * sqrt(1.0 + m[4] - m[0] - m[8]);
    q[0] = (m[2] - m[6]) / s;
    q[1] = (m[1] + m[3]) / s;
    q[2] = 0.25 * s;
    q[3] = (m[5] + m[7]) / s;
  } else {
    double s = 2.0 * sqrt(1.0 + m[8] - m[0] - m[4]);
    q[0] = (m[3] - m[1]) / s;
    q[1] = (m[2] + m[6]) / s;
    q[2] = (m[5] + m[7]) / s;
    q[3] = 0.25 * s;
  }
  if (q[0] < 0) {
    q[0] = -q[0];
    q[1] = -q[1];
    q[2] = -q[2];
    q[3] = -q[3];
  }
  q_norm(q);
}

double lmo_misorientation_deg(const double q1[4], const double q2[4], int sg) {
  sym_t s = sym_for(sg);
  return miso_deg(&s, q1, q2);
}

/* The reference's score-descending comparator, :958-962. */
typedef struct {
  double score;
  size_t row;
} lmo_score_row;

static int cmp_desc(const void *a, const void *b) {
  double sa = ((const lmo_score_row *)a)->score;
  double sb = ((const lmo_score_row *)b)->score;
  return (sa < sb) - (sa > sb);
}

int lmo_merge(const double *orients, const uint64_t *rows_in, const double *scores_in, int nr,
              double maxAngle, int sg, int nthreads, uint64_t *best_row, double *best_score,
              int32_t *cluster_size) {
  sym_t s = sym_for(sg);
  size_t *rowNrs = malloc((size_t)(nr > 0 ? nr : 1) * sizeof(size_t));
  double *sc = malloc((size_t)(nr > 0 ? nr : 1) * sizeof(double));
  for (int i = 0; i < nr; i++) {
    rowNrs[i] = (size_t)rows_in[i];
    sc[i] = scores_in[i];
  }
  /* :976-997 */
  if (nr > LMO_MERGE_CAP) {
    lmo_score_row *sr = malloc((size_t)nr * sizeof(lmo_score_row));
    for (int i = 0; i < nr; i++) {
      sr[i].score = sc[i];
      sr[i].row = rowNrs[i];
    }
    qsort(sr, (size_t)nr, sizeof(lmo_score_row), cmp_desc);
    for (int i = 0; i < LMO_MERGE_CAP; i++) {
      sc[i] = sr[i].score;
      rowNrs[i] = sr[i].row;
    }
    free(sr);
    nr = LMO_MERGE_CAP;
  }
  /* :1005-1011 */
  double *quats = malloc((size_t)(nr > 0 ? nr : 1) * 4 * sizeof(double));
  for (int i = 0; i < nr; i++) lmo_quat_from_matrix(orients + rowNrs[i] * 9, quats + 4 * (size_t)i);
  /* :1033-1059; misoDist[l][gi] (l > gi) = (float)GetMisOrientation(q_l, q_gi) */
  int *done = calloc((size_t)(nr > 0 ? nr : 1), sizeof(int));
  float *rowd = malloc((size_t)(nr > 0 ? nr : 1) * sizeof(float));
  int iterNr = 0;
  for (int gi = 0; gi < nr; gi++) {
    if (done[gi] != 0) continue;
    done[gi] = 1;
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 64)
    for (int l = gi + 1; l < nr; l++)
      if (done[l] == 0) rowd[l] = (float)miso_deg(&s, quats + 4 * (size_t)l, quats + 4 * (size_t)gi);
    /* the reference keeps the best row in an int (:1037) */
    int bestSol = (int)rowNrs[gi];
    double bestIntensity = sc[gi];
    for (int l = gi + 1; l < nr; l++) {
      if (done[l] > 0) continue;
      if (rowd[l] <= maxAngle) {
        done[l] = 1;
        done[gi]++;
        if (sc[l] > bestIntensity) {
          bestIntensity = sc[l];
          bestSol = (int)rowNrs[l];
        }
      }
    }
    best_row[iterNr] = (uint64_t)bestSol;
    best_score[iterNr] = bestIntensity;
    cluster_size[iterNr] = done[gi];
    iterNr++;
  }
  free(rowd);
  free(done);
  free(quats);
  free(sc);
  free(rowNrs);
  return iterNr;
}
