// lm_merge.cpp - merge of duplicate candidates (the consumer of the hit
// list), reference mergeDuplicateOrientations (src/LaueMatchingHeaders.h:
// 949-1065) and its quaternion helpers (:150-304).
//
// Steps: keep the 50000 best-scored candidates if there are more; unit
// quaternions of the candidates; pairwise misorientation under the crystal's
// point group (float32 matrix); greedy clustering in list order, each cluster
// taking its best-scored member.  Floating-point steps keep the reference's
// operation order because the clustering tests misorientation <= MaxAngle,
// and the point-group tables carry the reference's 5-decimal values.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <thread>
#include <vector>

#include "lm_internal.h"

namespace lm {
namespace {

constexpr double kRad2Deg = 57.2957795130823;
constexpr int kMergeCap = 50000;  // src/LaueMatchingHeaders.h:953

// Point-group rotations (w, x, y, z) for the space-group ranges triclinic
// (1-2), monoclinic (3-15), orthorhombic (16-74), tetragonal (75-142),
// trigonal (143-167), hexagonal (168-194), cubic (195-230)
// (src/LaueMatchingHeaders.h:83-148, 180-221).  The orthorhombic table keeps
// the reference's second entry (1, 1, 0, 0), normalised at use.
const double kTric[1][4] = {{1, 0, 0, 0}};
const double kMono[2][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}};
const double kOrt[4][4] = {{1, 0, 0, 0}, {1, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}};
const double kTet[8][4] = {{1, 0, 0, 0},           {0.70711, 0, 0, 0.70711},
                           {0, 0, 0, 1},           {0.70711, -0.0, -0.0, -0.70711},
                           {0, 1, 0, 0},           {0, 0, 1, 0},
                           {0, 0.70711, 0.70711, 0}, {0, -0.70711, 0.70711, 0}};
const double kTrig[6][4] = {{1, 0, 0, 0},         {0, 0.86603, -0.5, 0},
                            {0.5, 0, 0, 0.86603},  {0, 0, 1, 0},
                            {0.5, -0.0, -0.0, -0.86603}, {0, 0.86603, 0.5, 0}};
const double kHex[12][4] = {{1, 0, 0, 0},           {0.86603, 0, 0, 0.5},
                            {0.5, 0, 0, 0.86603},    {0, 0, 0, 1},
                            {0.5, -0.0, -0.0, -0.86603}, {0.86603, -0.0, -0.0, -0.5},
                            {0, 1, 0, 0},           {0, 0.86603, 0.5, 0},
                            {0, 0.5, 0.86603, 0},    {0, 0, 1, 0},
                            {0, -0.5, 0.86603, 0},   {0, -0.86603, 0.5, 0}};
const double kCub[24][4] = {
    {1, 0, 0, 0},              {0.70711, 0.70711, 0, 0},   {0, 1, 0, 0},
    {0.70711, -0.70711, 0, 0}, {0.70711, 0, 0.70711, 0},   {0, 0, 1, 0},
    {0.70711, 0, -0.70711, 0}, {0.70711, 0, 0, 0.70711},   {0, 0, 0, 1},
    {0.70711, 0, 0, -0.70711}, {0.5, 0.5, 0.5, 0.5},       {0.5, -0.5, -0.5, -0.5},
    {0.5, -0.5, 0.5, 0.5},     {0.5, 0.5, -0.5, -0.5},     {0.5, 0.5, -0.5, 0.5},
    {0.5, -0.5, 0.5, -0.5},    {0.5, -0.5, -0.5, 0.5},     {0.5, 0.5, 0.5, -0.5},
    {0, 0.70711, 0.70711, 0},  {0, -0.70711, 0.70711, 0},  {0, 0.70711, 0, 0.70711},
    {0, 0.70711, 0, -0.70711}, {0, 0, 0.70711, 0.70711},   {0, 0, 0.70711, -0.70711}};

struct SymSet {
  int n = 1;
  double q[24][4];
};

SymSet symmetry_for(int sg) {
  SymSet s;
  const double(*src)[4];
  if (sg <= 2) { s.n = 1; src = kTric; }
  else if (sg <= 15) { s.n = 2; src = kMono; }
  else if (sg <= 74) { s.n = 4; src = kOrt; }
  else if (sg <= 142) { s.n = 8; src = kTet; }
  else if (sg <= 167) { s.n = 6; src = kTrig; }
  else if (sg <= 194) { s.n = 12; src = kHex; }
  else { s.n = 24; src = kCub; }
  for (int i = 0; i < s.n; ++i)
    for (int j = 0; j < 4; ++j) s.q[i][j] = src[i][j];
  return s;
}

void normalize4(double q[4]) {
  const double nrm = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
  for (int i = 0; i < 4; ++i) q[i] /= nrm;
}

// Q = q (x) r, scalar part made non-negative, normalised.
void qmul(const double q[4], const double r[4], double Q[4]) {
  Q[0] = r[0] * q[0] - r[1] * q[1] - r[2] * q[2] - r[3] * q[3];
  Q[1] = r[1] * q[0] + r[0] * q[1] + r[3] * q[2] - r[2] * q[3];
  Q[2] = r[2] * q[0] + r[0] * q[2] + r[1] * q[3] - r[3] * q[1];
  Q[3] = r[3] * q[0] + r[0] * q[3] + r[2] * q[1] - r[1] * q[2];
  if (Q[0] < 0)
    for (int i = 0; i < 4; ++i) Q[i] = -Q[i];
  normalize4(Q);
}

// Symmetry-equivalent with the largest scalar part (first one on ties).
void fundamental(const SymSet &s, const double in[4], double out[4]) {
  double best[4] = {1, 0, 0, 0}, qt[4];
  double mc = -10000;
  for (int i = 0; i < s.n; ++i) {
    qmul(in, s.q[i], qt);
    if (mc < qt[0]) {
      mc = qt[0];
      std::memcpy(best, qt, sizeof best);
    }
  }
  std::memcpy(out, best, sizeof best);
  normalize4(out);
}

double misorientation(const SymSet &s, const double a[4], const double b[4]) {
  double fa[4], fb[4], inv[4], p[4], m[4];
  fundamental(s, a, fa);
  fundamental(s, b, fb);
  inv[0] = -fa[0];
  inv[1] = fa[1];
  inv[2] = fa[2];
  inv[3] = fa[3];
  qmul(inv, fb, p);
  fundamental(s, p, m);
  if (m[0] > 1) m[0] = 1;
  return 2 * (std::acos(m[0])) * kRad2Deg;
}

template <typename F>
void parallel_for(int n, int threads, F &&f) {
  if (threads <= 1 || n < 64) {
    for (int i = 0; i < n; ++i) f(i);
    return;
  }
  std::vector<std::thread> pool;
  std::atomic<int> next{0};
  for (int t = 0; t < threads; ++t)
    pool.emplace_back([&] {
      for (int i; (i = next.fetch_add(1)) < n;) f(i);
    });
  for (auto &th : pool) th.join();
}

}  // namespace

void quat_from_matrix(const double m[9], double q[4]) {
  const double tr = m[0] + m[4] + m[8];
  if (tr > 0) {
    const double s = 0.5 / std::sqrt(tr + 1.0);
    q[0] = 0.25 / s;
    q[1] = (m[7] - m[5]) * s;
    q[2] = (m[2] - m[6]) * s;
    q[3] = (m[3] - m[1]) * s;
  } else if (m[0] > m[4] && m[0] > m[8]) {
    const double s = 2.0 * std::sqrt(1.0 + m[0] - m[4] - m[8]);
    q[0] = (m[7] - m[5]) / s;
    q[1] = 0.25 * s;
    q[2] = (m[1] + m[3]) / s;
    q[3] = (m[2] + m[6]) / s;
  } else if (m[4] > m[8]) {
    const double s = 2.0 * std::sqrt(1.0 + m[4] - m[0] - m[8]);
    q[0] = (m[2] - m[6]) / s;
    q[1] = (m[1] + m[3]) / s;
    q[2] = 0.25 * s;
    q[3] = (m[5] + m[7]) / s;
  } else {
    const double s = 2.0 * std::sqrt(1.0 + m[8] - m[0] - m[4]);
    q[0] = (m[3] - m[1]) / s;
    q[1] = (m[2] + m[6]) / s;
    q[2] = (m[5] + m[7]) / s;
    q[3] = 0.25 * s;
  }
  if (q[0] < 0)
    for (int i = 0; i < 4; ++i) q[i] = -q[i];
  normalize4(q);
}

double misorientation_deg(const double a[4], const double b[4], int space_group) {
  return misorientation(symmetry_for(space_group), a, b);
}

int merge_candidates(const double *orients, uint64_t n_orients, const uint64_t *rows_in,
                     const double *scores_in, int nr, double max_angle, int space_group,
                     int threads, uint64_t *best_row, double *best_score, int *cluster_size) {
  std::vector<uint64_t> rows(rows_in, rows_in + nr);
  std::vector<double> scores(scores_in, scores_in + nr);
  for (int i = 0; i < nr; ++i)
    if (rows[i] >= n_orients) {
      set_error("merge: candidate row outside the orientation DB");
      return -1;
    }
  if (nr > kMergeCap) {
    // best-scored first, ties in list order (glibc qsort is a merge sort)
    std::vector<int> ord(nr);
    for (int i = 0; i < nr; ++i) ord[i] = i;
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return scores[a] > scores[b]; });
    std::vector<uint64_t> r2(kMergeCap);
    std::vector<double> s2(kMergeCap);
    for (int i = 0; i < kMergeCap; ++i) {
      r2[i] = rows[ord[i]];
      s2[i] = scores[ord[i]];
    }
    rows.swap(r2);
    scores.swap(s2);
    nr = kMergeCap;
  }
  const SymSet sym = symmetry_for(space_group);
  std::vector<double> q((size_t)nr * 4);
  for (int i = 0; i < nr; ++i) quat_from_matrix(orients + rows[i] * 9, &q[(size_t)i * 4]);
  std::vector<float> miso((size_t)nr * (nr > 0 ? nr - 1 : 0) / 2);
  parallel_for(nr, threads, [&](int i) {
    for (int j = 0; j < i; ++j)
      miso[(size_t)i * (i - 1) / 2 + j] =
          (float)misorientation(sym, &q[(size_t)i * 4], &q[(size_t)j * 4]);
  });
  std::vector<int> done(nr, 0);
  int k = 0;
  for (int i = 0; i < nr; ++i) {
    if (done[i]) continue;
    done[i] = 1;
    uint64_t br = rows[i];
    double bs = scores[i];
    for (int l = i + 1; l < nr; ++l) {
      if (done[l] > 0) continue;
      if (miso[(size_t)l * (l - 1) / 2 + i] <= max_angle) {
        done[l] = 1;
        done[i]++;
        if (scores[l] > bs) {
          bs = scores[l];
          br = rows[l];
        }
      }
    }
    best_row[k] = br;
    best_score[k] = bs;
    cluster_size[k] = done[i];
    ++k;
  }
  return k;
}

}  // namespace lm
