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
#include "lm_quat.h"

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

double misorientation(const SymSet &s, const double a[4], const double b[4]) {
  double fa[4], fb[4];
  quat_fundamental(s, a, fa);
  quat_fundamental(s, b, fb);
  return 2 * (std::acos(quat_miso_cos(s, fa, fb))) * kRad2Deg;
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
  quat_normalize(q);
}

double misorientation_deg(const double a[4], const double b[4], int space_group) {
  return misorientation(symmetry_for(space_group), a, b);
}

namespace {

// The candidates the merge sees: at most the 50000 best-scored (ties in list
// order, as glibc's merge-sort qsort keeps them), their unit quaternions and
// the fundamental-region reduction of each.
struct Prepared {
  std::vector<uint64_t> rows;
  std::vector<double> scores;
  std::vector<double> fq;  // nr x 4
  int nr = 0;
};

int prepare(const double *orients, uint64_t n_orients, const uint64_t *rows_in,
            const double *scores_in, int nr, const SymSet &sym, Prepared &P) {
  P.rows.assign(rows_in, rows_in + nr);
  P.scores.assign(scores_in, scores_in + nr);
  for (int i = 0; i < nr; ++i)
    if (P.rows[i] >= n_orients) {
      set_error("merge: candidate row outside the orientation DB");
      return -1;
    }
  if (nr > kMergeCap) {
    std::vector<int> ord(nr);
    for (int i = 0; i < nr; ++i) ord[i] = i;
    std::stable_sort(ord.begin(), ord.end(),
                     [&](int a, int b) { return P.scores[a] > P.scores[b]; });
    std::vector<uint64_t> r2(kMergeCap);
    std::vector<double> s2(kMergeCap);
    for (int i = 0; i < kMergeCap; ++i) {
      r2[i] = P.rows[ord[i]];
      s2[i] = P.scores[ord[i]];
    }
    P.rows.swap(r2);
    P.scores.swap(s2);
    nr = kMergeCap;
  }
  P.nr = nr;
  P.fq.resize((size_t)nr * 4);
  for (int i = 0; i < nr; ++i) {
    double q[4];
    quat_from_matrix(orients + P.rows[i] * 9, q);
    quat_fundamental(sym, q, &P.fq[(size_t)i * 4]);
  }
  return 0;
}

// The clustering test of the reference, (float)(2 acos(c) 180/pi) <= max_angle
// on the clamped cosine c in [0, 1], is monotone in c; returns the smallest
// double c that passes (so "c >= t" is the same test for every c), -1 when
// every c passes, 2 when none does, or NaN if the test is not monotone near
// the boundary (then the caller keeps the float matrix).
double cos_threshold(double max_angle) {
  auto pass = [&](double c) { return (float)(2 * (std::acos(c)) * kRad2Deg) <= max_angle; };
  if (!pass(1.0)) return 2.0;
  if (pass(0.0)) return -1.0;
  uint64_t lo, hi;  // bit patterns: pass(lo) false, pass(hi) true
  const double zero = 0.0, one = 1.0;
  std::memcpy(&lo, &zero, 8);
  std::memcpy(&hi, &one, 8);
  while (hi - lo > 1) {
    const uint64_t mid = lo + (hi - lo) / 2;
    double c;
    std::memcpy(&c, &mid, 8);
    (pass(c) ? hi : lo) = mid;
  }
  // every double within 2^16 representable steps on either side must agree
  for (uint64_t d = 0; d < (1u << 16); ++d) {
    double c;
    if (hi + d <= ((uint64_t)0x3ff0000000000000ull)) {
      const uint64_t b = hi + d;
      std::memcpy(&c, &b, 8);
      if (!pass(c)) return NAN;
    }
    if (lo >= d) {
      const uint64_t b = lo - d;
      std::memcpy(&c, &b, 8);
      if (pass(c)) return NAN;
    }
  }
  double t;
  std::memcpy(&t, &hi, 8);
  return t;
}

// Greedy clustering in list order over a pass test given as bit rows:
// bit l of row i (l > i) set when candidate l is within max_angle of i.
int greedy_bits(const Prepared &P, const std::vector<uint64_t> &bits, int W, uint64_t *best_row,
                double *best_score, int *cluster_size) {
  std::vector<uint64_t> done((size_t)W, 0);
  int k = 0;
  for (int i = 0; i < P.nr; ++i) {
    if ((done[i >> 6] >> (i & 63)) & 1) continue;
    done[i >> 6] |= 1ull << (i & 63);
    int size = 1;
    uint64_t br = P.rows[i];
    double bs = P.scores[i];
    const uint64_t *row = bits.data() + (size_t)i * W;
    for (int wd = (i + 1) >> 6; wd < W; ++wd) {
      uint64_t m = row[wd] & ~done[wd];
      done[wd] |= m;
      while (m) {
        const int l = wd * 64 + __builtin_ctzll(m);
        m &= m - 1;
        ++size;
        if (P.scores[l] > bs) {
          bs = P.scores[l];
          br = P.rows[l];
        }
      }
    }
    best_row[k] = br;
    best_score[k] = bs;
    cluster_size[k] = size;
    ++k;
  }
  return k;
}

}  // namespace

int merge_candidates(const double *orients, uint64_t n_orients, const uint64_t *rows_in,
                     const double *scores_in, int nr, double max_angle, int space_group,
                     int threads, uint64_t *best_row, double *best_score, int *cluster_size) {
  const SymSet sym = symmetry_for(space_group);
  Prepared P;
  if (prepare(orients, n_orients, rows_in, scores_in, nr, sym, P)) return -1;
  nr = P.nr;
  // pairwise misorientation (float, as the reference stores it), each pair
  // from the two candidates' fundamental-region reductions
  std::vector<float> miso((size_t)nr * (nr > 0 ? nr - 1 : 0) / 2);
  parallel_for(nr, threads, [&](int i) {
    for (int j = 0; j < i; ++j)
      miso[(size_t)i * (i - 1) / 2 + j] = (float)(
          2 * (std::acos(quat_miso_cos(sym, &P.fq[(size_t)i * 4], &P.fq[(size_t)j * 4]))) *
          kRad2Deg);
  });
  std::vector<int> done(nr, 0);
  int k = 0;
  for (int i = 0; i < nr; ++i) {
    if (done[i]) continue;
    done[i] = 1;
    uint64_t br = P.rows[i];
    double bs = P.scores[i];
    for (int l = i + 1; l < nr; ++l) {
      if (done[l] > 0) continue;
      if (miso[(size_t)l * (l - 1) / 2 + i] <= max_angle) {
        done[l] = 1;
        done[i]++;
        if (P.scores[l] > bs) {
          bs = P.scores[l];
          br = P.rows[l];
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

int merge_candidates_dev(hipStream_t st, const double *orients, uint64_t n_orients,
                         const uint64_t *rows_in, const double *scores_in, int nr,
                         double max_angle, int space_group, uint64_t *best_row,
                         double *best_score, int *cluster_size) {
  const SymSet sym = symmetry_for(space_group);
  const double t = cos_threshold(max_angle);
  if (std::isnan(t))  // keep the reference's float test on the host
    return merge_candidates(orients, n_orients, rows_in, scores_in, nr, max_angle, space_group,
                            (int)std::thread::hardware_concurrency(), best_row, best_score,
                            cluster_size);
  Prepared P;
  if (prepare(orients, n_orients, rows_in, scores_in, nr, sym, P)) return LM_ERR_INVALID;
  if (P.nr == 0) return 0;
  const int W = (P.nr + 63) / 64;
  std::vector<uint64_t> bits((size_t)P.nr * W);
  const int rc = merge_bits_dev(st, P.fq.data(), P.nr, sym, t, bits.data());
  if (rc != LM_OK) return rc;
  return greedy_bits(P, bits, W, best_row, best_score, cluster_size);
}

}  // namespace lm
