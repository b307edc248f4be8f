// lm_postproc.cpp - see lm_postproc.h.
//
// The merge and the quaternion / misorientation maths come from
// liblauematch (lm_merge_candidates, lm_quat_from_matrix,
// lm_misorientation_deg).  Floating-point steps keep the reference's
// operation order because their results feed threshold tests
// (nrSps >= MinNrSpots) and the printed columns.
#include "lm_postproc.h"

#include <omp.h>

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

constexpr double kHc = 1.2398419739;  // src/LaueMatchingHeaders.h:41
constexpr double kEps = 1E-12;

double angle_from(double s, double c) {
  return (s >= 0.0) ? std::acos(c) : 2.0 * M_PI - std::acos(c);
}

// Matrix -> (psi, phi, theta), src/LaueMatchingHeaders.h:306-330.
void mat_to_euler(const double m[9], double e[3]) {
  double psi, theta;
  const double phi = (std::fabs(m[8] - 1.0) < kEps) ? 0 : std::acos(m[8]);
  const double sph = std::sin(phi);
  if (std::fabs(sph) < kEps) {
    psi = 0.0;
    theta = (std::fabs(m[8] - 1.0) < kEps) ? angle_from(m[3], m[0]) : angle_from(-m[3], m[0]);
  } else {
    psi = (std::fabs(-m[5] / sph) <= 1.0) ? angle_from(m[2] / sph, -m[5] / sph)
                                           : angle_from(m[2] / sph, 1);
    theta = (std::fabs(m[7] / sph) <= 1.0) ? angle_from(m[6] / sph, m[7] / sph)
                                           : angle_from(m[6] / sph, 1);
  }
  e[0] = psi;
  e[1] = phi;
  e[2] = theta;
}

// (psi, phi, theta) -> matrix, src/LaueMatchingHeaders.h:332-347.
void euler_to_mat(const double e[3], double m[9]) {
  const double cps = std::cos(e[0]), cph = std::cos(e[1]), cth = std::cos(e[2]);
  const double sps = std::sin(e[0]), sph = std::sin(e[1]), sth = std::sin(e[2]);
  m[0] = cth * cps - sth * cph * sps;
  m[1] = -cth * cph * sps - sth * cps;
  m[2] = sph * sps;
  m[3] = cth * sps + sth * cph * cps;
  m[4] = cth * cph * cps - sth * sps;
  m[5] = -sph * cps;
  m[6] = sth * sph;
  m[7] = cth * sph;
  m[8] = cph;
}

void matmul3(const double a[9], const double b[9], double c[9]) {
  for (int r = 0; r < 3; ++r)
    for (int k = 0; k < 3; ++k)
      c[r * 3 + k] = a[r * 3 + 0] * b[0 * 3 + k] + a[r * 3 + 1] * b[1 * 3 + k] +
                     a[r * 3 + 2] * b[2 * 3 + k];
}

struct Geo {
  double recip[9], rotT[9], P[3], pxX, pxY, Elo, Ehi;
  int W, H;
};

// Continuous detector position of one reflection (the projection of
// src/LaueMatchingHeaders.h:458-486, pixels kept as doubles).
bool project_f(const double OM[9], const int32_t *hkl, const Geo &g, double *px, double *py,
               double qh[3]) {
  const double h = hkl[0], k = hkl[1], l = hkl[2];
  double q[3];
  for (int i = 0; i < 3; ++i) {
    double acc = 0;
    acc += OM[i * 3 + 0] * h;
    acc += OM[i * 3 + 1] * k;
    acc += OM[i * 3 + 2] * l;
    q[i] = acc;
  }
  const double qlen = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
  if (qlen == 0) return false;
  qh[0] = q[0] / qlen;
  qh[1] = q[1] / qlen;
  qh[2] = q[2] / qlen;
  const double d = qh[2];
  const double kf[3] = {0 - 2 * d * qh[0], 0 - 2 * d * qh[1], 1.0 - 2 * d * qh[2]};
  double xyz[3];
  for (int i = 0; i < 3; ++i) {
    double acc = 0;
    acc += g.rotT[i * 3 + 0] * kf[0];
    acc += g.rotT[i * 3 + 1] * kf[1];
    acc += g.rotT[i * 3 + 2] * kf[2];
    xyz[i] = acc;
  }
  if (xyz[2] <= 0) return false;
  const double X = xyz[0] * g.P[2] / xyz[2], Y = xyz[1] * g.P[2] / xyz[2];
  *px = ((X - g.P[0]) / g.pxX) + (0.5 * (g.W - 1));
  if (*px < 0 || *px > (g.W - 1)) return false;
  *py = ((Y - g.P[1]) / g.pxY) + (0.5 * (g.H - 1));
  if (*py < 0 || *py > (g.H - 1)) return false;
  const double E = kHc * qlen / (4 * M_PI * -qh[2]);
  return !(E < g.Elo || E > g.Ehi);
}

// Overlap of one orientation with the float image: q-hat dedup within 1e-6,
// at most maxSpots listed, result nrPos * sqrt(sum of positive pixels)
// (src/LaueMatchingHeaders.h:440-501, 807-947).  With `lines`, the positive
// spots are formatted as spot-file rows.
double overlap(const float *img, const double OM[9], const int32_t *hkls, const int *idx,
               int nidx, const Geo &g, int maxSpots, int *nrPos, int *nSim, std::string *lines,
               int imageNr, int grainNr) {
  std::vector<double> qs;
  qs.reserve(3 * (size_t)maxSpots);
  double result = 0;
  int pos = 0, sp = 0;
  char buf[256];
  for (int v = 0; v < nidx; ++v) {
    const int32_t *h = hkls + 3 * (idx ? idx[v] : v);
    double px, py, qh[3];
    if (!project_f(OM, h, g, &px, &py, qh)) continue;
    bool dup = false;
    for (int s = 0; s < sp; ++s)
      if ((std::fabs(qh[0] - qs[3 * s]) * 100000 < 0.1) &&
          (std::fabs(qh[1] - qs[3 * s + 1]) * 100000 < 0.1) &&
          (std::fabs(qh[2] - qs[3 * s + 2]) * 100000 < 0.1)) {
        dup = true;
        break;
      }
    if (dup) continue;
    qs.push_back(qh[0]);
    qs.push_back(qh[1]);
    qs.push_back(qh[2]);
    const float val = img[(size_t)((size_t)py * g.W + (size_t)px)];
    if (val > 0) {
      if (lines) {
        if (imageNr > 0)
          snprintf(buf, sizeof buf, "%d\t%d\t%d\t%d\t%d\t%d\t%5d\t%5d\t%lf\t%lf\t%lf\t%lf\n",
                   imageNr, grainNr, sp, h[0], h[1], h[2], (int)px, (int)py, qh[0], qh[1], qh[2],
                   (double)val);
        else
          snprintf(buf, sizeof buf, "%d\t%d\t%d\t%d\t%d\t%5d\t%5d\t%lf\t%lf\t%lf\t%lf\n",
                   grainNr, sp, h[0], h[1], h[2], (int)px, (int)py, qh[0], qh[1], qh[2],
                   (double)val);
        *lines += buf;
      }
      result += val;
      ++pos;
    }
    ++sp;
    if (sp == maxSpots) break;
  }
  if (nrPos) *nrPos = pos;
  if (nSim) *nSim = sp;
  return pos * std::sqrt(result);
}

Geo geo_from(const LmParams &p) {
  Geo g{};
  lm_geom lg;
  lm_geom_init(p.lattice, p.space_group, p.P, p.R, p.pxX, p.pxY, p.nrPxX, p.nrPxY, p.Elo, p.Ehi,
               &lg);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      g.recip[i * 3 + j] = lg.recip[i][j];
      g.rotT[i * 3 + j] = lg.rotT[i][j];
    }
  for (int i = 0; i < 3; ++i) g.P[i] = p.P[i];
  g.pxX = p.pxX;
  g.pxY = p.pxY;
  g.W = p.nrPxX;
  g.H = p.nrPxY;
  g.Elo = p.Elo;
  g.Ehi = p.Ehi;
  return g;
}

int emit(const PostInput &in, std::string &solOut, std::string &spotsOut) {
  const LmParams &p = *in.params;
  const lm_hits &h = *in.hits;
  // The merge sees nrResults entries: the collected hits, then zeros for
  // predicate passes with a zero score (src/LaueMatchingCPU.c:773-788).
  const int nr = (int)in.nr_results;
  std::vector<uint64_t> rows(nr, 0);
  std::vector<double> scores(nr, 0.0);
  for (uint64_t i = 0; i < h.count && (int)i < nr; ++i) {
    rows[i] = h.ids[i];
    scores[i] = h.score[i];
  }
  const int cap = nr < 50000 ? nr : 50000;
  std::vector<uint64_t> best(cap + 1);
  std::vector<double> bestScore(cap + 1);
  std::vector<int32_t> csize(cap + 1);
  int32_t nsol = 0;
  // the pair tests are O(nr^2): from a few thousand candidates on, the device
  // does them (same clusters)
  const bool on_device = in.ctx && nr >= 2048;
  if (nr > 0 &&
      (on_device
           ? lm_merge_candidates_dev(in.ctx, in.orients, in.n_orients, rows.data(), scores.data(),
                                     nr, p.max_angle, p.space_group, best.data(),
                                     bestScore.data(), csize.data(), &nsol)
           : lm_merge_candidates(in.orients, in.n_orients, rows.data(), scores.data(), nr,
                                 p.max_angle, p.space_group, in.num_procs, best.data(),
                                 bestScore.data(), csize.data(), &nsol)) != LM_OK) {
    fprintf(stderr, "merge failed: %s\n", lm_last_error());
    return -1;
  }
  const Geo g = geo_from(p);
  const size_t npx = (size_t)p.nrPxX * p.nrPxY;
  std::vector<float> fimg;
  const float *img = in.image_f;
  if (!img) {
    fimg.resize(npx);
    for (size_t i = 0; i < npx; ++i) fimg[i] = (float)in.image[i];
    img = fimg.data();
  }
  const int maxSpotsFit = p.max_spots * 3;  // src/LaueMatchingCPU.c:740
  std::vector<std::string> solLines(nsol), spotLines(nsol);
#pragma omp parallel for num_threads(in.num_procs) schedule(dynamic)
  for (int s = 0; s < nsol; ++s) {
    double OMbest[9], e[3], OMfit[9], q1[4], q2[4], OF[9];
    for (int k = 0; k < 9; ++k) OMbest[k] = in.orients[best[s] * 9 + k];
    mat_to_euler(OMbest, e);
    euler_to_mat(e, OMfit);
    matmul3(OMfit, g.recip, OF);
    // objective at the starting point of the refinement: the overlap over the
    // hkls that land on the detector there (src/LaueMatchingHeaders.h:1099-1108)
    std::vector<int> valid;
    valid.reserve(in.nhkls);
    for (int hh = 0; hh < in.nhkls; ++hh) {
      double px, py, qh[3];
      if (project_f(OF, in.hkls + 3 * hh, g, &px, &py, qh)) valid.push_back(hh);
    }
    const double mv = overlap(img, OF, in.hkls, valid.data(), (int)valid.size(), g, maxSpotsFit,
                              nullptr, nullptr, nullptr, 0, 0);
    lm_quat_from_matrix(OMbest, q1);
    lm_quat_from_matrix(OMfit, q2);
    int nrSps = 0, nSim = 0;
    std::string sl;
    overlap(img, OF, in.hkls, nullptr, in.nhkls, g, maxSpotsFit, &nrSps, &nSim, &sl, in.image_nr,
            s + 1);
    spotLines[s] = sl;
    if (nrSps >= p.min_nr_spots) {
      std::string row;
      char b[200];
      if (in.image_nr > 0) {
        snprintf(b, sizeof b, "%d\t", in.image_nr);
        row += b;
      }
      snprintf(b, sizeof b, "%d\t%d\t", s + 1, csize[s]);
      row += b;
      const double I = (mv / nrSps) * (mv / nrSps);
      snprintf(b, sizeof b, "%-13.4lf\t%-13.4lf\t%-13.4lf\t%d\t%d\t", I, nrSps * I, mv, nrSps,
               nSim);
      row += b;
      for (int k = 0; k < 9; ++k) {
        snprintf(b, sizeof b, "%-13.7lf\t\t", OF[k]);
        row += b;
      }
      for (int k = 0; k < 6; ++k) {
        snprintf(b, sizeof b, "%-13.7lf\t\t", p.lattice[k]);
        row += b;
      }
      for (int k = 0; k < 9; ++k) {
        snprintf(b, sizeof b, "%-13.7lf\t\t", OMfit[k]);
        row += b;
      }
      snprintf(b, sizeof b, "%-13.4lf\t%-13.7lf\t%llu\n", bestScore[s],
               lm_misorientation_deg(q1, q2, p.space_group), (unsigned long long)best[s]);
      row += b;
      solLines[s] = row;
    }
  }
  for (int s = 0; s < nsol; ++s) {
    spotsOut += spotLines[s];
    solOut += solLines[s];
  }
  return nsol;
}

}  // namespace

void write_solution_headers(FILE *sol, FILE *spots, bool image_column) {
  const char *pre = image_column ? "%ImageNr\t" : "%";
  fprintf(spots, "%sGrainNr\tSpotNr\th\tk\tl\tX\tY\tQhat[0]\tQhat[1]\tQhat[2]\tIntensity\n",
          pre);
  fprintf(sol,
          "%sGrainNr\tNumberOfSolutions\tIntensity\tNMatches*Intensity\tNMatches*"
          "sqrt(Intensity)\tNMatches\tNSpotsCalc\t"
          "Recip1\tRecip2\tRecip3\tRecip4\tRecip5\tRecip6\tRecip7\tRecip8\tRecip9\t"
          "LatticeParameterFit[a]\tLatticeParameterFit[b]\tLatticeParameterFit[c]\t"
          "LatticeParameterFit[alpha]\tLatticeParameterFit[beta]\tLatticeParameterFit[gamma]\t"
          "OrientMatrix0\tOrientMatrix1\tOrientMatrix2\tOrientMatrix3\tOrientMatrix4\t"
          "OrientMatrix5\tOrientMatrix6\tOrientMatrix7\tOrientMatrix8\t"
          "CoarseNMatches*sqrt(Intensity)\tmisOrientationPostRefinement[degrees]\t"
          "orientationRowNr\n",
          pre);
}

int write_solutions(const PostInput &in, const char *sol_path, const char *spots_path) {
  FILE *sol = fopen(sol_path, "w");
  if (!sol) {
    printf("Could not open file for writing solutions. Exiting.\n");
    return -1;
  }
  FILE *spots = fopen(spots_path, "w");
  if (!spots) {
    printf("Could not open file for writing spots. Exiting.\n");
    fclose(sol);
    return -1;
  }
  write_solution_headers(sol, spots, false);
  std::string solRows, spotRows;
  const int n = emit(in, solRows, spotRows);
  fputs(spotRows.c_str(), spots);
  fputs(solRows.c_str(), sol);
  fclose(spots);
  fclose(sol);
  return n;
}

int format_solutions(const PostInput &in, std::string *sol_rows, std::string *spot_rows) {
  return emit(in, *sol_rows, *spot_rows);
}
