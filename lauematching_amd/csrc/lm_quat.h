// lm_quat.h - quaternion steps of the duplicate merge, shared by the host
// merge (lm_merge.cpp) and the device pair kernel (lm_merge_dev.hip) so both
// run the same float64 operations in the same order (both are built without
// FP contraction; sqrt and division round correctly on both sides).
//
// Reference: QuatMultiply / BringDownToFundamentalRegionSym /
// GetMisOrientation (src/LaueMatchingHeaders.h:224-260).
#pragma once

#include <cmath>

#if defined(__HIPCC__) || defined(__HIP__)
#define LM_HD __host__ __device__
#else
#define LM_HD
#endif

namespace lm {

// Point-group rotations (w, x, y, z) of one crystal system.
struct SymSet {
  int n = 1;
  double q[24][4];
};

LM_HD inline void quat_normalize(double q[4]) {
  const double nrm = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
  for (int i = 0; i < 4; ++i) q[i] /= nrm;
}

// Q = q (x) r, scalar part made non-negative, normalised.
LM_HD inline void quat_mul(const double q[4], const double r[4], double Q[4]) {
  Q[0] = r[0] * q[0] - r[1] * q[1] - r[2] * q[2] - r[3] * q[3];
  Q[1] = r[1] * q[0] + r[0] * q[1] + r[3] * q[2] - r[2] * q[3];
  Q[2] = r[2] * q[0] + r[0] * q[2] + r[1] * q[3] - r[3] * q[1];
  Q[3] = r[3] * q[0] + r[0] * q[3] + r[2] * q[1] - r[1] * q[2];
  if (Q[0] < 0)
    for (int i = 0; i < 4; ++i) Q[i] = -Q[i];
  quat_normalize(Q);
}

// Symmetry-equivalent with the largest scalar part (first one on ties).
LM_HD inline void quat_fundamental(const SymSet &s, const double in[4], double out[4]) {
  double best[4] = {1, 0, 0, 0}, qt[4];
  double mc = -10000;
  for (int i = 0; i < s.n; ++i) {
    quat_mul(in, s.q[i], qt);
    if (mc < qt[0]) {
      mc = qt[0];
      for (int j = 0; j < 4; ++j) best[j] = qt[j];
    }
  }
  for (int j = 0; j < 4; ++j) out[j] = best[j];
  quat_normalize(out);
}

// Scalar part (clamped to 1) of the fundamental-region reduction of
// fa^-1 (x) fb, fa and fb the reductions of the two orientations; the
// misorientation angle is 2 acos of it.
LM_HD inline double quat_miso_cos(const SymSet &s, const double fa[4], const double fb[4]) {
  const double inv[4] = {-fa[0], fa[1], fa[2], fa[3]};
  double p[4], m[4];
  quat_mul(inv, fb, p);
  quat_fundamental(s, p, m);
  return m[0] > 1 ? 1.0 : m[0];
}

}  // namespace lm
