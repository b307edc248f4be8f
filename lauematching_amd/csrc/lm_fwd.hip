// lm_fwd.hip - forward simulation of the orientation DB into the resident
// spot table, and ForwardFile import / export.
//
// Reference behaviour (src/LaueMatchingCPU.c:591-667): for each orientation,
// thisOrient = orientation x recip; each hkl in file order is projected onto
// the detector; accepted reflections whose pixel is not yet taken by this
// orientation are appended until MaxNrLaueSpots spots are listed.
//
// Device design: one lane per orientation, the hkl loop is wave-uniform so the
// hkl triplets are scalar loads; each lane keeps its accepted pixels in an
// LDS column (slot-major, bank-conflict free) for the duplicate test.  The
// table is built in two passes over the DB: a count pass (spot counts only)
// then, once the host has turned the per-tile totals into offsets, a write
// pass that places every spot at its final position.  Both passes evaluate
// the same projection, so the counts of the two passes agree by construction.
//
// The float64 path is compiled with FP contraction off and keeps the
// reference's association order; HIP's double division and sqrt are the
// IEEE correctly rounded operations, as are the x86 SSE2 ones of the
// reference build.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "lm_gather.h"
#include "lm_internal.h"

#pragma clang fp contract(off)

namespace lm {
namespace {

constexpr int kFwdBlock = 128;
constexpr int kLdsSlots = 64;  // per-lane duplicate list kept in LDS when M <= 64
// With the kept lists in device memory (count pass, M <= 64), the first 32
// slots of each list stay in LDS and the rest go straight to the kept list:
// half the LDS per wave, twice the waves per CU.
constexpr int kLdsSlotsKeep = 32;
constexpr double kHc = 1.2398419739;  // src/LaueMatchingHeaders.h:41
constexpr double kFourPi = 4 * 3.14159265358979323846;

struct DevGeom {
  double rotT[9];
  double P[3];
  double pxX, pxY;
  double halfX, halfY;  // 0.5 * (N - 1)
  double W, H;
  double Elo, Ehi;
  int nrPxX;
  // division-free pre-test (plan_bit): q_z >= 0 rejects when Elo > 0;
  // energy bounds times 4 pi; detector-coordinate intervals (metres, before
  // the division by z) widened by a margin, used when pre_px is set
  int pre_sign, pre_px;
  double eloC, ehiC;
  double xlo, xhi, ylo, yhi;
  // cone pre-test (cone_bit), used when cone != 0: the lab direction c of
  // the detector's cone, the metric G0 = recip^T recip every orientation's
  // M^T M must match within tolG (then |q|^2 = h^T G0 h up to the per-hkl
  // margins of the plan)
  int cone;
  double cLab[3];
  double G0[6];  // (0,0) (1,1) (2,2) (0,1) (0,2) (1,2)
  double tolG;
};

// src/LaueMatchingCPU.c:604-636 for one (orientation, hkl).  Every value is
// computed as the reference computes it (same operations, same order); the
// reference's early exits become one conjunction at the end, so a wave runs
// the sequence without divergent branches (a rejected pair's later values
// may be inf or NaN and are then discarded).  The tests: qlen != 0, z > 0,
// -1 < t < N for both pixel coordinates (the (int) truncation kept within
// [0, N-1]), and not (E < Elo or E > Ehi).
__device__ __forceinline__ bool project64(const double *M, double h, double k, double l,
                                          const DevGeom &g, uint32_t *pix) {
  double q0 = 0, q1 = 0, q2 = 0;
  q0 += M[0] * h; q0 += M[1] * k; q0 += M[2] * l;
  q1 += M[3] * h; q1 += M[4] * k; q1 += M[5] * l;
  q2 += M[6] * h; q2 += M[7] * k; q2 += M[8] * l;
  const double qlen = __builtin_sqrt(q0 * q0 + q1 * q1 + q2 * q2);
  const double u0 = q0 / qlen, u1 = q1 / qlen, u2 = q2 / qlen;
  const double td = 2 * u2;
  const double k0 = 0 - td * u0, k1 = 0 - td * u1, k2 = 1.0 - td * u2;
  double x = 0, y = 0, z = 0;
  x += g.rotT[0] * k0; x += g.rotT[1] * k1; x += g.rotT[2] * k2;
  y += g.rotT[3] * k0; y += g.rotT[4] * k1; y += g.rotT[5] * k2;
  z += g.rotT[6] * k0; z += g.rotT[7] * k1; z += g.rotT[8] * k2;
  const double X = x * g.P[2] / z;
  const double Y = y * g.P[2] / z;
  const double tx = (X - g.P[0]) / g.pxX + g.halfX;
  const double ty = (Y - g.P[1]) / g.pxY + g.halfY;
  const double E = kHc * qlen / (kFourPi * (-u2));
  const bool ok = (qlen != 0) & (z > 0) & (tx > -1.0) & (tx < g.W) & (ty > -1.0) & (ty < g.H) &
                  !(E < g.Elo) & !(E > g.Ehi);
  // (int) of a value inside (-1, N) truncates toward zero as the reference's
  // cast does; rejected pairs convert 0
  const int ipx = (int)(ok ? tx : 0.0), ipy = (int)(ok ? ty : 0.0);
  *pix = (uint32_t)ipy * (uint32_t)g.nrPxX + (uint32_t)ipx;
  return ok;
}

// One pre-test evaluation of the plan (built on the host, fwd_plan): the hkl
// as doubles and the candidate bit (position in its 64-hkl chunk) for each
// sign of q_z.  A Friedel pair (h and -h in the same chunk) shares one
// evaluation: q(-h) is -q(h) exactly (IEEE products and sums round
// symmetrically), so |q|^2, |q_z| and K below are the same for both, and
// only the member with q_z < 0 can be accepted when Elo > 0.  bpos = 64
// (no bit) for an hkl without its partner in the chunk.
struct FwdEval {
  double h, k, l;
  uint64_t bits;  // bit when q_z < 0 (low word) / when q_z > 0 (high word); 64 = none
  // cone_bit bounds of this hkl: |q_z| in [ahi, alo], q_z (q.c) <= dcone
  double alo, ahi, dcone;
  double pad;     // 64 bytes: the int4 chunk table after the entries stays aligned
};

// The cone pre-test: a cheaper necessary condition than plan_bit, taken for
// the waves whose 64 orientations all have M^T M = G0 within tolG (any
// rotation of the lattice does; other matrices take plan_bit).  Then
// |q|^2 = h^T G0 h is a per-hkl constant up to a margin, and, with
// q = M (h, k, l):
//  * energy: the reference's E = hc|q|^2 / (4 pi |q_z|) to 1e-15 relative,
//    so E in [Elo, Ehi] needs |q_z| in [ahi, alo], bounds computed on the
//    host from |q|^2 with margins of 1e-9 relative and 1e-9 |q| absolute
//    (q_z is within a few 2^-53 |q| of the reference's);
//  * detector: every direction that lands inside the detector rectangle
//    (pixel coordinates -1 < t < N, the reference's test) lies within the
//    cone of half-angle g around the lab direction c, g the largest angle
//    between c and the rectangle's corners (the angle to c is quasiconvex
//    over the detector plane while g < 90 deg, so its maximum over the
//    rectangle is at a corner); cos g carries a margin of 1e-6.  With
//    k_f = e_z - 2 q_z q / |q|^2 (unit), k_f . c >= cos g reads
//    q_z (q . c) <= (c_z - cos g) |q|^2 / 2 = dcone (widened by the |q|^2
//    margin and 1e-9 |q|^2);
//  * sign: as in plan_bit, only the Friedel member with q_z < 0 can be
//    accepted (Elo > 0), and q_z = 0 is rejected by both.
// qc = q . c comes from v = M^T c, computed once per orientation.
// Returns the candidate bit, or 64 for none.  Branch-free.
__device__ __forceinline__ int cone_bit(const double *M, const double (&v)[3],
                                        const FwdEval &ev) {
  const double qz = __builtin_fma(M[8], ev.l, __builtin_fma(M[7], ev.k, M[6] * ev.h));
  const double qc = __builtin_fma(v[2], ev.l, __builtin_fma(v[1], ev.k, v[0] * ev.h));
  const double a = __builtin_fabs(qz);
  const bool rej = (a > ev.alo) | (a < ev.ahi) | (qz * qc > ev.dcone);
  const int bneg = (int)(uint32_t)ev.bits, bpos = (int)(uint32_t)(ev.bits >> 32);
  int bit = qz > 0 ? bpos : 64;
  bit = qz < 0 ? bneg : bit;
  return rej ? 64 : bit;
}

// A necessary condition for project64 to accept the member of the pair that
// the sign of q_z selects, with no division or square root: false only when
// the reference rejects it.  q and |q|^2 are computed exactly as project64
// computes them (same operations, same order), and each test carries a
// margin far above the rounding error of either side:
//  * sign: q-hat_z has the sign of q_z, so E = hc|q| / (4 pi (-q-hat_z)) is
//    negative or infinite for q_z >= 0 and E < Elo rejects (when Elo > 0);
//  * energy: E = hc |q|^2 / (4 pi |q_z|) mathematically; the reference's E is
//    within 7e-16 relative of it (six roundings), this side's products within
//    7e-16, so hc|q|^2 < Elo 4 pi |q_z| (1 - 1e-12) means E < Elo, and
//    likewise above Ehi;
//  * detector side: K = |q|^2 k_f = (-2 q_z q_x, -2 q_z q_y, |q|^2 - 2 q_z^2)
//    (|K| = |q|^2), lab z' = (rotT K)_z = |q|^2 z with an error below 20u
//    |q|^2 on both sides (u = 2^-53); z' < -1e-12 |q|^2 means z < 0;
//  * pixel bounds, when z' > 1e-3 |q|^2: the reference keeps -1 < t < N for
//    t = (x P2 / z - P0) / px + (N-1)/2, i.e. x P2 / z inside an interval of
//    detector coordinates; its X = x P2 / z is within |P2| 4.4e-9 of the exact
//    one for z >= 1e-3, and the interval is widened by |P2| 1e-7 + 1e-6 |px|,
//    so x' P2 outside z' [xlo, xhi] means t is outside (-1, N).
// Non-finite values make every rejecting comparison false: the pair goes to
// project64.  Returns the candidate bit, or 64 for none.  Branch-free.
// (project64 starts each component from 0: 0 + x differs from x only in the
// sign of a zero, which changes no comparison below, so the sums start from
// the first product here.)
__device__ __forceinline__ int plan_bit(const double *M, const FwdEval &ev, const DevGeom &g) {
  const double q0 = (M[0] * ev.h + M[1] * ev.k) + M[2] * ev.l;
  const double q1 = (M[3] * ev.h + M[4] * ev.k) + M[5] * ev.l;
  const double q2 = (M[6] * ev.h + M[7] * ev.k) + M[8] * ev.l;
  const double ql2 = q0 * q0 + q1 * q1 + q2 * q2;
  const double nq = __builtin_fabs(q2);
  const double A = kHc * ql2;
  bool rej = ql2 == 0;
  // eloC = Elo 4 pi (1 - 1e-12), ehiC = Ehi 4 pi (1 + 1e-12), rounded once more
  rej |= (nq > 1e-150) & ((A < g.eloC * nq) | (A > g.ehiC * nq));
  const double t = -2.0 * q2;
  const double K0 = t * q0, K1 = t * q1, K2 = __builtin_fma(t, q2, ql2);
  const double zz = __builtin_fma(g.rotT[6], K0, __builtin_fma(g.rotT[7], K1, g.rotT[8] * K2));
  const double xx = __builtin_fma(g.rotT[0], K0, __builtin_fma(g.rotT[1], K1, g.rotT[2] * K2));
  const double yy = __builtin_fma(g.rotT[3], K0, __builtin_fma(g.rotT[4], K1, g.rotT[5] * K2));
  rej |= zz < -1e-12 * ql2;
  const double xp = xx * g.P[2], yp = yy * g.P[2];
  const bool px = (g.pre_px != 0) & (zz > 1e-3 * ql2);
  rej |= px & ((xp < zz * g.xlo) | (xp > zz * g.xhi) | (yp < zz * g.ylo) | (yp > zz * g.yhi));
  // q_z == 0 (or NaN): the sign test rejects both members when Elo > 0
  const int bneg = (int)(uint32_t)ev.bits, bpos = (int)(uint32_t)(ev.bits >> 32);
  const int bzero = g.pre_sign ? 64 : bneg;
  int bit = q2 > 0 ? bpos : bzero;
  bit = q2 < 0 ? bneg : bit;
  return rej ? 64 : bit;
}

// Float32 geometry for the guarded path.
struct DevGeom32 {
  float rotT[9];
  float P[3];
  float inv_pxX, inv_pxY;  // 1/pitch (the bound covers the reciprocal's rounding)
  float halfX, halfY, W, H;
  float Elo, Ehi;
  float dist_px;           // max(P[2]/pxX, P[2]/pxY): pixels per unit of x/z
  float P0px, P1px;        // |P[0]|/pxX, |P[1]|/pxY
};

// Classification of one detector coordinate t against the reference's
// decisions: trunc(t) for -1 < t < N.  The decision points are -1, the
// integers k >= 1 (where trunc changes; trunc is 0 on (-1, 1)) and N.
// Returns 0 = clearly rejected, 1 = clearly accepted (*ip set), 2 = within e
// of a decision point (or not finite).  Every test compares one difference
// with e; the differences t - N, -1 - t, t - k, k + 1 - t are exact in float
// wherever they are small (Sterbenz), so no value of t falls between the
// reject, undecided and accept tests.
__device__ __forceinline__ int classify_px(float t, float n, float e, int *ip) {
  if (!(fabsf(t) < 1e30f)) return 2;
  if (t >= n) return (t - n <= e) ? 2 : 0;
  if (t <= -1.0f) return (-1.0f - t <= e) ? 2 : 0;
  float d;
  if (t >= 1.0f) {
    const float k = floorf(t);
    d = fminf(t - k, (k + 1.0f) - t);
  } else {
    d = fminf(t + 1.0f, 1.0f - t);
  }
  if (d <= e) return 2;
  *ip = (int)t;
  return 1;
}

// Float32 evaluation of one (orientation, hkl) with a forward-error bound
// (Higham-style, constants carrying a 2x slack over the derivation below).
// u = 2^-24.  With S = max row of sum |M_ij||h_j| and r = S/|q|:
//   q-hat components: <= 8u(1 + r) absolute      (q: gamma_4 S; /|q|)
//   kf = e_z - 2 qz q-hat: <= 32u(1 + r)
//   lab xyz = rotT kf: e_xyz <= 2 e_kf + 8u       (row sums of |rotT| <= sqrt 3)
//   t = (x P2/z - P0)/px + (N-1)/2:
//     e_t <= 2 [ (P2/px) e_xyz (1/|z| + |x|/z^2) + 8u (|t| + (N-1)/2 + |x P2/z|/px + |P0|/px) ]
//   E = hc |q| / (4 pi (-qz_hat)): relative e_E <= 2 [ 8u(1 + r) + e_kf/|qz_hat| + 8u ]
// Returns 0 = reject, 1 = accept (*pix set), 2 = undecided (recompute in FP64).
__device__ __forceinline__ int project32(const float *M, int hi, int ki, int li,
                                         const DevGeom32 &g, uint32_t nrPxX, uint32_t *pix) {
  constexpr float U = 5.9604645e-8f;
  const float h = (float)hi, k = (float)ki, l = (float)li;
  const float q0 = M[0] * h + M[1] * k + M[2] * l;
  const float q1 = M[3] * h + M[4] * k + M[5] * l;
  const float q2 = M[6] * h + M[7] * k + M[8] * l;
  const float s0 = fabsf(M[0] * h) + fabsf(M[1] * k) + fabsf(M[2] * l);
  const float s1 = fabsf(M[3] * h) + fabsf(M[4] * k) + fabsf(M[5] * l);
  const float s2 = fabsf(M[6] * h) + fabsf(M[7] * k) + fabsf(M[8] * l);
  const float ql2 = q0 * q0 + q1 * q1 + q2 * q2;
  if (hi == 0 && ki == 0 && li == 0) return 0;  // qlen == 0 exactly
  if (!(ql2 > 0.0f)) return 2;
  const float ql = sqrtf(ql2);
  const float r = fmaxf(s0, fmaxf(s1, s2)) / ql;
  const float u0 = q0 / ql, u1 = q1 / ql, u2 = q2 / ql;
  const float td = 2.0f * u2;
  const float k0 = -td * u0, k1 = -td * u1, k2 = 1.0f - td * u2;
  const float x = g.rotT[0] * k0 + g.rotT[1] * k1 + g.rotT[2] * k2;
  const float y = g.rotT[3] * k0 + g.rotT[4] * k1 + g.rotT[5] * k2;
  const float z = g.rotT[6] * k0 + g.rotT[7] * k1 + g.rotT[8] * k2;
  const float ekf = 32.0f * U * (1.0f + r);
  const float exyz = 2.0f * ekf + 8.0f * U;
  if (z < -exyz) return 0;
  if (z <= exyz) return 2;
  int verdict = 1;
  // energy window (inclusive): as for the pixel coordinates, each test
  // compares one difference (exact near the window edges) with eE
  const float su = -u2;
  if (fabsf(su) <= ekf) {
    verdict = 2;
  } else {
    constexpr float kHc4Pi = (float)(kHc / kFourPi);  // hc / (4 pi), keV nm
    const float E = kHc4Pi * ql / su;
    const float eE = 2.0f * (8.0f * U * (1.0f + r) + ekf / fabsf(su) + 8.0f * U) * fabsf(E) +
                     1e-6f * fabsf(E);
    const float dlo = E - g.Elo, dhi = g.Ehi - E;
    if (!(fabsf(E) < 1e30f)) {
      verdict = 2;
    } else {
      if (dlo < -eE || dhi < -eE) return 0;
      if (dlo <= eE || dhi <= eE) verdict = 2;
    }
  }
  const float iz = 1.0f / z;
  const float X = x * g.P[2] * iz, Y = y * g.P[2] * iz;
  const float tx = (X - g.P[0]) * g.inv_pxX + g.halfX;
  const float ty = (Y - g.P[1]) * g.inv_pxY + g.halfY;
  const float amp = g.dist_px * exyz * (iz + fmaxf(fabsf(x), fabsf(y)) * iz * iz);
  const float etx = 2.0f * (amp + 8.0f * U * (fabsf(tx) + g.halfX + fabsf(X) * g.inv_pxX + g.P0px));
  const float ety = 2.0f * (amp + 8.0f * U * (fabsf(ty) + g.halfY + fabsf(Y) * g.inv_pxY + g.P1px));
  int ipx = 0, ipy = 0;
  const int cx = classify_px(tx, g.W, etx, &ipx);
  if (cx == 0) return 0;
  const int cy = classify_px(ty, g.H, ety, &ipy);
  if (cy == 0) return 0;
  if (cx == 2 || cy == 2 || verdict == 2) return 2;
  *pix = (uint32_t)ipy * nrPxX + (uint32_t)ipx;
  return 1;
}

// thisOrient = orientation x recip (MatrixMultF33, src/LaueMatchingHeaders.h:
// 349-357: (a*b + c*d) + e*f), its float copy for the guarded path, and the
// cone pre-test's v = M^T c.  Returns true when M^T M = G0 within tolG (a
// rotation of the lattice: the cone pre-test applies; NaN fails).
template <bool GUARD>
__device__ __forceinline__ bool orient_load(const double *__restrict__ orients, uint64_t o,
                                            const double (&R)[9], const DevGeom &g,
                                            double (&M)[9], float (&M32)[9], double (&cv)[3]) {
  double t[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) t[i] = orients[o * 9 + i];
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    M[r * 3 + 0] = t[r * 3 + 0] * R[0] + t[r * 3 + 1] * R[3] + t[r * 3 + 2] * R[6];
    M[r * 3 + 1] = t[r * 3 + 0] * R[1] + t[r * 3 + 1] * R[4] + t[r * 3 + 2] * R[7];
    M[r * 3 + 2] = t[r * 3 + 0] * R[2] + t[r * 3 + 1] * R[5] + t[r * 3 + 2] * R[8];
  }
  if (GUARD) {
#pragma unroll
    for (int i = 0; i < 9; ++i) M32[i] = (float)M[i];
  }
  cv[0] = cv[1] = cv[2] = 0;
  if (!g.cone) return false;
  const double G[6] = {
      __builtin_fma(M[6], M[6], __builtin_fma(M[3], M[3], M[0] * M[0])),
      __builtin_fma(M[7], M[7], __builtin_fma(M[4], M[4], M[1] * M[1])),
      __builtin_fma(M[8], M[8], __builtin_fma(M[5], M[5], M[2] * M[2])),
      __builtin_fma(M[6], M[7], __builtin_fma(M[3], M[4], M[0] * M[1])),
      __builtin_fma(M[6], M[8], __builtin_fma(M[3], M[5], M[0] * M[2])),
      __builtin_fma(M[7], M[8], __builtin_fma(M[4], M[5], M[1] * M[2]))};
  bool fast = true;
#pragma unroll
  for (int i = 0; i < 6; ++i) fast &= __builtin_fabs(G[i] - g.G0[i]) <= g.tolG;
#pragma unroll
  for (int j = 0; j < 3; ++j)
    cv[j] = __builtin_fma(M[6 + j], g.cLab[2], __builtin_fma(M[3 + j], g.cLab[1], M[j] * g.cLab[0]));
  return fast;
}

// The spot list of one orientation (src/LaueMatchingCPU.c:604-667): hkls in
// file order, each accepted pixel not yet listed is appended, then
// spot(pix) is called, until ns reaches maxSpots.  List slot s < SLOTS is
// colL[s * kFwdBlock] (LDS), slot s >= SLOTS is colG[(s - SLOTS) * gStride].
// A lane enters with ns = maxSpots to take no part.
//
// hkls in chunks of at most 64 (chunks[c] = first hkl, first and last+1
// plan entry): first the division-free pre-test of the chunk's plan entries
// (wave-uniform loop, scalar loads) marks the candidates in a 64-bit mask,
// then each lane projects only its own candidates, in hkl order (ctz), so
// the wave runs as many projections as its busiest lane has candidates
// instead of one per hkl; first unseen pixel still wins.  Chunks are taken
// four at a time (up to 256 hkls): the pre-test fills four masks, then each
// lane walks the union of its four masks in hkl order, so the wave runs as
// many projections as its busiest lane has candidates in those hkls, rather
// than the sum over chunks of each chunk's busiest.  The next candidate's
// hkl is loaded while the current one is projected.  useCone
// (wave-uniform) selects cone_bit over plan_bit.  GUARD: float32
// projection, float64 recompute of undecided pairs (counted in nrec); nproj
// counts the projected pairs.
template <bool GUARD, int SLOTS, class Spot>
__device__ __forceinline__ void walk_hkls(const double (&M)[9], const float (&M32)[9],
                                          const double (&cv)[3], bool useCone, const DevGeom &g,
                                          const DevGeom32 &g32, const int32_t *__restrict__ hkls,
                                          const FwdEval *__restrict__ plan,
                                          const int4 *__restrict__ chunks, int nchunks,
                                          int maxSpots, int &ns, uint32_t &nrec, uint32_t &nproj,
                                          uint32_t *colL, uint32_t *colG, uint64_t gStride,
                                          Spot spot) {
  constexpr int kSuper = 4;
  for (int c0 = 0; c0 < nchunks; c0 += kSuper) {
    if (__all(ns >= maxSpots)) break;
    uint64_t cand[kSuper];
    int base[kSuper];
#pragma unroll
    for (int k = 0; k < kSuper; ++k) {
      cand[k] = 0;
      base[k] = 0;
      if (c0 + k < nchunks) {
        const int4 ch = chunks[c0 + k];
        base[k] = ch.x;
        uint64_t m = 0;
        if (useCone) {
#pragma unroll 4
          for (int e = ch.y; e < ch.z; ++e) {
            const int bit = cone_bit(M, cv, plan[e]);
            m |= bit < 64 ? 1ull << bit : 0ull;
          }
        } else {
#pragma unroll 2
          for (int e = ch.y; e < ch.z; ++e) {
            const int bit = plan_bit(M, plan[e], g);
            m |= bit < 64 ? 1ull << bit : 0ull;
          }
        }
        cand[k] = ns < maxSpots ? m : 0ull;
      }
    }
    // this lane's next candidate (hkl index, or -1): the lowest bit of its
    // first nonempty mask, which is then cleared; selects, no branches
    auto pop = [&]() -> int {
      const bool n0 = cand[0] != 0, n1 = cand[1] != 0, n2 = cand[2] != 0;
      const uint64_t c = n0 ? cand[0] : n1 ? cand[1] : n2 ? cand[2] : cand[3];
      const int b = n0 ? base[0] : n1 ? base[1] : n2 ? base[2] : base[3];
      const uint64_t r = c & (c - 1);
      cand[0] = n0 ? r : cand[0];
      cand[1] = !n0 && n1 ? r : cand[1];
      cand[2] = !n0 && !n1 && n2 ? r : cand[2];
      cand[3] = !n0 && !n1 && !n2 ? r : cand[3];
      return c ? b + (int)__builtin_ctzll(c) : -1;
    };
    int hi = pop();
    int hI = 0, kI = 0, lI = 0;
    if (hi >= 0) {
      hI = hkls[3 * hi + 0];
      kI = hkls[3 * hi + 1];
      lI = hkls[3 * hi + 2];
    }
    while (hi >= 0 && ns < maxSpots) {
      const int hn = pop();
      const int h0 = hI, k0 = kI, l0 = lI;
      if (hn >= 0) {
        hI = hkls[3 * hn + 0];
        kI = hkls[3 * hn + 1];
        lI = hkls[3 * hn + 2];
      }
      ++nproj;
      uint32_t pix = 0;
      bool ok;
      if (GUARD) {
        const int v = project32(M32, h0, k0, l0, g32, (uint32_t)g.nrPxX, &pix);
        if (v == 2) {
          ok = project64(M, (double)h0, (double)k0, (double)l0, g, &pix);
          ++nrec;
        } else {
          ok = v == 1;
        }
      } else {
        ok = project64(M, (double)h0, (double)k0, (double)l0, g, &pix);
      }
      if (ok) {
        // duplicate test over the whole list, LDS slots four at a time
        bool dup = false;
        const int nl = SLOTS ? min(ns, SLOTS) : 0;
        int s = 0;
        for (; s + 4 <= nl; s += 4)
          dup |= (colL[s * kFwdBlock] == pix) | (colL[(s + 1) * kFwdBlock] == pix) |
                 (colL[(s + 2) * kFwdBlock] == pix) | (colL[(s + 3) * kFwdBlock] == pix);
        for (; s < nl; ++s) dup |= colL[s * kFwdBlock] == pix;
        for (s = SLOTS; s < ns; ++s) dup |= colG[(uint64_t)(s - SLOTS) * gStride] == pix;
        if (!dup) {
          if (ns < SLOTS)
            colL[ns * kFwdBlock] = pix;
          else
            colG[(uint64_t)(ns - SLOTS) * gStride] = pix;
          spot(pix);
          ++ns;
        }
      }
      hi = hn;
    }
  }
}

// Sum of v over the wave, added to *dst by lane 0.
__device__ __forceinline__ void wave_add(uint32_t v, unsigned long long *dst) {
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d);
  if (__lane_id() == 0 && v) atomicAdd(dst, (unsigned long long)v);
}

// One lane per orientation.  WRITE=false: counts[o] = number of spots.
// WRITE=true: spots of orientation o go to entries[tile_off[tile] + in-tile
// prefix of counts], in spot order.  GUARD: float32 projection, float64
// recompute of undecided pairs (counted in *recomputed by the count pass).
// KEEP (count pass): the spot list of orientation o is also stored in
// keep[s * n + o], so the write pass (FROMKEEP) copies it instead of
// projecting again.
// SLOTS: list slots per lane kept in LDS (0, 32 or 64); slots past them live
// in `keep` (SLOTS = 32, count pass) or in `spill` (SLOTS = 0, M > 64).
template <bool WRITE, int SLOTS, bool GUARD, bool FROMKEEP = false>
__global__ void __launch_bounds__(kFwdBlock)
fwd_kernel(const double *__restrict__ orients, uint64_t n, const int32_t *__restrict__ hkls,
           int nhkls, double r00, double r01, double r02, double r10, double r11, double r12,
           double r20, double r21, double r22, DevGeom g, DevGeom32 g32, int maxSpots,
           int tileBits, uint16_t *__restrict__ counts, const uint64_t *__restrict__ tileOff,
           uint32_t *__restrict__ entries, uint32_t *__restrict__ spill,
           unsigned long long *__restrict__ recomputed, uint32_t *__restrict__ keep,
           const FwdEval *__restrict__ plan, const int4 *__restrict__ chunks, int nchunks) {
  __shared__ uint32_t lst[SLOTS ? SLOTS * kFwdBlock : 1];
  __shared__ uint32_t scan[kFwdBlock];
  const uint64_t o = (uint64_t)blockIdx.x * kFwdBlock + threadIdx.x;
  const bool live = o < n;
  const double R[9] = {r00, r01, r02, r10, r11, r12, r20, r21, r22};
  double M[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, cv[3] = {0, 0, 0};
  float M32[9];
  // the cone pre-test when every orientation of the wave is a rotation of
  // the lattice, else plan_bit
  bool fast = !live;
  if (live && !FROMKEEP) fast = orient_load<GUARD>(orients, o, R, g, M, M32, cv);
  const bool useCone = g.cone && !FROMKEEP && __all(fast);
  uint32_t nrec = 0, nproj = 0;
  uint32_t *const colL = lst + threadIdx.x;  // slot s at colL[s * kFwdBlock], s < SLOTS
  uint32_t *const colG = SLOTS == 0 ? spill + o : keep + (uint64_t)SLOTS * n + o;
  const uint64_t gStride = SLOTS == 0 ? (uint64_t)gridDim.x * kFwdBlock : n;
  auto get = [&](int s) -> uint32_t {
    return s < SLOTS ? colL[s * kFwdBlock] : colG[(uint64_t)(s - SLOTS) * gStride];
  };
  int ns = live ? 0 : maxSpots;
  if (FROMKEEP)
    ns = live ? (int)counts[o] : 0;
  else
    walk_hkls<GUARD, SLOTS>(M, M32, cv, useCone, g, g32, hkls, plan, chunks, nchunks, maxSpots,
                            ns, nrec, nproj, colL, colG, gStride, [](uint32_t) {});
  if (!WRITE) {
    if (live) counts[o] = (uint16_t)ns;
    if (keep && live)
      for (int s = 0; s < (SLOTS == kLdsSlotsKeep ? min(ns, SLOTS) : ns); ++s)
        keep[(uint64_t)s * n + o] = get(s);
    // recomputed[0]: pairs recomputed in float64; recomputed[1]: pairs that
    // passed the pre-test and were projected
    wave_add(nrec, recomputed);
    wave_add(nproj, recomputed + 1);
    return;
  }
  // In-tile offset of this orientation: counts of the tile's rows before this
  // block, then an exclusive scan over the block.
  const int c = live ? ns : 0;
  const uint64_t tile = o >> tileBits;
  const uint64_t tileFirst = tile << tileBits;
  const uint64_t blockFirst = (uint64_t)blockIdx.x * kFwdBlock;
  uint32_t before = 0;
  for (uint64_t r = tileFirst + threadIdx.x; r < blockFirst; r += kFwdBlock) before += counts[r];
  scan[threadIdx.x] = before;
  __syncthreads();
  for (int s = kFwdBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) scan[threadIdx.x] += scan[threadIdx.x + s];
    __syncthreads();
  }
  const uint32_t blockBase = scan[0];
  __syncthreads();
  scan[threadIdx.x] = (uint32_t)c;
  __syncthreads();
  for (int s = 1; s < kFwdBlock; s <<= 1) {
    uint32_t v = ((int)threadIdx.x >= s) ? scan[threadIdx.x - s] : 0;
    __syncthreads();
    scan[threadIdx.x] += v;
    __syncthreads();
  }
  const uint32_t excl = scan[threadIdx.x] - (uint32_t)c;
  if (live) {
    uint32_t *dst = entries + tileOff[tile] + blockBase + excl;
    for (int s = 0; s < c; ++s)
      dst[s] = FROMKEEP ? keep[(uint64_t)s * n + o] : get(s);
  }
}

// On-the-fly scoring (lm_score_fwd_batch_dev): each orientation's spot list
// is projected as the count pass projects it and then scored against B
// interleaved images, so no table is read or written.  The accumulators and
// the hit test are those of the table walk (score_kernel), so each image's
// hit records are the ones the table would give.  M <= 64: list slots 0..31
// in LDS, slots 32.. in spill[(s - 32) * (grid rows) + row in the launch].
// stats[0] pairs recomputed in float64, [1] pairs projected, [2] spots
// listed.
template <int B, bool GUARD>
__global__ void __launch_bounds__(kFwdBlock)
fwd_score_kernel(const double *__restrict__ orients, uint64_t n, uint64_t row0,
                 const int32_t *__restrict__ hkls, double r00, double r01, double r02, double r10,
                 double r11, double r12, double r20, double r21, double r22, DevGeom g,
                 DevGeom32 g32, int maxSpots, const FwdEval *__restrict__ plan,
                 const int4 *__restrict__ chunks, int nchunks, const uint8_t *__restrict__ img,
                 const Pred *__restrict__ pred, uint64_t *__restrict__ hits, uint64_t capPer,
                 uint32_t *__restrict__ counters, unsigned long long *__restrict__ stats,
                 uint32_t *__restrict__ spill) {
  __shared__ uint32_t lst[kLdsSlotsKeep * kFwdBlock];
  const uint64_t local = (uint64_t)blockIdx.x * kFwdBlock + threadIdx.x;
  const uint64_t o = row0 + local;
  const bool live = o < n;
  const int lane = __lane_id();
  const double R[9] = {r00, r01, r02, r10, r11, r12, r20, r21, r22};
  double M[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, cv[3] = {0, 0, 0};
  float M32[9];
  bool fast = !live;
  if (live) fast = orient_load<GUARD>(orients, o, R, g, M, M32, cv);
  const bool useCone = g.cone && __all(fast);
  uint32_t nrec = 0, nproj = 0;
  uint32_t *const colL = lst + threadIdx.x;
  uint32_t *const colG = spill + local;
  const uint64_t gStride = (uint64_t)gridDim.x * kFwdBlock;
  auto get = [&](int s) -> uint32_t {
    return s < kLdsSlotsKeep ? colL[s * kFwdBlock] : colG[(uint64_t)(s - kLdsSlotsKeep) * gStride];
  };
  int ns = live ? 0 : maxSpots;
  uint32_t acc1 = 0;
  Swar<B < 4 ? 4 : B> a;
#pragma unroll
  for (int j = 0; j < (B < 4 ? 1 : B / 4); ++j) a.s02[j] = a.s13[j] = a.cn[j] = 0;
  walk_hkls<GUARD, kLdsSlotsKeep>(M, M32, cv, useCone, g, g32, hkls, plan, chunks, nchunks,
                                  maxSpots, ns, nrec, nproj, colL, colG, gStride,
                                  [](uint32_t) {});
  // the listed pixels of the B images, read after the walk with four
  // gathers in flight (not one stalled gather per accepted spot)
  auto add = [&](const uint32_t (&w)[PxW<B>::W]) {
    if constexpr (B == 1)
      acc1 += w[0] ? (1u << kSumBits) + w[0] : 0u;
    else
      swar_px<B>(w, a);
  };
  const int nl = live ? ns : 0;
  int s = 0;
  for (; s + 4 <= nl; s += 4) {
    uint32_t w[4][PxW<B>::W];
#pragma unroll
    for (int q = 0; q < 4; ++q) load_px<B>(img, get(s + q), w[q]);
#pragma unroll
    for (int q = 0; q < 4; ++q) add(w[q]);
  }
  for (; s < nl; ++s) {
    uint32_t w[PxW<B>::W];
    load_px<B>(img, get(s), w);
    add(w);
  }
  wave_add(nrec, stats);
  wave_add(nproj, stats + 1);
  wave_add(live ? (uint32_t)ns : 0u, stats + 2);
  // src/LaueMatchingCPU.c:683-689 per image, as score_kernel decides it
#pragma unroll
  for (int b = 0; b < B; ++b) {
    uint32_t accb;
    if constexpr (B == 1)
      accb = acc1;
    else
      accb = swar_acc<B>(a, b);
    const uint32_t nsb = accb >> kSumBits;
    const uint32_t sum = accb & kSumMask;
    bool cand = false, zero_pass = false;
    if (live && b < pred->nimg) {
      if (nsb == 0)
        zero_pass = (0 >= pred->min_spots[b]) && (0.0 >= pred->min_int[b]);
      else
        cand = ((int)nsb >= pred->min_spots[b]) && ((double)sum >= pred->cand_thr[b]);
    }
    const uint64_t zm = __ballot(zero_pass);
    if (zm && lane == 0) atomicAdd(&counters[2 * b + 1], (uint32_t)__popcll(zm));
    append(cand, (o << 26) | (uint64_t)accb, hits + (uint64_t)b * capPer, &counters[2 * b],
           capPer);
  }
}

// Float64 in-order sums (src/LaueMatchingCPU.c:674-681) of the candidate
// rows of images with imScale != 1 on the on-the-fly path: each lane lists
// its row's spots again (float64: the guarded projection lists the same
// spots) and sums image b's pixels in list order.  xids[i] = b << 32 | row;
// the images lie back to back, npx pixels each.
__global__ void __launch_bounds__(kFwdBlock)
fwd_exact_kernel(const double *__restrict__ orients, const int32_t *__restrict__ hkls,
                 double r00, double r01, double r02, double r10, double r11, double r12,
                 double r20, double r21, double r22, DevGeom g, DevGeom32 g32, int maxSpots,
                 const FwdEval *__restrict__ plan, const int4 *__restrict__ chunks, int nchunks,
                 const uint64_t *__restrict__ xids, uint64_t ncand,
                 const uint8_t *__restrict__ imgs, uint64_t npx, const Pred *__restrict__ pred,
                 double *__restrict__ out) {
  __shared__ uint32_t lst[kLdsSlots * kFwdBlock];
  const uint64_t i = (uint64_t)blockIdx.x * kFwdBlock + threadIdx.x;
  const bool live = i < ncand;
  const uint64_t x = live ? xids[i] : 0;
  const uint32_t b = (uint32_t)(x >> 32);
  const uint64_t o = x & 0xffffffffull;
  const double R[9] = {r00, r01, r02, r10, r11, r12, r20, r21, r22};
  double M[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, cv[3] = {0, 0, 0};
  float M32[9];
  bool fast = !live;
  if (live) fast = orient_load<false>(orients, o, R, g, M, M32, cv);
  const bool useCone = g.cone && __all(fast);
  uint32_t nrec = 0, nproj = 0;
  int ns = live ? 0 : maxSpots;
  const double scale = live ? pred->scale[b] : 0.0;
  const uint8_t *const im = imgs + (uint64_t)b * npx;
  double tot = 0;
  walk_hkls<false, kLdsSlots>(M, M32, cv, useCone, g, g32, hkls, plan, chunks, nchunks, maxSpots,
                              ns, nrec, nproj, lst + threadIdx.x, nullptr, 0,
                              [&](uint32_t pix) {
                                const uint8_t raw = im[pix];
                                if (raw > 0) tot += (double)raw * scale;
                              });
  if (live) out[i] = tot;
}

// Per-tile sums of the counts.
__global__ void tile_sum_kernel(const uint16_t *__restrict__ counts, uint64_t n, int tileBits,
                                uint64_t *__restrict__ tileSum) {
  __shared__ uint64_t red[256];
  const uint64_t t = blockIdx.x;
  const uint64_t first = t << tileBits;
  uint64_t last = first + (1ull << tileBits);
  if (last > n) last = n;
  uint64_t s = 0;
  for (uint64_t r = first + threadIdx.x; r < last; r += 256) s += counts[r];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) tileSum[t] = red[0];
}

// ForwardFile records (chunk of rows [row0, row0+rows)) -> counts, with
// validation (count <= M, pixel inside the detector).
__global__ void import_count_kernel(const uint16_t *__restrict__ rec, uint64_t rows,
                                    uint64_t row0, int M, int W, int H,
                                    uint16_t *__restrict__ counts, uint32_t *__restrict__ bad) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows) return;
  const uint16_t *r = rec + i * (uint64_t)(1 + 2 * M);
  int c = r[0];
  bool ok = c <= M;
  if (ok)
    for (int s = 0; s < c; ++s)
      if (r[1 + 2 * s] >= W || r[2 + 2 * s] >= H) { ok = false; break; }
  if (!ok) atomicOr(bad, 1u);
  counts[row0 + i] = (uint16_t)(ok ? c : 0);
}

__global__ void import_write_kernel(const uint16_t *__restrict__ rec, uint64_t rows,
                                    uint64_t row0, int M, int W, int tileBits,
                                    const uint16_t *__restrict__ counts,
                                    const uint64_t *__restrict__ tileOff,
                                    uint32_t *__restrict__ entries) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows) return;
  const uint64_t o = row0 + i;
  const uint64_t tile = o >> tileBits;
  uint64_t off = tileOff[tile];
  for (uint64_t r = tile << tileBits; r < o; ++r) off += counts[r];
  const uint16_t *src = rec + i * (uint64_t)(1 + 2 * M);
  const int c = counts[o];
  for (int s = 0; s < c; ++s) entries[off + s] = (uint32_t)src[2 + 2 * s] * W + src[1 + 2 * s];
}

__global__ void export_kernel(uint16_t *__restrict__ rec, uint64_t rows, uint64_t row0, int M,
                              int W, int tileBits, const uint16_t *__restrict__ counts,
                              const uint64_t *__restrict__ tileOff,
                              const uint32_t *__restrict__ entries) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows) return;
  const uint64_t o = row0 + i;
  const uint64_t tile = o >> tileBits;
  uint64_t off = tileOff[tile];
  for (uint64_t r = tile << tileBits; r < o; ++r) off += counts[r];
  uint16_t *dst = rec + i * (uint64_t)(1 + 2 * M);
  const int c = counts[o];
  dst[0] = (uint16_t)c;
  for (int s = 0; s < M; ++s) {
    uint16_t x = 0, y = 0;
    if (s < c) {
      const uint32_t p = entries[off + s];
      x = (uint16_t)(p % (uint32_t)W);
      y = (uint16_t)(p / (uint32_t)W);
    }
    dst[1 + 2 * s] = x;
    dst[2 + 2 * s] = y;
  }
}

__device__ __forceinline__ uint64_t mix64(uint64_t x) {
  x += 0x9e3779b97f4a7c15ull;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}

// Position-tagged sum (mod 2^64) over the counts and the orientation-major
// entries: equal tables give equal sums, and any changed, moved or missing
// value changes the sum except with probability ~2^-64.
__global__ void checksum_kernel(const uint16_t *__restrict__ counts, uint64_t n,
                                const uint32_t *__restrict__ entries, uint64_t total,
                                unsigned long long *__restrict__ out) {
  uint64_t acc = 0;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    acc += mix64((i << 16) ^ counts[i]);
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride)
    acc += mix64(~(i << 32) ^ entries[i]);
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) acc += __shfl_xor(acc, d);
  if (__lane_id() == 0) atomicAdd(out, (unsigned long long)acc);
}

__global__ void selftest_kernel(const double *__restrict__ a, const double *__restrict__ b,
                                uint64_t n, double *__restrict__ q, double *__restrict__ s) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  q[i] = a[i] / b[i];
  s[i] = __builtin_sqrt(a[i]);
}

// Host constants of the cone pre-test (cone_bit); ok = false leaves every
// wave on plan_bit.
struct ConeSetup {
  bool ok = false;
  double cLab[3] = {0, 0, 0};
  double cosg = 0;   // cosine of the cone's half-angle, margin included
  double G0[6] = {0, 0, 0, 0, 0, 0};
  double G0max = 0, tolG = 0;
};

ConeSetup cone_setup(const Geom &g) {
  ConeSetup cs;
  const double P2 = g.P[2];
  if (!(g.Elo > 0) || !std::isfinite(g.Ehi) || !(g.Ehi >= g.Elo) || !std::isfinite(P2) ||
      P2 == 0 || !std::isfinite(g.P[0]) || !std::isfinite(g.P[1]) || !(g.pxX != 0) ||
      !(g.pxY != 0) || !std::isfinite(g.pxX) || !std::isfinite(g.pxY))
    return cs;
  // the detector rotation must be orthogonal: k_f . c = (rotT k_f) . c_det
  // is then the cosine between the lab direction and c
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0;
      for (int r = 0; r < 3; ++r) s += g.rotT[r * 3 + i] * g.rotT[r * 3 + j];
      if (!(std::fabs(s - (i == j ? 1.0 : 0.0)) <= 1e-12)) return cs;
    }
  // detector-frame directions (x/z, y/z, 1) of the rectangle's corners at
  // pixel coordinates -1 and N (x P2 / z = P0 + (t - (N-1)/2) px)
  double cd[4][3], c[3] = {0, 0, 0};
  int k = 0;
  for (int ix = 0; ix < 2; ++ix)
    for (int iy = 0; iy < 2; ++iy, ++k) {
      const double tx = ix ? (double)g.nrPxX : -1.0, ty = iy ? (double)g.nrPxY : -1.0;
      const double X = g.P[0] + (tx - 0.5 * (g.nrPxX - 1)) * g.pxX;
      const double Y = g.P[1] + (ty - 0.5 * (g.nrPxY - 1)) * g.pxY;
      const double v[3] = {X / P2, Y / P2, 1.0};
      const double nv = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
      for (int j = 0; j < 3; ++j) {
        cd[k][j] = v[j] / nv;
        c[j] += cd[k][j];
      }
    }
  const double nc = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
  if (!(nc > 0)) return cs;
  double cosg = 1.0;
  for (int j = 0; j < 3; ++j) c[j] /= nc;
  for (k = 0; k < 4; ++k) cosg = std::min(cosg, cd[k][0] * c[0] + cd[k][1] * c[1] + cd[k][2] * c[2]);
  // below 84 deg half-angle (the corner argument needs < 90 deg)
  if (!(cosg > 0.1)) return cs;
  cs.cosg = cosg - 1e-6;
  for (int i = 0; i < 3; ++i)
    cs.cLab[i] = g.rotT[0 * 3 + i] * c[0] + g.rotT[1 * 3 + i] * c[1] + g.rotT[2 * 3 + i] * c[2];
  const int ij[6][2] = {{0, 0}, {1, 1}, {2, 2}, {0, 1}, {0, 2}, {1, 2}};
  for (int e = 0; e < 6; ++e) {
    long double s = 0;
    for (int r = 0; r < 3; ++r)
      s += (long double)g.recip[r * 3 + ij[e][0]] * (long double)g.recip[r * 3 + ij[e][1]];
    cs.G0[e] = (double)s;
    cs.G0max = std::max(cs.G0max, std::fabs(cs.G0[e]));
  }
  if (!(cs.G0max > 0) || !std::isfinite(cs.G0max)) return cs;
  cs.tolG = 1e-9 * cs.G0max;
  cs.ok = true;
  return cs;
}

DevGeom dev_geom(const Geom &g) {
  DevGeom d{};
  for (int i = 0; i < 9; ++i) d.rotT[i] = g.rotT[i];
  for (int i = 0; i < 3; ++i) d.P[i] = g.P[i];
  d.pxX = g.pxX;
  d.pxY = g.pxY;
  d.halfX = 0.5 * (g.nrPxX - 1);
  d.halfY = 0.5 * (g.nrPxY - 1);
  d.W = (double)g.nrPxX;
  d.H = (double)g.nrPxY;
  d.Elo = g.Elo;
  d.Ehi = g.Ehi;
  d.nrPxX = g.nrPxX;
  // pre-test constants (plan_bit): products only, so each side's
  // rounding is a few units of 2^-53 against margins of 1e-12 and more
  d.pre_sign = g.Elo > 0;
  d.eloC = g.Elo * kFourPi * (1.0 - 1e-12);
  d.ehiC = g.Ehi * kFourPi * (1.0 + 1e-12);
  auto interval = [&](double P0, double px, double half, double N, double *lo, double *hi) {
    const double a = P0 + (-1.0 - half) * px, b = P0 + (N - half) * px;
    const double m = std::fabs(g.P[2]) * 1e-7 + std::fabs(px) * 1e-6 +
                     1e-12 * (std::fabs(a) + std::fabs(b));
    *lo = std::min(a, b) - m;
    *hi = std::max(a, b) + m;
    return std::isfinite(*lo) && std::isfinite(*hi) && px != 0;
  };
  const bool okx = interval(g.P[0], g.pxX, d.halfX, d.W, &d.xlo, &d.xhi);
  const bool oky = interval(g.P[1], g.pxY, d.halfY, d.H, &d.ylo, &d.yhi);
  d.pre_px = okx && oky && std::isfinite(g.P[2]);
  const ConeSetup cs = cone_setup(g);
  d.cone = cs.ok && d.pre_sign;
  for (int i = 0; i < 3; ++i) d.cLab[i] = cs.cLab[i];
  for (int i = 0; i < 6; ++i) d.G0[i] = cs.G0[i];
  d.tolG = cs.tolG;
  return d;
}

DevGeom32 dev_geom32(const Geom &g) {
  DevGeom32 d{};
  for (int i = 0; i < 9; ++i) d.rotT[i] = (float)g.rotT[i];
  for (int i = 0; i < 3; ++i) d.P[i] = (float)g.P[i];
  d.inv_pxX = (float)(1.0 / g.pxX);
  d.inv_pxY = (float)(1.0 / g.pxY);
  d.halfX = (float)(0.5 * (g.nrPxX - 1));
  d.halfY = (float)(0.5 * (g.nrPxY - 1));
  d.W = (float)g.nrPxX;
  d.H = (float)g.nrPxY;
  d.Elo = (float)g.Elo;
  d.Ehi = (float)g.Ehi;
  d.dist_px = (float)std::max(std::fabs(g.P[2]) / g.pxX, std::fabs(g.P[2]) / g.pxY);
  d.P0px = (float)(std::fabs(g.P[0]) / g.pxX);
  d.P1px = (float)(std::fabs(g.P[1]) / g.pxY);
  return d;
}

}  // namespace

// The pre-test plan: hkls in chunks of 64 in file order; within a chunk each
// hkl is evaluated once, together with the first unpaired -hkl after it in
// the same chunk when Elo > 0 (only the member with q_z < 0 can then be
// accepted).  Uploaded once per DB: plan entries, then one int4 per chunk
// (first hkl, first entry, last entry + 1, 0).
int ensure_fwd_plan(lm_db *db, hipStream_t st) {
  if (db->d_fwd_plan) return LM_OK;
  const int n = db->nhkls;
  const bool pair = db->geom.Elo > 0;
  const ConeSetup cs = cone_setup(db->geom);
  static_assert(sizeof(FwdEval) == 64, "plan entries keep the chunk table 16-byte aligned");
  std::vector<FwdEval> ev;
  std::vector<int4> ch;
  // Friedel mates in file order (the first unpaired -h after h), so that
  // chunks can end where no pair is split: a chunk is the longest run of at
  // most 64 hkls that splits no pair, or 64 hkls when every cut would
  std::vector<int> mate(n, -1);
  if (pair) {
    std::vector<bool> taken(n, false);
    for (int i = 0; i < n; ++i) {
      if (taken[i]) continue;
      taken[i] = true;
      const int32_t *a = &db->hkls[(size_t)3 * i];
      if ((a[0] | a[1] | a[2]) == 0) continue;
      for (int j = i + 1; j < n; ++j) {
        const int32_t *b = &db->hkls[(size_t)3 * j];
        if (!taken[j] && b[0] == -a[0] && b[1] == -a[1] && b[2] == -a[2]) {
          mate[i] = j;
          mate[j] = i;
          taken[j] = true;
          break;
        }
      }
    }
  }
  for (int s = 0; s < n;) {
    int e = std::min(n, s + 64);
    for (int cut = e; cut > s && e < n; --cut) {
      bool split = false;
      for (int i = s; i < cut && !split; ++i) split = mate[i] >= cut;
      if (!split) {
        e = cut;
        break;
      }
    }
    int4 c;
    c.x = s;
    c.y = (int)ev.size();
    bool used[64] = {};
    for (int i = s; i < e; ++i) {
      if (used[i - s]) continue;
      used[i - s] = true;
      const int32_t *a = &db->hkls[(size_t)3 * i];
      int partner = -1;
      for (int j = i + 1; pair && j < e && partner < 0; ++j) {
        const int32_t *b = &db->hkls[(size_t)3 * j];
        if (!used[j - s] && b[0] == -a[0] && b[1] == -a[1] && b[2] == -a[2] &&
            (a[0] | a[1] | a[2]) != 0)
          partner = j;
      }
      if (partner >= 0) used[partner - s] = true;
      FwdEval f;
      f.h = a[0];
      f.k = a[1];
      f.l = a[2];
      // without the sign test (Elo <= 0) the hkl is a candidate either way
      const uint64_t bneg = (uint64_t)(i - s);
      const uint64_t bpos = partner >= 0 ? (uint64_t)(partner - s) : (pair ? 64u : bneg);
      f.bits = bneg | (bpos << 32);
      f.alo = f.ahi = f.dcone = f.pad = 0;
      if (cs.ok) {
        // |q|^2 = h^T G0 h of a rotation of the lattice; eps covers
        // M^T M - G0 up to tolG per element: |h^T dG h| <= tolG (|h|_1)^2
        const long double hv[3] = {(long double)a[0], (long double)a[1], (long double)a[2]};
        const long double Gf[3][3] = {{cs.G0[0], cs.G0[3], cs.G0[4]},
                                      {cs.G0[3], cs.G0[1], cs.G0[5]},
                                      {cs.G0[4], cs.G0[5], cs.G0[2]}};
        long double q2 = 0;
        for (int r = 0; r < 3; ++r)
          for (int c = 0; c < 3; ++c) q2 += hv[r] * Gf[r][c] * hv[c];
        const double ql2 = (double)q2;
        if (!(ql2 > 0)) {
          f.alo = -1.0;  // |q| = 0: rejected (qlen == 0)
        } else {
          const double L1 = std::fabs((double)a[0]) + std::fabs((double)a[1]) +
                            std::fabs((double)a[2]);
          const double eps = cs.tolG * L1 * L1 / ql2 + 1e-12;
          const double ql = std::sqrt(ql2);
          f.alo = kHc * ql2 * (1 + eps) / (kFourPi * db->geom.Elo) * (1 + 1e-9) + 1e-9 * ql;
          f.ahi = kHc * ql2 * (1 - eps) / (kFourPi * db->geom.Ehi) * (1 - 1e-9) - 1e-9 * ql;
          const double kappa = cs.cLab[2] - cs.cosg;
          f.dcone = 0.5 * kappa * ql2 * (1 + (kappa > 0 ? eps : -eps)) + 1e-9 * ql2;
        }
      }
      ev.push_back(f);
    }
    c.z = (int)ev.size();
    c.w = 0;
    ch.push_back(c);
    s = e;
  }
  const size_t eb = ev.size() * sizeof(FwdEval), cb = ch.size() * sizeof(int4);
  void *d = nullptr;
  hipError_t e = hipMalloc(&d, std::max<size_t>(eb + cb, 64));
  if (e == hipSuccess && eb) e = hipMemcpyAsync(d, ev.data(), eb, hipMemcpyHostToDevice, st);
  if (e == hipSuccess && cb)
    e = hipMemcpyAsync((char *)d + eb, ch.data(), cb, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) {
    if (d) (void)hipFree(d);
    return hip_fail(e, "forward plan upload");
  }
  db->d_fwd_plan = d;
  db->fwd_plan_entries = (int)ev.size();
  db->fwd_plan_chunks = (int)ch.size();
  return LM_OK;
}

namespace {

void free_table(lm_db *db) {
  release_banded(db);
  Table &t = db->table;
  if (t.entries) (void)hipFree(t.entries);
  if (t.tile_off) (void)hipFree(t.tile_off);
  if (t.counts) (void)hipFree(t.counts);
  {
    std::lock_guard<std::mutex> lk(db->ctx->mu);
    db->ctx->in_use -= t.bytes;
  }
  t = Table{};
}

int reserve(lm_db *db, uint64_t bytes) {
  std::lock_guard<std::mutex> lk(db->ctx->mu);
  if (db->ctx->in_use + bytes > db->ctx->budget) {
    char msg[256];
    snprintf(msg, sizeof msg, "HBM budget exceeded: need %llu more bytes, %llu of %llu in use",
             (unsigned long long)bytes, (unsigned long long)db->ctx->in_use,
             (unsigned long long)db->ctx->budget);
    set_error(msg);
    return LM_ERR_BUDGET;
  }
  db->ctx->in_use += bytes;
  return LM_OK;
}

// Allocate counts + tile offsets for the DB's geometry.
int alloc_table_index(lm_db *db) {
  Table &t = db->table;
  const uint64_t npx = (uint64_t)db->geom.nrPxX * db->geom.nrPxY;
  t.pix_bits = pix_bits_for(npx);
  t.tile_bits = kTileBits;
  t.ntiles = (db->n + (1ull << t.tile_bits) - 1) >> t.tile_bits;
  const uint64_t bytes = (db->n + 64) * 2 + (t.ntiles + 1) * 8;
  int rc = reserve(db, bytes);
  if (rc) return rc;
  t.bytes = bytes;
  hipError_t e = hipMalloc(&t.counts, (db->n + 64) * 2);
  if (e == hipSuccess) e = hipMalloc(&t.tile_off, (t.ntiles + 1) * 8);
  if (e != hipSuccess) {
    free_table(db);
    return hip_fail(e, "hipMalloc(table index)");
  }
  return LM_OK;
}

// Per-tile totals -> offsets (host scan), then allocate the entry array.
int finish_offsets(lm_db *db, hipStream_t st) {
  Table &t = db->table;
  uint64_t *d_sum = nullptr;
  hipError_t e = hipMalloc(&d_sum, t.ntiles * 8);
  if (e != hipSuccess) return hip_fail(e, "hipMalloc(tile sums)");
  hipLaunchKernelGGL(tile_sum_kernel, dim3((unsigned)t.ntiles), dim3(256), 0, st, t.counts, db->n,
                     t.tile_bits, d_sum);
  std::vector<uint64_t> h(t.ntiles + 1);
  e = hipMemcpyAsync(h.data(), d_sum, t.ntiles * 8, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  (void)hipFree(d_sum);
  if (e != hipSuccess) return hip_fail(e, "tile sums");
  uint64_t acc = 0;
  for (uint64_t i = 0; i < t.ntiles; ++i) {
    uint64_t v = h[i];
    h[i] = acc;
    acc += v;
  }
  h[t.ntiles] = acc;
  t.total = acc;
  e = hipMemcpyAsync(t.tile_off, h.data(), (t.ntiles + 1) * 8, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return hip_fail(e, "tile offsets");
  const uint64_t ebytes = (acc + 64) * 4;
  int rc = reserve(db, ebytes);
  if (rc) return rc;
  t.bytes += ebytes;
  e = hipMalloc(&t.entries, ebytes);
  if (e != hipSuccess) return hip_fail(e, "hipMalloc(table entries)");
  e = hipMemsetAsync(t.entries + acc, 0, 64 * 4, st);
  if (e != hipSuccess) return hip_fail(e, "hipMemset(table slack)");
  return LM_OK;
}

}  // namespace

int forward_build(lm_db *db, int mode, lm_fwd_stats *stats) {
  if (mode != LM_FWD_FP64 && mode != LM_FWD_FP32_GUARD) {
    set_error("unknown forward mode");
    return LM_ERR_INVALID;
  }
  const bool guard = mode == LM_FWD_FP32_GUARD;
  hipStream_t st = db->ctx->stream;
  if (db->table.counts) free_table(db);
  int rc = ensure_fwd_plan(db, st);
  if (rc) return rc;
  rc = alloc_table_index(db);
  if (rc) return rc;
  Table &t = db->table;
  const FwdEval *plan = (const FwdEval *)db->d_fwd_plan;
  const int4 *chunks = (const int4 *)(plan + db->fwd_plan_entries);
  const int nch = db->fwd_plan_chunks;
  const DevGeom g = dev_geom(db->geom);
  const DevGeom32 g32 = dev_geom32(db->geom);
  const double *R = db->geom.recip;
  const bool ldsList = db->max_spots <= kLdsSlots;
  // Grid chunks keep each launch under 2^31 threads and bound the spill list.
  const uint64_t chunkBlocks = 1u << 18;  // 32M orientations per launch
  uint32_t *spill = nullptr;
  unsigned long long *d_rec = nullptr;
  // One projection pass when the HBM budget holds every orientation's spot
  // list (n * M words, slot-major per launch chunk) beside the table: the
  // count pass keeps the lists and the write pass copies them; otherwise
  // (no room for the lists, or then none for the table beside them) the
  // write pass projects again.
  uint32_t *keep = nullptr;
  const uint64_t keepBytes = ldsList ? db->n * (uint64_t)db->max_spots * 4 : 0;
  if (keepBytes && reserve(db, keepBytes) == LM_OK) {
    if (hipMalloc(&keep, keepBytes) != hipSuccess) {
      keep = nullptr;
      (void)hipGetLastError();
    }
    if (!keep) {
      std::lock_guard<std::mutex> lk(db->ctx->mu);
      db->ctx->in_use -= keepBytes;
    }
  }
  hipError_t e = hipMalloc(&d_rec, 16);
  if (e == hipSuccess) e = hipMemsetAsync(d_rec, 0, 16, st);
  if (e == hipSuccess && !ldsList)
    e = hipMalloc(&spill, chunkBlocks * kFwdBlock * (uint64_t)db->max_spots * 4);
  if (e != hipSuccess) {
    if (d_rec) (void)hipFree(d_rec);
    free_table(db);
    return hip_fail(e, "hipMalloc(forward scratch)");
  }
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  (void)hipEventRecord(e0, st);
  for (int pass = 0; pass < 2 && rc == LM_OK; ++pass) {
    if (pass == 1) {
      rc = finish_offsets(db, st);
      if (rc == LM_ERR_BUDGET && keep) {
        // no room for the table beside the kept lists: drop them and
        // project again in the write pass
        (void)hipStreamSynchronize(st);
        (void)hipFree(keep);
        keep = nullptr;
        {
          std::lock_guard<std::mutex> lk(db->ctx->mu);
          db->ctx->in_use -= keepBytes;
        }
        rc = finish_offsets(db, st);
      }
      if (rc) break;
    }
    for (uint64_t b0 = 0; b0 * kFwdBlock < db->n; b0 += chunkBlocks) {
      const uint64_t first = b0 * kFwdBlock;
      const uint64_t rows = std::min<uint64_t>(db->n - first, chunkBlocks * kFwdBlock);
      const unsigned blocks = (unsigned)((rows + kFwdBlock - 1) / kFwdBlock);
      // Chunks start on tile boundaries (chunkBlocks*kFwdBlock is a multiple
      // of the tile size), so the in-tile scan never crosses a chunk.
      const double *op = db->d_orients + first * 9;
      uint16_t *cp = t.counts + first;
      const uint64_t *tp = t.tile_off + (first >> t.tile_bits);
      uint32_t *kp = keep ? keep + first * (uint64_t)db->max_spots : nullptr;
#define LM_FWD(W, L, G)                                                                         \
  hipLaunchKernelGGL((fwd_kernel<W, L, G>), dim3(blocks), dim3(kFwdBlock), 0, st, op, rows,      \
                     db->d_hkls, db->nhkls, R[0], R[1], R[2], R[3], R[4], R[5], R[6], R[7], R[8], \
                     g, g32, db->max_spots, t.tile_bits, cp, tp, t.entries, spill, d_rec, kp,     \
                     plan, chunks, nch)
      const int sel = (pass << 2) | (ldsList << 1) | (guard ? 1 : 0);
      if (pass == 0 && keep && ldsList) {
        // count pass keeping the lists: 32 LDS slots, the rest in `keep`
        if (guard)
          LM_FWD(false, kLdsSlotsKeep, true);
        else
          LM_FWD(false, kLdsSlotsKeep, false);
      } else if (pass == 1 && keep) {
        // copy the kept lists (no projection and no list: SLOTS and GUARD
        // do not matter, false keeps the copy kernel free of LDS)
        hipLaunchKernelGGL((fwd_kernel<true, 0, false, true>), dim3(blocks), dim3(kFwdBlock), 0,
                           st, op, rows, db->d_hkls, db->nhkls, R[0], R[1], R[2], R[3], R[4],
                           R[5], R[6], R[7], R[8], g, g32, db->max_spots, t.tile_bits, cp, tp,
                           t.entries, spill, d_rec, kp, plan, chunks, nch);
      } else switch (sel) {
        case 0: LM_FWD(false, 0, false); break;
        case 1: LM_FWD(false, 0, true); break;
        case 2: LM_FWD(false, kLdsSlots, false); break;
        case 3: LM_FWD(false, kLdsSlots, true); break;
        case 4: LM_FWD(true, 0, false); break;
        case 5: LM_FWD(true, 0, true); break;
        case 6: LM_FWD(true, kLdsSlots, false); break;
        default: LM_FWD(true, kLdsSlots, true); break;
      }
#undef LM_FWD
      e = hipGetLastError();
      if (e != hipSuccess) {
        rc = hip_fail(e, "forward kernel launch");
        break;
      }
    }
  }
  (void)hipEventRecord(e1, st);
  unsigned long long nrec[2] = {0, 0};
  e = hipStreamSynchronize(st);
  if (e == hipSuccess) e = hipMemcpy(nrec, d_rec, 16, hipMemcpyDeviceToHost);
  if (spill) (void)hipFree(spill);
  if (keep) {
    (void)hipFree(keep);
    std::lock_guard<std::mutex> lk(db->ctx->mu);
    db->ctx->in_use -= keepBytes;
  }
  (void)hipFree(d_rec);
  if (rc == LM_OK && e != hipSuccess) rc = hip_fail(e, "forward simulation");
  float ms = 0;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) {
    free_table(db);
    return rc;
  }
  t.ready = true;
  rc = build_banded(db, true);
  if (rc) {
    free_table(db);
    return rc;
  }
  if (stats) {
    stats->n_orients = db->n;
    stats->n_spots = t.total;
    stats->n_pairs = db->n * (uint64_t)db->nhkls;
    stats->n_recomputed = nrec[0];
    stats->ms = ms;
    stats->n_projected = nrec[1];
  }
  return LM_OK;
}

int import_table(lm_db *db, const uint16_t *records, uint64_t count) {
  if (count != db->n || !records) {
    set_error("import_table: record count must equal the DB size");
    return LM_ERR_INVALID;
  }
  hipStream_t st = db->ctx->stream;
  if (db->table.counts) free_table(db);
  int rc = alloc_table_index(db);
  if (rc) return rc;
  Table &t = db->table;
  const int M = db->max_spots;
  const uint64_t rec = (uint64_t)(1 + 2 * M);
  const uint64_t chunk = std::max<uint64_t>(1, (256ull << 20) / (rec * 2));
  uint16_t *d_rec = nullptr;
  uint32_t *d_bad = nullptr;
  hipError_t e = hipMalloc(&d_rec, chunk * rec * 2);
  if (e == hipSuccess) e = hipMalloc(&d_bad, 4);
  if (e == hipSuccess) e = hipMemsetAsync(d_bad, 0, 4, st);
  for (int pass = 0; pass < 2 && e == hipSuccess && rc == LM_OK; ++pass) {
    if (pass == 1) {
      uint32_t bad = 0;
      e = hipMemcpyAsync(&bad, d_bad, 4, hipMemcpyDeviceToHost, st);
      if (e == hipSuccess) e = hipStreamSynchronize(st);
      if (e != hipSuccess) break;
      if (bad) {
        set_error("import_table: a record has count > MaxNrLaueSpots or a pixel outside the "
                  "detector");
        rc = LM_ERR_INVALID;
        break;
      }
      rc = finish_offsets(db, st);
      if (rc) break;
    }
    for (uint64_t r0 = 0; r0 < count && e == hipSuccess; r0 += chunk) {
      const uint64_t rows = std::min(chunk, count - r0);
      e = hipMemcpyAsync(d_rec, records + r0 * rec, rows * rec * 2, hipMemcpyHostToDevice, st);
      if (e != hipSuccess) break;
      const unsigned blocks = (unsigned)((rows + 255) / 256);
      if (pass == 0)
        hipLaunchKernelGGL(import_count_kernel, dim3(blocks), dim3(256), 0, st, d_rec, rows, r0, M,
                           db->geom.nrPxX, db->geom.nrPxY, t.counts, d_bad);
      else
        hipLaunchKernelGGL(import_write_kernel, dim3(blocks), dim3(256), 0, st, d_rec, rows, r0,
                           M, db->geom.nrPxX, t.tile_bits, t.counts, t.tile_off, t.entries);
      e = hipGetLastError();
      if (e == hipSuccess) e = hipStreamSynchronize(st);
    }
  }
  if (d_rec) (void)hipFree(d_rec);
  if (d_bad) (void)hipFree(d_bad);
  if (rc == LM_OK && e != hipSuccess) rc = hip_fail(e, "import_table");
  if (rc) {
    free_table(db);
    return rc;
  }
  t.ready = true;
  rc = build_banded(db, true);
  if (rc) {
    free_table(db);
    return rc;
  }
  return LM_OK;
}

int export_table(lm_db *db, uint64_t first, uint64_t count, uint16_t *out) {
  Table &t = db->table;
  if (!t.ready) {
    set_error("export_table: no table (run lm_db_forward or lm_db_import_table)");
    return LM_ERR_STATE;
  }
  if (first > db->n || count > db->n - first || (!out && count)) {
    set_error("export_table: row range outside the DB");
    return LM_ERR_INVALID;
  }
  hipStream_t st = db->ctx->stream;
  const int M = db->max_spots;
  const uint64_t rec = (uint64_t)(1 + 2 * M);
  const uint64_t chunk = std::max<uint64_t>(1, (256ull << 20) / (rec * 2));
  uint16_t *d_rec = nullptr;
  hipError_t e = hipMalloc(&d_rec, std::min(chunk, std::max<uint64_t>(count, 1)) * rec * 2);
  for (uint64_t r0 = 0; r0 < count && e == hipSuccess; r0 += chunk) {
    const uint64_t rows = std::min(chunk, count - r0);
    hipLaunchKernelGGL(export_kernel, dim3((unsigned)((rows + 255) / 256)), dim3(256), 0, st,
                       d_rec, rows, first + r0, M, db->geom.nrPxX, t.tile_bits, t.counts,
                       t.tile_off, t.entries);
    e = hipGetLastError();
    if (e == hipSuccess)
      e = hipMemcpyAsync(out + r0 * rec, d_rec, rows * rec * 2, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
  }
  if (d_rec) (void)hipFree(d_rec);
  if (e != hipSuccess) return hip_fail(e, "export_table");
  return LM_OK;
}

void release_table(lm_db *db) {
  if (db->table.counts) free_table(db);
}

int table_checksum(lm_db *db, uint64_t *sum) {
  const Table &t = db->table;
  if (!t.ready) {
    set_error("checksum: no table");
    return LM_ERR_STATE;
  }
  hipStream_t st = db->ctx->stream;
  unsigned long long *d = nullptr;
  hipError_t e = hipMalloc(&d, 8);
  if (e == hipSuccess) e = hipMemsetAsync(d, 0, 8, st);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(checksum_kernel, dim3(4096), dim3(256), 0, st, t.counts, db->n, t.entries,
                       t.total, d);
    e = hipGetLastError();
  }
  unsigned long long h = 0;
  if (e == hipSuccess) e = hipMemcpyAsync(&h, d, 8, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (d) (void)hipFree(d);
  if (e != hipSuccess) return hip_fail(e, "table checksum");
  *sum = h;
  return LM_OK;
}

int selftest_fp64(hipStream_t st, const double *a, const double *b, uint64_t n, double *q,
                  double *s) {
  double *d = nullptr;
  hipError_t e = hipMalloc(&d, std::max<uint64_t>(n, 1) * 32);
  if (e != hipSuccess) return hip_fail(e, "hipMalloc(selftest)");
  double *da = d, *db = d + n, *dq = d + 2 * n, *ds = d + 3 * n;
  e = hipMemcpyAsync(da, a, n * 8, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(db, b, n * 8, hipMemcpyHostToDevice, st);
  if (e == hipSuccess && n) {
    hipLaunchKernelGGL(selftest_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, da,
                       db, n, dq, ds);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(q, dq, n * 8, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipMemcpyAsync(s, ds, n * 8, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  (void)hipFree(d);
  if (e != hipSuccess) return hip_fail(e, "selftest_fp64");
  return LM_OK;
}


int launch_fwd_score(lm_db *db, int B, bool guard, const uint8_t *img, const Pred *pred,
                     uint64_t *hits, uint64_t capPer, uint32_t *counters,
                     unsigned long long *d_stats, uint32_t *spill, uint64_t spillRows,
                     hipStream_t st) {
  if (db->max_spots > kLdsSlots) {
    set_error("on-the-fly scoring keeps at most 64 spots per orientation (MaxNrLaueSpots)");
    return LM_ERR_INVALID;
  }
  int rc = ensure_fwd_plan(db, st);
  if (rc) return rc;
  const FwdEval *plan = (const FwdEval *)db->d_fwd_plan;
  const int4 *chunks = (const int4 *)(plan + db->fwd_plan_entries);
  const int nch = db->fwd_plan_chunks;
  const DevGeom g = dev_geom(db->geom);
  const DevGeom32 g32 = dev_geom32(db->geom);
  const double *R = db->geom.recip;
  // 32M orientations per launch, or as many as the spill of list slots
  // 32..M-1 holds (M > 32)
  const uint64_t chunkRows = db->max_spots > kLdsSlotsKeep
                                 ? spillRows / kFwdBlock * kFwdBlock
                                 : (uint64_t)(1u << 18) * kFwdBlock;
  if (chunkRows == 0 || (db->max_spots > kLdsSlotsKeep && !spill)) {
    set_error("on-the-fly scoring: no spill buffer for list slots past 32");
    return LM_ERR_INVALID;
  }
  for (uint64_t row0 = 0; row0 < db->n; row0 += chunkRows) {
    const uint64_t rows = std::min<uint64_t>(db->n - row0, chunkRows);
    const unsigned blocks = (unsigned)((rows + kFwdBlock - 1) / kFwdBlock);
#define LM_FS(BB, GG)                                                                           \
  hipLaunchKernelGGL((fwd_score_kernel<BB, GG>), dim3(blocks), dim3(kFwdBlock), 0, st,          \
                     db->d_orients, db->n, row0, db->d_hkls, R[0], R[1], R[2], R[3], R[4], R[5], \
                     R[6], R[7], R[8], g, g32, db->max_spots, plan, chunks, nch, img, pred, hits, \
                     capPer, counters, d_stats, spill)
#define LM_FS2(BB) \
  if (guard) LM_FS(BB, true); else LM_FS(BB, false)
    switch (B) {
      case 1: LM_FS2(1); break;
      case 4: LM_FS2(4); break;
      case 8: LM_FS2(8); break;
      case 16: LM_FS2(16); break;
      case 32: LM_FS2(32); break;
      case 64: LM_FS2(64); break;
      default: set_error("on-the-fly scoring: unsupported batch width"); return LM_ERR_INVALID;
    }
#undef LM_FS2
#undef LM_FS
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "on-the-fly scoring launch");
  }
  return LM_OK;
}

int launch_fwd_exact(lm_db *db, const uint64_t *xids, uint64_t k, const uint8_t *imgs,
                     uint64_t npx, const Pred *pred, double *out, hipStream_t st) {
  if (k == 0) return LM_OK;
  int rc = ensure_fwd_plan(db, st);
  if (rc) return rc;
  const FwdEval *plan = (const FwdEval *)db->d_fwd_plan;
  const int4 *chunks = (const int4 *)(plan + db->fwd_plan_entries);
  const DevGeom g = dev_geom(db->geom);
  const DevGeom32 g32 = dev_geom32(db->geom);
  const double *R = db->geom.recip;
  hipLaunchKernelGGL(fwd_exact_kernel, dim3((unsigned)((k + kFwdBlock - 1) / kFwdBlock)),
                     dim3(kFwdBlock), 0, st, db->d_orients, db->d_hkls, R[0], R[1], R[2], R[3],
                     R[4], R[5], R[6], R[7], R[8], g, g32, db->max_spots, plan, chunks,
                     db->fwd_plan_chunks, xids, k, imgs, npx, pred, out);
  const hipError_t e = hipGetLastError();
  return e == hipSuccess ? LM_OK : hip_fail(e, "on-the-fly exact pass launch");
}

}  // namespace lm
