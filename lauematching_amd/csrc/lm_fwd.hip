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
};

// Pixel from a detector coordinate: the reference keeps (int)t within
// [0, N-1] (src/LaueMatchingCPU.c:627-632), i.e. exactly -1 < t < N.
__device__ __forceinline__ bool to_pixel(double t, double n, int *ip) {
  if (!(t > -1.0 && t < n)) return false;
  *ip = (int)t;
  return true;
}

// src/LaueMatchingCPU.c:604-636 for one (orientation, hkl).
__device__ __forceinline__ bool project64(const double *M, double h, double k, double l,
                                          const DevGeom &g, uint32_t *pix) {
  double q0 = 0, q1 = 0, q2 = 0;
  q0 += M[0] * h; q0 += M[1] * k; q0 += M[2] * l;
  q1 += M[3] * h; q1 += M[4] * k; q1 += M[5] * l;
  q2 += M[6] * h; q2 += M[7] * k; q2 += M[8] * l;
  double qlen = __builtin_sqrt(q0 * q0 + q1 * q1 + q2 * q2);
  if (qlen == 0) return false;
  double u0 = q0 / qlen, u1 = q1 / qlen, u2 = q2 / qlen;
  double td = 2 * u2;
  double k0 = 0 - td * u0, k1 = 0 - td * u1, k2 = 1.0 - td * u2;
  double x = 0, y = 0, z = 0;
  x += g.rotT[0] * k0; x += g.rotT[1] * k1; x += g.rotT[2] * k2;
  y += g.rotT[3] * k0; y += g.rotT[4] * k1; y += g.rotT[5] * k2;
  z += g.rotT[6] * k0; z += g.rotT[7] * k1; z += g.rotT[8] * k2;
  if (z <= 0) return false;
  double X = x * g.P[2] / z;
  double Y = y * g.P[2] / z;
  int ipx, ipy;
  if (!to_pixel((X - g.P[0]) / g.pxX + g.halfX, g.W, &ipx)) return false;
  if (!to_pixel((Y - g.P[1]) / g.pxY + g.halfY, g.H, &ipy)) return false;
  double E = kHc * qlen / (kFourPi * (-u2));
  if (E < g.Elo || E > g.Ehi) return false;
  *pix = (uint32_t)ipy * (uint32_t)g.nrPxX + (uint32_t)ipx;
  return true;
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
};

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
  double M[9];
  if (live && !FROMKEEP) {
    double t[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) t[i] = orients[o * 9 + i];
    // MatrixMultF33 (src/LaueMatchingHeaders.h:349-357): (a*b + c*d) + e*f
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      M[r * 3 + 0] = t[r * 3 + 0] * r00 + t[r * 3 + 1] * r10 + t[r * 3 + 2] * r20;
      M[r * 3 + 1] = t[r * 3 + 0] * r01 + t[r * 3 + 1] * r11 + t[r * 3 + 2] * r21;
      M[r * 3 + 2] = t[r * 3 + 0] * r02 + t[r * 3 + 1] * r12 + t[r * 3 + 2] * r22;
    }
  }
  float M32[9];
  if (GUARD) {
#pragma unroll
    for (int i = 0; i < 9; ++i) M32[i] = (float)M[i];
  }
  uint32_t nrec = 0, nproj = 0;
  uint32_t *const colL = lst + threadIdx.x;  // slot s at colL[s * kFwdBlock], s < SLOTS
  uint32_t *const colG = SLOTS == 0 ? spill + o : keep + (uint64_t)SLOTS * n + o;
  const uint64_t gStride = SLOTS == 0 ? (uint64_t)gridDim.x * kFwdBlock : n;
  auto get = [&](int s) -> uint32_t {
    return s < SLOTS ? colL[s * kFwdBlock] : colG[(uint64_t)(s - SLOTS) * gStride];
  };
  auto put = [&](int s, uint32_t v) {
    if (s < SLOTS)
      colL[s * kFwdBlock] = v;
    else
      colG[(uint64_t)(s - SLOTS) * gStride] = v;
  };
  int ns = live ? 0 : maxSpots;
  if (FROMKEEP) ns = live ? (int)counts[o] : 0;
  // hkls in chunks of 64 (chunks[c] = first hkl, first and last+1 plan
  // entry): first the division-free pre-test of the chunk's plan entries
  // (wave-uniform loop, scalar loads) marks the candidates in a 64-bit mask,
  // then each lane projects only its own candidates, in hkl order (ctz), so
  // the wave runs as many projections as its busiest lane has candidates
  // instead of one per hkl; first unseen pixel still wins.
  // Chunks are taken four at a time (256 hkls): the pre-test fills four
  // masks, then each lane walks the union of its four masks in hkl order, so
  // the wave runs as many projections as its busiest lane has candidates in
  // the 256 hkls, rather than the sum over chunks of each chunk's busiest.
  constexpr int kSuper = 4;
  for (int c0 = 0; c0 < (FROMKEEP ? 0 : nchunks); c0 += kSuper) {
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
#pragma unroll 2
        for (int e = ch.y; e < ch.z; ++e) {
          const int bit = plan_bit(M, plan[e], g);
          m |= bit < 64 ? 1ull << bit : 0ull;
        }
        cand[k] = ns < maxSpots ? m : 0ull;
      }
    }
    for (;;) {
      // this lane's next candidate: lowest bit of its first nonempty mask
      int hi = -1;
#pragma unroll
      for (int k = kSuper - 1; k >= 0; --k)
        if (cand[k] != 0) hi = base[k] + __builtin_ctzll(cand[k]);
      bool taken = false;
#pragma unroll
      for (int k = 0; k < kSuper; ++k)
        if (!taken && cand[k] != 0) {
          cand[k] &= cand[k] - 1;
          taken = true;
        }
      if (hi < 0 || ns >= maxSpots) break;
      ++nproj;
      const int hI = hkls[3 * hi + 0], kI = hkls[3 * hi + 1], lI = hkls[3 * hi + 2];
      uint32_t pix = 0;
      bool ok;
      if (GUARD) {
        const int v = project32(M32, hI, kI, lI, g32, (uint32_t)g.nrPxX, &pix);
        if (v == 2) {
          ok = project64(M, (double)hI, (double)kI, (double)lI, g, &pix);
          ++nrec;
        } else {
          ok = v == 1;
        }
      } else {
        ok = project64(M, (double)hI, (double)kI, (double)lI, g, &pix);
      }
      if (ok) {
        bool dup = false;
        for (int s = 0; s < ns; ++s)
          if (get(s) == pix) { dup = true; break; }
        if (!dup) {
          put(ns, pix);
          ++ns;
        }
      }
    }
  }
  if (!WRITE) {
    if (live) counts[o] = (uint16_t)ns;
    if (keep && live)
      for (int s = 0; s < (SLOTS == kLdsSlotsKeep ? min(ns, SLOTS) : ns); ++s)
        keep[(uint64_t)s * n + o] = get(s);
    // recomputed[0]: pairs recomputed in float64; recomputed[1]: pairs that
    // passed the pre-test and were projected
    uint32_t w = nrec, c = nproj;
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) {
      w += __shfl_xor(w, d);
      c += __shfl_xor(c, d);
    }
    if (__lane_id() == 0) {
      if (w) atomicAdd(recomputed, (unsigned long long)w);
      if (c) atomicAdd(recomputed + 1, (unsigned long long)c);
    }
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
  std::vector<FwdEval> ev;
  std::vector<int4> ch;
  for (int s = 0; s < n; s += 64) {
    const int e = std::min(n, s + 64);
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
      ev.push_back(f);
    }
    c.z = (int)ev.size();
    c.w = 0;
    ch.push_back(c);
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

}  // namespace lm
