// lm_median.hip - the median-filter background of the preprocessing
// (compute_background, laue_index/preprocess.py:43-64), on the device.
//
// Without diplib the reference filters with
// scipy.ndimage.median_filter(bg, size=filter_radius), median_passes times
// (preprocess.py:61-65): a square window of filter_radius x filter_radius
// pixels whose first row and column lie filter_radius // 2 before the output
// pixel, the image extended by half-sample reflection (scipy's 'reflect',
// ... c b a | a b c ... | c b a ..., periodic with period 2n when the window
// is wider than the image: scipy's result while filter_radius < 8n on each
// axis, past which scipy reads outside its line buffers), and as output the
// element of 0-based rank
// (filter_radius^2) // 2 of the window's values (no averaging for even
// sizes).  This file computes exactly that.
//
// 1. Dense ranks.  The float64 values are sorted (radix sort of an
//    order-preserving 64-bit key, on the key bits that differ, with the
//    pixel index as payload); equal values share one rank, so the image
//    becomes a uint32 rank image over D distinct values plus a table of
//    those values.  The median of the ranks is the rank of the median, and
//    a pass's output holds ranks of the same table, so every pass works on
//    ranks and the table maps the last one back to float64.
// 2. The running median (Huang's): each column (or column segment) slides
//    its window down; per row the leaving row's filter_radius ranks leave an
//    LDS histogram of 16-bit counts (two per word), the entering row's
//    enter, and the median bin moves from where it was by wavefront scans
//    of 64 bins (DPP prefix sums; group sums of 64 bins from 4096 bins on).
//    D <= 4096 (integer-count detector frames): one wavefront per column,
//    no barrier (median_wave_kernel).  D <= 65536: one 256-thread workgroup
//    per column, the histogram over all ranks.  D > 65536: the histogram
//    over the rank's top 15 bits, and full-resolution counts of 32768 ranks
//    around the median (the band) give the low bits; the band is rebuilt
//    from the window only when the median's coarse bin leaves it.
//
// Limits: filter_radius <= 255 (a window of at most 65025 elements keeps
// every count in 16 bits) and at most 2^24 pixels (so ranks have at most 24
// bits, 9 of them from the band).  NaN values are not ordered (the
// reference's filters leave them undefined as well).
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <vector>

#include "lm_internal.h"
#include "lm_walk16.h"

namespace lm {
namespace {

constexpr int kThreads = 256;
constexpr int kMedThreads = 256;

unsigned grid_for(uint64_t n) {
  return (unsigned)std::min<uint64_t>((n + kThreads - 1) / kThreads, 1u << 20);
}

// float64 -> 64-bit key with the values' order (+0 for both zeros)
__device__ __forceinline__ uint64_t order_key(double v) {
  uint64_t b = v == 0.0 ? 0ull : (uint64_t)__double_as_longlong(v);
  return (b >> 63) ? ~b : (b | (1ull << 63));
}

__device__ __forceinline__ double key_value(uint64_t k) {
  const uint64_t b = (k >> 63) ? (k & ~(1ull << 63)) : ~k;
  return __longlong_as_double((long long)b);
}

__global__ void key_kernel(const double *__restrict__ v, uint64_t npx, uint64_t *__restrict__ keys,
                           uint32_t *__restrict__ idx) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < npx;
       i += (uint64_t)gridDim.x * blockDim.x) {
    keys[i] = order_key(v[i]);
    idx[i] = (uint32_t)i;
  }
}

// 1 where a sorted key starts a run of equal keys
__global__ void head_kernel(const uint64_t *__restrict__ sk, uint64_t n, uint32_t *__restrict__ f) {
  for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n;
       j += (uint64_t)gridDim.x * blockDim.x)
    f[j] = (j == 0 || sk[j] != sk[j - 1]) ? 1u : 0u;
}

// rank image and the table of distinct values (scan = inclusive sum of heads)
__global__ void rank_kernel(const uint64_t *__restrict__ sk, const uint32_t *__restrict__ sidx,
                            const uint32_t *__restrict__ scan, uint64_t n,
                            uint32_t *__restrict__ rank, double *__restrict__ table) {
  for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n;
       j += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t r = scan[j] - 1;
    rank[sidx[j]] = r;
    if (j == 0 || sk[j] != sk[j - 1]) table[r] = key_value(sk[j]);
  }
}

__global__ void unrank_kernel(const uint32_t *__restrict__ rank, const double *__restrict__ table,
                              uint32_t D, uint64_t npx, double *__restrict__ out) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < npx;
       i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = table[min(rank[i], D - 1)];  // ranks are < D; the clamp keeps reads in the table
}

// OR of key ^ keys[0] over the keys (one atomic per block): the bits in
// which they differ
__global__ void key_diff_kernel(const uint64_t *__restrict__ keys, uint64_t n,
                                unsigned long long *__restrict__ acc) {
  __shared__ unsigned long long part[kThreads / 64];
  const uint64_t k0 = keys[0];
  unsigned long long d = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    d |= keys[i] ^ k0;
#pragma unroll
  for (int s = 32; s > 0; s >>= 1) d |= __shfl_xor(d, s);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = d;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kThreads / 64; ++w) d |= part[w];
    if (d) atomicOr(acc, d);
  }
}

// scipy's 'reflect' extension: half-sample symmetric, period 2n
__device__ __forceinline__ int reflect(int i, int n) {
  const int p = 2 * n;
  i %= p;
  if (i < 0) i += p;
  return i < n ? i : p - 1 - i;
}

__device__ __forceinline__ uint32_t lane_value(uint32_t x, int l) {
  return (uint32_t)__builtin_amdgcn_readlane((int)x, l);
}

// Wave-uniform search for the bin holding the element of rank k in counts
// h(0..NB-1), grouped by 64.  On entry and on exit below = sum of h(j < m);
// on exit below <= k < below + h(m).  From the current bin it scans the rest
// of its group, then whole groups, then the bins of the group found.  With
// GROUPS the group sums g(0..NB/64-1) are kept (64 groups per step), else a
// group's sum is the scan of its bins.  Every loop ends at the first or last
// group.
template <bool GROUPS, class HF, class GF>
__device__ void seek(const HF &h, const GF &g, uint32_t NB, uint32_t k, uint32_t &m,
                     uint32_t &below) {
  const int lane = __lane_id();
  const int NG = (int)(NB >> 6);
  int grp;  // group holding the answer, below = count before its first bin
  if (below > k) {
    // down: bins m-1, m-2, ... of m's group, then groups before it
    const int g0 = (int)(m >> 6) << 6;
    const int bin = (int)m - 1 - lane;
    const uint32_t v = bin >= g0 ? h((uint32_t)bin) : 0u;
    const uint32_t p = wave_scan_dpp(v);
    const uint64_t hit = __ballot(p >= below - k);
    if (hit) {
      const int l = __builtin_ctzll(hit);
      m = (uint32_t)((int)m - 1 - l);
      below -= lane_value(p, l);
      return;
    }
    below -= lane_value(p, 63);
    int gg = (int)(m >> 6) - 1;
    for (;;) {
      if (gg < 0) {  // not reached while below counts the bins under m
        m = 0;
        below = 0;
        return;
      }
      if (!GROUPS) {  // group gg's bins, descending
        const uint32_t hv = h((uint32_t)(gg << 6) + 63 - lane);
        const uint32_t hp = wave_scan_dpp(hv);
        const uint64_t bhit = __ballot(hp >= below - k);
        if (bhit) {
          const int l = __builtin_ctzll(bhit);
          m = (uint32_t)(gg << 6) + 63 - l;
          below -= lane_value(hp, l);
          return;
        }
        below -= lane_value(hp, 63);
        --gg;
        continue;
      }
      const int gi = gg - lane;
      const uint32_t sv = gi >= 0 ? g((uint32_t)gi) : 0u;
      const uint32_t sp = wave_scan_dpp(sv);
      const uint64_t ghit = __ballot(sp >= below - k);
      if (ghit) {
        const int l = __builtin_ctzll(ghit);
        grp = gg - l;
        below -= lane_value(sp, l);
        break;
      }
      below -= lane_value(sp, 63);
      gg -= 64;
    }
  } else {
    if (below + h(m) > k) return;
    // up: bins m.. of m's group, then groups after it
    const int g0 = (int)(m >> 6) << 6;
    const int bin = g0 + lane;
    const uint32_t v = bin >= (int)m ? h((uint32_t)bin) : 0u;
    const uint32_t p = wave_scan_dpp(v);
    const uint64_t hit = __ballot(below + p > k);
    if (hit) {
      const int l = __builtin_ctzll(hit);
      m = (uint32_t)(g0 + l);
      below += lane_value(p, l) - lane_value(v, l);
      return;
    }
    below += lane_value(p, 63);
    int gg = (int)(m >> 6) + 1;
    for (;;) {
      if (gg >= NG) {  // not reached while the counts sum to more than k
        m = NB - 1;
        return;
      }
      if (!GROUPS) {  // group gg's bins
        const uint32_t hv = h((uint32_t)(gg << 6) + lane);
        const uint32_t hp = wave_scan_dpp(hv);
        const uint64_t bhit = __ballot(below + hp > k);
        if (bhit) {
          const int l = __builtin_ctzll(bhit);
          m = (uint32_t)(gg << 6) + l;
          below += lane_value(hp, l) - lane_value(hv, l);
          return;
        }
        below += lane_value(hp, 63);
        ++gg;
        continue;
      }
      const int gi = gg + lane;
      const uint32_t sv = gi < NG ? g((uint32_t)gi) : 0u;
      const uint32_t sp = wave_scan_dpp(sv);
      const uint64_t ghit = __ballot(below + sp > k);
      if (ghit) {
        const int l = __builtin_ctzll(ghit);
        grp = gg + l;
        below += lane_value(sp, l) - lane_value(sv, l);
        break;
      }
      below += lane_value(sp, 63);
      gg += 64;
    }
  }
  // the bins of group grp
  const uint32_t v = h((uint32_t)(grp << 6) + lane);
  const uint32_t p = wave_scan_dpp(v);
  const uint64_t hit = __ballot(below + p > k);
  const int l = hit ? __builtin_ctzll(hit) : 63;
  m = (uint32_t)(grp << 6) + l;
  below += lane_value(p, l) - lane_value(v, l);
}

// One workgroup per (column, row segment).  NBITS: histogram bins 2^NBITS
// over rank >> shift; FINE: the low `shift` bits of the rank come from a
// band of 2^15 full-resolution counts around the median, rebuilt from the
// window only when the median's coarse bin leaves it.  The
// rows a segment touches are reflected once into an LDS table; window scans
// give each thread one column (of 128 or 256) and stride over rows; the
// update loads run kDepth rows ahead of the histogram updates.  Every
// wavefront tracks the median bin itself (the same search over the same
// counts), so a row costs two barriers: after the updates and after the
// searches.
constexpr int kMaxSegRows = 1024;
constexpr int kDepth = 4;
constexpr uint32_t kBand = 1u << 15;

template <int NBITS, bool FINE>
__global__ void __launch_bounds__(kMedThreads)
median_column_kernel(const uint32_t *__restrict__ in, uint32_t *__restrict__ out, int W, int H,
                     int s, int seg_rows, int shift) {
  constexpr uint32_t NB = 1u << NBITS;
  constexpr bool GROUPS = NBITS >= 12;  // group sums pay once scans get long
  // frames with few distinct values (integer counts) crowd into few groups:
  // their group counts are summed per wavefront first
  constexpr bool AGG = GROUPS && NBITS <= 14;
  __shared__ uint32_t hist[NB / 2];
  __shared__ uint32_t grpc[GROUPS ? NB / 64 : 1];
  // FINE: full-resolution counts of ranks band_lo .. band_lo + kBand - 1
  // (16-bit, two per word), placed around the median's coarse bin
  __shared__ uint32_t band[FINE ? kBand / 2 : 1];
  __shared__ int rowtab[kMaxSegRows + 256];
  __shared__ int delta[kMedThreads / 64];

  uint32_t b = blockIdx.x;
  if ((gridDim.x & 7) == 0) b = (b & 7) * (gridDim.x >> 3) + (b >> 3);  // a band per XCD
  const int x = (int)(b % (uint32_t)W);
  const int y0 = (int)(b / (uint32_t)W) * seg_rows;
  if (y0 >= H) return;  // an empty last segment (whole workgroup)
  const int y1 = min(H, y0 + seg_rows);
  const int tid = threadIdx.x, lane = tid & 63;
  const int lo = s >> 1;
  const uint32_t k = (uint32_t)(s * s) >> 1;
  const uint32_t fmask = (1u << shift) - 1u;
  const auto hcount = [&](uint32_t j) { return (hist[j >> 1] >> ((j & 1u) << 4)) & 0xffffu; };
  const auto gcount = [&](uint32_t j) { return grpc[j]; };
  const auto bcount = [&](uint32_t j) { return (band[j >> 1] >> ((j & 1u) << 4)) & 0xffffu; };

  // rowtab[i] = image row of window row y0 - lo + i
  for (int i = tid; i < y1 - y0 + s; i += kMedThreads) rowtab[i] = reflect(y0 - lo + i, H);
  for (uint32_t i = tid; i < NB / 2; i += kMedThreads) hist[i] = 0;
  if (GROUPS)
    for (uint32_t i = tid; i < NB / 64; i += kMedThreads) grpc[i] = 0;
  // window scans: this thread's column sdx and first row sdy
  const int cw = s <= 128 ? 128 : 256, rw = kMedThreads / cw;
  const int sdx = tid % cw, sdy = tid / cw;
  const bool scol = sdx < s;
  const size_t wcol = scol ? (size_t)reflect(x - lo + sdx, W) : 0;
  __syncthreads();
  // the window of row y0
  if (scol)
    for (int dy = sdy; dy < s; dy += rw) {
      const uint32_t v = in[(size_t)rowtab[dy] * W + wcol] >> shift;
      atomicAdd(&hist[v >> 1], 1u << ((v & 1u) << 4));
      if (GROUPS) atomicAdd(&grpc[v >> 6], 1u);
    }
  __syncthreads();
  uint32_t m = 0, below = 0;
  const uint32_t nbc = kBand >> shift;  // coarse bins the band covers
  uint32_t c0 = 0, band_lo = 0x80000000u;  // no rank is in the band yet
  bool band_ok = false;
  seek<GROUPS>(hcount, gcount, NB, k, m, below);

  // this thread's updates per row: slots j = tid and tid + 256 of the s
  // removals (j < s) and s additions
  size_t col[2];
  bool act[2], add[2];
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const int j = tid + u * kMedThreads;
    act[u] = j < 2 * s;
    add[u] = j >= s;
    col[u] = act[u] ? (size_t)reflect(x - lo + (add[u] ? j - s : j), W) : 0;
  }
  // entering row y: window row y - 1 - lo leaves, y - 1 - lo + s enters
  const auto load_step = [&](int y, uint32_t (&v)[2]) {
    const int rl = rowtab[y - 1 - y0], re = rowtab[y - 1 - y0 + s];
#pragma unroll
    for (int u = 0; u < 2; ++u)
      if (act[u]) v[u] = in[(size_t)(add[u] ? re : rl) * W + col[u]];
  };
  const auto update = [&](const uint32_t (&v)[2]) {
    int c = 0;
    // AGG: the group of the wave's first lane gets one atomic for all its
    // lanes (one address per lane would serialize)
    uint32_t gf[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const uint32_t gl = act[u] ? (v[u] >> shift) >> 6 : 0xffffffffu;
      gf[u] = AGG ? (uint32_t)__builtin_amdgcn_readfirstlane((int)gl) : 0xffffffffu;
      if (AGG && gf[u] != 0xffffffffu) {
        const int net = __popcll(__ballot(gl == gf[u] && add[u])) -
                        __popcll(__ballot(gl == gf[u] && !add[u]));
        if (lane == 0 && net) atomicAdd(&grpc[gf[u]], (uint32_t)net);
      }
    }
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const uint32_t r = v[u], bin = r >> shift;
      if (act[u]) {
        const uint32_t inc = 1u << ((bin & 1u) << 4);
        atomicAdd(&hist[bin >> 1], add[u] ? inc : 0u - inc);
        if (GROUPS && (bin >> 6) != gf[u]) atomicAdd(&grpc[bin >> 6], add[u] ? 1u : 0xffffffffu);
        if (FINE) {
          const uint32_t o = r - band_lo;
          if (o < kBand) {
            const uint32_t binc = 1u << ((o & 1u) << 4);
            atomicAdd(&band[o >> 1], add[u] ? binc : 0u - binc);
          }
        }
      }
      c += __popcll(__ballot(act[u] && add[u] && bin < m));
      c -= __popcll(__ballot(act[u] && !add[u] && bin < m));
    }
    if (lane == 0) delta[tid >> 6] = c;
    __syncthreads();
    int dsum = 0;
#pragma unroll
    for (int w = 0; w < kMedThreads / 64; ++w) dsum += delta[w];
    below = (uint32_t)((int)below + dsum);
    seek<GROUPS>(hcount, gcount, NB, k, m, below);
  };
  // the output of row y (m and below current)
  const auto finish = [&](int y) {
    if (!FINE) {
      if (tid == 0) out[(size_t)y * W + x] = m;
      __syncthreads();
      return;
    }
    if (!band_ok || m - c0 >= nbc) {  // move the band to centre coarse bin m
      c0 = m > nbc / 2 ? m - nbc / 2 : 0;
      band_lo = c0 << shift;
      band_ok = true;
      for (uint32_t i = tid; i < kBand / 2; i += kMedThreads) band[i] = 0;
      __syncthreads();
      if (scol)
        for (int dy = sdy; dy < s; dy += rw) {
          const uint32_t o = in[(size_t)rowtab[y - y0 + dy] * W + wcol] - band_lo;
          if (o < kBand) atomicAdd(&band[o >> 1], 1u << ((o & 1u) << 4));
        }
      __syncthreads();
    }
    if (tid < 64) {  // the rank within coarse bin m: its 2^shift band counts
      const uint32_t kk = k - below, off = (m - c0) << shift, nf = fmask + 1u;
      uint32_t f = nf - 1, fb = 0;
      for (uint32_t j0 = 0; j0 < nf; j0 += 64) {
        const uint32_t v = j0 + lane < nf ? bcount(off + j0 + lane) : 0u;
        const uint32_t p = wave_scan_dpp(v);
        const uint64_t hit = __ballot(fb + p > kk);
        if (hit) {
          f = j0 + __builtin_ctzll(hit);
          break;
        }
        fb += lane_value(p, 63);
      }
      if (tid == 0) out[(size_t)y * W + x] = (m << shift) | f;
    }
    __syncthreads();
  };

  uint32_t ring[kDepth][2];
#pragma unroll
  for (int q = 0; q < kDepth; ++q)
    if (y0 + 1 + q < y1) load_step(y0 + 1 + q, ring[q]);
  finish(y0);
  for (int yb = y0 + 1; yb < y1; yb += kDepth) {
#pragma unroll
    for (int q = 0; q < kDepth; ++q) {
      const int y = yb + q;
      if (y < y1) {
        update(ring[q]);
        if (y + kDepth < y1) load_step(y + kDepth, ring[q]);
        finish(y);
      }
    }
  }
}

// The same running median with one wavefront per (column, row segment), for
// histograms of at most 2^12 bins: every update, search and output of a row
// is the wavefront's own, so rows need no barrier (LDS operations of one
// wavefront complete in order), and a CU holds a dozen columns at once.
// Lane l applies updates l, l + 64, ... of the 2 s per row.
constexpr int kWaveSlots = (2 * kMaxMedianSize + 63) / 64;  // 8

template <int NBITS>
__global__ void __launch_bounds__(64)
median_wave_kernel(const uint32_t *__restrict__ in, uint32_t *__restrict__ out, int W, int H,
                   int s, int seg_rows) {
  constexpr uint32_t NB = 1u << NBITS;
  static_assert(NBITS <= 12, "one wavefront per column for small histograms");
  // no group sums: the search scans at most 64 groups of bins, and a median
  // that moves far in one row is rare
  constexpr bool GROUPS = false;
  __shared__ uint32_t hist[NB / 2];
  __shared__ uint32_t grpc[1];
  __shared__ uint16_t rowtab[kMaxSegRows + 256];  // frames of at most 65535 rows
  uint32_t b = blockIdx.x;
  if ((gridDim.x & 7) == 0) b = (b & 7) * (gridDim.x >> 3) + (b >> 3);  // a band per XCD
  const int x = (int)(b % (uint32_t)W);
  const int y0 = (int)(b / (uint32_t)W) * seg_rows;
  if (y0 >= H) return;
  const int y1 = min(H, y0 + seg_rows);
  const int lane = threadIdx.x;
  const int lo = s >> 1;
  const uint32_t k = (uint32_t)(s * s) >> 1;
  const auto hcount = [&](uint32_t j) { return (hist[j >> 1] >> ((j & 1u) << 4)) & 0xffffu; };
  const auto gcount = [&](uint32_t j) { return grpc[j]; };
  const auto order = [] {  // LDS operations of the wavefront stay in program order
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
  };
  for (int i = lane; i < y1 - y0 + s; i += 64) rowtab[i] = (uint16_t)reflect(y0 - lo + i, H);
  for (uint32_t i = lane; i < NB / 2; i += 64) hist[i] = 0;
  order();
  for (int dx0 = 0; dx0 < s; dx0 += 64) {
    if (dx0 + lane >= s) continue;
    const size_t col = (size_t)reflect(x - lo + dx0 + lane, W);
    for (int dy = 0; dy < s; ++dy) {
      const uint32_t v = in[(size_t)rowtab[dy] * W + col];
      atomicAdd(&hist[v >> 1], 1u << ((v & 1u) << 4));
    }
  }
  order();
  uint32_t m = 0, below = 0;
  seek<GROUPS>(hcount, gcount, NB, k, m, below);
  if (lane == 0) out[(size_t)y0 * W + x] = m;

  size_t col[kWaveSlots];
  bool act[kWaveSlots], add[kWaveSlots];
#pragma unroll
  for (int u = 0; u < kWaveSlots; ++u) {
    const int j = lane + 64 * u;
    act[u] = j < 2 * s;
    add[u] = j >= s;
    col[u] = act[u] ? (size_t)reflect(x - lo + (add[u] ? j - s : j), W) : 0;
  }
  const int nslot = (2 * s + 63) / 64;  // slots any lane uses
  const auto load_step = [&](int y, uint32_t (&v)[kWaveSlots]) {
    const int rl = rowtab[y - 1 - y0], re = rowtab[y - 1 - y0 + s];
#pragma unroll
    for (int u = 0; u < kWaveSlots; ++u)
      if (u < nslot && act[u]) v[u] = in[(size_t)(add[u] ? re : rl) * W + col[u]];
  };
  uint32_t ring[kDepth][kWaveSlots];
#pragma unroll
  for (int q = 0; q < kDepth; ++q)
    if (y0 + 1 + q < y1) load_step(y0 + 1 + q, ring[q]);
  for (int yb = y0 + 1; yb < y1; yb += kDepth) {
#pragma unroll
    for (int q = 0; q < kDepth; ++q) {
      const int y = yb + q;
      if (y >= y1) break;
      int c = 0;
#pragma unroll
      for (int u = 0; u < kWaveSlots; ++u) {
        if (u >= nslot) break;
        const uint32_t bin = ring[q][u];
        const bool a = act[u];
        if (a) {
          const uint32_t inc = 1u << ((bin & 1u) << 4);
          atomicAdd(&hist[bin >> 1], add[u] ? inc : 0u - inc);
        }
        c += __popcll(__ballot(a && add[u] && bin < m));
        c -= __popcll(__ballot(a && !add[u] && bin < m));
      }
      if (y + kDepth < y1) load_step(y + kDepth, ring[q]);
      below = (uint32_t)((int)below + c);
      order();
      seek<GROUPS>(hcount, gcount, NB, k, m, below);
      if (lane == 0) out[(size_t)y * W + x] = m;
      order();
    }
  }
}

template <int NBITS, bool FINE>
void launch_median(hipStream_t st, const uint32_t *in, uint32_t *out, int W, int H, int s,
                   int seg_rows, int nseg, int shift) {
  hipLaunchKernelGGL((median_column_kernel<NBITS, FINE>), dim3((unsigned)(W * nseg)),
                     dim3(kMedThreads), 0, st, in, out, W, H, s, seg_rows, shift);
}

}  // namespace

hipError_t key_bit_range(hipStream_t st, const uint64_t *keys, uint64_t n,
                         unsigned long long *d_acc, int *lo, int *hi) {
  *lo = 0;
  *hi = 64;
  if (n < 2) {
    *hi = 0;
    return hipSuccess;
  }
  hipError_t e = hipMemsetAsync(d_acc, 0, 8, st);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(key_diff_kernel, dim3(std::min(grid_for(n), 1024u)), dim3(kThreads), 0, st,
                     keys, n, d_acc);
  unsigned long long d = 0;
  e = hipMemcpyAsync(&d, d_acc, 8, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return e;
  if (!d) {
    *hi = 0;  // all equal
    return hipSuccess;
  }
  *lo = __builtin_ctzll(d);
  *hi = 64 - __builtin_clzll(d);
  return hipSuccess;
}

int median_background_dev(lm_ctx *ctx, hipStream_t st, int W, int H, const double *d_img,
                          int size, int passes, double *d_bg) {
  const uint64_t npx = (uint64_t)W * H;
  if (W < 1 || H < 1 || npx > (1ull << 24) || size > kMaxMedianSize) {
    set_error("median background: image over 2^24 pixels or filter size over 255");
    return LM_ERR_INVALID;
  }
  if (size <= 0 || passes <= 0) {  // compute_background returns zeros
    const hipError_t e = hipMemsetAsync(d_bg, 0, npx * 8, st);
    return e == hipSuccess ? LM_OK : hip_fail(e, "median background");
  }
  // one allocation, carved: keys, sorted keys, table (8 B per pixel), index,
  // sorted index, scan, two rank images (4 B per pixel), sort/scan storage
  size_t sort_bytes = 0, scan_bytes = 0;
  hipError_t e = hipcub::DeviceRadixSort::SortPairs(
      nullptr, sort_bytes, (uint64_t *)nullptr, (uint64_t *)nullptr, (uint32_t *)nullptr,
      (uint32_t *)nullptr, (int)npx, 0, 64, st);
  if (e == hipSuccess)
    e = hipcub::DeviceScan::InclusiveSum(nullptr, scan_bytes, (uint32_t *)nullptr,
                                         (uint32_t *)nullptr, (int)npx, st);
  if (e != hipSuccess) return hip_fail(e, "median background buffers");
  const auto up = [](uint64_t n) { return (n + 255) & ~(uint64_t)255; };
  const uint64_t b8 = up(npx * 8), b4 = up(npx * 4);
  uint8_t *base = nullptr;  // kept by the context for the next call
  e = pool_get(ctx, ctx->median_pool, 0,
               3 * b8 + 5 * b4 + up(std::max(sort_bytes, scan_bytes)) + 256, (void **)&base);
  if (e != hipSuccess) return hip_fail(e, "median background buffers");
  uint64_t *keys = (uint64_t *)base, *skeys = (uint64_t *)(base + b8);
  double *table = (double *)(base + 2 * b8);
  uint32_t *idx = (uint32_t *)(base + 3 * b8), *sidx = (uint32_t *)(base + 3 * b8 + b4);
  uint32_t *scan = (uint32_t *)(base + 3 * b8 + 2 * b4);
  uint32_t *ra = (uint32_t *)(base + 3 * b8 + 3 * b4), *rb = (uint32_t *)(base + 3 * b8 + 4 * b4);
  void *tmp = base + 3 * b8 + 5 * b4;
  unsigned long long *acc =
      (unsigned long long *)(base + 3 * b8 + 5 * b4 + up(std::max(sort_bytes, scan_bytes)));
  const unsigned g = grid_for(npx);
  hipLaunchKernelGGL(key_kernel, dim3(g), dim3(kThreads), 0, st, d_img, npx, keys, idx);
  // sort only the key bits that differ (integer counts leave the low
  // mantissa bits zero: about half the radix passes)
  int lo = 0, hi = 64;
  e = key_bit_range(st, keys, npx, acc, &lo, &hi);
  if (e == hipSuccess && hi == 0) lo = 0, hi = 1;  // all keys equal: one pass keeps them
  if (e == hipSuccess)
    e = hipcub::DeviceRadixSort::SortPairs(tmp, sort_bytes, keys, skeys, idx, sidx, (int)npx, lo,
                                           hi, st);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(head_kernel, dim3(g), dim3(kThreads), 0, st, skeys, npx, ra);
    e = hipcub::DeviceScan::InclusiveSum(tmp, scan_bytes, ra, scan, (int)npx, st);
  }
  uint32_t D = 0;
  if (e == hipSuccess) {
    hipLaunchKernelGGL(rank_kernel, dim3(g), dim3(kThreads), 0, st, skeys, sidx, scan, npx, ra,
                       table);
    e = hipMemcpyAsync(&D, scan + npx - 1, 4, hipMemcpyDeviceToHost, st);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return hip_fail(e, "median background ranks");
  int R = 1;  // bits of the largest rank D - 1 (at least 1)
  while (R < 32 && ((uint64_t)(D - 1) >> R)) ++R;
  const int shift = R > 16 ? R - 15 : 0;  // over 16 bits: 15 coarse bits and a band
  // row segments: enough workgroups to fill the chip, each several windows tall
  const int want = (4096 + W - 1) / W;
  const int max_seg = std::max(1, H / std::max(2 * size, 64));
  const int nseg = std::max({1, std::min(want, max_seg), (H + kMaxSegRows - 1) / kMaxSegRows});
  const int seg_rows = (H + nseg - 1) / nseg;
  // frames with at most 2^12 distinct values: one wavefront per column, with
  // more segments (a CU holds a dozen such columns)
  const bool wave = shift == 0 && R <= 12 && H <= 65535;
  const int wseg = std::max(
      {1, std::min((8192 + W - 1) / W, max_seg), (H + kMaxSegRows - 1) / kMaxSegRows});
  const int wseg_rows = (H + wseg - 1) / wseg;
  uint32_t *src = ra, *dst = rb;
  for (int p = 0; p < passes; ++p) {
    if (wave) {
      const dim3 grid((unsigned)(W * wseg));
      if (R <= 8)
        hipLaunchKernelGGL((median_wave_kernel<8>), grid, dim3(64), 0, st, src, dst, W, H, size,
                           wseg_rows);
      else if (R <= 10)
        hipLaunchKernelGGL((median_wave_kernel<10>), grid, dim3(64), 0, st, src, dst, W, H, size,
                           wseg_rows);
      else
        hipLaunchKernelGGL((median_wave_kernel<12>), grid, dim3(64), 0, st, src, dst, W, H, size,
                           wseg_rows);
    } else if (shift > 0)
      launch_median<15, true>(st, src, dst, W, H, size, seg_rows, nseg, shift);
    else if (R <= 8)
      launch_median<8, false>(st, src, dst, W, H, size, seg_rows, nseg, 0);
    else if (R <= 10)
      launch_median<10, false>(st, src, dst, W, H, size, seg_rows, nseg, 0);
    else if (R <= 12)
      launch_median<12, false>(st, src, dst, W, H, size, seg_rows, nseg, 0);
    else if (R <= 14)
      launch_median<14, false>(st, src, dst, W, H, size, seg_rows, nseg, 0);
    else
      launch_median<16, false>(st, src, dst, W, H, size, seg_rows, nseg, 0);
    std::swap(src, dst);
  }
  hipLaunchKernelGGL(unrank_kernel, dim3(g), dim3(kThreads), 0, st, src, table, D, npx, d_bg);
  e = hipStreamSynchronize(st);  // the buffers are reused by the next call
  return e == hipSuccess ? LM_OK : hip_fail(e, "median background");
}

}  // namespace lm
