// lm_band.hip - pixel-banded copy of the spot table and the single-image
// compare kernel that reads it.
//
// Why: the compare step is one byte gather per spot.  On MI355X, 64 lanes
// gathering from random addresses of a 4 MiB image sustain about 260 G
// gathers/s, while 64 lanes confined to a 2 KiB window sustain about 1000 G/s
// (the L1/TA unit serves each distinct cache line of a wave instruction
// separately).  Reordering each super-tile's spots by pixel bucket makes
// consecutive lanes read neighbouring pixels, so a wave's gathers fall in a few
// cache lines; the per-orientation sums are then accumulated in LDS counters
// (one ds_add per nonzero spot) instead of per-lane registers.
//
// Exactness: the accumulator (nspots << 17 | sum of raw) is an integer, so
// the accumulation order does not matter; with imScale == 1 the reference
// totInt is that sum (src/LaueMatchingCPU.c:674-681).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>

#include "lm_internal.h"
#include "lm_walk16.h"

namespace lm {
namespace {

constexpr int kBuildThreads = 1024;
constexpr int kMaxBuckets = 16384;

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v, int lane) {
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t u = __shfl_up(v, d);
    if (lane >= d) v += u;
  }
  return v;
}

// One workgroup per super-tile: counting sort of the super-tile's entries by
// pixel bucket (two traversals of the orientation-major table: histogram,
// then scatter).  With chunk offsets, each 2^cb-pixel chunk's entry range is
// split into 2^ks pieces of balanced size (the LDS walk's work items).
__global__ void __launch_bounds__(kBuildThreads)
band_build_kernel(const uint16_t *__restrict__ counts, const uint64_t *__restrict__ tileOff,
                  const uint32_t *__restrict__ entries, uint64_t n, int tb, int pb, int bb,
                  int nbuckets, int nbands, uint32_t *__restrict__ bentries,
                  uint32_t *__restrict__ bandOff, int cb, int nchunks, int ks,
                  uint32_t *__restrict__ chunkOff) {
  __shared__ uint32_t hist[kMaxBuckets];
  __shared__ uint32_t wsum[kBuildThreads / 64];
  const int tid = threadIdx.x, lane = __lane_id(), wave = tid >> 6;
  const uint64_t t = blockIdx.x;
  const uint64_t first = t << tb;
  const uint64_t rows = std::min<uint64_t>(1ull << tb, n - first);
  const uint64_t base = tileOff[first >> kTileBits];
  const uint32_t pmask = (pb >= 32) ? 0xffffffffu : ((1u << pb) - 1u);
  for (int i = tid; i < nbuckets; i += kBuildThreads) hist[i] = 0;
  __syncthreads();
  for (int pass = 0; pass < 2; ++pass) {
    uint32_t run = 0;
    for (uint64_t k = 0; k < rows; k += kBuildThreads) {
      const uint64_t r = k + tid;
      const uint32_t c = r < rows ? counts[first + r] : 0u;
      const uint32_t incl = wave_incl_scan(c, lane);
      if (lane == 63) wsum[wave] = incl;
      __syncthreads();
      uint32_t before = 0, tot = 0;
      for (int w = 0; w < kBuildThreads / 64; ++w) {
        const uint32_t v = wsum[w];
        if (w < wave) before += v;
        tot += v;
      }
      const uint32_t excl = run + before + incl - c;
      for (uint32_t s = 0; s < c; ++s) {
        const uint32_t pix = entries[base + excl + s];
        const uint32_t b = pix >> bb;
        if (pass == 0) {
          atomicAdd(&hist[b], 1u);
        } else {
          const uint32_t pos = atomicAdd(&hist[b], 1u);
          bentries[base + pos] = ((uint32_t)r << pb) | (pix & pmask);
        }
      }
      run += tot;
      __syncthreads();
    }
    if (pass == 0) {
      // exclusive scan of hist (contiguous chunk per thread)
      const int per = (nbuckets + kBuildThreads - 1) / kBuildThreads;
      const int lo = std::min(nbuckets, tid * per), hi = std::min(nbuckets, lo + per);
      uint32_t s = 0;
      for (int i = lo; i < hi; ++i) s += hist[i];
      const uint32_t incl = wave_incl_scan(s, lane);
      if (lane == 63) wsum[wave] = incl;
      __syncthreads();
      uint32_t acc = incl - s;
      for (int w = 0; w < wave; ++w) acc += wsum[w];
      __syncthreads();
      for (int i = lo; i < hi; ++i) {
        const uint32_t v = hist[i];
        hist[i] = acc;
        acc += v;
      }
      __syncthreads();
      const int shift = pb - bb;
      uint32_t *bo = bandOff + t * (uint64_t)(nbands + 1);
      for (int j = tid; j <= nbands; j += kBuildThreads) {
        const int64_t bk = (int64_t)j << shift;
        bo[j] = (j == nbands || bk >= nbuckets) ? run : hist[bk];
      }
      if (chunkOff) {
        const uint64_t npieces = (uint64_t)nchunks << ks;
        uint32_t *co = chunkOff + t * (npieces + 1);
        auto start = [&](int c) -> uint32_t {
          const int64_t bk = (int64_t)c << (cb - bb);
          return (c == nchunks || bk >= nbuckets) ? run : hist[bk];
        };
        for (int j = tid; j <= nchunks; j += kBuildThreads) {
          if (j == nchunks) {
            co[npieces] = run;
            continue;
          }
          const uint32_t s0 = start(j), len = start(j + 1) - s0;
          for (int p = 0; p < (1 << ks); ++p)
            co[((uint64_t)j << ks) + p] = s0 + (uint32_t)(((uint64_t)len * (uint32_t)p) >> ks);
        }
      }
      __syncthreads();
    }
  }
}

// Single-image compare over the banded copy; one workgroup per super-tile
// (grid-stride).  cnt[r] = (nspots << 17) | sum of raw for row r.
//
// The super-tile's entries are walked as one flat range; lane l of group g
// takes entries g*THREADS*U + k*THREADS + l (k < U), so the 64 lanes of a wave
// read consecutive, pixel-sorted entries.  Each lane tracks the band of its
// entries incrementally (entries are sorted by pixel, so the band never
// decreases).  The next group's entry loads are issued after the current
// group's gathers, so they stay in flight across the gather wait and the LDS
// adds.
template <int TB, int THREADS, int U, bool WIDE = false>
__global__ void __launch_bounds__(THREADS)
score_band_kernel(const uint32_t *__restrict__ bentries, const uint64_t *__restrict__ tileOff,
                  const uint32_t *__restrict__ bandOff, int nbands, uint64_t n, uint64_t nsuper,
                  const uint8_t *__restrict__ img, const Pred *__restrict__ pred,
                  uint64_t *__restrict__ hits, uint64_t cap, uint32_t *__restrict__ counters) {
  constexpr int T = 1 << TB;
  constexpr int PB = 32 - TB;
  constexpr uint32_t PMASK = (1u << PB) - 1u;
  constexpr uint32_t STEP = THREADS * U;
  __shared__ uint32_t cnt[T];
  const int tid = threadIdx.x, lane = __lane_id();
  for (uint64_t t = blockIdx.x; t < nsuper; t += gridDim.x) {
    for (int i = tid; i < T; i += THREADS) cnt[i] = 0;
    __syncthreads();
    const uint64_t first = t << TB;
    const uint32_t *src = bentries + tileOff[first >> kTileBits];
    const uint32_t *bo = bandOff + t * (uint64_t)(nbands + 1);
    const uint32_t total = bo[nbands];
    int band = 0;
    uint32_t bend = bo[1];
    uint32_t e[U];
#pragma unroll
    for (int k = 0; k < U; ++k) {
      const uint32_t i = k * THREADS + tid;
      e[k] = i < total ? __builtin_nontemporal_load(src + i) : 0u;
    }
    for (uint32_t g0 = 0; g0 < total; g0 += STEP) {
      uint32_t raw[U];
      uint32_t row[U];
#pragma unroll
      for (int k = 0; k < U; ++k) {
        const uint32_t i = g0 + k * THREADS + tid;
        raw[k] = 0;
        row[k] = e[k] >> PB;
        if (i < total) {
          while (i >= bend) {
            ++band;
            bend = bo[band + 1];
          }
          const uint64_t px = ((uint64_t)band << PB) | (e[k] & PMASK);
          if constexpr (WIDE) {
            // the aligned word holding the pixel: a 4-byte gather per lane
            const uint32_t wd = *(const uint32_t *)(img + (px & ~3ull));
            raw[k] = (wd >> (8 * (uint32_t)(px & 3))) & 0xffu;
          } else {
            raw[k] = img[px];
          }
        }
      }
#pragma unroll
      for (int k = 0; k < U; ++k) {
        const uint32_t i = g0 + STEP + k * THREADS + tid;
        e[k] = i < total ? __builtin_nontemporal_load(src + i) : 0u;
      }
#pragma unroll
      for (int k = 0; k < U; ++k)
        if (raw[k]) atomicAdd(&cnt[row[k]], (1u << kSumBits) + raw[k]);
    }
    __syncthreads();
    for (int rr = tid; rr < T; rr += THREADS) {
      const uint64_t o = first + rr;
      const uint32_t acc = cnt[rr];
      const uint32_t ns = acc >> kSumBits, sum = acc & kSumMask;
      const bool live = o < n;
      const int min_spots = pred->min_spots[0];
      const double min_int = pred->min_int[0];
      // imScale != 1: a superset, decided by the float64 in-order re-sum
      const double cand_thr = pred->cand_thr[0];
      const bool zero_pass = live && ns == 0 && (0 >= min_spots) && (0.0 >= min_int);
      const bool cand = live && ns > 0 && (int)ns >= min_spots && (double)sum >= cand_thr;
      const uint64_t zm = __ballot(zero_pass);
      if (zm && lane == 0) atomicAdd(&counters[1], (uint32_t)__popcll(zm));
      const uint64_t m = __ballot(cand);
      if (m) {
        const int leader = __ffsll((unsigned long long)m) - 1;
        uint32_t b0 = 0;
        if (lane == leader) b0 = atomicAdd(&counters[0], (uint32_t)__popcll(m));
        b0 = __shfl(b0, leader);
        if (cand) {
          const uint64_t pos = (uint64_t)b0 + (uint64_t)__popcll(m & ((1ull << lane) - 1ull));
          if (pos < cap) hits[pos] = (o << 26) | (uint64_t)acc;
        }
      }
    }
    __syncthreads();
  }
}

// Same walk with 16-byte entry loads: lane l of group g takes the 4
// consecutive entries at 4*(g*THREADS*U + k*THREADS + l) (k < U), counted
// from the super-tile's first entry rounded down to a 16-byte boundary; the
// `mis` entries before the super-tile's start (and those past its end) are
// loaded but not used.  A quarter of the load instructions of the scalar walk
// for the same bytes in flight.
template <int TB, int THREADS, int U>
__global__ void __launch_bounds__(THREADS)
score_band_v4_kernel(const uint32_t *__restrict__ bentries, const uint64_t *__restrict__ tileOff,
                     const uint32_t *__restrict__ bandOff, int nbands, uint64_t n,
                     uint64_t nsuper, const uint8_t *__restrict__ img,
                     const Pred *__restrict__ pred, uint64_t *__restrict__ hits, uint64_t cap,
                     uint32_t *__restrict__ counters) {
  typedef uint32_t v4u __attribute__((ext_vector_type(4)));
  constexpr int T = 1 << TB;
  constexpr int PB = 32 - TB;
  constexpr uint32_t PMASK = (1u << PB) - 1u;
  constexpr uint32_t STEP = THREADS * U * 4;
  __shared__ uint32_t cnt[T];
  const int tid = threadIdx.x, lane = __lane_id();
  for (uint64_t t = blockIdx.x; t < nsuper; t += gridDim.x) {
    for (int i = tid; i < T; i += THREADS) cnt[i] = 0;
    __syncthreads();
    const uint64_t first = t << TB;
    const uint64_t base = tileOff[first >> kTileBits];
    const uint64_t a0 = base & ~3ull;
    const uint32_t mis = (uint32_t)(base - a0);
    const v4u *src = (const v4u *)(bentries + a0);
    const uint32_t *bo = bandOff + t * (uint64_t)(nbands + 1);
    const uint32_t end = bo[nbands] + mis;  // used entries: mis <= j < end
    int band = 0;
    uint32_t bend = bo[1] + mis;
    v4u e[U];
#pragma unroll
    for (int k = 0; k < U; ++k) {
      const uint32_t q = k * THREADS + tid;
      e[k] = 4 * q < end ? __builtin_nontemporal_load(src + q) : v4u{0u, 0u, 0u, 0u};
    }
    for (uint32_t g0 = 0; g0 < end; g0 += STEP) {
      uint32_t raw[U][4];
      uint32_t row[U][4];
#pragma unroll
      for (int k = 0; k < U; ++k) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const uint32_t j = g0 + 4 * (k * THREADS + tid) + v;
          const uint32_t ev = e[k][v];
          row[k][v] = ev >> PB;
          raw[k][v] = 0;
          if (j >= mis && j < end) {
            while (j >= bend) {
              ++band;
              bend = bo[band + 1] + mis;
            }
            raw[k][v] = img[((uint64_t)band << PB) | (ev & PMASK)];
          }
        }
      }
#pragma unroll
      for (int k = 0; k < U; ++k) {
        const uint32_t q = (g0 + STEP) / 4 + k * THREADS + tid;
        e[k] = 4 * q < end ? __builtin_nontemporal_load(src + q) : v4u{0u, 0u, 0u, 0u};
      }
#pragma unroll
      for (int k = 0; k < U; ++k)
#pragma unroll
        for (int v = 0; v < 4; ++v)
          if (raw[k][v]) atomicAdd(&cnt[row[k][v]], (1u << kSumBits) + raw[k][v]);
    }
    __syncthreads();
    for (int rr = tid; rr < T; rr += THREADS) {
      const uint64_t o = first + rr;
      const uint32_t acc = cnt[rr];
      const uint32_t ns = acc >> kSumBits, sum = acc & kSumMask;
      const bool live = o < n;
      const int min_spots = pred->min_spots[0];
      const double min_int = pred->min_int[0];
      // imScale != 1: a superset, decided by the float64 in-order re-sum
      const double cand_thr = pred->cand_thr[0];
      const bool zero_pass = live && ns == 0 && (0 >= min_spots) && (0.0 >= min_int);
      const bool cand = live && ns > 0 && (int)ns >= min_spots && (double)sum >= cand_thr;
      const uint64_t zm = __ballot(zero_pass);
      if (zm && lane == 0) atomicAdd(&counters[1], (uint32_t)__popcll(zm));
      const uint64_t m = __ballot(cand);
      if (m) {
        const int leader = __ffsll((unsigned long long)m) - 1;
        uint32_t b0 = 0;
        if (lane == leader) b0 = atomicAdd(&counters[0], (uint32_t)__popcll(m));
        b0 = __shfl(b0, leader);
        if (cand) {
          const uint64_t pos = (uint64_t)b0 + (uint64_t)__popcll(m & ((1ull << lane) - 1ull));
          if (pos < cap) hits[pos] = (o << 26) | (uint64_t)acc;
        }
      }
    }
    __syncthreads();
  }
}

// Single-image compare with the image staged through LDS instead of
// gathered from L2.  The byte gathers of score_band_kernel keep the texture
// address and data units busy (one wave instruction touches about a dozen
// cache lines); here each wave owns a 2^CB-byte slot of LDS and walks the
// super-tile chunk by chunk (chunk c = pixels [c << CB, (c+1) << CB), the
// entries of chunk c are chunkOff[c] .. chunkOff[c+1] of the pixel-sorted
// super-tile): coalesced 16-byte loads (2^CB / 1024 per lane) bring the
// chunk's pixels, EMAX coalesced loads per lane its entries, and each entry
// reads its pixel from LDS and adds into the shared row counters.  Wave w
// takes chunks w, w + NW, ...; D sets of loads rotate, so while one chunk is
// processed the next D - 1 are in flight.  All loads are buffer loads with a
// wave-uniform base and range: offsets past the range read zeros, so chunk
// indices past the last chunk are empty chunks, every wave runs the same
// straight-line loop and each wait covers exactly one set.  A set also loads
// the entry range of the chunk it will hold next, so refilling a set never
// waits on a scalar load.  DENSE adds a slow path for chunks with more than
// EMAX * 64 entries (the builder records the largest chunk).  LDS traffic of
// a wave is ordered, so its slot needs no barrier between chunks.
template <int TB, int CB, int EMAX, int D, int THREADS, bool DENSE, bool E16>
__global__ void __launch_bounds__(THREADS)
score_band_lds_kernel(const uint32_t *__restrict__ bentries, const uint64_t *__restrict__ tileOff,
                      const uint32_t *__restrict__ chunkOff, int nchunks, int ks, uint64_t n,
                      uint64_t nsuper, const uint8_t *__restrict__ img, uint64_t npx,
                      uint32_t spots256, const Pred *__restrict__ pred,
                      uint64_t *__restrict__ hits, uint64_t cap, uint32_t *__restrict__ counters,
                      uint64_t *__restrict__ cands, uint64_t candCap,
                      const uint32_t *__restrict__ chunkNz) {
  typedef uint32_t v4u __attribute__((ext_vector_type(4)));
  constexpr int T = 1 << TB;
  constexpr int NW = THREADS / 64;
  constexpr int PB = 32 - TB;
  constexpr int C = 1 << CB;
  constexpr int V = C / (64 * 16);  // 16-byte image loads per lane per chunk
  static_assert(V >= 1 && V * 64 * 16 == C, "chunk must be a multiple of 1 KiB");
  static_assert(!E16 || EMAX % 4 == 0, "16-byte entry loads hold 4 entries per lane");
  // 2^16-row super-tiles: 16-bit sums, two rows per counter word
  constexpr bool S16 = TB >= kSum16TileBits;
  constexpr int NCNT = S16 ? T / 2 : T;
  static_assert(NCNT * 4 + NW * C <= 163840, "LDS");
  constexpr int kRawBufferWord3 = 0x00020000;  // gfx9 raw buffer, 32-bit data format
  constexpr int kNonTemporal = 2;              // buffer load aux bit 1: nt
  __shared__ __attribute__((aligned(16))) uint32_t cnt[NCNT];
  __shared__ v4u slot[NW][C / 16];
  const int tid = threadIdx.x, lane = __lane_id();
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const uint8_t *mine = (const uint8_t *)&slot[w][0];
  const __amdgpu_buffer_rsrc_t irs =
      __builtin_amdgcn_make_buffer_rsrc((void *)img, (short)0, (int)(uint32_t)npx, kRawBufferWord3);
  // image chunks with no nonzero pixel (chunkNz[c] == 0, from
  // chunk_nz_kernel on this image): their spots add nothing, so their entry
  // ranges are taken as empty and never loaded; flags past the last chunk
  // read 0.  The work items are pieces: chunk c's entries are pieces
  // c << ks .. (c << ks) + 2^ks - 1 (nchunks counts pieces here)
  const __amdgpu_buffer_rsrc_t frs = __builtin_amdgcn_make_buffer_rsrc(
      (void *)chunkNz, (short)0, (int)(((uint32_t)nchunks >> ks) * 4u), kRawBufferWord3);
  // 16-bit counters: the sum of each row's pixels, or its count of nonzero
  // pixels when that count is the selective test: MinNrSpots >= 2 and a row
  // expects at most MinNrSpots / 4 nonzero pixels (nonzero fraction x mean
  // spots per row, spots256 = 256 x that mean), as on sparse images
  const uint32_t pk0 = pred->pk_thr[0];
  const uint32_t kmin = pk0 >> 23;
  const bool count_mode = S16 && kmin >= 2u &&
                          (uint64_t)counters[kNzSlot] * 4u * spots256 <= npx * kmin * 256u;
  // counters start at zero; each epilogue clears the words it reads, so the
  // next super-tile starts without a clearing pass
  for (int i = tid; i < NCNT / 4; i += THREADS) ((v4u *)cnt)[i] = v4u{0u, 0u, 0u, 0u};
  __syncthreads();
  for (uint64_t t = blockIdx.x; t < nsuper; t += gridDim.x) {
    const uint64_t first = t << TB;
    const uint32_t *co = chunkOff + t * (uint64_t)(nchunks + 1);
    const uint32_t total = co[nchunks];
    // E16: the entry buffer starts at the 16-byte boundary at or below the
    // super-tile's first entry (mis entries earlier), so 16-byte entry loads
    // are aligned
    const uint64_t e0 = tileOff[first >> kTileBits];
    const uint32_t mis = E16 ? (uint32_t)(e0 & 3u) : 0u;
    const __amdgpu_buffer_rsrc_t ers = __builtin_amdgcn_make_buffer_rsrc(
        (void *)(bentries + (e0 - mis)), (short)0, (int)((total + mis) * 4u), kRawBufferWord3);
    // chunk offsets as a buffer too: index nchunks holds `total`; indices are
    // clamped to it, so chunks past the end are empty
    const __amdgpu_buffer_rsrc_t crs = __builtin_amdgcn_make_buffer_rsrc(
        (void *)co, (short)0, (int)((uint32_t)(nchunks + 1) * 4u), kRawBufferWord3);
    auto coff = [&](int c) {
      const uint32_t cc = (uint32_t)(c < nchunks ? c : nchunks);
      return __builtin_amdgcn_raw_buffer_load_b32(crs, cc * 4u, 0, 0);
    };
    v4u pim[D][V];
    uint32_t pe[D][EMAX];
    uint32_t beg[D], end[D];
    // The wave's chunks w, w + NW, ... (j-th: chunk w + NW j) go in rounds of
    // RC = the largest multiple of D up to 64: lane l holds the entry range of
    // the round's chunk l (one load per round), so a set's entry range comes
    // from a lane read, not from a load the set's entry loads would wait on.
    // Round r + 1's ranges are loaded when round r starts; inside round r the
    // fetches read round r's ranges only, and the last D consumes of the round
    // (unrolled) fetch round r + 1's first chunks from the ranges loaded a
    // round earlier, so the pipeline runs through round boundaries without a
    // drain and every wait stays an exact count.
    constexpr int RC = (64 / D) * D;
    const int per_wave = (nchunks + NW - 1) / NW;
    const int nround = (per_wave + RC - 1) / RC;
    auto load_round = [&](int r, uint32_t &vb, uint32_t &ve) {
      const int cl = w + NW * (r * RC + (lane < RC ? lane : RC - 1));
      vb = coff(cl);
      const uint32_t e1 = coff(cl + 1);
      const uint32_t nz = __builtin_amdgcn_raw_buffer_load_b32(
          frs, ((uint32_t)(cl < nchunks ? cl : nchunks) >> ks) * 4u, 0, 0);
      ve = nz ? e1 : vb;
    };
    uint32_t cur_b, cur_e, nxt_b, nxt_e;
    load_round(0, cur_b, cur_e);
    {
      // chunk j of the wave, entry range from lane jj of (vb, ve)
      auto fetch = [&](int d, int j, int jj, uint32_t vb, uint32_t ve) {
        const bool in = j < per_wave;
        beg[d] = in ? (uint32_t)__builtin_amdgcn_readlane((int)vb, jj) : total;
        end[d] = in ? (uint32_t)__builtin_amdgcn_readlane((int)ve, jj) : total;
        // a chunk with no entries to walk (none, or an all-zero image chunk)
        // needs no pixels: its loads read chunk 0 again, which stays cached
        const uint32_t ib = beg[d] != end[d] ? ((uint32_t)(w + NW * j) >> ks) << CB : 0u;
#pragma unroll
        for (int v = 0; v < V; ++v)
          pim[d][v] = __builtin_amdgcn_raw_buffer_load_b128(irs, ib + 16u * (v * 64 + lane), 0, 0);
        if constexpr (E16) {
          // lane l holds entries a + 4l .. a + 4l + 3 (+ 256 per load), a the
          // chunk's first entry rounded down to a 16-byte boundary; the
          // descriptor ends at the chunk's last entry, so loads past it read
          // zeros instead of the next chunk's entries.  Entries are read once:
          // non-temporal loads, so the L2 keeps the image chunks
          const uint32_t a = (beg[d] + mis) & ~3u;
          const __amdgpu_buffer_rsrc_t crs16 = __builtin_amdgcn_make_buffer_rsrc(
              (void *)(bentries + (e0 - mis) + a), (short)0, (int)((end[d] + mis - a) * 4u),
              kRawBufferWord3);
#pragma unroll
          for (int q = 0; q < EMAX / 4; ++q) {
            const v4u v = __builtin_amdgcn_raw_buffer_load_b128(crs16, (q * 256 + 4 * lane) * 4u,
                                                                0, kNonTemporal);
            pe[d][4 * q + 0] = v.x;
            pe[d][4 * q + 1] = v.y;
            pe[d][4 * q + 2] = v.z;
            pe[d][4 * q + 3] = v.w;
          }
        } else {
#pragma unroll
          for (int k = 0; k < EMAX; ++k)
            pe[d][k] =
                __builtin_amdgcn_raw_buffer_load_b32(ers, (beg[d] + k * 64 + lane) * 4u, 0,
                                                     kNonTemporal);
        }
        // keep the sets' loads in issue order, so the wait before a set
        // counts only the later sets
        __builtin_amdgcn_sched_barrier(0);
      };
      auto consume = [&](int d, int j) {
#pragma unroll
        for (int v = 0; v < V; ++v) slot[w][v * 64 + lane] = pim[d][v];
        __builtin_amdgcn_wave_barrier();
        const uint32_t cbeg = beg[d], cend = end[d];
        // buffer index of this lane's k-th entry, and of the first entry the
        // loads above did not cover
        const uint32_t a = E16 ? ((cbeg + mis) & ~3u) : cbeg + mis;
        auto idx = [&](int k) {
          return E16 ? a + (k >> 2) * 256 + 4 * lane + (k & 3) : a + k * 64 + lane;
        };
        // all pixel reads first (lanes outside the chunk read a masked,
        // in-slot byte and drop it), then the counter adds.  In range:
        // cbeg + mis <= idx(k) < cend + mis, i.e. (idx(k) - cbeg - mis),
        // wrapping below the chunk, < cend - cbeg; idx(k) - idx(0) is a
        // compile-time constant, so each test is one add and one compare.
        const uint32_t rel0 = idx(0) - (cbeg + mis), len = cend - cbeg;
        uint32_t raw[EMAX];
#pragma unroll
        for (int k = 0; k < EMAX; ++k) {
          const uint32_t r = mine[pe[d][k] & (C - 1)];
          raw[k] = rel0 + (idx(k) - idx(0)) < len ? r : 0u;
        }
        auto add = [&](uint32_t e, uint32_t r) {
          const uint32_t row = e >> PB;
          if constexpr (S16)
            atomicAdd(&cnt[row >> 1], (count_mode ? 1u : r) << ((row & 1u) << 4));
          else
            atomicAdd(&cnt[row], (1u << kSumBits) + r);
        };
#pragma unroll
        for (int k = 0; k < EMAX; ++k)
          if (raw[k]) add(pe[d][k], raw[k]);
        if constexpr (DENSE) {
          for (uint32_t i = a + EMAX * 64 + lane; i < cend + mis; i += 64) {
            const uint32_t e = __builtin_amdgcn_raw_buffer_load_b32(ers, i * 4u, 0, kNonTemporal);
            const uint32_t r = mine[e & (C - 1)];
            if (r) add(e, r);
          }
        }
        __builtin_amdgcn_wave_barrier();
      };
#pragma unroll
      for (int d = 0; d < D; ++d) fetch(d, d, d, cur_b, cur_e);
      for (int r = 0; r < nround; ++r) {
        load_round(r + 1, nxt_b, nxt_e);
        const int j0 = r * RC;
        for (int j = j0; j < j0 + RC - D && j < per_wave; j += D) {
#pragma unroll
          for (int d = 0; d < D; ++d) {
            consume(d, j + d);
            fetch(d, j + d + D, j + d + D - j0, cur_b, cur_e);
          }
        }
        if (j0 + RC - D >= per_wave) break;  // every later chunk is empty
#pragma unroll
        for (int d = 0; d < D; ++d) {
          consume(d, j0 + RC - D + d);
          fetch(d, j0 + RC + d, d, nxt_b, nxt_e);
        }
        cur_b = nxt_b;
        cur_e = nxt_e;
      }
    }
    __syncthreads();
    if constexpr (S16) {
      sum16_epilogue<THREADS>(cnt, first, n, pred, pk0, count_mode, counters, cands, candCap);
      continue;
    }
    for (int rr = tid; rr < T; rr += THREADS) {
      const uint64_t o = first + rr;
      const uint32_t acc = cnt[rr];
      cnt[rr] = 0;
      const uint32_t ns = acc >> kSumBits, sum = acc & kSumMask;
      const bool live = o < n;
      const int min_spots = pred->min_spots[0];
      const double min_int = pred->min_int[0];
      // imScale != 1: a superset, decided by the float64 in-order re-sum
      const double cand_thr = pred->cand_thr[0];
      const bool zero_pass = live && ns == 0 && (0 >= min_spots) && (0.0 >= min_int);
      const bool cand = live && ns > 0 && (int)ns >= min_spots && (double)sum >= cand_thr;
      const uint64_t zm = __ballot(zero_pass);
      if (zm && lane == 0) atomicAdd(&counters[1], (uint32_t)__popcll(zm));
      const uint64_t m = __ballot(cand);
      if (m) {
        const int leader = __ffsll((unsigned long long)m) - 1;
        uint32_t b0 = 0;
        if (lane == leader) b0 = atomicAdd(&counters[0], (uint32_t)__popcll(m));
        b0 = __shfl(b0, leader);
        if (cand) {
          const uint64_t pos = (uint64_t)b0 + (uint64_t)__popcll(m & ((1ull << lane) - 1ull));
          if (pos < cap) hits[pos] = (o << 26) | (uint64_t)acc;
        }
      }
    }
    __syncthreads();
  }
}

// Per image chunk of 2^cb pixels (the LDS walk's chunks), whether any pixel
// is nonzero (flags[c] = 0 or 1; the walk skips the entries of all-zero
// chunks, whose spots add nothing), and the image's count of nonzero pixels
// (counters[kNzSlot], which picks sums or counts for the 16-bit counters).
// One wave per chunk in turn, 16 bytes per lane and load; one atomic per
// workgroup, launched with few workgroups so the atomics do not queue on one
// address.  Bytes past npx read as zero.
__global__ void __launch_bounds__(256)
chunk_nz_kernel(const uint8_t *__restrict__ img, uint64_t npx, int cb, uint32_t nchunks,
                uint32_t *__restrict__ flags, uint32_t *__restrict__ counters) {
  typedef uint32_t v4u __attribute__((ext_vector_type(4)));
  const __amdgpu_buffer_rsrc_t rs =
      __builtin_amdgcn_make_buffer_rsrc((void *)img, (short)0, (int)(uint32_t)npx, 0x00020000);
  const int lane = __lane_id();
  const uint32_t wid = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t nw = (gridDim.x * blockDim.x) >> 6;
  const uint32_t per = (1u << cb) / (64u * 16u);  // 16-byte loads per lane and chunk
  uint32_t c = 0;
  for (uint32_t ch = wid; ch < nchunks; ch += nw) {
    uint32_t any = 0;
    for (uint32_t k = 0; k < per; ++k) {
      const uint32_t off = (ch << cb) + 16u * (k * 64u + (uint32_t)lane);
      const v4u v = __builtin_amdgcn_raw_buffer_load_b128(rs, off, 0, 0);
      for (int q = 0; q < 4; ++q) {
        const uint32_t x = v[q];
        // nonzero bytes of x (bit 7 of each byte of ((x & 0x7f..) + 0x7f..) | x)
        c += __popc((((x & 0x7f7f7f7fu) + 0x7f7f7f7fu) | x) & 0x80808080u);
        any |= x;
      }
    }
    const bool nz = __ballot(any != 0) != 0;
    if (lane == 0) flags[ch] = nz ? 1u : 0u;
  }
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) c += __shfl_xor(c, d);
  __shared__ uint32_t wsum[4];
  if (lane == 0) wsum[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(&counters[kNzSlot], wsum[0] + wsum[1] + wsum[2] + wsum[3]);
}

// Exact recount of the candidate rows of the 16-bit-sum walk, one wave per
// row (grid-stride over the list): the row's spots are read from the
// orientation-major table (its offset is the tile offset plus the counts of
// the rows before it in the tile), and nspots (nonzero pixels) and the sum
// are counted again; rows passing nspots >= MinNrSpots and sum >= MinIntensity
// (src/LaueMatchingCPU.c:683-689) are appended as (row << 26 | nspots << 17 |
// sum), the record of the other walks.
__global__ void __launch_bounds__(256)
recount_kernel(const uint16_t *__restrict__ counts, const uint64_t *__restrict__ tileOff,
               const uint32_t *__restrict__ entries, uint64_t n, const uint8_t *__restrict__ img,
               const Pred *__restrict__ pred, const uint64_t *__restrict__ cands,
               uint64_t candCap, uint64_t *__restrict__ hits, uint64_t cap,
               uint32_t *__restrict__ counters) {
  const int lane = __lane_id();
  const uint64_t wid = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  const uint64_t ncand = std::min<uint64_t>(counters[kCandSlot], candCap);
  const int min_spots = pred->min_spots[0];
  // the integer sum against the candidate threshold (MinIntensity itself
  // when imScale == 1; a superset otherwise, decided by the in-order re-sum)
  const double cand_thr = pred->cand_thr[0];
  auto wave_sum = [](uint32_t v) {
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) v += __shfl_xor(v, d);
    return v;
  };
  for (uint64_t c = wid; c < ncand; c += nw) {
    const uint64_t o = cands[c];
    if (o >= n) continue;  // never written by the walk (wave-uniform)
    const uint64_t t0 = (o >> kTileBits) << kTileBits;
    const uint32_t r = (uint32_t)(o - t0);
    uint32_t part = 0;
    for (uint32_t i = lane; i < r; i += 64) part += counts[t0 + i];
    const uint64_t off = tileOff[o >> kTileBits] + wave_sum(part);
    const uint32_t no = counts[o];
    uint32_t ns = 0, sum = 0;
    for (uint32_t s0 = 0; s0 < no; s0 += 64) {
      const uint32_t raw = s0 + lane < no ? (uint32_t)img[entries[off + s0 + lane]] : 0u;
      ns += (uint32_t)__popcll(__ballot(raw != 0));
      sum += wave_sum(raw);
    }
    if (lane == 0 && ns > 0 && (int)ns >= min_spots && (double)sum >= cand_thr) {
      const uint32_t pos = atomicAdd(&counters[0], 1u);
      if (pos < cap) hits[pos] = (o << 26) | ((uint64_t)ns << kSumBits) | sum;
    }
  }
}

// Largest number of entries in one chunk of a banded copy.
__global__ void chunk_max_kernel(const uint32_t *__restrict__ chunkOff, uint64_t nsuper,
                                 int nchunks, uint32_t *__restrict__ out) {
  const uint64_t total = nsuper * (uint64_t)nchunks;
  uint32_t m = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t t = i / (uint64_t)nchunks;
    const uint64_t j = i - t * (uint64_t)nchunks;
    const uint32_t *co = chunkOff + t * (uint64_t)(nchunks + 1);
    m = max(m, co[j + 1] - co[j]);
  }
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, d));
  if (__lane_id() == 0) atomicMax(out, m);
}

// Batch compare over a banded copy: B images interleaved pixel-major
// ([pixel][B] bytes), one B-byte gather per spot, LDS counters per
// (image, row) in image-major planes cnt[b*T + row].  Same flat, pipelined
// walk of the super-tile's pixel-sorted entries as the single-image kernel;
// since every workgroup sweeps its bands in pixel order, the interleaved
// rows being gathered are shared through L2 by the workgroups in flight.
typedef uint32_t u32x4b __attribute__((ext_vector_type(4)));

template <int B>
__device__ __forceinline__ void gather_b(const uint8_t *__restrict__ img, uint64_t pix,
                                         uint32_t (&w)[B / 4]) {
  if constexpr (B == 4) {
    w[0] = *(const uint32_t *)(img + pix * 4);
  } else if constexpr (B == 8) {
    const uint2 v = *(const uint2 *)(img + pix * 8);
    w[0] = v.x;
    w[1] = v.y;
  } else {
    const u32x4b *q = (const u32x4b *)(img + pix * B);
#pragma unroll
    for (int j = 0; j < B / 16; ++j) {
      const u32x4b v = q[j];
      w[4 * j] = v.x;
      w[4 * j + 1] = v.y;
      w[4 * j + 2] = v.z;
      w[4 * j + 3] = v.w;
    }
  }
}

template <int TB, int B, int U>
__global__ void __launch_bounds__((1 << TB) < 1024 ? (1 << TB) : 1024)
score_band_batch_kernel(const uint32_t *__restrict__ bentries,
                        const uint64_t *__restrict__ tileOff,
                        const uint32_t *__restrict__ bandOff, int nbands, uint64_t n,
                        uint64_t nsuper, const uint8_t *__restrict__ img,
                        const Pred *__restrict__ pred, uint64_t *__restrict__ hits,
                        uint64_t capPer, uint32_t *__restrict__ counters) {
  constexpr int T = 1 << TB;
  constexpr int THREADS = T < 1024 ? T : 1024;
  constexpr int PB = 32 - TB;
  constexpr uint32_t PMASK = (1u << PB) - 1u;
  constexpr uint32_t STEP = THREADS * U;
  static_assert(B * T * 4 <= 131072, "LDS counters exceed 128 KiB");
  __shared__ uint32_t cnt[B * T];
  const int tid = threadIdx.x, lane = __lane_id();
  const int nimg = pred->nimg;
  for (uint64_t t = blockIdx.x; t < nsuper; t += gridDim.x) {
    for (int i = tid; i < B * T; i += THREADS) cnt[i] = 0;
    __syncthreads();
    const uint64_t first = t << TB;
    const uint32_t *src = bentries + tileOff[first >> kTileBits];
    const uint32_t *bo = bandOff + t * (uint64_t)(nbands + 1);
    const uint32_t total = bo[nbands];
    int band = 0;
    uint32_t bend = bo[1];
    uint32_t e[U];
#pragma unroll
    for (int k = 0; k < U; ++k) {
      const uint32_t i = k * THREADS + tid;
      e[k] = i < total ? __builtin_nontemporal_load(src + i) : 0u;
    }
    for (uint32_t g0 = 0; g0 < total; g0 += STEP) {
      uint32_t w[U][B / 4];
      uint32_t row[U];
      bool ok[U];
#pragma unroll
      for (int k = 0; k < U; ++k) {
        const uint32_t i = g0 + k * THREADS + tid;
        row[k] = e[k] >> PB;
        ok[k] = i < total;
        if (ok[k]) {
          while (i >= bend) {
            ++band;
            bend = bo[band + 1];
          }
          gather_b<B>(img, ((uint64_t)band << PB) | (e[k] & PMASK), w[k]);
        }
      }
#pragma unroll
      for (int k = 0; k < U; ++k) {
        const uint32_t i = g0 + STEP + k * THREADS + tid;
        e[k] = i < total ? __builtin_nontemporal_load(src + i) : 0u;
      }
#pragma unroll
      for (int k = 0; k < U; ++k) {
        if (!ok[k]) continue;
        // most pixels are zero: test four images at a time
#pragma unroll
        for (int j = 0; j < B / 4; ++j) {
          const uint32_t v = w[k][j];
          if (!v) continue;
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const uint32_t raw = (v >> (8 * q)) & 0xffu;
            if (raw) atomicAdd(&cnt[(4 * j + q) * T + row[k]], (1u << kSumBits) + raw);
          }
        }
      }
    }
    __syncthreads();
    for (int rr = tid; rr < T; rr += THREADS) {
      const uint64_t o = first + rr;
      const bool live = o < n;
      for (int b = 0; b < nimg; ++b) {
        const uint32_t acc = cnt[b * T + rr];
        const uint32_t ns = acc >> kSumBits, sum = acc & kSumMask;
        const int min_spots = pred->min_spots[b];
        const bool zero_pass = live && ns == 0 && (0 >= min_spots) && (0.0 >= pred->min_int[b]);
        const bool cand = live && ns > 0 && (int)ns >= min_spots && (double)sum >= pred->cand_thr[b];
        const uint64_t zm = __ballot(zero_pass);
        if (zm && lane == 0) atomicAdd(&counters[2 * b + 1], (uint32_t)__popcll(zm));
        const uint64_t m = __ballot(cand);
        if (m) {
          const int leader = __ffsll((unsigned long long)m) - 1;
          uint32_t base = 0;
          if (lane == leader) base = atomicAdd(&counters[2 * b], (uint32_t)__popcll(m));
          base = __shfl(base, leader);
          if (cand) {
            const uint64_t pos = (uint64_t)base + (uint64_t)__popcll(m & ((1ull << lane) - 1ull));
            if (pos < capPer) hits[(uint64_t)b * capPer + pos] = (o << 26) | acc;
          }
        }
      }
    }
    __syncthreads();
  }
}

// The batch walk with 16-bit sums: one LDS word holds the sums of two
// images -- images 4j and 4j+2 in word 2j, images 4j+1 and 4j+3 in word
// 2j+1, as the interleaved pixel word of images 4j..4j+3 splits with the
// masks 0x00ff00ff and 0x00ff00ff << 8 -- so a spot costs B/2 LDS adds
// instead of B, and B images fit 2^tb rows with B * 2^tb <= 65536 (twice
// the rows of score_band_batch_kernel).  The sums are exact (255 *
// MaxNrLaueSpots < 2^16: a half-word never carries into its neighbour); a
// (row, image) pair whose sum reaches max(ceil(cand_thr), MinNrSpots, 1)
// (each nonzero pixel adds at least 1) can pass and is listed as
// b << 32 | row for the exact recount from the orientation-major table
// (nib_recount_kernel), which also counts its spots.  Needs MinIntensity > 0
// for every image (else every row with a spot is a candidate).
template <int TB, int B, int U>
__global__ void __launch_bounds__((1 << TB) < 1024 ? (1 << TB) : 1024)
score_band_batch16_kernel(const uint32_t *__restrict__ bentries,
                          const uint64_t *__restrict__ tileOff,
                          const uint32_t *__restrict__ bandOff, int nbands, uint64_t n,
                          uint64_t nsuper, const uint8_t *__restrict__ img,
                          const Pred *__restrict__ pred, uint64_t *__restrict__ cands,
                          uint64_t candCap, uint32_t *__restrict__ counters) {
  constexpr int T = 1 << TB;
  constexpr int THREADS = T < 1024 ? T : 1024;
  constexpr int PB = 32 - TB;
  constexpr uint32_t PMASK = (1u << PB) - 1u;
  constexpr uint32_t STEP = THREADS * U;
  constexpr int NP = B / 2;  // counter words per row
  static_assert(B >= 4 && B % 4 == 0, "four images per pixel word");
  static_assert(NP * T * 4 <= 131072, "LDS counters exceed 128 KiB");
  __shared__ uint32_t cnt[NP * T];
  const int tid = threadIdx.x, lane = __lane_id();
  const int nimg = pred->nimg;
  for (uint64_t t = blockIdx.x; t < nsuper; t += gridDim.x) {
    for (int i = tid; i < NP * T; i += THREADS) cnt[i] = 0;
    __syncthreads();
    const uint64_t first = t << TB;
    const uint32_t *src = bentries + tileOff[first >> kTileBits];
    const uint32_t *bo = bandOff + t * (uint64_t)(nbands + 1);
    const uint32_t total = bo[nbands];
    int band = 0;
    uint32_t bend = bo[1];
    uint32_t e[U];
#pragma unroll
    for (int k = 0; k < U; ++k) {
      const uint32_t i = k * THREADS + tid;
      e[k] = i < total ? __builtin_nontemporal_load(src + i) : 0u;
    }
    for (uint32_t g0 = 0; g0 < total; g0 += STEP) {
      uint32_t w[U][B / 4];
      uint32_t row[U];
      bool ok[U];
#pragma unroll
      for (int k = 0; k < U; ++k) {
        const uint32_t i = g0 + k * THREADS + tid;
        row[k] = e[k] >> PB;
        ok[k] = i < total;
        if (ok[k]) {
          while (i >= bend) {
            ++band;
            bend = bo[band + 1];
          }
          gather_b<B>(img, ((uint64_t)band << PB) | (e[k] & PMASK), w[k]);
        }
      }
#pragma unroll
      for (int k = 0; k < U; ++k) {
        const uint32_t i = g0 + STEP + k * THREADS + tid;
        e[k] = i < total ? __builtin_nontemporal_load(src + i) : 0u;
      }
#pragma unroll
      for (int k = 0; k < U; ++k) {
        if (!ok[k]) continue;
#pragma unroll
        for (int j = 0; j < B / 4; ++j) {
          const uint32_t v = w[k][j];
          const uint32_t lo = v & 0x00ff00ffu, hi = (v >> 8) & 0x00ff00ffu;
          if (lo) atomicAdd(&cnt[(2 * j) * T + row[k]], lo);
          if (hi) atomicAdd(&cnt[(2 * j + 1) * T + row[k]], hi);
        }
      }
    }
    __syncthreads();
    for (int rr = tid; rr < T; rr += THREADS) {
      const uint64_t o = first + rr;
      const bool live = o < n;
      for (int b = 0; b < nimg; ++b) {
        const uint32_t word = cnt[(2 * (b >> 2) + (b & 1)) * T + rr];
        const uint32_t sum = (b & 2) ? word >> 16 : word & 0xffffu;
        const uint32_t pk = pred->pk_thr[b];
        const uint32_t nthr = (pk >> 23) > 1u ? (pk >> 23) : 1u, sthr = pk & 0x7fffffu;
        const bool cand = live && sum >= nthr && sum >= sthr;
        const uint64_t m = __ballot(cand);
        if (m) {
          const int leader = __ffsll((unsigned long long)m) - 1;
          uint32_t base = 0;
          if (lane == leader) base = atomicAdd(&counters[kCandSlot], (uint32_t)__popcll(m));
          base = __shfl(base, leader);
          if (cand) {
            const uint64_t pos = (uint64_t)base + (uint64_t)__popcll(m & ((1ull << lane) - 1ull));
            if (pos < candCap) cands[pos] = ((uint64_t)b << 32) | o;
          }
        }
      }
    }
    __syncthreads();
  }
}

int reserve_bytes(lm_db *db, uint64_t bytes) {
  std::lock_guard<std::mutex> lk(db->ctx->mu);
  if (db->ctx->in_use + bytes > db->ctx->budget) {
    set_error("HBM budget exceeded by the banded table copy");
    return LM_ERR_BUDGET;
  }
  db->ctx->in_use += bytes;
  return LM_OK;
}

void release_one(lm_db *db, Banded &b) {
  if (b.entries) (void)hipFree(b.entries);
  if (b.band_off) (void)hipFree(b.band_off);
  if (b.chunk_off) (void)hipFree(b.chunk_off);
  if (b.bytes) {
    std::lock_guard<std::mutex> lk(db->ctx->mu);
    db->ctx->in_use -= b.bytes;
  }
  b = Banded{};
}

int build_one(lm_db *db, Banded &b, int tb, bool chunks) {
  release_one(db, b);
  if (tb == 0) return LM_OK;
  const Table &t = db->table;
  const uint64_t npx = (uint64_t)db->geom.nrPxX * db->geom.nrPxY;
  b.tb = tb;
  b.pb = 32 - tb;
  b.nbands = (int)((npx + (1ull << b.pb) - 1) >> b.pb);
  int bb = kBucketBits;
  while (((npx + (1ull << bb) - 1) >> bb) > (uint64_t)kMaxBuckets) ++bb;
  if (bb > b.pb) {
    set_error("detector too large for the banded table");
    return LM_ERR_INVALID;
  }
  const int nbuckets = (int)((npx + (1ull << bb) - 1) >> bb);
  b.nsuper = (db->n + (1ull << tb) - 1) >> tb;
  const uint64_t ebytes = (t.total + 64) * 4;
  const uint64_t obytes = b.nsuper * (uint64_t)(b.nbands + 1) * 4;
  // chunk offsets for the LDS-staged walk (chunks must be whole buckets)
  const int cb = wave_chunk_bits(tb);
  chunks = chunks && bb <= cb;
  b.nchunks = chunks ? (int)((npx + (1ull << cb) - 1) >> cb) : 0;
  // pieces per chunk: as set, or the fewest that bring the mean piece to at
  // most 512 entries (the walk's fast loads hold 8 or 12 entries per lane;
  // a small detector puts thousands of entries in each chunk)
  b.ks = 0;
  if (chunks && db->banded_split >= 0) {
    b.ks = db->banded_split;
  } else if (chunks) {
    const double mean = (double)t.total / (double)std::max<uint64_t>(db->n, 1) *
                        (double)(1ull << tb) / (double)b.nchunks;
    while (b.ks < kMaxSplitBits && mean > 512.0 * (double)(1 << b.ks)) ++b.ks;
  }
  const uint64_t cbytes =
      chunks ? b.nsuper * (((uint64_t)b.nchunks << b.ks) + 1) * 4 : 0;
  int rc = reserve_bytes(db, ebytes + obytes + cbytes);
  if (rc) {
    b = Banded{};
    return rc;
  }
  b.bytes = ebytes + obytes + cbytes;
  hipError_t e = hipMalloc(&b.entries, ebytes);
  if (e == hipSuccess) e = hipMalloc(&b.band_off, obytes);
  if (e == hipSuccess && chunks) e = hipMalloc(&b.chunk_off, cbytes);
  hipStream_t st = db->ctx->stream;
  if (e == hipSuccess) e = hipMemsetAsync(b.entries + t.total, 0, 64 * 4, st);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(band_build_kernel, dim3((unsigned)b.nsuper), dim3(kBuildThreads), 0, st,
                       t.counts, t.tile_off, t.entries, db->n, tb, b.pb, bb, nbuckets, b.nbands,
                       b.entries, b.band_off, cb, b.nchunks, b.ks, b.chunk_off);
    e = hipGetLastError();
  }
  if (e == hipSuccess && chunks) {
    uint32_t *d_max = nullptr;
    e = hipMalloc(&d_max, 4);
    if (e == hipSuccess) e = hipMemsetAsync(d_max, 0, 4, st);
    if (e == hipSuccess) {
      hipLaunchKernelGGL(chunk_max_kernel, dim3(1024), dim3(256), 0, st, b.chunk_off, b.nsuper,
                         b.nchunks << b.ks, d_max);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(&b.max_chunk, d_max, 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (d_max) (void)hipFree(d_max);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) {
    release_one(db, b);
    return hip_fail(e, "build_banded");
  }
  b.ready = true;
  return LM_OK;
}

}  // namespace

void release_banded(lm_db *db) {
  ++db->generation;  // any captured launch that read the old table is stale
  release_one(db, db->banded);
  for (Banded &b : db->banded_bk) release_one(db, b);
}

int build_banded(lm_db *db, bool best_effort) {
  release_banded(db);
  if (db->banded_tb != 0 && (db->banded_tb < 13 || db->banded_tb > kSum16TileBits)) {
    set_error("banded super-tile bits must be 0 or 13..16");
    return LM_ERR_INVALID;
  }
  if (db->banded_tb == kSum16TileBits && db->max_spots > kSum16MaxSpots) {
    set_error("2^16-row super-tiles keep 16-bit sums: MaxNrLaueSpots must be at most 257");
    return LM_ERR_INVALID;
  }
  for (const int8_t tb : db->width_tb)
    if (tb != 0 && (tb < 9 || tb > 13)) {
      set_error("batch super-tile bits must be 0 or 9..13");
      return LM_ERR_INVALID;
    }
  int rc = build_one(db, db->banded, db->banded_tb, db->banded_tb >= 14);
  // without room for a copy (best effort), scoring reads the orientation-major
  // table instead, which gives the same results
  if (rc == LM_ERR_BUDGET && best_effort) rc = LM_OK;
  // one batch copy per super-tile size some width takes
  for (int tb = 9; tb <= 13 && rc == LM_OK; ++tb) {
    bool used = false;
    for (const int8_t w : db->width_tb) used |= w == tb;
    if (!used) continue;
    rc = build_one(db, db->banded_bk[tb - 9], tb, false);
    if (rc == LM_ERR_BUDGET && best_effort) rc = LM_OK;
  }
  if (rc) release_banded(db);
  return rc;
}

uint64_t cand_capacity(const lm_db *db) {
  return std::min<uint64_t>(std::max<uint64_t>(db->n / 16, 4096), 1ull << 22);
}

int launch_banded(lm_db *db, Scratch &sc, const uint8_t *img, uint64_t cap, hipStream_t st) {
  const Banded &b = db->banded;
  const Table &t = db->table;
  hipError_t e;
  const unsigned grid = (unsigned)std::min<uint64_t>(b.nsuper, 1u << 20);
#define LM_SB(KERNEL, TB, TH, U)                                                                \
  hipLaunchKernelGGL((KERNEL<TB, TH, U>), dim3(grid), dim3(TH), 0, st, b.entries, t.tile_off,   \
                     b.band_off, b.nbands, db->n, b.nsuper, img, sc.pred, sc.hits, cap,         \
                     sc.counters)
  const uint64_t npx = (uint64_t)db->geom.nrPxX * db->geom.nrPxY;
  // the LDS-staged walk loads the image in 16-byte pieces; an image that is
  // not 16-byte aligned (a caller's device pointer) takes the gather walk
  // 2^16-row super-tiles have only the LDS-staged walk (with a 16-byte
  // aligned image; the caller takes the table walk otherwise)
  const bool s16 = b.tb >= kSum16TileBits;
  const int L = s16 && !(db->banded_loads == 3 || db->banded_loads >= 5) ? 3 : db->banded_loads;
  // (2^13-row super-tiles have no chunk offsets: the LDS walk settings read
  // them with the byte-gather walk)
  const bool lds_walk = (L == 3 || L >= 5) && b.tb >= 14;
  if (s16 && (((uintptr_t)img & 15) || !sc.cands)) {
    set_error("2^16-row walk needs a 16-byte aligned image and a candidate list");
    return LM_ERR_INVALID;
  }
  const bool lds_ok = lds_walk && ((uintptr_t)img & 15) == 0;
  if (lds_ok && (!sc.chunk_nz || sc.chunk_nz_cap < (uint64_t)b.nchunks)) {
    set_error("LDS-staged walk needs the chunk flag buffer");
    return LM_ERR_INVALID;
  }
  // the LDS walk skips the image's all-zero chunks, and the 16-bit-counter
  // walk picks sums or spot counts from the image's nonzero fraction: both
  // counted first on the device
  if (lds_ok)
    hipLaunchKernelGGL(chunk_nz_kernel,
                       dim3((unsigned)std::min<uint64_t>(((uint64_t)b.nchunks + 3) / 4, 128)),
                       dim3(256), 0, st, img, npx, wave_chunk_bits(b.tb), (uint32_t)b.nchunks,
                       sc.chunk_nz, sc.counters);
  // events around the compare kernel alone (not inside a captured graph)
  const bool timed = !sc.use_graphs && sc.ew0 && sc.ew1;
  if (timed) (void)hipEventRecord(sc.ew0, st);
  auto walk_done = [&] {
    if (timed) {
      (void)hipEventRecord(sc.ew1, st);
      sc.walk_timed = true;
    }
  };
  if (lds_ok) {
    if (!b.chunk_off || b.tb < 14 || b.tb > kSum16TileBits) {
      set_error("LDS-staged walk needs 2^14..2^16 super-tiles and chunk offsets");
      return LM_ERR_INVALID;
    }
    // shapes: entry loads per lane per chunk, chunk sets in flight per wave,
    // threads per workgroup (1024: 128 VGPRs per lane; 512: up to 256)
    // chunks larger than EMAX * 64 entries need the kernel's slow path
    // 16-byte entry loads start up to 3 entries before a chunk, so a chunk
    // fits the fast loads when it has at most EM * 64 - 3 entries
    constexpr bool kE16 = true;
    const uint32_t slack = kE16 ? 3u : 0u;
    // mean spots per row (x 256, rounded up): picks sums or counts per image
    const uint64_t rows = std::max<uint64_t>(db->n, 1);
    const uint32_t spots256 = (uint32_t)((t.total * 256u + rows - 1) / rows);
#define LM_LDS(TB, EM, DD, TH)                                                                   \
  if (b.max_chunk + slack <= (uint32_t)(EM)*64u)                                                 \
    hipLaunchKernelGGL((score_band_lds_kernel<TB, wave_chunk_bits(TB), EM, DD, TH, false, kE16>), \
                       dim3(grid), dim3(TH), 0, st, b.entries, t.tile_off, b.chunk_off,           \
                       b.nchunks << b.ks, b.ks, db->n, b.nsuper, img, npx, spots256, sc.pred,     \
                       sc.hits, cap, sc.counters, sc.cands, sc.cand_cap, sc.chunk_nz);           \
  else                                                                                            \
    hipLaunchKernelGGL((score_band_lds_kernel<TB, wave_chunk_bits(TB), EM, DD, TH, true, kE16>),  \
                       dim3(grid), dim3(TH), 0, st, b.entries, t.tile_off, b.chunk_off,           \
                       b.nchunks << b.ks, b.ks, db->n, b.nsuper, img, npx, spots256, sc.pred,     \
                       sc.hits, cap, sc.counters, sc.cands, sc.cand_cap, sc.chunk_nz)
#define LM_LDS_ALL(TB, E1, E2)               \
  switch (L) {                               \
    case 3: LM_LDS(TB, E1, 2, 1024); break;  \
    case 5: LM_LDS(TB, E2, 2, 1024); break;  \
    case 6: LM_LDS(TB, E1, 3, 1024); break;  \
    default: LM_LDS(TB, E2, 3, 1024); break; \
  }
    // twice the rows per super-tile, twice the entries per chunk
    if (b.tb == 16) {
      LM_LDS_ALL(16, 8, 12);
    } else if (b.tb == 15) {
      LM_LDS_ALL(15, 4, 8);
    } else {
      LM_LDS_ALL(14, 4, 8);
    }
#undef LM_LDS_ALL
#undef LM_LDS
    walk_done();
    if (s16) {
      hipLaunchKernelGGL(recount_kernel, dim3(256), dim3(256), 0, st, t.counts, t.tile_off,
                         t.entries, db->n, img, sc.pred, sc.cands, sc.cand_cap, sc.hits, cap,
                         sc.counters);
    }
    e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "banded score kernel launch");
    return LM_OK;
  }
  // 4-byte gathers read whole words of the image: it must be 4-byte aligned
  // and its size a multiple of 4 (else the byte-gather walk is used)
  int loads = (L == 3 || L >= 5) ? 0 : L;
  if (loads == 4 && (((uintptr_t)img & 3) || (npx & 3))) loads = 0;
  const int key = b.tb * 10 + loads;
#define LM_SBW(TB, TH)                                                                           \
  hipLaunchKernelGGL((score_band_kernel<TB, TH, 8, true>), dim3(grid), dim3(TH), 0, st, b.entries, \
                     t.tile_off, b.band_off, b.nbands, db->n, b.nsuper, img, sc.pred, sc.hits,   \
                     cap, sc.counters)
  switch (key) {
    case 150: LM_SB(score_band_kernel, 15, 1024, 8); break;
    case 154: LM_SBW(15, 1024); break;
    case 144: LM_SBW(14, 1024); break;
    case 134: LM_SBW(13, 512); break;
    case 151: LM_SB(score_band_v4_kernel, 15, 1024, 2); break;
    case 152: LM_SB(score_band_v4_kernel, 15, 1024, 4); break;
    case 140: LM_SB(score_band_kernel, 14, 1024, 8); break;
    case 141: LM_SB(score_band_v4_kernel, 14, 1024, 2); break;
    case 142: LM_SB(score_band_v4_kernel, 14, 1024, 4); break;
    case 130: LM_SB(score_band_kernel, 13, 512, 8); break;
    case 131: LM_SB(score_band_v4_kernel, 13, 512, 2); break;
    case 132: LM_SB(score_band_v4_kernel, 13, 512, 4); break;
    default:
      set_error("no banded kernel for this super-tile size and load width");
      return LM_ERR_INVALID;
  }
#undef LM_SB
#undef LM_SBW
  walk_done();
  e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "banded score kernel launch");
  return LM_OK;
}

bool banded_batch_fits(const lm_db *db, int B) {
  if (B < 4 || B > 64) return false;
  const int tb = db->width_tb[width_log2(B)];
  if (tb < 9 || tb > 13) return false;
  const Banded &b = db->banded_bk[tb - 9];
  // 32-bit counters: B * 2^tb <= 32768; 16-bit sums: B * 2^tb <= 65536
  const uint64_t lim = ((db->batch16_widths >> width_log2(B)) & 1) ? 65536 : 32768;
  return b.ready && (uint64_t)B * (1ull << b.tb) <= lim;
}

bool banded_batch16(const lm_db *db, int B) {
  return banded_batch_fits(db, B) && ((db->batch16_widths >> width_log2(B)) & 1);
}

int launch_banded_batch16(lm_db *db, Scratch &sc, int B, const uint8_t *img, hipStream_t st) {
  const Banded &b = db->banded_bk[db->width_tb[width_log2(B)] - 9];
  const Table &t = db->table;
  const unsigned grid = (unsigned)std::min<uint64_t>(b.nsuper, 1u << 20);
  const unsigned threads = (1u << b.tb) < 1024 ? (1u << b.tb) : 1024;
#define LM_B16(TB, BB)                                                                           \
  hipLaunchKernelGGL((score_band_batch16_kernel<TB, BB, 4>), dim3(grid), dim3(threads), 0, st,   \
                     b.entries, t.tile_off, b.band_off, b.nbands, db->n, b.nsuper, img, sc.pred, \
                     sc.cands, sc.cand_cap, sc.counters)
  switch (b.tb * 100 + B) {
    case 1304: LM_B16(13, 4); break;
    case 1308: LM_B16(13, 8); break;
    case 1204: LM_B16(12, 4); break;
    case 1208: LM_B16(12, 8); break;
    case 1216: LM_B16(12, 16); break;
    case 1104: LM_B16(11, 4); break;
    case 1108: LM_B16(11, 8); break;
    case 1116: LM_B16(11, 16); break;
    case 1132: LM_B16(11, 32); break;
    case 1004: LM_B16(10, 4); break;
    case 1008: LM_B16(10, 8); break;
    case 1016: LM_B16(10, 16); break;
    case 1032: LM_B16(10, 32); break;
    case 1064: LM_B16(10, 64); break;
    case 904: LM_B16(9, 4); break;
    case 908: LM_B16(9, 8); break;
    case 916: LM_B16(9, 16); break;
    case 932: LM_B16(9, 32); break;
    case 964: LM_B16(9, 64); break;
    default:
      set_error("no 16-bit batch kernel for this batch width and super-tile size");
      return LM_ERR_INVALID;
  }
#undef LM_B16
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "16-bit batch kernel launch");
  return LM_OK;
}

int launch_banded_batch(lm_db *db, Scratch &sc, int B, const uint8_t *img, uint64_t capPer,
                        hipStream_t st) {
  const Banded &b = db->banded_bk[db->width_tb[width_log2(B)] - 9];
  const Table &t = db->table;
  const unsigned grid = (unsigned)std::min<uint64_t>(b.nsuper, 1u << 20);
  const unsigned threads = (1u << b.tb) < 1024 ? (1u << b.tb) : 1024;
#define LM_BB(TB, BB)                                                                            \
  hipLaunchKernelGGL((score_band_batch_kernel<TB, BB, 4>), dim3(grid), dim3(threads), 0, st,     \
                     b.entries, t.tile_off, b.band_off, b.nbands, db->n, b.nsuper, img, sc.pred, \
                     sc.hits, capPer, sc.counters)
  const int key = b.tb * 100 + B;
  switch (key) {
    case 1304: LM_BB(13, 4); break;
    case 1204: LM_BB(12, 4); break;
    case 1208: LM_BB(12, 8); break;
    case 1104: LM_BB(11, 4); break;
    case 1108: LM_BB(11, 8); break;
    case 1116: LM_BB(11, 16); break;
    case 1004: LM_BB(10, 4); break;
    case 1008: LM_BB(10, 8); break;
    case 1016: LM_BB(10, 16); break;
    case 1032: LM_BB(10, 32); break;
    case 904: LM_BB(9, 4); break;
    case 908: LM_BB(9, 8); break;
    case 916: LM_BB(9, 16); break;
    case 932: LM_BB(9, 32); break;
    case 964: LM_BB(9, 64); break;
    default:
      set_error("no banded batch kernel for this batch width and super-tile size");
      return LM_ERR_INVALID;
  }
#undef LM_BB
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "banded batch kernel launch");
  return LM_OK;
}

}  // namespace lm
