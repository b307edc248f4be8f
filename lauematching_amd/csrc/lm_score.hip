// lm_score.hip - compare/score of uint8 images against the resident spot
// table, with hit compaction.
//
// Reference behaviour (src/LaueMatchingCPU.c:668-689, 772-788): for each
// orientation, over its listed spots: raw = image_u8[y*W + x]; if raw > 0,
// totInt += raw*imScale (float64, table order) and nSpots++.  A hit is
// nSpots >= MinNrSpots && totInt >= MinIntensity, with score
// totInt*sqrt(nSpots); hits are collected in ascending orientation order and
// a zero score is not collected.
//
// Device design.  One workgroup per tile of 256 orientations, one lane per
// orientation: the tile's spot entries are streamed into LDS with coalesced
// loads, then every lane walks its own spots from LDS, gathers the image byte
// (the image stays hot in L2 / Infinity Cache) and keeps the packed integer
// accumulator (nspots << 17 | sum of raw) in a register.  B images can be
// scored in the same pass: they are interleaved pixel-major so one gather of
// B bytes serves all of them, and each lane keeps B accumulators (or B / 16
// lanes keep 16 each).  For 128 and 256 images the interleave holds 4-bit
// upper bounds of the pixels instead (half the bytes per image, so 256
// images share one cache line per spot); their walk keeps exact spot counts
// and bounded sums, and the (row, image) pairs that can pass are recounted
// from the uint8 images.
//
// Exactness.  The integer sum is exact.  When imScale == 1 (the quantizer's
// scale for any image whose maximum is 255) totInt equals that sum and the
// predicate is decided on the device.  Otherwise the device keeps a
// conservative superset of candidates and a second kernel recomputes their
// float64 sums in table order, which the host then tests exactly.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "lm_gather.h"
#include "lm_internal.h"

namespace lm {
namespace {

constexpr int kTile = 256;          // orientations per workgroup (= 1 << tile bits)
constexpr int kStageWords = 12288;  // 48 KB of staged entries per workgroup

// Packed integer form of "nspots >= min_spots && (double)sum >= thr" for
// nspots in 1..255 and integer sums below 2^23: the smallest passing nspots
// (0..256) in bits 23..31, the smallest passing sum in bits 0..22 (a sum
// never reaches 2^23 - 1, so that value means "never", as for a NaN or an
// unreachable threshold).
static uint32_t pack_threshold(int min_spots, double thr) {
  const uint32_t ns = min_spots <= 0 ? 0u : (min_spots > 256 ? 256u : (uint32_t)min_spots);
  constexpr uint32_t kNever = (1u << 23) - 1u;
  uint32_t sum;
  if (thr <= 0.0)
    sum = 0u;
  else if (!(thr < (double)kNever))  // NaN, +inf, or beyond any sum
    sum = kNever;
  else
    sum = (uint32_t)std::ceil(thr);
  return (ns << 23) | sum;
}

__host__ __device__ __forceinline__ bool bit(const Pred &p, int b) {
  return (p.exact[b >> 5] >> (b & 31)) & 1u;
}

template <int B>
__device__ __forceinline__ void accum_px(const uint32_t (&w)[PxW<B>::W], uint32_t (&acc)[B]) {
#pragma unroll
  for (int b = 0; b < B; ++b) {
    const uint32_t raw = B == 1 ? w[0] : (w[b >> 2] >> (8 * (b & 3))) & 0xffu;
    acc[b] += raw ? (1u << kSumBits) + raw : 0u;
  }
}

// Accumulate the c spots at src[0..c) (LDS or global), four gathers in flight
// (two for 128 images: 128 accumulators plus 2 x 32 words in flight).
template <int B>
__device__ __forceinline__ void accum_spots(const uint8_t *__restrict__ img, const uint32_t *src,
                                            uint32_t c, uint32_t (&acc)[B]) {
  uint32_t s = 0;
  if constexpr (B >= 128) {
    for (; s + 2 <= c; s += 2) {
      uint32_t w0[PxW<B>::W], w1[PxW<B>::W];
      const uint32_t p0 = src[s], p1 = src[s + 1];
      load_px<B>(img, p0, w0);
      load_px<B>(img, p1, w1);
      accum_px<B>(w0, acc);
      accum_px<B>(w1, acc);
    }
  }
  for (; s + 4 <= c; s += 4) {
    uint32_t w0[PxW<B>::W], w1[PxW<B>::W], w2[PxW<B>::W], w3[PxW<B>::W];
    const uint32_t p0 = src[s], p1 = src[s + 1], p2 = src[s + 2], p3 = src[s + 3];
    load_px<B>(img, p0, w0);
    load_px<B>(img, p1, w1);
    load_px<B>(img, p2, w2);
    load_px<B>(img, p3, w3);
    accum_px<B>(w0, acc);
    accum_px<B>(w1, acc);
    accum_px<B>(w2, acc);
    accum_px<B>(w3, acc);
  }
  for (; s < c; ++s) {
    uint32_t w[PxW<B>::W];
    load_px<B>(img, src[s], w);
    accum_px<B>(w, acc);
  }
}

template <int B>
__device__ __forceinline__ void swar_spots(const uint8_t *__restrict__ img, const uint32_t *src,
                                           uint32_t c, Swar<B> &a) {
  constexpr int G = B >= 128 ? 2 : 4;  // gathers in flight
  uint32_t s = 0;
  for (; s + G <= c; s += G) {
    uint32_t w[G][PxW<B>::W];
#pragma unroll
    for (int g = 0; g < G; ++g) load_px<B>(img, src[s + g], w[g]);
#pragma unroll
    for (int g = 0; g < G; ++g) swar_px<B>(w[g], a);
  }
  for (; s < c; ++s) {
    uint32_t w[PxW<B>::W];
    load_px<B>(img, src[s], w);
    swar_px<B>(w, a);
  }
}

// counters layout: [2*b] = hits written for image b, [2*b+1] = zero-spot rows
// passing the predicate for image b.
template <int B, bool SW>
__global__ void __launch_bounds__(kTile)
score_kernel(const uint16_t *__restrict__ counts, const uint64_t *__restrict__ tileOff,
             const uint32_t *__restrict__ entries, uint64_t n, uint64_t ntiles,
             const uint8_t *__restrict__ img, const Pred *__restrict__ pred,
             uint64_t *__restrict__ hits, uint64_t capPer, uint32_t *__restrict__ counters) {
  __shared__ uint32_t stage[kStageWords];
  __shared__ uint32_t wsum[kTile / 64];
  const int tid = threadIdx.x;
  const int lane = __lane_id();
  const int wave = tid >> 6;
  for (uint64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const uint64_t o = t * kTile + tid;
    const uint32_t c = o < n ? counts[o] : 0u;
    // exclusive scan of c over the workgroup
    uint32_t incl = c;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t v = __shfl_up(incl, d);
      if (lane >= d) incl += v;
    }
    if (lane == 63) wsum[wave] = incl;
    const uint64_t base = tileOff[t];
    const uint64_t tot = tileOff[t + 1] - base;
    __syncthreads();
    uint32_t excl = incl - c;
    for (int w = 0; w < wave; ++w) excl += wsum[w];
    const bool staged = tot <= (uint64_t)kStageWords;
    if (staged) {
      // 16-byte loads from an aligned start; words outside [base, base+tot) dropped.
      const uint64_t a0 = base & ~3ull;
      const uint64_t nvec = (base + tot - a0 + 3) >> 2;
      const u32x4 *src = (const u32x4 *)(entries + a0);
      for (uint64_t i = tid; i < nvec; i += kTile) {
        const u32x4 v = __builtin_nontemporal_load(src + i);
        const int64_t j = (int64_t)(a0 + 4 * i) - (int64_t)base;
        if (j >= 0 && j < (int64_t)tot) stage[j] = v.x;
        if (j + 1 >= 0 && j + 1 < (int64_t)tot) stage[j + 1] = v.y;
        if (j + 2 >= 0 && j + 2 < (int64_t)tot) stage[j + 2] = v.z;
        if (j + 3 >= 0 && j + 3 < (int64_t)tot) stage[j + 3] = v.w;
      }
    }
    __syncthreads();
    // predicate and compaction of image b's accumulator
    auto finish = [&](int b, uint32_t accb) {
      const uint32_t ns = accb >> kSumBits;
      const uint32_t sum = accb & kSumMask;
      const bool live = o < n;
      bool cand = false, zero_pass = false;
      if (live && b < pred->nimg) {
        if (ns == 0) {
          zero_pass = (0 >= pred->min_spots[b]) && (0.0 >= pred->min_int[b]);
        } else {
          cand = ((int)ns >= pred->min_spots[b]) && ((double)sum >= pred->cand_thr[b]);
        }
      }
      const uint64_t zm = __ballot(zero_pass);
      if (zm && lane == 0) atomicAdd(&counters[2 * b + 1], (uint32_t)__popcll(zm));
      const uint64_t rec = (o << 26) | (uint64_t)accb;  // image b's own region
      append(cand, rec, hits + (uint64_t)b * capPer, &counters[2 * b], capPer);
    };
    if constexpr (SW) {
      Swar<B> a;
#pragma unroll
      for (int j = 0; j < B / 4; ++j) a.s02[j] = a.s13[j] = a.cn[j] = 0;
      if (staged)
        swar_spots<B>(img, stage + excl, c, a);
      else
        swar_spots<B>(img, entries + base + excl, c, a);
#pragma unroll
      for (int b = 0; b < B; ++b) finish(b, swar_acc<B>(a, b));
    } else {
      uint32_t acc[B];
#pragma unroll
      for (int b = 0; b < B; ++b) acc[b] = 0;
      if (staged)
        accum_spots<B>(img, stage + excl, c, acc);
      else
        accum_spots<B>(img, entries + base + excl, c, acc);
#pragma unroll
      for (int b = 0; b < B; ++b) finish(b, acc[b]);
    }
    __syncthreads();
  }
}

// One row of the grouped walk: 16 images of each of the row's c spots.
template <int B>
__device__ __forceinline__ void group_row(const uint32_t *sp, uint32_t c, const uint8_t *simg,
                                          Swar<16> &a) {
  uint32_t s = 0;
  for (; s + 4 <= c; s += 4) {
    uint32_t w[4][4];
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      const u32x4 v = *(const u32x4 *)(simg + (uint64_t)sp[s + g] * B);
      w[g][0] = v.x;
      w[g][1] = v.y;
      w[g][2] = v.z;
      w[g][3] = v.w;
    }
#pragma unroll
    for (int g = 0; g < 4; ++g) swar_px<16>(w[g], a);
  }
  for (; s < c; ++s) {
    const u32x4 v = *(const u32x4 *)(simg + (uint64_t)sp[s] * B);
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
    swar_px<16>(w, a);
  }
}

// Batches of B = 64 or 128 images with L = B / 16 lanes per orientation:
// lane `sub` of a group gathers the 16 bytes of images 16*sub .. 16*sub+15
// of each spot (one 16-byte load) and keeps 12 packed accumulators.  A wave
// instruction then touches 64 / L distinct lines, each fully used, instead
// of 64 lines at B bytes each.  1024 threads per workgroup; a 256-orientation
// tile is done in L / 4 passes over its LDS-staged entries.  The epilogue
// first builds per-lane masks of passing images, so waves without a hit
// (nearly all) skip the per-image compaction.
template <int B>
__global__ void __launch_bounds__(1024)
score_group_kernel(const uint16_t *__restrict__ counts, const uint64_t *__restrict__ tileOff,
                   const uint32_t *__restrict__ entries, uint64_t n, uint64_t ntiles,
                   const uint8_t *__restrict__ img, const Pred *__restrict__ pred,
                   uint64_t *__restrict__ hits, uint64_t capPer, uint32_t *__restrict__ counters) {
  constexpr int L = B / 16;
  constexpr int OPP = 1024 / L;       // orientations per pass
  constexpr int PASSES = kTile / OPP;
  static_assert(PASSES * OPP == kTile, "tile must split into passes");
  __shared__ uint32_t stage[kStageWords];
  __shared__ uint32_t wsum[kTile / 64];
  __shared__ uint32_t rowExcl[kTile], rowCnt[kTile];
  const int tid = threadIdx.x;
  const int lane = __lane_id();
  const int wave = tid >> 6;
  const int sub = tid % L, rl = tid / L;
  const int nimg = pred->nimg;
  // this lane's 16 images: packed thresholds, images in the batch, images
  // for which a row with no spot passes
  uint32_t thr[16];
#pragma unroll
  for (int b = 0; b < 16; ++b) thr[b] = pred->pk_thr[16 * sub + b];
  const int nmine = nimg - 16 * sub;
  const uint32_t inmask = nmine >= 16 ? 0xffffu : (nmine <= 0 ? 0u : (1u << nmine) - 1u);
  const uint32_t zok = (pred->zero_ok[sub >> 1] >> (16 * (sub & 1))) & inmask;
  for (uint64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const uint64_t base = tileOff[t];
    const uint64_t tot = tileOff[t + 1] - base;
    if (tid < kTile) {
      const uint64_t o = t * kTile + tid;
      const uint32_t c = o < n ? counts[o] : 0u;
      uint32_t incl = c;
#pragma unroll
      for (int d = 1; d < 64; d <<= 1) {
        const uint32_t v = __shfl_up(incl, d);
        if (lane >= d) incl += v;
      }
      if (lane == 63) wsum[wave] = incl;
      rowCnt[tid] = c;
      rowExcl[tid] = incl - c;
    }
    __syncthreads();
    if (tid < kTile) {
      uint32_t add = 0;
      for (int w = 0; w < wave; ++w) add += wsum[w];
      rowExcl[tid] += add;
    }
    const bool staged = tot <= (uint64_t)kStageWords;
    if (staged) {
      const uint64_t a0 = base & ~3ull;
      const uint64_t nvec = (base + tot - a0 + 3) >> 2;
      const u32x4 *src = (const u32x4 *)(entries + a0);
      for (uint64_t i = tid; i < nvec; i += 1024) {
        const u32x4 v = __builtin_nontemporal_load(src + i);
        const int64_t j = (int64_t)(a0 + 4 * i) - (int64_t)base;
        if (j >= 0 && j < (int64_t)tot) stage[j] = v.x;
        if (j + 1 >= 0 && j + 1 < (int64_t)tot) stage[j + 1] = v.y;
        if (j + 2 >= 0 && j + 2 < (int64_t)tot) stage[j + 2] = v.z;
        if (j + 3 >= 0 && j + 3 < (int64_t)tot) stage[j + 3] = v.w;
      }
    }
    __syncthreads();
    for (int pass = 0; pass < PASSES; ++pass) {
      const int r = pass * OPP + rl;
      const uint64_t o = t * kTile + r;
      const uint32_t c = rowCnt[r], excl = rowExcl[r];
      const uint8_t *simg = img + 16 * sub;
      Swar<16> a;
#pragma unroll
      for (int j = 0; j < 4; ++j) a.s02[j] = a.s13[j] = a.cn[j] = 0;
      // (two calls, so the staged one reads its indices from LDS: through
      // one generic pointer they would be flat loads, and every gather would
      // wait for all earlier ones)
      if (staged)
        group_row<B>(stage + excl, c, simg, a);
      else
        group_row<B>(entries + base + excl, c, simg, a);
      // masks of this lane's 16 images that pass (cand) or pass with no spot
      uint32_t cmask = 0, zmask = 0;
      if (o < n) {
#pragma unroll
        for (int b = 0; b < 16; ++b) {
          const uint32_t accb = swar_acc<16>(a, b);
          const uint32_t ns = accb >> kSumBits, sum = accb & kSumMask;
          if (ns == 0)
            zmask |= 1u << b;
          else if (ns >= (thr[b] >> 23) && sum >= (thr[b] & 0x7fffffu))
            cmask |= 1u << b;
        }
        zmask &= zok;
        cmask &= inmask;
      }
      if (__ballot((cmask | zmask) != 0)) {
        for (int ss = 0; ss < L; ++ss) {
#pragma unroll
          for (int b = 0; b < 16; ++b) {
            const int ib = 16 * ss + b;
            const bool mine_sub = sub == ss;
            const uint64_t zm = __ballot(mine_sub && ((zmask >> b) & 1u));
            if (zm && lane == 0) atomicAdd(&counters[2 * ib + 1], (uint32_t)__popcll(zm));
            const bool cand = mine_sub && ((cmask >> b) & 1u);
            const uint64_t rec = (o << 26) | (uint64_t)swar_acc<16>(a, b);
            append(cand, rec, hits + (uint64_t)ib * capPer, &counters[2 * ib], capPer);
          }
        }
      }
    }
    __syncthreads();
  }
}

// Batches of B = 128 or 256 images read as 4-bit upper bounds.  The batch
// walk is limited by its random line fetches from the interleaved images, so
// halving the bytes per pixel serves twice the images per fetch.  Nibble
// q = ceil(v / 17) of pixel value v (0..15) is zero exactly when v is, and
// 17 q >= v, so per (row, image) the nibble walk gives the exact count of
// nonzero spots and an upper bound 17 * sum(q) of the raw sum: a row can
// pass only if count >= MinNrSpots and sum(q) >= ceil(ceil(cand_thr) / 17).
// Those (row, image) pairs are candidates; nib_recount_kernel recounts them
// exactly from the uint8 images.  Layout [pixel][B / 2] bytes; byte k of a
// pixel's 16-byte group g holds image 32 g + 2 k in its low nibble and
// 32 g + 2 k + 1 in its high nibble.
//
// Threads: a wave per (32-image group g, 64 pixel quads); a lane loads one
// 4-pixel word of each of its 32 images (256 contiguous bytes per load
// instruction) and writes each pixel's 16 nibble bytes with one store.
template <int B>
__global__ void nibble_kernel(const uint8_t *__restrict__ src, const Pred *__restrict__ pred,
                              uint64_t nquads, uint64_t npx, uint8_t *__restrict__ dst) {
  constexpr int PB = B / 2;
  const int g = threadIdx.x >> 6;
  const uint64_t q = (uint64_t)blockIdx.x * 64 + (threadIdx.x & 63);
  if (q >= nquads) return;
  const int nimg = pred->nimg;
  uint32_t v[32];
#pragma unroll
  for (int j = 0; j < 32; ++j) {
    const int b = 32 * g + j;
    v[j] = b < nimg ? *(const uint32_t *)(src + (uint64_t)b * npx + 4 * q) : 0u;
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    uint32_t w[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      uint32_t word = 0;
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const uint32_t x = ((v[8 * m + k] >> (8 * i)) & 0xffu) + 16u;
        word |= ((x * 241u) >> 12) << (4 * k);  // floor(x / 17) for x < 4096
      }
      w[m] = word;
    }
    *(u32x4 *)(dst + (4 * q + i) * PB + 16 * g) = u32x4{w[0], w[1], w[2], w[3]};
  }
}

// One pixel word (8 nibbles, images 0..7 of the word) into the byte sums of
// the even (bl) and odd (bh) images and the nibble counts of nonzero values
// (cn): at most 15 words between flushes.
__device__ __forceinline__ void nib_word(uint32_t x, uint32_t &bl, uint32_t &bh, uint32_t &cn) {
  bl += x & 0x0f0f0f0fu;
  bh += (x >> 4) & 0x0f0f0f0fu;
  // bit 3 of each nibble of ((x & 7) + 7) | x is set exactly when it is nonzero
  cn += ((((x & 0x77777777u) + 0x77777777u) | x) >> 3) & 0x11111111u;
}

// The 16-bit accumulator fields of a word's 8 images come in pairs (low
// field, high field) = (0, 4), (2, 6), (1, 5), (3, 7) for p = 0..3.
__host__ __device__ constexpr int nib_pair_lo(int p) { return p == 0 ? 0 : (p == 1 ? 2 : (p == 2 ? 1 : 3)); }

// One row of the nibble walk: the row's c spots in groups of four loads in
// flight (past the row's end a group reloads the row's last spot and adds
// zero, so the loads need no branch and each wait counts exactly), byte
// sums and nibble counts of the lane's WPL words (8 images each) over at
// most 12 spots, then widened into the paired 16-bit sums S and counts C.
template <int PB, int WPL>
__device__ __forceinline__ void nib_row(const uint32_t *sp, uint32_t c, const uint8_t *simg,
                                        uint32_t (&S)[4 * WPL], uint32_t (&C)[4 * WPL]) {
  typedef uint32_t vw __attribute__((ext_vector_type(WPL)));
  uint32_t s = 0;
  while (s < c) {
    uint32_t bl[WPL], bh[WPL], cn[WPL];
#pragma unroll
    for (int w = 0; w < WPL; ++w) bl[w] = bh[w] = cn[w] = 0;
    for (int grp = 0; grp < 3 && s < c; ++grp, s += 4) {
      uint32_t idx[4];
#pragma unroll
      for (int g = 0; g < 4; ++g) idx[g] = sp[min(s + g, c - 1u)];
      vw v[4];
#pragma unroll
      for (int g = 0; g < 4; ++g) v[g] = *(const vw *)(simg + (uint64_t)idx[g] * PB);
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        const bool in = s + g < c;
#pragma unroll
        for (int w = 0; w < WPL; ++w) nib_word(in ? v[g][w] : 0u, bl[w], bh[w], cn[w]);
      }
    }
#pragma unroll
    for (int w = 0; w < WPL; ++w) {
      S[4 * w + 0] += bl[w] & 0x00ff00ffu;
      S[4 * w + 1] += (bl[w] >> 8) & 0x00ff00ffu;
      S[4 * w + 2] += bh[w] & 0x00ff00ffu;
      S[4 * w + 3] += (bh[w] >> 8) & 0x00ff00ffu;
      C[4 * w + 0] += cn[w] & 0x000f000fu;
      C[4 * w + 1] += (cn[w] >> 8) & 0x000f000fu;
      C[4 * w + 2] += (cn[w] >> 4) & 0x000f000fu;
      C[4 * w + 3] += (cn[w] >> 12) & 0x000f000fu;
    }
  }
}

// L = B / IPL lanes per orientation, each owning IPL images (one IPL/2-byte
// load per spot); the rest as score_group_kernel.  Per lane IPL/2 words of
// paired 16-bit nibble sums and as many of paired counts; the paired
// thresholds live in LDS, the count thresholds are read only for rows whose
// sums pass.  IPL = 16 keeps the lane's state small enough for two 1024-thread
// workgroups per CU (8 waves per SIMD), IPL = 32 needs 128 VGPRs (4 waves).
// Candidates go to cands[] as image << 32 | row, counted at
// counters[kCandSlot].
template <int B, int IPL>
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(IPL == 16 ? 8 : 4)))
score_nib_kernel(const uint16_t *__restrict__ counts, const uint64_t *__restrict__ tileOff,
                 const uint32_t *__restrict__ entries, uint64_t n, uint64_t ntiles,
                 const uint8_t *__restrict__ nib, const Pred *__restrict__ pred,
                 uint64_t *__restrict__ cands, uint64_t cap, uint32_t *__restrict__ counters) {
  constexpr int L = B / IPL;
  constexpr int PB = B / 2;
  constexpr int WPL = IPL / 8;  // 32-bit words (8 images) per lane and spot
  constexpr int NF = 4 * WPL;   // paired accumulator words per lane
  constexpr int OPP = 1024 / L;
  constexpr int PASSES = kTile / OPP;
  static_assert(PASSES * OPP == kTile, "tile must split into passes");
  __shared__ uint32_t stage[kStageWords];
  __shared__ uint32_t wsum[kTile / 64];
  __shared__ uint32_t rowExcl[kTile], rowCnt[kTile];
  __shared__ uint32_t sthr[L * NF], cthr[L * NF];
  const int tid = threadIdx.x;
  const int lane = __lane_id();
  const int wave = tid >> 6;
  const int sub = tid % L, rl = tid / L;
  if (tid < L * NF) {
    const int k = tid % NF, w = k >> 2, p = k & 3;
    const int b = IPL * (tid / NF) + 8 * w + nib_pair_lo(p);
    sthr[tid] = (uint32_t)pred->nib_sum[b] | ((uint32_t)pred->nib_sum[b + 4] << 16);
    cthr[tid] = (uint32_t)pred->nib_cnt[b] | ((uint32_t)pred->nib_cnt[b + 4] << 16);
  }
  const uint32_t *mys = sthr + NF * sub, *myc = cthr + NF * sub;
  for (uint64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const uint64_t base = tileOff[t];
    const uint64_t tot = tileOff[t + 1] - base;
    if (tid < kTile) {
      const uint64_t o = t * kTile + tid;
      const uint32_t c = o < n ? counts[o] : 0u;
      uint32_t incl = c;
#pragma unroll
      for (int d = 1; d < 64; d <<= 1) {
        const uint32_t v = __shfl_up(incl, d);
        if (lane >= d) incl += v;
      }
      if (lane == 63) wsum[wave] = incl;
      rowCnt[tid] = c;
      rowExcl[tid] = incl - c;
    }
    __syncthreads();
    if (tid < kTile) {
      uint32_t add = 0;
      for (int w = 0; w < wave; ++w) add += wsum[w];
      rowExcl[tid] += add;
    }
    const bool staged = tot <= (uint64_t)kStageWords;
    if (staged) {
      const uint64_t a0 = base & ~3ull;
      const uint64_t nvec = (base + tot - a0 + 3) >> 2;
      const u32x4 *src = (const u32x4 *)(entries + a0);
      for (uint64_t i = tid; i < nvec; i += 1024) {
        const u32x4 v = __builtin_nontemporal_load(src + i);
        const int64_t j = (int64_t)(a0 + 4 * i) - (int64_t)base;
        if (j >= 0 && j < (int64_t)tot) stage[j] = v.x;
        if (j + 1 >= 0 && j + 1 < (int64_t)tot) stage[j + 1] = v.y;
        if (j + 2 >= 0 && j + 2 < (int64_t)tot) stage[j + 2] = v.z;
        if (j + 3 >= 0 && j + 3 < (int64_t)tot) stage[j + 3] = v.w;
      }
    }
    __syncthreads();
    for (int pass = 0; pass < PASSES; ++pass) {
      const int r = pass * OPP + rl;
      const uint64_t o = t * kTile + r;
      const uint32_t c = rowCnt[r], excl = rowExcl[r];
      const uint8_t *simg = nib + (IPL / 2) * sub;
      uint32_t S[NF], C[NF];
#pragma unroll
      for (int k = 0; k < NF; ++k) S[k] = C[k] = 0;
      // (LDS-staged indices through their own call: see score_group_kernel)
      if (staged)
        nib_row<PB, WPL>(stage + excl, c, simg, S, C);
      else
        nib_row<PB, WPL>(entries + base + excl, c, simg, S, C);
      // fields < 2^15 and thresholds <= 0x7fff: bit 15 of each field of
      // (a | 0x8000) - t is set exactly when a >= t
      uint32_t any = 0;
#pragma unroll
      for (int k = 0; k < NF; ++k) any |= ((S[k] | 0x80008000u) - mys[k]) & 0x80008000u;
      if (o < n && any) {
#pragma unroll
        for (int w = 0; w < WPL; ++w)
#pragma unroll
          for (int p = 0; p < 4; ++p) {
            const int k = 4 * w + p;
            const int b = IPL * sub + 8 * w + nib_pair_lo(p);
            const uint32_t g = ((S[k] | 0x80008000u) - mys[k]) &
                               ((C[k] | 0x80008000u) - myc[k]) & 0x80008000u;
            if (g & 0x8000u) {
              const uint32_t pos = atomicAdd(&counters[kCandSlot], 1u);
              if (pos < cap) cands[pos] = ((uint64_t)b << 32) | o;
            }
            if (g & 0x80000000u) {
              const uint32_t pos = atomicAdd(&counters[kCandSlot], 1u);
              if (pos < cap) cands[pos] = ((uint64_t)(b + 4) << 32) | o;
            }
          }
      }
    }
    __syncthreads();
  }
}

// Exact count and sum of the nibble walk's candidates from the uint8 images
// ([nimg][npx]), one lane per candidate; rows passing nspots >= MinNrSpots
// and sum >= cand_thr get the usual hit record in their image's region.
__global__ void nib_recount_kernel(const uint16_t *__restrict__ counts,
                                   const uint64_t *__restrict__ tileOff,
                                   const uint32_t *__restrict__ entries,
                                   const uint64_t *__restrict__ cands, uint64_t cap,
                                   const uint8_t *__restrict__ imgs, uint64_t npx,
                                   const Pred *__restrict__ pred, uint64_t *__restrict__ hits,
                                   uint64_t capPer, uint32_t *__restrict__ counters) {
  const uint64_t ncand = min((uint64_t)counters[kCandSlot], cap);
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < ncand;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t x = cands[i];
    const uint32_t b = (uint32_t)(x >> 32);
    const uint64_t o = x & 0xffffffffull;
    const uint64_t t = o / kTile;
    uint64_t off = tileOff[t];
    for (uint64_t r = t * kTile; r < o; ++r) off += counts[r];
    const uint32_t c = counts[o];
    const uint8_t *im = imgs + (uint64_t)b * npx;
    uint32_t ns = 0, sum = 0;
    for (uint32_t s = 0; s < c; ++s) {
      const uint32_t v = im[entries[off + s]];
      ns += v != 0;
      sum += v;
    }
    if (ns > 0 && (int)ns >= pred->min_spots[b] && (double)sum >= pred->cand_thr[b]) {
      const uint32_t pos = atomicAdd(&counters[2 * b], 1u);
      if (pos < capPer) hits[(uint64_t)b * capPer + pos] = (o << 26) | ((uint64_t)ns << kSumBits) | sum;
    }
  }
}

// In-order float64 scoring, one lane per orientation: totInt accumulates the
// pixel values > 0 in table order exactly like the reference loops
// (src/LaueMatchingCPU.c:647-653 for the float64 image of the DoFwd=1
// branch, :674-681 for raw*imScale of the uint8 image).  Hits with a
// nonzero score go to hits[] (id << 26 | nspots) and vals[] (totInt).
template <typename T>
__global__ void __launch_bounds__(kTile)
score_seq_kernel(const uint16_t *__restrict__ counts, const uint64_t *__restrict__ tileOff,
                 const uint32_t *__restrict__ entries, uint64_t n, uint64_t ntiles,
                 const T *__restrict__ img, double scale, int min_spots, double min_int,
                 uint64_t *__restrict__ hits, double *__restrict__ vals, uint64_t cap,
                 uint32_t *__restrict__ counters) {
  __shared__ uint32_t stage[kStageWords];
  __shared__ uint32_t wsum[kTile / 64];
  const int tid = threadIdx.x;
  const int lane = __lane_id();
  const int wave = tid >> 6;
  for (uint64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const uint64_t o = t * kTile + tid;
    const uint32_t c = o < n ? counts[o] : 0u;
    uint32_t incl = c;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t v = __shfl_up(incl, d);
      if (lane >= d) incl += v;
    }
    if (lane == 63) wsum[wave] = incl;
    const uint64_t base = tileOff[t];
    const uint64_t tot = tileOff[t + 1] - base;
    __syncthreads();
    uint32_t excl = incl - c;
    for (int w = 0; w < wave; ++w) excl += wsum[w];
    const bool staged = tot <= (uint64_t)kStageWords;
    if (staged)
      for (uint64_t i = tid; i < tot; i += kTile) stage[i] = entries[base + i];
    __syncthreads();
    double acc = 0;
    int ns = 0;
    for (uint32_t s = 0; s < c; ++s) {
      const uint32_t pix = staged ? stage[excl + s] : entries[base + excl + s];
      const T v = img[pix];
      if (v > 0) {
        if constexpr (sizeof(T) == 1)
          acc += (double)v * scale;
        else
          acc += (double)v;
        ++ns;
      }
    }
    const bool live = o < n;
    const bool pass = live && ns >= min_spots && acc >= min_int;
    const uint64_t zm = __ballot(pass && ns == 0);
    if (zm && lane == 0) atomicAdd(&counters[1], (uint32_t)__popcll(zm));
    const bool keep = pass && ns > 0;
    const uint64_t m = __ballot(keep);
    if (m) {
      const int leader = __ffsll((unsigned long long)m) - 1;
      uint32_t b0 = 0;
      if (lane == leader) b0 = atomicAdd(&counters[0], (uint32_t)__popcll(m));
      b0 = __shfl(b0, leader);
      if (keep) {
        const uint64_t pos = (uint64_t)b0 + (uint64_t)__popcll(m & ((1ull << lane) - 1ull));
        if (pos < cap) {
          hits[pos] = (o << 26) | (uint64_t)ns;
          vals[pos] = acc;
        }
      }
    }
    __syncthreads();
  }
}

// [nimg][npx] -> [npx][B] (B >= nimg; missing images read as zero).  nimg
// comes from the device predicate, so one captured launch serves every batch
// of width B.
template <int B>
__global__ void interleave_kernel(const uint8_t *__restrict__ src, const Pred *__restrict__ pred,
                                  uint64_t npx, uint8_t *__restrict__ dst) {
  const uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npx) return;
  const int nimg = pred->nimg;
  uint8_t v[B];
#pragma unroll
  for (int b = 0; b < B; ++b) v[b] = b < nimg ? src[(uint64_t)b * npx + p] : 0;
  if constexpr (B == 4) {
    *(uint32_t *)(dst + p * 4) = v[0] | (v[1] << 8) | (v[2] << 16) | ((uint32_t)v[3] << 24);
  } else {
#pragma unroll
    for (int q = 0; q < B / 8; ++q) {
      uint2 w;
      w.x = v[8 * q + 0] | (v[8 * q + 1] << 8) | (v[8 * q + 2] << 16) |
            ((uint32_t)v[8 * q + 3] << 24);
      w.y = v[8 * q + 4] | (v[8 * q + 5] << 8) | (v[8 * q + 6] << 16) |
            ((uint32_t)v[8 * q + 7] << 24);
      *(uint2 *)(dst + p * B + 8 * q) = w;
    }
  }
}

// The same interleave for B >= 16 with 4-byte aligned images and a pixel
// count divisible by 4: a thread takes 4 pixels of 16 images (one 4-byte
// load per image), transposes the 4 x 16 bytes in registers and stores each
// pixel's 16 bytes with one 16-byte store; the B / 16 threads of a pixel
// quad are neighbouring lanes, so each store instruction writes whole
// B-byte pixel rows.
template <int B>
__global__ void interleave4_kernel(const uint8_t *__restrict__ src, const Pred *__restrict__ pred,
                                   uint64_t nquads, uint64_t npx, uint8_t *__restrict__ dst) {
  constexpr int G = B / 16;
  const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t q = tid / G;
  const int g = (int)(tid % G);
  if (q >= nquads) return;
  const int nimg = pred->nimg;
  uint32_t v[16];
#pragma unroll
  for (int j = 0; j < 16; ++j) {
    const int b = 16 * g + j;
    v[j] = b < nimg ? *(const uint32_t *)(src + (uint64_t)b * npx + 4 * q) : 0u;
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    uint32_t w[4];
#pragma unroll
    for (int k = 0; k < 4; ++k)
      w[k] = ((v[4 * k + 0] >> (8 * i)) & 0xffu) | (((v[4 * k + 1] >> (8 * i)) & 0xffu) << 8) |
             (((v[4 * k + 2] >> (8 * i)) & 0xffu) << 16) |
             (((v[4 * k + 3] >> (8 * i)) & 0xffu) << 24);
    *(u32x4 *)(dst + (4 * q + i) * B + 16 * g) = u32x4{w[0], w[1], w[2], w[3]};
  }
}

// Packs the hit records of every image of a batch back to back (image b's
// region starts after the records of images 0..b-1), so the host copies them
// with one transfer.  One workgroup per image; counts <= capPer by the time
// this runs.
__global__ void pack_hits_kernel(const uint64_t *__restrict__ hits, uint64_t capPer,
                                 const uint32_t *__restrict__ counters, int nimg,
                                 uint64_t *__restrict__ out) {
  const int b = blockIdx.x;
  if (b >= nimg) return;
  uint64_t off = 0;
  for (int j = 0; j < b; ++j) off += std::min<uint64_t>(counters[2 * j], capPer);
  const uint64_t n = std::min<uint64_t>(counters[2 * b], capPer);
  for (uint64_t i = threadIdx.x; i < n; i += blockDim.x) out[off + i] = hits[b * capPer + i];
}

// Exact float64 sums of candidates (src/LaueMatchingCPU.c:674-681 order).
// Image b's pixel p is img[b * bstep + p * stride] (interleaved batches:
// bstep 1, stride B; separate images: bstep npx, stride 1).
__global__ void exact_kernel(const uint16_t *__restrict__ counts,
                             const uint64_t *__restrict__ tileOff,
                             const uint32_t *__restrict__ entries, const uint64_t *__restrict__ xids,
                             uint64_t ncand, const uint8_t *__restrict__ img, uint32_t stride,
                             uint64_t bstep, const Pred *__restrict__ pred,
                             double *__restrict__ out) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ncand) return;
  const uint64_t x = xids[i];
  const uint32_t b = (uint32_t)(x >> 32);
  const uint64_t o = x & 0xffffffffull;
  const uint64_t t = o / kTile;
  uint64_t off = tileOff[t];
  for (uint64_t r = t * kTile; r < o; ++r) off += counts[r];
  const int c = counts[o];
  const double scale = pred->scale[b];
  const uint8_t *im = img + (uint64_t)b * bstep;
  double tot = 0;
  for (int s = 0; s < c; ++s) {
    const uint8_t raw = im[(uint64_t)entries[off + s] * stride];
    if (raw > 0) tot += (double)raw * scale;
  }
  out[i] = tot;
}

// The same float64 in-order sums for the hit records of a single image
// (imScale != 1), in the launch sequence right after its walk: vals[i] for
// hits[i], i < min(counters[0], cap), so the host fetches them with the
// records instead of sending the rows back.
__global__ void exact_hits_kernel(const uint16_t *__restrict__ counts,
                                  const uint64_t *__restrict__ tileOff,
                                  const uint32_t *__restrict__ entries,
                                  const uint64_t *__restrict__ hits, uint64_t cap,
                                  const uint32_t *__restrict__ counters,
                                  const uint8_t *__restrict__ img, const Pred *__restrict__ pred,
                                  double *__restrict__ vals) {
  const uint64_t k = min((uint64_t)counters[0], cap);
  const double scale = pred->scale[0];
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < k;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t o = (hits[i] >> 26) & 0xffffffffull;
    const uint64_t t = o / kTile;
    uint64_t off = tileOff[t];
    for (uint64_t r = t * kTile; r < o; ++r) off += counts[r];
    const int c = counts[o];
    double tot = 0;
    for (int s = 0; s < c; ++s) {
      const uint8_t raw = img[entries[off + s]];
      if (raw > 0) tot += (double)raw * scale;
    }
    vals[i] = tot;
  }
}

// A call's results for the host, written by the device straight into the
// pinned buffer (its device address `out`): the counters and, for one image,
// the first min(hits, cap) hit records and, with `exact`, their float64 sums.
// One stream wait then replaces the copies.
__global__ void publish_kernel(const uint32_t *__restrict__ counters,
                               const uint64_t *__restrict__ hits, const double *__restrict__ vals,
                               uint64_t cap, int single, int exact, uint32_t *__restrict__ out) {
  for (int i = threadIdx.x; i < kCounterWords; i += blockDim.x) out[i] = counters[i];
  if (single) {
    const uint64_t k = min((uint64_t)counters[0], cap);
    uint64_t *pre = (uint64_t *)(out + kCounterWords);
    for (uint64_t i = threadIdx.x; i < k; i += blockDim.x) pre[i] = hits[i];
    if (exact) {
      double *hv = (double *)(pre + kHitPrefix);
      for (uint64_t i = threadIdx.x; i < k; i += blockDim.x) hv[i] = vals[i];
    }
  }
}

// Fill kernel for the launch counters (kept as a kernel so that a captured
// launch sequence holds no memset node).
__global__ void zero_kernel(uint32_t *__restrict__ p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// Captured launch sequences kept per scratch (one per DB, source buffer,
// width and capacity in use).
constexpr size_t kMaxGraphs = 32;

// Captured sequences reference the scratch buffers; drop them before any of
// those buffers is reallocated.
void drop_graphs(Scratch &sc) {
  for (auto &g : sc.graphs) (void)hipGraphExecDestroy(g.exec);
  sc.graphs.clear();
}

}  // namespace

// Moves one scratch buffer from old_b to new_b bytes in the scratch's total
// and charges the part beyond `prepaid` to the context's HBM budget.
int scratch_account(Scratch &sc, uint64_t old_b, uint64_t new_b) {
  const uint64_t before = sc.bytes, after = sc.bytes - old_b + new_b;
  if (sc.ctx) {
    const uint64_t cb = before > sc.prepaid ? before - sc.prepaid : 0;
    const uint64_t ca = after > sc.prepaid ? after - sc.prepaid : 0;
    std::lock_guard<std::mutex> lk(sc.ctx->mu);
    if (ca > cb && sc.ctx->in_use + (ca - cb) > sc.ctx->budget) {
      set_error("HBM budget exceeded by scoring buffers");
      return LM_ERR_BUDGET;
    }
    sc.ctx->in_use = sc.ctx->in_use + ca - cb;
  }
  sc.bytes = after;
  return LM_OK;
}

namespace {

// Replaces *buf (cap elements of `elem` bytes) by a buffer of `need`
// elements: captured sequences that used the old one are dropped, the bytes
// charged first (LM_ERR_BUDGET leaves the old buffer in place).
template <typename T>
int regrow(Scratch &sc, T **buf, uint64_t *cap, uint64_t need, uint64_t elem, const char *what) {
  if (*cap >= need) return LM_OK;
  int rc = scratch_account(sc, *cap * elem, need * elem);
  if (rc) return rc;
  drop_graphs(sc);
  if (*buf) (void)hipFree(*buf);
  *buf = nullptr;
  *cap = 0;
  hipError_t e = hipMalloc((void **)buf, need * elem);
  if (e != hipSuccess) {
    (void)scratch_account(sc, need * elem, 0);
    return hip_fail(e, what);
  }
  *cap = need;
  return LM_OK;
}

int ensure_hits(Scratch &sc, uint64_t need) {
  return regrow(sc, &sc.hits, &sc.hit_cap, need, 8, "hipMalloc(hit buffer)");
}

// (a captured single-image sequence may write vals)
int ensure_vals(Scratch &sc, uint64_t need) {
  return regrow(sc, &sc.vals, &sc.val_cap, need, 8, "hipMalloc(hit values)");
}

int ensure_cands(Scratch &sc, uint64_t need) {
  return regrow(sc, &sc.cands, &sc.cand_cap, need, 8, "hipMalloc(candidate rows)");
}

int ensure_batch(Scratch &sc, uint64_t need) {
  return regrow(sc, &sc.batch, &sc.batch_cap, need, 1, "hipMalloc(interleaved batch)");
}

int launch_score(lm_db *db, Scratch &sc, int B, const uint8_t *img, uint64_t capPer,
                 hipStream_t st) {
  const Table &t = db->table;
  const unsigned grid = (unsigned)std::min<uint64_t>(t.ntiles, 1u << 20);
  // packed four-image accumulators need at most 255 spots per row
  const bool sw = db->max_spots <= 255;
#define LM_LAUNCH(BB, SWW)                                                                      \
  hipLaunchKernelGGL((score_kernel<BB, SWW>), dim3(grid), dim3(kTile), 0, st, t.counts,        \
                     t.tile_off, t.entries, db->n, t.ntiles, img, sc.pred, sc.hits, capPer,      \
                     sc.counters)
#define LM_LAUNCH2(BB) \
  if (sw) LM_LAUNCH(BB, true); else LM_LAUNCH(BB, false)
  switch (B) {
    case 1: LM_LAUNCH(1, false); break;
    case 4: LM_LAUNCH2(4); break;
    case 8: LM_LAUNCH2(8); break;
    case 16: LM_LAUNCH2(16); break;
    case 32: LM_LAUNCH2(32); break;
    case 64:
      if (sw && ((db->group_widths >> 6) & 1))
        hipLaunchKernelGGL((score_group_kernel<64>), dim3(grid), dim3(1024), 0, st, t.counts,
                           t.tile_off, t.entries, db->n, t.ntiles, img, sc.pred, sc.hits, capPer,
                           sc.counters);
      else
        LM_LAUNCH2(64);
      break;
    case 128:
      if (sw && ((db->group_widths >> 7) & 1))
        hipLaunchKernelGGL((score_group_kernel<128>), dim3(grid), dim3(1024), 0, st, t.counts,
                           t.tile_off, t.entries, db->n, t.ntiles, img, sc.pred, sc.hits, capPer,
                           sc.counters);
      else
        LM_LAUNCH(128, false);
      break;
    default: set_error("unsupported batch width"); return LM_ERR_INVALID;
  }
#undef LM_LAUNCH2
#undef LM_LAUNCH
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "score kernel launch");
  return LM_OK;
}

int launch_interleave(int B, const uint8_t *src, const Pred *pred, uint64_t npx, uint8_t *dst,
                      hipStream_t st) {
  if (B >= 16 && (npx & 3) == 0 && ((uintptr_t)src & 3) == 0) {
    const uint64_t nquads = npx / 4;
    const unsigned g4 = (unsigned)((nquads * (B / 16) + 255) / 256);
#define LM_IL4(BB)                                                                             \
  hipLaunchKernelGGL(interleave4_kernel<BB>, dim3(g4), dim3(256), 0, st, src, pred, nquads, npx, \
                     dst)
    switch (B) {
      case 16: LM_IL4(16); break;
      case 32: LM_IL4(32); break;
      case 64: LM_IL4(64); break;
      default: LM_IL4(128); break;
    }
#undef LM_IL4
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "interleave kernel launch");
    return LM_OK;
  }
  const unsigned grid = (unsigned)((npx + 255) / 256);
#define LM_IL(BB) \
  hipLaunchKernelGGL(interleave_kernel<BB>, dim3(grid), dim3(256), 0, st, src, pred, npx, dst)
  switch (B) {
    case 4: LM_IL(4); break;
    case 8: LM_IL(8); break;
    case 16: LM_IL(16); break;
    case 32: LM_IL(32); break;
    case 64: LM_IL(64); break;
    case 128: LM_IL(128); break;
    default: set_error("unsupported batch width"); return LM_ERR_INVALID;
  }
#undef LM_IL
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "interleave kernel launch");
  return LM_OK;
}

// Nibble layout of a batch of B = 128 or 256 images: needs 4-byte aligned
// images and a pixel count divisible by 4 (checked by the caller).
int launch_nibble_pack(int B, const uint8_t *src, const Pred *pred, uint64_t npx, uint8_t *dst,
                       hipStream_t st) {
  const uint64_t nquads = npx / 4;
  const unsigned grid = (unsigned)((nquads + 63) / 64);
  if (B == 256)
    hipLaunchKernelGGL(nibble_kernel<256>, dim3(grid), dim3(64 * 8), 0, st, src, pred, nquads,
                       npx, dst);
  else
    hipLaunchKernelGGL(nibble_kernel<128>, dim3(grid), dim3(64 * 4), 0, st, src, pred, nquads,
                       npx, dst);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "nibble kernel launch");
  return LM_OK;
}

// Images per lane of the nibble walk (measured at 100M orientations: 256
// images 34.8 ms with 16, 37.2 ms with 32; 128 images 28.1 ms with 32,
// 30.6 ms with 16).
template <int B>
constexpr int nib_ipl() { return B == 256 ? 16 : 32; }

// Nibble walk over the table, then the exact recount of its candidates.
int launch_nibble_score(lm_db *db, Scratch &sc, int B, const uint8_t *imgs, uint64_t npx,
                        uint64_t capPer, hipStream_t st) {
  const Table &t = db->table;
  // a few workgroups per CU looping over the tiles: the kernel's start-up
  // (its register spills among it) is paid per workgroup, not per tile
  const unsigned grid = (unsigned)std::min<uint64_t>(t.ntiles, 4096);
  if (B == 256)
    hipLaunchKernelGGL((score_nib_kernel<256, nib_ipl<256>()>), dim3(grid), dim3(1024), 0, st, t.counts,
                       t.tile_off, t.entries, db->n, t.ntiles, sc.batch, sc.pred, sc.cands,
                       sc.cand_cap, sc.counters);
  else
    hipLaunchKernelGGL((score_nib_kernel<128, nib_ipl<128>()>), dim3(grid), dim3(1024), 0, st, t.counts,
                       t.tile_off, t.entries, db->n, t.ntiles, sc.batch, sc.pred, sc.cands,
                       sc.cand_cap, sc.counters);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "nibble score kernel launch");
  hipLaunchKernelGGL(nib_recount_kernel, dim3(512), dim3(256), 0, st, t.counts, t.tile_off,
                     t.entries, sc.cands, sc.cand_cap, imgs, npx, sc.pred, sc.hits, capPer,
                     sc.counters);
  e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "nibble recount launch");
  return LM_OK;
}

int fill_hits(lm_hits &h, uint64_t m) {
  if (!m) return LM_OK;
  h.ids = (uint64_t *)malloc(m * 8);
  h.nspots = (int32_t *)malloc(m * 4);
  h.totint = (double *)malloc(m * 8);
  h.score = (double *)malloc(m * 8);
  if (!h.ids || !h.nspots || !h.totint || !h.score) {
    lm_hits_free(&h);
    set_error("out of host memory for hits");
    return LM_ERR_NOMEM;
  }
  return LM_OK;
}

}  // namespace

int batch_width(int nimg) {
  if (nimg <= 1) return 1;
  if (nimg <= 4) return 4;
  int B = 8;
  while (B < nimg) B <<= 1;
  return B;
}

int width_log2(int B) {
  int k = 0;
  while ((1 << k) < B) ++k;
  return k;
}

int scratch_init(Scratch &s) {
  if (s.pred) return LM_OK;
  hipError_t e = hipMalloc(&s.counters, kCounterWords * 4);
  if (e == hipSuccess) e = hipMalloc(&s.pred, sizeof(Pred));
  if (e == hipSuccess) e = hipHostMalloc(&s.pred_host, sizeof(Pred), hipHostMallocDefault);
  if (e == hipSuccess)
    // coherent: the device's writes (publish_kernel) reach host memory
    // without passing through the device caches
    e = hipHostMalloc(&s.host_cnt, kCounterWords * 4 + kHitPrefix * 16,
                      hipHostMallocCoherent | hipHostMallocMapped);
  // (without a device address the results come back by copies)
  if (e == hipSuccess && hipHostGetDevicePointer((void **)&s.host_cnt_dev, s.host_cnt, 0) !=
                             hipSuccess) {
    (void)hipGetLastError();
    s.host_cnt_dev = nullptr;
  }
  if (e == hipSuccess) e = hipEventCreate(&s.ev0);
  if (e == hipSuccess) e = hipEventCreate(&s.ev1);
  if (e == hipSuccess) e = hipEventCreate(&s.ew0);
  if (e == hipSuccess) e = hipEventCreate(&s.ew1);
  if (e != hipSuccess) {
    scratch_release(s);
    return hip_fail(e, "scratch buffers");
  }
  return LM_OK;
}

void scratch_release(Scratch &s) {
  (void)scratch_account(s, s.bytes, 0);
  for (auto &g : s.graphs) (void)hipGraphExecDestroy(g.exec);
  if (s.hits) (void)hipFree(s.hits);
  if (s.counters) (void)hipFree(s.counters);
  if (s.cands) (void)hipFree(s.cands);
  if (s.chunk_nz) (void)hipFree(s.chunk_nz);
  if (s.vals) (void)hipFree(s.vals);
  if (s.batch) (void)hipFree(s.batch);
  if (s.img) (void)hipFree(s.img);
  if (s.fwd_spill) (void)hipFree(s.fwd_spill);
  if (s.pred) (void)hipFree(s.pred);
  if (s.pred_host) (void)hipHostFree(s.pred_host);
  if (s.host_cnt) (void)hipHostFree(s.host_cnt);
  if (s.ev0) (void)hipEventDestroy(s.ev0);
  if (s.ev1) (void)hipEventDestroy(s.ev1);
  if (s.ew0) (void)hipEventDestroy(s.ew0);
  if (s.ew1) (void)hipEventDestroy(s.ew1);
  lm_ctx *ctx = s.ctx;
  const uint64_t prepaid = s.prepaid;
  s = Scratch{};
  s.ctx = ctx;  // the owner may use the scratch again
  s.prepaid = prepaid;
}

int ensure_chunk_flags(Scratch &s, uint64_t n) {
  return regrow(s, &s.chunk_nz, &s.chunk_nz_cap, std::max<uint64_t>(n, 1), 4,
                "hipMalloc(chunk flags)");
}

int scratch_images(Scratch &s, uint64_t bytes) {
  if (s.img_cap >= bytes) return LM_OK;
  // (the staging buffer is never captured in a graph)
  int rc = scratch_account(s, s.img_cap, bytes);
  if (rc) return rc;
  if (s.img) (void)hipFree(s.img);
  s.img = nullptr;
  s.img_cap = 0;
  hipError_t e = hipMalloc(&s.img, bytes);
  if (e != hipSuccess) {
    (void)scratch_account(s, bytes, 0);
    return hip_fail(e, "hipMalloc(image staging)");
  }
  s.img_cap = bytes;
  return LM_OK;
}

uint64_t scratch_bytes_for(const lm_db *db, int max_batch) {
  const uint64_t npx = (uint64_t)db->geom.nrPxX * db->geom.nrPxY;
  const int B = batch_width(std::min(std::max(max_batch, 1), kMaxBatch));
  uint64_t b = (1ull << 16) * 8 * (uint64_t)B;              // hit records
  b += (uint64_t)8 * (1ull << 22);                           // candidate rows (their cap)
  b += (1ull << 16) * 8;                                     // float64 sums
  if (B > 1) b += npx * (uint64_t)std::min(B, kMaxByteBatch);  // interleaved batch
  b += ((uint64_t)db->banded.nchunks + 1) * 4;               // chunk flags
  return b;
}

int score_seq(lm_db *db, Scratch &sc, const void *d_img, bool f64, double scale,
              int min_nr_spots, double min_intensity, hipStream_t st, lm_hits *out) {
  const Table &t = db->table;
  if (!t.ready && db->htable.ready)
    return score_seq_host(db, sc, d_img, f64, scale, min_nr_spots, min_intensity, st, out);
  if (!t.ready) {
    set_error("score: no table (run lm_db_forward or lm_db_import_table)");
    return LM_ERR_STATE;
  }
  if (!d_img || !out) {
    set_error("score: null argument");
    return LM_ERR_INVALID;
  }
  memset(out, 0, sizeof *out);
  int rc = scratch_init(sc);
  if (rc) return rc;
  uint64_t cap = std::max<uint64_t>(sc.hit_cap, 1u << 16);
  uint32_t cnt[2] = {0, 0};
  float ms = 0;
  const unsigned grid = (unsigned)std::min<uint64_t>(t.ntiles, 1u << 20);
  for (int attempt = 0; attempt < 2; ++attempt) {
    rc = ensure_hits(sc, cap);
    if (rc == LM_OK) rc = ensure_vals(sc, cap);
    if (rc) break;
    hipLaunchKernelGGL(zero_kernel, dim3(1), dim3(kCounterWords), 0, st, sc.counters,
                       kCounterWords);
    (void)hipEventRecord(sc.ev0, st);
    if (f64)
      hipLaunchKernelGGL(score_seq_kernel<double>, dim3(grid), dim3(kTile), 0, st, t.counts,
                         t.tile_off, t.entries, db->n, t.ntiles, (const double *)d_img, scale,
                         min_nr_spots, min_intensity, sc.hits, sc.vals, cap, sc.counters);
    else
      hipLaunchKernelGGL(score_seq_kernel<uint8_t>, dim3(grid), dim3(kTile), 0, st, t.counts,
                         t.tile_off, t.entries, db->n, t.ntiles, (const uint8_t *)d_img, scale,
                         min_nr_spots, min_intensity, sc.hits, sc.vals, cap, sc.counters);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) { rc = hip_fail(e, "score kernel launch"); break; }
    (void)hipEventRecord(sc.ev1, st);
    e = hipMemcpyAsync(cnt, sc.counters, 8, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) { rc = hip_fail(e, "score kernel"); break; }
    (void)hipEventElapsedTime(&ms, sc.ev0, sc.ev1);
    if (cnt[0] <= cap) break;
    if (attempt == 1) { set_error("score: hit list overflow"); rc = LM_ERR_OVERFLOW; break; }
    cap = cnt[0] + cnt[0] / 4 + 1024;
  }
  if (rc) return rc;
  const uint64_t k = cnt[0];
  std::vector<uint64_t> rec(k);
  std::vector<double> val(k);
  if (k) {
    hipError_t e = hipMemcpyAsync(rec.data(), sc.hits, k * 8, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(val.data(), sc.vals, k * 8, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(e, "hit copy");
  }
  std::vector<uint32_t> idx(k);
  for (uint64_t i = 0; i < k; ++i) idx[i] = (uint32_t)i;
  std::sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) { return rec[a] < rec[b]; });
  out->nr_results = (uint64_t)cnt[1] + k;
  out->count = k;
  out->kernel_ms = ms;
  out->walk_ms = ms;
  rc = fill_hits(*out, k);
  if (rc) return rc;
  for (uint64_t j = 0; j < k; ++j) {
    const uint64_t r = rec[idx[j]];
    const int ns = (int)(r & 0x3ffffffull);
    out->ids[j] = r >> 26;
    out->nspots[j] = ns;
    out->totint[j] = val[idx[j]];
    out->score[j] = val[idx[j]] * std::sqrt((double)ns);  // src/LaueMatchingCPU.c:688
  }
  return LM_OK;
}

// The device predicate of a call: per image its thresholds, imScale and the
// integer candidate tests the kernels apply; entries past nimg never pass.
// Uploads the predicate unless the device copy already holds it and was
// uploaded on this stream (stream order puts that upload before this call's
// kernels): frames scored with the same thresholds and scale reuse it.
static hipError_t upload_pred(Scratch &sc, const Pred &p, hipStream_t st) {
  if (sc.pred_current && sc.pred_stream == st && memcmp(sc.pred_host, &p, sizeof(Pred)) == 0)
    return hipSuccess;
  memcpy(sc.pred_host, &p, sizeof(Pred));
  const hipError_t e = hipMemcpyAsync(sc.pred, sc.pred_host, sizeof(Pred), hipMemcpyHostToDevice, st);
  sc.pred_current = e == hipSuccess;
  sc.pred_stream = st;
  return e;
}

static void build_pred(Pred &pred, int nimg, const double *im_scales, const int *min_spots,
                       const double *min_int) {
  memset(&pred, 0, sizeof pred);
  pred.nimg = nimg;
  for (int b = 0; b < kMaxBatch; ++b) {
    pred.cand_thr[b] = INFINITY;
    pred.nib_sum[b] = pred.nib_cnt[b] = 0x7fff;
  }
  for (int b = 0; b < nimg; ++b) {
    const double s = im_scales[b];
    pred.min_spots[b] = min_spots[b];
    pred.min_int[b] = min_int[b];
    pred.scale[b] = s;
    if (s == 1.0) {
      // totInt == sum exactly: sum >= min_int decided on the device
      pred.cand_thr[b] = min_int[b];
      pred.exact[b >> 5] |= 1u << (b & 31);
    } else {
      // superset: the sequential float64 sum is within n*2^-53 relative of sum*s
      pred.cand_thr[b] = min_int[b] <= 0 ? -INFINITY : min_int[b] / s * (1.0 - 1e-12);
    }
    pred.pk_thr[b] = pack_threshold(min_spots[b], pred.cand_thr[b]);
    // 4-bit test: a sum >= cand_thr is an integer >= smin = ceil(cand_thr),
    // and 17 * sum(q) >= sum, so sum(q) >= ceil(smin / 17); sums never
    // exceed 255 * 511 < 2^17
    const double ct = pred.cand_thr[b];
    if (ct <= 0)
      pred.nib_sum[b] = 0;
    else if (ct < 131072.0)
      pred.nib_sum[b] = (uint16_t)std::min<uint32_t>(((uint32_t)std::ceil(ct) + 16u) / 17u, 0x7fffu);
    pred.nib_cnt[b] = (uint16_t)std::min(std::max(min_spots[b], 0), 0x7fff);
    if (0 >= min_spots[b] && 0.0 >= min_int[b]) pred.zero_ok[b >> 5] |= 1u << (b & 31);
  }
}

int score_images(lm_db *db, Scratch &sc, int nimg, const uint8_t *d_imgs,
                 const double *im_scales, const int *min_spots, const double *min_int,
                 hipStream_t st, lm_hits *outs) {
  const Table &t = db->table;
  if (!t.ready && db->htable.ready)
    return score_images_host(db, sc, nimg, d_imgs, im_scales, min_spots, min_int, st, outs);
  if (!t.ready) {
    set_error("score: no table (run lm_db_forward or lm_db_import_table)");
    return LM_ERR_STATE;
  }
  if (nimg < 1 || nimg > kMaxBatch || !d_imgs || !im_scales || !outs) {
    set_error("score: need 1..256 images");
    return LM_ERR_INVALID;
  }
  for (int b = 0; b < nimg; ++b)
    if (!(im_scales[b] > 0) || !std::isfinite(im_scales[b])) {
      set_error("score: imScale must be positive and finite");
      return LM_ERR_INVALID;
    }
  int rc = scratch_init(sc);
  if (rc) return rc;
  const uint64_t npx = (uint64_t)db->geom.nrPxX * db->geom.nrPxY;
  const int B = batch_width(nimg);
  // 4-bit bounds for 128 / 256 images (when enabled for the width): every
  // image needs MinIntensity > 0 (else every row is a candidate), and the
  // packing kernel 4-byte aligned images of a pixel count divisible by 4
  bool nib = B >= 128 && ((db->nib_widths >> width_log2(B)) & 1) && (npx & 3) == 0 &&
             ((uintptr_t)d_imgs & 3) == 0;
  for (int b = 0; b < nimg && nib; ++b)
    if (!(min_int[b] > 0)) nib = false;
  // more images than the uint8 walks take, without the 4-bit walk: two calls
  // (and widths the autotuner found cheaper as two calls of half the width)
  auto halves = [&](int h = kMaxByteBatch) -> int {
    int r = score_images(db, sc, h, d_imgs, im_scales, min_spots, min_int, st, outs);
    if (r) return r;
    r = score_images(db, sc, nimg - h, d_imgs + (uint64_t)h * npx, im_scales + h, min_spots + h,
                     min_int + h, st, outs + h);
    if (r) {
      for (int b = 0; b < h; ++b) lm_hits_free(&outs[b]);
      return r;
    }
    // every image reports the device time of both calls
    const double k = outs[0].kernel_ms + outs[h].kernel_ms, w = outs[0].walk_ms + outs[h].walk_ms;
    for (int b = 0; b < nimg; ++b) {
      outs[b].kernel_ms = k;
      outs[b].walk_ms = w;
    }
    return LM_OK;
  };
  if (B > kMaxByteBatch && !nib) return halves();
  if (B >= 8 && B <= 64 && ((db->split_widths >> width_log2(B)) & 1) && nimg > B / 2)
    return halves(B / 2);
  // 2..4 images (width 4) one at a time, when the autotuner found single
  // calls faster than the width-4 launch
  if (B == 4 && ((db->split_widths >> 2) & 1)) {
    double k = 0, w = 0;
    for (int b = 0; b < nimg; ++b) {
      const int r = score_images(db, sc, 1, d_imgs + (uint64_t)b * npx, im_scales + b,
                                 min_spots + b, min_int + b, st, outs + b);
      if (r) {
        for (int c = 0; c < b; ++c) lm_hits_free(&outs[c]);
        return r;
      }
      k += outs[b].kernel_ms;
      w += outs[b].walk_ms;
    }
    for (int b = 0; b < nimg; ++b) {
      outs[b].kernel_ms = k;
      outs[b].walk_ms = w;
    }
    return LM_OK;
  }
  Pred built;
  build_pred(built, nimg, im_scales, min_spots, min_int);
  hipError_t e = upload_pred(sc, built, st);
  if (e != hipSuccess) return hip_fail(e, "predicate upload");
  const Pred &pred = *sc.pred_host;
  uint64_t capPer = std::max<uint64_t>(sc.hit_cap / (uint64_t)B, 1u << 16);
  // kernel: 0 lane-per-orientation over the table, 1 single-image banded
  // copy, 2 batch banded copy, 3 4-bit bounds over the table, 4 batch
  // banded copy with 16-bit sums (3 and 4 recount their candidates)
  int kind = 0;
  // the 16-bit-sum walk keeps rows whose sum reaches MinIntensity as
  // candidates: with MinIntensity <= 0 that is every row, so the table walk
  // is taken instead
  const bool sum16 = db->banded.ready && db->banded.tb >= kSum16TileBits;
  // (imScale != 1: the walk keeps a candidate superset, re-summed in order
  // in float64 below, as for batches)
  if (B == 1 && db->banded.ready &&
      (!sum16 || (min_int[0] > 0 && ((uintptr_t)d_imgs & 15) == 0)))
    kind = 1;
  if (kind == 1 && sum16) {
    rc = ensure_cands(sc, cand_capacity(db));
    if (rc) return rc;
  }
  if (kind == 1 && db->banded.chunk_off) {
    rc = ensure_chunk_flags(sc, (uint64_t)db->banded.nchunks);
    if (rc) return rc;
  }
  if (B > 1 && banded_batch_fits(db, B) && ((db->batch_widths >> width_log2(B)) & 1)) {
    // 4: the batch copy with 16-bit sums and the exact recount of its
    // candidates, when every MinIntensity > 0 (else the table walk)
    bool pos = true;
    for (int b = 0; b < nimg; ++b) pos &= min_int[b] > 0;
    kind = banded_batch16(db, B) ? (pos ? 4 : 0) : 2;
  }
  if (kind == 4) {
    rc = ensure_cands(sc, std::min<uint64_t>(
                              1u << 22, std::max<uint64_t>(cand_capacity(db), db->n * B / 64)));
    if (rc) return rc;
  }
  // 3: 4-bit bounds of 128 / 256 images over the table plus the exact recount
  if (nib) {
    kind = 3;
    rc = ensure_cands(sc, std::min<uint64_t>(
                              1u << 22, std::max<uint64_t>(cand_capacity(db), db->n * B / 64)));
    if (rc) return rc;
  }
  if (B > 1) {
    // (uint8 interleave up to 128 images; 256 images only as nibbles)
    rc = ensure_batch(sc, npx * std::min(B, kMaxByteBatch));
    if (rc) return rc;
  }
  const uint8_t *img = B > 1 ? sc.batch : d_imgs;
  // a single image with imScale != 1: the float64 re-sum of its hits runs in
  // the launch sequence (its values come back with the records)
  const bool dev_exact = nimg == 1 && !bit(pred, 0);
  // interleave (B > 1), counter fill, score kernel: enqueued directly, or
  // replayed from a captured graph keyed by everything the arguments use
  auto sequence = [&](uint64_t cp) -> int {
    if (kind == 3) {
      const int r = launch_nibble_pack(B, d_imgs, sc.pred, npx, sc.batch, st);
      if (r) return r;
    } else if (B > 1) {
      const int r = launch_interleave(B, d_imgs, sc.pred, npx, sc.batch, st);
      if (r) return r;
    }
    hipLaunchKernelGGL(zero_kernel, dim3(1), dim3(kCounterWords), 0, st, sc.counters,
                       kCounterWords);
    if (kind == 3) return launch_nibble_score(db, sc, B, d_imgs, npx, cp, st);
    if (kind == 2) return launch_banded_batch(db, sc, B, img, cp, st);
    if (kind == 4) {
      const int r = launch_banded_batch16(db, sc, B, img, st);
      if (r) return r;
      hipLaunchKernelGGL(nib_recount_kernel, dim3(512), dim3(256), 0, st, t.counts, t.tile_off,
                         t.entries, sc.cands, sc.cand_cap, d_imgs, npx, sc.pred, sc.hits, cp,
                         sc.counters);
      const hipError_t le = hipGetLastError();
      return le == hipSuccess ? LM_OK : hip_fail(le, "batch recount launch");
    }
    const int r = kind == 1 ? launch_banded(db, sc, img, cp, st) : launch_score(db, sc, B, img, cp, st);
    if (r || !dev_exact) return r;
    // one thread per record the buffer can hold (the count is only known on
    // the device; a captured sequence keeps this grid for its capacity)
    const unsigned xg = (unsigned)std::min<uint64_t>((cp + 255) / 256, 65535);
    hipLaunchKernelGGL(exact_hits_kernel, dim3(xg), dim3(256), 0, st, t.counts, t.tile_off,
                       t.entries, sc.hits, cp, sc.counters, img, sc.pred, sc.vals);
    const hipError_t le = hipGetLastError();
    return le == hipSuccess ? LM_OK : hip_fail(le, "exact pass launch");
  };
  auto run = [&](uint64_t cp) -> int {
    if (!sc.use_graphs) return sequence(cp);
    const int gkind = kind + (dev_exact ? 16 : 0);
    hipGraphExec_t exec = nullptr;
    // sequences captured on an older table or layout of this DB can never
    // be replayed again: drop them; the hit is moved to the back (recently
    // used), and past kMaxGraphs the least recently used is dropped
    for (size_t i = 0; i < sc.graphs.size();) {
      if (sc.graphs[i].db == db && sc.graphs[i].generation != db->generation) {
        (void)hipGraphExecDestroy(sc.graphs[i].exec);
        sc.graphs.erase(sc.graphs.begin() + i);
      } else {
        ++i;
      }
    }
    for (size_t i = 0; i < sc.graphs.size(); ++i) {
      const auto &g = sc.graphs[i];
      if (g.db == db && g.src == d_imgs && g.hits == sc.hits && g.width == B && g.cap_per == cp &&
          g.kind == gkind) {
        exec = g.exec;
        std::rotate(sc.graphs.begin() + i, sc.graphs.begin() + i + 1, sc.graphs.end());
        break;
      }
    }
    if (!exec) {
      if (sc.graphs.size() >= kMaxGraphs) {
        (void)hipGraphExecDestroy(sc.graphs.front().exec);
        sc.graphs.erase(sc.graphs.begin());
      }
      hipGraph_t graph = nullptr;
      hipError_t ge = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
      if (ge != hipSuccess) return hip_fail(ge, "hipStreamBeginCapture");
      const int r = sequence(cp);
      ge = hipStreamEndCapture(st, &graph);
      if (r) {
        if (graph) (void)hipGraphDestroy(graph);
        return r;
      }
      if (ge == hipSuccess) ge = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      if (graph) (void)hipGraphDestroy(graph);
      if (ge != hipSuccess) return hip_fail(ge, "graph capture");
      sc.graphs.push_back({db, d_imgs, sc.hits, db->generation, B, cp, gkind, exec});
    }
    ++sc.graph_launches;
    const hipError_t le = hipGraphLaunch(exec, st);
    return le == hipSuccess ? LM_OK : hip_fail(le, "hipGraphLaunch");
  };
  float ms = 0, walk_ms = 0;
  std::vector<uint32_t> cnt(kCounterWords);
  uint64_t *const host_pre = (uint64_t *)(sc.host_cnt + kCounterWords);
  double *const host_vals = (double *)(host_pre + kHitPrefix);
  // at most three runs: a candidate overflow switches to the table walk
  // once, a hit overflow regrows the hit buffer once
  bool grown = false;
  for (;;) {
    rc = ensure_hits(sc, capPer * (uint64_t)B);  // a width-B launch serves any nimg <= B
    if (rc == LM_OK && dev_exact) rc = ensure_vals(sc, capPer);
    if (rc) return rc;
    (void)hipEventRecord(sc.ev0, st);
    sc.walk_timed = false;
    rc = run(capPer);
    if (rc) return rc;
    (void)hipEventRecord(sc.ev1, st);
    // the counters and, for one image, the first hit records in one wait
    if (sc.host_cnt_dev) {
      hipLaunchKernelGGL(publish_kernel, dim3(1), dim3(256), 0, st, sc.counters, sc.hits,
                         sc.vals, std::min(capPer, kHitPrefix), nimg == 1 ? 1 : 0,
                         dev_exact ? 1 : 0, sc.host_cnt_dev);
      e = hipGetLastError();
    } else {
      e = hipMemcpyAsync(sc.host_cnt, sc.counters, kCounterWords * 4, hipMemcpyDeviceToHost,
                         st);
      if (e == hipSuccess && nimg == 1)
        e = hipMemcpyAsync(host_pre, sc.hits, std::min(capPer, kHitPrefix) * 8,
                           hipMemcpyDeviceToHost, st);
      if (e == hipSuccess && dev_exact)
        e = hipMemcpyAsync(host_vals, sc.vals, std::min(capPer, kHitPrefix) * 8,
                           hipMemcpyDeviceToHost, st);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(e, "score kernel");
    memcpy(cnt.data(), sc.host_cnt, kCounterWords * 4);
    (void)hipEventElapsedTime(&ms, sc.ev0, sc.ev1);
    walk_ms = ms;
    if (sc.walk_timed) (void)hipEventElapsedTime(&walk_ms, sc.ew0, sc.ew1);
    if (kind == 1 && sum16 && cnt[kCandSlot] > sc.cand_cap) {
      kind = 0;  // more candidates than the list holds: rerun over the table
      continue;
    }
    if (kind == 4 && cnt[kCandSlot] > sc.cand_cap) {
      kind = 0;  // more candidates than the list holds: rerun over the table
      continue;
    }
    if (kind == 3 && cnt[kCandSlot] > sc.cand_cap) {
      // the same for the 4-bit walk: uint8 walks (two calls above 128 images)
      if (B > kMaxByteBatch) return halves();
      kind = 0;
      continue;
    }
    uint64_t need = 0;
    for (int b = 0; b < nimg; ++b) need = std::max<uint64_t>(need, cnt[2 * b]);
    if (need <= capPer) break;
    if (grown) {
      set_error("score: hit list overflow");
      return LM_ERR_OVERFLOW;
    }
    grown = true;
    capPer = need + need / 4 + 1024;  // rerun once with a buffer that fits
  }
  // every image's hit records in one host buffer, copied with one wait
  std::vector<uint64_t> first(nimg + 1, 0);
  for (int b = 0; b < nimg; ++b) first[b + 1] = first[b] + cnt[2 * b];
  std::vector<uint64_t> all(first[nimg]);
  std::vector<double> dvals(dev_exact ? first[1] : 0);  // device order, as `all`
  if (nimg == 1 && first[1] && first[1] <= std::min(capPer, kHitPrefix)) {
    memcpy(all.data(), host_pre, first[1] * 8);  // already here
    if (dev_exact) memcpy(dvals.data(), host_vals, first[1] * 8);
  } else if (nimg == 1 && first[1]) {
    e = hipMemcpyAsync(all.data(), sc.hits, first[1] * 8, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess && dev_exact)
      e = hipMemcpyAsync(dvals.data(), sc.vals, first[1] * 8, hipMemcpyDeviceToHost, st);
  } else if (first[nimg]) {
    rc = ensure_vals(sc, first[nimg]);
    if (rc) return rc;
    uint64_t *packed = (uint64_t *)sc.vals;
    hipLaunchKernelGGL(pack_hits_kernel, dim3(nimg), dim3(256), 0, st, sc.hits, capPer,
                       sc.counters, nimg, packed);
    e = hipGetLastError();
    if (e == hipSuccess)
      e = hipMemcpyAsync(all.data(), packed, first[nimg] * 8, hipMemcpyDeviceToHost, st);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return hip_fail(e, "hit copy");
  // images with imScale != 1: the float64 in-order sums of all their
  // candidates in one launch (candidate = image << 32 | row; the hit buffer
  // is free again once the records are on the host)
  std::vector<uint64_t> xids;
  std::vector<uint64_t> xfirst(nimg + 1, 0);
  std::vector<double> xtot;
  if (dev_exact) {
    // sort the records with their device-computed sums
    const uint64_t k = first[1];
    std::vector<uint32_t> ord(k);
    for (uint64_t i = 0; i < k; ++i) ord[i] = (uint32_t)i;
    std::sort(ord.begin(), ord.end(), [&](uint32_t a, uint32_t b2) { return all[a] < all[b2]; });
    std::vector<uint64_t> sorted(k);
    xtot.resize(k);
    for (uint64_t i = 0; i < k; ++i) {
      sorted[i] = all[ord[i]];
      xtot[i] = dvals[ord[i]];
    }
    all.swap(sorted);
    xfirst[1] = k;
  } else {
    for (int b = 0; b < nimg; ++b) std::sort(all.begin() + first[b], all.begin() + first[b + 1]);
  }
  for (int b = 0; b < nimg && !dev_exact; ++b) {
    if (!bit(pred, b))
      for (uint64_t i = first[b]; i < first[b + 1]; ++i)
        xids.push_back(((uint64_t)b << 32) | ((all[i] >> 26) & 0xffffffffull));
    xfirst[b + 1] = xids.size();
  }
  if (!dev_exact) xtot.resize(xids.size());
  if (!xids.empty()) {
    const uint64_t k = xids.size();
    rc = ensure_vals(sc, k);
    if (rc) return rc;
    e = hipMemcpyAsync(sc.hits, xids.data(), k * 8, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) {
      // image b's bytes: separate images for the 4-bit walk and single
      // images, interleaved for the uint8 batch walks
      const bool sep = kind == 3 || kind == 4;
      const uint8_t *xi = sep ? d_imgs : img;
      const uint32_t stride = sep ? 1u : (uint32_t)B;
      const uint64_t bstep = sep ? npx : (B > 1 ? 1u : 0u);
      hipLaunchKernelGGL(exact_kernel, dim3((unsigned)((k + 127) / 128)), dim3(128), 0, st,
                         t.counts, t.tile_off, t.entries, sc.hits, k, xi, stride, bstep, sc.pred,
                         sc.vals);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(xtot.data(), sc.vals, k * 8, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(e, "exact pass");
  }
  for (int b = 0; b < nimg; ++b) memset(&outs[b], 0, sizeof(lm_hits));
  for (int b = 0; b < nimg && rc == LM_OK; ++b) {
    lm_hits &h = outs[b];
    h.kernel_ms = ms;
    h.walk_ms = walk_ms;
    const uint64_t k = cnt[2 * b];
    const uint64_t *rec = all.data() + first[b];
    std::vector<double> tot(k);
    for (uint64_t i = 0; i < k; ++i)
      tot[i] = bit(pred, b) ? (double)(rec[i] & kSumMask) : xtot[xfirst[b] + i];
    // final predicate and arrays (ascending id)
    uint64_t m = 0;
    for (uint64_t i = 0; i < k; ++i)
      if (tot[i] >= min_int[b]) ++m;
    h.nr_results = (uint64_t)cnt[2 * b + 1] + m;
    h.count = m;
    rc = fill_hits(h, m);
    if (rc) break;
    uint64_t j = 0;
    for (uint64_t i = 0; i < k; ++i) {
      if (!(tot[i] >= min_int[b])) continue;
      const uint32_t packed = (uint32_t)(rec[i] & 0x3ffffffull);
      const int ns = (int)(packed >> kSumBits);
      h.ids[j] = (rec[i] >> 26) & 0xffffffffull;
      h.nspots[j] = ns;
      h.totint[j] = tot[i];
      h.score[j] = tot[i] * std::sqrt((double)ns);  // src/LaueMatchingCPU.c:688
      ++j;
    }
  }
  if (rc)
    for (int b = 0; b < nimg; ++b) lm_hits_free(&outs[b]);
  return rc;
}

// ---- on-the-fly scoring: no table ------------------------------------------
//
// nimg device images (back to back) scored by projecting every orientation
// in the same launch (launch_fwd_score), in groups of at most 64 images, one
// projection pass per group; then the table path's hit handling: counters,
// packed records sorted per image, the float64 in-order re-sums of the
// candidates of images with imScale != 1 (their spots listed again by
// launch_fwd_exact), and the final predicate.
int score_images_fwd(lm_db *db, Scratch &sc, int nimg, const uint8_t *d_imgs,
                     const double *im_scales, const int *min_spots, const double *min_int,
                     int mode, hipStream_t st, lm_hits *outs, lm_fwd_stats *stats) {
  if (nimg < 1 || nimg > kMaxBatch || !d_imgs || !im_scales || !min_spots || !min_int || !outs) {
    set_error("score_fwd: need 1..256 images and their thresholds");
    return LM_ERR_INVALID;
  }
  if (mode != LM_FWD_FP64 && mode != LM_FWD_FP32_GUARD) {
    set_error("score_fwd: unknown forward mode");
    return LM_ERR_INVALID;
  }
  for (int b = 0; b < nimg; ++b)
    if (!(im_scales[b] > 0) || !std::isfinite(im_scales[b])) {
      set_error("score_fwd: imScale must be positive and finite");
      return LM_ERR_INVALID;
    }
  const uint64_t npx = (uint64_t)db->geom.nrPxX * db->geom.nrPxY;
  constexpr int kGroup = 64;
  if (nimg > kGroup) {
    lm_fwd_stats acc{}, one{};
    for (int b0 = 0; b0 < nimg; b0 += kGroup) {
      const int k = std::min(kGroup, nimg - b0);
      const int r = score_images_fwd(db, sc, k, d_imgs + (uint64_t)b0 * npx, im_scales + b0,
                                     min_spots + b0, min_int + b0, mode, st, outs + b0, &one);
      if (r) {
        for (int b = 0; b < b0; ++b) lm_hits_free(&outs[b]);
        return r;
      }
      acc.n_orients = one.n_orients;
      acc.n_spots = one.n_spots;
      acc.n_pairs += one.n_pairs;
      acc.n_recomputed += one.n_recomputed;
      acc.n_projected += one.n_projected;
      acc.ms += one.ms;
    }
    if (stats) *stats = acc;
    return LM_OK;
  }
  int rc = scratch_init(sc);
  if (rc) return rc;
  const int B = batch_width(nimg);
  Pred built;
  build_pred(built, nimg, im_scales, min_spots, min_int);
  hipError_t e = upload_pred(sc, built, st);
  if (e != hipSuccess) return hip_fail(e, "predicate upload");
  const Pred &pred = *sc.pred_host;
  uint64_t capPer = std::max<uint64_t>(sc.hit_cap / (uint64_t)B, 1u << 16);
  if (B > 1) {
    rc = ensure_batch(sc, npx * B);
    if (rc) return rc;
  }
  // the three statistics counters live in the (otherwise unused) candidate list
  rc = ensure_cands(sc, 4);
  if (rc) return rc;
  unsigned long long *d_stats = (unsigned long long *)sc.cands;
  // list slots 32..M-1 of 4M orientations per launch (512 MiB at M = 64)
  constexpr uint64_t kSpillRows = 1u << 22;
  const uint64_t spillRows = std::min<uint64_t>(kSpillRows, db->n);
  if (db->max_spots > 32) {
    rc = regrow(sc, &sc.fwd_spill, &sc.fwd_spill_cap,
                (spillRows + 256) * (uint64_t)(db->max_spots - 32), 4,
                "hipMalloc(on-the-fly list spill)");
    if (rc) return rc;
  }
  const uint8_t *img = B > 1 ? sc.batch : d_imgs;
  std::vector<uint32_t> cnt(kCounterWords);
  uint64_t hstat[3] = {0, 0, 0};
  float ms = 0;
  bool grown = false;
  for (;;) {
    rc = ensure_hits(sc, capPer * (uint64_t)B);
    if (rc) return rc;
    (void)hipEventRecord(sc.ev0, st);
    if (B > 1) {
      rc = launch_interleave(B, d_imgs, sc.pred, npx, sc.batch, st);
      if (rc) return rc;
    }
    hipLaunchKernelGGL(zero_kernel, dim3(1), dim3(kCounterWords), 0, st, sc.counters,
                       kCounterWords);
    hipLaunchKernelGGL(zero_kernel, dim3(1), dim3(64), 0, st, (uint32_t *)sc.cands, 6);
    rc = launch_fwd_score(db, B, mode == LM_FWD_FP32_GUARD, img, sc.pred, sc.hits, capPer,
                          sc.counters, d_stats, sc.fwd_spill, spillRows + 256, st);
    if (rc) return rc;
    (void)hipEventRecord(sc.ev1, st);
    e = hipMemcpyAsync(sc.host_cnt, sc.counters, kCounterWords * 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(hstat, d_stats, 24, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(e, "on-the-fly scoring");
    memcpy(cnt.data(), sc.host_cnt, kCounterWords * 4);
    (void)hipEventElapsedTime(&ms, sc.ev0, sc.ev1);
    uint64_t need = 0;
    for (int b = 0; b < nimg; ++b) need = std::max<uint64_t>(need, cnt[2 * b]);
    if (need <= capPer) break;
    if (grown) {
      set_error("score_fwd: hit list overflow");
      return LM_ERR_OVERFLOW;
    }
    grown = true;
    capPer = need + need / 4 + 1024;  // rerun once with a buffer that fits
  }
  std::vector<uint64_t> first(nimg + 1, 0);
  for (int b = 0; b < nimg; ++b) first[b + 1] = first[b] + cnt[2 * b];
  std::vector<uint64_t> all(first[nimg]);
  if (first[nimg]) {
    rc = ensure_vals(sc, first[nimg]);
    if (rc) return rc;
    uint64_t *packed = (uint64_t *)sc.vals;
    hipLaunchKernelGGL(pack_hits_kernel, dim3(nimg), dim3(256), 0, st, sc.hits, capPer,
                       sc.counters, nimg, packed);
    e = hipGetLastError();
    if (e == hipSuccess)
      e = hipMemcpyAsync(all.data(), packed, first[nimg] * 8, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(e, "hit copy");
  }
  for (int b = 0; b < nimg; ++b) std::sort(all.begin() + first[b], all.begin() + first[b + 1]);
  // images with imScale != 1: float64 in-order sums of their candidates
  std::vector<uint64_t> xids;
  std::vector<uint64_t> xfirst(nimg + 1, 0);
  for (int b = 0; b < nimg; ++b) {
    if (!bit(pred, b))
      for (uint64_t i = first[b]; i < first[b + 1]; ++i)
        xids.push_back(((uint64_t)b << 32) | ((all[i] >> 26) & 0xffffffffull));
    xfirst[b + 1] = xids.size();
  }
  std::vector<double> xtot(xids.size());
  if (!xids.empty()) {
    const uint64_t k = xids.size();
    rc = ensure_vals(sc, k);
    if (rc) return rc;
    // (the hit buffer holds at least k records and is free again)
    e = hipMemcpyAsync(sc.hits, xids.data(), k * 8, hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return hip_fail(e, "candidate upload");
    rc = launch_fwd_exact(db, sc.hits, k, d_imgs, npx, sc.pred, sc.vals, st);
    if (rc) return rc;
    e = hipMemcpyAsync(xtot.data(), sc.vals, k * 8, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(e, "on-the-fly exact pass");
  }
  for (int b = 0; b < nimg; ++b) memset(&outs[b], 0, sizeof(lm_hits));
  for (int b = 0; b < nimg && rc == LM_OK; ++b) {
    lm_hits &h = outs[b];
    h.kernel_ms = ms;
    h.walk_ms = ms;
    const uint64_t k = cnt[2 * b];
    const uint64_t *rec = all.data() + first[b];
    std::vector<double> tot(k);
    for (uint64_t i = 0; i < k; ++i)
      tot[i] = bit(pred, b) ? (double)(rec[i] & kSumMask) : xtot[xfirst[b] + i];
    uint64_t m = 0;
    for (uint64_t i = 0; i < k; ++i)
      if (tot[i] >= min_int[b]) ++m;
    h.nr_results = (uint64_t)cnt[2 * b + 1] + m;
    h.count = m;
    rc = fill_hits(h, m);
    if (rc) break;
    uint64_t j = 0;
    for (uint64_t i = 0; i < k; ++i) {
      if (!(tot[i] >= min_int[b])) continue;
      const uint32_t packed = (uint32_t)(rec[i] & 0x3ffffffull);
      const int ns = (int)(packed >> kSumBits);
      h.ids[j] = (rec[i] >> 26) & 0xffffffffull;
      h.nspots[j] = ns;
      h.totint[j] = tot[i];
      h.score[j] = tot[i] * std::sqrt((double)ns);  // src/LaueMatchingCPU.c:688
      ++j;
    }
  }
  if (rc) {
    for (int b = 0; b < nimg; ++b) lm_hits_free(&outs[b]);
    return rc;
  }
  if (stats) {
    stats->n_orients = db->n;
    stats->n_spots = hstat[2];
    stats->n_pairs = db->n * (uint64_t)db->nhkls;
    stats->n_recomputed = hstat[0];
    stats->n_projected = hstat[1];
    stats->ms = ms;
  }
  return LM_OK;
}

// ---- scoring that allocates nothing -----------------------------------------
//
// One image scored with every device buffer carved from `scratch`, a
// file-scope device array each GPU holds from module load until the process
// ends; host code uses the address hipGetSymbolAddress gives for the current
// device (noalloc_init, called by lm_ctx_create), never the array's name.
// Calls on one GPU share the array: overlapping calls on one GPU race.

// Layout (256-byte aligned pieces): counters | predicate | chunk flags |
// float64 sums | hit records | candidate rows.
constexpr uint64_t kNaCounters = 4096;
constexpr uint64_t kNaPred = (sizeof(Pred) + 255) / 256 * 256;
constexpr uint64_t kNaChunks = 1u << 16;     // chunk flags of the LDS walk
constexpr uint64_t kNaHits = 1u << 16;       // hit records (and float64 sums)
constexpr uint64_t kNaCands = 1u << 21;      // candidate rows of the 16-bit walk
constexpr uint64_t kNaOffPred = kNaCounters;
constexpr uint64_t kNaOffChunks = kNaOffPred + kNaPred;
constexpr uint64_t kNaOffVals = kNaOffChunks + kNaChunks * 4;
constexpr uint64_t kNaOffHits = kNaOffVals + kNaHits * 8;
constexpr uint64_t kNaOffCands = kNaOffHits + kNaHits * 8;
constexpr uint64_t kNaBytes = kNaOffCands + kNaCands * 8;  // about 17.3 MiB

static __device__ __align__(256) uint8_t scratch[kNaBytes];

namespace {

constexpr int kNaMaxDevices = 64;
std::mutex g_na_mu;
void *g_na_ptr[kNaMaxDevices] = {};

// Image 0's entries of the predicate; every other image never passes.
struct PredOne {
  double min_int, cand_thr, scale;
  int min_spots;
  uint32_t exact, pk_thr, zero_ok;
  uint32_t nib_sum, nib_cnt;
};

// Writes a whole Pred from the kernel argument (no host staging buffer).
__global__ void pred_one_kernel(PredOne p, Pred *__restrict__ d) {
  for (int b = threadIdx.x; b < kMaxBatch; b += blockDim.x) {
    const bool z = b == 0;
    d->min_int[b] = z ? p.min_int : 0.0;
    d->cand_thr[b] = z ? p.cand_thr : INFINITY;
    d->scale[b] = z ? p.scale : 0.0;
    d->min_spots[b] = z ? p.min_spots : 0;
    d->pk_thr[b] = z ? p.pk_thr : 0u;
    d->nib_sum[b] = z ? (uint16_t)p.nib_sum : (uint16_t)0x7fff;
    d->nib_cnt[b] = z ? (uint16_t)p.nib_cnt : (uint16_t)0x7fff;
    if (b < kMaxBatch / 32) {
      d->exact[b] = z ? p.exact : 0u;
      d->zero_ok[b] = z ? p.zero_ok : 0u;
    }
  }
  if (threadIdx.x == 0) d->nimg = 1;
}

// In-place heap sort of keys with their values (no allocation).
void sort_pairs(uint64_t *key, double *val, uint64_t n) {
  auto sift = [&](uint64_t i, uint64_t m) {
    for (;;) {
      uint64_t c = 2 * i + 1;
      if (c >= m) return;
      if (c + 1 < m && key[c + 1] > key[c]) ++c;
      if (key[i] >= key[c]) return;
      std::swap(key[i], key[c]);
      std::swap(val[i], val[c]);
      i = c;
    }
  };
  for (uint64_t i = n / 2; i-- > 0;) sift(i, n);
  for (uint64_t m = n; m > 1; --m) {
    std::swap(key[0], key[m - 1]);
    std::swap(val[0], val[m - 1]);
    sift(0, m - 1);
  }
}

}  // namespace

int noalloc_init(int device) {
  if (device < 0 || device >= kNaMaxDevices) {
    set_error("device index out of range for the scratch table");
    return LM_ERR_INVALID;
  }
  std::lock_guard<std::mutex> lk(g_na_mu);
  if (g_na_ptr[device]) return LM_OK;
  void *ptr = nullptr;
  hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
  g_na_ptr[device] = ptr;
  return LM_OK;
}

int score_noalloc(lm_db *db, const uint8_t *d_img, double im_scale, int min_sp, double min_int,
                  hipStream_t st, uint64_t cap, uint64_t *ids, int32_t *nspots, double *totint,
                  double *score, uint64_t *count, uint64_t *nr_results) {
  const Table &t = db->table;
  if (!t.ready) {
    set_error("score: no table (run lm_db_forward or lm_db_import_table)");
    return LM_ERR_STATE;
  }
  if (!d_img || !count || !nr_results || (cap && (!ids || !nspots || !totint || !score))) {
    set_error("lm_score_dev_noalloc: bad arguments");
    return LM_ERR_INVALID;
  }
  if (!(im_scale > 0) || !std::isfinite(im_scale)) {
    set_error("score: imScale must be positive and finite");
    return LM_ERR_INVALID;
  }
  if (db->banded.ready && db->banded.chunk_off && (uint64_t)db->banded.nchunks > kNaChunks) {
    set_error("lm_score_dev_noalloc: the detector has more than 65536 pixel chunks "
              "(at most 2^27 pixels); use lm_score_dev");
    return LM_ERR_INVALID;
  }
  uint8_t *base = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_na_mu);
    base = (uint8_t *)g_na_ptr[db->ctx->device];
  }
  if (!base) {
    set_error("lm_score_dev_noalloc: scratch address not set up (lm_ctx_create)");
    return LM_ERR_STATE;
  }
  Scratch sc;  // views into `scratch`; no events, no graphs, not charged
  sc.counters = (uint32_t *)base;
  sc.pred = (Pred *)(base + kNaOffPred);
  sc.chunk_nz = (uint32_t *)(base + kNaOffChunks);
  sc.chunk_nz_cap = kNaChunks;
  sc.vals = (double *)(base + kNaOffVals);
  sc.val_cap = kNaHits;
  sc.hits = (uint64_t *)(base + kNaOffHits);
  sc.hit_cap = kNaHits;
  sc.cands = (uint64_t *)(base + kNaOffCands);
  sc.cand_cap = std::min<uint64_t>(cand_capacity(db), kNaCands);
  {
    Pred pred;
    build_pred(pred, 1, &im_scale, &min_sp, &min_int);
    PredOne p1;
    p1.min_int = pred.min_int[0];
    p1.cand_thr = pred.cand_thr[0];
    p1.scale = pred.scale[0];
    p1.min_spots = pred.min_spots[0];
    p1.exact = pred.exact[0];
    p1.pk_thr = pred.pk_thr[0];
    p1.zero_ok = pred.zero_ok[0];
    p1.nib_sum = pred.nib_sum[0];
    p1.nib_cnt = pred.nib_cnt[0];
    hipLaunchKernelGGL(pred_one_kernel, dim3(1), dim3(256), 0, st, p1, sc.pred);
  }
  const bool exact = im_scale == 1.0;
  const bool sum16 = db->banded.ready && db->banded.tb >= kSum16TileBits;
  int kind = (db->banded.ready && (!sum16 || (min_int > 0 && ((uintptr_t)d_img & 15) == 0)))
                 ? 1 : 0;
  uint32_t cnt[kCounterWords];
  for (;;) {
    hipLaunchKernelGGL(zero_kernel, dim3(1), dim3(kCounterWords), 0, st, sc.counters,
                       kCounterWords);
    int rc = kind == 1 ? launch_banded(db, sc, d_img, kNaHits, st)
                       : launch_score(db, sc, 1, d_img, kNaHits, st);
    if (rc) return rc;
    if (!exact) {
      hipLaunchKernelGGL(exact_hits_kernel, dim3(64), dim3(256), 0, st, t.counts, t.tile_off,
                         t.entries, sc.hits, kNaHits, sc.counters, d_img, sc.pred, sc.vals);
    }
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
      e = hipMemcpyAsync(cnt, sc.counters, sizeof cnt, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(e, "score kernel");
    if (kind == 1 && sum16 && cnt[kCandSlot] > sc.cand_cap) {
      kind = 0;  // more candidates than the list holds: the table walk
      continue;
    }
    break;
  }
  const uint64_t k = cnt[0];
  if (k > kNaHits) {
    set_error("lm_score_dev_noalloc: more than 65536 candidate hits; use lm_score_dev");
    return LM_ERR_OVERFLOW;
  }
  if (k > cap) {
    *count = k;
    set_error("lm_score_dev_noalloc: the output arrays hold fewer entries than the "
              "candidate hits (count holds the number needed)");
    return LM_ERR_OVERFLOW;
  }
  hipError_t e = hipSuccess;
  if (k) e = hipMemcpyAsync(ids, sc.hits, k * 8, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess && k && !exact)
    e = hipMemcpyAsync(totint, sc.vals, k * 8, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return hip_fail(e, "hit copy");
  // ascending rows (records are row << 26 | nspots << 17 | sum), then the
  // reference predicate on the float64 sums, compacted in place
  if (exact)
    std::sort(ids, ids + k);
  else
    sort_pairs(ids, totint, k);
  uint64_t m = 0;
  for (uint64_t i = 0; i < k; ++i) {
    const uint64_t rec = ids[i];
    const double tot = exact ? (double)(rec & kSumMask) : totint[i];
    if (!(tot >= min_int)) continue;
    const int ns = (int)((uint32_t)(rec & 0x3ffffffull) >> kSumBits);
    ids[m] = (rec >> 26) & 0xffffffffull;
    nspots[m] = ns;
    totint[m] = tot;
    score[m] = tot * std::sqrt((double)ns);  // src/LaueMatchingCPU.c:688
    ++m;
  }
  *count = m;
  *nr_results = (uint64_t)cnt[1] + m;
  return LM_OK;
}

}  // namespace lm
