// lm_gather.h - device helpers shared by the scoring walks and the
// on-the-fly forward scoring: one gather of B interleaved image bytes per
// spot, packed (SWAR) per-image accumulation, wave-aggregated hit append.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "lm_internal.h"

namespace lm {

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// Words holding one pixel of B interleaved images.
template <int B>
struct PxW {
  static constexpr int W = B < 4 ? 1 : B / 4;
};

// One gather per spot serves all B images: B bytes at img + pix*B
// (1 byte, 4 bytes, 8 bytes or B/16 16-byte loads).
template <int B>
__device__ __forceinline__ void load_px(const uint8_t *__restrict__ img, uint32_t pix,
                                        uint32_t (&w)[PxW<B>::W]) {
  if constexpr (B == 1) {
    w[0] = img[pix];
  } else if constexpr (B == 4) {
    w[0] = *(const uint32_t *)(img + (uint64_t)pix * 4);
  } else if constexpr (B == 8) {
    const uint2 v = *(const uint2 *)(img + (uint64_t)pix * 8);
    w[0] = v.x;
    w[1] = v.y;
  } else {
    static_assert(B % 16 == 0, "batch width must be 1, 4, 8 or a multiple of 16");
    const u32x4 *p = (const u32x4 *)(img + (uint64_t)pix * B);
#pragma unroll
    for (int q = 0; q < B / 16; ++q) {
      const u32x4 v = p[q];
      w[4 * q + 0] = v.x;
      w[4 * q + 1] = v.y;
      w[4 * q + 2] = v.z;
      w[4 * q + 3] = v.w;
    }
  }
}

// Packed (SWAR) accumulation of four images per 32-bit word: for the word w
// holding pixel bytes x0..x3 of images 4j..4j+3, s02 gathers x0 and x2 in two
// 16-bit sums, s13 gathers x1 and x3, and cn counts the nonzero bytes in four
// 8-bit fields.  Valid while a row has at most 255 spots (8-bit counts; sums
// stay below 255 * 255 < 2^16).  A byte x is nonzero exactly when bit 7 of
// ((x & 0x7f) + 0x7f) | x is set.
template <int B>
struct Swar {
  static constexpr int J = B / 4;
  uint32_t s02[J], s13[J], cn[J];
};

template <int B>
__device__ __forceinline__ void swar_px(const uint32_t (&w)[PxW<B>::W], Swar<B> &a) {
#pragma unroll
  for (int j = 0; j < B / 4; ++j) {
    const uint32_t x = w[j];
    a.s02[j] += x & 0x00ff00ffu;
    a.s13[j] += (x >> 8) & 0x00ff00ffu;
    const uint32_t nz = ((x & 0x7f7f7f7fu) + 0x7f7f7f7fu) | x;
    a.cn[j] += (nz >> 7) & 0x01010101u;
  }
}

// Image b's packed accumulator (nspots << 17 | sum) from the SWAR fields.
template <int B>
__device__ __forceinline__ uint32_t swar_acc(const Swar<B> &a, int b) {
  const int j = b >> 2, q = b & 3;
  const uint32_t s = (q & 1) ? a.s13[j] : a.s02[j];
  const uint32_t sum = (q & 2) ? (s >> 16) : (s & 0xffffu);
  const uint32_t ns = (a.cn[j] >> (8 * q)) & 0xffu;
  return (ns << kSumBits) | sum;
}

// Wave-aggregated append: one atomic per wave, lanes write at base + rank.
__device__ __forceinline__ void append(bool flag, uint64_t rec, uint64_t *__restrict__ out,
                                       uint32_t *__restrict__ counter, uint64_t cap) {
  const uint64_t m = __ballot(flag);
  if (m == 0) return;
  const int lane = __lane_id();
  const int leader = __ffsll((unsigned long long)m) - 1;
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(m));
  base = __shfl(base, leader);
  if (flag) {
    const uint64_t below = m & ((1ull << lane) - 1ull);
    const uint64_t pos = (uint64_t)base + (uint64_t)__popcll(below);
    if (pos < cap) out[pos] = rec;
  }
}

}  // namespace lm
