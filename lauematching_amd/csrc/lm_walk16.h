// lm_walk16.h - the end of a 2^16-row super-tile in the single-image walks
// that keep 16-bit row counters in LDS (two rows per 32-bit word).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "lm_internal.h"

namespace lm {

// Rows that can pass go to the candidate list for the exact recount, and
// the counters are left cleared for the next super-tile; every thread of
// the workgroup calls it after the walk's barrier, and it ends with one.
// Sum mode: a row can pass only if its sum reaches MinIntensity and the spot
// count it bounds (each nonzero pixel adds at least 1) can reach MinNrSpots;
// count mode: the counter is the exact spot count, tested against
// MinNrSpots, and the sum is recounted.
template <int THREADS>
__device__ __forceinline__ void sum16_epilogue(uint32_t *cnt, uint64_t first, uint64_t n,
                                               const Pred *__restrict__ pred, uint32_t pk0,
                                               bool count_mode, uint32_t *__restrict__ counters,
                                               uint64_t *__restrict__ cands, uint64_t candCap) {
  typedef uint32_t v4u __attribute__((ext_vector_type(4)));
  constexpr int T = 1 << kSum16TileBits;
  constexpr int NCNT = T / 2;
  const int tid = threadIdx.x, lane = __lane_id();
  const uint32_t sthr = count_mode ? 0u : (pk0 & 0x7fffffu);
  const uint32_t nthr = (pk0 >> 23) > 1u ? (pk0 >> 23) : 1u;
  const bool zok = pred->zero_ok[0] & 1u;
  if (!zok) {
    // 8 rows per 16-byte LDS read, one max test for the 8 (rows that pass
    // are rare), and the words cleared for the next super-tile
    const uint32_t thr = sthr > nthr ? sthr : nthr;
    for (int i = tid; i < NCNT / 4; i += THREADS) {
      const v4u q = ((v4u *)cnt)[i];
      ((v4u *)cnt)[i] = v4u{0u, 0u, 0u, 0u};
      uint32_t m = 0;
#pragma unroll
      for (int k = 0; k < 4; ++k) m = max(m, max(q[k] & 0xffffu, q[k] >> 16));
      if (m >= thr) {
#pragma unroll
        for (int k = 0; k < 4; ++k)
#pragma unroll
          for (int h = 0; h < 2; ++h) {
            const uint32_t sum = (q[k] >> (16 * h)) & 0xffffu;
            const uint64_t o = first + 8u * (uint32_t)i + 2u * k + h;
            if (o < n && sum >= thr) {
              const uint32_t pos = atomicAdd(&counters[kCandSlot], 1u);
              if (pos < candCap) cands[pos] = o;
            }
          }
      }
    }
    __syncthreads();
    return;
  }
  for (int rr = tid; rr < T; rr += THREADS) {
    const uint64_t o = first + rr;
    const uint32_t sum = (cnt[rr >> 1] >> ((rr & 1) << 4)) & 0xffffu;
    const bool live = o < n;
    // rows with no spot pass only when both thresholds are <= 0
    const uint64_t zm = __ballot(live && sum == 0);
    if (zm && lane == 0) atomicAdd(&counters[1], (uint32_t)__popcll(zm));
    const bool cand = live && sum >= nthr && sum >= sthr;
    const uint64_t m = __ballot(cand);
    if (m) {
      const int leader = __ffsll((unsigned long long)m) - 1;
      uint32_t b0 = 0;
      if (lane == leader) b0 = atomicAdd(&counters[kCandSlot], (uint32_t)__popcll(m));
      b0 = __shfl(b0, leader);
      if (cand) {
        const uint64_t pos = (uint64_t)b0 + (uint64_t)__popcll(m & ((1ull << lane) - 1ull));
        if (pos < candCap) cands[pos] = o;
      }
    }
  }
  __syncthreads();
  for (int i = tid; i < NCNT / 4; i += THREADS) ((v4u *)cnt)[i] = v4u{0u, 0u, 0u, 0u};
  __syncthreads();
}

// Inclusive prefix sum over the 64 lanes of a wavefront: DPP row shifts
// within each row of 16 lanes, then row broadcasts of lanes 15 and 31.
__device__ __forceinline__ uint32_t wave_scan_dpp(uint32_t x) {
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xf, 0xf, false);  // row_shr:1
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xf, 0xf, false);  // row_shr:2
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xf, 0xf, false);  // row_shr:4
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xf, 0xf, false);  // row_shr:8
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xa, 0xf, false);  // row_bcast:15
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xc, 0xf, false);  // row_bcast:31
  return x;
}

}  // namespace lm
