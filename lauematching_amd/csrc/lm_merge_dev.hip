// lm_merge_dev.hip - pair tests of the duplicate merge on the device
// (mergeDuplicateOrientations, src/LaueMatchingHeaders.h:969-1065: every
// pair of candidates is compared by misorientation before the greedy pass).
//
// One lane per pair (l, i), l > i: the misorientation cosine of the two
// fundamental-region quaternions (lm_quat.h, the host's operations in the
// host's order) against the smallest passing cosine the host derived from
// its own acos, so the bit equals the reference's float test.  A wave covers
// 64 consecutive l of one row i and writes one 64-bit word with a ballot.
#include <hip/hip_runtime.h>

#include <vector>

#include "lm_internal.h"

namespace lm {
namespace {

constexpr int kWavesPerBlock = 4;

__global__ void __launch_bounds__(64 * kWavesPerBlock)
merge_bits_kernel(const double *__restrict__ fq, int n, int W, SymSet sym, double cos_min,
                  uint64_t *__restrict__ bits) {
  const int i = blockIdx.y;
  const int wd = blockIdx.x * kWavesPerBlock + (int)(threadIdx.x >> 6);
  const int lane = (int)(threadIdx.x & 63);
  if (wd >= W) return;  // whole waves
  const int l = wd * 64 + lane;
  bool pass = false;
  if (wd * 64 + 63 > i && l > i && l < n) {
    double fa[4], fb[4];
    for (int k = 0; k < 4; ++k) {
      fa[k] = fq[(size_t)l * 4 + k];
      fb[k] = fq[(size_t)i * 4 + k];
    }
    pass = quat_miso_cos(sym, fa, fb) >= cos_min;
  }
  const uint64_t m = __ballot(pass);
  if (lane == 0) bits[(size_t)i * W + wd] = m;
}

}  // namespace

int merge_bits_dev(hipStream_t st, const double *fq, int n, const SymSet &sym, double cos_min,
                   uint64_t *bits_host) {
  if (n <= 0) return LM_OK;
  if (n > 65535) {
    set_error("merge: at most 65535 candidates on the device");
    return LM_ERR_INVALID;
  }
  const int W = (n + 63) / 64;
  const size_t nbits = (size_t)n * W;
  double *d_fq = nullptr;
  uint64_t *d_bits = nullptr;
  hipError_t e = hipMalloc(&d_fq, (size_t)n * 4 * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&d_bits, nbits * 8);
  if (e == hipSuccess)
    e = hipMemcpyAsync(d_fq, fq, (size_t)n * 4 * sizeof(double), hipMemcpyHostToDevice, st);
  if (e == hipSuccess) {
    const dim3 grid((unsigned)((W + kWavesPerBlock - 1) / kWavesPerBlock), (unsigned)n);
    hipLaunchKernelGGL(merge_bits_kernel, grid, dim3(64 * kWavesPerBlock), 0, st, d_fq, n, W,
                       sym, cos_min, d_bits);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(bits_host, d_bits, nbits * 8, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (d_fq) (void)hipFree(d_fq);
  if (d_bits) (void)hipFree(d_bits);
  return e == hipSuccess ? LM_OK : hip_fail(e, "merge pair kernel");
}

}  // namespace lm
