// lm_prep.hip - the image preprocessing before the compare, on the device
// (preprocess_image, laue_index/preprocess.py:221-337; thresholds
// laue_index/thresholds.py:29-110; the uint8 quantization of
// src/LaueMatchingCPU.c:355-377).
//
// Pipeline (one call, one stream):
//   1. v = max(raw - background, 0)                      (preprocess.py:263);
//      without a background, the median-filter background of lm_median.hip
//   2. threshold t: fixed, numpy-linear percentile of all pixels, or the
//      adaptive noise floor median(nz) + 4 * 1.4826 * MAD(nz), >= 1
//      (thresholds.py:36-47); order statistics come from a radix sort of the
//      float64 bit patterns (all values are >= +0, so the bit order is the
//      value order), then the interpolation is done on the host in numpy's
//      operation order
//   3. u16 = (uint16) v where v > t, else 0                   (:283, :295)
//   4. 8-connected components of u16 > 0 by union-find (links go from the
//      larger root index to the smaller with a compare-and-swap, so parents
//      only decrease and concurrent finds terminate); per component the
//      area, bounding-box corner and exact integer sums of w, x*w, y*w
//      (w = u16) with 64-bit atomics                       (:158-195)
//   5. components with area < min_area are zeroed; centres of mass
//      ((sum x w - x0 sum w) / sum w + x0, as ndimage.center_of_mass on the
//      bounding box computes them exactly for integer weights); sigma from
//      the nearest pair of centres                          (:198-218, :323)
//   6. float64 Gaussian blur, radius int(4 sigma + 0.5), scipy's 'reflect'
//      edges, along y then x                                         (:329)
//   7. uint8 quantization against the image maximum (lm_quantize_u8).
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <vector>

#include "lm_internal.h"

namespace lm {
namespace {

constexpr int kThreads = 256;

unsigned blocks_for(uint64_t n) {
  return (unsigned)std::min<uint64_t>((n + kThreads - 1) / kThreads, 1u << 20);
}

// grid of the reductions: grid-stride loops, one atomic per block (one
// address taking an atomic per wave of a 4M-pixel image serializes for
// hundreds of microseconds)
constexpr unsigned kReduceBlocks = 1024;
unsigned reduce_blocks(uint64_t n) { return std::min(blocks_for(n), kReduceBlocks); }

// sum / maximum over a block of kThreads (every thread gets the result)
template <typename T, typename Op>
__device__ T block_reduce(T v, Op op) {
  __shared__ T part[kThreads / 64];
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) v = op(v, __shfl_xor(v, d));
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
  __syncthreads();
  T r = part[0];
#pragma unroll
  for (int w = 1; w < kThreads / 64; ++w) r = op(r, part[w]);
  __syncthreads();
  return r;
}

__device__ __forceinline__ uint64_t dbits(double x) {
  return (uint64_t)__double_as_longlong(x);
}

// 1. background subtraction (+0 for every non-positive difference)
__global__ void bgsub_kernel(const double *__restrict__ raw, const double *__restrict__ bg,
                             uint64_t npx, double *__restrict__ out) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < npx;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const double v = raw[i] - (bg ? bg[i] : 0.0);
    out[i] = v > 0 ? v : 0.0;
  }
}

// 2. sort keys of every pixel: mode 0 its value, 1 its value if nonzero,
// 2 |v - med| if nonzero; zero pixels of modes 1 and 2 get the largest key,
// so that after the sort the nonzero pixels' keys come first; *count gets
// the number of nonzero pixels (one atomic per block)
__global__ void keys_kernel(const double *__restrict__ v, uint64_t npx, int mode, double med,
                            uint64_t *__restrict__ keys, unsigned long long *__restrict__ count) {
  unsigned long long c = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < npx;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const double x = v[i];
    if (mode == 0) {
      keys[i] = dbits(x);
    } else {
      c += x > 0;
      keys[i] = x > 0 ? dbits(mode == 1 ? x : fabs(x - med)) : ~0ull;
    }
  }
  if (mode == 0) return;  // (uniform: no block reduction)
  c = block_reduce(c, [](unsigned long long a, unsigned long long b) { return a + b; });
  if (threadIdx.x == 0 && c) atomicAdd(count, c);
}

// 3. threshold and uint16 cast (C conversion through a 64-bit integer)
// (*above counts the pixels > t, thresholds.py:29-32 keeps exactly those)
__global__ void threshold_kernel(const double *__restrict__ v, uint64_t npx, double t,
                                 uint16_t *__restrict__ u16, unsigned long long *above) {
  unsigned long long c = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < npx;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const double x = v[i];
    c += x > t;
    u16[i] = x > t ? (uint16_t)(int64_t)x : (uint16_t)0;
  }
  c = block_reduce(c, [](unsigned long long a, unsigned long long b) { return a + b; });
  if (threadIdx.x == 0 && c) atomicAdd(above, c);
}

// 4. connected components
__device__ __forceinline__ int cc_find(const int32_t *p, int x) {
  int q = p[x];
  while (q != x) {
    x = q;
    q = p[x];
  }
  return x;
}

__device__ void cc_unite(int32_t *p, int a, int b) {
  for (;;) {
    a = cc_find(p, a);
    b = cc_find(p, b);
    if (a == b) return;
    if (a > b) {
      const int t = a;
      a = b;
      b = t;
    }
    const int old = atomicCAS(&p[b], b, a);
    if (old == b) return;
    b = old;  // b stopped being a root: retry from its new parent
  }
}

__global__ void cc_init_kernel(const uint16_t *__restrict__ u16, uint64_t npx,
                               int32_t *__restrict__ p) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < npx;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = u16[i] ? (int32_t)i : -1;
}

__global__ void cc_merge_kernel(const uint16_t *__restrict__ u16, int W, int H, int32_t *p) {
  const uint64_t npx = (uint64_t)W * H;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < npx;
       i += (uint64_t)gridDim.x * blockDim.x) {
    if (!u16[i]) continue;
    const int x = (int)(i % W), y = (int)(i / W);
    // W, NW, N, NE: every 8-neighbour pair is visited from one side
    if (x > 0 && u16[i - 1]) cc_unite(p, (int)i, (int)i - 1);
    if (y > 0) {
      const int up = (int)i - W;
      if (x > 0 && u16[up - 1]) cc_unite(p, (int)i, up - 1);
      if (u16[up]) cc_unite(p, (int)i, up);
      if (x + 1 < W && u16[up + 1]) cc_unite(p, (int)i, up + 1);
    }
  }
}

__global__ void cc_roots_kernel(const int32_t *__restrict__ p, uint64_t npx, int32_t *__restrict__ cid,
                                unsigned int *__restrict__ ncomp) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < npx;
       i += (uint64_t)gridDim.x * blockDim.x) {
    if (p[i] == (int32_t)i) cid[i] = (int32_t)atomicAdd(ncomp, 1u);
  }
}

struct CompStats {
  unsigned int *area;
  unsigned long long *sw, *sx, *sy;
  int *minx, *miny;
};

__global__ void comp_init_kernel(CompStats c, unsigned int n) {
  const unsigned int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  c.area[i] = 0;
  c.sw[i] = c.sx[i] = c.sy[i] = 0;
  c.minx[i] = INT_MAX;
  c.miny[i] = INT_MAX;
}

__global__ void comp_stats_kernel(const uint16_t *__restrict__ u16, const int32_t *p,
                                  const int32_t *__restrict__ cid, int W, uint64_t npx,
                                  CompStats c) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < npx;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const unsigned long long w = u16[i];
    if (!w) continue;
    const int k = cid[cc_find(p, (int)i)];
    const unsigned long long x = i % W, y = i / W;
    atomicAdd(&c.area[k], 1u);
    atomicAdd(&c.sw[k], w);
    atomicAdd(&c.sx[k], x * w);
    atomicAdd(&c.sy[k], y * w);
    atomicMin(&c.minx[k], (int)x);
    atomicMin(&c.miny[k], (int)y);
  }
}

// 5. keep the pixels of kept components, as float64
__global__ void filter_kernel(const uint16_t *__restrict__ u16, const int32_t *p,
                              const int32_t *__restrict__ cid, const uint8_t *__restrict__ keep,
                              uint64_t npx, double *__restrict__ out) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < npx;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint16_t w = u16[i];
    out[i] = (w && keep[cid[cc_find(p, (int)i)]]) ? (double)w : 0.0;
  }
}

// 6. one pass of the separable blur: out = sum_j w[j] in[reflect(c + j - r)]
// along x (axis 1) or y (axis 0)
__device__ __forceinline__ int reflect_index(int i, int n) {
  const int period = 2 * n;
  int m = i % period;
  if (m < 0) m += period;
  return m < n ? m : period - 1 - m;
}

__global__ void blur_kernel(const double *__restrict__ in, int W, int H, int along_x,
                            const double *__restrict__ w, int r, double *__restrict__ out) {
  const uint64_t npx = (uint64_t)W * H;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < npx;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const int x = (int)(i % W), y = (int)(i / W);
    double acc = 0.0;
    if (along_x) {
      const double *row = in + (uint64_t)y * W;
      for (int j = 0; j <= 2 * r; ++j) acc += w[j] * row[reflect_index(x + j - r, W)];
    } else {
      for (int j = 0; j <= 2 * r; ++j) acc += w[j] * in[(uint64_t)reflect_index(y + j - r, H) * W + x];
    }
    out[i] = acc;
  }
}

// 7. maximum (non-negative doubles: bit order) and uint8 quantization
__global__ void max_kernel(const double *__restrict__ v, uint64_t npx,
                           unsigned long long *__restrict__ mx) {
  unsigned long long m = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < npx;
       i += (uint64_t)gridDim.x * blockDim.x)
    m = std::max<unsigned long long>(m, dbits(v[i]));
  m = block_reduce(m, [](unsigned long long a, unsigned long long b) { return a > b ? a : b; });
  if (threadIdx.x == 0) atomicMax(mx, m);
}

__global__ void quantize_kernel(const double *__restrict__ v, uint64_t npx,
                                const unsigned long long *__restrict__ mxb,
                                uint8_t *__restrict__ out) {
  const double mx = __longlong_as_double((long long)*mxb);
  const double inv = (mx > 0) ? 255.0 / mx : 0.0;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < npx;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const double x = v[i];
    uint8_t q = 0;
    if (x > 0) {
      const double s = x * inv;
      q = (s >= 255.0) ? 255 : (s < 1.0) ? 1 : (uint8_t)s;
    }
    out[i] = q;
  }
}

// numpy's linear interpolation between two order statistics
// (numpy lib _lerp: a + (b - a) t, or b - (b - a)(1 - t) from t >= 0.5)
double lerp_np(double a, double b, double t) {
  const double d = b - a;
  return t >= 0.5 ? b - d * (1 - t) : a + d * t;
}

// Device buffers of one call: the i-th buffer a call asks for is slot i of
// the context's preprocessing pool (kept and reused by the next call).
struct Bufs {
  lm_ctx *ctx;
  size_t next = 0;
  template <typename T>
  hipError_t get(T **p, uint64_t count) {
    void *q = nullptr;
    const hipError_t e =
        pool_get(ctx, ctx->prep_pool, next++, std::max<uint64_t>(count, 1) * sizeof(T), &q);
    *p = (T *)q;
    return e;
  }
};

// Order statistics k0 and k1 (0-based) of the keys, after sorting them on
// the bits in which they differ.
hipError_t order_stats(hipStream_t st, Bufs &B, uint64_t *keys, uint64_t n, uint64_t k0,
                       uint64_t k1, double *v0, double *v1) {
  uint64_t *sorted = nullptr;
  unsigned long long *acc = nullptr;
  hipError_t e = B.get(&sorted, n);
  if (e == hipSuccess) e = B.get(&acc, 1);
  int lo = 0, hi = 64;
  if (e == hipSuccess) e = key_bit_range(st, keys, n, acc, &lo, &hi);
  if (e == hipSuccess && hi == 0) lo = 0, hi = 1;  // all keys equal
  size_t tmp_bytes = 0;
  if (e == hipSuccess)
    e = hipcub::DeviceRadixSort::SortKeys(nullptr, tmp_bytes, keys, sorted, (int)n, lo, hi, st);
  void *tmp = nullptr;
  if (e == hipSuccess) e = B.get((uint8_t **)&tmp, tmp_bytes);
  if (e == hipSuccess)
    e = hipcub::DeviceRadixSort::SortKeys(tmp, tmp_bytes, keys, sorted, (int)n, lo, hi, st);
  uint64_t b0 = 0, b1 = 0;
  if (e == hipSuccess) e = hipMemcpyAsync(&b0, sorted + k0, 8, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipMemcpyAsync(&b1, sorted + k1, 8, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  std::memcpy(v0, &b0, 8);
  std::memcpy(v1, &b1, 8);
  return e;
}

// median of the n smallest of nsort keys (numpy: the mean of the two middle
// values when n is even)
hipError_t median_of(hipStream_t st, Bufs &B, uint64_t *keys, uint64_t nsort, uint64_t n,
                     double *med) {
  double a, b;
  const uint64_t h = n / 2;
  const hipError_t e = order_stats(st, B, keys, nsort, n % 2 ? h : h - 1, h, &a, &b);
  *med = n % 2 ? b : (a + b) / 2.0;
  return e;
}

// Smallest distance between two of the points (sorted by x, then a sweep
// that stops once the x gap alone exceeds the best distance).
double nearest_pair(std::vector<std::pair<double, double>> pts) {
  std::sort(pts.begin(), pts.end());
  double best = INFINITY;
  for (size_t i = 0; i < pts.size(); ++i)
    for (size_t j = i + 1; j < pts.size(); ++j) {
      const double dx = pts[j].first - pts[i].first;
      if (dx > best) break;
      const double dy = pts[j].second - pts[i].second;
      const double d = std::sqrt(dx * dx + dy * dy);
      if (d < best) best = d;
    }
  return best;
}

}  // namespace

int preprocess(lm_ctx *ctx, int W, int H, const double *raw, const double *background,
               const lm_prep_config &cfg, uint8_t *u8_host, uint8_t *d_u8_out, double *blurred,
               lm_prep_result &res) {
  hipStream_t st = ctx->stream;
  const uint64_t npx = (uint64_t)W * H;
  std::memset(&res, 0, sizeof res);
  Bufs B{ctx};
  double *d_raw = nullptr, *d_bg = nullptr, *d_v = nullptr, *d_f = nullptr, *d_t = nullptr;
  uint16_t *d_u16 = nullptr;
  int32_t *d_p = nullptr, *d_cid = nullptr;
  uint64_t *d_keys = nullptr;
  unsigned long long *d_cnt = nullptr;
  uint8_t *d_u8 = d_u8_out;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  const bool median_bg = !background && cfg.filter_radius > 0 && cfg.median_passes > 0;
  hipError_t e = B.get(&d_raw, npx);
  if (e == hipSuccess && (background || median_bg)) e = B.get(&d_bg, npx);
  if (e == hipSuccess) e = B.get(&d_v, npx);
  if (e == hipSuccess) e = B.get(&d_u16, npx);
  if (e == hipSuccess) e = B.get(&d_keys, npx);
  if (e == hipSuccess) e = B.get(&d_cnt, 3);  // [2]: pixels above the threshold
  if (e == hipSuccess && !d_u8) e = B.get(&d_u8, npx);
  if (e == hipSuccess) e = hipEventCreate(&ev0);
  if (e == hipSuccess) e = hipEventCreate(&ev1);
  if (e != hipSuccess) {
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    return hip_fail(e, "preprocess buffers");
  }
  struct EvGuard {
    hipEvent_t a, b;
    ~EvGuard() {
      (void)hipEventDestroy(a);
      (void)hipEventDestroy(b);
    }
  } evg{ev0, ev1};
  const unsigned g = blocks_for(npx);
  e = hipMemcpyAsync(d_raw, raw, npx * 8, hipMemcpyHostToDevice, st);
  if (e == hipSuccess && background) e = hipMemcpyAsync(d_bg, background, npx * 8, hipMemcpyHostToDevice, st);
  if (e != hipSuccess) return hip_fail(e, "preprocess upload");
  (void)hipEventRecord(ev0, st);
  if (median_bg) {  // no background given: compute_background (preprocess.py:255-266)
    const int rc = median_background_dev(ctx, st, W, H, d_raw, cfg.filter_radius,
                                         cfg.median_passes, d_bg);
    if (rc != LM_OK) return rc;
  }
  hipLaunchKernelGGL(bgsub_kernel, dim3(g), dim3(kThreads), 0, st, d_raw, d_bg, npx, d_v);

  // 2. threshold (thresholds.py:36-47, 92-110)
  double t = 0;
  const int method = cfg.threshold_method == LM_THRESH_OTSU ? LM_THRESH_ADAPTIVE
                                                            : cfg.threshold_method;
  if (method == LM_THRESH_FIXED) {
    t = cfg.threshold_value;
  } else if (method == LM_THRESH_PERCENTILE) {
    hipLaunchKernelGGL(keys_kernel, dim3(g), dim3(kThreads), 0, st, d_v, npx, 0, 0.0, d_keys,
                       d_cnt);
    const double q = cfg.threshold_percentile / 100.0;
    const double vi = (double)(npx - 1) * q;
    double prev = std::floor(vi), next = prev + 1;
    if (vi >= (double)(npx - 1)) prev = next = -1;
    if (vi < 0) prev = next = 0;
    const double gamma = vi - prev;
    const uint64_t k0 = prev < 0 ? npx - 1 : (uint64_t)prev;
    const uint64_t k1 = next < 0 ? npx - 1 : std::min<uint64_t>((uint64_t)next, npx - 1);
    double a, b;
    e = order_stats(st, B, d_keys, npx, k0, k1, &a, &b);
    if (e != hipSuccess) return hip_fail(e, "percentile");
    t = lerp_np(a, b, gamma);
  } else if (method == LM_THRESH_ADAPTIVE) {
    e = hipMemsetAsync(d_cnt, 0, 16, st);
    hipLaunchKernelGGL(keys_kernel, dim3(reduce_blocks(npx)), dim3(kThreads), 0, st, d_v, npx, 1,
                       0.0, d_keys, d_cnt);
    unsigned long long nz = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&nz, d_cnt, 8, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(e, "noise floor");
    if (nz == 0) {
      t = 1.0;
    } else {
      double med, m2;
      e = median_of(st, B, d_keys, npx, nz, &med);
      if (e == hipSuccess) e = hipMemsetAsync(d_cnt, 0, 16, st);
      if (e == hipSuccess) {
        hipLaunchKernelGGL(keys_kernel, dim3(reduce_blocks(npx)), dim3(kThreads), 0, st, d_v,
                           npx, 2, med, d_keys, d_cnt);
        e = median_of(st, B, d_keys, npx, nz, &m2);
      }
      if (e != hipSuccess) return hip_fail(e, "noise floor");
      const double mad = 1.4826 * m2;
      t = std::max(med + 4.0 * mad, 1.0);
    }
  } else {
    set_error("preprocess: unknown threshold method");
    return LM_ERR_INVALID;
  }
  res.threshold = t;
  if (e == hipSuccess) e = hipMemsetAsync(d_cnt + 2, 0, 8, st);
  if (e != hipSuccess) return hip_fail(e, "threshold");
  hipLaunchKernelGGL(threshold_kernel, dim3(reduce_blocks(npx)), dim3(kThreads), 0, st, d_v, npx,
                     t, d_u16, d_cnt + 2);

  // 4. components
  e = B.get(&d_p, npx);
  if (e == hipSuccess) e = B.get(&d_cid, npx);
  if (e == hipSuccess) e = hipMemsetAsync(d_cnt, 0, 16, st);
  if (e != hipSuccess) return hip_fail(e, "components");
  hipLaunchKernelGGL(cc_init_kernel, dim3(g), dim3(kThreads), 0, st, d_u16, npx, d_p);
  hipLaunchKernelGGL(cc_merge_kernel, dim3(g), dim3(kThreads), 0, st, d_u16, W, H, d_p);
  hipLaunchKernelGGL(cc_roots_kernel, dim3(g), dim3(kThreads), 0, st, d_p, npx, d_cid,
                     (unsigned int *)d_cnt);
  unsigned int ncomp = 0;
  unsigned long long above = 0;
  e = hipMemcpyAsync(&ncomp, d_cnt, 4, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipMemcpyAsync(&above, d_cnt + 2, 8, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return hip_fail(e, "components");
  res.n_components = (int32_t)ncomp;
  res.n_above = above;
  std::vector<uint8_t> keep(ncomp ? ncomp : 1, 0);
  std::vector<std::pair<double, double>> centres;
  if (ncomp) {
    CompStats c{};
    e = B.get(&c.area, ncomp);
    if (e == hipSuccess) e = B.get(&c.sw, ncomp);
    if (e == hipSuccess) e = B.get(&c.sx, ncomp);
    if (e == hipSuccess) e = B.get(&c.sy, ncomp);
    if (e == hipSuccess) e = B.get(&c.minx, ncomp);
    if (e == hipSuccess) e = B.get(&c.miny, ncomp);
    if (e != hipSuccess) return hip_fail(e, "component stats");
    hipLaunchKernelGGL(comp_init_kernel, dim3((ncomp + 255) / 256), dim3(256), 0, st, c, ncomp);
    hipLaunchKernelGGL(comp_stats_kernel, dim3(g), dim3(kThreads), 0, st, d_u16, d_p, d_cid, W,
                       npx, c);
    std::vector<unsigned int> area(ncomp);
    std::vector<unsigned long long> sw(ncomp), sx(ncomp), sy(ncomp);
    std::vector<int> mnx(ncomp), mny(ncomp);
    e = hipMemcpyAsync(area.data(), c.area, ncomp * 4ull, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(sw.data(), c.sw, ncomp * 8ull, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(sx.data(), c.sx, ncomp * 8ull, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(sy.data(), c.sy, ncomp * 8ull, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(mnx.data(), c.minx, ncomp * 4ull, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(mny.data(), c.miny, ncomp * 4ull, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(e, "component stats");
    for (unsigned int k = 0; k < ncomp; ++k) {
      if ((int64_t)area[k] < (int64_t)cfg.min_area) continue;
      keep[k] = 1;
      // ndimage.center_of_mass on the bounding box: integer sums are exact
      const double x0 = mnx[k], y0 = mny[k];
      const double w = (double)sw[k];
      const double cx = (double)(sx[k] - (unsigned long long)mnx[k] * sw[k]) / w + x0;
      const double cy = (double)(sy[k] - (unsigned long long)mny[k] * sw[k]) / w + y0;
      centres.emplace_back(cx, cy);
    }
  }
  res.n_kept = (int32_t)centres.size();
  e = B.get(&d_f, npx);
  if (e == hipSuccess) e = B.get(&d_t, npx);
  if (e != hipSuccess) return hip_fail(e, "blur buffers");
  if (centres.empty()) {
    // no component (kept): the blurred image is zero (preprocess.py:287-313)
    e = hipMemsetAsync(d_f, 0, npx * 8, st);
    if (e != hipSuccess) return hip_fail(e, "preprocess");
  } else {
    uint8_t *d_keep = nullptr;
    e = B.get(&d_keep, ncomp);
    if (e == hipSuccess) e = hipMemcpyAsync(d_keep, keep.data(), ncomp, hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return hip_fail(e, "component filter");
    hipLaunchKernelGGL(filter_kernel, dim3(g), dim3(kThreads), 0, st, d_u16, d_p, d_cid, d_keep,
                       npx, d_t);
    // sigma (preprocess.py:198-218)
    double sigma;
    if (centres.size() < 2) {
      sigma = 3.0;
    } else {
      const double min_px = nearest_pair(centres);
      double delta;
      if (cfg.px_x > 0 && cfg.distance > 0)
        delta = (cfg.distance * std::tan(cfg.orientation_spacing * (M_PI / 180.0))) / cfg.px_x;
      else
        delta = min_px;
      sigma = delta > 0 ? 0.25 * std::min(min_px, delta) : 0.25 * min_px;
      sigma = std::max(sigma, 1.0);
    }
    res.sigma = sigma;
    // scipy.ndimage.gaussian_filter1d weights: radius int(4 sd + 0.5),
    // exp(-0.5 / sigma^2 * x^2) normalised by their sum
    const int r = (int)(4.0 * sigma + 0.5);
    std::vector<double> wts(2 * r + 1);
    const double s2 = sigma * sigma;
    double sum = 0;
    for (int j = -r; j <= r; ++j) {
      wts[j + r] = std::exp(-0.5 / s2 * (double)(j * j));
      sum += wts[j + r];
    }
    for (double &x : wts) x /= sum;
    double *d_w = nullptr;
    e = B.get(&d_w, wts.size());
    if (e == hipSuccess)
      e = hipMemcpyAsync(d_w, wts.data(), wts.size() * 8, hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return hip_fail(e, "blur weights");
    hipLaunchKernelGGL(blur_kernel, dim3(g), dim3(kThreads), 0, st, d_t, W, H, 0, d_w, r, d_v);
    hipLaunchKernelGGL(blur_kernel, dim3(g), dim3(kThreads), 0, st, d_v, W, H, 1, d_w, r, d_f);
  }
  // 7. quantization
  e = hipMemsetAsync(d_cnt, 0, 16, st);
  if (e != hipSuccess) return hip_fail(e, "quantize");
  hipLaunchKernelGGL(max_kernel, dim3(reduce_blocks(npx)), dim3(kThreads), 0, st, d_f, npx, d_cnt);
  hipLaunchKernelGGL(quantize_kernel, dim3(g), dim3(kThreads), 0, st, d_f, npx, d_cnt, d_u8);
  (void)hipEventRecord(ev1, st);
  unsigned long long mxb = 0;
  e = hipMemcpyAsync(&mxb, d_cnt, 8, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess && u8_host) e = hipMemcpyAsync(u8_host, d_u8, npx, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess && blurred) e = hipMemcpyAsync(blurred, d_f, npx * 8, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return hip_fail(e, "preprocess");
  double mx;
  std::memcpy(&mx, &mxb, 8);
  res.im_scale = mx > 0 ? mx / 255.0 : 1.0;
  float ms = 0;
  (void)hipEventElapsedTime(&ms, ev0, ev1);
  res.ms = ms;
  return LM_OK;
}

}  // namespace lm
