// lm_api.cpp - the extern "C" entry points of liblauematch.so.
//
// Host-side pieces that must be bit-identical to the reference host code are
// computed here in plain double arithmetic in the reference's association
// order: the detector rotation (src/LaueMatchingCPU.c:224-239), the
// reciprocal basis (src/LaueMatchingHeaders.h:379-436) and the uint8
// quantization (src/LaueMatchingCPU.c:355-377).

#include <hip/hip_runtime.h>
#include <immintrin.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>
#include <algorithm>

#include "lm_internal.h"

namespace lm {

static thread_local std::string g_err;

void set_error(const std::string &msg) { g_err = msg; }

// set by pool_get when a buffer would pass the context's HBM budget: the
// hipErrorOutOfMemory it returns is then reported as LM_ERR_BUDGET
thread_local bool g_budget_fail = false;

int hip_fail(hipError_t e, const char *what) {
  if (e == hipErrorOutOfMemory && g_budget_fail) {
    g_budget_fail = false;
    g_err = std::string(what) + ": HBM budget exceeded by working buffers";
    return LM_ERR_BUDGET;
  }
  g_err = std::string(what) + ": " + hipGetErrorString(e);
  return LM_ERR_HIP;
}

int pix_bits_for(uint64_t npx) {
  int b = 0;
  while ((1ull << b) < npx) ++b;
  return b;
}

namespace {

constexpr double kDeg2Rad = 0.0174532925199433;  // src/LaueMatchingHeaders.h:31
constexpr double kEps = 1E-12;                   // src/LaueMatchingHeaders.h:42

double zero_small(double v) { return std::fabs(v) < kEps ? 0 : v; }

// Reciprocal basis: real-space basis vectors, then b_i x b_j * 2pi/V
// (rhombohedral setting for space groups 146/148/155/160/161/166/167).
void recip_basis(const double lat[6], int sgnum, double out[9]) {
  const double a = lat[0], b = lat[1], c = lat[2];
  const bool rh = sgnum == 146 || sgnum == 148 || sgnum == 155 || sgnum == 160 || sgnum == 161 ||
                  sgnum == 166 || sgnum == 167;
  const double ca = std::cos(lat[3] * kDeg2Rad), cb = std::cos(lat[4] * kDeg2Rad);
  const double cg = std::cos(lat[5] * kDeg2Rad), sg = std::sin(lat[5] * kDeg2Rad);
  const double phi = std::sqrt(1.0 - ca * ca - cb * cb - cg * cg + 2 * ca * cb * cg);
  const double vol = a * b * c * phi;
  const double pv = (2 * M_PI) / vol;
  double A[3], B[3], C[3];
  if (!rh) {
    A[0] = zero_small(a);
    A[1] = 0.0;
    A[2] = 0.0;
    B[0] = b * cg;
    B[1] = zero_small(b * sg);
    B[2] = 0;
    C[0] = c * cb;
    C[1] = c * (ca - cb * cg) / sg;
    C[2] = zero_small(c * phi / sg);
  } else {
    const double p = std::sqrt(1.0 + 2 * ca), q = std::sqrt(1.0 - ca);
    const double u = (a / 3.0) * (p - q), w = (a / 3.0) * (p + 2 * q);
    A[0] = w; A[1] = u; A[2] = u;
    B[0] = u; B[1] = w; B[2] = u;
    C[0] = u; C[1] = u; C[2] = w;
  }
  out[0] = zero_small((B[1] * C[2] - B[2] * C[1]) * pv);
  out[3] = zero_small((B[2] * C[0] - B[0] * C[2]) * pv);
  out[6] = zero_small((B[0] * C[1] - B[1] * C[0]) * pv);
  out[1] = zero_small((C[1] * A[2] - C[2] * A[1]) * pv);
  out[4] = zero_small((C[2] * A[0] - C[0] * A[2]) * pv);
  out[7] = zero_small((C[0] * A[1] - C[1] * A[0]) * pv);
  out[2] = zero_small((A[1] * B[2] - A[2] * B[1]) * pv);
  out[5] = zero_small((A[2] * B[0] - A[0] * B[2]) * pv);
  out[8] = zero_small((A[0] * B[1] - A[1] * B[0]) * pv);
}

// Rodrigues rotation for the rotation vector r, returned transposed.
void detector_rotT(const double r[3], double rotT[9]) {
  const double ang = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
  const double v0 = r[0] / ang, v1 = r[1] / ang, v2 = r[2] / ang;
  const double c = std::cos(ang), s = std::sin(ang), omc = 1 - std::cos(ang);
  double R[9];
  R[0] = c + omc * (v0 * v0);
  R[1] = omc * v0 * v1 - s * v2;
  R[2] = omc * v0 * v2 + s * v1;
  R[3] = omc * v1 * v0 + s * v2;
  R[4] = c + omc * (v1 * v1);
  R[5] = omc * v1 * v2 - s * v0;
  R[6] = omc * v2 * v0 - s * v1;
  R[7] = omc * v2 * v1 + s * v0;
  R[8] = c + omc * (v2 * v2);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) rotT[i * 3 + j] = R[j * 3 + i];
}

int check_ctx(lm_ctx *ctx) {
  if (!ctx) {
    set_error("null context");
    return LM_ERR_INVALID;
  }
  hipError_t e = hipSetDevice(ctx->device);
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  return LM_OK;
}

int stage_host_images(lm_db *db, int nimg, const uint8_t *const *imgs, uint8_t **dptr) {
  const uint64_t npx = (uint64_t)db->geom.nrPxX * db->geom.nrPxY;
  int rc = scratch_images(db->scratch, npx * (uint64_t)nimg);
  if (rc) return rc;
  for (int b = 0; b < nimg; ++b) {
    if (!imgs[b]) {
      set_error("null image pointer");
      return LM_ERR_INVALID;
    }
    hipError_t e = hipMemcpyAsync(db->scratch.img + (uint64_t)b * npx, imgs[b], npx,
                                  hipMemcpyHostToDevice, db->ctx->stream);
    if (e != hipSuccess) return hip_fail(e, "image upload");
  }
  *dptr = db->scratch.img;
  return LM_OK;
}

// Direct scoring call on the DB's own buffers with one threshold pair.
int score_uniform(lm_db *db, int nimg, const uint8_t *d, const double *scales, int min_sp,
                  double min_int, hipStream_t st, lm_hits *outs) {
  int sp[kMaxBatch];
  double mi[kMaxBatch];
  for (int b = 0; b < kMaxBatch; ++b) {
    sp[b] = min_sp;
    mi[b] = min_int;
  }
  return score_images(db, db->scratch, nimg, d, scales, sp, mi, st, outs);
}

// The uint8 quantization (src/LaueMatchingCPU.c:355-377): mx = the largest
// pixel (strict >, from 0), then x > 0 -> v = x * (255 / mx) clamped to
// [1, 255] and truncated, else 0.  The scalar loops are the definition.  The
// AVX2 loops compute the same IEEE operations four pixels at a time: the
// maximum as (x > m) ? x : m (what vmaxps/vmaxpd compute, NaN skipped like
// the scalar test), v = x * inv, min(v, 255) then max(., 1), truncation, and
// a mask for x > 0.  They are taken only when mx is finite: then v is finite
// wherever x > 0 and the two agree on every input, NaN, signed zeros and
// subnormals included (with mx = +inf the scalar code truncates a NaN).
template <class T>
double max_scalar(const T *p, uint64_t n) {
  double mx = 0;
  for (uint64_t i = 0; i < n; ++i)
    if ((double)p[i] > mx) mx = p[i];
  return mx;
}

template <class T>
void map_scalar(const T *p, uint64_t n, double inv, uint8_t *out) {
  for (uint64_t i = 0; i < n; ++i) {
    const double x = p[i];
    if (x > 0) {
      const double v = x * inv;
      out[i] = (v >= 255.0) ? 255 : (v < 1.0) ? 1 : (uint8_t)v;
    } else {
      out[i] = 0;
    }
  }
}

__attribute__((target("avx2"))) double max_avx2(const float *p, uint64_t n) {
  __m256 m0 = _mm256_setzero_ps(), m1 = m0, m2 = m0, m3 = m0;
  uint64_t i = 0;
  for (; i + 32 <= n; i += 32) {
    m0 = _mm256_max_ps(_mm256_loadu_ps(p + i), m0);
    m1 = _mm256_max_ps(_mm256_loadu_ps(p + i + 8), m1);
    m2 = _mm256_max_ps(_mm256_loadu_ps(p + i + 16), m2);
    m3 = _mm256_max_ps(_mm256_loadu_ps(p + i + 24), m3);
  }
  alignas(32) float lane[32];
  _mm256_store_ps(lane, m0);
  _mm256_store_ps(lane + 8, m1);
  _mm256_store_ps(lane + 16, m2);
  _mm256_store_ps(lane + 24, m3);
  double mx = max_scalar(lane, 32);
  const double t = max_scalar(p + i, n - i);
  return t > mx ? t : mx;
}

__attribute__((target("avx2"))) double max_avx2(const double *p, uint64_t n) {
  __m256d m0 = _mm256_setzero_pd(), m1 = m0, m2 = m0, m3 = m0;
  uint64_t i = 0;
  for (; i + 16 <= n; i += 16) {
    m0 = _mm256_max_pd(_mm256_loadu_pd(p + i), m0);
    m1 = _mm256_max_pd(_mm256_loadu_pd(p + i + 4), m1);
    m2 = _mm256_max_pd(_mm256_loadu_pd(p + i + 8), m2);
    m3 = _mm256_max_pd(_mm256_loadu_pd(p + i + 12), m3);
  }
  alignas(32) double lane[16];
  _mm256_store_pd(lane, m0);
  _mm256_store_pd(lane + 4, m1);
  _mm256_store_pd(lane + 8, m2);
  _mm256_store_pd(lane + 12, m3);
  double mx = max_scalar(lane, 16);
  const double t = max_scalar(p + i, n - i);
  return t > mx ? t : mx;
}

// four pixels (as doubles) -> four int32 levels
__attribute__((target("avx2"))) inline __m128i level4(__m256d x, __m256d inv) {
  __m256d v = _mm256_mul_pd(x, inv);
  v = _mm256_max_pd(_mm256_min_pd(v, _mm256_set1_pd(255.0)), _mm256_set1_pd(1.0));
  v = _mm256_and_pd(v, _mm256_cmp_pd(x, _mm256_setzero_pd(), _CMP_GT_OQ));
  return _mm256_cvttpd_epi32(v);
}

__attribute__((target("avx2"))) inline void store16(uint8_t *o, __m128i a, __m128i b, __m128i c,
                                                   __m128i d) {
  const __m128i w = _mm_packus_epi16(_mm_packus_epi32(a, b), _mm_packus_epi32(c, d));
  _mm_storeu_si128((__m128i *)o, w);
}

__attribute__((target("avx2"))) void map_avx2(const float *p, uint64_t n, double inv,
                                              uint8_t *out) {
  const __m256d vi = _mm256_set1_pd(inv);
  uint64_t i = 0;
  for (; i + 16 <= n; i += 16) {
    const __m256 a = _mm256_loadu_ps(p + i), b = _mm256_loadu_ps(p + i + 8);
    store16(out + i, level4(_mm256_cvtps_pd(_mm256_castps256_ps128(a)), vi),
            level4(_mm256_cvtps_pd(_mm256_extractf128_ps(a, 1)), vi),
            level4(_mm256_cvtps_pd(_mm256_castps256_ps128(b)), vi),
            level4(_mm256_cvtps_pd(_mm256_extractf128_ps(b, 1)), vi));
  }
  map_scalar(p + i, n - i, inv, out + i);
}

__attribute__((target("avx2"))) void map_avx2(const double *p, uint64_t n, double inv,
                                              uint8_t *out) {
  const __m256d vi = _mm256_set1_pd(inv);
  uint64_t i = 0;
  for (; i + 16 <= n; i += 16)
    store16(out + i, level4(_mm256_loadu_pd(p + i), vi), level4(_mm256_loadu_pd(p + i + 4), vi),
            level4(_mm256_loadu_pd(p + i + 8), vi), level4(_mm256_loadu_pd(p + i + 12), vi));
  map_scalar(p + i, n - i, inv, out + i);
}

bool have_avx2() {
  static const bool yes = __builtin_cpu_supports("avx2");
  return yes;
}

template <class T>
void quantize(const T *img, uint64_t npx, uint8_t *out, double *im_scale) {
  const bool vec = have_avx2();
  const double mx = vec ? max_avx2(img, npx) : max_scalar(img, npx);
  const double scale = (mx > 0) ? mx / 255.0 : 1.0;
  const double inv = (mx > 0) ? 255.0 / mx : 0.0;
  if (vec && mx < HUGE_VAL)
    map_avx2(img, npx, inv, out);
  else
    map_scalar(img, npx, inv, out);
  if (im_scale) *im_scale = scale;
}

}  // namespace
}  // namespace lm

using namespace lm;

namespace lm {

hipError_t pool_get(lm_ctx *ctx, DevPool &p, size_t slot, uint64_t bytes, void **out) {
  if (slot >= p.ptr.size()) {
    p.ptr.resize(slot + 1, nullptr);
    p.cap.resize(slot + 1, 0);
  }
  bytes = std::max<uint64_t>((bytes + 255) & ~(uint64_t)255, 256);
  if (p.cap[slot] < bytes) {
    if (ctx->in_use - p.cap[slot] + bytes > ctx->budget) {
      set_error("HBM budget exceeded by working buffers");
      g_budget_fail = true;
      return hipErrorOutOfMemory;
    }
    if (p.ptr[slot]) (void)hipFree(p.ptr[slot]);  // the previous call has finished with it
    ctx->in_use -= p.cap[slot];
    p.ptr[slot] = nullptr;
    p.cap[slot] = 0;
    const hipError_t e = hipMalloc(&p.ptr[slot], bytes);
    if (e != hipSuccess) {
      p.ptr[slot] = nullptr;
      return e;
    }
    p.cap[slot] = bytes;
    ctx->in_use += bytes;
  }
  *out = p.ptr[slot];
  return hipSuccess;
}

void ctx_ref(lm_ctx *ctx) { ctx->refs.fetch_add(1); }

void ctx_unref(lm_ctx *ctx) {
  if (ctx->refs.fetch_sub(1) != 1) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  pool_release(ctx, ctx->prep_pool);
  pool_release(ctx, ctx->median_pool);
  pool_release(ctx, ctx->io_pool);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

void pool_release(lm_ctx *ctx, DevPool &p) {
  for (size_t i = 0; i < p.ptr.size(); ++i) {
    if (p.ptr[i]) (void)hipFree(p.ptr[i]);
    ctx->in_use -= p.cap[i];
  }
  p = DevPool{};
}

}  // namespace lm

extern "C" {

const char *lm_last_error(void) { return g_err.c_str(); }

const char *lm_version(void) { return "lauematch-mi355x 0.1 (gfx950)"; }

int lm_ctx_create(int device, uint64_t budget, lm_ctx **out) {
  if (!out) {
    set_error("null output pointer");
    return LM_ERR_INVALID;
  }
  *out = nullptr;
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDeviceCount");
  if (device < 0 || device >= ndev) {
    set_error("device index out of range");
    return LM_ERR_INVALID;
  }
  e = hipSetDevice(device);
  if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
  size_t freeb = 0, totalb = 0;
  e = hipMemGetInfo(&freeb, &totalb);
  if (e != hipSuccess) return hip_fail(e, "hipMemGetInfo");
  // address of this device's copy of the no-allocation scoring scratch
  int rc = noalloc_init(device);
  if (rc) return rc;
  lm_ctx *c = new (std::nothrow) lm_ctx;
  if (!c) {
    set_error("out of host memory");
    return LM_ERR_NOMEM;
  }
  c->device = device;
  c->budget = budget ? budget : (uint64_t)(0.9 * (double)freeb);
  e = hipStreamCreate(&c->stream);
  if (e != hipSuccess) {
    delete c;
    return hip_fail(e, "hipStreamCreate");
  }
  *out = c;
  return LM_OK;
}

void lm_ctx_destroy(lm_ctx *ctx) {
  if (ctx) ctx_unref(ctx);  // DBs and servers still made on it keep it until they go
}

uint64_t lm_ctx_bytes_in_use(const lm_ctx *ctx) { return ctx ? ctx->in_use : 0; }

int lm_geom_init(const double lattice[6], int32_t space_group, const double P[3],
                 const double R[3], double pxX, double pxY, int32_t nrPxX, int32_t nrPxY,
                 double Elo, double Ehi, lm_geom *g) {
  if (!lattice || !P || !R || !g) {
    set_error("lm_geom_init: null argument");
    return LM_ERR_INVALID;
  }
  if (nrPxX <= 0 || nrPxY <= 0 || nrPxX > 65535 || nrPxY > 65535 || lattice[0] == 0.0) {
    set_error("lm_geom_init: detector size must be 1..65535 and lattice a nonzero");
    return LM_ERR_INVALID;
  }
  double rb[9], rt[9];
  recip_basis(lattice, space_group, rb);
  detector_rotT(R, rt);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      g->recip[i][j] = rb[i * 3 + j];
      g->rotT[i][j] = rt[i * 3 + j];
    }
  for (int i = 0; i < 3; ++i) g->P[i] = P[i];
  g->pxX = pxX;
  g->pxY = pxY;
  g->nrPxX = nrPxX;
  g->nrPxY = nrPxY;
  g->Elo = Elo;
  g->Ehi = Ehi;
  return LM_OK;
}

int lm_quantize_u8(const double *img, uint64_t npx, uint8_t *out, double *im_scale) {
  if ((!img || !out) && npx) {
    set_error("lm_quantize_u8: null buffer");
    return LM_ERR_INVALID;
  }
  quantize(img, npx, out, im_scale);
  return LM_OK;
}

int lm_quantize_u8_f32(const float *img, uint64_t npx, uint8_t *out, double *im_scale) {
  if ((!img || !out) && npx) {
    set_error("lm_quantize_u8_f32: null buffer");
    return LM_ERR_INVALID;
  }
  // every float is a double exactly: the same operations as lm_quantize_u8
  // on the widened pixels, without the float64 copy
  quantize(img, npx, out, im_scale);
  return LM_OK;
}

int lm_db_create(lm_ctx *ctx, const double *orients, uint64_t n, const int32_t *hkls,
                 int32_t nhkls, const lm_geom *geom, int32_t max_spots, lm_db **out) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  if (!out || !geom || (n && !orients) || nhkls < 0 || (nhkls && !hkls)) {
    set_error("lm_db_create: bad arguments");
    return LM_ERR_INVALID;
  }
  if (max_spots < 1 || max_spots > kMaxSpots) {
    set_error("lm_db_create: MaxNrLaueSpots must be 1..511");
    return LM_ERR_INVALID;
  }
  if (n >= (1ull << 32)) {
    set_error("lm_db_create: at most 2^32-1 orientations per DB");
    return LM_ERR_INVALID;
  }
  if (geom->nrPxX <= 0 || geom->nrPxY <= 0 || geom->nrPxX > 65535 || geom->nrPxY > 65535 ||
      (uint64_t)geom->nrPxX * geom->nrPxY > (1ull << 31)) {
    set_error("lm_db_create: detector size out of range");
    return LM_ERR_INVALID;
  }
  *out = nullptr;
  const uint64_t bytes = n * 72 + (uint64_t)nhkls * 12 + 64;
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (ctx->in_use + bytes > ctx->budget) {
      set_error("lm_db_create: HBM budget exceeded");
      return LM_ERR_BUDGET;
    }
    ctx->in_use += bytes;
  }
  lm_db *db = new (std::nothrow) lm_db;
  if (!db) {
    set_error("out of host memory");
    return LM_ERR_NOMEM;
  }
  db->ctx = ctx;
  db->n = n;
  db->nhkls = nhkls;
  db->max_spots = max_spots;
  // single-image layout: the image staged through LDS over 2^16-row
  // super-tiles with 16-bit sums (2^15 rows and 32-bit counters when
  // 255 * MaxNrLaueSpots does not fit 16 bits)
  db->banded_tb = max_spots <= kSum16MaxSpots ? kSum16TileBits : 15;
  db->banded_loads = 3;
  db->scratch.ctx = ctx;  // direct calls' buffers count against the budget
  db->hkls.assign(hkls, hkls + (size_t)nhkls * 3);
  Geom &g = db->geom;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      g.recip[i * 3 + j] = geom->recip[i][j];
      g.rotT[i * 3 + j] = geom->rotT[i][j];
    }
  for (int i = 0; i < 3; ++i) g.P[i] = geom->P[i];
  g.pxX = geom->pxX;
  g.pxY = geom->pxY;
  g.nrPxX = geom->nrPxX;
  g.nrPxY = geom->nrPxY;
  g.Elo = geom->Elo;
  g.Ehi = geom->Ehi;
  hipError_t e = hipMalloc(&db->d_orients, n * 72 + 8);
  if (e == hipSuccess) e = hipMalloc(&db->d_hkls, (uint64_t)nhkls * 12 + 16);
  if (e == hipSuccess && n)
    e = hipMemcpyAsync(db->d_orients, orients, n * 72, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess && nhkls)
    e = hipMemcpyAsync(db->d_hkls, hkls, (uint64_t)nhkls * 12, hipMemcpyHostToDevice,
                       ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    rc = hip_fail(e, "lm_db_create upload");
    {
      std::lock_guard<std::mutex> lk(ctx->mu);
      ctx->in_use -= bytes;
    }
    if (db->d_orients) (void)hipFree(db->d_orients);
    if (db->d_hkls) (void)hipFree(db->d_hkls);
    delete db;
    return rc;
  }
  ctx_ref(ctx);
  *out = db;
  return LM_OK;
}

void lm_db_destroy(lm_db *db) {
  if (!db) return;
  (void)hipSetDevice(db->ctx->device);
  (void)hipStreamSynchronize(db->ctx->stream);
  release_table(db);
  release_host_table(db);
  if (db->d_fwd_plan) (void)hipFree(db->d_fwd_plan);
  if (db->d_orients) (void)hipFree(db->d_orients);
  if (db->d_hkls) (void)hipFree(db->d_hkls);
  scratch_release(db->scratch);
  {
    std::lock_guard<std::mutex> lk(db->ctx->mu);
    db->ctx->in_use -= db->n * 72 + (uint64_t)db->nhkls * 12 + 64;
  }
  lm_ctx *ctx = db->ctx;
  delete db;
  ctx_unref(ctx);
}

uint64_t lm_db_size(const lm_db *db) { return db ? db->n : 0; }
uint64_t lm_db_table_bytes(const lm_db *db) { return db && db->table.ready ? db->table.bytes : 0; }
uint64_t lm_db_table_spots(const lm_db *db) {
  if (!db) return 0;
  return db->table.ready ? db->table.total : (db->htable.ready ? db->htable.total : 0);
}

int lm_db_forward(lm_db *db, int32_t mode, lm_fwd_stats *stats) {
  if (!db) {
    set_error("null db");
    return LM_ERR_INVALID;
  }
  int rc = check_ctx(db->ctx);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(db->mu);
  release_host_table(db);
  rc = forward_build(db, mode, stats);
  // a table larger than the HBM budget stays on the host, scored in chunks
  if (rc == LM_ERR_BUDGET) rc = forward_host(db, mode, stats);
  return rc;
}

int lm_db_set_option(lm_db *db, int32_t option, int64_t value) {
  if (!db) {
    set_error("null db");
    return LM_ERR_INVALID;
  }
  int rc = check_ctx(db->ctx);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(db->mu);
  if (option == LM_OPT_BANDED_TILE_BITS) {
    if (value != 0 && (value < 13 || value > kSum16TileBits)) {
      set_error("LM_OPT_BANDED_TILE_BITS must be 0 or 13..16");
      return LM_ERR_INVALID;
    }
    if (value == kSum16TileBits && db->max_spots > kSum16MaxSpots) {
      set_error("LM_OPT_BANDED_TILE_BITS 16 keeps 16-bit sums: MaxNrLaueSpots must be <= 257");
      return LM_ERR_INVALID;
    }
    db->banded_tb = (int)value;
    if (db->table.ready) return build_banded(db);
    return LM_OK;
  }
  if (option == LM_OPT_BATCH_TILE_BITS) {
    if (value != 0 && (value < 9 || value > 13)) {
      set_error("LM_OPT_BATCH_TILE_BITS must be 0 or 9..13");
      return LM_ERR_INVALID;
    }
    db->batch_tb = (int)value;
    // every width the copy's LDS counters hold (B x 2^bits <= 32768) takes it
    for (int k = 0; k < 9; ++k)
      db->width_tb[k] = (value && k >= 2 && k <= 6 && (1ll << k) * (1ll << value) <= 32768)
                            ? (int8_t)value : (int8_t)0;
    if (db->table.ready) return build_banded(db);
    return LM_OK;
  }
  if (option == LM_OPT_BATCH_WIDTH_TILE_BITS) {
    const int64_t k = value >> 8, tb = value & 0xff;
    const int64_t lim = (k >= 2 && k <= 6 && ((db->batch16_widths >> k) & 1)) ? 65536 : 32768;
    if (k < 2 || k > 6 || (tb != 0 && (tb < 9 || tb > 13 || (1ll << k) * (1ll << tb) > lim))) {
      set_error("LM_OPT_BATCH_WIDTH_TILE_BITS: (log2 width 2..6) << 8 | bits (0, or 9..13 "
                "with width x 2^bits <= 32768, or <= 65536 with 16-bit sums)");
      return LM_ERR_INVALID;
    }
    db->width_tb[k] = (int8_t)tb;
    if (db->table.ready) return build_banded(db);
    return LM_OK;
  }
  if (option == LM_OPT_BATCH_WIDTHS) {
    if (value < 0 || value > 0x7f) {
      set_error("LM_OPT_BATCH_WIDTHS is a mask of bits 0..6");
      return LM_ERR_INVALID;
    }
    db->batch_widths = (uint32_t)value;
    ++db->generation;
    return LM_OK;
  }
  if (option == LM_OPT_BANDED_SPLIT) {
    if (value < -1 || value > kMaxSplitBits) {
      set_error("LM_OPT_BANDED_SPLIT must be -1 (from the mean chunk) or 0..3");
      return LM_ERR_INVALID;
    }
    db->banded_split = (int)value;
    if (db->table.ready) return build_banded(db);
    return LM_OK;
  }
  if (option == LM_OPT_BANDED_LOADS) {
    if (value < 0 || value > 7) {
      set_error("LM_OPT_BANDED_LOADS must be 0..7");
      return LM_ERR_INVALID;
    }
    db->banded_loads = (int)value;
    ++db->generation;  // captured launches name the kernel
    return LM_OK;
  }
  if (option == LM_OPT_USE_GRAPHS) {
    db->scratch.use_graphs = value != 0;
    return LM_OK;
  }
  if (option == LM_OPT_BATCH_GROUPED) {
    if (value < 0 || value > 0xc0 || (value & 0x3f)) {
      set_error("LM_OPT_BATCH_GROUPED is a mask of bits 6 and 7 (64 and 128 images)");
      return LM_ERR_INVALID;
    }
    db->group_widths = (uint32_t)value;
    ++db->generation;
    return LM_OK;
  }
  if (option == LM_OPT_BATCH_SUM16) {
    if (value < 0 || (value & ~(int64_t)0x7c)) {
      set_error("LM_OPT_BATCH_SUM16 is a mask of bits 2..6 (batches of 4..64 images)");
      return LM_ERR_INVALID;
    }
    db->batch16_widths = (uint32_t)value;
    ++db->generation;
    return LM_OK;
  }
  if (option == LM_OPT_BATCH_SPLIT) {
    if (value < 0 || (value & ~(int64_t)0x7c)) {
      set_error("LM_OPT_BATCH_SPLIT is a mask of bits 2..6 (batches of 4..64 images)");
      return LM_ERR_INVALID;
    }
    db->split_widths = (uint32_t)value;
    ++db->generation;
    return LM_OK;
  }
  if (option == LM_OPT_BATCH_NIBBLE) {
    if (value < 0 || value > 0x180 || (value & 0x7f)) {
      set_error("LM_OPT_BATCH_NIBBLE is a mask of bits 7 and 8 (128 and 256 images)");
      return LM_ERR_INVALID;
    }
    db->nib_widths = (uint32_t)value;
    ++db->generation;
    return LM_OK;
  }
  set_error("unknown option");
  return LM_ERR_INVALID;
}

int lm_db_get_option(lm_db *db, int32_t option, int64_t *value) {
  if (!db || !value) {
    set_error("lm_db_get_option: bad arguments");
    return LM_ERR_INVALID;
  }
  std::lock_guard<std::mutex> lk(db->mu);
  switch (option) {
    case LM_OPT_BANDED_TILE_BITS: *value = db->banded_tb; return LM_OK;
    case LM_OPT_BATCH_TILE_BITS: *value = db->batch_tb; return LM_OK;
    case LM_OPT_BATCH_WIDTHS: *value = db->batch_widths; return LM_OK;
    case LM_OPT_USE_GRAPHS: *value = db->scratch.use_graphs ? 1 : 0; return LM_OK;
    case LM_OPT_BANDED_LOADS: *value = db->banded_loads; return LM_OK;
    case LM_OPT_BANDED_SPLIT:  // the copy's split when it has one, else the setting
      *value = db->banded.ready && db->banded.chunk_off ? db->banded.ks : db->banded_split;
      return LM_OK;
    case LM_OPT_TABLE_ON_HOST: *value = db->htable.ready ? 1 : 0; return LM_OK;
    case LM_OPT_BATCH_GROUPED: *value = db->group_widths; return LM_OK;
    case LM_OPT_BATCH_NIBBLE: *value = db->nib_widths; return LM_OK;
    case LM_OPT_BATCH_SUM16: *value = db->batch16_widths; return LM_OK;
    case LM_OPT_BATCH_SPLIT: *value = db->split_widths; return LM_OK;
    default: set_error("unknown option"); return LM_ERR_INVALID;
  }
}

int lm_db_autotune(lm_db *db, const uint8_t *const *images, int32_t nimg, int32_t max_batch,
                   int32_t reps, lm_tune_report *out) {
  if (!db || !images || nimg < 1 || !out || max_batch < 1 || reps < 1) {
    set_error("lm_db_autotune: bad arguments");
    return LM_ERR_INVALID;
  }
  for (int i = 0; i < nimg; ++i)
    if (!images[i]) {
      set_error("lm_db_autotune: null image");
      return LM_ERR_INVALID;
    }
  int rc = check_ctx(db->ctx);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(db->mu);
  if (!db->table.ready) {
    set_error(db->htable.ready
                  ? "lm_db_autotune: the table is kept on the host (larger than the HBM budget)"
                  : "lm_db_autotune: no table (run lm_db_forward or lm_db_import_table first)");
    return LM_ERR_STATE;
  }
  return autotune(db, images, nimg, max_batch, reps, out);
}

int lm_db_table_checksum(lm_db *db, uint64_t *sum) {
  if (!db || !sum) {
    set_error("lm_db_table_checksum: bad arguments");
    return LM_ERR_INVALID;
  }
  int rc = check_ctx(db->ctx);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(db->mu);
  if (db->htable.ready) {
    *sum = checksum_host(db);
    return LM_OK;
  }
  return table_checksum(db, sum);
}

int lm_db_import_table(lm_db *db, const uint16_t *records, uint64_t count) {
  if (!db) {
    set_error("null db");
    return LM_ERR_INVALID;
  }
  int rc = check_ctx(db->ctx);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(db->mu);
  release_host_table(db);
  rc = import_table(db, records, count);
  if (rc == LM_ERR_BUDGET) rc = import_host(db, records, count);
  return rc;
}

int lm_db_export_table(lm_db *db, uint64_t first, uint64_t count, uint16_t *out) {
  if (!db) {
    set_error("null db");
    return LM_ERR_INVALID;
  }
  int rc = check_ctx(db->ctx);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(db->mu);
  if (db->htable.ready) {
    if (first > db->n || count > db->n - first || (count && !out)) {
      set_error("export_table: rows out of range");
      return LM_ERR_INVALID;
    }
    return export_host(db, first, count, out);
  }
  return export_table(db, first, count, out);
}

int lm_score(lm_db *db, const uint8_t *img_u8, double im_scale, int32_t min_nr_spots,
             double min_intensity, lm_hits *out) {
  return lm_score_batch(db, 1, &img_u8, &im_scale, min_nr_spots, min_intensity, out);
}

int lm_score_dev(lm_db *db, const uint8_t *d_img_u8, double im_scale, int32_t min_nr_spots,
                 double min_intensity, void *stream, lm_hits *out) {
  return lm_score_batch_dev(db, 1, d_img_u8, &im_scale, min_nr_spots, min_intensity, stream,
                            out);
}

int lm_score_dev_noalloc(lm_db *db, const uint8_t *d_img_u8, double im_scale,
                         int32_t min_nr_spots, double min_intensity, void *stream, uint64_t cap,
                         uint64_t *ids, int32_t *nspots, double *totint, double *score,
                         uint64_t *count, uint64_t *nr_results) {
  if (!db) {
    set_error("lm_score_dev_noalloc: null db");
    return LM_ERR_INVALID;
  }
  int rc = check_ctx(db->ctx);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(db->mu);
  if (db->htable.ready) {
    set_error("lm_score_dev_noalloc: the table is kept on the host (larger than the HBM "
              "budget); use lm_score_dev");
    return LM_ERR_STATE;
  }
  hipStream_t st = stream ? (hipStream_t)stream : db->ctx->stream;
  return score_noalloc(db, d_img_u8, im_scale, min_nr_spots, min_intensity, st, cap, ids, nspots,
                       totint, score, count, nr_results);
}

int lm_score_batch(lm_db *db, int32_t nimg, const uint8_t *const *imgs_u8,
                   const double *im_scales, int32_t min_nr_spots, double min_intensity,
                   lm_hits *outs) {
  if (!db || !imgs_u8 || nimg < 1) {
    set_error("lm_score_batch: bad arguments");
    return LM_ERR_INVALID;
  }
  int rc = check_ctx(db->ctx);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(db->mu);
  if (nimg > kMaxBatch) {
    set_error("lm_score_batch: at most 256 images per call");
    return LM_ERR_INVALID;
  }
  uint8_t *d = nullptr;
  rc = stage_host_images(db, nimg, imgs_u8, &d);
  if (rc) return rc;
  return score_uniform(db, nimg, d, im_scales, min_nr_spots, min_intensity, db->ctx->stream,
                       outs);
}

int lm_score_batch_each(lm_db *db, int32_t nimg, const uint8_t *const *imgs_u8,
                        const double *im_scales, const int32_t *min_nr_spots,
                        const double *min_intensity, lm_hits *outs) {
  if (!db || !imgs_u8 || !im_scales || !min_nr_spots || !min_intensity || !outs || nimg < 1) {
    set_error("lm_score_batch_each: bad arguments");
    return LM_ERR_INVALID;
  }
  int rc = check_ctx(db->ctx);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(db->mu);
  if (nimg > kMaxBatch) {
    set_error("lm_score_batch_each: at most 256 images per call");
    return LM_ERR_INVALID;
  }
  uint8_t *d = nullptr;
  rc = stage_host_images(db, nimg, imgs_u8, &d);
  if (rc) return rc;
  int sp[kMaxBatch];
  double mi[kMaxBatch];
  for (int b = 0; b < kMaxBatch; ++b) {
    sp[b] = b < nimg ? min_nr_spots[b] : 0;
    mi[b] = b < nimg ? min_intensity[b] : 0.0;
  }
  return score_images(db, db->scratch, nimg, d, im_scales, sp, mi, db->ctx->stream, outs);
}

int lm_score_f64(lm_db *db, const double *img_f64, int32_t min_nr_spots, double min_intensity,
                 lm_hits *out) {
  if (!db || !img_f64 || !out) {
    set_error("lm_score_f64: bad arguments");
    return LM_ERR_INVALID;
  }
  int rc = check_ctx(db->ctx);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(db->mu);
  const uint64_t bytes = (uint64_t)db->geom.nrPxX * db->geom.nrPxY * 8;
  rc = scratch_images(db->scratch, bytes);
  if (rc) return rc;
  hipError_t e = hipMemcpyAsync(db->scratch.img, img_f64, bytes, hipMemcpyHostToDevice,
                                db->ctx->stream);
  if (e != hipSuccess) return hip_fail(e, "image upload");
  return score_seq(db, db->scratch, db->scratch.img, true, 1.0, min_nr_spots, min_intensity,
                   db->ctx->stream, out);
}

int lm_score_batch_dev(lm_db *db, int32_t nimg, const uint8_t *d_imgs, const double *im_scales,
                       int32_t min_nr_spots, double min_intensity, void *stream, lm_hits *outs) {
  if (!db || !d_imgs || nimg < 1) {
    set_error("lm_score_batch_dev: bad arguments");
    return LM_ERR_INVALID;
  }
  int rc = check_ctx(db->ctx);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(db->mu);
  hipStream_t st = stream ? (hipStream_t)stream : db->ctx->stream;
  return score_uniform(db, nimg, d_imgs, im_scales, min_nr_spots, min_intensity, st, outs);
}

int lm_score_fwd_batch_dev(lm_db *db, int32_t nimg, const uint8_t *d_imgs,
                           const double *im_scales, const int32_t *min_nr_spots,
                           const double *min_intensity, int32_t mode, void *stream,
                           lm_hits *outs, lm_fwd_stats *stats) {
  if (!db || !d_imgs || nimg < 1 || nimg > kMaxBatch || !im_scales || !min_nr_spots ||
      !min_intensity || !outs) {
    set_error("lm_score_fwd_batch_dev: bad arguments");
    return LM_ERR_INVALID;
  }
  int rc = check_ctx(db->ctx);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(db->mu);
  hipStream_t st = stream ? (hipStream_t)stream : db->ctx->stream;
  std::vector<int> sp(min_nr_spots, min_nr_spots + nimg);
  return score_images_fwd(db, db->scratch, nimg, d_imgs, im_scales, sp.data(), min_intensity,
                          mode, st, outs, stats);
}

int lm_selftest_fp64(lm_ctx *ctx, const double *a, const double *b, uint64_t n, double *q,
                     double *s) {
  int rc = check_ctx(ctx);
  if (rc) return rc;
  if (n && (!a || !b || !q || !s)) {
    set_error("lm_selftest_fp64: null buffer");
    return LM_ERR_INVALID;
  }
  return selftest_fp64(ctx->stream, a, b, n, q, s);
}

int lm_merge_candidates(const double *orients, uint64_t n_orients, const uint64_t *rows,
                        const double *scores, int32_t nr, double max_angle, int32_t space_group,
                        int32_t threads, uint64_t *best_row, double *best_score,
                        int32_t *cluster_size, int32_t *n_clusters) {
  if (nr < 0 || (nr && (!orients || !rows || !scores || !best_row || !best_score ||
                        !cluster_size)) || !n_clusters) {
    set_error("lm_merge_candidates: bad arguments");
    return LM_ERR_INVALID;
  }
  std::vector<int> sizes((size_t)std::min(nr, 50000));
  const int k = merge_candidates(orients, n_orients, rows, scores, nr, max_angle, space_group,
                                 threads, best_row, best_score, sizes.data());
  if (k < 0) return LM_ERR_INVALID;
  for (int i = 0; i < k; ++i) cluster_size[i] = sizes[i];
  *n_clusters = k;
  return LM_OK;
}

int lm_merge_candidates_dev(lm_ctx *ctx, const double *orients, uint64_t n_orients,
                            const uint64_t *rows, const double *scores, int32_t nr,
                            double max_angle, int32_t space_group, uint64_t *best_row,
                            double *best_score, int32_t *cluster_size, int32_t *n_clusters) {
  if (!ctx || nr < 0 ||
      (nr && (!orients || !rows || !scores || !best_row || !best_score || !cluster_size)) ||
      !n_clusters) {
    set_error("lm_merge_candidates_dev: bad arguments");
    return LM_ERR_INVALID;
  }
  int rc = check_ctx(ctx);
  if (rc) return rc;
  std::vector<int> sizes((size_t)std::min(nr, 50000));
  const int k = merge_candidates_dev(ctx->stream, orients, n_orients, rows, scores, nr,
                                     max_angle, space_group, best_row, best_score, sizes.data());
  if (k < 0) return k;  // an LM_ERR_* code
  for (int i = 0; i < k; ++i) cluster_size[i] = sizes[i];
  *n_clusters = k;
  return LM_OK;
}

int lm_preprocess(lm_ctx *ctx, int32_t nr_px_x, int32_t nr_px_y, const double *raw,
                  const double *background, const lm_prep_config *cfg, uint8_t *u8_host,
                  void *d_u8, double *blurred, lm_prep_result *res) {
  if (!ctx || !raw || !cfg || !res || nr_px_x < 1 || nr_px_y < 1 ||
      (uint64_t)nr_px_x * (uint64_t)nr_px_y >= (1ull << 31)) {
    set_error("lm_preprocess: bad arguments");
    return LM_ERR_INVALID;
  }
  if (cfg->threshold_method < LM_THRESH_ADAPTIVE || cfg->threshold_method > LM_THRESH_OTSU ||
      (cfg->threshold_method == LM_THRESH_PERCENTILE &&
       !(cfg->threshold_percentile >= 0 && cfg->threshold_percentile <= 100))) {
    set_error("lm_preprocess: threshold method or percentile out of range");
    return LM_ERR_INVALID;
  }
  if (!background && cfg->filter_radius > 0 && cfg->median_passes > 0 &&
      (cfg->filter_radius > kMaxMedianSize ||
       (uint64_t)nr_px_x * (uint64_t)nr_px_y > (1ull << 24))) {
    set_error("lm_preprocess: median background needs filter_radius <= 255 and at most 2^24 "
              "pixels");
    return LM_ERR_INVALID;
  }
  int rc = check_ctx(ctx);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(ctx->mu);
  const int r = preprocess(ctx, nr_px_x, nr_px_y, raw, background, *cfg, u8_host,
                           (uint8_t *)d_u8, blurred, *res);
  if (r == LM_ERR_BUDGET) {  // a frame too large for the budget keeps nothing
    (void)hipStreamSynchronize(ctx->stream);
    pool_release(ctx, ctx->prep_pool);
    pool_release(ctx, ctx->median_pool);
  }
  return r;
}

int lm_median_background(lm_ctx *ctx, int32_t nr_px_x, int32_t nr_px_y, const double *image,
                         int32_t filter_radius, int32_t median_passes, double *background,
                         double *ms) {
  if (!ctx || !image || !background || nr_px_x < 1 || nr_px_y < 1 ||
      (uint64_t)nr_px_x * (uint64_t)nr_px_y > (1ull << 24) || filter_radius > kMaxMedianSize) {
    set_error("lm_median_background: bad arguments (at most 2^24 pixels, filter_radius <= 255)");
    return LM_ERR_INVALID;
  }
  int rc = check_ctx(ctx);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(ctx->mu);
  const uint64_t npx = (uint64_t)nr_px_x * (uint64_t)nr_px_y;
  hipStream_t st = ctx->stream;
  double *d_img = nullptr, *d_bg = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
  hipError_t e = pool_get(ctx, ctx->io_pool, 0, npx * 8, (void **)&d_img);
  if (e == hipSuccess) e = pool_get(ctx, ctx->io_pool, 1, npx * 8, (void **)&d_bg);
  if (e == hipSuccess) e = hipEventCreate(&ev[0]);
  if (e == hipSuccess) e = hipEventCreate(&ev[1]);
  if (e == hipSuccess) e = hipMemcpyAsync(d_img, image, npx * 8, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipEventRecord(ev[0], st);
  rc = e == hipSuccess ? LM_OK : hip_fail(e, "lm_median_background");
  if (rc == LM_OK)
    rc = median_background_dev(ctx, st, nr_px_x, nr_px_y, d_img, filter_radius, median_passes,
                               d_bg);
  if (rc == LM_OK) {
    e = hipEventRecord(ev[1], st);
    if (e == hipSuccess)
      e = hipMemcpyAsync(background, d_bg, npx * 8, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    float t = 0;
    if (e == hipSuccess && ms) e = hipEventElapsedTime(&t, ev[0], ev[1]);
    if (ms) *ms = t;
    if (e != hipSuccess) rc = hip_fail(e, "lm_median_background");
  }
  for (hipEvent_t x : ev)
    if (x) (void)hipEventDestroy(x);
  if (rc == LM_ERR_BUDGET) {  // a frame too large for the budget keeps nothing
    (void)hipStreamSynchronize(st);
    pool_release(ctx, ctx->io_pool);
    pool_release(ctx, ctx->median_pool);
  }
  return rc;
}

void lm_quat_from_matrix(const double m[9], double q[4]) { quat_from_matrix(m, q); }

double lm_misorientation_deg(const double q1[4], const double q2[4], int32_t space_group) {
  return misorientation_deg(q1, q2, space_group);
}

void lm_hits_free(lm_hits *h) {
  if (!h) return;
  free(h->ids);
  free(h->nspots);
  free(h->totint);
  free(h->score);
  h->ids = nullptr;
  h->nspots = nullptr;
  h->totint = nullptr;
  h->score = nullptr;
  h->count = 0;
}

}  // extern "C"
