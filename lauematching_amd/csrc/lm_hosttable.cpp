// lm_hosttable.cpp - spot tables larger than the HBM budget.
//
// The reference GPU program copies the table to the device in chunks of at
// most 4 GB sized from the free device memory and scores each chunk
// (src/LaueMatchingGPU.cu:660-749).  Here, when the context's budget cannot
// hold the resident table, the table is kept in host memory in its compact
// form (counts, per-tile offsets, pixel entries) and every scoring call
// streams it through one device staging buffer, chunk by chunk, running the
// ordinary scoring kernels on each chunk as if it were a DB of its own rows.
// Chunks are whole 256-row tiles in row order, so the concatenated hit lists
// are ascending and equal to the resident table's.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <vector>

#include "lm_internal.h"

namespace lm {
namespace {

constexpr uint64_t kTile = 1ull << kTileBits;
// the reference's chunk bound (src/LaueMatchingGPU.cu:664-671)
constexpr uint64_t kMaxChunkBytes = 4ull << 30;

int charge(lm_ctx *ctx, uint64_t bytes) {
  std::lock_guard<std::mutex> lk(ctx->mu);
  if (ctx->in_use + bytes > ctx->budget) {
    set_error("HBM budget exceeded by the table staging buffer");
    return LM_ERR_BUDGET;
  }
  ctx->in_use += bytes;
  return LM_OK;
}

void uncharge(lm_ctx *ctx, uint64_t bytes) {
  std::lock_guard<std::mutex> lk(ctx->mu);
  ctx->in_use -= bytes;
}

uint64_t available(const lm_ctx *ctx) {
  lm_ctx *c = const_cast<lm_ctx *>(ctx);
  std::lock_guard<std::mutex> lk(c->mu);
  return c->budget > c->in_use ? c->budget - c->in_use : 0;
}

// A DB over rows [r0, r0 + rows) of db: same geometry, hkls, plan and
// layout options; no banded copies; its own (empty) table.  Shares device
// buffers of db, so it is never passed to lm_db_destroy.
std::unique_ptr<lm_db> view_of(lm_db *db, uint64_t r0, uint64_t rows) {
  std::unique_ptr<lm_db> v(new lm_db);
  v->ctx = db->ctx;
  v->n = rows;
  v->nhkls = db->nhkls;
  v->max_spots = db->max_spots;
  v->geom = db->geom;
  v->d_orients = db->d_orients + r0 * 9;
  v->d_hkls = db->d_hkls;
  v->d_fwd_plan = db->d_fwd_plan;
  v->fwd_plan_entries = db->fwd_plan_entries;
  v->fwd_plan_chunks = db->fwd_plan_chunks;
  v->banded_tb = 0;
  v->batch_tb = 0;
  v->batch_widths = db->batch_widths;
  v->group_widths = db->group_widths;
  v->nib_widths = db->nib_widths;
  return v;
}

void free_stage(lm_db *db) {
  HostTable &h = db->htable;
  if (h.d_counts) (void)hipFree(h.d_counts);
  if (h.d_tile_off) (void)hipFree(h.d_tile_off);
  if (h.d_entries) (void)hipFree(h.d_entries);
  h.d_counts = nullptr;
  h.d_tile_off = nullptr;
  h.d_entries = nullptr;
  if (h.bytes) uncharge(db->ctx, h.bytes);
  h.bytes = 0;
}

// Staging for one chunk: up to 4 GB, half of the budget left, and at least
// one full tile of M spots per row.
int stage_init(lm_db *db) {
  HostTable &h = db->htable;
  const uint64_t ntiles = h.tile_off.size() - 1;
  const uint64_t min_entries = kTile * (uint64_t)db->max_spots + 64;
  uint64_t target = std::min<uint64_t>(kMaxChunkBytes, available(db->ctx) / 2);
  uint64_t cap_entries = std::max<uint64_t>(target / 4, min_entries);
  cap_entries = std::min<uint64_t>(cap_entries, h.total + 64);
  cap_entries = std::max<uint64_t>(cap_entries, std::min<uint64_t>(min_entries, h.total + 64));
  uint64_t cap_tiles = std::min<uint64_t>(ntiles, std::max<uint64_t>(cap_entries / 64, 1));
  const uint64_t bytes = cap_entries * 4 + cap_tiles * kTile * 2 + (cap_tiles + 1) * 8 + 256;
  int rc = charge(db->ctx, bytes);
  if (rc) return rc;
  h.bytes = bytes;
  hipError_t e = hipMalloc(&h.d_entries, cap_entries * 4);
  if (e == hipSuccess) e = hipMalloc(&h.d_counts, cap_tiles * kTile * 2 + 128);
  if (e == hipSuccess) e = hipMalloc(&h.d_tile_off, (cap_tiles + 1) * 8);
  if (e != hipSuccess) {
    free_stage(db);
    return hip_fail(e, "hipMalloc(table staging)");
  }
  h.cap_entries = cap_entries - 64;  // 64 words of slack after the last entry
  h.cap_tiles = cap_tiles;
  return LM_OK;
}

// Next chunk starting at tile t0: whole tiles while they fit the staging.
uint64_t chunk_tiles(const lm_db *db, uint64_t t0) {
  const HostTable &h = db->htable;
  const uint64_t ntiles = h.tile_off.size() - 1;
  uint64_t t1 = t0;
  while (t1 < ntiles && t1 - t0 < h.cap_tiles && h.tile_off[t1 + 1] - h.tile_off[t0] <= h.cap_entries)
    ++t1;
  return t1;
}

// Uploads tiles [t0, t1) into the staging and points the view's table at it.
int stage_chunk(lm_db *db, lm_db *v, uint64_t t0, uint64_t t1, hipStream_t st) {
  HostTable &h = db->htable;
  const uint64_t r0 = t0 * kTile, r1 = std::min<uint64_t>(t1 * kTile, db->n);
  const uint64_t e0 = h.tile_off[t0], e1 = h.tile_off[t1];
  std::vector<uint64_t> off(t1 - t0 + 1);
  for (uint64_t t = t0; t <= t1; ++t) off[t - t0] = h.tile_off[t] - e0;
  hipError_t e = hipMemcpyAsync(h.d_counts, h.counts.data() + r0, (r1 - r0) * 2,
                                hipMemcpyHostToDevice, st);
  if (e == hipSuccess)
    e = hipMemcpyAsync(h.d_tile_off, off.data(), off.size() * 8, hipMemcpyHostToDevice, st);
  if (e == hipSuccess && e1 > e0)
    e = hipMemcpyAsync(h.d_entries, h.entries.data() + e0, (e1 - e0) * 4, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemsetAsync(h.d_entries + (e1 - e0), 0, 64 * 4, st);
  if (e == hipSuccess) e = hipMemsetAsync(h.d_counts + (r1 - r0), 0, 64 * 2, st);
  if (e != hipSuccess) return hip_fail(e, "table chunk upload");
  Table &t = v->table;
  t.entries = h.d_entries;
  t.tile_off = h.d_tile_off;
  t.counts = h.d_counts;
  t.total = e1 - e0;
  t.ntiles = t1 - t0;
  t.tile_bits = kTileBits;
  t.pix_bits = pix_bits_for((uint64_t)db->geom.nrPxX * db->geom.nrPxY);
  t.bytes = 0;  // the staging is charged to the parent DB
  t.ready = true;
  v->n = r1 - r0;
  v->d_orients = db->d_orients + r0 * 9;
  ++v->generation;
  return LM_OK;
}

// Appends a chunk's hits (rows offset by r0) to the accumulated result.
int append_hits(lm_hits &acc, lm_hits &part, uint64_t r0) {
  const uint64_t m = acc.count + part.count;
  if (part.count) {
    uint64_t *ids = (uint64_t *)realloc(acc.ids, m * 8);
    if (ids) acc.ids = ids;
    int32_t *ns = (int32_t *)realloc(acc.nspots, m * 4);
    if (ns) acc.nspots = ns;
    double *ti = (double *)realloc(acc.totint, m * 8);
    if (ti) acc.totint = ti;
    double *sc = (double *)realloc(acc.score, m * 8);
    if (sc) acc.score = sc;
    if (!ids || !ns || !ti || !sc) {
      set_error("out of host memory for hits");
      return LM_ERR_NOMEM;
    }
    for (uint64_t i = 0; i < part.count; ++i) {
      acc.ids[acc.count + i] = part.ids[i] + r0;
      acc.nspots[acc.count + i] = part.nspots[i];
      acc.totint[acc.count + i] = part.totint[i];
      acc.score[acc.count + i] = part.score[i];
    }
  }
  acc.count = m;
  acc.nr_results += part.nr_results;
  acc.kernel_ms += part.kernel_ms;
  acc.walk_ms += part.walk_ms;
  lm_hits_free(&part);
  return LM_OK;
}

}  // namespace

void release_host_table(lm_db *db) {
  free_stage(db);
  db->htable = HostTable{};
}

int forward_host(lm_db *db, int mode, lm_fwd_stats *stats) {
  release_host_table(db);
  HostTable &h = db->htable;
  const uint64_t n = db->n;
  const uint64_t ntiles = (n + kTile - 1) / kTile;
  h.counts.assign(n + 64, 0);
  h.tile_off.assign(ntiles + 1, 0);
  // rows per chunk: the chunk's table, kept lists and index in half of what
  // the budget has left (at M words per row each at most), whole tiles
  const uint64_t per_row = 2 + 8 * (uint64_t)db->max_spots + 8;
  uint64_t rows = std::max<uint64_t>(kTile, available(db->ctx) / 2 / per_row / kTile * kTile);
  hipStream_t st = db->ctx->stream;
  int prc = ensure_fwd_plan(db, st);  // the chunk views share the DB's plan
  if (prc) return prc;
  lm_fwd_stats tot{};
  tot.n_orients = n;
  tot.n_pairs = n * (uint64_t)db->nhkls;
  for (uint64_t r0 = 0; r0 < n;) {
    const uint64_t cnt = std::min(rows, n - r0);
    auto v = view_of(db, r0, cnt);
    lm_fwd_stats s{};
    int rc = forward_build(v.get(), mode, &s);
    if (rc == LM_ERR_BUDGET && rows > kTile) {
      rows = std::max<uint64_t>(kTile, rows / 2 / kTile * kTile);
      continue;
    }
    if (rc) {
      release_host_table(db);
      return rc;
    }
    const Table &t = v->table;
    const uint64_t t0 = r0 / kTile;
    std::vector<uint64_t> off(t.ntiles + 1);
    h.entries.resize(h.total + t.total);
    hipError_t e = hipMemcpyAsync(h.counts.data() + r0, t.counts, cnt * 2, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess)
      e = hipMemcpyAsync(off.data(), t.tile_off, off.size() * 8, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess && t.total)
      e = hipMemcpyAsync(h.entries.data() + h.total, t.entries, t.total * 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
      release_table(v.get());
      release_host_table(db);
      return hip_fail(e, "table chunk download");
    }
    for (uint64_t k = 0; k <= t.ntiles; ++k) h.tile_off[t0 + k] = h.total + off[k];
    h.total += t.total;
    tot.n_spots += s.n_spots;
    tot.n_recomputed += s.n_recomputed;
    tot.n_projected += s.n_projected;
    tot.ms += s.ms;
    release_table(v.get());
    r0 += cnt;
  }
  int rc = stage_init(db);
  if (rc) {
    release_host_table(db);
    return rc;
  }
  h.ready = true;
  ++db->generation;
  if (stats) *stats = tot;
  return LM_OK;
}

int import_host(lm_db *db, const uint16_t *records, uint64_t count) {
  release_host_table(db);
  HostTable &h = db->htable;
  const int M = db->max_spots, W = db->geom.nrPxX, H = db->geom.nrPxY;
  const uint64_t rec = (uint64_t)(1 + 2 * M);
  const uint64_t ntiles = (count + kTile - 1) / kTile;
  h.counts.assign(count + 64, 0);
  h.tile_off.assign(ntiles + 1, 0);
  uint64_t total = 0;
  for (uint64_t o = 0; o < count; ++o) {
    const uint16_t *r = records + o * rec;
    bool ok = r[0] <= M;
    for (int s = 0; ok && s < r[0]; ++s) ok = r[1 + 2 * s] < W && r[2 + 2 * s] < H;
    if (!ok) {
      release_host_table(db);
      set_error("import_table: a record's count exceeds MaxNrLaueSpots or a pixel lies "
                "outside the detector");
      return LM_ERR_INVALID;
    }
    h.counts[o] = r[0];
    total += r[0];
  }
  h.entries.resize(total);
  uint64_t k = 0;
  for (uint64_t o = 0; o < count; ++o) {
    if (o % kTile == 0) h.tile_off[o / kTile] = k;
    const uint16_t *r = records + o * rec;
    for (int s = 0; s < r[0]; ++s)
      h.entries[k++] = (uint32_t)r[2 + 2 * s] * (uint32_t)W + r[1 + 2 * s];
  }
  h.tile_off[ntiles] = k;
  h.total = k;
  int rc = stage_init(db);
  if (rc) {
    release_host_table(db);
    return rc;
  }
  h.ready = true;
  ++db->generation;
  return LM_OK;
}

int export_host(lm_db *db, uint64_t first, uint64_t count, uint16_t *out) {
  const HostTable &h = db->htable;
  const int M = db->max_spots, W = db->geom.nrPxX;
  const uint64_t rec = (uint64_t)(1 + 2 * M);
  uint64_t k = h.tile_off[first / kTile];
  for (uint64_t r = first / kTile * kTile; r < first; ++r) k += h.counts[r];
  for (uint64_t i = 0; i < count; ++i) {
    uint16_t *d = out + i * rec;
    const int c = h.counts[first + i];
    memset(d, 0, rec * 2);
    d[0] = (uint16_t)c;
    for (int s = 0; s < c; ++s, ++k) {
      d[1 + 2 * s] = (uint16_t)(h.entries[k] % (uint32_t)W);
      d[2 + 2 * s] = (uint16_t)(h.entries[k] / (uint32_t)W);
    }
  }
  return LM_OK;
}

// The position-tagged sum of the device checksum (checksum_kernel), on the host.
uint64_t checksum_host(const lm_db *db) {
  auto mix = [](uint64_t x) {
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
  };
  const HostTable &h = db->htable;
  uint64_t acc = 0;
  for (uint64_t i = 0; i < db->n; ++i) acc += mix((i << 16) ^ h.counts[i]);
  for (uint64_t i = 0; i < h.total; ++i) acc += mix(~(i << 32) ^ h.entries[i]);
  return acc;
}

int score_images_host(lm_db *db, Scratch &sc, int nimg, const uint8_t *d_imgs,
                      const double *im_scales, const int *min_spots, const double *min_int,
                      hipStream_t st, lm_hits *outs) {
  // captured sequences key on the DB object: the chunk view is rebuilt per
  // call, so chunks run eagerly
  const bool graphs = sc.use_graphs;
  sc.use_graphs = false;
  for (int b = 0; b < nimg; ++b) memset(&outs[b], 0, sizeof(lm_hits));
  std::vector<lm_hits> part(nimg);
  auto v = view_of(db, 0, 0);
  int rc = LM_OK;
  const uint64_t ntiles = db->htable.tile_off.size() - 1;
  for (uint64_t t0 = 0; t0 < ntiles && rc == LM_OK;) {
    const uint64_t t1 = chunk_tiles(db, t0);
    if (t1 == t0) {  // (a tile always fits the staging; kept as a guard)
      set_error("table staging smaller than one tile");
      rc = LM_ERR_STATE;
      break;
    }
    rc = stage_chunk(db, v.get(), t0, t1, st);
    if (rc == LM_OK)
      rc = score_images(v.get(), sc, nimg, d_imgs, im_scales, min_spots, min_int, st, part.data());
    for (int b = 0; b < nimg && rc == LM_OK; ++b) rc = append_hits(outs[b], part[b], t0 * kTile);
    t0 = t1;
  }
  (void)hipStreamSynchronize(st);  // the staging is reused by the next call
  sc.use_graphs = graphs;
  if (rc)
    for (int b = 0; b < nimg; ++b) lm_hits_free(&outs[b]);
  return rc;
}

int score_seq_host(lm_db *db, Scratch &sc, const void *d_img, bool f64, double scale,
                   int min_nr_spots, double min_intensity, hipStream_t st, lm_hits *out) {
  memset(out, 0, sizeof(lm_hits));
  lm_hits part{};
  auto v = view_of(db, 0, 0);
  int rc = LM_OK;
  const uint64_t ntiles = db->htable.tile_off.size() - 1;
  for (uint64_t t0 = 0; t0 < ntiles && rc == LM_OK;) {
    const uint64_t t1 = chunk_tiles(db, t0);
    if (t1 == t0) {
      set_error("table staging smaller than one tile");
      rc = LM_ERR_STATE;
      break;
    }
    rc = stage_chunk(db, v.get(), t0, t1, st);
    if (rc == LM_OK)
      rc = score_seq(v.get(), sc, d_img, f64, scale, min_nr_spots, min_intensity, st, &part);
    if (rc == LM_OK) rc = append_hits(*out, part, t0 * kTile);
    t0 = t1;
  }
  (void)hipStreamSynchronize(st);
  if (rc) lm_hits_free(out);
  return rc;
}

}  // namespace lm
