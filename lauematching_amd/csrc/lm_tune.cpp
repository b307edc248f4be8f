// lm_tune.cpp - autotuner for the scoring layouts of one DB.
//
// The compare step can read three layouts of the same spot table:
//   - the orientation-major table: one lane per orientation (up to 128
//     images), B/16 lanes of 16 images (64, 128), or 4-bit bounds of the
//     pixels with an exact recount of the candidates (128, 256),
//   - the single-image banded copy (super-tiles of 2^13 .. 2^16 rows),
//   - the batch banded copy (super-tiles of 2^9..2^13 rows; a batch of width
//     B fits when B * 2^tb <= 32768, its LDS counters filling <= 128 KiB, or
//     B * 2^tb <= 65536 with 16-bit sums and an exact recount).
// Which one is fastest depends on the image density, the table's spots per
// orientation and the batch width, so the tuner times each candidate on the
// caller's sample images (device time of the whole launch sequence, best of
// `reps`) and keeps the winners: the single-image super-tile size and entry
// loads, the batch copy's super-tile size (least per-image time summed over
// the batch widths) and, per batch width, whether it reads the batch copy or
// the orientation-major table, and with which walk.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "lm_internal.h"

namespace lm {
namespace {

int time_batch(lm_db *db, const uint8_t *d_imgs, int B, int reps, double *best_ms) {
  std::vector<double> ones(B, 1.0), thr(B, INFINITY);
  std::vector<int> sp(B, 1 << 20);  // no hits: time the scan, not the collection
  std::vector<lm_hits> outs(B);
  double best = INFINITY;
  for (int r = 0; r < reps; ++r) {
    const int rc = score_images(db, db->scratch, B, d_imgs, ones.data(), sp.data(), thr.data(),
                                db->ctx->stream, outs.data());
    if (rc) return rc;
    best = std::min(best, outs[0].kernel_ms);
    for (auto &h : outs) lm_hits_free(&h);
  }
  *best_ms = best;
  return LM_OK;
}

}  // namespace

int autotune(lm_db *db, const uint8_t *const *imgs, int nimg, int max_batch, int reps,
             lm_tune_report *rep) {
  const uint64_t npx = (uint64_t)db->geom.nrPxX * db->geom.nrPxY;
  memset(rep, 0, sizeof *rep);
  // stage kMaxBatch images: the samples, repeated
  int rc = scratch_images(db->scratch, npx * kMaxBatch);
  if (rc) return rc;
  uint8_t *d = db->scratch.img;
  hipStream_t st = db->ctx->stream;
  for (int i = 0; i < kMaxBatch; ++i) {
    const hipError_t e = hipMemcpyAsync(d + (uint64_t)i * npx, imgs[i % nimg], npx,
                                        hipMemcpyHostToDevice, st);
    if (e != hipSuccess) return hip_fail(e, "autotune staging");
  }
  if (hipStreamSynchronize(st) != hipSuccess) return hip_fail(hipGetLastError(), "autotune");
  const int saved_tb = db->banded_tb, saved_btb = db->batch_tb, saved_loads = db->banded_loads,
            saved_bsplit = db->banded_split;
  const uint32_t saved_mask = db->batch_widths, saved_group = db->group_widths,
                 saved_nib = db->nib_widths, saved_s16 = db->batch16_widths,
                 saved_split = db->split_widths;
  int8_t saved_wtb[9];
  memcpy(saved_wtb, db->width_tb, sizeof saved_wtb);
  // every width that tb's counters hold: 32-bit counters B x 2^tb <= 32768,
  // 16-bit sums B x 2^tb <= 65536
  auto set_all_widths = [&](int tb, int lim = 32768) {
    for (int k = 0; k < 9; ++k)
      db->width_tb[k] = (tb && k >= 2 && k <= 6 && (1 << k) * (1 << tb) <= lim) ? (int8_t)tb : 0;
  };
  auto restore = [&](int code) {
    db->banded_tb = saved_tb;
    db->batch_tb = saved_btb;
    memcpy(db->width_tb, saved_wtb, sizeof saved_wtb);
    db->batch_widths = saved_mask;
    db->banded_loads = saved_loads;
    db->group_widths = saved_group;
    db->nib_widths = saved_nib;
    db->batch16_widths = saved_s16;
    db->split_widths = saved_split;
    db->banded_split = saved_bsplit;
    const int r = build_banded(db);
    return code ? code : r;
  };
  // single image: the table, then 2^13 .. 2^16 super-tiles x entry loads
  db->batch_tb = 0;
  set_all_widths(0);
  db->banded_tb = 0;
  db->banded_split = -1;
  rc = build_banded(db);
  if (rc) return restore(rc);
  rc = time_batch(db, d, 1, reps, &rep->ms_single_table);
  if (rc) return restore(rc);
  double best1 = rep->ms_single_table;
  int best_tb = 0, best_loads = 0, best_split = -1;
  // (2^16 rows keep 16-bit sums: only the LDS-staged walks, only when a
  // row's sum fits).  The LDS walks are timed with every split of the image
  // chunks' entries into 1..8 pieces (ms_single keeps each walk's best);
  // the gather walks read no chunk offsets and are timed once.
  const int max_tb = db->max_spots <= kSum16MaxSpots ? kSum16TileBits : 15;
  double split_ms[4][8][kMaxSplitBits + 1] = {};
  for (int i = 0; 13 + i <= max_tb; ++i) {
    db->banded_tb = 13 + i;
    for (int ks = 0; ks <= (db->banded_tb >= 14 ? kMaxSplitBits : 0); ++ks) {
      db->banded_split = ks;
      rc = build_banded(db);
      if (rc) return restore(rc);
      for (int loads = 0; loads < 8; ++loads) {
        const bool lds_walk = loads == 3 || loads >= 5;
        if (lds_walk && (!db->banded.chunk_off || db->banded_tb == 13)) continue;
        if (!lds_walk && (db->banded_tb == kSum16TileBits || ks > 0)) continue;
        db->banded_loads = loads;
        ++db->generation;
        double ms = 0;
        rc = time_batch(db, d, 1, reps, &ms);
        if (rc) return restore(rc);
        split_ms[i][loads][ks] = ms;
        double &cell = rep->ms_single[i][loads];
        if (cell == 0 || ms < cell) cell = ms;
        // a later layout must be faster by more than the timing noise (1%)
        if (ms < 0.99 * best1) {
          best1 = ms;
          best_tb = 13 + i;
          best_loads = loads;
          best_split = lds_walk ? ks : -1;
        }
      }
    }
  }
  if (best_tb)
    for (int ks = 0; ks <= kMaxSplitBits; ++ks)
      rep->ms_split[ks] = split_ms[best_tb - 13][best_loads][ks];
  db->banded_split = best_split;
  rep->banded_split = best_split;
  // batch widths 4..max_batch
  int nw = 0;
  for (int B = 4; B <= std::min(max_batch, kMaxBatch); B <<= 1) rep->widths[nw++] = B;
  rep->n_widths = nw;
  db->banded_tb = 0;  // room for the batch copies
  db->batch_tb = 0;
  db->batch16_widths = 0;
  db->split_widths = 0;
  set_all_widths(0);
  rc = build_banded(db);
  if (rc) return restore(rc);
  // the table walks: one lane per orientation (up to 128 images), grouped
  // lanes (64, 128), 4-bit bounds (128, 256); tbl[w] keeps the fastest,
  // grouped_mask / nibble_mask record which
  uint32_t gmask = 0, nmask = 0;
  std::vector<double> tbl(nw, INFINITY);
  for (int w = 0; w < nw; ++w) {
    const int B = rep->widths[w];
    const uint32_t bitB = 1u << width_log2(B);
    int pick = -1;  // 0 lane, 1 grouped, 2 nibble
    db->group_widths = 0;
    db->nib_widths = 0;
    ++db->generation;
    if (B <= kMaxByteBatch) {
      rc = time_batch(db, d, B, reps, &rep->ms_lane[w]);
      if (rc) return restore(rc);
      tbl[w] = rep->ms_lane[w];
      pick = 0;
    }
    if (B == 64 || B == 128) {
      db->group_widths = bitB;
      ++db->generation;
      rc = time_batch(db, d, B, reps, &rep->ms_grouped[w]);
      if (rc) return restore(rc);
      db->group_widths = 0;
      if (rep->ms_grouped[w] < tbl[w]) {
        tbl[w] = rep->ms_grouped[w];
        pick = 1;
      }
    }
    if (B >= 128) {
      db->nib_widths = bitB;
      ++db->generation;
      rc = time_batch(db, d, B, reps, &rep->ms_nibble[w]);
      if (rc) return restore(rc);
      db->nib_widths = 0;
      if (rep->ms_nibble[w] < tbl[w]) {
        tbl[w] = rep->ms_nibble[w];
        pick = 2;
      }
    }
    if (pick == 1) gmask |= bitB;
    if (pick == 2) nmask |= bitB;
  }
  db->group_widths = gmask;
  db->nib_widths = nmask;
  ++db->generation;
  db->batch_widths = 0x7f;
  for (int tb = 9; tb <= 13; ++tb) {
    db->batch_tb = tb;
    db->batch16_widths = 0;
    set_all_widths(tb);
    rc = build_banded(db);
    if (rc) return restore(rc);
    for (int w = 0; w < nw; ++w) {
      const int B = rep->widths[w];
      if (!banded_batch_fits(db, B)) continue;
      rc = time_batch(db, d, B, reps, &rep->ms_banded[w][tb - 9]);
      if (rc) return restore(rc);
    }
    // the same copy read with 16-bit sums (twice the rows per width)
    db->batch16_widths = 0x7c;
    set_all_widths(tb, 65536);
    ++db->generation;
    for (int w = 0; w < nw; ++w) {
      const int B = rep->widths[w];
      if (!banded_batch16(db, B)) continue;
      rc = time_batch(db, d, B, reps, &rep->ms_banded16[w][tb - 9]);
      if (rc) return restore(rc);
    }
    db->batch16_widths = 0;
  }
  // per width: the fastest batch copy when it beats the table walks (each
  // width may keep a different super-tile size and counter width; the copies
  // are built once)
  uint32_t mask = 0, s16 = 0;
  int8_t wtb[9] = {};
  std::vector<double> wms(nw, 0.0);
  for (int w = 0; w < nw; ++w) {
    const int k = width_log2(rep->widths[w]);
    double best = tbl[w];
    wms[w] = tbl[w];
    for (int tb = 9; tb <= 13; ++tb)
      for (int v = 0; v < 2; ++v) {
        const double ms = v ? rep->ms_banded16[w][tb - 9] : rep->ms_banded[w][tb - 9];
        if (ms > 0 && ms < best) {
          best = ms;
          wms[w] = ms;
          wtb[k] = (int8_t)tb;
          if (v)
            s16 |= 1u << k;
          else
            s16 &= ~(1u << k);
        }
      }
    rep->batch_tile_bits_w[w] = wtb[k];
    if (wtb[k]) mask |= 1u << k;
  }
  rep->sum16_mask = s16;
  // a width served faster by two calls of half the images (the narrower
  // width's layout staying in L2 where the wider one's does not)
  uint32_t split = 0;
  // width 4 (2..4 images) against four single-image calls
  if (nw > 0 && rep->widths[0] == 4 && 4.0 * best1 < wms[0]) {
    wms[0] = 4.0 * best1;
    split |= 1u << 2;
  }
  for (int w = 1; w < nw; ++w) {
    const int B = rep->widths[w];
    if (B < 8 || B > 64 || rep->widths[w - 1] * 2 != B) continue;
    if (2.0 * wms[w - 1] < wms[w]) {
      wms[w] = 2.0 * wms[w - 1];
      split |= 1u << width_log2(B);
    }
  }
  rep->split_mask = split;
  int best_btb = 0;  // reported: the copy of the narrowest width that keeps one
  for (int w = nw - 1; w >= 0; --w)
    if (wtb[width_log2(rep->widths[w])]) best_btb = wtb[width_log2(rep->widths[w])];
  rep->grouped_mask = gmask;
  rep->nibble_mask = nmask;
  rep->banded_tile_bits = best_tb;
  rep->banded_loads = best_loads;
  rep->batch_tile_bits = best_btb;
  rep->batch_widths_mask = mask;
  db->banded_tb = best_tb;
  db->banded_loads = best_loads;
  db->batch_tb = best_btb;
  memcpy(db->width_tb, wtb, sizeof wtb);
  db->batch_widths = mask;
  db->batch16_widths = s16;
  db->split_widths = split;
  // cost table for the server's batch-size choice
  for (double &v : db->width_ms) v = 0;
  db->width_ms[0] = best1;
  for (int w = 0; w < nw; ++w) db->width_ms[width_log2(rep->widths[w])] = wms[w];
  rc = build_banded(db);  // also bumps the generation: no stale captured launch
  if (rc) return restore(rc);
  return LM_OK;
}

}  // namespace lm
