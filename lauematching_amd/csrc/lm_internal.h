// lm_internal.h - shared definitions of the HIP library (not part of the ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "lauematch.h"

namespace lm {

// Spot table resident in HBM.
//
// Orientations are grouped in tiles of T = 2^tile_bits consecutive rows.  A
// tile's spots are stored contiguously as 32-bit entries
//     entry = (row_in_tile << pix_bits) | (y * nrPxX + x)
// in table order (row-major within the tile, spot order within a row), so a
// single streaming read of the entry array feeds the compare kernel and the
// owner of each entry is known without a search.  tile_off[t] is the index
// of the first entry of tile t (tile_off[ntiles] = total).  counts[o] is the
// per-orientation spot count (the ForwardFile's word 0).
struct Table {
  uint32_t *entries = nullptr;  // total + 64 words of slack for 16-byte loads
  uint64_t *tile_off = nullptr; // ntiles + 1
  uint16_t *counts = nullptr;   // n
  uint64_t total = 0;
  uint64_t ntiles = 0;
  uint64_t bytes = 0;
  int tile_bits = 0;
  int pix_bits = 0;
  bool ready = false;
};

// Pixel-banded copy of the table used by the single-image compare kernel.
//
// Orientations are grouped in super-tiles of 2^tb rows.  A super-tile's spots
// (the same spots as its rows in Table, same total) are reordered by pixel
// bucket (2^kBucketBits consecutive pixels per bucket), so consecutive entries
// hit neighbouring pixels, and encoded as
//     entry = (row_in_super_tile << pb) | (pixel & (2^pb - 1)),  pb = 32 - tb
// band j of a super-tile holds the pixels [j << pb, (j+1) << pb);
// band_off[t * (nbands+1) + j] is the offset of band j from the super-tile's
// first entry, which is Table::tile_off at the super-tile's first 256-row tile.
struct Banded {
  uint32_t *entries = nullptr;   // same length as Table::entries
  uint32_t *band_off = nullptr;  // nsuper * (nbands + 1)
  uint64_t nsuper = 0;
  int tb = 0;
  int pb = 0;
  int nbands = 0;
  uint64_t bytes = 0;
  bool ready = false;
};
constexpr int kBucketBits = 10;

struct Geom {
  double recip[9];
  double rotT[9];
  double P[3];
  double pxX, pxY;
  int nrPxX, nrPxY;
  double Elo, Ehi;
};

}  // namespace lm

struct lm_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  uint64_t budget = 0;
  uint64_t in_use = 0;
  std::mutex mu;
};

struct lm_db {
  lm_ctx *ctx = nullptr;
  uint64_t n = 0;
  int nhkls = 0;
  int max_spots = 0;
  lm::Geom geom{};
  double *d_orients = nullptr;  // n * 9
  int32_t *d_hkls = nullptr;    // nhkls * 3
  lm::Table table;
  lm::Banded banded;
  int banded_tb = 14;             // super-tile bits for the banded copy (0 = no copy)
  // scratch for scoring
  uint64_t *d_hits = nullptr;     // packed hit records
  uint64_t hit_cap = 0;
  uint32_t *d_counters = nullptr; // [0] = hits written, [1] = nr_results, [2..] per-image
  uint64_t counter_slots = 0;
  uint8_t *d_batch = nullptr;     // interleaved image batch [npx][B]
  uint64_t batch_cap = 0;
  uint8_t *d_img = nullptr;       // staging for host images
  uint64_t img_cap = 0;
  double *d_vals = nullptr;       // float64 sums of in-order hits
  uint64_t val_cap = 0;
  std::mutex mu;
};

namespace lm {

// Bits of the packed per-orientation accumulator: (nspots << kSumBits) | sum.
constexpr int kSumBits = 17;
constexpr uint32_t kSumMask = (1u << kSumBits) - 1u;
constexpr int kMaxSpots = 511;

void set_error(const std::string &msg);
int hip_fail(hipError_t e, const char *what);

int pix_bits_for(uint64_t npx);
// Orientations per table tile = 1 << kTileBits (one scoring workgroup).
constexpr int kTileBits = 8;

// lm_fwd.hip
int forward_build(lm_db *db, int mode, lm_fwd_stats *stats);
int import_table(lm_db *db, const uint16_t *records, uint64_t count);
int export_table(lm_db *db, uint64_t first, uint64_t count, uint16_t *out);
void release_table(lm_db *db);
int selftest_fp64(hipStream_t st, const double *a, const double *b, uint64_t n, double *q,
                  double *s);

// lm_score.hip
int score_images(lm_db *db, int nimg, const uint8_t *d_imgs, const double *im_scales,
                 int min_nr_spots, double min_intensity, hipStream_t stream, lm_hits *outs);
// Build / drop the pixel-banded copy (after forward or import).
int build_banded(lm_db *db);
void release_banded(lm_db *db);
int launch_banded(lm_db *db, const uint8_t *img, int min_spots, double min_int, uint64_t cap,
                  hipStream_t st);
// One image against the banded copy; imScale must be 1 (integer sums).
int score_banded(lm_db *db, const uint8_t *d_img, int min_nr_spots, double min_intensity,
                 hipStream_t stream, lm_hits *out);
// lm_merge.cpp
void quat_from_matrix(const double m[9], double q[4]);
double misorientation_deg(const double a[4], const double b[4], int space_group);
int merge_candidates(const double *orients, uint64_t n_orients, const uint64_t *rows,
                     const double *scores, int nr, double max_angle, int space_group,
                     int threads, uint64_t *best_row, double *best_score, int *cluster_size);
// One image, float64 sums accumulated in table order on the device
// (f64 = image of doubles; otherwise uint8 values times scale).
int score_seq(lm_db *db, const void *d_img, bool f64, double scale, int min_nr_spots,
              double min_intensity, hipStream_t stream, lm_hits *out);

}  // namespace lm
