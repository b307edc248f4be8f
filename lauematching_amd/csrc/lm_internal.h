// lm_internal.h - shared definitions of the HIP library (not part of the ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <atomic>
#include <mutex>
#include <string>
#include <vector>

#include "lauematch.h"
#include "lm_quat.h"

namespace lm {

// Spot table resident in HBM.
//
// Orientations are grouped in tiles of T = 2^tile_bits consecutive rows.  A
// tile's spots are stored contiguously as 32-bit pixel indices y*nrPxX + x in
// table order (row after row, spot order within a row).  tile_off[t] is the
// index of the first entry of tile t (tile_off[ntiles] = total); counts[o] is
// the per-orientation spot count (the ForwardFile's word 0).
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
  uint32_t *chunk_off = nullptr; // nsuper * ((nchunks << ks) + 1): offsets of the pieces of
                                 // the 2^wave_chunk_bits(tb)-pixel chunks
  uint64_t nsuper = 0;
  int tb = 0;
  int pb = 0;
  int nbands = 0;
  int nchunks = 0;
  int ks = 0;                    // each chunk's entries split into 2^ks balanced pieces
  uint32_t max_chunk = 0;        // largest number of entries in one piece
  uint64_t bytes = 0;
  bool ready = false;
};
constexpr int kBucketBits = 10;
constexpr int kMaxSplitBits = 3;  // at most 8 pieces per chunk of the LDS walk
// LDS-staged walk: pixels per wave chunk (log2) for a super-tile of 2^tb rows;
// 16 chunk slots and 2^tb row counters fill at most 160 KiB of LDS.
constexpr int wave_chunk_bits(int tb) { return tb >= 15 ? 11 : 12; }

struct Geom {
  double recip[9];
  double rotT[9];
  double P[3];
  double pxX, pxY;
  int nrPxX, nrPxY;
  double Elo, Ehi;
};

constexpr int kMaxBatch = 256;
// Widest batch the uint8 interleaved walks take; wider ones read 4-bit bounds.
constexpr int kMaxByteBatch = 128;

// Per-image predicate of one scoring launch, read by the kernels from device
// memory (so a captured launch sequence stays valid for new thresholds).
struct Pred {
  double min_int[kMaxBatch];
  double cand_thr[kMaxBatch];  // candidate threshold on the integer sum
  double scale[kMaxBatch];     // imScale of image b
  int min_spots[kMaxBatch];
  uint32_t exact[kMaxBatch / 32]; // bit b: image b has imScale == 1
  // The same candidate test in integers, for kernels whose lanes hold
  // different images (per-lane loads of the doubles above would cost a
  // vector load per image and row): for nspots > 0 a row passes exactly when
  // nspots >= pk_thr >> 23 and sum >= pk_thr & 0x7fffff (both clamped so
  // that "never" stays never for nspots <= 255 and sums < 2^23); zero_ok
  // bit b says a row with no spot passes (0 >= min_spots and 0 >= min_int).
  uint32_t pk_thr[kMaxBatch];
  uint32_t zero_ok[kMaxBatch / 32];
  // The 4-bit walk's candidate test: sum of ceil(v / 17) >= nib_sum and
  // nonzero count >= nib_cnt (both <= 0x7fff; 0x7fff = never for images
  // outside the batch).
  uint16_t nib_sum[kMaxBatch];
  uint16_t nib_cnt[kMaxBatch];
  int nimg;
};

// Device buffers of one scoring stream.  A DB owns one for direct calls;
// each serving worker owns its own so batches on different streams never
// share buffers.
struct Scratch {
  uint64_t *hits = nullptr;  // packed hit records, cap_per * nimg
  uint64_t hit_cap = 0;
  uint32_t *counters = nullptr;  // kCounterWords
  uint64_t *cands = nullptr;     // candidate rows of the 16-bit-sum walk
  uint64_t cand_cap = 0;
  uint32_t *chunk_nz = nullptr;  // per image chunk of the LDS walk: any nonzero pixel
  uint64_t chunk_nz_cap = 0;
  double *vals = nullptr;        // float64 sums of in-order hits
  uint64_t val_cap = 0;
  uint8_t *batch = nullptr;      // interleaved images [npx][B]
  uint64_t batch_cap = 0;
  uint8_t *img = nullptr;        // staging of host images
  uint64_t img_cap = 0;
  uint32_t *fwd_spill = nullptr; // on-the-fly scoring: list slots past 32 (words)
  uint64_t fwd_spill_cap = 0;
  Pred *pred = nullptr;          // device copy of the predicate
  Pred *pred_host = nullptr;     // pinned host copy
  // the device copy holds *pred_host, uploaded on pred_stream (a call on the
  // same stream with the same predicate skips the upload)
  bool pred_current = false;
  hipStream_t pred_stream = nullptr;
  // pinned host copies of the counters and of the first hit records of a
  // single-image call, fetched together after the launch sequence
  uint32_t *host_cnt = nullptr;
  uint32_t *host_cnt_dev = nullptr;  // its device address (kernels write results there)
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // around the compare kernel alone (eager single-image banded launches)
  hipEvent_t ew0 = nullptr, ew1 = nullptr;
  bool walk_timed = false;
  // Captured launch sequences (interleave + counter fill + score kernel),
  // keyed by every pointer and size their kernel arguments depend on.
  bool use_graphs = false;
  struct Graph {
    const void *db, *src, *hits;
    uint64_t generation;
    int width;  // batch width B (the images themselves come from the predicate)
    uint64_t cap_per;
    int kind;
    hipGraphExec_t exec;
  };
  std::vector<Graph> graphs;
  uint64_t graph_launches = 0;
  // Device bytes of the buffers above, charged to ctx's HBM budget beyond
  // `prepaid` (bytes the owner already reserved, e.g. a serving worker's
  // share); null ctx = not charged.
  lm_ctx *ctx = nullptr;
  uint64_t bytes = 0;
  uint64_t prepaid = 0;
};

int scratch_init(Scratch &s);
void scratch_release(Scratch &s);
// Worst-case bytes of the scratch buffers of scoring calls on db at widths up
// to max_batch, as first allocated (the hit buffer grows past its initial
// 2^16 records per image only for calls with more hits, charged then).
uint64_t scratch_bytes_for(const lm_db *db, int max_batch);
int scratch_images(Scratch &s, uint64_t bytes);
// chunk flag buffer of the single-image LDS walk (at least n words)
int ensure_chunk_flags(Scratch &s, uint64_t n);

// Spot table kept in host memory when the HBM budget cannot hold it
// (lm_hosttable.cpp): compact counts, per-tile offsets and pixel entries,
// streamed through one device staging buffer per scoring call.
struct HostTable {
  bool ready = false;
  std::vector<uint16_t> counts;    // n + 64
  std::vector<uint64_t> tile_off;  // ntiles + 1, absolute
  std::vector<uint32_t> entries;   // total
  uint64_t total = 0;
  uint16_t *d_counts = nullptr;    // staging of one chunk (charged to the budget)
  uint64_t *d_tile_off = nullptr;
  uint32_t *d_entries = nullptr;
  uint64_t cap_tiles = 0, cap_entries = 0, bytes = 0;
};

// Device buffers a context keeps between calls of one kind (preprocessing,
// median background): slot i holds the i-th buffer such a call asks for and
// grows only when a call needs more, so repeated frames allocate nothing.
// Callers hold the context's mutex; the bytes count against its HBM budget.
struct DevPool {
  std::vector<void *> ptr;
  std::vector<uint64_t> cap;
};

}  // namespace lm

struct lm_ctx {
  int device = 0;
  // the caller's reference plus one per DB and server made on it: the
  // context is released when the last of them is destroyed, in any order
  std::atomic<int> refs{1};
  hipStream_t stream = nullptr;
  uint64_t budget = 0;
  uint64_t in_use = 0;
  std::mutex mu;
  lm::DevPool prep_pool;    // lm_preprocess
  lm::DevPool median_pool;  // the median background's ranks and sort storage
  lm::DevPool io_pool;      // lm_median_background's image and result
};

struct lm_db {
  lm_ctx *ctx = nullptr;
  uint64_t n = 0;
  int nhkls = 0;
  int max_spots = 0;
  lm::Geom geom{};
  double *d_orients = nullptr;  // n * 9
  int32_t *d_hkls = nullptr;    // nhkls * 3
  std::vector<int32_t> hkls;    // host copy (nhkls * 3), for the forward plan
  void *d_fwd_plan = nullptr;   // forward pre-test plan: entries, then chunks
  int fwd_plan_entries = 0, fwd_plan_chunks = 0;
  lm::Table table;
  lm::Banded banded;            // single-image copy
  lm::Banded banded_bk[5];      // batch copies of 2^9..2^13-row super-tiles (B x T LDS counters)
  int banded_tb = 16;           // super-tile bits for the banded copy (0 = no copy)
  int banded_loads = 3;         // single-image walk (LM_OPT_BANDED_LOADS)
  int banded_split = -1;        // log2 pieces per chunk of the LDS walk (-1 = from the mean)
  int batch_tb = 0;             // LM_OPT_BATCH_TILE_BITS: the copy every width takes by default
  int8_t width_tb[9] = {};      // per width 2^k: super-tile bits of its batch copy (0 = none)
  uint32_t batch_widths = 0x7f; // bit log2(B): batches of width B use the batch copy
  double width_ms[9] = {};      // autotuned device ms per launch at width 2^k (0 = unknown)
  uint32_t group_widths = 0xc0; // bit log2(B): table batches of width B use grouped lanes
  uint32_t nib_widths = 0x180;  // bit log2(B): batches of width B (128, 256) use 4-bit bounds
  uint32_t batch16_widths = 0;  // bit log2(B): the batch copy read with 16-bit sums
  uint32_t split_widths = 0;    // bit log2(B): scored as two calls of B/2 images (B = 4: singles)
  uint64_t generation = 0;      // bumped whenever the table or its banded copy changes
  lm::Scratch scratch;          // buffers of direct (non-serving) calls
  lm::HostTable htable;         // the table when it does not fit the budget
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
// Device counters of a scoring call: per image b, hits at [2b] and rows
// passing with no spot at [2b + 1]; candidate rows of the 16-bit-counter walk
// at kCandSlot, nonzero pixels of its image at kNzSlot.
constexpr int kCandSlot = 2 * kMaxBatch;
constexpr int kNzSlot = 2 * kMaxBatch + 1;
constexpr int kCounterWords = 2 * kMaxBatch + 2;
// Hit records of a single-image call copied back with its counters (one
// round trip when there are at most this many).
constexpr uint64_t kHitPrefix = 4096;

// Single-image banded copies of 2^16-row super-tiles keep a 16-bit counter
// per row (two rows per LDS word): the sum of its pixels (exact while
// 255 * max_spots < 2^16), or, for sparse images, its count of nonzero
// pixels; rows that can pass on that counter are recounted from the table.
constexpr int kSum16TileBits = 16;
constexpr int kSum16MaxSpots = 257;

// Orientations per table tile = 1 << kTileBits (one scoring workgroup).
constexpr int kTileBits = 8;

// lm_fwd.hip
int forward_build(lm_db *db, int mode, lm_fwd_stats *stats);
int import_table(lm_db *db, const uint16_t *records, uint64_t count);
int export_table(lm_db *db, uint64_t first, uint64_t count, uint16_t *out);
void release_table(lm_db *db);
int selftest_fp64(hipStream_t st, const double *a, const double *b, uint64_t n, double *q,
                  double *s);
int table_checksum(lm_db *db, uint64_t *sum);

// lm_fwd.hip: the forward pre-test plan of the DB's hkls (built once)
int ensure_fwd_plan(lm_db *db, hipStream_t st);
// lm_fwd.hip: on-the-fly scoring, no table: every orientation projected
// (float64, or float32 with the guard band) and scored against B = 1, 4, 8,
// 16, 32 or 64 interleaved images (img) with the table walks' accumulators,
// predicate and hit records; d_stats[0..2] += pairs recomputed in float64,
// pairs projected, spots listed.  MaxNrLaueSpots <= 64.
// List slots 32..M-1 go to spill (at least spillRows * (M - 32) words;
// launches then take at most spillRows orientations).
int launch_fwd_score(lm_db *db, int B, bool guard, const uint8_t *img, const Pred *pred,
                     uint64_t *hits, uint64_t capPer, uint32_t *counters,
                     unsigned long long *d_stats, uint32_t *spill, uint64_t spillRows,
                     hipStream_t st);
// lm_fwd.hip: float64 in-order sums of k candidates (image << 32 | row) of
// images laid out back to back, each row's spots listed again.
int launch_fwd_exact(lm_db *db, const uint64_t *xids, uint64_t k, const uint8_t *imgs,
                     uint64_t npx, const Pred *pred, double *out, hipStream_t st);

// lm_hosttable.cpp: tables larger than the HBM budget, kept on the host and
// scored in chunks (forward_host / import_host are taken when the resident
// build returns LM_ERR_BUDGET)
int forward_host(lm_db *db, int mode, lm_fwd_stats *stats);
int import_host(lm_db *db, const uint16_t *records, uint64_t count);
int export_host(lm_db *db, uint64_t first, uint64_t count, uint16_t *out);
uint64_t checksum_host(const lm_db *db);
void release_host_table(lm_db *db);
int score_images_host(lm_db *db, Scratch &sc, int nimg, const uint8_t *d_imgs,
                      const double *im_scales, const int *min_spots, const double *min_int,
                      hipStream_t st, lm_hits *outs);
int score_seq_host(lm_db *db, Scratch &sc, const void *d_img, bool f64, double scale,
                   int min_nr_spots, double min_intensity, hipStream_t st, lm_hits *out);

// lm_band.hip
// best_effort: a copy the HBM budget cannot hold is left out (scoring then
// reads the orientation-major table) instead of failing.
int build_banded(lm_db *db, bool best_effort = false);
void release_banded(lm_db *db);
int launch_banded(lm_db *db, Scratch &sc, const uint8_t *img, uint64_t cap, hipStream_t st);
// Candidate list of the 16-bit-sum walk: rows / 16, at least 4096, at most
// 2^22 (32 MiB); a call with more candidates reruns over the table.
uint64_t cand_capacity(const lm_db *db);
bool banded_batch_fits(const lm_db *db, int B);
// Width B reads its batch copy with 16-bit sums (LM_OPT_BATCH_SUM16):
// candidates to cands (b << 32 | row) for the exact recount.
bool banded_batch16(const lm_db *db, int B);
int launch_banded_batch16(lm_db *db, Scratch &sc, int B, const uint8_t *img, hipStream_t st);
int launch_banded_batch(lm_db *db, Scratch &sc, int B, const uint8_t *img, uint64_t capPer,
                        hipStream_t st);

// lm_score.hip
// Score nimg device images laid out back to back ([nimg][npx]) with
// per-image thresholds; outs[b] receives image b's ascending hit list.
int score_images(lm_db *db, Scratch &sc, int nimg, const uint8_t *d_imgs,
                 const double *im_scales, const int *min_spots, const double *min_int,
                 hipStream_t stream, lm_hits *outs);
// One image, float64 sums accumulated in table order on the device
// (f64 = image of doubles; otherwise uint8 values times scale).
int score_seq(lm_db *db, Scratch &sc, const void *d_img, bool f64, double scale,
              int min_nr_spots, double min_intensity, hipStream_t stream, lm_hits *out);
// Images scored with every orientation projected on the fly (no table;
// mode LM_FWD_FP64 or LM_FWD_FP32_GUARD), groups of at most 64 images per
// projection pass; the same hits as score_images on the table that
// lm_db_forward(mode) builds.
int score_images_fwd(lm_db *db, Scratch &sc, int nimg, const uint8_t *d_imgs,
                     const double *im_scales, const int *min_spots, const double *min_int,
                     int mode, hipStream_t stream, lm_hits *outs, lm_fwd_stats *stats);

// Scoring without allocation (lm_score_dev_noalloc): every device buffer is
// carved from a file-scope device array whose address noalloc_init (called
// by lm_ctx_create with the device current) looks up once per device.
int noalloc_init(int device);
int score_noalloc(lm_db *db, const uint8_t *d_img, double im_scale, int min_sp, double min_int,
                  hipStream_t st, uint64_t cap, uint64_t *ids, int32_t *nspots, double *totint,
                  double *score, uint64_t *count, uint64_t *nr_results);

int width_log2(int B);       // smallest k with 2^k >= B
int batch_width(int nimg);   // kernel width for nimg images: 1, 4, 8, 16, 32, 64

// lm_tune.cpp
int autotune(lm_db *db, const uint8_t *const *imgs, int nimg, int max_batch, int reps,
             lm_tune_report *rep);

// lm_prep.hip: the preprocessing pipeline of lm_preprocess (u8 to the host
// and/or the device buffer d_u8; blurred optional).
int preprocess(lm_ctx *ctx, int W, int H, const double *raw, const double *background,
               const lm_prep_config &cfg, uint8_t *u8_host, uint8_t *d_u8, double *blurred,
               lm_prep_result &res);

// lm_api.cpp: slot `slot` of pool p with at least `bytes` (ctx->mu held by
// the caller); past the HBM budget hipErrorOutOfMemory, which hip_fail then
// reports as LM_ERR_BUDGET.
hipError_t pool_get(lm_ctx *ctx, DevPool &p, size_t slot, uint64_t bytes, void **out);
void ctx_ref(lm_ctx *ctx);
void ctx_unref(lm_ctx *ctx);  // releases the context at the last reference
void pool_release(lm_ctx *ctx, DevPool &p);

// lm_median.hip: scipy.ndimage.median_filter(size=size), `passes` times, of
// the W x H float64 device image d_img into d_bg (zeros when size <= 0 or
// passes <= 0); synchronizes st before returning.
constexpr int kMaxMedianSize = 255;
// Bits [lo, hi) in which n 64-bit device keys differ (hi = 0 when all are
// equal), so a radix sort can skip the others; d_acc: 8 bytes of device
// memory; synchronizes st.
hipError_t key_bit_range(hipStream_t st, const uint64_t *keys, uint64_t n,
                         unsigned long long *d_acc, int *lo, int *hi);
int median_background_dev(lm_ctx *ctx, hipStream_t st, int W, int H, const double *d_img,
                          int size, int passes, double *d_bg);

// lm_merge.cpp
void quat_from_matrix(const double m[9], double q[4]);
double misorientation_deg(const double a[4], const double b[4], int space_group);
int merge_candidates(const double *orients, uint64_t n_orients, const uint64_t *rows,
                     const double *scores, int nr, double max_angle, int space_group,
                     int threads, uint64_t *best_row, double *best_score, int *cluster_size);
// Same clustering with the pair tests on the device (stream st).
int merge_candidates_dev(hipStream_t st, const double *orients, uint64_t n_orients,
                         const uint64_t *rows, const double *scores, int nr, double max_angle,
                         int space_group, uint64_t *best_row, double *best_score,
                         int *cluster_size);

// lm_merge_dev.hip: bit l of row i (l > i, rows of W = ceil(n / 64) words)
// set when the clamped misorientation cosine of candidates l and i (their
// fundamental-region quaternions fq, n x 4) is at least cos_min.
int merge_bits_dev(hipStream_t st, const double *fq, int n, const SymSet &sym, double cos_min,
                   uint64_t *bits_host);

}  // namespace lm
