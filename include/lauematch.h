/*
 * lauematch.h - C ABI of the MI355X-native LaueMatching indexing hot path.
 *
 * The reference has no library ABI: its hot path lives inside the main()
 * functions of LaueMatchingCPU / LaueMatchingGPU / LaueMatchingGPUStream.
 * These entry points replace the in-process seams of those programs:
 *
 *   - the OpenMP forward simulation that fills the ForwardFile table
 *     (src/LaueMatchingCPU.c:591-667, src/LaueMatchingGPU.cu:463-617,
 *     src/LaueMatchingGPUStream.cu:836-952)        -> lm_db_forward
 *   - reading / writing the ForwardFile table
 *     (src/LaueMatchingCPU.c:446-486, 548-576, 693-726)
 *                                                  -> lm_db_import_table,
 *                                                     lm_db_export_table
 *   - the uint8 image quantization (src/LaueMatchingCPU.c:355-377)
 *                                                  -> lm_quantize_u8
 *   - the compare/score step with hit collection in ascending orientation
 *     order (src/LaueMatchingCPU.c:668-689, 772-788; the CUDA kernel
 *     `compare` and its chunk loop, src/LaueMatchingGPU.cu:45-81, 660-792)
 *                                                  -> lm_score, lm_score_dev
 *   - the streaming daemon's per-image GPU work
 *     (src/LaueMatchingGPUStream.cu:1167-1257)     -> lm_score_batch and the
 *                                                     lm_server_* request loop
 *
 * Conventions: every call returns LM_OK (0) or a negative LM_ERR_* code and
 * never exits the process; lm_last_error() gives the calling thread's last
 * message.  Host inputs stay owned by the caller; result arrays inside
 * lm_hits are owned by the library until lm_hits_free().  Units are those of
 * the reference parameter file: lattice in nm and degrees, P_Array and pixel
 * pitch in metres, R_Array a rotation vector in radians, energies in keV.
 */
#ifndef LAUEMATCH_H
#define LAUEMATCH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LM_OK 0
#define LM_ERR_INVALID (-1)  /* bad argument or inconsistent sizes            */
#define LM_ERR_HIP (-2)      /* HIP runtime error                             */
#define LM_ERR_NOMEM (-3)    /* host or device allocation failed              */
#define LM_ERR_BUDGET (-4)   /* request exceeds the context's HBM budget      */
#define LM_ERR_OVERFLOW (-5) /* hit list exceeded its capacity                */
#define LM_ERR_STATE (-6)    /* object not ready (e.g. no table yet)          */

/* Forward-simulation modes for lm_db_forward. */
#define LM_FWD_FP64 0       /* float64, exact reference operation order          */
#define LM_FWD_FP32_GUARD 1 /* float32 with a float64 recompute near boundaries  */

typedef struct lm_ctx lm_ctx;
typedef struct lm_db lm_db;

/* Geometry of one detector + crystal, computed on the host exactly as the
 * reference does (src/LaueMatchingCPU.c:224-239, src/LaueMatchingHeaders.h:379-436). */
typedef struct lm_geom {
  double recip[3][3]; /* reciprocal basis B (1/nm), reference calcRecipArray  */
  double rotT[3][3];  /* transpose of the detector rotation (Rodrigues)       */
  double P[3];        /* P_Array (m)                                          */
  double pxX, pxY;    /* pixel pitch (m)                                      */
  int32_t nrPxX, nrPxY;
  double Elo, Ehi;    /* inclusive energy window (keV)                        */
} lm_geom;

typedef struct lm_fwd_stats {
  uint64_t n_orients;    /* orientations simulated                             */
  uint64_t n_spots;      /* spots in the table                                 */
  uint64_t n_pairs;      /* (orientation, hkl) pairs evaluated                 */
  uint64_t n_recomputed; /* pairs recomputed in float64 (FP32_GUARD only)      */
  double ms;             /* device time of the simulation                      */
  uint64_t n_projected;  /* pairs past the division-free pre-test, projected
                            in full (FP64, or FP32 with the guard band)       */
} lm_fwd_stats;

/* Compact result of scoring one image against one DB. */
typedef struct lm_hits {
  uint64_t nr_results; /* orientations passing the predicate (reference
                          nrResults, src/LaueMatchingCPU.c:683-689)           */
  uint64_t count;      /* collected hits (score != 0), ascending id,
                          src/LaueMatchingCPU.c:781-788                       */
  uint64_t *ids;       /* orientation row numbers                            */
  int32_t *nspots;     /* spots with intensity > 0                           */
  double *totint;      /* float64 intensity sum, record order                */
  double *score;       /* totint * sqrt(nspots)                              */
  double kernel_ms;    /* device time of the scoring launch sequence         */
  double walk_ms;      /* device time of the compare kernel alone (single
                          image over the banded copy, eager launches; else
                          equal to kernel_ms)                                */
} lm_hits;

/* ---- context -------------------------------------------------------------- */

/* Bind device `device`; hbm_budget_bytes = 0 means 90% of free device memory.
 * No reference counterpart (the reference selects device 0 implicitly). */
int lm_ctx_create(int device, uint64_t hbm_budget_bytes, lm_ctx **out);
/* Releases the caller's context; DBs and servers made on it keep it (its
 * stream and budget) until the last of them is destroyed, in any order. */
void lm_ctx_destroy(lm_ctx *ctx);
/* Bytes of device memory currently held by the context's DBs and buffers. */
uint64_t lm_ctx_bytes_in_use(const lm_ctx *ctx);
const char *lm_last_error(void);
const char *lm_version(void);

/* ---- host-side geometry ----------------------------------------------------- */

/* Fill g from parameter-file values (reference main(), src/LaueMatchingCPU.c:
 * 224-239 and :418-419). */
int lm_geom_init(const double lattice[6], int32_t space_group, const double P[3],
                 const double R[3], double pxX, double pxY, int32_t nrPxX, int32_t nrPxY,
                 double Elo, double Ehi, lm_geom *g);

/* uint8 quantization of a float64 image, src/LaueMatchingCPU.c:355-377.
 * Writes out[npx] and *im_scale (= max/255, or 1 when max <= 0). */
int lm_quantize_u8(const double *img, uint64_t npx, uint8_t *out, double *im_scale);
/* The same quantization of a float32 image (the streaming daemon's frames,
 * src/LaueMatchingGPUStream.cu:299-368): identical to lm_quantize_u8 on the
 * pixels widened to float64, without the float64 copy. */
int lm_quantize_u8_f32(const float *img, uint64_t npx, uint8_t *out, double *im_scale);

/* ---- orientation DB + spot table ------------------------------------------------ */

/* Upload n orientations (row-major 3x3 float64, the 72-byte records of the
 * orientation DB file, src/LaueMatchingCPU.c:241-287) and the HKL list
 * (src/LaueMatchingCPU.c:289-312).  max_spots = MaxNrLaueSpots (1..511). */
int lm_db_create(lm_ctx *ctx, const double *orients, uint64_t n, const int32_t *hkls,
                 int32_t nhkls, const lm_geom *geom, int32_t max_spots, lm_db **out);
void lm_db_destroy(lm_db *db);
uint64_t lm_db_size(const lm_db *db);
/* Device bytes of the resident spot table (0 before forward/import). */
uint64_t lm_db_table_bytes(const lm_db *db);
/* Spots in the table (sum of per-orientation counts). */
uint64_t lm_db_table_spots(const lm_db *db);

/* Layout options (no reference counterpart; used by the autotuner).
 * LM_OPT_BANDED_TILE_BITS: orientations per super-tile of the pixel-banded
 * table copy used by single-image scoring, as log2 (13..16); 0 drops the
 * copy (scoring then reads the orientation-major table only).  Default 16
 * (15 when MaxNrLaueSpots > 257).  Applies at
 * once if a table exists.  16 (MaxNrLaueSpots <= 257 only) keeps a 16-bit
 * counter per row — the sum of its pixels, or for sparse images (under a
 * quarter of the pixels nonzero) with MinNrSpots >= 2 its count of nonzero
 * pixels — and recounts the rows that can pass on it; calls with MinIntensity
 * <= 0, an image that is not 16-byte aligned, or more such rows than n/16
 * (at least 4096, at most 2^22) read the orientation-major table.
 * LM_OPT_BATCH_TILE_BITS: super-tile size (log2, 9..13) of a second banded
 * copy used by batches of width B (4..64) with B * 2^bits <= 32768 (the
 * per-(image, orientation) counters of a super-tile fill at most 128 KiB of
 * LDS); 0 (default) drops it.  Applies at once if a table exists.
 * LM_OPT_BATCH_WIDTHS: bit k set = batches of width 2^k read the batch copy
 * (when it fits); clear = they read the orientation-major table.  Default
 * 0x7f. */
#define LM_OPT_BANDED_TILE_BITS 1
#define LM_OPT_BATCH_TILE_BITS 2
#define LM_OPT_BATCH_WIDTHS 3
/* LM_OPT_USE_GRAPHS: 1 = direct scoring calls on this DB replay captured
 * launch sequences (hipGraph), keyed by every pointer and size they use. */
#define LM_OPT_USE_GRAPHS 4
/* LM_OPT_BANDED_LOADS: the single-image banded walk: 0 = eight 4-byte entry
 * loads per lane in flight, 1 = two 16-byte loads (8 entries), 2 = four
 * 16-byte loads (16 entries), 3 = the image staged through LDS in per-wave
 * chunks (2^14..2^16 super-tiles), 4 = as 0 with each pixel read through a
 * 4-byte gather of its aligned word; 3, 5, 6, 7 = the LDS walk with 4, 8, 4, 8
 * entries per lane (8, 12, 8, 12 for 2^16 super-tiles) and 2, 2, 3, 3 chunk
 * sets in flight per wave.  2^16 super-tiles always take the LDS walk (3 for
 * 0, 1, 2, 4).  Default 3. */
#define LM_OPT_BANDED_LOADS 5
/* LM_OPT_BATCH_GROUPED: bit 6 / bit 7 set = table batches of 64 / 128 images
 * give each orientation B/16 lanes of 16 images each (one 16-byte load per
 * lane and spot) instead of one lane holding all B.  Default 0xc0. */
#define LM_OPT_BATCH_GROUPED 6
/* LM_OPT_BATCH_NIBBLE: bit 7 / bit 8 set = batches of 128 / 129..256 images
 * read 4-bit upper bounds of the pixels (ceil(v / 17): zero exactly when v
 * is, so spot counts stay exact), interleaved as B/2 bytes per pixel, and
 * recount the (row, image) pairs that can pass from the uint8 images.  Needs
 * MinIntensity > 0 for every image, 4-byte aligned images and a pixel count
 * divisible by 4; otherwise (and when more pairs can pass than the candidate
 * list holds) 128 images take the uint8 walks and wider batches are scored
 * as two calls of at most 128.  Default 0x180. */
#define LM_OPT_BATCH_NIBBLE 7
/* LM_OPT_TABLE_ON_HOST (read only): 1 when the spot table did not fit the
 * context's HBM budget and is kept in host memory (compact: counts and 4
 * bytes per spot); every scoring call then copies it to the device in chunks
 * of at most 4 GB and half the free budget and scores each chunk, with the
 * same results (src/LaueMatchingGPU.cu:660-749 streams the ForwardFile the
 * same way).  Such a DB cannot be autotuned, served by lm_server or scored
 * by lm_score_dev_noalloc.  A banded copy that does not fit is left out
 * (scoring reads the orientation-major table). */
#define LM_OPT_TABLE_ON_HOST 8
/* LM_OPT_BATCH_WIDTH_TILE_BITS (set only): value = (log2 B) << 8 | bits gives
 * batches of width B (4..64) their own batch copy of 2^bits-row super-tiles
 * (9..13, B x 2^bits <= 32768; 0 = none), so different widths can keep
 * different copies (the autotuner sets these); LM_OPT_BATCH_TILE_BITS sets
 * every fitting width at once. */
#define LM_OPT_BATCH_WIDTH_TILE_BITS 9
/* LM_OPT_BATCH_SUM16: bit k set (k = 2..6) = batches of width 2^k read their
 * batch copy with 16-bit sums, two images per LDS counter word, so the copy
 * may have B x 2^bits <= 65536 (set this before LM_OPT_BATCH_WIDTH_TILE_BITS);
 * the (row, image) pairs whose sum can pass are recounted exactly from the
 * table.  Calls where some MinIntensity <= 0 read the table instead. */
#define LM_OPT_BATCH_SUM16 10
/* LM_OPT_BATCH_SPLIT: bit k set (k = 3..6) = a batch of width 2^k (more than
 * 2^(k-1) images) is scored as two calls of at most 2^(k-1) images, when the
 * narrower width's layout serves two calls faster than one; bit 2 = 2..4
 * images (width 4) are scored one at a time (the autotuner sets both from
 * its timings). */
#define LM_OPT_BATCH_SPLIT 11
/* LM_OPT_BANDED_SPLIT: the LDS-staged walk's work items are pieces of the
 * entries of one image chunk: 2^value balanced pieces per chunk (0..3), or -1
 * (default) for the fewest that bring the mean piece to at most 512 entries
 * (small detectors put thousands of a super-tile's entries in each chunk).
 * Reading it gives the split of the current copy. */
#define LM_OPT_BANDED_SPLIT 12
int lm_db_set_option(lm_db *db, int32_t option, int64_t value);
int lm_db_get_option(lm_db *db, int32_t option, int64_t *value);

/* Autotuner (no reference counterpart): times every scoring layout on the
 * sample images (nimg >= 1 host uint8 images of the DB's detector, repeated
 * to fill a batch) for one image and for batch widths 4, 8, ..., max_batch,
 * best of `reps` runs each, then sets LM_OPT_BANDED_TILE_BITS,
 * LM_OPT_BANDED_LOADS, LM_OPT_BATCH_TILE_BITS, LM_OPT_BATCH_WIDTHS,
 * LM_OPT_BATCH_GROUPED and LM_OPT_BATCH_NIBBLE to the fastest.  Needs a
 * table; the timings are device milliseconds per launch sequence. */
typedef struct lm_tune_report {
  int32_t n_widths;
  int32_t widths[7];          /* batch widths timed (4 .. max_batch <= 256)      */
  double ms_single_table;     /* one image: orientation-major table               */
  double ms_single[4][8];     /* one image: banded 2^13 / 2^14 / 2^15 / 2^16 x
                                 loads 0..7, the LDS walks at their best split
                                 (0 = not available)                             */
  double ms_lane[7];          /* per width: orientation-major table               */
  double ms_grouped[7];       /* per width: table, grouped lanes (64, 128 only)   */
  double ms_banded[7][5];     /* per width: batch copy 2^9..2^13 (0 = no fit)     */
  double ms_nibble[7];        /* per width: 4-bit bounds (128, 256 only)          */
  int32_t banded_tile_bits;   /* chosen LM_OPT_BANDED_TILE_BITS                   */
  int32_t banded_loads;       /* chosen LM_OPT_BANDED_LOADS                       */
  int32_t batch_tile_bits;    /* chosen LM_OPT_BATCH_TILE_BITS                    */
  uint32_t batch_widths_mask; /* chosen LM_OPT_BATCH_WIDTHS                       */
  uint32_t grouped_mask;      /* chosen LM_OPT_BATCH_GROUPED                      */
  uint32_t nibble_mask;       /* chosen LM_OPT_BATCH_NIBBLE                       */
  int32_t batch_tile_bits_w[7]; /* per width: super-tile bits of the batch copy it
                                   keeps (0 = a table walk)                        */
  double ms_banded16[7][5];   /* per width: batch copy 2^9..2^13 read with 16-bit
                                 sums (0 = no fit)                               */
  uint32_t sum16_mask;        /* chosen LM_OPT_BATCH_SUM16                        */
  uint32_t split_mask;        /* chosen LM_OPT_BATCH_SPLIT                        */
  int32_t banded_split;       /* chosen LM_OPT_BANDED_SPLIT (-1: no LDS walk kept) */
  double ms_split[4];         /* the kept single-image layout with 2^0..2^3 pieces
                                 per chunk (0 = not timed)                      */
} lm_tune_report;
int lm_db_autotune(lm_db *db, const uint8_t *const *images, int32_t nimg, int32_t max_batch,
                   int32_t reps, lm_tune_report *out);

/* Build the resident spot table on the device (replaces the DoFwd=1 forward
 * simulation: src/LaueMatchingCPU.c:591-667, in the GPU programs
 * src/LaueMatchingGPU.cu:463-617 and src/LaueMatchingGPUStream.cu:836-952).
 * mode = LM_FWD_FP64 or LM_FWD_FP32_GUARD; stats may be NULL. */
int lm_db_forward(lm_db *db, int32_t mode, lm_fwd_stats *stats);

/* Replace the table from ForwardFile records (the DoFwd=0 read,
 * src/LaueMatchingCPU.c:446-486, 548-576): count records of (1+2M) uint16
 * (word 0 = spot count, then x,y pairs), orientation-major, covering rows
 * [0, count) with count == lm_db_size(db).  Counts > M or pixels outside the
 * detector are rejected with LM_ERR_INVALID. */
int lm_db_import_table(lm_db *db, const uint16_t *records, uint64_t count);

/* 64-bit position-tagged checksum of the table (counts and spot lists in
 * table order), for comparing tables built different ways (e.g. LM_FWD_FP64
 * vs LM_FWD_FP32_GUARD) at sizes where exporting them is impractical.  No
 * reference counterpart. */
int lm_db_table_checksum(lm_db *db, uint64_t *sum);

/* Write rows [first, first+count) as ForwardFile records (zero padded),
 * byte-identical to the reference writer (src/LaueMatchingCPU.c:693-726). */
int lm_db_export_table(lm_db *db, uint64_t first, uint64_t count, uint16_t *out);

/* ---- scoring -------------------------------------------------------------------- */

/* Score a host uint8 image (nrPxX x nrPxY of the DB geometry, row-major,
 * X fastest) with the reference predicate nspots >= min_nr_spots &&
 * totint >= min_intensity, totint = sequential float64 sum of raw*im_scale
 * in table order (the DoFwd=0 compare, src/LaueMatchingCPU.c:668-689; the
 * GPU programs' compare kernel, src/LaueMatchingGPU.cu:45-81, 722-775).
 * Results in ascending orientation order, as the reference collects them
 * (src/LaueMatchingCPU.c:772-788). */
int lm_score(lm_db *db, const uint8_t *img_u8, double im_scale, int32_t min_nr_spots,
             double min_intensity, lm_hits *out);

/* Score a float64 image the way the reference's DoFwd=1 branch does:
 * nspots counts pixels > 0, totint is their sequential float64 sum in table
 * order (src/LaueMatchingCPU.c:647-653, 683-689; the GPU programs reach it
 * through src/LaueMatchingGPU.cu:463-617). */
int lm_score_f64(lm_db *db, const double *img_f64, int32_t min_nr_spots, double min_intensity,
                 lm_hits *out);

/* Same with the image already in device memory (d_img_u8 is a device
 * pointer); stream is a hipStream_t (NULL = the context's stream). */
int lm_score_dev(lm_db *db, const uint8_t *d_img_u8, double im_scale, int32_t min_nr_spots,
                 double min_intensity, void *stream, lm_hits *out);

/* lm_score_dev with no allocation: no hipMalloc, hipFree or other allocating
 * call, for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid it).  No counterpart in the reference.
 * Every device buffer of the call is carved from a fixed 17.3 MiB array that
 * each GPU holds from the library's load until the process ends (hit records
 * up to 65536 per call, candidate rows up to 2^21, detectors of at most 2^27
 * pixels; beyond these the call fails with LM_ERR_OVERFLOW / LM_ERR_INVALID
 * and lm_score_dev serves).  Calls that overlap on one GPU share that array
 * and race: the caller serialises them.  Results go to the caller's host
 * arrays ids/nspots/totint/score of cap entries, ascending ids, identical to
 * lm_score_dev's; cap must hold the call's candidate records (else
 * LM_ERR_OVERFLOW with *count = the number needed).  No kernel timing. */
int lm_score_dev_noalloc(lm_db *db, const uint8_t *d_img_u8, double im_scale,
                         int32_t min_nr_spots, double min_intensity, void *stream, uint64_t cap,
                         uint64_t *ids, int32_t *nspots, double *totint, double *score,
                         uint64_t *count, uint64_t *nr_results);

/* Score nimg (1..256) host images against the DB in one pass over the table.
 * outs[i] receives image i's hits, identical to lm_score on that image.  No
 * reference counterpart: the reference programs score one image per pass
 * over the table (src/LaueMatchingGPU.cu:660-749,
 * src/LaueMatchingGPUStream.cu:1167-1257). */
int lm_score_batch(lm_db *db, int32_t nimg, const uint8_t *const *imgs_u8,
                   const double *im_scales, int32_t min_nr_spots, double min_intensity,
                   lm_hits *outs);

/* Same with thresholds per image (min_nr_spots[i], min_intensity[i]), as
 * requests of different clients share one batch in the server.  No
 * counterpart in the reference, whose programs take one threshold pair per
 * run (src/LaueMatchingCPU.c:683-689). */
int lm_score_batch_each(lm_db *db, int32_t nimg, const uint8_t *const *imgs_u8,
                        const double *im_scales, const int32_t *min_nr_spots,
                        const double *min_intensity, lm_hits *outs);

/* Device-resident batch: d_imgs points to nimg device images laid out back to
 * back (nimg * nrPxX * nrPxY bytes).  No reference counterpart. */
int lm_score_batch_dev(lm_db *db, int32_t nimg, const uint8_t *d_imgs, const double *im_scales,
                       int32_t min_nr_spots, double min_intensity, void *stream, lm_hits *outs);

/* On-the-fly scoring: the forward simulation of src/LaueMatchingCPU.c:591-667
 * (mode LM_FWD_FP64 or LM_FWD_FP32_GUARD) and the compare of :668-689 in one
 * pass, with no spot table read or written -- each orientation's spot list
 * is projected and scored against the images as its spots are found.  Images
 * as in lm_score_batch_dev (1..256, device-resident, back to back), with
 * thresholds per image; groups of up to 64 images share one projection pass.
 * Each image's hits equal those lm_score_batch_dev returns from the table
 * lm_db_forward(mode) builds (no table is needed: the DB may have none).
 * stats (may be NULL): pairs, projected and recomputed pairs summed over the
 * passes, n_spots the spots listed per pass, ms the device time.  Needs
 * MaxNrLaueSpots <= 64.  The reference reads a precomputed ForwardFile
 * instead (src/LaueMatchingGPU.cu:660-749); no counterpart there. */
int lm_score_fwd_batch_dev(lm_db *db, int32_t nimg, const uint8_t *d_imgs,
                           const double *im_scales, const int32_t *min_nr_spots,
                           const double *min_intensity, int32_t mode, void *stream,
                           lm_hits *outs, lm_fwd_stats *stats);

void lm_hits_free(lm_hits *h);

/* ---- serving: continuous batching over resident DBs ---------------------------------- */

/* Replaces the streaming daemon's GPU side (src/LaueMatchingGPUStream.cu:103-292
 * bounded queue, :1040-1075 per-stream buffers, :1167-1257 request loop with its
 * serialising "GPU gate").  Requests wait in pinned admission slots; each worker
 * owns a HIP stream and takes the oldest request's DB with up to max_batch of that
 * DB's pending requests, so requests arriving while others run join the next
 * batch and batches of different workers overlap on the device.  Each request's
 * result equals lm_score on that image alone.  DBs must be added before the first
 * submit and must not be rebuilt while the server runs. */
typedef struct lm_server lm_server;
typedef struct lm_server_config {
  int32_t max_batch;    /* images per launch, 1..256 (0 = 32)                     */
  int32_t workers;      /* concurrent batches / HIP streams (0 = 2)              */
  int32_t max_inflight; /* admitted, not yet completed requests (0 = 2*workers*max_batch) */
  int32_t use_graphs;   /* replay captured launch sequences (hipGraph)           */
  int32_t batch_after_ms; /* batch only when the pending images of a DB would take
                             longer than this one at a time (autotuned DBs;
                             0 = four single-image launches, < 0 = always batch) */
} lm_server_config;
typedef struct lm_server_stats {
  uint64_t completed;
  uint64_t batches;
  double mean_batch;    /* images per launch                                     */
  double p50_ms, p99_ms, max_ms; /* submit -> result latency                      */
  double images_per_s;  /* completed / (last completion - first submit)           */
  uint64_t graph_launches;
} lm_server_stats;

int lm_server_create(lm_ctx *ctx, const lm_server_config *cfg, lm_server **out);
int lm_server_add_db(lm_server *s, lm_db *db, int32_t *db_id);
/* Copies the image (db geometry, uint8) into an admission slot, blocking while
 * all slots are in use; reserves the serving buffers from the context's HBM
 * budget on the first call (LM_ERR_BUDGET if they do not fit). */
int lm_server_submit(lm_server *s, int32_t db_id, const uint8_t *img_u8, double im_scale,
                     int32_t min_nr_spots, double min_intensity, uint64_t *ticket);
/* Blocks until the ticket's result is ready; the caller frees it with lm_hits_free. */
int lm_server_wait(lm_server *s, uint64_t ticket, lm_hits *out, double *latency_ms);
/* Non-blocking form: *ready = 1 and the result in *out if it is done (the
 * ticket is then collected), else *ready = 0.  Unknown or already collected
 * tickets are LM_ERR_INVALID for both calls. */
int lm_server_poll(lm_server *s, uint64_t ticket, lm_hits *out, double *latency_ms,
                   int32_t *ready);
int lm_server_get_stats(lm_server *s, lm_server_stats *stats);
void lm_server_reset_stats(lm_server *s);
/* Finishes every admitted request, then releases the server. */
void lm_server_destroy(lm_server *s);

/* ---- merge of duplicate candidates ------------------------------------------------- */

/* Cluster the nr candidates (orientation rows + coarse scores, in hit-list
 * order) by misorientation <= max_angle degrees under the point group of
 * space_group, greedily in list order; each cluster reports its best-scored
 * row and score and its size (mergeDuplicateOrientations,
 * src/LaueMatchingHeaders.h:969-1065, including the cap at the 50000
 * best-scored candidates).  Output arrays hold min(nr, 50000) entries;
 * *n_clusters receives the cluster count.  threads <= 1 runs serially. */
int lm_merge_candidates(const double *orients, uint64_t n_orients, const uint64_t *rows,
                        const double *scores, int32_t nr, double max_angle, int32_t space_group,
                        int32_t threads, uint64_t *best_row, double *best_score,
                        int32_t *cluster_size, int32_t *n_clusters);

/* The same merge with the pairwise misorientation tests on the device of
 * ctx (replaces the O(nr^2) loop of src/LaueMatchingHeaders.h:969-1065); the
 * result is identical to lm_merge_candidates.  The greedy pass stays on the
 * host over the device's pair bits (nr x ceil(nr/64) words, 313 MB at the
 * 50000 cap). */
int lm_merge_candidates_dev(lm_ctx *ctx, const double *orients, uint64_t n_orients,
                            const uint64_t *rows, const double *scores, int32_t nr,
                            double max_angle, int32_t space_group, uint64_t *best_row,
                            double *best_score, int32_t *cluster_size, int32_t *n_clusters);

/* Unit quaternion (w >= 0) of a row-major rotation matrix
 * (OrientMat2Quat, src/LaueMatchingHeaders.h:262-296). */
void lm_quat_from_matrix(const double m[9], double q[4]);

/* Misorientation angle (degrees) between two unit quaternions under the
 * point group of space_group (GetMisOrientation, src/LaueMatchingHeaders.h:247-260). */
double lm_misorientation_deg(const double q1[4], const double q2[4], int32_t space_group);

/* ---- image preprocessing ----------------------------------------------------------- */

/* The preprocessing before the compare (preprocess_image,
 * laue_index/preprocess.py:221-337, thresholds laue_index/thresholds.py:29-110)
 * on the device: background subtraction max(raw - background, 0); a
 * threshold (pixels <= t become 0) that is fixed, a percentile of all pixels
 * (numpy's linear interpolation), or adaptive (median of the nonzero pixels
 * + 4 * 1.4826 * their MAD, at least 1; also used for "otsu", as the
 * reference does without scikit-image); the cast to uint16; 8-connected
 * components of the nonzero pixels, those with fewer than min_area pixels
 * removed; the Gaussian sigma max(0.25 * min(nearest centre-of-mass
 * distance, distance * tan(orientation_spacing) / px), 1) (3 with fewer than
 * two components); a float64 Gaussian blur (radius int(4 sigma + 0.5),
 * reflecting edges) and the uint8 quantization of lm_quantize_u8.  With a
 * NULL background the background is lm_median_background of the raw image
 * with cfg's filter_radius and median_passes (preprocess.py:255-266: no
 * background file, so it is computed; zero when either is <= 0).  No
 * counterpart runs on the GPU in the reference. */
#define LM_THRESH_ADAPTIVE 0
#define LM_THRESH_PERCENTILE 1
#define LM_THRESH_FIXED 2
#define LM_THRESH_OTSU 3
typedef struct lm_prep_config {
  int32_t threshold_method;    /* LM_THRESH_*                                       */
  double threshold_value;      /* LM_THRESH_FIXED                                   */
  double threshold_percentile; /* LM_THRESH_PERCENTILE (0..100)                     */
  int32_t min_area;            /* components with fewer pixels are removed          */
  double px_x;                 /* pixel size (mm), detector distance (mm) and the   */
  double distance;             /* DB's orientation spacing (degrees) bound sigma    */
  double orientation_spacing;
  int32_t filter_radius;       /* background median filter size (0: none)          */
  int32_t median_passes;       /* and passes, used when background is NULL          */
} lm_prep_config;
typedef struct lm_prep_result {
  double threshold;      /* threshold applied                                  */
  double sigma;          /* Gaussian sigma (pixels; 0 when nothing was kept)   */
  int32_t n_components;  /* 8-connected components before the area filter      */
  int32_t n_kept;        /* components kept                                    */
  double im_scale;       /* imScale of the uint8 image                         */
  double ms;             /* device time of the whole pipeline                  */
  uint64_t n_above;      /* pixels above the threshold (the ones it keeps)     */
} lm_prep_result;
/* raw and background: nrPxY x nrPxX float64 host images (background may be
 * NULL); the uint8 image goes to u8_host and/or d_u8 (a device buffer of the
 * context's device), blurred (optional, host) receives the float64 image. */
int lm_preprocess(lm_ctx *ctx, int32_t nr_px_x, int32_t nr_px_y, const double *raw,
                  const double *background, const lm_prep_config *cfg, uint8_t *u8_host,
                  void *d_u8, double *blurred, lm_prep_result *res);

/* The median-filter background of a raw frame (compute_background,
 * laue_index/preprocess.py:43-64) on the device, as the reference computes
 * it without diplib: scipy.ndimage.median_filter(bg, size=filter_radius)
 * applied median_passes times (:61-65), i.e. per pixel the element of rank
 * filter_radius^2 / 2 (0-based, rounded down) of the filter_radius x
 * filter_radius window whose first row and column lie filter_radius / 2
 * before the pixel, edges extended by half-sample reflection.  Zeros when
 * filter_radius <= 0 or median_passes <= 0 (:53-54).  image and background
 * are nr_px_y x nr_px_x float64 host images; ms (optional) receives the
 * device time.  Limits: filter_radius <= 255, at most 2^24 pixels.  diplib's
 * MedianFilter (:56-59), the reference's first choice when it is installed,
 * uses a circular window instead and is not provided. */
int lm_median_background(lm_ctx *ctx, int32_t nr_px_x, int32_t nr_px_y, const double *image,
                         int32_t filter_radius, int32_t median_passes, double *background,
                         double *ms);

/* ---- diagnostics ------------------------------------------------------------------ */

/* Evaluate on the device, with the forward simulation's arithmetic, q[i] =
 * a[i] / b[i] and s[i] = sqrt(a[i]) for n host inputs (results to host).  No
 * reference counterpart: it lets callers verify that the device float64
 * division and square root round like the host's (the property the
 * bit-exact ForwardFile relies on). */
int lm_selftest_fp64(lm_ctx *ctx, const double *a, const double *b, uint64_t n, double *q,
                     double *s);

#ifdef __cplusplus
}
#endif
#endif
