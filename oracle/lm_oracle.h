/*
 * lm_oracle.h - CPU oracle for the LaueMatching indexing hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain C99 + OpenMP restatement of the
 * reference LaueMatchingCPU hot loop, used as the checker for the HIP path.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product library (lauematching_amd) never links or calls it.
 *
 * Every function cites the reference file:line it restates
 * (paths relative to the AdvancedPhotonSource/LaueMatching repository).
 *
 * Arithmetic contract: IEEE binary64, left-to-right, no FMA contraction
 * (built with -ffp-contract=off), correctly rounded sqrt/division from libm.
 */
#ifndef LM_ORACLE_H
#define LM_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Detector / energy geometry, units as in the reference parameter file:
 * P in metres, pixel pitch in metres, energies in keV, recip in 1/nm. */
typedef struct {
  double recip[3][3];        /* B matrix from the lattice (calcRecipArray)  */
  double rotT[3][3];         /* transpose of the detector rotation matrix   */
  double P[3];               /* P_Array                                     */
  double pxX, pxY;           /* PxX, PxY                                    */
  int nrPxX, nrPxY;          /* NrPxX, NrPxY                                */
  double Elo, Ehi;           /* energy window, inclusive                    */
} lmo_geom;

/* src/LaueMatchingCPU.c:224-239: Rodrigues rotation from R_Array, transposed. */
void lmo_detector_rotT(const double rArr[3], double rotT[3][3]);

/* src/LaueMatchingHeaders.h:379-436 (calcRecipArray, zeroOut at :359). */
void lmo_calc_recip(const double lat[6], int spaceGroup, double recip[3][3]);

/* src/LaueMatchingHeaders.h:349-357 (MatrixMultF33): res = m * n. */
void lmo_matmul33(const double m[3][3], const double n[3][3], double res[3][3]);

/* One (orientation, hkl) projection step, src/LaueMatchingCPU.c:604-636.
 * thisOrient = orientation * recip.  Returns 1 if the reflection lands on the
 * detector inside the energy window, writing the pixel; also returns qhat and
 * the energy when the pointers are non-NULL (for geometry tests). */
int lmo_project(const double thisOrient[3][3], const int hkl[3], const lmo_geom *g,
                int *ipx, int *ipy, double qhat_out[3], double *E_out);

/* Forward simulation of orientations [0, n) into the ForwardFile table
 * (n x (1+2M) uint16, orientation-major; word 0 = count, then x,y pairs,
 * unused words zero), src/LaueMatchingCPU.c:591-667.  If img_f64 is non-NULL
 * the DoFwd=1 scoring against the float64 image (:647-653, :683-689) is done
 * as well and written to matched (n doubles, 0 = not a hit); returns the
 * number of hits, or the number of orientations processed when img is NULL. */
long lmo_forward(const double *orients, size_t n, const int *hkls, int nhkls,
                 const lmo_geom *g, int maxNrSpots, uint16_t *table,
                 const double *img_f64, int minNrSpots, double minIntensity,
                 double *matched, int nthreads);

/* src/LaueMatchingCPU.c:355-377: uint8 quantization, returns imScale. */
double lmo_quantize_u8(const double *img, size_t npx, uint8_t *out);

/* DoFwd=0 compare against the uint8 image, src/LaueMatchingCPU.c:668-689.
 * matched[o] = totInt*sqrt(nSpots) for hits, else 0.  Optional per-orientation
 * outputs nspots/totint (may be NULL).  Returns nrResults. */
long lmo_compare_u8(const uint16_t *table, size_t n, int maxNrSpots,
                    const uint8_t *img, int nrPxX, double imScale,
                    int minNrSpots, double minIntensity, double *matched,
                    int32_t *nspots, double *totint, int nthreads);

/* src/LaueMatchingCPU.c:772-788: ascending collection of matched != 0.
 * Writes up to cap entries; returns the number of hits found. */
long lmo_collect(const double *matched, size_t n, uint64_t *rowNrs, double *mA,
                 long cap);

/* ---- merge of duplicate candidates (lm_merge_oracle.c) ---------------------- */

/* OrientMat2Quat, src/LaueMatchingHeaders.h:262-296 (row-major matrix). */
void lmo_quat_from_matrix(const double m[9], double q[4]);

/* GetMisOrientation, src/LaueMatchingHeaders.h:248-260, with the point group
 * of space group sg (MakeSymmetries :182-220); degrees. */
double lmo_misorientation_deg(const double q1[4], const double q2[4], int sg);

/* mergeDuplicateOrientations, src/LaueMatchingHeaders.h:949-1065: the cap at
 * the 50000 best-scored candidates (qsort), float32 misorientations, greedy
 * clustering in list order with `<= maxAngle` and the strict `>` best-score
 * rule.  Outputs hold min(nr, 50000) entries; returns the cluster count. */
int lmo_merge(const double *orients, const uint64_t *rows, const double *scores, int nr,
              double maxAngle, int sg, int nthreads, uint64_t *best_row, double *best_score,
              int32_t *cluster_size);

/* One pass of scipy.ndimage.median_filter(size=s), mode 'reflect', as
 * compute_background calls it (laue_index/preprocess.py:59-63), on a frame of
 * dense ranks rk[H][W] (0..D-1, increasing with the value): out = the rank
 * of each window's element of 0-based order s*s/2.  s <= 255.  Returns 0,
 * -1 on bad arguments, -2 when a buffer could not be allocated. */
int lmo_median_ranks(const uint32_t *rk, int W, int H, int s, uint32_t D, uint32_t *out);

#ifdef __cplusplus
}
#endif
#endif
