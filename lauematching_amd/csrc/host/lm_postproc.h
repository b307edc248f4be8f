// lm_postproc.h - CPU post-processing of the ascending hit list: merge of
// duplicate candidates, orientation write-out, spot lists.
//
// Follows the reference consumers of the hit list:
//   mergeDuplicateOrientations  (src/LaueMatchingHeaders.h:969-1065)
//   fitAndWriteOrientations     (src/LaueMatchingHeaders.h:1072-1153)
//   writeCalcOverlap            (src/LaueMatchingHeaders.h:807-947)
// NLopt is not available to this build, so the orientation refinement
// (FitOrientation, src/LaueMatchingHeaders.h:696-801) is not run: every
// solution is written at its coarse orientation with the overlap objective
// evaluated there (the starting point the optimizer would improve from),
// lattice parameters unchanged and the post-refinement misorientation that
// of the Euler round trip.
#pragma once

#include <cstdint>
#include <cstdio>
#include <string>

#include "lauematch.h"
#include "lm_params.h"

struct PostInput {
  const double *orients = nullptr;  // the orientation DB (N x 9)
  uint64_t n_orients = 0;
  const lm_hits *hits = nullptr;    // ascending hit list
  uint64_t nr_results = 0;          // predicate passes (reference nrResults)
  const LmParams *params = nullptr;
  const int32_t *hkls = nullptr;
  int nhkls = 0;
  const double *image = nullptr;    // float64 image (converted to float for fitting)
  const float *image_f = nullptr;   // or directly a float image
  int num_procs = 1;
  lm_ctx *ctx = nullptr;            // set: large candidate lists merge on its device
  int image_nr = 0;                 // > 0: streaming output with an ImageNr column
};

// Merge + write into new files with the reference headers.  Returns the
// number of unique orientations, or -1 if a file cannot be opened.
int write_solutions(const PostInput &in, const char *sol_path, const char *spots_path);

// Merge + the rows of both files as text, in grain order, for the caller to
// append (streaming daemon: several frames are formatted at once and written
// in order).  Returns the number of unique orientations, or -1.
int format_solutions(const PostInput &in, std::string *sol_rows, std::string *spot_rows);

// Header lines of the two output files (src/LaueMatchingCPU.c:753-770).
void write_solution_headers(FILE *sol, FILE *spots, bool image_column);

