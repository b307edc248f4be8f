// lm_params.h - parameter file of the LaueMatching programs.
//
// Same keys, prefix matching, value syntax and C defaults as the reference
// readers (src/LaueMatchingCPU.c:60-209, src/LaueMatchingGPUStream.cu:576-717).
#pragma once

#include <string>

struct LmParams {
  double lattice[6] = {0, 0, 0, 0, 0, 0};
  double tol_latc[6] = {0, 0, 0, 0, 0, 0};
  double tol_c_over_a = 0;
  int space_group = 225;
  double P[3] = {0, 0, 0};
  double R[3] = {0, 0, 0};
  double pxX = 0, pxY = 0;
  int nrPxX = 0, nrPxY = 0;
  double Elo = 5, Ehi = 30;
  int max_spots = 500;
  int min_nr_spots = 5;
  int min_good_spots = 5;
  int do_fwd = 1;
  double min_intensity = 1000.0;
  double max_angle = 2.0;
  unsigned long long batch_size = 1000000;
  std::string forward_file;
  std::string result_dir = "results_stream";  // daemon default
  bool nelder_mead = false;
};

// Returns 0 on success; -1 if the file cannot be opened; -2 if NrPxX/NrPxY
// or the lattice constant a is missing (the reference's validation).
int lm_read_params(const char *path, LmParams *p);
