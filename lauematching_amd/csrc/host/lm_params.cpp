// lm_params.cpp - parameter-file reader (see lm_params.h).
#include "lm_params.h"

#include <cstdio>
#include <cstring>

namespace {

bool starts(const char *line, const char *key) { return strncmp(line, key, strlen(key)) == 0; }

}  // namespace

int lm_read_params(const char *path, LmParams *p) {
  FILE *f = fopen(path, "r");
  if (!f) return -1;
  char line[1000], tag[1000], word[1000];
  while (fgets(line, sizeof line, f)) {
    // Keys in the order the reference tests them; the first prefix match wins.
    if (starts(line, "LatticeParameter")) {
      sscanf(line, "%999s %lf %lf %lf %lf %lf %lf", tag, &p->lattice[0], &p->lattice[1],
             &p->lattice[2], &p->lattice[3], &p->lattice[4], &p->lattice[5]);
    } else if (starts(line, "P_Array")) {
      sscanf(line, "%999s %lf %lf %lf", tag, &p->P[0], &p->P[1], &p->P[2]);
    } else if (starts(line, "R_Array")) {
      sscanf(line, "%999s %lf %lf %lf", tag, &p->R[0], &p->R[1], &p->R[2]);
    } else if (starts(line, "tol_c_over_a")) {
      sscanf(line, "%999s %lf", tag, &p->tol_c_over_a);
    } else if (starts(line, "PxX")) {
      sscanf(line, "%999s %lf", tag, &p->pxX);
    } else if (starts(line, "PxY")) {
      sscanf(line, "%999s %lf", tag, &p->pxY);
    } else if (starts(line, "Elo")) {
      sscanf(line, "%999s %lf", tag, &p->Elo);
    } else if (starts(line, "Ehi")) {
      sscanf(line, "%999s %lf", tag, &p->Ehi);
    } else if (starts(line, "DoFwd")) {
      sscanf(line, "%999s %d", tag, &p->do_fwd);
    } else if (starts(line, "NrPxX")) {
      sscanf(line, "%999s %d", tag, &p->nrPxX);
    } else if (starts(line, "NrPxY")) {
      sscanf(line, "%999s %d", tag, &p->nrPxY);
    } else if (starts(line, "MaxNrLaueSpots")) {
      sscanf(line, "%999s %d", tag, &p->max_spots);
    } else if (starts(line, "BatchSize")) {
      sscanf(line, "%999s %llu", tag, &p->batch_size);
      if (p->batch_size == 0) p->batch_size = 1000000;
    } else if (starts(line, "MinNrSpots")) {
      sscanf(line, "%999s %d", tag, &p->min_nr_spots);
    } else if (starts(line, "SpaceGroup")) {
      sscanf(line, "%999s %d", tag, &p->space_group);
    } else if (starts(line, "MinIntensity")) {
      sscanf(line, "%999s %lf", tag, &p->min_intensity);
    } else if (starts(line, "MaxAngle")) {
      sscanf(line, "%999s %lf", tag, &p->max_angle);
    } else if (starts(line, "ForwardFile")) {
      if (sscanf(line, "%999s %999s", tag, word) == 2) p->forward_file = word;
    } else if (starts(line, "tol_LatC")) {
      sscanf(line, "%999s %lf %lf %lf %lf %lf %lf", tag, &p->tol_latc[0], &p->tol_latc[1],
             &p->tol_latc[2], &p->tol_latc[3], &p->tol_latc[4], &p->tol_latc[5]);
    } else if (starts(line, "Optimizer")) {
      if (sscanf(line, "%999s %999s", tag, word) == 2)
        p->nelder_mead = strncmp(word, "NelderMead", 10) == 0;
    } else if (starts(line, "ResultDir")) {
      if (sscanf(line, "%999s %999s", tag, word) == 2) p->result_dir = word;
    } else if (starts(line, "MinGoodSpots")) {
      sscanf(line, "%999s %d", tag, &p->min_good_spots);
    }
  }
  fclose(f);
  if (p->tol_c_over_a != 0)
    for (double &v : p->tol_latc) v = 0;
  if (p->nrPxX <= 0 || p->nrPxY <= 0 || p->lattice[0] == 0.0) return -2;
  return 0;
}
