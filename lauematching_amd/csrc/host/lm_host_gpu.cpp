// lm_host_gpu.cpp - bin/LaueMatchingGPU: the single-image indexing program.
//
//   LaueMatchingGPU params.txt orient.bin hkls.csv image.bin nCPUs
//
// Same command line, parameter file, input formats and ForwardFile handling as
// the reference program (src/LaueMatchingGPU.cu:107-792, following
// src/LaueMatchingCPU.c:48-733), with the hot path on the MI355X through
// liblauematch.so:
//   DoFwd=1 (or a missing ForwardFile): forward simulation on the GPU, the
//     table written to ForwardFile, the float64 image scored
//     (src/LaueMatchingCPU.c:595-667, 683-689);
//   DoFwd=0: the ForwardFile table imported to HBM, the uint8-quantized image
//     scored (src/LaueMatchingCPU.c:668-689).
// The ascending hit list (src/LaueMatchingCPU.c:772-788) is written to
// <image>.hits.txt ("rowNr nSpots totInt score", %.17g) and handed to the
// CPU post-processing (lm_postproc.h).
#include <omp.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "lauematch.h"
#include "lm_hostio.h"
#include "lm_params.h"
#include "lm_postproc.h"

namespace {

void usage() {
  puts("LaueMatching on the GPU (MI355X)\n"
       "Arguments: \n"
       "* \tparameterFile (text)\n"
       "* \tbinary files for candidate orientation list [double]\n"
       "* \ttext of valid_hkls (preferably sorted on f2), space separated\n"
       "* \tbinary file for image [double]\n"
       "* \tnumber of CPU cores to use: \n\n"
       "Parameter file with the following parameters: \n"
       "\t\t* LatticeParameter (in nm and degrees),\n"
       "\t\t* tol_latC (in %, 6 values),\n"
       "\t\t* tol_c_over_a (in %, 1 value),\n"
       "\t\t* SpaceGroup,\n"
       "\t\t* P_Array, R_Array, PxX, PxY, NrPxX, NrPxY,\n"
       "\t\t* Elo, Ehi, MaxNrLaueSpots, ForwardFile, DoFwd,\n"
       "\t\t* MinNrSpots, MinIntensity, MaxAngle.\n");
}

int fail(const char *what) {
  fprintf(stderr, "FATAL: %s: %s\n", what, lm_last_error());
  return 1;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc != 6) {
    usage();
    return 0;
  }
  const double t_start = omp_get_wtime();
  puts("Reading parameter file");
  fflush(stdout);
  LmParams p;
  int prc = lm_read_params(argv[1], &p);
  if (prc == -1) {
    printf("Could not open parameter file %s.\n", argv[1]);
    return 1;
  }
  if (prc == -2) {
    fprintf(stderr, "FATAL: NrPxX/NrPxY or LatticeParameter missing (%d x %d). "
            "Check the parameter file.\n", p.nrPxX, p.nrPxY);
    return 1;
  }
  puts("Parameters read");
  const int numProcs = atoi(argv[5]);
  if (numProcs < 1) {
    fprintf(stderr, "FATAL: invalid thread count '%s' (need >= 1).\n", argv[5]);
    return 1;
  }
  puts("Reading orientations");
  double t0 = omp_get_wtime();
  HostBuffer orient;
  if (read_file(argv[2], &orient) != 0) {
    puts("Could not read orientation file, exiting.");
    return 1;
  }
  const uint64_t n = (uint64_t)((double)orient.size / (double)(9 * sizeof(double)));
  printf("%llu Orientations %s, took %lf seconds, now reading hkls\n", (unsigned long long)n,
         orient.mapped ? "mapped into memory" : "read", omp_get_wtime() - t0);
  std::vector<int32_t> hkls;
  if (read_hkls(argv[3], &hkls) != 0) {
    printf("Could not read hkl file %s.\n", argv[3]);
    return 1;
  }
  const int nhkls = (int)(hkls.size() / 3);
  const size_t npx = (size_t)p.nrPxX * p.nrPxY;
  printf("%d hkls read. \nNow reading image file %s, number of pixels %zu. \nminNrSpots: %d, "
         "MinIntensity: %d\n", nhkls, argv[4], npx, p.min_nr_spots, (int)p.min_intensity);
  std::vector<double> image;
  int irc = read_image_f64(argv[4], npx, &image);
  if (irc != 0) {
    printf(irc == -1 ? "Could not read image file %s.\n"
                     : "Error: Unexpected end of file while reading image %s.\n", argv[4]);
    return 1;
  }
  size_t nonzero = 0;
  for (double v : image) nonzero += v > 0;
  printf("Pixels with intensity: %zu\n", nonzero);
  std::vector<uint8_t> u8(npx);
  double imScale = 1.0;
  lm_quantize_u8(image.data(), npx, u8.data(), &imScale);

  t0 = omp_get_wtime();
  lm_ctx *ctx = nullptr;
  if (lm_ctx_create(0, 0, &ctx) != LM_OK) return fail("GPU context");
  printf("GPU context init: %lf s\n", omp_get_wtime() - t0);
  lm_geom g;
  if (lm_geom_init(p.lattice, p.space_group, p.P, p.R, p.pxX, p.pxY, p.nrPxX, p.nrPxY, p.Elo,
                   p.Ehi, &g) != LM_OK)
    return fail("geometry");
  lm_db *db = nullptr;
  if (lm_db_create(ctx, (const double *)orient.data, n, hkls.data(), nhkls, &g, p.max_spots,
                   &db) != LM_OK)
    return fail("orientation DB upload");
  printf("Now running using %d threads.\n", numProcs);
  const double start_time = omp_get_wtime();
  int doFwd = p.do_fwd;
  if (doFwd == 0) {
    printf("Trying to see if the forward simulation exists. Looking for %s file.\n",
           p.forward_file.c_str());
    if (access(p.forward_file.c_str(), R_OK) != 0) {
      printf("Could not read the forward simulation file. Running in simulation mode!\n");
      doFwd = 1;
    } else {
      printf("%s file was found. Will not do forward simulation.\n", p.forward_file.c_str());
    }
  } else {
    printf("Forward simulation was requested, will be saved to %s.\n", p.forward_file.c_str());
  }
  lm_hits hits{};
  if (doFwd) {
    lm_fwd_stats st{};
    if (lm_db_forward(db, LM_FWD_FP64, &st) != LM_OK) return fail("forward simulation");
    printf("GPU forward simulation: %llu spots, %lf s\n", (unsigned long long)st.n_spots,
           st.ms / 1e3);
    if (!p.forward_file.empty() &&
        write_forward_file(p.forward_file.c_str(), db, n, p.max_spots) != 0)
      printf("Could not write to output file %s\n", p.forward_file.c_str());
    if (lm_score_f64(db, image.data(), p.min_nr_spots, p.min_intensity, &hits) != LM_OK)
      return fail("scoring");
  } else {
    double tr = omp_get_wtime();
    HostBuffer fwd;
    const uint64_t need = n * (uint64_t)(1 + 2 * p.max_spots) * 2;
    if (read_file(p.forward_file.c_str(), &fwd) != 0 || fwd.size < need) {
      printf("Error: Failed to read forward simulation file.\n");
      return 1;
    }
    printf("Time elapsed to read the forward simulation: %lf s\n", omp_get_wtime() - tr);
    if (lm_db_import_table(db, (const uint16_t *)fwd.data, n) != LM_OK)
      return fail("ForwardFile import");
    if (lm_score(db, u8.data(), imScale, p.min_nr_spots, p.min_intensity, &hits) != LM_OK)
      return fail("scoring");
  }
  printf("NrMatches: %llu\n", (unsigned long long)hits.nr_results);
  const double time2 = omp_get_wtime() - start_time;
  printf("Finished comparing, time elapsed after comparing with forward simulation: %lf "
         "seconds.\nSearching for unique solutions.\n", time2);
  fflush(stdout);

  std::string hfn = std::string(argv[4]) + ".hits.txt";
  if (FILE *hf = fopen(hfn.c_str(), "w")) {
    fprintf(hf, "%%rowNr\tnSpots\ttotInt\tscore\n");
    for (uint64_t i = 0; i < hits.count; ++i)
      fprintf(hf, "%llu\t%d\t%.17g\t%.17g\n", (unsigned long long)hits.ids[i], hits.nspots[i],
              hits.totint[i], hits.score[i]);
    fclose(hf);
  }
  PostInput in;
  in.orients = (const double *)orient.data;
  in.n_orients = n;
  in.hits = &hits;
  in.nr_results = hits.nr_results;
  in.params = &p;
  in.hkls = hkls.data();
  in.nhkls = nhkls;
  in.image = image.data();
  in.num_procs = numProcs;
  in.ctx = ctx;
  in.image_nr = 0;
  const std::string base = argv[4];
  int rc = write_solutions(in, (base + ".solutions.txt").c_str(), (base + ".spots.txt").c_str());
  lm_hits_free(&hits);
  lm_db_destroy(db);
  lm_ctx_destroy(ctx);
  printf("Initial solutions: %llu Unique Orientations: %d\nTotal time: %lf seconds.\n",
         (unsigned long long)in.nr_results, rc < 0 ? 0 : rc, omp_get_wtime() - t_start);
  return rc < 0 ? 1 : 0;
}
