/*
 * lmo_cpu.c - command-line driver for the CPU oracle (TEST INFRASTRUCTURE).
 *
 * Mirrors the hot-path part of the reference LaueMatchingCPU main()
 * (src/LaueMatchingCPU.c:48-733 and the collection at :772-788):
 *
 *   lmo_cpu params.txt orient.bin hkls.csv image.bin nCPUs [hits_out.txt]
 *
 * DoFwd=1 (or a missing ForwardFile) runs the forward simulation, scores the
 * float64 image and writes the ForwardFile; DoFwd=0 reads the ForwardFile and
 * scores the uint8-quantized image.  Prints the reference's hot-path timer
 * ("time2") and nrResults, and optionally writes the ascending hit list
 * ("rowNr score" with %.17g).  Merge, fit and the solutions writer are not
 * part of this driver.  Environment LMO_MAX_ORIENTS bounds the number of
 * orientations processed (a prefix of the DB) for timed samples.
 */
#define _GNU_SOURCE
#include "lm_oracle.h"

#include <fcntl.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <unistd.h>

static void *read_all(const char *fn, size_t *sz) {
  FILE *f = fopen(fn, "rb");
  if (!f) return NULL;
  fseek(f, 0, SEEK_END);
  long n = ftell(f);
  rewind(f);
  void *buf = malloc(n > 0 ? (size_t)n : 1);
  if (buf && n > 0 && fread(buf, 1, (size_t)n, f) != (size_t)n) {
    free(buf);
    buf = NULL;
  }
  fclose(f);
  *sz = (size_t)(n > 0 ? n : 0);
  return buf;
}

int main(int argc, char **argv) {
  if (argc != 6 && argc != 7) {
    puts("usage: lmo_cpu params.txt orient.bin hkls.csv image.bin nCPUs [hits_out.txt]");
    return 0;
  }
  lmo_params p;
  if (lmo_parse_params(argv[1], &p) != 0) {
    fprintf(stderr, "bad parameter file %s\n", argv[1]);
    return 1;
  }
  lmo_geom g;
  lmo_geom_from_params(&p, &g);
  size_t osz;
  double *orients = (double *)read_all(argv[2], &osz);
  if (!orients) {
    fprintf(stderr, "cannot read %s\n", argv[2]);
    return 1;
  }
  /* src/LaueMatchingCPU.c:254 */
  size_t n = (size_t)((double)osz / (double)(9 * sizeof(double)));
  const char *lim = getenv("LMO_MAX_ORIENTS");
  if (lim && strtoull(lim, NULL, 10) > 0 && strtoull(lim, NULL, 10) < n) n = strtoull(lim, NULL, 10);
  FILE *hf = fopen(argv[3], "r");
  if (!hf) {
    fprintf(stderr, "cannot read %s\n", argv[3]);
    return 1;
  }
  int cap = 200000, nh = 0; /* MaxNHKLS, src/LaueMatchingHeaders.h:35 */
  int *hkls = calloc((size_t)cap * 3, sizeof(int));
  char line[1000];
  while (nh < cap && fgets(line, sizeof line, hf)) {
    sscanf(line, "%d %d %d", &hkls[3 * nh], &hkls[3 * nh + 1], &hkls[3 * nh + 2]);
    nh++;
  }
  fclose(hf);
  size_t npx = (size_t)p.nrPxX * p.nrPxY, isz;
  double *img = (double *)read_all(argv[4], &isz);
  if (!img || isz < npx * sizeof(double)) {
    fprintf(stderr, "cannot read image %s\n", argv[4]);
    return 1;
  }
  uint8_t *u8 = malloc(npx);
  double imScale = lmo_quantize_u8(img, npx, u8);
  int ncpu = atoi(argv[5]);
  if (ncpu < 1) {
    fprintf(stderr, "invalid thread count\n");
    return 1;
  }
  int M = p.maxNrSpots;
  size_t rec = (size_t)(1 + 2 * M);
  double *matched = calloc(n, sizeof(double));
  int doFwd = p.doFwd;
  if (doFwd == 0 && access(p.forwardFile, R_OK) != 0) doFwd = 1; /* :421-438 */
  uint16_t *table = malloc(n * rec * sizeof(uint16_t));
  if (!table || !matched) {
    fprintf(stderr, "allocation failed\n");
    return 1;
  }
  long nres;
  double t0;
  if (doFwd) {
    t0 = omp_get_wtime();
    nres = lmo_forward(orients, n, hkls, nh, &g, M, table, img, p.minNrSpots, p.minIntensity,
                       matched, ncpu);
    int fd = open(p.forwardFile, O_CREAT | O_WRONLY, S_IRUSR | S_IWUSR);
    if (fd >= 0) {
      size_t bytes = n * rec * sizeof(uint16_t), done = 0;
      while (done < bytes) {
        ssize_t w = pwrite(fd, (char *)table + done, bytes - done, (off_t)done);
        if (w <= 0) break;
        done += (size_t)w;
      }
      close(fd);
    }
  } else {
    int fd = open(p.forwardFile, O_RDONLY);
    size_t bytes = n * rec * sizeof(uint16_t), done = 0;
    while (fd >= 0 && done < bytes) {
      ssize_t r = pread(fd, (char *)table + done, bytes - done, (off_t)done);
      if (r <= 0) break;
      done += (size_t)r;
    }
    if (fd >= 0) close(fd);
    if (done < bytes) {
      fprintf(stderr, "short ForwardFile read\n");
      return 1;
    }
    t0 = omp_get_wtime();
    nres = lmo_compare_u8(table, n, M, u8, p.nrPxX, imScale, p.minNrSpots, p.minIntensity,
                          matched, NULL, NULL, ncpu);
  }
  double time2 = omp_get_wtime() - t0;
  printf("Finished comparing, time elapsed after comparing with forward simulation: %lf "
         "seconds.\n",
         time2);
  printf("orientations %zu hkls %d threads %d doFwd %d imScale %.17g\n", n, nh, ncpu, doFwd,
         imScale);
  printf("Initial solutions: %ld\n", nres);
  if (argc == 7) {
    FILE *o = fopen(argv[6], "w");
    for (size_t i = 0; o && i < n; i++)
      if (matched[i] != 0) fprintf(o, "%zu %.17g\n", i, matched[i]);
    if (o) fclose(o);
  }
  free(table);
  free(matched);
  free(u8);
  free(img);
  free(hkls);
  free(orients);
  return 0;
}
