// lm_hostio.h - file formats of the LaueMatching programs (host side).
//
//   orientation DB : N x 9 float64 row-major, no header; N = floor(size/72)
//                    (src/LaueMatchingCPU.c:241-287; mmap for /dev/shm paths)
//   HKL list       : text, first three integers per line, at most 200000 lines
//                    (src/LaueMatchingCPU.c:289-312)
//   image          : nrPxX*nrPxY float64, row-major, X fastest
//                    (src/LaueMatchingCPU.c:314-353)
//   ForwardFile    : N x (1+2M) uint16 records, orientation-major, no header;
//                    written without O_TRUNC, read by mmap for /dev/shm paths
//                    (src/LaueMatchingCPU.c:421-486, 548-576, 693-726)
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "lauematch.h"

struct HostBuffer {
  void *data = nullptr;
  size_t size = 0;
  bool mapped = false;
  HostBuffer() = default;
  HostBuffer(const HostBuffer &) = delete;
  HostBuffer &operator=(const HostBuffer &) = delete;
  ~HostBuffer();
};

bool is_shm_path(const char *path);
// Whole file, mmap'ed read-only for /dev/shm paths, read into memory otherwise.
int read_file(const char *path, HostBuffer *out);
int read_hkls(const char *path, std::vector<int32_t> *hkls);
int read_image_f64(const char *path, size_t npx, std::vector<double> *img);
// Rows [0, n) of the DB's table, as records, at byte offset 0 of path.
int write_forward_file(const char *path, lm_db *db, uint64_t n, int max_spots);
