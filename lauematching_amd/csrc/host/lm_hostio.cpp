// lm_hostio.cpp - see lm_hostio.h.
#include "lm_hostio.h"

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>

HostBuffer::~HostBuffer() {
  if (!data) return;
  if (mapped)
    munmap(data, size);
  else
    free(data);
}

bool is_shm_path(const char *path) { return strncmp(path, "/dev/shm", 8) == 0; }

int read_file(const char *path, HostBuffer *out) {
  int fd = open(path, O_RDONLY);
  if (fd < 0) return -1;
  struct stat st;
  if (fstat(fd, &st) != 0) {
    close(fd);
    return -1;
  }
  out->size = (size_t)st.st_size;
  if (is_shm_path(path) && out->size) {
    void *p = mmap(nullptr, out->size, PROT_READ, MAP_SHARED, fd, 0);
    close(fd);
    if (p == MAP_FAILED) return -1;
    out->data = p;
    out->mapped = true;
    return 0;
  }
  out->data = malloc(out->size ? out->size : 1);
  if (!out->data) {
    close(fd);
    return -1;
  }
  size_t got = 0;
  while (got < out->size) {
    ssize_t r = pread(fd, (char *)out->data + got, out->size - got, (off_t)got);
    if (r <= 0) break;
    got += (size_t)r;
  }
  close(fd);
  return got == out->size ? 0 : -1;
}

int read_hkls(const char *path, std::vector<int32_t> *hkls) {
  FILE *f = fopen(path, "r");
  if (!f) return -1;
  char line[1000];
  hkls->clear();
  const int cap = 200000;  // MaxNHKLS, src/LaueMatchingHeaders.h:35
  int n = 0;
  while (fgets(line, sizeof line, f)) {
    if (n >= cap) {
      printf("Warning: more than %d HKLs in file, truncating.\n", cap);
      break;
    }
    int h = 0, k = 0, l = 0;
    sscanf(line, "%d %d %d", &h, &k, &l);
    hkls->push_back(h);
    hkls->push_back(k);
    hkls->push_back(l);
    ++n;
  }
  fclose(f);
  return 0;
}

int read_image_f64(const char *path, size_t npx, std::vector<double> *img) {
  FILE *f = fopen(path, "rb");
  if (!f) return -1;
  img->assign(npx, 0.0);
  size_t got = fread(img->data(), sizeof(double), npx, f);
  fclose(f);
  return got == npx ? 0 : -2;
}

int write_forward_file(const char *path, lm_db *db, uint64_t n, int max_spots) {
  int fd = open(path, O_CREAT | O_WRONLY, S_IRUSR | S_IWUSR);
  if (fd < 0) return -1;
  const uint64_t rec = (uint64_t)(1 + 2 * max_spots);
  const uint64_t chunk = 1u << 20;
  std::vector<uint16_t> buf(std::min<uint64_t>(chunk, n) * rec);
  int rc = 0;
  for (uint64_t r0 = 0; r0 < n && rc == 0; r0 += chunk) {
    const uint64_t rows = std::min(chunk, n - r0);
    if (lm_db_export_table(db, r0, rows, buf.data()) != LM_OK) {
      rc = -2;
      break;
    }
    const size_t bytes = rows * rec * 2;
    size_t done = 0;
    while (done < bytes) {
      ssize_t w = pwrite(fd, (const char *)buf.data() + done, bytes - done,
                         (off_t)(r0 * rec * 2 + done));
      if (w < 0) {
        rc = -3;
        break;
      }
      done += (size_t)w;
    }
  }
  if (fsync(fd) < 0) fprintf(stderr, "WARNING: fsync(%s) failed: %s\n", path, strerror(errno));
  close(fd);
  return rc;
}
