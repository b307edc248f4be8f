// lm_host_stream.cpp - bin/LaueMatchingGPUStream: the streaming daemon.
//
//   LaueMatchingGPUStream params.txt orient.bin hkls.csv nCPUs
//                         [params2.txt orient2.bin hkls2.csv ...]
//
// Contract of the reference daemon (src/LaueMatchingGPUStream.cu:536-1329):
//   * one detector size per daemon (NrPxX x NrPxY from the parameter file);
//   * TCP port 60517 (LM_STREAM_PORT overrides it for tests), frames of
//     uint16 image_num (little endian) + float32[NrPxX*NrPxY] (wire v1,
//     scripts/laue_image_server.py:217-219), 5 s receive timeout, a bounded
//     queue that blocks the sender (backpressure);
//   * results appended to <ResultDir>/solutions.txt and spots.txt with an
//     ImageNr column (none for image 0, as in the reference), solutions
//     written when the overlap has at least MinGoodSpots spots;
//   * first SIGINT/SIGTERM: stop accepting, finish every received image, close
//     the files; a second signal exits at once.
// Wire v2 (no reference counterpart; the next port, 60518 by default): each
// frame carries its own header (24 bytes, little endian): magic "LMF2",
// uint32 image number, uint16 DB index, uint16 pixel type (0 float32,
// 1 float64, 2 uint16, 3 uint8), uint32 width, uint32 height, uint32 0; then
// width * height pixels.  DB index k selects the k-th (params, orientations,
// hkls) triplet of the command line; width and height must be that DB's
// detector.  A malformed header closes the connection.  Results of DB 0 go
// to <ResultDir>/solutions.txt and spots.txt as in v1, those of DB k > 0 to
// <ResultDir of k>/solutions_db<k>.txt and spots_db<k>.txt.  v1 frames keep
// going to DB 0 unchanged.
// The per-image GPU work goes through the continuous-batching server
// (lm_server_*): frames that arrive while others are scored join the next
// batch, instead of the reference's one-kernel-at-a-time GPU gate.  Scores
// follow the CPU DoFwd=0 semantics: the frame is converted to float64 and
// uint8-quantized (src/LaueMatchingCPU.c:355-377) before scoring.
#include <arpa/inet.h>
#include <netinet/in.h>
#include <omp.h>
#include <signal.h>
#include <sys/socket.h>
#include <sys/stat.h>
#include <unistd.h>

#include <atomic>
#include <cerrno>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "lauematch.h"
#include "lm_hostio.h"
#include "lm_params.h"
#include "lm_postproc.h"

namespace {

constexpr int kPort = 60517;  // src/LaueMatchingGPUStream.cu:42
constexpr int kMaxBatch = 256;  // images per launch of the server
constexpr uint32_t kV2Magic = 0x32464D4Cu;  // "LMF2"

std::atomic<bool> g_running{true};
std::atomic<int> g_listen_fd{-1}, g_listen_fd2{-1};

void on_signal(int sig) {
  if (g_running.exchange(false)) {
    const char msg[] = "\nCaught signal, requesting shutdown...\n";
    (void)!write(STDOUT_FILENO, msg, sizeof msg - 1);
    for (const int fd : {g_listen_fd.load(), g_listen_fd2.load()})
      if (fd >= 0) shutdown(fd, SHUT_RDWR);
  } else {
    const char msg[] = "\nForced exit.\n";
    (void)!write(STDOUT_FILENO, msg, sizeof msg - 1);
    _exit(1);
  }
  (void)sig;
}

struct Job {
  uint32_t image_num;
  int db;
  uint64_t ticket;
  int type;                  // pixel type of raw: 0 float32, 1 float64, 2 uint16, 3 uint8
  std::vector<uint8_t> raw;  // the frame's pixels as received (for the fit), from the pool
  double t_arrival;          // frame fully received
  double t_submit;           // quantized, handed to the server
};

// One DB of the daemon: its parameter file, orientations, hkls, table and
// output files.
struct DbSlot {
  LmParams p, fitp;
  HostBuffer orient;
  uint64_t n = 0;
  std::vector<int32_t> hkls;
  int nhkls = 0;
  lm_db *db = nullptr;
  int32_t server_id = 0;
  size_t npx = 0;
  FILE *sol = nullptr, *spots = nullptr;
};

// A scored frame with hits, waiting for the merge and fit.  seq orders the
// result rows: they are appended in the order the frames were submitted.
struct PostJob {
  uint64_t seq;
  Job job;
  lm_hits hits;
};

struct PostDone {
  Job job;
  int nsol;
  std::string sol, spots;
};

struct Daemon {
  std::vector<std::unique_ptr<DbSlot>> dbs;
  lm_server *server = nullptr;
  std::mutex mu;
  std::condition_variable cv;
  std::deque<Job> jobs;
  int open_clients = 0;
  // received frames waiting for a quantize worker; `unsubmitted` counts them
  // (under mu) until their job is queued or dropped
  struct Frame {
    int db;
    uint32_t num;
    int type;
    std::vector<uint8_t> raw;
    double t_arrival;
    double recv_ms;
  };
  std::mutex in_mu;
  std::condition_variable in_cv;
  std::deque<Frame> intake;
  bool in_closed = false;
  int unsubmitted = 0;
  std::vector<std::thread> handlers;
  // Frame buffers handed from the connection handlers to the fit and back.
  // They are kept for reuse (no allocation, zero fill or page faults per
  // frame) up to pool_bytes in all; a handler that needs more waits until a
  // frame is done, which holds the sender back as the reference's bounded
  // frame queue does (src/LaueMatchingGPUStream.cu:45).  Frames stay in
  // their buffer from arrival until the fit, so the frames in flight -- the
  // rate times that time -- must fit: the default 8 GiB holds the server's
  // 512 admitted frames of a 2048 x 2048 float32 detector
  // (LM_STREAM_FRAME_POOL_GB overrides it).
  size_t pool_bytes = size_t(8) << 30;
  std::mutex pool_mu;
  std::condition_variable pool_cv;
  std::vector<std::vector<uint8_t>> pool;  // free buffers
  size_t held = 0;                         // capacity of every buffer, free or in use
  std::vector<uint8_t> take_buffer(size_t bytes) {
    std::unique_lock<std::mutex> lk(pool_mu);
    for (;;) {
      // the smallest free buffer that holds the frame; one more than twice
      // its size is only taken when no new buffer fits under the cap
      int best = -1;
      for (int i = 0; i < (int)pool.size(); ++i)
        if (pool[i].capacity() >= bytes &&
            (best < 0 || pool[i].capacity() < pool[best].capacity()))
          best = i;
      const bool room = held + bytes <= pool_bytes || held == 0;
      if (best >= 0 && (pool[best].capacity() <= 2 * bytes || !room)) {
        std::vector<uint8_t> b = std::move(pool[best]);
        pool.erase(pool.begin() + best);
        b.resize(bytes);
        return b;
      }
      if (room) {
        held += bytes;
        lk.unlock();
        std::vector<uint8_t> b(bytes);
        return b;
      }
      int small = -1;  // a free buffer too small for the frame: make room
      for (int i = 0; i < (int)pool.size(); ++i)
        if (pool[i].capacity() < bytes) small = i;
      if (small >= 0) {
        held -= pool[small].capacity();
        pool.erase(pool.begin() + small);
        continue;
      }
      pool_cv.wait(lk);
    }
  }
  void give_buffer(std::vector<uint8_t> &&b) {
    {
      std::lock_guard<std::mutex> lk(pool_mu);
      if (b.capacity()) pool.push_back(std::move(b));
    }
    pool_cv.notify_one();
  }
  // host time spent per stage over all frames (ns): payload receive (from
  // its header on, so it includes waiting for the sender), quantization,
  // hand-over to the server (the copy into its admission slot)
  std::atomic<int64_t> ns_recv{0}, ns_quant{0}, ns_submit{0};
  std::atomic<uint64_t> frames{0};
  // merge + fit workers and the in-order writer of their rows
  std::mutex post_mu;
  std::condition_variable post_cv;
  std::deque<PostJob> post;
  bool post_closed = false;
  std::mutex out_mu;
  std::map<uint64_t, PostDone> done;
  uint64_t next_out = 0;
};

bool recv_all(int fd, void *buf, size_t len) {
  size_t got = 0;
  while (got < len) {
    const ssize_t n = recv(fd, (char *)buf + got, len - got, MSG_WAITALL);
    if (n < 0) {
      if (errno == EINTR) continue;
      if (errno == EAGAIN || errno == EWOULDBLOCK) {
        if (!g_running) return false;
        continue;
      }
      return false;
    }
    if (n == 0) return false;
    got += (size_t)n;
  }
  return true;
}

constexpr size_t kPixelBytes[4] = {4, 8, 2, 1};

// The uint8 quantization of a frame (src/LaueMatchingCPU.c:355-377) from its
// pixels as received.  Float frames go through lm_quantize_u8(_f32).  An
// integer frame maps each value v to the level the quantizer gives v when
// the frame's largest value is mx; that map is the quantization of the
// values 0..mx themselves (largest mx), so one call of the library's
// quantizer on them gives the table and the scale.  Returns true, with u8
// left unwritten, for a uint8 frame whose table is the identity (largest
// value 255, or none above 0): its levels are its own bytes.
bool quantize_frame(int type, const uint8_t *raw, size_t npx, uint8_t *u8, double *scale) {
  if (type == 0) {
    lm_quantize_u8_f32((const float *)raw, npx, u8, scale);
    return false;
  }
  if (type == 1) {
    lm_quantize_u8((const double *)raw, npx, u8, scale);
    return false;
  }
  uint32_t mx = 0;
  if (type == 2) {
    const uint16_t *v = (const uint16_t *)raw;
    for (size_t i = 0; i < npx; ++i) mx = v[i] > mx ? v[i] : mx;
  } else {
    for (size_t i = 0; i < npx; ++i) mx = raw[i] > mx ? raw[i] : mx;
  }
  std::vector<float> vals(mx + 1);
  for (uint32_t v = 0; v <= mx; ++v) vals[v] = (float)v;
  std::vector<uint8_t> lut(mx + 1);
  lm_quantize_u8_f32(vals.data(), mx + 1, lut.data(), scale);
  if (type == 2) {
    const uint16_t *v = (const uint16_t *)raw;
    for (size_t i = 0; i < npx; ++i) u8[i] = lut[v[i]];
    return false;
  }
  bool identity = true;
  for (uint32_t v = 0; v <= mx; ++v) identity = identity && lut[v] == v;
  if (identity) return true;
  for (size_t i = 0; i < npx; ++i) u8[i] = lut[raw[i]];
  return false;
}

// Quantize one frame, submit it to the server for DB k and queue it for the
// hit list and post-processing.  raw moves into the job.
bool submit_frame(Daemon *d, int k, uint32_t num, int type, std::vector<uint8_t> &raw,
                  std::vector<uint8_t> &u8, double t_arrival, double recv_ms) {
  DbSlot &s = *d->dbs[k];
  u8.resize(s.npx);
  double scale = 1.0;
  const double t_q0 = omp_get_wtime();
  const bool same = quantize_frame(type, raw.data(), s.npx, u8.data(), &scale);
  const double t_submit = omp_get_wtime();
  d->ns_quant += (int64_t)((t_submit - t_q0) * 1e9);
  uint64_t ticket = 0;
  if (lm_server_submit(d->server, s.server_id, same ? raw.data() : u8.data(), scale,
                       s.p.min_nr_spots,
                       s.p.min_intensity, &ticket) != LM_OK) {
    fprintf(stderr, "submit failed for image %u: %s\n", num, lm_last_error());
    d->give_buffer(std::move(raw));
    {
      std::lock_guard<std::mutex> lk(d->mu);
      --d->unsubmitted;
    }
    d->cv.notify_one();
    return false;
  }
  const double t_handed = omp_get_wtime();
  d->ns_submit += (int64_t)((t_handed - t_submit) * 1e9);
  ++d->frames;
  printf("Received image %u for DB %d (%zu bytes wire, %zu pixels; receive %.2f ms, queued "
         "%.2f ms, quantize %.2f ms, hand-over %.2f ms)\n", num, k, raw.size(), s.npx, recv_ms,
         (t_q0 - t_arrival) * 1e3, (t_submit - t_q0) * 1e3, (t_handed - t_submit) * 1e3);
  fflush(stdout);
  {
    std::lock_guard<std::mutex> lk(d->mu);
    d->jobs.push_back(Job{num, k, ticket, type, std::move(raw), t_arrival, t_submit});
    --d->unsubmitted;
  }
  d->cv.notify_one();
  return true;
}

// Hands a received frame to the quantize workers, so that the connection
// goes back to receiving at once.
void queue_frame(Daemon *d, int k, uint32_t num, int type, std::vector<uint8_t> &&raw,
                 double t_arrival, double recv_ms) {
  {
    std::lock_guard<std::mutex> lk(d->mu);
    ++d->unsubmitted;
  }
  {
    std::lock_guard<std::mutex> lk(d->in_mu);
    d->intake.push_back(Daemon::Frame{k, num, type, std::move(raw), t_arrival, recv_ms});
  }
  d->in_cv.notify_one();
}

// Quantize worker: quantizes received frames and submits them to the server.
void quantize_worker(Daemon *d) {
  std::vector<uint8_t> u8;
  for (;;) {
    Daemon::Frame f;
    {
      std::unique_lock<std::mutex> lk(d->in_mu);
      d->in_cv.wait(lk, [&] { return !d->intake.empty() || d->in_closed; });
      if (d->intake.empty()) return;
      f = std::move(d->intake.front());
      d->intake.pop_front();
    }
    submit_frame(d, f.db, f.num, f.type, f.raw, u8, f.t_arrival, f.recv_ms);
  }
}

// v1: uint16 image number + float32 pixels of DB 0's detector
// (scripts/laue_image_server.py:217-219, src/LaueMatchingGPUStream.cu:295-383)
void serve_v1(Daemon *d, int fd) {
  const size_t npx = d->dbs[0]->npx;
  printf("Client handler started (socket %d), wire_payload=%zu bytes\n", fd, npx * 4);
  fflush(stdout);
  while (g_running) {
    uint8_t hdr[2];
    if (!recv_all(fd, hdr, 2)) break;
    const double t0 = omp_get_wtime();
    std::vector<uint8_t> buf = d->take_buffer(npx * 4);
    if (!recv_all(fd, buf.data(), npx * 4)) {
      d->give_buffer(std::move(buf));
      break;
    }
    const double t1 = omp_get_wtime();
    d->ns_recv += (int64_t)((t1 - t0) * 1e9);
    const uint16_t num = (uint16_t)(hdr[0] | (hdr[1] << 8));
    queue_frame(d, 0, num, 0, std::move(buf), t1, (t1 - t0) * 1e3);
  }
}

// v2: per-frame header (see the top of this file)
void serve_v2(Daemon *d, int fd) {
  printf("Client handler started (socket %d), wire v2\n", fd);
  fflush(stdout);
  while (g_running) {
    uint8_t h[24];
    if (!recv_all(fd, h, sizeof h)) break;
    auto u32 = [&](int o) {
      return (uint32_t)h[o] | ((uint32_t)h[o + 1] << 8) | ((uint32_t)h[o + 2] << 16) |
             ((uint32_t)h[o + 3] << 24);
    };
    const uint32_t magic = u32(0), num = u32(4), W = u32(12), H = u32(16), resv = u32(20);
    const int k = h[8] | (h[9] << 8);
    const int type = h[10] | (h[11] << 8);
    if (magic != kV2Magic || resv != 0 || type > 3 || k >= (int)d->dbs.size() ||
        (int64_t)W != d->dbs[k]->p.nrPxX || (int64_t)H != d->dbs[k]->p.nrPxY) {
      fprintf(stderr, "wire v2: bad frame header (magic %08x, DB %d, type %d, %u x %u); "
              "closing the connection\n", magic, k, type, W, H);
      break;
    }
    const size_t npx = (size_t)W * H;
    const double t0 = omp_get_wtime();
    std::vector<uint8_t> buf = d->take_buffer(npx * kPixelBytes[type]);
    if (!recv_all(fd, buf.data(), buf.size())) {
      d->give_buffer(std::move(buf));
      break;
    }
    const double t1 = omp_get_wtime();
    d->ns_recv += (int64_t)((t1 - t0) * 1e9);
    queue_frame(d, k, num, type, std::move(buf), t1, (t1 - t0) * 1e3);
  }
}

void handle_client(Daemon *d, int fd, bool v2) {
  timeval to{5, 0};
  setsockopt(fd, SOL_SOCKET, SO_RCVTIMEO, &to, sizeof to);
  if (v2)
    serve_v2(d, fd);
  else
    serve_v1(d, fd);
  close(fd);
  {
    std::lock_guard<std::mutex> lk(d->mu);
    --d->open_clients;
  }
  d->cv.notify_one();
  printf("Client handler finished (socket %d)\n", fd);
  fflush(stdout);
}

void accept_loop(Daemon *d, int lfd, bool v2) {
  while (g_running) {
    sockaddr_in cli{};
    socklen_t len = sizeof cli;
    const int fd = accept(lfd, (sockaddr *)&cli, &len);
    if (fd < 0) {
      if (!g_running) break;
      if (errno == EINTR) continue;
      break;
    }
    std::lock_guard<std::mutex> lk(d->mu);
    ++d->open_clients;
    d->handlers.emplace_back(handle_client, d, fd, v2);
  }
}

// Sample frames for the autotuner: the spots of a few rows of the DB's own
// table painted as 5 x 5 blobs on an empty detector (sparse, like
// preprocessed frames), so the layouts are timed on this DB's geometry.
void sample_frames(DbSlot &s, std::vector<std::vector<uint8_t>> &frames) {
  const int M = s.p.max_spots, W = s.p.nrPxX, H = s.p.nrPxY;
  std::vector<uint16_t> rec((size_t)(1 + 2 * M));
  uint64_t x = 0x9e3779b97f4a7c15ull;
  auto next = [&] {
    x ^= x << 13;
    x ^= x >> 7;
    x ^= x << 17;
    return x;
  };
  for (int f = 0; f < 3; ++f) {
    std::vector<uint8_t> img((size_t)W * H, 0);
    for (int g = 0; g < 5 && s.n; ++g) {
      if (lm_db_export_table(s.db, next() % s.n, 1, rec.data()) != LM_OK) break;
      for (int k = 0; k < rec[0]; ++k) {
        const int cx = rec[1 + 2 * k], cy = rec[2 + 2 * k];
        const int amp = 60 + (int)(next() % 196);
        for (int dy = -2; dy <= 2; ++dy)
          for (int dx = -2; dx <= 2; ++dx) {
            const int px = cx + dx, py = cy + dy;
            if (px < 0 || py < 0 || px >= W || py >= H) continue;
            const int v = amp >> (std::abs(dx) + std::abs(dy));
            uint8_t &t = img[(size_t)py * W + px];
            if (v > t) t = (uint8_t)v;
          }
      }
    }
    frames.push_back(std::move(img));
  }
}

// Times the scoring layouts on this DB (single image and batches up to the
// server's width) and keeps the fastest; LM_STREAM_AUTOTUNE=0 keeps the
// default layouts.
int autotune_slot(DbSlot &s) {
  const char *env = getenv("LM_STREAM_AUTOTUNE");
  if (env && atoi(env) == 0) return 0;
  std::vector<std::vector<uint8_t>> frames;
  sample_frames(s, frames);
  std::vector<const uint8_t *> ptrs;
  for (auto &f : frames) ptrs.push_back(f.data());
  lm_tune_report rep{};
  const double t0 = omp_get_wtime();
  if (lm_db_autotune(s.db, ptrs.data(), (int)ptrs.size(), kMaxBatch, 2, &rep) != LM_OK) {
    fprintf(stderr, "FATAL: autotune: %s\n", lm_last_error());
    return 1;
  }
  printf("Autotuned scoring layouts in %.1f s (single image: 2^%d-row super-tiles, walk %d)\n",
         omp_get_wtime() - t0, rep.banded_tile_bits, rep.banded_loads);
  fflush(stdout);
  return 0;
}

// Loads DB k of the command line: parameters, orientations, hkls, the spot
// table (ForwardFile or the GPU forward simulation) and its output files.
int load_slot(lm_ctx *ctx, Daemon &d, int k, const char *par, const char *ori, const char *hkl) {
  auto s = std::make_unique<DbSlot>();
  const int prc = lm_read_params(par, &s->p);
  if (prc == -1) {
    printf("Could not open parameter file %s.\n", par);
    return 1;
  }
  if (prc == -2) {
    fprintf(stderr, "FATAL: NrPxX/NrPxY or LatticeParameter missing in %s.\n", par);
    return 1;
  }
  if (read_file(ori, &s->orient) != 0) {
    printf("Could not read orientation file %s, exiting.\n", ori);
    return 1;
  }
  s->n = (uint64_t)((double)s->orient.size / (double)(9 * sizeof(double)));
  if (read_hkls(hkl, &s->hkls) != 0) {
    printf("Could not read hkl file %s.\n", hkl);
    return 1;
  }
  s->nhkls = (int)(s->hkls.size() / 3);
  s->npx = (size_t)s->p.nrPxX * s->p.nrPxY;
  printf("DB %d: %llu orientations, %d hkls, image %d x %d\n", k, (unsigned long long)s->n,
         s->nhkls, s->p.nrPxX, s->p.nrPxY);
  lm_geom g;
  if (lm_geom_init(s->p.lattice, s->p.space_group, s->p.P, s->p.R, s->p.pxX, s->p.pxY,
                   s->p.nrPxX, s->p.nrPxY, s->p.Elo, s->p.Ehi, &g) != LM_OK ||
      lm_db_create(ctx, (const double *)s->orient.data, s->n, s->hkls.data(), s->nhkls, &g,
                   s->p.max_spots, &s->db) != LM_OK) {
    fprintf(stderr, "FATAL: %s\n", lm_last_error());
    return 1;
  }
  bool fwd = s->p.do_fwd != 0 || access(s->p.forward_file.c_str(), R_OK) != 0;
  if (!fwd) {
    HostBuffer table;
    const uint64_t need = s->n * (uint64_t)(1 + 2 * s->p.max_spots) * 2;
    if (read_file(s->p.forward_file.c_str(), &table) != 0 || table.size < need ||
        lm_db_import_table(s->db, (const uint16_t *)table.data, s->n) != LM_OK) {
      printf("Could not use the forward simulation file %s; simulating.\n",
             s->p.forward_file.c_str());
      fwd = true;
    }
  }
  if (fwd) {
    lm_fwd_stats st{};
    if (lm_db_forward(s->db, LM_FWD_FP64, &st) != LM_OK) {
      fprintf(stderr, "FATAL: forward simulation: %s\n", lm_last_error());
      return 1;
    }
    printf("GPU forward simulation: %llu spots in %.1f ms\n", (unsigned long long)st.n_spots,
           st.ms);
    if (!s->p.forward_file.empty())
      write_forward_file(s->p.forward_file.c_str(), s->db, s->n, s->p.max_spots);
  }
  if (autotune_slot(*s) != 0) return 1;
  if (lm_server_add_db(d.server, s->db, &s->server_id) != LM_OK) {
    fprintf(stderr, "FATAL: %s\n", lm_last_error());
    return 1;
  }
  mkdir(s->p.result_dir.c_str(), 0755);
  const std::string sfx = k == 0 ? "" : "_db" + std::to_string(k);
  const std::string solFN = s->p.result_dir + "/solutions" + sfx + ".txt";
  const std::string spotFN = s->p.result_dir + "/spots" + sfx + ".txt";
  s->sol = fopen(solFN.c_str(), "w");
  s->spots = fopen(spotFN.c_str(), "w");
  if (!s->sol || !s->spots) {
    printf("Could not open output files in %s.\n", s->p.result_dir.c_str());
    return 1;
  }
  write_solution_headers(s->sol, s->spots, true);
  fflush(s->sol);
  fflush(s->spots);
  s->fitp = s->p;
  s->fitp.min_nr_spots = s->p.min_good_spots;  // src/LaueMatchingGPUStream.cu:1173
  d.dbs.push_back(std::move(s));
  return 0;
}

int listen_on(int port) {
  const int lfd = socket(AF_INET, SOCK_STREAM, 0);
  int one = 1;
  setsockopt(lfd, SOL_SOCKET, SO_REUSEADDR, &one, sizeof one);
  sockaddr_in addr{};
  addr.sin_family = AF_INET;
  addr.sin_addr.s_addr = INADDR_ANY;
  addr.sin_port = htons((uint16_t)port);
  if (lfd < 0 || bind(lfd, (sockaddr *)&addr, sizeof addr) < 0 || listen(lfd, 16) < 0) {
    perror("bind/listen");
    if (lfd >= 0) close(lfd);
    return -1;
  }
  return lfd;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc < 5 || (argc - 5) % 3 != 0) {
    puts("LaueMatching streaming daemon on the GPU (MI355X)\n"
         "Arguments: parameterFile orientationFile hklFile nCPUs\n"
         "           [parameterFile2 orientationFile2 hklFile2 ...]\n"
         "Receives frames on TCP port 60517: uint16 image number + float32 pixels (DB 0),\n"
         "and on port 60518 frames with a header naming the DB, size and pixel type.\n");
    return 0;
  }
  const double t_init = omp_get_wtime();
  Daemon d;
  puts("Reading parameter file");
  fflush(stdout);
  const int numProcs = atoi(argv[4]);
  if (numProcs < 1) {
    fprintf(stderr, "FATAL: invalid thread count '%s'.\n", argv[4]);
    return 1;
  }
  lm_ctx *ctx = nullptr;
  if (lm_ctx_create(0, 0, &ctx) != LM_OK) {
    fprintf(stderr, "FATAL: %s\n", lm_last_error());
    return 1;
  }
  // frames admitted beyond the ones on the device wait in pinned slots; a
  // sender blocks when all are taken (the reference bounds its queue at 32
  // frames, src/LaueMatchingGPUStream.cu:45, with one image per launch)
  lm_server_config cfg{};
  cfg.max_batch = kMaxBatch;
  cfg.workers = 2;
  cfg.max_inflight = 2 * kMaxBatch;
  cfg.use_graphs = 1;
  if (lm_server_create(ctx, &cfg, &d.server) != LM_OK) {
    fprintf(stderr, "FATAL: %s\n", lm_last_error());
    return 1;
  }
  if (load_slot(ctx, d, 0, argv[1], argv[2], argv[3])) return 1;
  for (int a = 5, k = 1; a + 2 < argc; a += 3, ++k)
    if (load_slot(ctx, d, k, argv[a], argv[a + 1], argv[a + 2])) return 1;

  struct sigaction sa {};
  sa.sa_handler = on_signal;
  sigemptyset(&sa.sa_mask);
  sa.sa_flags = 0;  // no SA_RESTART: blocking calls return EINTR
  sigaction(SIGINT, &sa, nullptr);
  sigaction(SIGTERM, &sa, nullptr);

  const char *penv = getenv("LM_STREAM_PORT");
  const int port = penv ? atoi(penv) : kPort;
  const int lfd = listen_on(port), lfd2 = listen_on(port + 1);
  if (lfd < 0 || lfd2 < 0) return 1;
  g_listen_fd = lfd;
  g_listen_fd2 = lfd2;
  if (const char *pe = getenv("LM_STREAM_FRAME_POOL_GB"))
    if (atof(pe) > 0) d.pool_bytes = (size_t)(atof(pe) * (double)(1ull << 30));
  const int nQuant = std::max(2, std::min(8, numProcs / 2));
  std::vector<std::thread> quantizers;
  for (int i = 0; i < nQuant; ++i) quantizers.emplace_back(quantize_worker, &d);
  std::thread acceptor(accept_loop, &d, lfd, false);
  std::thread acceptor2(accept_loop, &d, lfd2, true);
  const DbSlot &d0 = *d.dbs[0];
  printf("\n=== LaueMatchingGPUStream ready ===\n"
         "  Port:          %d (wire v1, DB 0), %d (wire v2, %zu DBs)\n"
         "  Orientations:  %llu\n"
         "  Image size:    %d x %d (%.1f MB)\n"
         "  Batching:      up to %d images per launch, %d concurrent batches\n"
         "  Init time:     %.2f s\n"
         "  Waiting for images...\n\n",
         port, port + 1, d.dbs.size(), (unsigned long long)d0.n, d0.p.nrPxX, d0.p.nrPxY,
         d0.npx * 4 / 1e6, cfg.max_batch, cfg.workers, omp_get_wtime() - t_init);
  fflush(stdout);

  // Merge + fit workers: the rows are appended in submission order by
  // whichever worker completes the next one.  A frame that is alone in the
  // post-processing stage is formatted by a team of threads (its latency);
  // under load each worker formats its own frame on one thread, so no
  // OpenMP team is started whose idle threads would spin on the cores the
  // connection handlers need.
  const int nPost = std::min(8, numProcs);
  const int teamThreads = std::min(4, numProcs);
  int postActive = 0;
  auto post_worker = [&] {
    for (;;) {
      PostJob pj;
      int threads = 1;
      {
        std::unique_lock<std::mutex> lk(d.post_mu);
        d.post_cv.wait(lk, [&] { return !d.post.empty() || d.post_closed; });
        if (d.post.empty()) return;
        pj = std::move(d.post.front());
        d.post.pop_front();
        if (postActive == 0 && d.post.empty()) threads = teamThreads;
        ++postActive;
      }
      DbSlot &s = *d.dbs[pj.job.db];
      const size_t npx = s.npx;
      std::vector<float> conv;
      PostInput in;
      if (pj.job.type == 0) {
        in.image_f = (const float *)pj.job.raw.data();
      } else if (pj.job.type == 1) {
        in.image = (const double *)pj.job.raw.data();
      } else {
        conv.resize(npx);
        const uint8_t *r = pj.job.raw.data();
        if (pj.job.type == 2)
          for (size_t i = 0; i < npx; ++i) conv[i] = ((const uint16_t *)r)[i];
        else
          for (size_t i = 0; i < npx; ++i) conv[i] = r[i];
        in.image_f = conv.data();
      }
      in.orients = (const double *)s.orient.data;
      in.n_orients = s.n;
      in.hits = &pj.hits;
      in.nr_results = pj.hits.nr_results;
      in.params = &s.fitp;
      in.hkls = s.hkls.data();
      in.nhkls = s.nhkls;
      in.num_procs = threads;
      in.ctx = ctx;
      in.image_nr = (int)pj.job.image_num;
      PostDone r;
      r.nsol = format_solutions(in, &r.sol, &r.spots);
      lm_hits_free(&pj.hits);
      {
        std::lock_guard<std::mutex> lk(d.post_mu);
        --postActive;
      }
      r.job = std::move(pj.job);
      std::lock_guard<std::mutex> lk(d.out_mu);
      d.done.emplace(pj.seq, std::move(r));
      for (auto it = d.done.find(d.next_out); it != d.done.end();
           it = d.done.find(d.next_out)) {
        PostDone &o = it->second;
        DbSlot &t = *d.dbs[o.job.db];
        fputs(o.spots.c_str(), t.spots);
        fputs(o.sol.c_str(), t.sol);
        fflush(t.sol);
        fflush(t.spots);
        if (o.nsol >= 0)
          printf("[Image %u] %d unique orientations found\n", o.job.image_num, o.nsol);
        printf("[Image %u] Total: %.1f ms after arrival\n", o.job.image_num,
               (omp_get_wtime() - o.job.t_arrival) * 1e3);
        fflush(stdout);
        d.give_buffer(std::move(o.job.raw));
        d.done.erase(it);
        ++d.next_out;
      }
    }
  };
  std::vector<std::thread> posters;
  for (int i = 0; i < nPost; ++i) posters.emplace_back(post_worker);

  // Hit lists in submission order: each is reported when the server completes
  // it; frames with matches go on to the merge and fit.
  uint64_t processed = 0, seq = 0;
  for (;;) {
    Job job;
    {
      std::unique_lock<std::mutex> lk(d.mu);
      d.cv.wait_for(lk, std::chrono::milliseconds(100), [&] { return !d.jobs.empty(); });
      if (d.jobs.empty()) {
        if (!g_running && d.open_clients == 0 && d.unsubmitted == 0) break;
        continue;
      }
      job = std::move(d.jobs.front());
      d.jobs.pop_front();
    }
    lm_hits hits{};
    double lat = 0;
    if (lm_server_wait(d.server, job.ticket, &hits, &lat) != LM_OK) {
      fprintf(stderr, "[Image %u] scoring failed: %s\n", job.image_num, lm_last_error());
      d.give_buffer(std::move(job.raw));
      continue;
    }
    // latencies from the frame's arrival: the hit list is ready when the
    // server completed it (quantization + the server's queue and batch)
    printf("[Image %u] GPU: %.1f ms (queue + batch), %llu matches; hit list %.1f ms after "
           "arrival\n", job.image_num, lat, (unsigned long long)hits.nr_results,
           (job.t_submit - job.t_arrival) * 1e3 + lat);
    ++processed;
    if (hits.nr_results == 0) {
      printf("[Image %u] No matches, skipping fitting.\n", job.image_num);
      printf("[Image %u] Total: %.1f ms after arrival\n", job.image_num,
             (omp_get_wtime() - job.t_arrival) * 1e3);
      fflush(stdout);
      lm_hits_free(&hits);
      d.give_buffer(std::move(job.raw));
      continue;
    }
    fflush(stdout);
    {
      std::lock_guard<std::mutex> lk(d.post_mu);
      d.post.push_back(PostJob{seq++, std::move(job), hits});
    }
    d.post_cv.notify_one();
  }
  {
    std::lock_guard<std::mutex> lk(d.post_mu);
    d.post_closed = true;
  }
  d.post_cv.notify_all();
  for (auto &t : posters) t.join();
  printf("\nShutting down...\n");
  shutdown(lfd, SHUT_RDWR);
  shutdown(lfd2, SHUT_RDWR);
  close(lfd);
  close(lfd2);
  acceptor.join();
  acceptor2.join();
  for (auto &t : d.handlers) t.join();
  {
    std::lock_guard<std::mutex> lk(d.in_mu);
    d.in_closed = true;
  }
  d.in_cv.notify_all();
  for (auto &t : quantizers) t.join();
  lm_server_stats ss{};
  lm_server_get_stats(d.server, &ss);
  printf("Processed %llu images; latency p50 %.1f ms, p99 %.1f ms; %.1f images/s; "
         "%.2f images per launch\n", (unsigned long long)processed, ss.p50_ms, ss.p99_ms,
         ss.images_per_s, ss.mean_batch);
  if (d.frames)
    printf("Host time per frame: receive %.2f ms (from its header), quantize %.2f ms, "
           "hand-over to the server %.2f ms\n", d.ns_recv / 1e6 / d.frames,
           d.ns_quant / 1e6 / d.frames, d.ns_submit / 1e6 / d.frames);
  lm_server_destroy(d.server);
  for (auto &s : d.dbs) {
    fclose(s->sol);
    fclose(s->spots);
    lm_db_destroy(s->db);
  }
  d.dbs.clear();
  lm_ctx_destroy(ctx);
  printf("LaueMatchingGPUStream exited cleanly.\n");
  return 0;
}
