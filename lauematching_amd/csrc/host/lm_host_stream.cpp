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
#include <deque>
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
  std::vector<float> image;  // the frame's pixels (for the fit), from the pool
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

struct Daemon {
  std::vector<std::unique_ptr<DbSlot>> dbs;
  lm_server *server = nullptr;
  std::mutex mu;
  std::condition_variable cv;
  std::deque<Job> jobs;
  int open_clients = 0;
  std::vector<std::thread> handlers;
  // frame buffers handed from the connection handlers to the result loop and
  // back, so a frame is received into a buffer that the fit then reads
  std::mutex pool_mu;
  std::vector<std::vector<float>> pool;
  std::vector<float> take_buffer(size_t npx) {
    std::lock_guard<std::mutex> lk(pool_mu);
    std::vector<float> b;
    if (!pool.empty()) {
      b = std::move(pool.back());
      pool.pop_back();
    }
    b.resize(npx);
    return b;
  }
  void give_buffer(std::vector<float> &&b) {
    std::lock_guard<std::mutex> lk(pool_mu);
    if (pool.size() < 64) pool.push_back(std::move(b));
  }
};

bool recv_all(int fd, void *buf, size_t len) {
  size_t got = 0;
  while (got < len) {
    const ssize_t n = recv(fd, (char *)buf + got, len - got, 0);
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

// Quantize one frame (src/LaueMatchingCPU.c:355-377) -- from its float64
// pixels when the wire carried them (f64 non-null), else from the float32
// pixels, which is the same arithmetic -- submit it to the server for DB k and
// queue it for post-processing (which reads the float32 pixels, as the
// reference daemon's fit does).  img moves into the job.
bool submit_frame(Daemon *d, int k, uint32_t num, std::vector<float> &img, const double *f64,
                  std::vector<uint8_t> &u8, double t_arrival) {
  DbSlot &s = *d->dbs[k];
  u8.resize(s.npx);
  double scale = 1.0;
  if (f64)
    lm_quantize_u8(f64, s.npx, u8.data(), &scale);
  else
    lm_quantize_u8_f32(img.data(), s.npx, u8.data(), &scale);
  const double t_submit = omp_get_wtime();
  uint64_t ticket = 0;
  if (lm_server_submit(d->server, s.server_id, u8.data(), scale, s.p.min_nr_spots,
                       s.p.min_intensity, &ticket) != LM_OK) {
    fprintf(stderr, "submit failed for image %u: %s\n", num, lm_last_error());
    return false;
  }
  printf("Received image %u for DB %d (%zu pixels)\n", num, k, s.npx);
  fflush(stdout);
  {
    std::lock_guard<std::mutex> lk(d->mu);
    d->jobs.push_back(Job{num, k, ticket, std::move(img), t_arrival, t_submit});
  }
  d->cv.notify_one();
  return true;
}

// v1: uint16 image number + float32 pixels of DB 0's detector
// (scripts/laue_image_server.py:217-219, src/LaueMatchingGPUStream.cu:295-383)
void serve_v1(Daemon *d, int fd) {
  const size_t npx = d->dbs[0]->npx;
  std::vector<uint8_t> u8;
  printf("Client handler started (socket %d), wire_payload=%zu bytes\n", fd, npx * 4);
  fflush(stdout);
  while (g_running) {
    uint8_t hdr[2];
    if (!recv_all(fd, hdr, 2)) break;
    std::vector<float> buf = d->take_buffer(npx);
    if (!recv_all(fd, buf.data(), npx * 4)) {
      d->give_buffer(std::move(buf));
      break;
    }
    const uint16_t num = (uint16_t)(hdr[0] | (hdr[1] << 8));
    if (!submit_frame(d, 0, num, buf, nullptr, u8, omp_get_wtime())) break;
  }
}

// v2: per-frame header (see the top of this file)
void serve_v2(Daemon *d, int fd) {
  std::vector<uint8_t> raw;
  std::vector<double> f64;
  std::vector<uint8_t> u8;
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
    static const size_t kSize[4] = {4, 8, 2, 1};
    if (magic != kV2Magic || resv != 0 || type > 3 || k >= (int)d->dbs.size() ||
        (int64_t)W != d->dbs[k]->p.nrPxX || (int64_t)H != d->dbs[k]->p.nrPxY) {
      fprintf(stderr, "wire v2: bad frame header (magic %08x, DB %d, type %d, %u x %u); "
              "closing the connection\n", magic, k, type, W, H);
      break;
    }
    const size_t npx = (size_t)W * H;
    raw.resize(npx * kSize[type]);
    if (!recv_all(fd, raw.data(), raw.size())) break;
    const double t_arrival = omp_get_wtime();
    std::vector<float> img = d->take_buffer(npx);
    if (type == 1) f64.resize(npx);
    for (size_t i = 0; i < npx; ++i) {
      const uint8_t *q = raw.data() + i * kSize[type];
      switch (type) {
        case 0: { float v; memcpy(&v, q, 4); img[i] = v; break; }
        case 1: { double v; memcpy(&v, q, 8); f64[i] = v; img[i] = (float)v; break; }
        case 2: img[i] = (float)(q[0] | (q[1] << 8)); break;
        default: img[i] = (float)q[0]; break;
      }
    }
    if (!submit_frame(d, k, num, img, type == 1 ? f64.data() : nullptr, u8, t_arrival)) break;
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

  uint64_t processed = 0;
  for (;;) {
    Job job;
    {
      std::unique_lock<std::mutex> lk(d.mu);
      d.cv.wait_for(lk, std::chrono::milliseconds(100), [&] { return !d.jobs.empty(); });
      if (d.jobs.empty()) {
        if (!g_running && d.open_clients == 0) break;
        continue;
      }
      job = std::move(d.jobs.front());
      d.jobs.pop_front();
    }
    DbSlot &s = *d.dbs[job.db];
    lm_hits hits{};
    double lat = 0;
    if (lm_server_wait(d.server, job.ticket, &hits, &lat) != LM_OK) {
      fprintf(stderr, "[Image %u] scoring failed: %s\n", job.image_num, lm_last_error());
      continue;
    }
    // latencies from the frame's arrival: the hit list is ready when the
    // server completed it (quantization + the server's queue and batch)
    printf("[Image %u] GPU: %.1f ms (queue + batch), %llu matches; hit list %.1f ms after "
           "arrival\n", job.image_num, lat, (unsigned long long)hits.nr_results,
           (job.t_submit - job.t_arrival) * 1e3 + lat);
    if (hits.nr_results == 0) {
      printf("[Image %u] No matches, skipping fitting.\n", job.image_num);
    } else {
      PostInput in;
      in.orients = (const double *)s.orient.data;
      in.n_orients = s.n;
      in.hits = &hits;
      in.nr_results = hits.nr_results;
      in.params = &s.fitp;
      in.hkls = s.hkls.data();
      in.nhkls = s.nhkls;
      in.image_f = job.image.data();
      in.num_procs = numProcs;
      in.ctx = ctx;
      in.image_nr = (int)job.image_num;
      const int ns = append_solutions(in, s.sol, s.spots);
      printf("[Image %u] %d unique orientations found\n", job.image_num, ns);
    }
    printf("[Image %u] Total: %.1f ms after arrival\n", job.image_num,
           (omp_get_wtime() - job.t_arrival) * 1e3);
    fflush(stdout);
    d.give_buffer(std::move(job.image));
    lm_hits_free(&hits);
    ++processed;
  }
  printf("\nShutting down...\n");
  shutdown(lfd, SHUT_RDWR);
  shutdown(lfd2, SHUT_RDWR);
  close(lfd);
  close(lfd2);
  acceptor.join();
  acceptor2.join();
  for (auto &t : d.handlers) t.join();
  lm_server_stats ss{};
  lm_server_get_stats(d.server, &ss);
  printf("Processed %llu images; latency p50 %.1f ms, p99 %.1f ms; %.1f images/s; "
         "%.2f images per launch\n", (unsigned long long)processed, ss.p50_ms, ss.p99_ms,
         ss.images_per_s, ss.mean_batch);
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
