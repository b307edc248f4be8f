// lm_host_stream.cpp - bin/LaueMatchingGPUStream: the streaming daemon.
//
//   LaueMatchingGPUStream params.txt orient.bin hkls.csv nCPUs
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

std::atomic<bool> g_running{true};
std::atomic<int> g_listen_fd{-1};

void on_signal(int sig) {
  if (g_running.exchange(false)) {
    const char msg[] = "\nCaught signal, requesting shutdown...\n";
    (void)!write(STDOUT_FILENO, msg, sizeof msg - 1);
    const int fd = g_listen_fd.load();
    if (fd >= 0) shutdown(fd, SHUT_RDWR);
  } else {
    const char msg[] = "\nForced exit.\n";
    (void)!write(STDOUT_FILENO, msg, sizeof msg - 1);
    _exit(1);
  }
  (void)sig;
}

struct Job {
  uint16_t image_num;
  uint64_t ticket;
  std::vector<float> image;
  double t_recv;
};

struct Daemon {
  LmParams p;
  lm_server *server = nullptr;
  int32_t db_id = 0;
  size_t npx = 0;
  std::mutex mu;
  std::condition_variable cv;
  std::deque<Job> jobs;
  int open_clients = 0;
  std::vector<std::thread> handlers;
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

void handle_client(Daemon *d, int fd) {
  timeval to{5, 0};
  setsockopt(fd, SOL_SOCKET, SO_RCVTIMEO, &to, sizeof to);
  std::vector<float> buf(d->npx);
  std::vector<double> f64(d->npx);
  std::vector<uint8_t> u8(d->npx);
  printf("Client handler started (socket %d), wire_payload=%zu bytes\n", fd, d->npx * 4);
  fflush(stdout);
  while (g_running) {
    uint8_t hdr[2];
    if (!recv_all(fd, hdr, 2)) break;
    if (!recv_all(fd, buf.data(), d->npx * 4)) break;
    const uint16_t num = (uint16_t)(hdr[0] | (hdr[1] << 8));
    for (size_t i = 0; i < d->npx; ++i) f64[i] = buf[i];
    double scale = 1.0;
    lm_quantize_u8(f64.data(), d->npx, u8.data(), &scale);
    uint64_t ticket = 0;
    if (lm_server_submit(d->server, d->db_id, u8.data(), scale, d->p.min_nr_spots,
                         d->p.min_intensity, &ticket) != LM_OK) {
      fprintf(stderr, "submit failed for image %u: %s\n", num, lm_last_error());
      break;
    }
    printf("Received image %u (%zu bytes wire, %zu pixels)\n", num, d->npx * 4, d->npx);
    fflush(stdout);
    {
      std::lock_guard<std::mutex> lk(d->mu);
      d->jobs.push_back(Job{num, ticket, buf, omp_get_wtime()});
    }
    d->cv.notify_one();
  }
  close(fd);
  {
    std::lock_guard<std::mutex> lk(d->mu);
    --d->open_clients;
  }
  d->cv.notify_one();
  printf("Client handler finished (socket %d)\n", fd);
  fflush(stdout);
}

void accept_loop(Daemon *d, int lfd) {
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
    d->handlers.emplace_back(handle_client, d, fd);
  }
}

}  // namespace

int main(int argc, char **argv) {
  if (argc != 5) {
    puts("LaueMatching streaming daemon on the GPU (MI355X)\n"
         "Arguments: parameterFile orientationFile hklFile nCPUs\n"
         "Receives frames on TCP port 60517: uint16 image number + float32 pixels.\n");
    return 0;
  }
  const double t_init = omp_get_wtime();
  Daemon d;
  puts("Reading parameter file");
  fflush(stdout);
  const int prc = lm_read_params(argv[1], &d.p);
  if (prc == -1) {
    printf("Could not open parameter file %s.\n", argv[1]);
    return 1;
  }
  if (prc == -2) {
    fprintf(stderr, "FATAL: NrPxX/NrPxY or LatticeParameter missing.\n");
    return 1;
  }
  const int numProcs = atoi(argv[4]);
  if (numProcs < 1) {
    fprintf(stderr, "FATAL: invalid thread count '%s'.\n", argv[4]);
    return 1;
  }
  HostBuffer orient;
  if (read_file(argv[2], &orient) != 0) {
    puts("Could not read orientation file, exiting.");
    return 1;
  }
  const uint64_t n = (uint64_t)((double)orient.size / (double)(9 * sizeof(double)));
  std::vector<int32_t> hkls;
  if (read_hkls(argv[3], &hkls) != 0) {
    printf("Could not read hkl file %s.\n", argv[3]);
    return 1;
  }
  const int nhkls = (int)(hkls.size() / 3);
  d.npx = (size_t)d.p.nrPxX * d.p.nrPxY;
  printf("%llu orientations, %d hkls, image %d x %d\n", (unsigned long long)n, nhkls, d.p.nrPxX,
         d.p.nrPxY);
  lm_ctx *ctx = nullptr;
  lm_geom g;
  lm_db *db = nullptr;
  if (lm_ctx_create(0, 0, &ctx) != LM_OK ||
      lm_geom_init(d.p.lattice, d.p.space_group, d.p.P, d.p.R, d.p.pxX, d.p.pxY, d.p.nrPxX,
                   d.p.nrPxY, d.p.Elo, d.p.Ehi, &g) != LM_OK ||
      lm_db_create(ctx, (const double *)orient.data, n, hkls.data(), nhkls, &g, d.p.max_spots,
                   &db) != LM_OK) {
    fprintf(stderr, "FATAL: %s\n", lm_last_error());
    return 1;
  }
  bool fwd = d.p.do_fwd != 0 || access(d.p.forward_file.c_str(), R_OK) != 0;
  if (!fwd) {
    HostBuffer table;
    const uint64_t need = n * (uint64_t)(1 + 2 * d.p.max_spots) * 2;
    if (read_file(d.p.forward_file.c_str(), &table) != 0 || table.size < need ||
        lm_db_import_table(db, (const uint16_t *)table.data, n) != LM_OK) {
      printf("Could not use the forward simulation file %s; simulating.\n",
             d.p.forward_file.c_str());
      fwd = true;
    }
  }
  if (fwd) {
    lm_fwd_stats st{};
    if (lm_db_forward(db, LM_FWD_FP64, &st) != LM_OK) {
      fprintf(stderr, "FATAL: forward simulation: %s\n", lm_last_error());
      return 1;
    }
    printf("GPU forward simulation: %llu spots in %.1f ms\n", (unsigned long long)st.n_spots,
           st.ms);
    if (!d.p.forward_file.empty())
      write_forward_file(d.p.forward_file.c_str(), db, n, d.p.max_spots);
  }
  lm_server_config cfg{};
  cfg.max_batch = 16;
  cfg.workers = 2;
  cfg.max_inflight = 32;  // MAX_QUEUE_SIZE, src/LaueMatchingGPUStream.cu:45
  cfg.use_graphs = 1;
  if (lm_server_create(ctx, &cfg, &d.server) != LM_OK ||
      lm_server_add_db(d.server, db, &d.db_id) != LM_OK) {
    fprintf(stderr, "FATAL: %s\n", lm_last_error());
    return 1;
  }
  mkdir(d.p.result_dir.c_str(), 0755);
  const std::string solFN = d.p.result_dir + "/solutions.txt";
  const std::string spotFN = d.p.result_dir + "/spots.txt";
  FILE *sol = fopen(solFN.c_str(), "w");
  FILE *spots = fopen(spotFN.c_str(), "w");
  if (!sol || !spots) {
    printf("Could not open output files in %s.\n", d.p.result_dir.c_str());
    return 1;
  }
  write_solution_headers(sol, spots, true);
  fflush(sol);
  fflush(spots);

  struct sigaction sa {};
  sa.sa_handler = on_signal;
  sigemptyset(&sa.sa_mask);
  sa.sa_flags = 0;  // no SA_RESTART: blocking calls return EINTR
  sigaction(SIGINT, &sa, nullptr);
  sigaction(SIGTERM, &sa, nullptr);

  const char *penv = getenv("LM_STREAM_PORT");
  const int port = penv ? atoi(penv) : kPort;
  const int lfd = socket(AF_INET, SOCK_STREAM, 0);
  int one = 1;
  setsockopt(lfd, SOL_SOCKET, SO_REUSEADDR, &one, sizeof one);
  sockaddr_in addr{};
  addr.sin_family = AF_INET;
  addr.sin_addr.s_addr = INADDR_ANY;
  addr.sin_port = htons((uint16_t)port);
  if (lfd < 0 || bind(lfd, (sockaddr *)&addr, sizeof addr) < 0 || listen(lfd, 16) < 0) {
    perror("bind/listen");
    return 1;
  }
  g_listen_fd = lfd;
  std::thread acceptor(accept_loop, &d, lfd);
  printf("\n=== LaueMatchingGPUStream ready ===\n"
         "  Port:          %d\n"
         "  Orientations:  %llu\n"
         "  Image size:    %d x %d (%.1f MB)\n"
         "  Batching:      up to %d images per launch, %d concurrent batches\n"
         "  Init time:     %.2f s\n"
         "  Waiting for images...\n\n",
         port, (unsigned long long)n, d.p.nrPxX, d.p.nrPxY, d.npx * 4 / 1e6, cfg.max_batch,
         cfg.workers, omp_get_wtime() - t_init);
  fflush(stdout);

  LmParams fitp = d.p;
  fitp.min_nr_spots = d.p.min_good_spots;  // src/LaueMatchingGPUStream.cu:1173
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
    lm_hits hits{};
    double lat = 0;
    if (lm_server_wait(d.server, job.ticket, &hits, &lat) != LM_OK) {
      fprintf(stderr, "[Image %u] scoring failed: %s\n", job.image_num, lm_last_error());
      continue;
    }
    printf("[Image %u] GPU: %.1f ms (queue + batch), %llu matches\n", job.image_num, lat,
           (unsigned long long)hits.nr_results);
    if (hits.nr_results == 0) {
      printf("[Image %u] No matches, skipping fitting.\n", job.image_num);
    } else {
      PostInput in;
      in.orients = (const double *)orient.data;
      in.n_orients = n;
      in.hits = &hits;
      in.nr_results = hits.nr_results;
      in.params = &fitp;
      in.hkls = hkls.data();
      in.nhkls = nhkls;
      in.image_f = job.image.data();
      in.num_procs = numProcs;
      in.ctx = ctx;
      in.image_nr = job.image_num;
      const int ns = append_solutions(in, sol, spots);
      printf("[Image %u] %d unique orientations found\n", job.image_num, ns);
    }
    printf("[Image %u] Total: %.1f ms\n", job.image_num, (omp_get_wtime() - job.t_recv) * 1e3);
    fflush(stdout);
    lm_hits_free(&hits);
    ++processed;
  }
  printf("\nShutting down...\n");
  shutdown(lfd, SHUT_RDWR);
  close(lfd);
  acceptor.join();
  for (auto &t : d.handlers) t.join();
  lm_server_stats ss{};
  lm_server_get_stats(d.server, &ss);
  printf("Processed %llu images; latency p50 %.1f ms, p99 %.1f ms; %.1f images/s; "
         "%.2f images per launch\n", (unsigned long long)processed, ss.p50_ms, ss.p99_ms,
         ss.images_per_s, ss.mean_batch);
  lm_server_destroy(d.server);
  fclose(sol);
  fclose(spots);
  lm_db_destroy(db);
  lm_ctx_destroy(ctx);
  printf("LaueMatchingGPUStream exited cleanly.\n");
  return 0;
}
