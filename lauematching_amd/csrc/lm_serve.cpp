// lm_serve.cpp - continuous-batching request loop over resident DBs.
//
// Replaces the streaming daemon's GPU side (src/LaueMatchingGPUStream.cu:
// 103-292 queue, 1040-1075 stream contexts, 1167-1257 request loop), where a
// "GPU gate" event serialised one image per kernel launch.  Here:
//   * requests are admitted into pinned host slots (backpressure when all
//     slots are taken, like the reference's bounded queue);
//   * W worker threads each own a HIP stream and scoring buffers; a free
//     worker takes the oldest pending request's DB and up to max_batch of
//     that DB's pending requests in arrival order, so requests that arrive
//     while others run join the next batch, and batches of different workers
//     overlap on the device;
//   * every request gets exactly the result a lone lm_score call gives
//     (each image's result is independent of its batch mates).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "lm_internal.h"

namespace lm {
namespace {

using Clock = std::chrono::steady_clock;

struct Request {
  uint64_t ticket = 0;
  int db = 0;
  int slot = -1;
  double scale = 1.0;
  int min_spots = 0;
  double min_int = 0;
  Clock::time_point t_submit;
};

struct Done {
  lm_hits hits{};
  int status = LM_OK;
  double latency_ms = 0;
};

}  // namespace
}  // namespace lm

struct lm_server {
  lm_ctx *ctx = nullptr;
  lm_server_config cfg{};
  std::vector<lm_db *> dbs;
  uint64_t max_npx = 0;
  // admission slots: pinned host copies of submitted images
  std::vector<uint8_t *> slots;
  std::vector<int> free_slots;
  uint64_t slot_bytes = 0;
  uint64_t reserved = 0;  // device bytes reserved from the context budget
  uint64_t scratch_prepaid = 0;  // per worker: scoring buffers within `reserved`
  std::deque<lm::Request> pending;
  std::unordered_map<uint64_t, lm::Done> done;
  uint64_t next_ticket = 1;
  std::unordered_set<uint64_t> open_tickets;  // submitted, result not yet collected
  bool stop = false;
  int busy = 0;  // workers with a batch on the device
  std::mutex mu;
  std::condition_variable cv_work, cv_done, cv_slot;
  std::vector<std::thread> workers;
  // statistics
  std::vector<double> lat_ms;
  uint64_t batches = 0, batched_images = 0;
  lm::Clock::time_point first_submit, last_done;
  bool have_first = false;
  uint64_t graph_launches = 0;
};

namespace lm {
namespace {

// How many of the c pending requests of one DB to score in one launch.
// A batch of nimg images runs at width batch_width(nimg), so its cost is
// that width's autotuned device time.  While the pending images would finish
// within `after_ms` one at a time, they are taken one at a time (a small
// batch costs several single launches and would only add latency at light
// load); past that, the count with the best images per millisecond is taken
// (fewer than c when padding to the next width costs more than it gains; the
// rest go to the next free worker).  Without timings, all c are taken.
int pick_batch(const lm_db *db, int c, int after_ms) {
  const double one = db->width_ms[0];
  if (c <= 1 || one <= 0) return c;
  const double budget = after_ms > 0 ? after_ms : (after_ms == 0 ? 4.0 * one : 0.0);
  if (c * one <= budget) return 1;
  auto cost = [&](int n) { return db->width_ms[width_log2(batch_width(n))]; };
  int best = 1;
  double best_rate = 1.0 / one;
  std::vector<int> cand;
  for (int w = 4; w <= c; w <<= 1) cand.push_back(w);
  cand.push_back(c);
  for (int n : cand) {
    const double ms = cost(n);
    if (ms <= 0) continue;  // width not timed
    const double rate = n / ms;
    if (rate > best_rate) {
      best_rate = rate;
      best = n;
    }
  }
  return best;
}

void worker_loop(lm_server *s, int wid) {
  (void)hipSetDevice(s->ctx->device);
  hipStream_t st = nullptr;
  if (hipStreamCreate(&st) != hipSuccess) return;
  Scratch sc;
  sc.use_graphs = s->cfg.use_graphs != 0;
  // the worker's share of the reservation made at server start covers its
  // buffers as first allocated; growth beyond it is charged to the budget
  sc.ctx = s->ctx;
  sc.prepaid = s->scratch_prepaid;
  uint8_t *stage = nullptr;
  uint64_t stage_cap = 0;
  uint64_t seen_graphs = 0;
  std::vector<Request> batch;
  for (;;) {
    batch.clear();
    int dbi = -1;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      // while another worker's batch is on the device, let the queue fill
      // (up to a full batch) until that batch is done, instead of taking
      // the few requests that arrived meanwhile one by one
      s->cv_work.wait(lk, [&] {
        if (s->stop) return true;
        if (s->pending.empty()) return false;
        return s->busy == 0 || (int)s->pending.size() >= s->cfg.max_batch;
      });
      if (s->pending.empty()) break;  // stop requested and queue drained
      dbi = s->pending.front().db;
      int same = 0;
      for (const auto &r : s->pending) same += r.db == dbi;
      const int take =
          pick_batch(s->dbs[dbi], std::min(same, s->cfg.max_batch), s->cfg.batch_after_ms);
      for (auto it = s->pending.begin(); it != s->pending.end() && (int)batch.size() < take;) {
        if (it->db == dbi) {
          batch.push_back(*it);
          it = s->pending.erase(it);
        } else {
          ++it;
        }
      }
      ++s->busy;
    }
    lm_db *db = s->dbs[dbi];
    const int nb = (int)batch.size();
    const uint64_t npx = (uint64_t)db->geom.nrPxX * db->geom.nrPxY;
    std::vector<lm_hits> outs(nb);
    int rc = LM_OK;
    if (stage_cap < npx * nb) {
      if (stage) (void)hipFree(stage);
      stage = nullptr;
      stage_cap = 0;
      hipError_t e = hipMalloc(&stage, npx * s->cfg.max_batch);
      if (e != hipSuccess)
        rc = hip_fail(e, "hipMalloc(server staging)");
      else
        stage_cap = npx * s->cfg.max_batch;
    }
    std::vector<double> scales(nb);
    std::vector<int> msp(nb);
    std::vector<double> mint(nb);
    for (int i = 0; i < nb && rc == LM_OK; ++i) {
      hipError_t e = hipMemcpyAsync(stage + (uint64_t)i * npx, s->slots[batch[i].slot], npx,
                                    hipMemcpyHostToDevice, st);
      if (e != hipSuccess) rc = hip_fail(e, "server image upload");
      scales[i] = batch[i].scale;
      msp[i] = batch[i].min_spots;
      mint[i] = batch[i].min_int;
    }
    if (rc == LM_OK) {
      rc = score_images(db, sc, nb, stage, scales.data(), msp.data(), mint.data(), st,
                          outs.data());
    }
    const auto now = Clock::now();
    {
      std::lock_guard<std::mutex> lk(s->mu);
      for (int i = 0; i < nb; ++i) {
        Done d;
        d.status = rc;
        if (rc == LM_OK) d.hits = outs[i];
        d.latency_ms =
            std::chrono::duration<double, std::milli>(now - batch[i].t_submit).count();
        s->done[batch[i].ticket] = d;
        s->free_slots.push_back(batch[i].slot);
        if (rc == LM_OK) s->lat_ms.push_back(d.latency_ms);
      }
      s->batches += 1;
      s->batched_images += nb;
      s->last_done = now;
      s->graph_launches += sc.graph_launches - seen_graphs;
      seen_graphs = sc.graph_launches;
      --s->busy;
    }
    s->cv_done.notify_all();
    s->cv_slot.notify_all();
    s->cv_work.notify_all();  // a waiting worker takes what queued meanwhile
  }
  scratch_release(sc);
  if (stage) (void)hipFree(stage);
  (void)hipStreamDestroy(st);
  (void)wid;
}

}  // namespace
}  // namespace lm

using namespace lm;

extern "C" {

int lm_server_create(lm_ctx *ctx, const lm_server_config *cfg, lm_server **out) {
  if (!ctx || !out) {
    set_error("lm_server_create: bad arguments");
    return LM_ERR_INVALID;
  }
  lm_server *s = new (std::nothrow) lm_server;
  if (!s) {
    set_error("out of host memory");
    return LM_ERR_NOMEM;
  }
  s->ctx = ctx;
  if (cfg) s->cfg = *cfg;
  if (s->cfg.max_batch <= 0) s->cfg.max_batch = 32;
  if (s->cfg.max_batch > kMaxBatch) s->cfg.max_batch = kMaxBatch;
  if (s->cfg.workers <= 0) s->cfg.workers = 2;
  if (s->cfg.max_inflight <= 0) s->cfg.max_inflight = 2 * s->cfg.workers * s->cfg.max_batch;
  ctx_ref(ctx);
  *out = s;
  return LM_OK;
}

int lm_server_add_db(lm_server *s, lm_db *db, int32_t *db_id) {
  if (!s || !db || !db_id) {
    set_error("lm_server_add_db: bad arguments");
    return LM_ERR_INVALID;
  }
  if (!db->table.ready) {
    set_error("lm_server_add_db: the DB has no table");
    return LM_ERR_STATE;
  }
  std::lock_guard<std::mutex> lk(s->mu);
  if (!s->workers.empty()) {
    set_error("lm_server_add_db: add every DB before the first submit");
    return LM_ERR_STATE;
  }
  const uint64_t npx = (uint64_t)db->geom.nrPxX * db->geom.nrPxY;
  s->max_npx = std::max(s->max_npx, npx);
  s->dbs.push_back(db);
  *db_id = (int32_t)s->dbs.size() - 1;
  return LM_OK;
}

static int server_start(lm_server *s) {
  // device reservation per worker: the staging buffer for a full batch of
  // the largest detector plus the worst case, over the DBs, of the scoring
  // buffers (hit records, candidate rows, interleaved batch, chunk flags)
  uint64_t scratch = 0;
  for (const lm_db *db : s->dbs) scratch = std::max(scratch, scratch_bytes_for(db, s->cfg.max_batch));
  s->scratch_prepaid = scratch;
  const uint64_t per_worker = s->max_npx * s->cfg.max_batch + scratch;
  const uint64_t need = per_worker * s->cfg.workers;
  {
    std::lock_guard<std::mutex> lk(s->ctx->mu);
    if (s->ctx->in_use + need > s->ctx->budget) {
      set_error("lm_server: HBM budget exceeded by the serving buffers");
      return LM_ERR_BUDGET;
    }
    s->ctx->in_use += need;
  }
  s->reserved = need;
  s->slot_bytes = s->max_npx;
  for (int i = 0; i < s->cfg.max_inflight; ++i) {
    void *p = nullptr;
    hipError_t e = hipHostMalloc(&p, s->slot_bytes, hipHostMallocDefault);
    if (e != hipSuccess) return hip_fail(e, "hipHostMalloc(admission slot)");
    s->slots.push_back((uint8_t *)p);
    s->free_slots.push_back(i);
  }
  for (int w = 0; w < s->cfg.workers; ++w) s->workers.emplace_back(worker_loop, s, w);
  return LM_OK;
}

int lm_server_submit(lm_server *s, int32_t db_id, const uint8_t *img_u8, double im_scale,
                     int32_t min_nr_spots, double min_intensity, uint64_t *ticket) {
  if (!s || !img_u8 || !ticket || db_id < 0 || db_id >= (int32_t)s->dbs.size()) {
    set_error("lm_server_submit: bad arguments");
    return LM_ERR_INVALID;
  }
  if (!(im_scale > 0)) {
    set_error("lm_server_submit: imScale must be positive");
    return LM_ERR_INVALID;
  }
  const auto t0 = Clock::now();
  std::unique_lock<std::mutex> lk(s->mu);
  if (s->workers.empty()) {
    const int rc = server_start(s);
    if (rc) return rc;
  }
  s->cv_slot.wait(lk, [&] { return !s->free_slots.empty(); });
  const int slot = s->free_slots.back();
  s->free_slots.pop_back();
  lm_db *db = s->dbs[db_id];
  const uint64_t npx = (uint64_t)db->geom.nrPxX * db->geom.nrPxY;
  lk.unlock();
  std::memcpy(s->slots[slot], img_u8, npx);
  lk.lock();
  Request r;
  r.ticket = s->next_ticket++;
  s->open_tickets.insert(r.ticket);
  r.db = db_id;
  r.slot = slot;
  r.scale = im_scale;
  r.min_spots = min_nr_spots;
  r.min_int = min_intensity;
  r.t_submit = t0;
  if (!s->have_first) {
    s->first_submit = t0;
    s->have_first = true;
  }
  s->pending.push_back(r);
  *ticket = r.ticket;
  lk.unlock();
  s->cv_work.notify_one();
  return LM_OK;
}

int lm_server_wait(lm_server *s, uint64_t ticket, lm_hits *out, double *latency_ms) {
  if (!s || !out) {
    set_error("lm_server_wait: bad arguments");
    return LM_ERR_INVALID;
  }
  std::unique_lock<std::mutex> lk(s->mu);
  if (!s->open_tickets.count(ticket)) {
    set_error("lm_server_wait: unknown or already collected ticket");
    return LM_ERR_INVALID;
  }
  s->cv_done.wait(lk, [&] { return s->done.count(ticket) > 0; });
  Done d = s->done[ticket];
  s->done.erase(ticket);
  s->open_tickets.erase(ticket);
  *out = d.hits;
  if (latency_ms) *latency_ms = d.latency_ms;
  if (d.status != LM_OK) set_error("lm_server: batch failed");
  return d.status;
}

int lm_server_poll(lm_server *s, uint64_t ticket, lm_hits *out, double *latency_ms,
                   int32_t *ready) {
  if (!s || !out || !ready) {
    set_error("lm_server_poll: bad arguments");
    return LM_ERR_INVALID;
  }
  std::lock_guard<std::mutex> lk(s->mu);
  if (!s->open_tickets.count(ticket)) {
    set_error("lm_server_poll: unknown or already collected ticket");
    return LM_ERR_INVALID;
  }
  auto it = s->done.find(ticket);
  if (it == s->done.end()) {
    *ready = 0;
    return LM_OK;
  }
  Done d = it->second;
  s->done.erase(it);
  s->open_tickets.erase(ticket);
  *ready = 1;
  *out = d.hits;
  if (latency_ms) *latency_ms = d.latency_ms;
  if (d.status != LM_OK) set_error("lm_server: batch failed");
  return d.status;
}

int lm_server_get_stats(lm_server *s, lm_server_stats *st) {
  if (!s || !st) {
    set_error("lm_server_get_stats: bad arguments");
    return LM_ERR_INVALID;
  }
  std::lock_guard<std::mutex> lk(s->mu);
  memset(st, 0, sizeof *st);
  std::vector<double> v = s->lat_ms;
  std::sort(v.begin(), v.end());
  st->completed = v.size();
  st->batches = s->batches;
  st->mean_batch = s->batches ? (double)s->batched_images / (double)s->batches : 0;
  if (!v.empty()) {
    auto pct = [&](double q) {
      const size_t i = std::min(v.size() - 1, (size_t)std::ceil(q * v.size()) - (q > 0 ? 1 : 0));
      return v[i];
    };
    st->p50_ms = pct(0.50);
    st->p99_ms = pct(0.99);
    st->max_ms = v.back();
    const double span =
        std::chrono::duration<double>(s->last_done - s->first_submit).count();
    st->images_per_s = span > 0 ? (double)v.size() / span : 0;
  }
  st->graph_launches = s->graph_launches;
  return LM_OK;
}

void lm_server_reset_stats(lm_server *s) {
  if (!s) return;
  std::lock_guard<std::mutex> lk(s->mu);
  s->lat_ms.clear();
  s->batches = 0;
  s->batched_images = 0;
  s->have_first = false;
}

void lm_server_destroy(lm_server *s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  for (auto &t : s->workers) t.join();
  for (auto &kv : s->done) lm_hits_free(&kv.second.hits);
  for (uint8_t *p : s->slots) (void)hipHostFree(p);
  if (s->reserved) {
    std::lock_guard<std::mutex> lk(s->ctx->mu);
    s->ctx->in_use -= s->reserved;
  }
  lm_ctx *ctx = s->ctx;
  delete s;
  ctx_unref(ctx);
}

}  // extern "C"
