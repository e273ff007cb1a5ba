// cs_serve.cpp -- a centrality request server with continuous batching.
//
// Requests (shortest- or simplest-path centrality over any registered graph,
// each with its own thresholds and source subset) are admitted first-come
// first-served while their device memory fits the HBM budget and fewer than
// `max_active` are on the device. Admitted shortest-path requests are served
// by one worker in back-to-back batched launches ("slices"): each slice takes a
// share of the remaining sources of every active request, sized by a running
// estimate of the per-source cost so that one slice lasts about `slice_ms`.
// A request that arrives while a slice runs joins the next one; a request
// completes when its last sources finish, and leaves. Simplest-path requests
// run on a second worker and stream, so the two kernel families overlap on the
// device; since an angular launch lasts about as long as its slowest source,
// each simplest slice packs every active simplest request into one launch.
// Results are copied out and unpacked by the worker that finishes a request.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <thread>
#include <unordered_map>

#include "cs_api.hpp"

using Clock = std::chrono::steady_clock;

namespace cs {
namespace {

// Sources of one simplest-path slice at most (about the angular kernel's lane budget).
constexpr uint64_t kSimplestSlice = 262144;

double ms_between(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

struct Req {
  int64_t id = 0;
  cs_graph *h = nullptr;
  int kind = CS_REQ_SHORTEST;
  CallSpec spec;                 // the first threshold chunk
  std::vector<CallSpec> chunks;  // one per kMaxT thresholds, each searching to the largest cutoff
  SourceSet ss;
  std::vector<uint8_t> elig;
  double *out = nullptr;
  uint32_t *reach = nullptr;
  bool betw = false;
  double betw_scale = 1.0;
  uint32_t N = 0, T = 0, M = 0;
  size_t dev_bytes = 0;
  uint32_t reach_words() const { return std::max<uint32_t>(T, kMaxT); }
  // device state (valid once `ready`)
  bool ready = false;
  uint32_t *d_src = nullptr;
  float *d_wt = nullptr, *d_cyc = nullptr;
  uint8_t *d_el = nullptr;
  double *d_acc = nullptr;
  uint32_t *d_reach = nullptr;
  DevGraph *dg = nullptr;
  DevEdges *de = nullptr;
  // progress
  uint32_t handed = 0;    // sources given to launches
  uint32_t finished = 0;  // sources whose launch completed
  uint32_t launches = 0;
  int status = 0;  // 0 queued, 1 running, 2 done, < 0 error
  bool cancelled = false;  // no more sources are handed out; completes once none is in flight
  int builder = -1;        // the worker that admitted it and builds its device state
  std::string msg;
  Clock::time_point t_submit, t_admit, t_done;

  uint32_t n() const { return (uint32_t)ss.src.size(); }
};

void free_req_device(Req &r) {
  for (void *q : {(void *)r.d_src, (void *)r.d_wt, (void *)r.d_cyc, (void *)r.d_el, (void *)r.d_acc,
                  (void *)r.d_reach})
    if (q) (void)hipFree(q);
  r.d_src = nullptr;
  r.d_wt = r.d_cyc = nullptr;
  r.d_el = nullptr;
  r.d_acc = nullptr;
  r.d_reach = nullptr;
}

template <typename T>
T *dev_copy(const T *src, size_t n) {
  void *q = nullptr;
  hip_check(hipMalloc(&q, sizeof(T) * std::max<size_t>(n, 1)), "hipMalloc request");
  if (n) hip_check(hipMemcpy(q, src, sizeof(T) * n, hipMemcpyHostToDevice), "upload request");
  return static_cast<T *>(q);
}

}  // namespace
}  // namespace cs

using namespace cs;

struct cs_server {
  cs_server_config cfg{};
  int device = 0;
  std::mutex mu;
  std::condition_variable cv_work;  // new requests or shutdown
  std::condition_variable cv_done;  // a request finished
  std::deque<Req *> queue;          // submitted, not yet admitted
  std::vector<Req *> active;        // admitted, device state being built or ready
  std::unordered_map<int64_t, std::unique_ptr<Req>> reqs;
  int64_t next_id = 0;
  size_t used = 0;
  bool stop = false;
  double ms_per_src[2] = {0.0, 0.0};  // running per-source cost, shortest / simplest
  std::thread workers[2];
  bool dead[2] = {false, false};      // a worker that ended without a shutdown
  std::string dead_msg[2];
  std::atomic<int> fault[2] = {0, 0};  // failure injected into a worker's next slice (tests)

  // Moves queued requests to `active` while the budget allows (FIFO). Called
  // with `mu` held; returns the newly admitted requests for the caller to
  // build device state for outside the lock.
  std::vector<Req *> admit_locked(int who) {
    std::vector<Req *> fresh;
    while (!queue.empty()) {
      Req *r = queue.front();
      if (dead[r->kind]) {
        queue.pop_front();
        fail_locked(r, CS_EINTERNAL, dead_msg[r->kind]);
        continue;
      }
      if (active.size() >= cfg.max_active) break;
      if (used + r->dev_bytes > cfg.hbm_budget_bytes) break;
      queue.pop_front();
      used += r->dev_bytes;
      r->status = 1;
      r->t_admit = Clock::now();
      r->builder = who;
      active.push_back(r);
      fresh.push_back(r);
    }
    return fresh;
  }

  void build_device(Req &r) {
    Graph &g = r.h->g;
    {
      std::lock_guard<std::mutex> lk(g.mu);
      r.dg = &device_graph(g, device);
      r.de = &device_edges(g, *r.dg, r.spec.speed);
    }
    const uint32_t n = r.n();
    r.d_src = dev_copy(r.ss.src.data(), n);
    r.d_wt = dev_copy(r.ss.wt.data(), n);
    if (r.kind == CS_REQ_SHORTEST) r.d_cyc = dev_copy(r.ss.cyc.data(), n);
    r.d_el = dev_copy(r.elig.data(), r.elig.size());
    hip_check(hipMalloc(&r.d_acc, sizeof(double) * std::max<size_t>((size_t)r.N * r.M * r.T, 1)), "hipMalloc acc");
    hip_check(hipMemset(r.d_acc, 0, sizeof(double) * (size_t)r.N * r.M * r.T), "memset acc");
    hip_check(hipMalloc(&r.d_reach, 4 * r.reach_words()), "hipMalloc reach");
    hip_check(hipMemset(r.d_reach, 0, 4 * r.reach_words()), "memset reach");
  }

  void finish(Req &r) {
    std::vector<double> hacc((size_t)r.N * r.M * r.T);
    std::vector<uint32_t> hreach(r.reach_words());
    if (!hacc.empty())
      hip_check(hipMemcpy(hacc.data(), r.d_acc, sizeof(double) * hacc.size(), hipMemcpyDeviceToHost), "copy out");
    hip_check(hipMemcpy(hreach.data(), r.d_reach, 4 * hreach.size(), hipMemcpyDeviceToHost), "copy out");
    unpack_acc_strided(hacc.data(), r.N, r.M, r.T, r.T, r.out,
                       r.kind == CS_REQ_SHORTEST ? result_order(r.h->g) : nullptr);
    if (r.betw) {
      if (r.kind == CS_REQ_SHORTEST)
        scale_rows(r.out, r.N, r.T, 5, 6, r.betw_scale);
      else
        scale_rows(r.out, r.N, r.T, 3, 3, r.betw_scale);
    }
    if (r.reach)
      for (uint32_t t = 0; t < r.T; ++t) r.reach[t] = hreach[t];
  }

  // A request that never reached the device fails (mu held).
  void fail_locked(Req *r, int err, const std::string &msg) {
    r->status = -err;
    r->msg = msg;
    r->t_done = Clock::now();
    cv_done.notify_all();
  }

  // Marks requests done (or failed), releasing their budget.
  void complete(const std::vector<Req *> &done, int err, const std::string &msg) {
    std::lock_guard<std::mutex> lk(mu);
    for (Req *r : done) {
      r->status = err ? -err : 2;
      r->msg = msg;
      r->t_done = Clock::now();
      used -= r->dev_bytes;
      active.erase(std::remove(active.begin(), active.end(), r), active.end());
    }
    cv_done.notify_all();
    cv_work.notify_all();
  }

  bool admissible_locked() const {
    return !queue.empty() && active.size() < cfg.max_active && used + queue.front()->dev_bytes <= cfg.hbm_budget_bytes;
  }

  // A cancelled request of this kind whose launched sources have all finished.
  Req *cancelled_idle_locked(int kind) const {
    for (Req *r : active)
      if (r->kind == kind && r->cancelled && r->ready && r->finished == r->handed) return r;
    return nullptr;
  }

  bool has_work_locked(int kind) const {
    if (admissible_locked() || cancelled_idle_locked(kind)) return true;
    for (const Req *r : active)
      if (r->kind == kind && r->ready && !r->cancelled && r->handed < r->n()) return true;
    return false;
  }

  // Runs one worker. Whatever ends it other than a shutdown (a failure
  // outside the per-request handling, down to a HIP error before its loop)
  // fails every queued and active request of its kind, and later submissions
  // of that kind fail at once: no request waits forever.
  void worker(int kind) {
    std::string why = "server worker stopped";
    try {
      worker_loop(kind);
      return;
    } catch (const Error &e) {
      why = e.msg;
    } catch (const std::exception &e) {
      why = e.what();
    } catch (...) {
      why = "unknown failure";
    }
    std::vector<Req *> orphans;
    {
      std::lock_guard<std::mutex> lk(mu);
      if (stop) return;
      dead[kind] = true;
      dead_msg[kind] = "server worker failed: " + why;
      for (auto it = queue.begin(); it != queue.end();) {
        if ((*it)->kind == kind) {
          fail_locked(*it, CS_EINTERNAL, dead_msg[kind]);
          it = queue.erase(it);
        } else {
          ++it;
        }
      }
      // its kind's requests on the device, and any it admitted but had not built
      // (those the other worker is building, it fails itself)
      for (Req *r : active)
        if ((r->kind == kind && r->ready) || (!r->ready && r->builder == kind)) orphans.push_back(r);
    }
    for (Req *r : orphans) free_req_device(*r);
    complete(orphans, CS_EINTERNAL, "server worker failed: " + why);
  }

  // The failure a caught exception becomes for the requests it affects.
  static void describe(std::exception_ptr ep, int &err, std::string &msg) {
    try {
      std::rethrow_exception(ep);
    } catch (const Error &e) {
      err = e.code;
      msg = e.msg;
    } catch (const std::bad_alloc &) {
      err = CS_ENOMEM;
      msg = "host allocation failed";
    } catch (const std::exception &e) {
      err = CS_EINTERNAL;
      msg = e.what();
    } catch (...) {
      err = CS_EINTERNAL;
      msg = "unknown failure";
    }
  }

  void worker_loop(int kind) {
    hip_check(hipSetDevice(device), "hipSetDevice");
    hipStream_t stream;
    hip_check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "stream");
    struct StreamGuard {
      hipStream_t s;
      ~StreamGuard() { (void)hipStreamDestroy(s); }
    } stream_guard{stream};
    for (;;) {
      std::vector<Req *> fresh;
      std::vector<Req *> slice_reqs;
      std::vector<uint32_t> slice_take;
      Req *to_cancel = nullptr;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv_work.wait(lk, [&] { return stop || has_work_locked(kind); });
        if (stop) break;
        fresh = admit_locked(kind);
        to_cancel = cancelled_idle_locked(kind);
      }
      if (fault[kind].load() == 3) {
        fault[kind] = 0;
        throw std::runtime_error("injected worker failure");
      }
      if (to_cancel) {
        free_req_device(*to_cancel);
        complete({to_cancel}, CS_ECANCELED, "request cancelled");
      }
      // device state of newly admitted requests (any kind)
      for (Req *r : fresh) {
        try {
          build_device(*r);
          std::unique_lock<std::mutex> lk(mu);
          if (dead[r->kind]) {
            const std::string why = dead_msg[r->kind];
            lk.unlock();
            free_req_device(*r);
            complete({r}, CS_EINTERNAL, why);
            continue;
          }
          r->ready = true;
        } catch (...) {
          int err = 0;
          std::string msg;
          describe(std::current_exception(), err, msg);
          free_req_device(*r);
          complete({r}, err, msg);
        }
      }
      if (!fresh.empty()) cv_work.notify_all();
      // choose this slice: a fair share of every active request's remaining sources
      {
        std::lock_guard<std::mutex> lk(mu);
        std::vector<Req *> mine;
        for (Req *r : active)
          if (r->kind == kind && r->ready && !r->cancelled && r->handed < r->n()) mine.push_back(r);
        if (mine.empty()) continue;
        const double per = ms_per_src[kind] > 0.0 ? ms_per_src[kind] : 0.004;
        // an angular launch lasts about as long as its slowest source whatever
        // its width, so a simplest slice takes everything up to the lane budget
        uint64_t target = kind == CS_REQ_SHORTEST ? (uint64_t)std::clamp(cfg.slice_ms / per, 4096.0, 1048576.0)
                                                  : kSimplestSlice;
        std::vector<uint32_t> take(mine.size(), 0);
        uint64_t left = target;
        for (int pass = 0; pass < 2 && left > 0; ++pass) {
          const uint64_t share = std::max<uint64_t>(1, left / mine.size());
          for (size_t i = 0; i < mine.size() && left > 0; ++i) {
            const uint32_t rem = mine[i]->n() - mine[i]->handed - take[i];
            const uint32_t t = (uint32_t)std::min<uint64_t>({rem, pass == 0 ? share : left, left});
            take[i] += t;
            left -= t;
          }
        }
        for (size_t i = 0; i < mine.size(); ++i) {
          if (!take[i]) continue;
          slice_reqs.push_back(mine[i]);
          slice_take.push_back(take[i]);
          mine[i]->handed += take[i];
          mine[i]->launches += 1;
        }
      }
      if (slice_reqs.empty()) continue;
      int err = 0;
      std::string msg;
      const auto t0 = Clock::now();
      uint64_t n_src = 0;
      try {
        cs_stats st{};
        if (const int f = fault[kind].exchange(0)) {
          if (f == 1) throw std::runtime_error("injected slice failure");
          if (f == 2) throw std::bad_alloc();
        }
        if (kind == CS_REQ_SHORTEST) {
          std::vector<JobSpec> specs;
          std::vector<uint2> items;
          uint32_t NT = 1;
          for (size_t i = 0; i < slice_reqs.size(); ++i) {
            Req &r = *slice_reqs[i];
            // one job per threshold chunk, each filling its rows of the request's accumulators
            for (size_t c = 0; c < r.chunks.size(); ++c) {
              const uint32_t t0 = (uint32_t)c * kMaxT;
              JobSpec js{&r.h->g, r.dg, r.de, r.chunks[c], r.d_src, r.d_wt, r.d_cyc, r.n(), r.d_el,
                         r.d_acc + (size_t)t0 * r.N, r.d_reach + t0, 1.0f};
              js.tstride = r.T;
              const uint32_t job = (uint32_t)specs.size();
              specs.push_back(js);
              NT = std::max(NT, r.chunks[c].T);
              const uint32_t first = r.handed - slice_take[i];
              for (uint32_t k = 0; k < slice_take[i]; ++k) items.push_back(make_uint2(job, first + k));
            }
          }
          n_src = items.size();
          launch_shortest_batch(specs, NT, default_tuning(), stream, st, &items);
        } else {
          std::vector<AngSpec> specs;
          std::vector<uint2> items;
          for (size_t i = 0; i < slice_reqs.size(); ++i) {
            Req &r = *slice_reqs[i];
            for (size_t c = 0; c < r.chunks.size(); ++c) {
              const uint32_t t0 = (uint32_t)c * kMaxT;
              AngSpec as{&r.h->g, r.dg, r.de, r.chunks[c], r.d_src, r.d_wt, r.d_el, r.d_acc + (size_t)t0 * r.N,
                         r.d_reach + t0};
              as.tstride = r.T;
              const uint32_t job = (uint32_t)specs.size();
              specs.push_back(as);
              const uint32_t first = r.handed - slice_take[i];
              for (uint32_t k = 0; k < slice_take[i]; ++k) items.push_back(make_uint2(job, first + k));
            }
          }
          n_src = items.size();
          launch_simplest_jobs(specs, &items, 0, stream, st);
        }
      } catch (...) {
        describe(std::current_exception(), err, msg);
      }
      const double ms = ms_between(t0, Clock::now());
      std::vector<Req *> done;
      {
        std::lock_guard<std::mutex> lk(mu);
        if (n_src) {
          const double per = ms / (double)n_src;
          ms_per_src[kind] = ms_per_src[kind] > 0.0 ? 0.7 * ms_per_src[kind] + 0.3 * per : per;
        }
        for (size_t i = 0; i < slice_reqs.size(); ++i) {
          Req *r = slice_reqs[i];
          r->finished += slice_take[i];
          if (err || r->finished == r->n() || (r->cancelled && r->finished == r->handed)) done.push_back(r);
          if (err) r->handed = r->n();
        }
      }
      for (Req *r : done) {
        int e2 = err;
        std::string m2 = msg;
        if (!err && r->cancelled) {
          e2 = CS_ECANCELED;
          m2 = "request cancelled";
        } else if (!err) {
          try {
            finish(*r);
          } catch (...) {
            describe(std::current_exception(), e2, m2);
          }
        }
        free_req_device(*r);
        complete({r}, e2, m2);
      }
    }
  }
};

namespace {

int64_t submit_impl(cs_server *s, cs_graph *h, int kind, const cs_params *p, double *out, uint32_t *reach) {
  if (!s || !h || !out) fail(CS_EINVAL, "server, graph and out must not be NULL");
  if (kind != CS_REQ_SHORTEST && kind != CS_REQ_SIMPLEST) fail(CS_EINVAL, "unknown request kind");
  Graph &g = h->g;
  if (kind == CS_REQ_SIMPLEST) check_simplest_graph(g);
  check_params(g, p);
  if (kind == CS_REQ_SHORTEST) run_shortest_checks(g, p);
  auto r = std::make_unique<Req>();
  r->h = h;
  r->kind = kind;
  for (uint32_t t0 = 0; t0 < p->n_thresholds; t0 += kMaxT)
    r->chunks.push_back(make_spec(p, t0, chunk_len(p->n_thresholds, t0)));
  r->spec = r->chunks[0];
  r->ss = source_set(g, p);
  if (kind == CS_REQ_SHORTEST) {
    to_device_sources(g, r->ss);
    r->elig = to_device_u8(g, p->eligible);
  } else {
    r->elig.assign(p->eligible, p->eligible + g.n);  // the angular search keeps the original order
  }
  r->out = out;
  r->reach = reach;
  r->betw = p->compute_betweenness != 0;
  r->betw_scale = p->betw_scale;
  r->N = g.n;
  r->T = p->n_thresholds;
  r->M = kind == CS_REQ_SHORTEST ? 7 : 4;
  r->dev_bytes = (size_t)r->N * r->M * r->T * 8 + (size_t)r->n() * 12 + g.n + 4 * r->reach_words();
  if (r->dev_bytes > s->cfg.hbm_budget_bytes)
    fail(CS_ENOMEM, "request needs " + std::to_string(r->dev_bytes >> 20) + " MiB, above the server's HBM budget");
  r->t_submit = Clock::now();
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->stop) fail(CS_EBUSY, "server is shutting down");
  if (s->dead[kind]) fail(CS_EINTERNAL, s->dead_msg[kind]);
  r->id = s->next_id++;
  const int64_t id = r->id;
  Req *raw = r.get();
  s->reqs.emplace(id, std::move(r));
  if (raw->n() == 0) {
    raw->status = 2;  // nothing to search: zero results
    raw->t_admit = raw->t_done = raw->t_submit;
    std::fill(out, out + (size_t)raw->N * raw->M * raw->T, 0.0);
    if (reach)
      for (uint32_t t = 0; t < raw->T; ++t) reach[t] = 0;
    s->cv_done.notify_all();
  } else {
    s->queue.push_back(raw);
    s->cv_work.notify_all();
  }
  return id;
}

void fill_info(const Req &r, cs_request_info *info) {
  if (!info) return;
  info->status = r.status;
  const auto now = Clock::now();
  info->queued_ms = r.status >= 1 || r.status < 0 ? ms_between(r.t_submit, r.t_admit) : ms_between(r.t_submit, now);
  info->latency_ms = (r.status == 2 || r.status < 0) ? ms_between(r.t_submit, r.t_done) : ms_between(r.t_submit, now);
  info->sources = r.n();
  info->launches = r.launches;
}

}  // namespace

extern "C" {

int cs_server_create(const cs_server_config *cfg, cs_server **out) {
  return guarded_call([&] {
    if (!out) fail(CS_EINVAL, "out must not be NULL");
    auto s = std::make_unique<cs_server>();
    s->device = current_device();
    if (cfg) s->cfg = *cfg;
    if (s->cfg.hbm_budget_bytes == 0) {
      size_t free_b = 0, total_b = 0;
      hip_check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
      s->cfg.hbm_budget_bytes = free_b / 4;
    }
    if (s->cfg.max_active == 0) s->cfg.max_active = 64;
    if (!(s->cfg.slice_ms > 0.0)) s->cfg.slice_ms = 100.0;
    cs_server *raw = s.get();
    raw->workers[0] = std::thread([raw] { raw->worker(CS_REQ_SHORTEST); });
    raw->workers[1] = std::thread([raw] { raw->worker(CS_REQ_SIMPLEST); });
    *out = s.release();
  });
}

void cs_server_destroy(cs_server *s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
    for (Req *r : s->queue) {
      r->status = -CS_EBUSY;
      r->msg = "server shut down before the request was admitted";
    }
    s->queue.clear();
  }
  s->cv_work.notify_all();
  for (auto &w : s->workers)
    if (w.joinable()) w.join();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    for (Req *r : s->active) {
      free_req_device(*r);
      r->status = -CS_EBUSY;
      r->msg = "server shut down before the request completed";
    }
    s->active.clear();
    s->cv_done.notify_all();
  }
  delete s;
  (void)hipGetLastError();
}

int64_t cs_server_submit(cs_server *s, cs_graph *g, int kind, const cs_params *p, double *out,
                         uint32_t *reach_totals) {
  int64_t id = -1;
  const int rc = guarded_call([&] { id = submit_impl(s, g, kind, p, out, reach_totals); });
  return rc == CS_OK ? id : -(int64_t)rc;
}

int cs_server_poll(cs_server *s, int64_t id, cs_request_info *info) {
  return guarded_call([&] {
    if (!s) fail(CS_EINVAL, "server must not be NULL");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) fail(CS_EINVAL, "unknown request id " + std::to_string(id));
    fill_info(*it->second, info);
  });
}

int cs_server_wait(cs_server *s, int64_t id, cs_request_info *info) { return cs_server_wait_for(s, id, -1.0, info); }

int cs_server_wait_for(cs_server *s, int64_t id, double timeout_ms, cs_request_info *info) {
  int status = CS_OK;
  std::string msg;
  const int rc = guarded_call([&] {
    if (!s) fail(CS_EINVAL, "server must not be NULL");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) fail(CS_EINVAL, "unknown request id " + std::to_string(id));
    Req &r = *it->second;
    auto finished = [&] { return r.status == 2 || r.status < 0; };
    if (timeout_ms < 0.0) {
      s->cv_done.wait(lk, finished);
    } else if (!s->cv_done.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), finished)) {
      fill_info(r, info);
      status = CS_ETIMEDOUT;
      msg = "request " + std::to_string(id) + " not finished after " + std::to_string(timeout_ms) + " ms";
      return;
    }
    fill_info(r, info);
    status = r.status < 0 ? -r.status : CS_OK;
    msg = r.msg;
  });
  if (rc != CS_OK) return rc;
  if (status != CS_OK) return guarded_call([&] { fail(status, msg); });
  return CS_OK;
}

int cs_server_cancel(cs_server *s, int64_t id) {
  return guarded_call([&] {
    if (!s) fail(CS_EINVAL, "server must not be NULL");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) fail(CS_EINVAL, "unknown request id " + std::to_string(id));
    Req *r = it->second.get();
    if (r->status == 2 || r->status < 0) return;  // already over
    auto q = std::find(s->queue.begin(), s->queue.end(), r);
    if (q != s->queue.end()) {
      s->queue.erase(q);
      s->fail_locked(r, CS_ECANCELED, "request cancelled");
      return;
    }
    r->cancelled = true;  // its worker completes it once no source of it is in flight
    s->cv_work.notify_all();
  });
}

int cs_server_inject_fault(cs_server *s, int kind, int what) {
  return guarded_call([&] {
    if (!s) fail(CS_EINVAL, "server must not be NULL");
    if (kind != CS_REQ_SHORTEST && kind != CS_REQ_SIMPLEST) fail(CS_EINVAL, "unknown request kind");
    if (what < 0 || what > 3) fail(CS_EINVAL, "unknown fault");
    s->fault[kind] = what;
    s->cv_work.notify_all();
  });
}

int cs_server_release(cs_server *s, int64_t id) {
  return guarded_call([&] {
    if (!s) fail(CS_EINVAL, "server must not be NULL");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) fail(CS_EINVAL, "unknown request id " + std::to_string(id));
    if (it->second->status != 2 && it->second->status >= 0) fail(CS_EBUSY, "request still in progress");
    s->reqs.erase(it);
  });
}

}  // extern "C"
