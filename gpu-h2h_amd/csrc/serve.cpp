// Serving: continuous batching of query requests over the resident index.
//
// A scheduler thread owns `streams` device lanes (a HIP stream, pinned host
// staging and device buffers of max_request queries each).  Whenever a lane
// is free it packs the oldest pending requests into one launch -- as many as
// fit the lane and the in-flight admission budget (max_inflight queries on
// the device at once) -- so requests that arrive while others run are
// admitted as soon as a lane and budget free up, and the lanes' copies and
// kernels overlap on the device.  Completion is detected with events; each
// request's slice of the packed result is copied to its own output buffer.
// A request's answer does not depend on what it was packed with: every query
// is answered independently from the read-only index.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <unordered_map>

#include "h2h_internal.h"

namespace h2h {

struct Request {
  uint64_t ticket = 0;
  std::vector<int32_t> x, y;
  int32_t* out = nullptr;
  int64_t n = 0;
  double submit_us = 0, start_us = 0, done_us = 0;
  int state = 0;  // 0 pending, 1 running, 2 done, 3 cancelled
  bool discard = false;
};

struct Lane {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  int32_t *hx = nullptr, *hy = nullptr, *hout = nullptr;  // pinned
  DevArray dx, dy, dout;
  std::vector<std::shared_ptr<Request>> batch;
  std::vector<int64_t> offs;
  int64_t count = 0;
  bool busy = false;
};

struct Server {
  h2h_ctx* ctx = nullptr;
  h2h_server_cfg cfg{};
  std::vector<Lane> lanes;
  std::mutex mu;
  std::condition_variable cv_sched, cv_done;
  std::deque<std::shared_ptr<Request>> pending;
  std::unordered_map<uint64_t, std::shared_ptr<Request>> all;
  uint64_t next_ticket = 1;
  int64_t inflight = 0;
  bool stop = false;
  std::thread thr;
  std::chrono::steady_clock::time_point t0;
  std::string fatal;

  double now_us() const {
    return std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count();
  }

  void run();
  ~Server();
};

void Server::run() {
  (void)hipSetDevice(ctx->device);
  std::unique_lock<std::mutex> lk(mu);
  for (;;) {
    // 1. retire finished lanes
    bool any_busy = false;
    for (auto& L : lanes) {
      if (!L.busy) continue;
      if (hipEventQuery(L.done) == hipErrorNotReady) {
        any_busy = true;
        continue;
      }
      const double t = now_us();
      for (size_t i = 0; i < L.batch.size(); i++) {
        auto& r = L.batch[i];
        if (!r->discard) std::memcpy(r->out, L.hout + L.offs[i], r->n * sizeof(int32_t));
        r->done_us = t;
        r->state = r->discard ? 3 : 2;
      }
      inflight -= L.count;
      L.batch.clear();
      L.offs.clear();
      L.count = 0;
      L.busy = false;
      cv_done.notify_all();
    }
    if (stop && !any_busy) break;
    // 2. admit pending requests onto free lanes
    if (!stop) {
      for (auto& L : lanes) {
        if (L.busy || pending.empty()) continue;
        int64_t cnt = 0;
        while (!pending.empty()) {
          auto& r = pending.front();
          if (r->state == 3) {  // cancelled while pending
            pending.pop_front();
            continue;
          }
          if (cnt + r->n > cfg.max_request) break;
          if (cnt > 0 && inflight + cnt + r->n > cfg.max_inflight) break;
          L.batch.push_back(r);
          L.offs.push_back(cnt);
          std::memcpy(L.hx + cnt, r->x.data(), r->n * sizeof(int32_t));
          std::memcpy(L.hy + cnt, r->y.data(), r->n * sizeof(int32_t));
          cnt += r->n;
          r->state = 1;
          r->start_us = now_us();
          pending.pop_front();
        }
        if (cnt == 0) continue;
        L.count = cnt;
        L.busy = true;
        inflight += cnt;
        hipError_t e = hipMemcpyAsync(L.dx.p, L.hx, cnt * sizeof(int32_t), hipMemcpyHostToDevice, L.stream);
        if (e == hipSuccess) e = hipMemcpyAsync(L.dy.p, L.hy, cnt * sizeof(int32_t), hipMemcpyHostToDevice, L.stream);
        h2h_status st = H2H_OK;
        if (e == hipSuccess)
          st = launch_query(ctx, cfg.kind, cfg.exec, L.dx.as<int32_t>(), L.dy.as<int32_t>(), L.dout.as<int32_t>(), cnt,
                            L.stream);
        if (e == hipSuccess && st == H2H_OK)
          e = hipMemcpyAsync(L.hout, L.dout.p, cnt * sizeof(int32_t), hipMemcpyDeviceToHost, L.stream);
        if (e == hipSuccess) e = hipEventRecord(L.done, L.stream);
        if (e != hipSuccess || st != H2H_OK) {
          fatal = e != hipSuccess ? hipGetErrorString(e) : ctx->err;
          for (auto& r : L.batch) r->state = 3;
          (void)hipStreamSynchronize(L.stream);
          inflight -= cnt;
          L.batch.clear();
          L.offs.clear();
          L.count = 0;
          L.busy = false;
          cv_done.notify_all();
        }
        any_busy = any_busy || L.busy;
      }
    }
    if (any_busy) {
      // device work outstanding: poll at a short interval, wake early on submit
      lk.unlock();
      std::this_thread::sleep_for(std::chrono::microseconds(20));
      lk.lock();
    } else {
      cv_sched.wait_for(lk, std::chrono::milliseconds(50), [&] { return stop || !pending.empty(); });
    }
  }
  for (auto& r : pending) r->state = 3;
  pending.clear();
  cv_done.notify_all();
}

Server::~Server() {
  {
    std::lock_guard<std::mutex> lk(mu);
    stop = true;
  }
  cv_sched.notify_all();
  if (thr.joinable()) thr.join();
  for (auto& L : lanes) {
    if (L.stream) (void)hipStreamSynchronize(L.stream);
    if (L.done) (void)hipEventDestroy(L.done);
    if (L.stream) (void)hipStreamDestroy(L.stream);
    if (L.hx) (void)hipHostFree(L.hx);
    if (L.hy) (void)hipHostFree(L.hy);
    if (L.hout) (void)hipHostFree(L.hout);
  }
}

}  // namespace h2h

h2h_ctx::h2h_ctx() = default;

h2h_ctx::~h2h_ctx() {
  server.reset();
  for (hipEvent_t e : {ev_fork, ev_cl[0], ev_cl[1], ev_cl[2], ev_join2})
    if (e) (void)hipEventDestroy(e);
  if (aux_stream) (void)hipStreamDestroy(aux_stream);
  if (stream) (void)hipStreamDestroy(stream);
}

using namespace h2h;

extern "C" {

static h2h_status server_start_body(h2h_ctx* ctx, const h2h_server_cfg* cfg) {
  if (!ctx || !cfg) return H2H_ERR_ARG;
  if (ctx->device < 0) return set_err(ctx, H2H_ERR_NOGPU, "host-only context");
  if (!ctx->labels_ready) return set_err(ctx, H2H_ERR_STATE, "build labels before serving");
  if (ctx->server) return set_err(ctx, H2H_ERR_STATE, "server already running");
  if (cfg->streams < 1 || cfg->streams > 8 || cfg->max_request < 1 || cfg->max_inflight < cfg->max_request ||
      (cfg->kind != H2H_QUERY_LCA && cfg->kind != H2H_QUERY_NOLCA) ||
      (cfg->exec != H2H_EXEC_PLAIN && cfg->exec != H2H_EXEC_BATCHED))
    return set_err(ctx, H2H_ERR_ARG, "bad server configuration");
  (void)hipSetDevice(ctx->device);
  auto S = std::make_unique<Server>();
  S->ctx = ctx;
  S->cfg = *cfg;
  S->lanes.resize(cfg->streams);
  const size_t bytes = cfg->max_request * sizeof(int32_t);
  for (auto& L : S->lanes) {
    H2H_HIP(ctx, hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking));
    H2H_HIP(ctx, hipEventCreateWithFlags(&L.done, hipEventDisableTiming));
    H2H_HIP(ctx, hipHostMalloc((void**)&L.hx, bytes, hipHostMallocDefault));
    H2H_HIP(ctx, hipHostMalloc((void**)&L.hy, bytes, hipHostMallocDefault));
    H2H_HIP(ctx, hipHostMalloc((void**)&L.hout, bytes, hipHostMallocDefault));
    h2h_status st;
    if ((st = dev_alloc(ctx, L.dx, bytes))) return st;
    if ((st = dev_alloc(ctx, L.dy, bytes))) return st;
    if ((st = dev_alloc(ctx, L.dout, bytes))) return st;
  }
  S->t0 = std::chrono::steady_clock::now();
  Server* raw = S.get();
  S->thr = std::thread([raw] { raw->run(); });
  ctx->server = std::move(S);
  return H2H_OK;
}

h2h_status h2h_server_start(h2h_ctx* ctx, const h2h_server_cfg* cfg) {
  return guarded(ctx, [&] { return server_start_body(ctx, cfg); });
}

static h2h_status submit_body(h2h_ctx* ctx, const int32_t* x, const int32_t* y, int32_t* out, int64_t n, uint64_t* ticket) {
  if (!ctx || !ticket || n < 1 || !x || !y || !out) return H2H_ERR_ARG;
  Server* S = ctx->server.get();
  if (!S) return H2H_ERR_STATE;
  if (n > S->cfg.max_request) return H2H_ERR_ADMIT;
  for (int64_t i = 0; i < n; i++)
    if (x[i] < 0 || x[i] >= ctx->n || y[i] < 0 || y[i] >= ctx->n) return H2H_ERR_ARG;
  auto r = std::make_shared<Request>();
  r->x.assign(x, x + n);
  r->y.assign(y, y + n);
  r->out = out;
  r->n = n;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->stop) return H2H_ERR_CANCELLED;
    r->ticket = S->next_ticket++;
    r->submit_us = S->now_us();
    S->all[r->ticket] = r;
    S->pending.push_back(r);
    *ticket = r->ticket;
  }
  S->cv_sched.notify_one();
  return H2H_OK;
}

h2h_status h2h_submit(h2h_ctx* ctx, const int32_t* x, const int32_t* y, int32_t* out, int64_t n, uint64_t* ticket) {
  return guarded(ctx, [&] { return submit_body(ctx, x, y, out, n, ticket); });
}

h2h_status h2h_wait(h2h_ctx* ctx, uint64_t ticket, int timeout_ms, h2h_req_stats* stats) {
  if (!ctx) return H2H_ERR_ARG;
  Server* S = ctx->server.get();
  if (!S) return H2H_ERR_STATE;
  std::unique_lock<std::mutex> lk(S->mu);
  auto it = S->all.find(ticket);
  if (it == S->all.end()) return H2H_ERR_ARG;
  auto r = it->second;
  auto ready = [&] { return r->state >= 2; };
  if (timeout_ms < 0) S->cv_done.wait(lk, ready);
  else if (!S->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready)) return H2H_ERR_TIMEOUT;
  if (stats) {
    stats->submit_us = r->submit_us;
    stats->start_us = r->start_us;
    stats->done_us = r->done_us;
    stats->n = r->n;
  }
  S->all.erase(ticket);
  return r->state == 2 ? H2H_OK : H2H_ERR_CANCELLED;
}

h2h_status h2h_cancel(h2h_ctx* ctx, uint64_t ticket) {
  if (!ctx) return H2H_ERR_ARG;
  Server* S = ctx->server.get();
  if (!S) return H2H_ERR_STATE;
  std::lock_guard<std::mutex> lk(S->mu);
  auto it = S->all.find(ticket);
  if (it == S->all.end()) return H2H_ERR_ARG;
  auto& r = it->second;
  if (r->state == 0) {
    r->state = 3;
    S->cv_done.notify_all();
  } else if (r->state == 1) {
    r->discard = true;  // results of the running launch are dropped
  }
  return H2H_OK;
}

h2h_status h2h_server_stop(h2h_ctx* ctx) {
  if (!ctx) return H2H_ERR_ARG;
  if (!ctx->server) return H2H_ERR_STATE;
  ctx->server.reset();
  return H2H_OK;
}

}  // extern "C"
