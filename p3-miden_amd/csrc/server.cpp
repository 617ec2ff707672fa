// Prove-request server: continuous batching of independent PCS proofs on one GPU.
//
// Each request is the pcs_trace flow of p3g_prove (commit_traces, observe the
// root, sample two extension points, open_with_channel) over its own traces.
// The server owns W worker contexts, each with its own stream and allocation
// cache, so up to W requests are on the device at once and their kernels,
// copies and host-side transcript work overlap.  Requests are admitted in
// arrival order as soon as a worker is free AND the request's device footprint
// fits in the HBM budget next to the requests already in flight; a request
// that does not fit lets later, smaller ones go first (backfill) until it has
// waited longer than the backfill window, after which nothing overtakes it.
// The workers draw device memory from one shared block pool (blocks change
// hands between worker streams through events), so a block freed by one
// request is reused by the next whichever worker runs it, and the pool never
// holds more than the device can give.  A proof depends only on its request,
// never on what runs beside it, so each result is byte-identical to running
// the request alone.
#include <atomic>
#include <chrono>
#include <cstdio>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <utility>
#include <vector>

#include "../../include/p3gpu.h"
#include "handles.hpp"

using namespace p3g;

namespace {

using clk = std::chrono::steady_clock;

struct Request {
  uint64_t ticket = 0;
  std::string sig;  // hash, parameters, input kind and trace shapes: requests with one sig batch together
  uint32_t hash = 0;
  p3g_pcs_params params{};
  std::vector<p3g_matrix> traces;
  uint32_t flags = 0;
  uint64_t footprint = 0;
  clk::time_point t_submit, t_admit, t_done;
  bool admitted = false, done = false;
  int status = P3G_OK;
  std::atomic<int> cancel{0};  // set by p3g_server_cancel while the request runs
  bool inject_alloc_failure = false;
  uint32_t worker = 0;
  uint64_t root[4] = {0, 0, 0, 0}, z[4] = {0, 0, 0, 0};
  std::vector<uint64_t> fields, comms;
  std::string error;
  // device copies of host traces made ahead of admission by the upload stage
  uint64_t trace_bytes = 0;         // host trace bytes (0 for device inputs)
  bool reserved = false;            // footprint counted in flight ahead of admission (upload stage)
  bool staging = false;             // the upload stage is enqueuing its copies
  bool staged = false;              // copies enqueued: staged_ptrs and staged_ev are set
  std::vector<void*> staged_ptrs;   // one per trace (nullptr for an empty one)
  hipEvent_t staged_ev = nullptr;   // recorded on the upload stream after the copies
};

}  // namespace

struct p3g_server {
  int device = 0;
  uint64_t budget = 0;
  double backfill_ms = 0;
  std::vector<p3g_ctx*> ctxs;
  SharedPool pool;
  std::vector<std::thread> threads;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Request*> queue;
  std::map<uint64_t, Request*> reqs;
  uint64_t next_ticket = 1;
  uint64_t inflight_bytes = 0, peak_inflight_bytes = 0;
  uint32_t running = 0, peak_running = 0;
  uint64_t admitted = 0, completed = 0, batches = 0, batched_requests = 0;
  uint32_t max_batch = 1;
  double batch_window_ms = 0;
  bool stop = false;
  clk::time_point t0;
  // upload stage: copies the host traces of up to upload_ahead queued requests into
  // pool blocks on its own stream while the workers prove earlier requests
  uint32_t upload_ahead = 0;
  hipStream_t upload_stream = nullptr;
  std::thread stager;
  uint32_t n_staged = 0;  // staged or staging, not yet admitted

  void worker_loop(uint32_t w);
  void stager_loop();
  // gives a staged request's copies back to the pool (s: the stream whose work last used them)
  void unstage_locked(Request* r, hipStream_t s);
  // the device-pointer form of a staged request's traces
  std::vector<p3g_matrix> staged_traces(const Request* r) const;
  std::deque<Request*>::iterator pick_locked();
  double batch_wait_locked(const Request* r) const;
  void take_batch_locked(Request* first, std::vector<Request*>& batch);
  void run_one(p3g_ctx* ctx, Request* r);
  void run_batch(p3g_ctx* ctx, std::vector<Request*>& batch);
};

// First queued request that fits (queue.end() if none); the queue head may be
// overtaken only while it is younger than the backfill window.
std::deque<Request*>::iterator p3g_server::pick_locked() {
  const auto now = clk::now();
  for (auto it = queue.begin(); it != queue.end(); ++it) {
    Request* r = *it;
    // the upload stage reserved this one's memory already: it only needs a worker
    if (r->reserved || inflight_bytes + r->footprint <= budget) return it;
    const double waited = std::chrono::duration<double, std::milli>(now - r->t_submit).count();
    if (waited > backfill_ms) return queue.end();
  }
  return queue.end();
}

// Milliseconds to hold a picked request so that more of its signature can join
// its batch: while it is younger than the batch window and fewer than
// max_batch requests of its signature are queued.  0 = take it now.
double p3g_server::batch_wait_locked(const Request* r) const {
  if (batch_window_ms <= 0 || max_batch <= 1) return 0;
  const double age = std::chrono::duration<double, std::milli>(clk::now() - r->t_submit).count();
  if (age >= batch_window_ms) return 0;
  uint32_t same = 0;
  for (const Request* q : queue)
    if (q->sig == r->sig && ++same >= max_batch) return 0;
  return batch_window_ms - age;
}

// Later queued requests with the first one's signature join its batch while
// the batch stays under max_batch and the budget holds.
void p3g_server::take_batch_locked(Request* first, std::vector<Request*>& batch) {
  batch.assign(1, first);
  if (max_batch <= 1) return;
  uint64_t extra = first->footprint;  // not yet counted in inflight_bytes
  for (auto it = queue.begin(); it != queue.end() && batch.size() < max_batch;) {
    Request* r = *it;
    if (r->sig == first->sig && inflight_bytes + extra + r->footprint <= budget) {
      extra += r->footprint;
      batch.push_back(r);
      it = queue.erase(it);
    } else {
      ++it;
    }
  }
}

std::vector<p3g_matrix> p3g_server::staged_traces(const Request* r) const {
  std::vector<p3g_matrix> t = r->traces;
  for (size_t i = 0; i < t.size(); i++)
    if (r->staged_ptrs[i]) t[i].data = static_cast<const uint64_t*>(r->staged_ptrs[i]);
  return t;
}

void p3g_server::unstage_locked(Request* r, hipStream_t s) {
  for (void* p : r->staged_ptrs)
    if (p) pool.release(p, s);
  r->staged_ptrs.clear();
  if (r->staged_ev) (void)hipEventDestroy(r->staged_ev);
  r->staged_ev = nullptr;
  r->staged = false;
}

void p3g_server::run_one(p3g_ctx* ctx, Request* r) {
  p3g_transcript* tx = nullptr;
  int rc = p3g_transcript_create(r->hash == P3G_HASH_POSEIDON2_GL12 ? P3G_CHALLENGER_DUPLEX_P2 : P3G_CHALLENGER_KECCAK,
                                 &tx);
  // the proof stops at its next phase boundary once the request is canceled
  ctx->c.cancel = &r->cancel;
  if (r->inject_alloc_failure) ctx->c.fail_alloc_after = 2;
  if (rc == P3G_OK) {
    if (r->staged && hipStreamWaitEvent(ctx->c.stream, r->staged_ev, 0) == hipSuccess) {
      // the traces were copied ahead: prove from the device copies
      const std::vector<p3g_matrix> t = staged_traces(r);
      rc = p3g_prove(ctx, r->hash, &r->params, t.data(), t.size(), P3G_DEVICE_PTRS, tx, r->root, r->z);
    } else {
      (void)hipGetLastError();
      rc = p3g_prove(ctx, r->hash, &r->params, r->traces.data(), r->traces.size(), r->flags & P3G_DEVICE_PTRS, tx,
                     r->root, r->z);
    }
  }
  if (r->staged) {
    std::lock_guard<std::mutex> g(mu);
    unstage_locked(r, ctx->c.stream);
  }
  ctx->c.cancel = nullptr;
  ctx->c.fail_alloc_after = -1;
  if (rc == P3G_OK && r->cancel.load()) rc = P3G_E_CANCELED;  // canceled after its last checkpoint
  if (rc == P3G_OK) {
    r->fields = std::move(tx->tx.fields);  // the transcript is freed below
    r->comms = std::move(tx->tx.comms);
  } else {
    r->error = p3g_last_error();
  }
  p3g_transcript_free(tx);
  r->status = rc;
}

void p3g_server::run_batch(p3g_ctx* ctx, std::vector<Request*>& batch) {
  const size_t k = batch.size(), t = batch[0]->traces.size();
  std::vector<p3g_transcript*> txs(k, nullptr);
  std::vector<p3g_matrix> flat;
  int rc = P3G_OK;
  for (size_t j = 0; j < k && rc == P3G_OK; j++) {
    rc = p3g_transcript_create(batch[j]->hash == P3G_HASH_POSEIDON2_GL12 ? P3G_CHALLENGER_DUPLEX_P2
                                                                          : P3G_CHALLENGER_KECCAK,
                               &txs[j]);
    flat.insert(flat.end(), batch[j]->traces.begin(), batch[j]->traces.end());
  }
  std::vector<uint64_t> roots(4 * k), zs(4 * k);
  // a failure injected into one member fails the lock-step batch, as a real one would
  for (auto* r : batch)
    if (r->inject_alloc_failure) ctx->c.fail_alloc_after = 2;
  // members copied ahead by the upload stage: when all were, prove from the copies
  bool all_staged = true;
  for (auto* r : batch) all_staged = all_staged && r->staged;
  uint32_t in_flags = batch[0]->flags & P3G_DEVICE_PTRS;
  if (all_staged) {
    flat.clear();
    for (auto* r : batch) {
      if (hipStreamWaitEvent(ctx->c.stream, r->staged_ev, 0) != hipSuccess) rc = P3G_E_DEVICE;
      const std::vector<p3g_matrix> tr = staged_traces(r);
      flat.insert(flat.end(), tr.begin(), tr.end());
    }
    in_flags = P3G_DEVICE_PTRS;
  }
  if (rc == P3G_OK)
    rc = p3g_prove_many(ctx, batch[0]->hash, &batch[0]->params, flat.data(), k, t, in_flags, txs.data(), roots.data(),
                        zs.data());
  ctx->c.fail_alloc_after = -1;
  {
    std::lock_guard<std::mutex> g(mu);
    for (auto* r : batch)
      if (r->staged) unstage_locked(r, ctx->c.stream);
  }
  if (rc != P3G_OK) {
    // a result depends only on its own request: when the batch fails, every member
    // runs alone, so that only the requests that fail by themselves report an error
    for (size_t j = 0; j < k; j++) p3g_transcript_free(txs[j]);
    for (Request* r : batch) run_one(ctx, r);
    return;
  }
  for (size_t j = 0; j < k; j++) {
    Request* r = batch[j];
    // a member canceled while the batch ran keeps no result (the others are unaffected)
    r->status = r->cancel.load() ? P3G_E_CANCELED : P3G_OK;
    if (r->status == P3G_OK) {
      memcpy(r->root, roots.data() + 4 * j, 32);
      memcpy(r->z, zs.data() + 4 * j, 32);
      r->fields = std::move(txs[j]->tx.fields);  // the transcript is freed below
      r->comms = std::move(txs[j]->tx.comms);
    } else {
      r->error = "request canceled";
    }
    p3g_transcript_free(txs[j]);
  }
}

void p3g_server::worker_loop(uint32_t w) {
  (void)hipSetDevice(device);
  p3g_ctx* ctx = ctxs[w];
  std::vector<Request*> batch;
  for (;;) {
    {
      std::unique_lock<std::mutex> lk(mu);
      Request* r = nullptr;
      for (;;) {
        if (stop) return;
        auto it = pick_locked();
        if (it != queue.end()) {
          const double hold = batch_wait_locked(*it);
          if (hold <= 0) {
            r = *it;
            queue.erase(it);
            break;
          }
          // inside the batch window: let more requests of this shape arrive
          cv_work.wait_for(lk, std::chrono::duration<double, std::milli>(hold));
          continue;
        }
        // re-check on every completion and submission; a timed wait lets an
        // aged queue head stop the backfill
        cv_work.wait_for(lk, std::chrono::milliseconds(5));
      }
      take_batch_locked(r, batch);
      const auto now = clk::now();
      for (auto* b : batch) {
        b->admitted = true;
        b->t_admit = now;
        b->worker = w;
        if (!b->reserved) inflight_bytes += b->footprint;
      }
      // a member whose copies the upload stage is enqueuing right now: wait for them
      // (the worker would otherwise copy the same traces itself)
      for (auto* b : batch) {
        cv_work.wait(lk, [&] { return !b->staging; });
        if (b->reserved) n_staged--;
      }
      running += (uint32_t)batch.size();
      admitted += batch.size();
      if (batch.size() > 1) {
        batches++;
        batched_requests += batch.size();
      }
      if (inflight_bytes > peak_inflight_bytes) peak_inflight_bytes = inflight_bytes;
      if (running > peak_running) peak_running = running;
    }
    if (batch.size() == 1) run_one(ctx, batch[0]);
    else run_batch(ctx, batch);
    {
      std::lock_guard<std::mutex> g(mu);
      const auto now = clk::now();
      for (auto* b : batch) {
        b->done = true;
        b->t_done = now;
        inflight_bytes -= b->footprint;
      }
      running -= (uint32_t)batch.size();
      completed += batch.size();
    }
    cv_done.notify_all();
    cv_work.notify_all();
  }
}

void p3g_server::stager_loop() {
  (void)hipSetDevice(device);
  for (;;) {
    Request* r = nullptr;
    {
      std::unique_lock<std::mutex> lk(mu);
      for (;;) {
        if (stop) return;
        if (n_staged < upload_ahead) {
          // the first queued request with host traces and no copies yet, in arrival
          // order, once its copies fit in the budget beside what is held already
          for (Request* q : queue) {
            if (q->reserved || (q->flags & P3G_DEVICE_PTRS) || !q->trace_bytes) continue;
            if (inflight_bytes + q->footprint <= budget) r = q;
            break;
          }
        }
        if (r) break;
        cv_work.wait_for(lk, std::chrono::milliseconds(5));
      }
      // admitted as far as memory goes: its whole footprint is reserved now
      r->reserved = true;
      r->staging = true;
      n_staged++;
      inflight_bytes += r->footprint;
      if (inflight_bytes > peak_inflight_bytes) peak_inflight_bytes = inflight_bytes;
    }
    // enqueue the copies outside the lock; a failure leaves the request to its worker
    std::vector<void*> ptrs(r->traces.size(), nullptr);
    hipEvent_t ev = nullptr;
    bool ok = true;
    try {
      for (size_t i = 0; i < r->traces.size(); i++) {
        const uint64_t bytes = 8 * r->traces[i].height * r->traces[i].width;
        if (!bytes) continue;
        ptrs[i] = pool.alloc(bytes, upload_stream);
        if (hipMemcpyAsync(ptrs[i], r->traces[i].data, bytes, hipMemcpyHostToDevice, upload_stream) != hipSuccess)
          ok = false;
      }
      if (ok && (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess ||
                 hipEventRecord(ev, upload_stream) != hipSuccess))
        ok = false;
    } catch (...) {
      ok = false;
    }
    if (!ok) (void)hipGetLastError();
    {
      std::lock_guard<std::mutex> g(mu);
      r->staging = false;
      r->staged_ptrs = ptrs;
      r->staged_ev = ev;
      r->staged = true;
      // without copies (a failure) the worker uploads the traces itself, within the
      // reservation
      if (!ok || r->cancel.load()) unstage_locked(r, upload_stream);
      if (r->cancel.load() && !r->admitted) {
        // canceled while its copies were being enqueued: complete the cancellation
        for (auto q = queue.begin(); q != queue.end(); ++q)
          if (*q == r) {
            queue.erase(q);
            break;
          }
        r->status = P3G_E_CANCELED;
        r->error = "request canceled";
        r->done = true;
        r->t_admit = r->t_done = clk::now();
        n_staged--;
        inflight_bytes -= r->footprint;
      }
    }
    cv_work.notify_all();
    cv_done.notify_all();
  }
}

extern "C" {

uint64_t p3g_request_footprint(const p3g_request* req) {
  if (!req || !req->traces || req->k == 0) return 0;
  const uint64_t B = 1ull << req->params.log_blowup;
  uint64_t trace = 0, lde = 0, L = 0, hmin = ~0ull, hmax = 0;
  for (uint64_t i = 0; i < req->k; i++) {
    const uint64_t n = req->traces[i].height, w = req->traces[i].width;
    trace += 8 * n * w;
    lde += 8 * n * B * w;
    if (n * B > L) L = n * B;
    if (n < hmin) hmin = n;
    if (n > hmax) hmax = n;
  }
  const uint64_t W = req->hash_kind == P3G_HASH_POSEIDON2_GL12 ? 12 : 25;
  uint64_t bytes = lde + 64 * L;                        // LDEs and digest layers
  if (!(req->flags & P3G_DEVICE_PTRS)) bytes += trace;  // device copies of host traces
  if (hmin != hmax) bytes += 2 * L * W * 8;             // lifted sponge states
  bytes += 112 * L + 32 * (L / B);                      // DEEP evaluations, weights, FRI layers and trees
  return bytes + bytes / 8 + (64ull << 20);             // allocator rounding and tables
}

int p3g_server_set_batch_window(p3g_server* s, double batch_window_ms) {
  if (!s || !(batch_window_ms >= 0) || batch_window_ms > 10000) return P3G_E_INVAL;
  std::lock_guard<std::mutex> g(s->mu);
  s->batch_window_ms = batch_window_ms;
  s->cv_work.notify_all();
  return P3G_OK;
}

int p3g_server_set_upload_ahead(p3g_server* s, uint32_t requests) {
  if (!s || requests > 16) return P3G_E_INVAL;
  {
    std::lock_guard<std::mutex> g(s->mu);
    if (requests && !s->upload_stream) {
      (void)hipSetDevice(s->device);
      if (hipStreamCreateWithFlags(&s->upload_stream, hipStreamNonBlocking) != hipSuccess) {
        (void)hipGetLastError();
        s->upload_stream = nullptr;
        return P3G_E_DEVICE;
      }
      s->stager = std::thread([s] { s->stager_loop(); });
    }
    s->upload_ahead = requests;
  }
  s->cv_work.notify_all();
  return P3G_OK;
}

int p3g_server_create(int device, uint32_t workers, uint64_t hbm_budget_bytes, double backfill_ms, uint32_t max_batch,
                      p3g_server** out) {
  if (!out || workers == 0 || workers > 16 || hbm_budget_bytes == 0 || max_batch == 0 || max_batch > 4096)
    return P3G_E_INVAL;
  *out = nullptr;
  p3g_server* s = new p3g_server();
  s->device = device;
  s->budget = hbm_budget_bytes;
  s->backfill_ms = backfill_ms;
  s->max_batch = max_batch;
  s->t0 = clk::now();
  for (uint32_t w = 0; w < workers; w++) {
    p3g_ctx* c = nullptr;
    int rc = p3g_ctx_create(device, &c);
    if (rc != P3G_OK) {
      for (auto* x : s->ctxs) p3g_ctx_destroy(x);
      delete s;
      return rc;
    }
    c->c.pool = &s->pool;
    s->ctxs.push_back(c);
  }
  // the pool may hold what the device has free now, less a margin for tables
  // and the callers' own buffers; the admission budget never exceeds it
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
    (void)hipGetLastError();
    free_b = (size_t)hbm_budget_bytes;
  }
  const size_t margin = (size_t)4 << 30;
  s->pool.limit = free_b > 2 * margin ? free_b - margin : free_b / 2;
  if (s->budget > s->pool.limit) s->budget = s->pool.limit;
  for (uint32_t w = 0; w < workers; w++) s->threads.emplace_back([s, w] { s->worker_loop(w); });
  *out = s;
  return P3G_OK;
}

int p3g_server_submit(p3g_server* s, const p3g_request* req, uint64_t* ticket) {
  if (!s || !req || !ticket || !req->traces || req->k == 0) return P3G_E_INVAL;
  if (req->hash_kind > P3G_HASH_KECCAK_GL) return P3G_E_INVAL;
  // a request is checked alone here, so that a bad one never reaches a lock-step batch
  // (where its error would be reported to its batch-mates too)
  for (uint64_t i = 0; i < req->k; i++)
    if (!req->traces[i].data && req->traces[i].height * req->traces[i].width) return P3G_E_INVAL;
  {
    const p3g_pcs_params& q = req->params;
    try {
      check_pcs_params(PcsParams{q.log_blowup, q.log_folding_arity, q.log_final_degree, q.folding_pow_bits,
                                 q.deep_pow_bits, q.num_queries, q.query_pow_bits});
    } catch (const ApiError&) {
      return P3G_E_INVAL;
    }
  }
  const uint64_t fp = p3g_request_footprint(req);
  if (fp > s->budget) return P3G_E_RANGE;
  Request* r = new Request();
  r->hash = req->hash_kind;
  r->params = req->params;
  r->traces.assign(req->traces, req->traces + req->k);
  r->flags = req->flags;
  r->inject_alloc_failure = (req->flags & P3G_INJECT_ALLOC_FAILURE) != 0;
  r->footprint = fp;
  if (!(req->flags & P3G_DEVICE_PTRS))
    for (uint64_t i = 0; i < req->k; i++) r->trace_bytes += 8 * req->traces[i].height * req->traces[i].width;
  {
    char buf[96];
    const p3g_pcs_params& q = req->params;
    snprintf(buf, sizeof(buf), "%u/%u/%u.%u.%u.%u.%u.%u.%u", req->hash_kind, req->flags & P3G_DEVICE_PTRS, q.log_blowup,
             q.log_folding_arity, q.log_final_degree, q.folding_pow_bits, q.deep_pow_bits, q.num_queries,
             q.query_pow_bits);
    r->sig = buf;
    for (uint64_t i = 0; i < req->k; i++)
      r->sig += ":" + std::to_string(req->traces[i].height) + "x" + std::to_string(req->traces[i].width);
  }
  {
    std::lock_guard<std::mutex> g(s->mu);
    r->ticket = s->next_ticket++;
    r->t_submit = clk::now();
    s->reqs[r->ticket] = r;
    s->queue.push_back(r);
    *ticket = r->ticket;
  }
  s->cv_work.notify_all();
  return P3G_OK;
}

static double ms_since(const p3g_server* s, clk::time_point t) {
  return std::chrono::duration<double, std::milli>(t - s->t0).count();
}

int p3g_server_wait(p3g_server* s, uint64_t ticket, p3g_result_info* info) {
  if (!s) return P3G_E_INVAL;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return P3G_E_RANGE;
  Request* r = it->second;
  s->cv_done.wait(lk, [&] { return r->done || s->stop; });
  if (!r->done) return P3G_E_INVAL;
  if (info) {
    info->status = r->status;
    memcpy(info->root, r->root, 32);
    memcpy(info->eval_points, r->z, 32);
    info->num_fields = r->fields.size();
    info->num_commitments = r->comms.size() / 4;
    info->submit_ms = ms_since(s, r->t_submit);
    info->admit_ms = ms_since(s, r->t_admit);
    info->done_ms = ms_since(s, r->t_done);
    info->worker = r->worker;
    info->footprint_bytes = r->footprint;
  }
  return r->status;
}

int p3g_server_poll(p3g_server* s, uint64_t ticket, int* done) {
  if (!s || !done) return P3G_E_INVAL;
  std::lock_guard<std::mutex> g(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return P3G_E_RANGE;
  *done = it->second->done ? 1 : 0;
  return P3G_OK;
}

int p3g_server_proof(p3g_server* s, uint64_t ticket, uint64_t* fields, uint64_t* commitments) {
  if (!s) return P3G_E_INVAL;
  std::lock_guard<std::mutex> g(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return P3G_E_RANGE;
  Request* r = it->second;
  if (!r->done) return P3G_E_INVAL;
  if (r->status != P3G_OK) return r->status;
  if (fields && !r->fields.empty()) memcpy(fields, r->fields.data(), r->fields.size() * 8);
  if (commitments && !r->comms.empty()) memcpy(commitments, r->comms.data(), r->comms.size() * 8);
  return P3G_OK;
}

int p3g_server_cancel(p3g_server* s, uint64_t ticket) {
  if (!s) return P3G_E_INVAL;
  {
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return P3G_E_RANGE;
    Request* r = it->second;
    if (r->done) return P3G_OK;  // finished first: its result stands
    if (!r->admitted && r->staging) {
      r->cancel.store(1);  // the upload stage completes the cancellation when its copies are enqueued
    } else if (!r->admitted) {
      for (auto q = s->queue.begin(); q != s->queue.end(); ++q)
        if (*q == r) {
          s->queue.erase(q);
          break;
        }
      if (r->staged) s->unstage_locked(r, s->upload_stream);
      if (r->reserved) {
        s->n_staged--;
        s->inflight_bytes -= r->footprint;
      }
      r->status = P3G_E_CANCELED;
      r->error = "request canceled";
      r->done = true;
      r->t_admit = r->t_done = clk::now();
    } else {
      r->cancel.store(1);
    }
  }
  s->cv_done.notify_all();
  s->cv_work.notify_all();
  return P3G_OK;
}

int p3g_server_release(p3g_server* s, uint64_t ticket) {
  if (!s) return P3G_E_INVAL;
  std::lock_guard<std::mutex> g(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return P3G_E_RANGE;
  if (!it->second->done) return P3G_E_INVAL;
  delete it->second;
  s->reqs.erase(it);
  return P3G_OK;
}

int p3g_server_stats(p3g_server* s, p3g_server_counters* out) {
  if (!s || !out) return P3G_E_INVAL;
  std::lock_guard<std::mutex> g(s->mu);
  out->admitted = s->admitted;
  out->completed = s->completed;
  out->queued = s->queue.size();
  out->running = s->running;
  out->batches = s->batches;
  out->batched_requests = s->batched_requests;
  out->inflight_bytes = s->inflight_bytes;
  out->peak_inflight_bytes = s->peak_inflight_bytes;
  out->peak_running = s->peak_running;
  out->budget_bytes = s->budget;
  std::lock_guard<std::mutex> g2(s->pool.mu);
  out->n_malloc = s->pool.n_malloc;
  out->n_free = s->pool.n_free;
  out->malloc_ms = s->pool.malloc_ms;
  out->free_ms = s->pool.free_ms;
  return P3G_OK;
}

int p3g_server_trim(p3g_server* s) {
  if (!s) return P3G_E_INVAL;
  s->pool.trim();
  return P3G_OK;
}

void p3g_server_destroy(p3g_server* s) {
  if (!s) return;
  {
    std::unique_lock<std::mutex> lk(s->mu);
    // queued requests are dropped; running ones (and copies being enqueued) finish first
    s->cv_work.wait(lk, [&] {
      for (auto* r : s->queue)
        if (r->staging) return false;
      return true;
    });
    for (auto* r : s->queue) {
      if (r->staged) s->unstage_locked(r, s->upload_stream);
      if (r->reserved) s->inflight_bytes -= r->footprint;
      r->status = P3G_E_INVAL;
      r->done = true;
    }
    s->queue.clear();
    s->cv_done.wait(lk, [&] { return s->running == 0; });
    s->stop = true;
  }
  s->cv_work.notify_all();
  s->cv_done.notify_all();
  for (auto& t : s->threads) t.join();
  if (s->stager.joinable()) s->stager.join();
  if (s->upload_stream) {
    (void)hipStreamSynchronize(s->upload_stream);
    s->pool.forget_stream(s->upload_stream);
    (void)hipStreamDestroy(s->upload_stream);
  }
  for (auto& kv : s->reqs) delete kv.second;
  for (auto* c : s->ctxs) p3g_ctx_destroy(c);
  delete s;
}

int p3g_host_alloc(uint64_t bytes, void** out) {
  if (!out) return P3G_E_INVAL;
  *out = nullptr;
  if (hipHostMalloc(out, bytes ? bytes : 8, hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    return P3G_E_NOMEM;
  }
  return P3G_OK;
}

int p3g_host_free(void* p) {
  if (p && hipHostFree(p) != hipSuccess) {
    (void)hipGetLastError();
    return P3G_E_DEVICE;
  }
  return P3G_OK;
}

}  // extern "C"
