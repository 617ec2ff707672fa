// Device context: stream, Poseidon2 constants, cached power tables and a
// size-bucketed cache of device allocations (the prover allocates the same
// shapes proof after proof, so blocks are recycled instead of hipMalloc'd).
#include <cstring>
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <exception>
#include <functional>
#include <thread>

#include "host_pool.hpp"
#include "internal.hpp"

namespace p3g {

// ------------------------------------------------------------------ host thread pool
namespace {
HostPool& host_pool() {
  static HostPool p([] {
    const unsigned hw = std::thread::hardware_concurrency();
    return std::min(16u, hw > 1 ? hw : 1u);
  }());
  return p;
}
}  // namespace

void parallel_for(size_t n, const std::function<void(size_t)>& f) { host_pool().run(n, f); }

static double ms_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

void* SharedPool::take(std::multimap<size_t, Block>::iterator it, hipStream_t s) {
  Block b = it->second;
  free_blocks.erase(it);
  cached_bytes -= b.sz;
  if (b.s && b.s != s) {
    hipError_t e = hipSuccess;
    if (b.ev) {
      e = hipStreamWaitEvent(s, b.ev, 0);  // exactly the previous owner's work
      (void)hipEventDestroy(b.ev);         // not recycled while the wait is pending
      b.ev = nullptr;
    } else {
      // s waits for everything queued on the releasing stream so far
      hipEvent_t ev = nullptr;
      e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
      if (e == hipSuccess) e = hipEventRecord(ev, b.s);
      if (e == hipSuccess) e = hipStreamWaitEvent(s, ev, 0);
      if (ev) (void)hipEventDestroy(ev);
    }
    if (e != hipSuccess) {
      (void)hipGetLastError();
      (void)hipStreamSynchronize(b.s);  // without the event: wait for the old owner here
    }
    n_cross_stream++;
  }
  if (b.ev) spare_events.push_back(b.ev);
  live[b.p] = b.sz;
  live_bytes += b.sz;
  return b.p;
}

void SharedPool::free_one(std::multimap<size_t, Block>::iterator it) {
  Block b = it->second;
  free_blocks.erase(it);
  cached_bytes -= b.sz;
  const auto t0 = std::chrono::steady_clock::now();
  if (b.ev) {
    (void)hipEventSynchronize(b.ev);
    spare_events.push_back(b.ev);
  } else if (b.s) {
    (void)hipStreamSynchronize(b.s);
  }
  (void)hipFree(b.p);
  n_free++;
  free_ms += ms_since(t0);
}

// Block sizes are rounded up to 16 classes per octave (at most 6.25% over the
// request; 256-byte steps below 4 KiB), so requests whose sizes differ a little,
// such as the same phase of batches of 63 and 64 jobs, map to the same cached
// blocks instead of each growing the pool.
static size_t size_class(size_t bytes) {
  if (bytes <= 4096) return bytes == 0 ? 256 : (bytes + 255) & ~size_t(255);
  int e = 63 - __builtin_clzll((unsigned long long)bytes);
  const size_t step = size_t(1) << (e - 4);
  return (bytes + step - 1) & ~(step - 1);
}

// The first block of [lo, hi) that another stream can take without waiting
// for work queued after the block's release: one of stream s or with no
// pending work, else one with an exact release event, else one whose stream
// has drained.  hi if there is none.
std::multimap<size_t, SharedPool::Block>::iterator SharedPool::pick(std::multimap<size_t, Block>::iterator lo,
                                                                    std::multimap<size_t, Block>::iterator hi,
                                                                    hipStream_t s) {
  std::multimap<size_t, Block>::iterator evented = hi;
  for (auto c = lo; c != hi; ++c) {
    if (!c->second.s || c->second.s == s) return c;
    if (c->second.ev && evented == hi) evented = c;
  }
  if (evented != hi) return evented;
  hipStream_t busy[16];
  int nbusy = 0;
  for (auto c = lo; c != hi; ++c) {
    const hipStream_t bs = c->second.s;
    if (std::find(busy, busy + nbusy, bs) != busy + nbusy) continue;
    if (hipStreamQuery(bs) == hipSuccess) {
      for (auto& kv : free_blocks)
        if (kv.second.s == bs && !kv.second.ev) kv.second.s = nullptr;  // all of bs's work is done
      return c;
    }
    (void)hipGetLastError();  // hipErrorNotReady
    if (nbusy < 16) busy[nbusy++] = bs;
  }
  return hi;
}

void* SharedPool::alloc(size_t bytes, hipStream_t s) {
  bytes = size_class(bytes);
  std::lock_guard<std::mutex> g(mu);
  const auto lo = free_blocks.lower_bound(bytes);
  // a cached block up to 25% larger serves a request; for large requests (whose
  // hipMalloc costs milliseconds and whose sizes vary continuously between the
  // traces of a served mix) up to twice as large
  const auto hi = free_blocks.upper_bound(bytes >= LARGE_BLOCK_BYTES ? 2 * bytes : bytes + bytes / 4);
  auto c = pick(lo, hi, s);
  if (c != hi) return take(c, s);
  // what fits belongs to busy streams: waiting for one could hold this stream
  // behind another context's long kernels, so grow the pool while the limit
  // allows; at the limit take an oversized block that needs no such wait, and
  // only then wait
  if (live_bytes + cached_bytes + bytes > limit) {
    c = pick(lo, free_blocks.end(), s);
    if (c != free_blocks.end()) return take(c, s);
    if (lo != free_blocks.end()) return take(lo, s);
    while (!free_blocks.empty() && live_bytes + cached_bytes + bytes > limit) free_one(std::prev(free_blocks.end()));
  }
  void* p = nullptr;
  const auto t0 = std::chrono::steady_clock::now();
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    while (!free_blocks.empty()) free_one(std::prev(free_blocks.end()));
    e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      throw ApiError(E_NOMEM, "device allocation failed");
    }
  }
  n_malloc++;
  malloc_ms += ms_since(t0);
  live[p] = bytes;
  live_bytes += bytes;
  return p;
}

void SharedPool::release(void* p, hipStream_t s) {
  if (!p) return;
  std::lock_guard<std::mutex> g(mu);
  auto it = live.find(p);
  if (it == live.end()) return;
  hipEvent_t ev = nullptr;
  if (it->second >= EXACT_EVENT_BYTES) {
    if (!spare_events.empty()) {
      ev = spare_events.back();
      spare_events.pop_back();
    } else if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
      (void)hipGetLastError();
      ev = nullptr;
    }
    if (ev && hipEventRecord(ev, s) != hipSuccess) {
      (void)hipGetLastError();
      spare_events.push_back(ev);
      ev = nullptr;  // the block falls back to the stream-based rules
    }
  }
  free_blocks.emplace(it->second, Block{p, it->second, s, ev});
  cached_bytes += it->second;
  live_bytes -= it->second;
  live.erase(it);
}

void SharedPool::forget_stream(hipStream_t s) {
  std::lock_guard<std::mutex> g(mu);
  for (auto& kv : free_blocks)
    if (kv.second.s == s) {
      kv.second.s = nullptr;
      if (kv.second.ev) {
        spare_events.push_back(kv.second.ev);  // completed: the stream was synchronized
        kv.second.ev = nullptr;
      }
    }
}

void SharedPool::trim() {
  std::lock_guard<std::mutex> g(mu);
  while (!free_blocks.empty()) free_one(std::prev(free_blocks.end()));
}

SharedPool::~SharedPool() {
  trim();
  for (auto& kv : live) (void)hipFree(kv.first);
  for (hipEvent_t e : spare_events) (void)hipEventDestroy(e);
}

uint64_t* Ctx::host_stage(size_t words) {
  if (words <= pinned_words) return pinned;
  if (pinned) (void)hipHostFree(pinned);
  pinned = nullptr;
  pinned_words = 0;
  size_t want = words + words / 2;
  if (hipHostMalloc((void**)&pinned, want * 8, hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    pinned = nullptr;
    throw ApiError(E_NOMEM, "pinned host allocation failed");
  }
  pinned_words = want;
  return pinned;
}

constexpr size_t UPLOAD_RING_BYTES = size_t(8) << 20;

void Ctx::upload_async(void* dst, const void* src, size_t bytes, hipStream_t s) {
  if (bytes == 0) return;
  const size_t need = (bytes + 255) & ~size_t(255);
  if (need > UPLOAD_RING_BYTES) {  // rare: a large image goes straight from the caller's memory
    P3G_CHECK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
    P3G_CHECK_HIP(hipStreamSynchronize(s));
    return;
  }
  std::lock_guard<std::mutex> g(up_mu);
  if (!up_ring) {
    if (hipHostMalloc((void**)&up_ring, UPLOAD_RING_BYTES, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      up_ring = nullptr;
      throw ApiError(E_NOMEM, "pinned host allocation failed");
    }
    up_cap = UPLOAD_RING_BYTES;
    up_head = 0;
  }
  // retire finished segments from the front without waiting
  while (!up_inflight.empty() && hipEventQuery(up_inflight.front().ev) == hipSuccess) {
    up_spare.push_back(up_inflight.front().ev);
    up_inflight.pop_front();
  }
  const size_t pos = up_head + need <= up_cap ? up_head : 0;
  auto overlaps = [&] {
    for (const UpSeg& f : up_inflight)
      if (f.start < pos + need && pos < f.end) return true;
    return false;
  };
  // oldest first: every earlier copy completes no later than the ones after it
  while (overlaps()) {
    UpSeg f = up_inflight.front();
    up_inflight.pop_front();
    const hipError_t e = hipEventSynchronize(f.ev);
    up_spare.push_back(f.ev);
    P3G_CHECK_HIP(e);
  }
  memcpy(up_ring + pos, src, bytes);
  hipEvent_t ev;
  if (!up_spare.empty()) {
    ev = up_spare.back();
    up_spare.pop_back();
  } else {
    P3G_CHECK_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  }
  const hipError_t e1 = hipMemcpyAsync(dst, up_ring + pos, bytes, hipMemcpyHostToDevice, s);
  const hipError_t e2 = e1 == hipSuccess ? hipEventRecord(ev, s) : e1;
  if (e2 != hipSuccess) {
    // the copy may still be queued: keep the region fenced by the stream itself
    (void)hipStreamSynchronize(s);
    up_spare.push_back(ev);
    P3G_CHECK_HIP(e2);
  }
  up_inflight.push_back(UpSeg{pos, pos + need, ev});
  up_head = pos + need;
}

void* Ctx::upload(const void* src, size_t bytes, hipStream_t s) {
  void* d = alloc(bytes ? bytes : 8);
  try {
    upload_async(d, src, bytes, s);
  } catch (...) {
    release(d);
    throw;
  }
  return d;
}

void Ctx::read_back(void* dst, const void* dsrc, size_t bytes, hipStream_t s) {
  if (bytes == 0) {
    P3G_CHECK_HIP(hipStreamSynchronize(s));
    return;
  }
  if (bytes > rb_bytes) {
    if (rb_pinned) (void)hipHostFree(rb_pinned);
    rb_pinned = nullptr;
    rb_bytes = 0;
    const size_t want = std::max<size_t>(bytes, 64 << 10);
    if (hipHostMalloc((void**)&rb_pinned, want, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      rb_pinned = nullptr;
      throw ApiError(E_NOMEM, "pinned host allocation failed");
    }
    rb_bytes = want;
  }
  P3G_CHECK_HIP(hipMemcpyAsync(rb_pinned, dsrc, bytes, hipMemcpyDeviceToHost, s));
  P3G_CHECK_HIP(hipStreamSynchronize(s));
  memcpy(dst, rb_pinned, bytes);
}

void Ctx::free_upload_ring() {
  std::lock_guard<std::mutex> g(up_mu);
  for (const UpSeg& f : up_inflight) {
    (void)hipEventSynchronize(f.ev);
    (void)hipEventDestroy(f.ev);
  }
  up_inflight.clear();
  for (hipEvent_t ev : up_spare) (void)hipEventDestroy(ev);
  up_spare.clear();
  if (up_ring) (void)hipHostFree(up_ring);
  up_ring = nullptr;
  up_cap = up_head = 0;
  if (rb_pinned) (void)hipHostFree(rb_pinned);
  rb_pinned = nullptr;
  rb_bytes = 0;
}

void* Ctx::alloc(size_t bytes) {
  if (fail_alloc_after >= 0 && fail_alloc_after-- == 0)
    throw ApiError(E_NOMEM, "device allocation failed (injected for a failure-handling test)");
  if (bytes == 0) bytes = 8;
  bytes = (bytes + 255) & ~size_t(255);
  if (pool) {
    void* p = pool->alloc(bytes, stream);
    std::lock_guard<std::mutex> g(mu);
    live_blocks[p] = bytes;
    live_bytes += bytes;
    if (live_bytes > peak_live_bytes) peak_live_bytes = live_bytes;
    return p;
  }
  std::lock_guard<std::mutex> g(mu);
  auto it = free_blocks.lower_bound(bytes);
  // reuse a cached block when it is not more than 25% larger than requested
  if (it != free_blocks.end() && it->first <= bytes + bytes / 4) {
    void* p = it->second;
    size_t sz = it->first;
    free_blocks.erase(it);
    cached_bytes -= sz;
    live_blocks[p] = sz;
    live_bytes += sz;
    if (live_bytes > peak_live_bytes) peak_live_bytes = live_bytes;
    return p;
  }
  void* p = nullptr;
  const auto t0 = std::chrono::steady_clock::now();
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) {
    // drop the cache and retry once
    (void)hipGetLastError();
    for (auto& kv : free_blocks) (void)hipFree(kv.second);
    free_blocks.clear();
    cached_bytes = 0;
    e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      throw ApiError(E_NOMEM, "device allocation failed");
    }
  }
  n_malloc++;
  malloc_ms += ms_since(t0);
  live_blocks[p] = bytes;
  live_bytes += bytes;
  if (live_bytes > peak_live_bytes) peak_live_bytes = live_bytes;
  return p;
}

void Ctx::release(void* p) {
  if (!p) return;
  std::lock_guard<std::mutex> g(mu);
  auto it = live_blocks.find(p);
  if (it == live_blocks.end()) return;
  if (pool) {
    live_bytes -= it->second;
    live_blocks.erase(it);
    pool->release(p, stream);
    return;
  }
  free_blocks.emplace(it->second, p);
  cached_bytes += it->second;
  live_bytes -= it->second;
  live_blocks.erase(it);
}

void Ctx::trim() {
  std::lock_guard<std::mutex> g(mu);
  const auto t0 = std::chrono::steady_clock::now();
  n_free += free_blocks.size();
  for (auto& kv : free_blocks) (void)hipFree(kv.second);
  free_ms += ms_since(t0);
  free_blocks.clear();
  cached_bytes = 0;
}

const uint64_t* Ctx::pow_table(uint64_t base, uint32_t bits, uint64_t lo_mul) {
  auto key = std::make_tuple(base, bits, lo_mul);
  {
    std::lock_guard<std::mutex> g(mu);
    auto it = pow_tables.find(key);
    if (it != pow_tables.end()) return it->second;
  }
  uint64_t hl = pow_table_hi_len(bits);
  std::vector<uint64_t> host(POW_LO + hl);
  uint64_t v = lo_mul;
  for (uint32_t i = 0; i < POW_LO; i++) {
    host[i] = v;
    v = gl_mul(v, base);
  }
  uint64_t step = gl_pow(base, POW_LO), h = 1;
  for (uint64_t i = 0; i < hl; i++) {
    host[POW_LO + i] = h;
    h = gl_mul(h, step);
  }
  uint64_t* d = nullptr;
  P3G_CHECK_HIP(hipMalloc(&d, host.size() * 8));
  P3G_CHECK_HIP(hipMemcpy(d, host.data(), host.size() * 8, hipMemcpyHostToDevice));
  std::lock_guard<std::mutex> g(mu);
  pow_tables[key] = d;
  return d;
}

Tree::~Tree() {
  if (ctx) {
    for (auto& m : mats)
      if (m.owned) ctx->release(m.ptr);
    if (salt.owned) ctx->release(salt.ptr);
    if (owns_digests) ctx->release(digests);
  }
}

void check_pcs_params(const PcsParams& p) {
  if (p.log_arity < 1 || p.log_arity > 3) throw ApiError(E_INVAL, "invalid folding arity: log_arity must be 1, 2 or 3");
  if (p.log_blowup == 0) throw ApiError(E_INVAL, "log_blowup must be > 0");
  if (uint64_t(p.log_final_degree) + p.log_blowup > 64) throw ApiError(E_INVAL, "log_final_degree + log_blowup exceeds 64");
  if (p.num_queries == 0) throw ApiError(E_INVAL, "num_queries must be > 0");
  if (p.folding_pow_bits >= MAX_POW_BITS || p.deep_pow_bits >= MAX_POW_BITS || p.query_pow_bits >= MAX_POW_BITS)
    throw ApiError(E_INVAL, "proof-of-work bits must be below 32");
}

size_t fri_num_rounds(const PcsParams& p, unsigned log_domain) {
  unsigned lmf = p.log_final_degree + p.log_blowup;
  unsigned diff = log_domain > lmf ? log_domain - lmf : 0;
  return (diff + p.log_arity - 1) / p.log_arity;
}

size_t fri_final_degree(const PcsParams& p, unsigned log_domain) {
  size_t rounds = fri_num_rounds(p, log_domain);
  // the reference subtracts in usize here (fri/mod.rs:84-92) and fails when the
  // last round folds past the domain (an arity too large for a tiny domain)
  if (rounds * p.log_arity > log_domain)
    throw ApiError(E_INVAL, "FRI folding overshoots the LDE domain: log_arity too large for this height");
  unsigned lfs = log_domain - (unsigned)(rounds * p.log_arity);
  unsigned e = lfs > p.log_blowup ? lfs - p.log_blowup : 0;
  return (size_t)1 << e;
}

}  // namespace p3g
