// Device context: stream, Poseidon2 constants, cached power tables and a
// size-bucketed cache of device allocations (the prover allocates the same
// shapes proof after proof, so blocks are recycled instead of hipMalloc'd).
#include "internal.hpp"

namespace p3g {

void* Ctx::alloc(size_t bytes) {
  if (bytes == 0) bytes = 8;
  bytes = (bytes + 255) & ~size_t(255);
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
  free_blocks.emplace(it->second, p);
  cached_bytes += it->second;
  live_bytes -= it->second;
  live_blocks.erase(it);
}

void Ctx::trim() {
  std::lock_guard<std::mutex> g(mu);
  for (auto& kv : free_blocks) (void)hipFree(kv.second);
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
    if (owns_digests) ctx->release(digests);
  }
}

size_t fri_num_rounds(const PcsParams& p, unsigned log_domain) {
  unsigned lmf = p.log_final_degree + p.log_blowup;
  unsigned diff = log_domain > lmf ? log_domain - lmf : 0;
  return (diff + p.log_arity - 1) / p.log_arity;
}

size_t fri_final_degree(const PcsParams& p, unsigned log_domain) {
  size_t rounds = fri_num_rounds(p, log_domain);
  unsigned lfs = log_domain - (unsigned)(rounds * p.log_arity);
  unsigned e = lfs > p.log_blowup ? lfs - p.log_blowup : 0;
  return (size_t)1 << e;
}

}  // namespace p3g
