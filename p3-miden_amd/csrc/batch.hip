// Batched lifted commit of many traces of uneven heights and widths.
//
// Same result as commit_traces (p3-miden-lifted-stark/src/prover/commit.rs:175-228):
// per trace a coset LDE on the nested shift GENERATOR^(2^(log n_max - log n_j)),
// bit-reversed rows, then one aligned LMCS tree over all LDEs.  The batch is
// planned once per shape:
//   - all device buffers (inputs, LDEs, sponge states, digest layers, launch
//     descriptors, coset scale tables) are allocated and the descriptors
//     uploaded when the plan is made;
//   - every NTT pass index is ONE grouped launch over all traces (block ranges
//     map to traces through a prefix table), so a 12-trace batch issues as
//     many NTT launches as its tallest trace instead of twelve times that;
//   - leaf hashing runs per height group with lifted state inheritance, the
//     compression layers follow;
//   - the whole launch sequence has no host interaction, so it is captured
//     once into a hipGraph and replayed for every batch of that shape;
//   - the NTT stage grouping (8-stage register-blocked passes vs 6- or 4-stage
//     LDS passes) and the leaf-hash workgroup size are chosen by timing the
//     candidates on the plan's own shape (cached per shape for the process).
#include <algorithm>
#include <chrono>
#include <cstring>
#include <map>
#include <mutex>

#include "internal.hpp"
#include "batch.hpp"

namespace p3g {

// ---- kernels and helpers defined in lde.hip / lmcs.hip
extern void launch_ntt_grouped(uint32_t mode, bool reg, uint32_t K, const void* descs, const uint64_t* starts,
                               uint32_t nd, uint64_t blocks, hipStream_t s);
extern void launch_ntt_single(uint32_t mode, bool reg, const void* args, uint64_t blocks, hipStream_t s);
extern std::vector<uint32_t> ntt_pass_split(uint32_t logn, uint32_t kmax);
extern bool ntt_reg_ok(uint32_t K, uint32_t logC);
extern void launch_leaf_raw(uint32_t hash, int threads, const void* mats, uint32_t nmats, const uint64_t* prev,
                            uint32_t prev_shift, uint64_t* out_states, uint64_t* out_digests, uint64_t h,
                            const P2Consts* p2, hipStream_t s);
extern void launch_compress_layers(uint32_t hash, uint64_t* digests, const std::vector<uint64_t>& layer_off,
                                   uint64_t height, const P2Consts* p2, hipStream_t s);
extern size_t ntt_args_size();
extern void make_ntt_args(void* out, const uint64_t* src, uint64_t* dst, const uint64_t* tw, const uint64_t* scale,
                          uint64_t scale_stride, uint64_t n, uint32_t logn, uint32_t s0, uint32_t K, uint32_t w,
                          uint32_t logC, uint32_t ncol, uint32_t logB, uint32_t mode);
extern size_t leaf_mat_size();
extern void make_leaf_mat(void* out, const uint64_t* ptr, uint64_t width);

namespace {
constexpr uint32_t NTT_INV_FIRST = 0, NTT_INV = 1, NTT_FWD_FIRST = 2, NTT_FWD = 3;

struct PassDesc {
  uint32_t trace, mode, s0, K;
  bool reg;  // register-blocked kernel (K stages on 2^K x 16 tiles)
  uint64_t blocks;
};

struct TuneChoice {
  uint32_t kmax;
  int leaf_threads;
  bool grouped;
};
std::mutex g_tune_mu;
std::map<std::string, TuneChoice> g_tune_cache;
}  // namespace

CommitPlan::~CommitPlan() {
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  if (ctx) {
    for (auto& t : traces) ctx->release(t.ptr);
    for (auto& m : ldes) ctx->release(m.ptr);
    ctx->release(digests);
    ctx->release(states[0]);
    ctx->release(states[1]);
    ctx->release(blob);
    for (auto* sc : scale_owned) ctx->release(sc);
  }
}

std::string CommitPlan::shape_key() const {
  std::string k = std::to_string(hash) + "/" + std::to_string(log_blowup);
  for (auto& t : traces) k += ":" + std::to_string(t.height) + "x" + std::to_string(t.width);
  return k;
}

void CommitPlan::build_launches(uint32_t kmax, int leaf_threads, bool grouped) {
  launches.clear();
  this->kmax = kmax;
  this->leaf_threads = leaf_threads;
  this->grouped = grouped;
  const uint32_t B = 1u << log_blowup;
  // ---- NTT passes per trace
  std::vector<std::vector<PassDesc>> inv(traces.size()), fwd(traces.size());
  size_t max_inv = 0, max_fwd = 0;
  for (size_t j = 0; j < traces.size(); j++) {
    const uint64_t n = traces[j].height, w = traces[j].width;
    const unsigned logn = log2_exact(n);
    uint32_t logC = 0;
    while ((2u << logC) <= 16 && (2ull << logC) <= w) logC++;
    const uint64_t ncol = (w + (1ull << logC) - 1) >> logC;
    auto add = [&](std::vector<PassDesc>& v, uint32_t mode, uint32_t s0, uint32_t K, uint64_t nblk) {
      v.push_back(PassDesc{(uint32_t)j, mode, s0, K, ntt_reg_ok(K, logC), (n >> K) * ncol * nblk});
    };
    const std::vector<uint32_t> split = ntt_pass_split(logn, kmax);
    uint32_t s0 = 0;
    for (size_t i = 0; i < split.size(); i++) {
      add(inv[j], i == 0 ? NTT_INV_FIRST : NTT_INV, s0, split[i], 1);
      add(fwd[j], i == 0 ? NTT_FWD_FIRST : NTT_FWD, s0, split[i], i == 0 ? 1 : B);
      s0 += split[i];
    }
    max_inv = std::max(max_inv, inv[j].size());
    max_fwd = std::max(max_fwd, fwd[j].size());
  }
  // ---- descriptor blob layout: [ntt descriptors | starts | leaf mats | scale tables (already placed)]
  std::vector<uint8_t> host;
  auto reserve_bytes = [&](size_t n) {
    size_t off = (host.size() + 255) & ~size_t(255);
    host.resize(off + n);
    return off;
  };
  struct Pending {
    Launch l;
    size_t desc_off, start_off, mats_off;
  };
  std::vector<Pending> pend;
  auto group = [&](std::vector<std::vector<PassDesc>>& per, size_t depth) {
    for (size_t i = 0; i < depth; i++)
      // variants: the generic LDS kernel (0) and register-blocked kernels by K (5..8)
      for (int variant : {0, 5, 6, 7, 8}) {
        std::vector<PassDesc> g;
        for (auto& v : per)
          if (i < v.size() && (v[i].reg ? (int)v[i].K : 0) == variant) g.push_back(v[i]);
        if (g.empty()) continue;
        if (!grouped) {
          for (const PassDesc& d : g) {
            Pending ps{};
            ps.l.kind = Launch::NTT_SINGLE;
            ps.l.mode = d.mode;
            ps.l.reg = d.reg;
            ps.l.K = d.K;
            ps.l.nd = 1;
            ps.l.blocks = d.blocks;
            ps.l.args.resize(ntt_args_size());
            const DevMat& tr = traces[d.trace];
            uint32_t logC = 0;
            while ((2u << logC) <= 16 && (2ull << logC) <= tr.width) logC++;
            const uint32_t ncol = (uint32_t)((tr.width + (1ull << logC) - 1) >> logC);
            const bool inverse = d.mode <= NTT_INV;
            make_ntt_args(ps.l.args.data(), tr.ptr, ldes[d.trace].ptr, inverse ? tw_inv[d.trace] : tw_fwd[d.trace],
                          scale_dev[d.trace], scale_stride[d.trace], tr.height, log2_exact(tr.height), d.s0, d.K,
                          (uint32_t)tr.width, logC, ncol, log_blowup, d.mode);
            pend.push_back(std::move(ps));
          }
          continue;
        }
        Pending pd{};
        pd.l.kind = Launch::NTT;
        pd.l.mode = g[0].mode;
        pd.l.reg = variant != 0;
        pd.l.K = (uint32_t)variant;
        pd.l.nd = (uint32_t)g.size();
        pd.desc_off = reserve_bytes(ntt_args_size() * g.size());
        pd.start_off = reserve_bytes(8 * (g.size() + 1));
        uint64_t acc = 0;
        for (size_t k = 0; k < g.size(); k++) {
          const PassDesc& d = g[k];
          const DevMat& tr = traces[d.trace];
          const unsigned logn = log2_exact(tr.height);
          uint32_t logC = 0;
          while ((2u << logC) <= 16 && (2ull << logC) <= tr.width) logC++;
          const uint32_t ncol = (uint32_t)((tr.width + (1ull << logC) - 1) >> logC);
          const bool inverse = d.mode <= NTT_INV;
          make_ntt_args(host.data() + pd.desc_off + k * ntt_args_size(), tr.ptr, ldes[d.trace].ptr,
                        inverse ? tw_inv[d.trace] : tw_fwd[d.trace], scale_dev[d.trace], scale_stride[d.trace],
                        tr.height, logn, d.s0, d.K, (uint32_t)tr.width, logC, ncol, log_blowup, d.mode);
          memcpy(host.data() + pd.start_off + 8 * k, &acc, 8);
          acc += d.blocks;
        }
        memcpy(host.data() + pd.start_off + 8 * g.size(), &acc, 8);
        pd.l.blocks = acc;
        pend.push_back(pd);
      }
  };
  group(inv, max_inv);
  group(fwd, max_fwd);
  // ---- leaf hashing by height group
  {
    size_t i = 0;
    int cur = -1;
    uint64_t cur_h = 0;
    while (i < ldes.size()) {
      size_t j = i;
      while (j < ldes.size() && ldes[j].height == ldes[i].height) j++;
      Pending pd{};
      pd.l.kind = Launch::LEAF;
      pd.l.nd = (uint32_t)(j - i);
      pd.mats_off = reserve_bytes(leaf_mat_size() * (j - i));
      for (size_t k = i; k < j; k++)
        make_leaf_mat(host.data() + pd.mats_off + (k - i) * leaf_mat_size(), ldes[k].ptr, ldes[k].width);
      const uint64_t h = ldes[i].height;
      const bool final = (j == ldes.size());
      pd.l.h = h;
      pd.l.prev = cur >= 0 ? cur : -1;
      pd.l.prev_shift = cur >= 0 ? log2_exact(h / cur_h) : 0;
      const int nxt = cur >= 0 ? cur ^ 1 : 0;
      pd.l.out_state = final ? -1 : nxt;
      pd.l.final = final;
      pend.push_back(pd);
      cur = nxt;
      cur_h = h;
      i = j;
    }
  }
  // ---- upload the blob (descriptors reference plan-owned buffers only)
  ctx->release(blob);
  blob = (uint8_t*)ctx->alloc(host.size() ? host.size() : 8);
  P3G_CHECK_HIP(hipMemcpyAsync(blob, host.data(), host.size(), hipMemcpyHostToDevice, ctx->stream));
  P3G_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  for (auto& pd : pend) {
    Launch l = std::move(pd.l);
    if (l.kind == Launch::NTT_SINGLE) {
      // descriptor travels by value
    } else if (l.kind == Launch::NTT) {
      l.descs = blob + pd.desc_off;
      l.starts = reinterpret_cast<const uint64_t*>(blob + pd.start_off);
    } else {
      l.descs = blob + pd.mats_off;
    }
    launches.push_back(l);
  }
  if (exec) {
    (void)hipGraphExecDestroy(exec);
    exec = nullptr;
  }
  if (graph) {
    (void)hipGraphDestroy(graph);
    graph = nullptr;
  }
}

void CommitPlan::enqueue(hipStream_t s) {
  for (auto& l : launches) {
    if (l.kind == Launch::NTT) {
      launch_ntt_grouped(l.mode, l.reg, l.K, l.descs, l.starts, l.nd, l.blocks, s);
    } else if (l.kind == Launch::NTT_SINGLE) {
      launch_ntt_single(l.mode, l.reg, l.args.data(), l.blocks, s);
    } else {
      launch_leaf_raw(hash, leaf_threads, l.descs, l.nd, l.prev >= 0 ? states[l.prev] : nullptr, l.prev_shift,
                      l.out_state >= 0 ? states[l.out_state] : nullptr, l.final ? digests : nullptr, l.h,
                      ctx->dev_p2, s);
    }
  }
  launch_compress_layers(hash, digests, layer_off, height, ctx->dev_p2, s);
}

void CommitPlan::capture(hipStream_t s) {
  if (exec) return;
  P3G_CHECK_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  try {
    enqueue(s);
  } catch (...) {
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(s, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  P3G_CHECK_HIP(hipStreamEndCapture(s, &graph));
  P3G_CHECK_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
}

double CommitPlan::time_once(hipStream_t s) {
  hipEvent_t a, b;
  P3G_CHECK_HIP(hipEventCreate(&a));
  P3G_CHECK_HIP(hipEventCreate(&b));
  P3G_CHECK_HIP(hipEventRecord(a, s));
  enqueue(s);
  P3G_CHECK_HIP(hipEventRecord(b, s));
  P3G_CHECK_HIP(hipEventSynchronize(b));
  float ms = 0;
  P3G_CHECK_HIP(hipEventElapsedTime(&ms, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return ms;
}

void CommitPlan::autotune(hipStream_t s) {
  const std::string key = shape_key();
  {
    std::lock_guard<std::mutex> g(g_tune_mu);
    auto it = g_tune_cache.find(key);
    if (it != g_tune_cache.end()) {
      build_launches(it->second.kmax, it->second.leaf_threads, it->second.grouped);
      return;
    }
  }
  TuneChoice best{8, 128, true};
  double best_ms = 1e30;
  auto trial = [&](TuneChoice c) {
    build_launches(c.kmax, c.leaf_threads, c.grouped);
    time_once(s);  // warm
    const double ms = time_once(s);
    tune_log += "kmax=" + std::to_string(c.kmax) + " leaf=" + std::to_string(c.leaf_threads) +
                (c.grouped ? " grouped" : " per-trace") + ": " + std::to_string(ms) + " ms\n";
    if (ms < best_ms) {
      best_ms = ms;
      best = c;
    }
  };
  // NTT launch structure and stage grouping first, then the leaf workgroup size
  for (bool grouped : {true, false})
    for (uint32_t k : {8u, 6u}) trial(TuneChoice{k, 128, grouped});
  const TuneChoice ntt_best = best;
  for (int t : {64, 256}) trial(TuneChoice{ntt_best.kmax, t, ntt_best.grouped});
  build_launches(best.kmax, best.leaf_threads, best.grouped);
  std::lock_guard<std::mutex> g(g_tune_mu);
  g_tune_cache[key] = best;
}

CommitPlan* make_commit_plan(Ctx& ctx, uint32_t hash, const std::vector<uint64_t>& heights,
                             const std::vector<uint64_t>& widths, unsigned log_blowup, bool use_graph,
                             bool autotune, bool grouped) {
  if (heights.empty() || heights.size() != widths.size()) throw ApiError(E_EMPTY, "at least one trace required");
  for (size_t i = 0; i < heights.size(); i++) {
    if (heights[i] == 0 || (heights[i] & (heights[i] - 1))) throw ApiError(E_NOT_POW2, "trace height must be a power of two");
    if (i && heights[i] < heights[i - 1]) throw ApiError(E_UNSORTED, "traces must be sorted by height");
    if (widths[i] == 0) throw ApiError(E_INVAL, "zero-width trace");
  }
  if (log2_exact(heights.back()) + log_blowup > 32) throw ApiError(E_RANGE, "LDE height exceeds the field two-adicity");
  CommitPlan* p = new CommitPlan();
  try {
    p->ctx = &ctx;
    p->hash = hash;
    p->log_blowup = log_blowup;
    p->use_graph = use_graph;
    const unsigned lmax = log2_exact(heights.back());
    const uint32_t B = 1u << log_blowup;
    for (size_t i = 0; i < heights.size(); i++) {
      DevMat t;
      t.height = heights[i];
      t.width = widths[i];
      t.ptr = (uint64_t*)ctx.alloc(t.height * t.width * 8);
      P3G_CHECK_HIP(hipMemsetAsync(t.ptr, 0, t.height * t.width * 8, ctx.stream));
      p->traces.push_back(t);
      DevMat m;
      m.height = t.height << log_blowup;
      m.width = t.width;
      m.ptr = (uint64_t*)ctx.alloc(m.height * m.width * 8);
      p->ldes.push_back(m);
      // tables: twiddles (shared per log n) and the B coset scale tables of this trace
      const unsigned logn = log2_exact(t.height);
      const uint32_t twbits = logn ? logn - 1 : 0;
      p->tw_inv.push_back(ctx.pow_table(gl_inv(gl_two_adic_generator(logn)), twbits));
      p->tw_fwd.push_back(ctx.pow_table(gl_two_adic_generator(logn), twbits));
      uint64_t shift = GENERATOR;
      for (unsigned k = logn; k < lmax; k++) shift = gl_mul(shift, shift);
      const uint64_t stride = POW_LO + pow_table_hi_len(logn);
      std::vector<uint64_t> host(stride * B);
      const uint64_t ninv = gl_inv_pow2(logn);
      const uint64_t om = gl_two_adic_generator(logn + log_blowup);
      uint64_t ct = shift;
      for (uint32_t c = 0; c < B; c++) {
        uint64_t* tab = host.data() + c * stride;
        uint64_t v = ninv;
        for (uint32_t k = 0; k < POW_LO; k++) {
          tab[k] = v;
          v = gl_mul(v, ct);
        }
        uint64_t step = gl_pow(ct, POW_LO), h = 1;
        for (uint64_t k = 0; k < pow_table_hi_len(logn); k++) {
          tab[POW_LO + k] = h;
          h = gl_mul(h, step);
        }
        ct = gl_mul(ct, om);
      }
      uint64_t* sc = (uint64_t*)ctx.alloc(host.size() * 8);
      P3G_CHECK_HIP(hipMemcpyAsync(sc, host.data(), host.size() * 8, hipMemcpyHostToDevice, ctx.stream));
      P3G_CHECK_HIP(hipStreamSynchronize(ctx.stream));
      p->scale_dev.push_back(sc);
      p->scale_stride.push_back(stride);
      p->scale_owned.push_back(sc);
    }
    p->height = p->ldes.back().height;
    p->depth = log2_exact(p->height);
    p->layer_off.resize(p->depth + 1);
    uint64_t off = 0;
    for (unsigned l = 0; l <= p->depth; l++) {
      p->layer_off[l] = off;
      off += p->height >> l;
    }
    p->digests = (uint64_t*)ctx.alloc(off * 32);
    const uint64_t W = hash_width(hash);
    if (heights.front() != heights.back()) {
      p->states[0] = (uint64_t*)ctx.alloc(p->height * W * 8);
      p->states[1] = (uint64_t*)ctx.alloc(p->height * W * 8);
    }
    if (autotune) p->autotune(ctx.stream);
    else p->build_launches(8, 128, grouped);
  } catch (...) {
    delete p;
    throw;
  }
  return p;
}

void CommitPlan::run(hipStream_t s) {
  if (use_graph) {
    capture(s);
    P3G_CHECK_HIP(hipGraphLaunch(exec, s));
  } else {
    enqueue(s);
  }
  ctx->read_back(root, digests + layer_off[depth] * 4, 32, s);
}

Tree* CommitPlan::tree_view() {
  Tree* t = new Tree();
  t->ctx = ctx;
  t->hash = hash;
  t->alignment = hash_alignment(hash);
  for (auto& m : ldes) {
    DevMat v = m;
    v.owned = false;
    t->mats.push_back(v);
  }
  t->height = height;
  t->depth = depth;
  t->digests = digests;
  t->owns_digests = false;
  t->layer_off = layer_off;
  memcpy(t->root, root, 32);
  return t;
}

}  // namespace p3g

namespace p3g {

// Synthetic canonical field elements from a counter-based generator:
// v = splitmix64(seed + i * golden), redrawn from v itself when v >= p.
__global__ void fill_random_kernel(uint64_t seed, uint64_t n, uint64_t* out) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    uint64_t z = seed + (i + 1) * 0x9e3779b97f4a7c15ull;
    for (;;) {
      z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
      z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
      z = z ^ (z >> 31);
      if (z < P) break;
      z += 0x9e3779b97f4a7c15ull;
    }
    out[i] = z;
  }
}

void fill_random_device(Ctx& ctx, uint64_t seed, uint64_t n, uint64_t* out, hipStream_t s) {
  (void)ctx;
  if (n == 0) return;
  uint64_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(fill_random_kernel, dim3((uint32_t)blocks), dim3(256), 0, s, seed, n, out);
  P3G_CHECK_HIP(hipGetLastError());
}

}  // namespace p3g
