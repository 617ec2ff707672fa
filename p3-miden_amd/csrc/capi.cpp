// extern "C" entry points declared in include/p3gpu.h.  Every call catches
// library and HIP failures and turns them into status codes; nothing throws
// across the boundary.
#include <array>
#include <cstddef>
#include <cstring>
#include <string>

#include "../../include/p3gpu.h"
#include "handles.hpp"

using namespace p3g;

namespace p3g {
void fill_random_device(Ctx& ctx, uint64_t seed, uint64_t n, uint64_t* out, hipStream_t s);
uint64_t calib_valu_insts_per_wave(unsigned kind, uint32_t iters);
double calib_valu(Ctx& ctx, unsigned kind, uint32_t iters, uint32_t blocks);
void calib_copy(Ctx& ctx, uint64_t bytes, uint32_t reps, double* kernel_ms, double* memcpy_ms);
void field_check(Ctx& ctx, uint32_t op, uint32_t e, const uint64_t* a, const uint64_t* b, const uint64_t* c,
                 uint64_t n, uint64_t* out, uint64_t* wave_flags);
}

static thread_local std::string g_last_error;

// the caller's parameters, checked once for every open and prove path
static PcsParams pcs_params(const p3g_pcs_params* q) {
  const PcsParams p{q->log_blowup,   q->log_folding_arity, q->log_final_degree, q->folding_pow_bits,
                    q->deep_pow_bits, q->num_queries,       q->query_pow_bits};
  check_pcs_params(p);
  return p;
}

template <class F>
static int guard(F&& f) {
  try {
    f();
    return P3G_OK;
  } catch (const ApiError& e) {
    g_last_error = e.msg;
    return e.code;
  } catch (const DeviceError& e) {
    g_last_error = std::string(e.what) + ": " + hipGetErrorString(e.err);
    return P3G_E_DEVICE;
  } catch (const std::bad_alloc&) {
    g_last_error = "host allocation failed";
    return P3G_E_NOMEM;
  } catch (...) {
    g_last_error = "unknown failure";
    return P3G_E_INVAL;
  }
}

extern "C" {

const char* p3g_status_string(int status) {
  switch (status) {
    case P3G_OK: return "ok";
    case P3G_E_INVAL: return "invalid argument";
    case P3G_E_NOT_POW2: return "height is not a power of two";
    case P3G_E_UNSORTED: return "matrices are not sorted by height";
    case P3G_E_NOMEM: return "out of device memory";
    case P3G_E_DEVICE: return "device error";
    case P3G_E_EMPTY: return "empty batch";
    case P3G_E_TRANSCRIPT: return "transcript error";
    case P3G_E_RANGE: return "index or size out of range";
    case P3G_E_CANCELED: return "request canceled";
    default: return "unknown status";
  }
}

const char* p3g_last_error(void) { return g_last_error.c_str(); }

static int ctx_create(int device, hipStream_t external, p3g_ctx** out) {
  if (!out) return P3G_E_INVAL;
  *out = nullptr;
  p3g_ctx* c = new p3g_ctx();
  int rc = guard([&] {
    c->c.device = device;
    // the NTT kernels multiply by 16th roots of unity as shifts
    if (gl_pow(2, OMEGA16_LOG2) != gl_two_adic_generator(4))
      throw ApiError(P3G_E_INVAL, "omega_16 is not 2^OMEGA16_LOG2");
    P3G_CHECK_HIP(hipSetDevice(device));
    if (external) {
      int sdev = -1;
      P3G_CHECK_HIP(hipStreamGetDevice(external, &sdev));
      if (sdev != device) throw ApiError(P3G_E_INVAL, "stream belongs to another device");
      c->c.stream = external;
      c->c.owns_stream = false;
    } else {
      P3G_CHECK_HIP(hipStreamCreateWithFlags(&c->c.stream, hipStreamNonBlocking));
    }
    lde_scratch_init(device);  // the no-alloc LDE's table array on this device
    p2_generate_constants(2025, &c->c.host_p2);
    P3G_CHECK_HIP(hipMalloc(&c->c.dev_p2, sizeof(P2Consts)));
    P3G_CHECK_HIP(hipMemcpy(c->c.dev_p2, &c->c.host_p2, sizeof(P2Consts), hipMemcpyHostToDevice));
  });
  if (rc != P3G_OK) {
    // undo what was set up before the failure: the constants' block and a stream this call created
    if (c->c.dev_p2) (void)hipFree(c->c.dev_p2);
    if (c->c.stream && c->c.owns_stream) (void)hipStreamDestroy(c->c.stream);
    delete c;
    return rc;
  }
  *out = c;
  return P3G_OK;
}

int p3g_ctx_create(int device, p3g_ctx** out) { return ctx_create(device, nullptr, out); }

int p3g_ctx_create_on_stream(int device, void* stream, p3g_ctx** out) {
  if (!stream) return P3G_E_INVAL;  // the null stream would mark released blocks as idle
  return ctx_create(device, static_cast<hipStream_t>(stream), out);
}

void p3g_ctx_destroy(p3g_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->c.device);
  if (ctx->c.stream) (void)hipStreamSynchronize(ctx->c.stream);
  ctx->c.trim();
  if (ctx->c.pool) {
    std::vector<void*> held;
    for (auto& kv : ctx->c.live_blocks) held.push_back(kv.first);
    for (void* p : held) ctx->c.release(p);  // back to the shared pool, which frees them
    // the stream is synchronized above and destroyed below: its blocks have no pending work
    ctx->c.pool->forget_stream(ctx->c.stream);
  } else {
    for (auto& kv : ctx->c.live_blocks) (void)hipFree(kv.first);
  }
  for (auto& kv : ctx->c.pow_tables) (void)hipFree(kv.second);
  if (ctx->c.dev_p2) (void)hipFree(ctx->c.dev_p2);
  if (ctx->c.pinned) (void)hipHostFree(ctx->c.pinned);
  ctx->c.free_upload_ring();
  if (ctx->c.stream && ctx->c.owns_stream) (void)hipStreamDestroy(ctx->c.stream);
  delete ctx;
}

int p3g_ctx_memory(p3g_ctx* ctx, uint64_t* live, uint64_t* cached, uint64_t* peak_live, int reset_peak) {
  if (!ctx) return P3G_E_INVAL;
  std::lock_guard<std::mutex> g(ctx->c.mu);
  if (live) *live = ctx->c.live_bytes;
  if (cached) *cached = ctx->c.cached_bytes;
  if (peak_live) *peak_live = ctx->c.peak_live_bytes;
  if (reset_peak) ctx->c.peak_live_bytes = ctx->c.live_bytes;
  return P3G_OK;
}

int p3g_ctx_synchronize(p3g_ctx* ctx) {
  if (!ctx) return P3G_E_INVAL;
  return guard([&] { P3G_CHECK_HIP(hipStreamSynchronize(ctx->c.stream)); });
}

int p3g_ctx_trim(p3g_ctx* ctx) {
  if (!ctx) return P3G_E_INVAL;
  return guard([&] {
    P3G_CHECK_HIP(hipStreamSynchronize(ctx->c.stream));
    ctx->c.trim();
  });
}

int p3g_profile_enable(p3g_ctx* ctx, int on) {
  if (!ctx) return P3G_E_INVAL;
  ctx->c.profiling = on != 0;
  return P3G_OK;
}

int p3g_profile_reset(p3g_ctx* ctx) {
  if (!ctx) return P3G_E_INVAL;
  return guard([&] {
    P3G_CHECK_HIP(hipStreamSynchronize(ctx->c.stream));
    std::lock_guard<std::mutex> g(ctx->c.mu);
    for (auto& kv : ctx->c.prof)
      for (auto& ev : kv.second) {
        (void)hipEventDestroy(ev.first);
        (void)hipEventDestroy(ev.second);
      }
    ctx->c.prof.clear();
  });
}

int p3g_profile_read(p3g_ctx* ctx, const char* name, double* total_ms, uint64_t* launches) {
  if (!ctx || !name || !total_ms || !launches) return P3G_E_INVAL;
  return guard([&] {
    P3G_CHECK_HIP(hipStreamSynchronize(ctx->c.stream));
    std::lock_guard<std::mutex> g(ctx->c.mu);
    double tot = 0;
    uint64_t n = 0;
    auto it = ctx->c.prof.find(name);
    if (it != ctx->c.prof.end())
      for (auto& ev : it->second) {
        float ms = 0;
        P3G_CHECK_HIP(hipEventElapsedTime(&ms, ev.first, ev.second));
        tot += ms;
        n++;
      }
    *total_ms = tot;
    *launches = n;
  });
}

int p3g_calibrate_valu(p3g_ctx* ctx, uint32_t kind, uint32_t iters, uint32_t blocks, double* ms,
                       uint64_t* insts_per_wave) {
  if (!ctx || !ms || !insts_per_wave) return P3G_E_INVAL;
  return guard([&] {
    P3G_CHECK_HIP(hipSetDevice(ctx->c.device));
    *ms = calib_valu(ctx->c, kind, iters, blocks);
    *insts_per_wave = calib_valu_insts_per_wave(kind, iters);
  });
}

int p3g_calibrate_copy(p3g_ctx* ctx, uint64_t bytes, uint32_t reps, double* kernel_ms, double* memcpy_ms) {
  if (!ctx || !kernel_ms || !memcpy_ms) return P3G_E_INVAL;
  return guard([&] {
    P3G_CHECK_HIP(hipSetDevice(ctx->c.device));
    calib_copy(ctx->c, bytes, reps, kernel_ms, memcpy_ms);
  });
}

int p3g_field_check(p3g_ctx* ctx, uint32_t op, uint32_t e, const uint64_t* a, const uint64_t* b, const uint64_t* c,
                    uint64_t n, uint64_t* out, uint64_t* wave_flags) {
  if (!ctx || !a || !b || !out) return P3G_E_INVAL;
  return guard([&] {
    P3G_CHECK_HIP(hipSetDevice(ctx->c.device));
    field_check(ctx->c, op, e, a, b, c, n, out, wave_flags);
  });
}

int p3g_alloc(p3g_ctx* ctx, uint64_t bytes, void** dptr) {
  if (!ctx || !dptr) return P3G_E_INVAL;
  return guard([&] { *dptr = ctx->c.alloc(bytes); });
}

int p3g_free(p3g_ctx* ctx, void* dptr) {
  if (!ctx) return P3G_E_INVAL;
  return guard([&] {
    P3G_CHECK_HIP(hipStreamSynchronize(ctx->c.stream));
    ctx->c.release(dptr);
  });
}

int p3g_upload(p3g_ctx* ctx, void* dst, const void* src, uint64_t bytes) {
  if (!ctx || (!dst && bytes) || (!src && bytes)) return P3G_E_INVAL;
  return guard([&] {
    P3G_CHECK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->c.stream));
    P3G_CHECK_HIP(hipStreamSynchronize(ctx->c.stream));
  });
}

int p3g_download(p3g_ctx* ctx, void* dst, const void* src, uint64_t bytes) {
  if (!ctx || (!dst && bytes) || (!src && bytes)) return P3G_E_INVAL;
  return guard([&] {
    P3G_CHECK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->c.stream));
    P3G_CHECK_HIP(hipStreamSynchronize(ctx->c.stream));
  });
}

int p3g_fill_random(uint64_t seed, uint64_t n, uint64_t* host_out) {
  if (!host_out && n) return P3G_E_INVAL;
  SmallRng rng(seed);
  for (uint64_t i = 0; i < n; i++) host_out[i] = rng.next_gl();
  return P3G_OK;
}

// ---- helpers for host/device matrix inputs
static uint64_t* to_device(Ctx& c, const uint64_t* src, uint64_t words) {
  uint64_t* d = (uint64_t*)c.alloc(words * sizeof(uint64_t));
  if (words) {
    P3G_CHECK_HIP(hipMemcpyAsync(d, src, words * 8, hipMemcpyHostToDevice, c.stream));
    P3G_CHECK_HIP(hipStreamSynchronize(c.stream));
  }
  return d;
}

int p3g_coset_lde_batch(p3g_ctx* ctx, const uint64_t* trace, uint64_t n, uint64_t w, uint32_t added_bits,
                        uint64_t shift, uint32_t flags, uint64_t* out) {
  if (!ctx || !trace || !out) return P3G_E_INVAL;
  return guard([&] {
    Ctx& c = ctx->c;
    P3G_CHECK_HIP(hipSetDevice(c.device));
    if (shift >= P) throw ApiError(P3G_E_INVAL, "shift is not a canonical field element");
    const bool dev = flags & P3G_DEVICE_PTRS;
    const uint64_t N = n << added_bits;
    const uint64_t* dtrace = dev ? trace : to_device(c, trace, n * w);
    uint64_t* dout = dev ? out : (uint64_t*)c.alloc(N * w * 8);
    uint64_t* tmp = nullptr;
    try {
      lde_coset_batch(c, dtrace, n, w, added_bits, shift, dout, c.stream);
      if (!(flags & P3G_BITREV_ROWS)) {
        // natural order: undo the bit reversal into a temporary, then swap back
        tmp = (uint64_t*)c.alloc(N * w * 8);
        bit_reverse_rows(c, dout, tmp, N, w, c.stream);
        P3G_CHECK_HIP(hipMemcpyAsync(dout, tmp, N * w * 8, hipMemcpyDeviceToDevice, c.stream));
      }
      if (!dev) P3G_CHECK_HIP(hipMemcpyAsync(out, dout, N * w * 8, hipMemcpyDeviceToHost, c.stream));
      P3G_CHECK_HIP(hipStreamSynchronize(c.stream));
    } catch (...) {
      if (!dev) {
        c.release((void*)dtrace);
        c.release(dout);
      }
      c.release(tmp);
      throw;
    }
    c.release(tmp);
    if (!dev) {
      c.release((void*)dtrace);
      c.release(dout);
    }
  });
}

int p3g_coset_lde_batch_noalloc(p3g_ctx* ctx, const uint64_t* d_trace, uint64_t n, uint64_t w, uint32_t added_bits,
                                uint64_t shift, uint64_t* d_out) {
  if (!ctx || !d_trace || !d_out) return P3G_E_INVAL;
  return guard([&] {
    Ctx& c = ctx->c;
    P3G_CHECK_HIP(hipSetDevice(c.device));
    if (shift >= P) throw ApiError(P3G_E_INVAL, "shift is not a canonical field element");
    lde_coset_batch_noalloc(c, d_trace, n, w, added_bits, shift, d_out, c.stream);
  });
}

static_assert(sizeof(p3g_air_op) == sizeof(AirOpIn) && offsetof(p3g_air_op, imm) == offsetof(AirOpIn, imm),
              "p3g_air_op layout");

int p3g_eval_constraints(p3g_ctx* ctx, const p3g_air_op* ops, uint64_t nops, const p3g_matrix* main,
                         const p3g_matrix* aux, const uint64_t* periodic, uint64_t periodic_rows,
                         uint64_t periodic_cols, const uint64_t* publics, uint64_t num_publics,
                         const uint64_t* ext_inputs, uint64_t num_ext_inputs, uint64_t n, uint32_t log_d,
                         uint64_t shift, const uint64_t alpha[2], uint64_t* accumulator) {
  if (!ctx || !ops || !main || !alpha || !accumulator) return P3G_E_INVAL;
  if ((num_publics && !publics) || (num_ext_inputs && !ext_inputs) || (periodic_rows && periodic_cols && !periodic))
    return P3G_E_INVAL;
  return guard([&] {
    Ctx& c = ctx->c;
    P3G_CHECK_HIP(hipSetDevice(c.device));
    if (alpha[0] >= P || alpha[1] >= P) throw ApiError(P3G_E_INVAL, "non-canonical alpha");
    DevMat m{const_cast<uint64_t*>(main->data), main->height, main->width, false};
    DevMat x{};
    if (aux) x = DevMat{const_cast<uint64_t*>(aux->data), aux->height, aux->width, false};
    std::vector<ef2> ext(num_ext_inputs);
    for (uint64_t i = 0; i < num_ext_inputs; i++) ext[i] = ef2{ext_inputs[2 * i], ext_inputs[2 * i + 1]};
    eval_constraints(c, reinterpret_cast<const AirOpIn*>(ops), nops, m, aux ? &x : nullptr, periodic, periodic_rows,
                     periodic_cols, publics, num_publics, ext.data(), num_ext_inputs, n, log_d, shift,
                     ef2{alpha[0], alpha[1]}, reinterpret_cast<ef2*>(accumulator), c.stream);
  });
}

int p3g_air_compile(const p3g_air_op* ops, uint64_t nops, uint64_t* code_bytes) {
  if (!ops || !code_bytes) return P3G_E_INVAL;
  return guard([&] { *code_bytes = air_compile(reinterpret_cast<const AirOpIn*>(ops), nops); });
}

uint32_t p3g_lmcs_alignment(uint32_t hash_kind) { return hash_alignment(hash_kind); }

static std::vector<DevMat> import_mats(Ctx& c, const p3g_matrix* mats, uint64_t k, uint32_t flags) {
  std::vector<DevMat> v;
  const bool dev = flags & P3G_DEVICE_PTRS;
  try {
    for (uint64_t i = 0; i < k; i++) {
      if (!mats[i].data && mats[i].height * mats[i].width) throw ApiError(P3G_E_INVAL, "null matrix data");
      DevMat m;
      m.height = mats[i].height;
      m.width = mats[i].width;
      if (dev) {
        m.ptr = const_cast<uint64_t*>(mats[i].data);
        m.owned = (flags & P3G_TAKE_OWNERSHIP) != 0;
      } else {
        m.ptr = to_device(c, mats[i].data, m.height * m.width);
        m.owned = true;
      }
      v.push_back(m);
    }
  } catch (...) {
    for (auto& m : v)
      if (!dev) c.release(m.ptr);
    throw;
  }
  return v;
}

static void check_hash(uint32_t h) {
  if (h != HASH_P2 && h != HASH_KECCAK) throw ApiError(P3G_E_INVAL, "unknown hash kind");
}

int p3g_lmcs_build(p3g_ctx* ctx, uint32_t hash_kind, const p3g_matrix* mats, uint64_t k, uint32_t alignment,
                   uint32_t flags, p3g_tree** out) {
  if (!ctx || !out || (!mats && k)) return P3G_E_INVAL;
  *out = nullptr;
  return guard([&] {
    Ctx& c = ctx->c;
    check_hash(hash_kind);
    P3G_CHECK_HIP(hipSetDevice(c.device));
    if (k == 0) throw ApiError(P3G_E_EMPTY, "cannot commit empty batch");
    std::vector<DevMat> dm = import_mats(c, mats, k, flags);
    Tree* t = nullptr;
    try {
      t = lmcs_build(c, hash_kind, dm, alignment, c.stream);
    } catch (...) {
      for (auto& m : dm)
        if (m.owned && !(flags & P3G_DEVICE_PTRS)) c.release(m.ptr);
      throw;
    }
    *out = new p3g_tree{t};
  });
}

int p3g_lmcs_build_salted(p3g_ctx* ctx, uint32_t hash_kind, const p3g_matrix* mats, uint64_t k, uint32_t alignment,
                          uint32_t flags, const p3g_matrix* salt, p3g_tree** out) {
  if (!ctx || !out || (!mats && k) || !salt) return P3G_E_INVAL;
  *out = nullptr;
  return guard([&] {
    Ctx& c = ctx->c;
    check_hash(hash_kind);
    P3G_CHECK_HIP(hipSetDevice(c.device));
    if (k == 0) throw ApiError(P3G_E_EMPTY, "cannot commit empty batch");
    if (salt->width == 0 || salt->height != mats[k - 1].height)
      throw ApiError(P3G_E_INVAL, "salt must have the tree height and a non-zero width");
    std::vector<DevMat> dm = import_mats(c, mats, k, flags);
    std::vector<DevMat> ds;
    Tree* t = nullptr;
    try {
      ds = import_mats(c, salt, 1, flags);
      t = lmcs_build(c, hash_kind, dm, alignment, c.stream, true, &ds[0]);
    } catch (...) {
      for (auto& m : dm)
        if (m.owned && !(flags & P3G_DEVICE_PTRS)) c.release(m.ptr);
      for (auto& m : ds)
        if (m.owned && !(flags & P3G_DEVICE_PTRS)) c.release(m.ptr);
      throw;
    }
    *out = new p3g_tree{t};
  });
}

int p3g_tree_salt(const p3g_tree* t, uint64_t index, uint64_t* out, uint64_t cap, uint64_t* len) {
  if (!t || !len) return P3G_E_INVAL;
  return guard([&] {
    const Tree& tr = *t->t;
    *len = tr.salt.ptr ? tr.salt.width : 0;
    if (!tr.salt.ptr) return;
    if (index >= tr.height) throw ApiError(P3G_E_RANGE, "leaf index out of range");
    if (out) {
      if (cap < tr.salt.width) throw ApiError(P3G_E_RANGE, "output buffer too small");
      P3G_CHECK_HIP(hipSetDevice(tr.ctx->device));
      std::vector<uint64_t> raw;
      gather_to_host(*tr.ctx, {{tr.salt.ptr + index * tr.salt.width, tr.salt.width}}, raw, tr.ctx->stream);
      memcpy(out, raw.data(), tr.salt.width * 8);
    }
  });
}

int p3g_lmcs_hash(uint32_t hash_kind, const uint64_t* elems, const uint64_t* row_widths, uint64_t nrows,
                  uint64_t out[4]) {
  if (!out || (nrows && !row_widths)) return P3G_E_INVAL;
  return guard([&] {
    check_hash(hash_kind);
    uint64_t st[25] = {};
    uint64_t off = 0;
    for (uint64_t r = 0; r < nrows; r++) {
      host_sponge_absorb(hash_kind, st, elems + off, row_widths[r]);
      off += row_widths[r];
    }
    memcpy(out, st, 32);
  });
}

int p3g_lmcs_compress(uint32_t hash_kind, const uint64_t left[4], const uint64_t right[4], uint64_t out[4]) {
  if (!left || !right || !out) return P3G_E_INVAL;
  return guard([&] {
    check_hash(hash_kind);
    host_compress(hash_kind, left, right, out);
  });
}

int p3g_tree_root(const p3g_tree* t, uint64_t out[4]) {
  if (!t || !out) return P3G_E_INVAL;
  memcpy(out, t->t->root, 32);
  return P3G_OK;
}

uint64_t p3g_tree_height(const p3g_tree* t) { return t ? t->t->height : 0; }
uint32_t p3g_tree_alignment(const p3g_tree* t) { return t ? t->t->alignment : 0; }
uint64_t p3g_tree_num_matrices(const p3g_tree* t) { return t ? t->t->mats.size() : 0; }

int p3g_tree_widths(const p3g_tree* t, uint64_t* out) {
  if (!t || !out) return P3G_E_INVAL;
  for (size_t i = 0; i < t->t->mats.size(); i++) out[i] = aligned_len(t->t->mats[i].width, t->t->alignment);
  return P3G_OK;
}

int p3g_tree_rows(p3g_tree* t, uint64_t index, uint64_t* out, uint64_t cap, uint64_t* len) {
  if (!t || !len) return P3G_E_INVAL;
  return guard([&] {
    P3G_CHECK_HIP(hipSetDevice(t->t->ctx->device));
    std::vector<uint64_t> rows;
    tree_rows(*t->t, index, rows);
    *len = rows.size();
    if (out) {
      if (cap < rows.size()) throw ApiError(P3G_E_RANGE, "output buffer too small");
      memcpy(out, rows.data(), rows.size() * 8);
    }
  });
}

int p3g_tree_layer(p3g_tree* t, uint32_t layer, uint64_t* out) {
  if (!t || !out) return P3G_E_INVAL;
  return guard([&] {
    Tree& tr = *t->t;
    if (layer > tr.depth) throw ApiError(P3G_E_RANGE, "layer out of range");
    P3G_CHECK_HIP(hipSetDevice(tr.ctx->device));
    P3G_CHECK_HIP(hipMemcpyAsync(out, tr.digests + tr.layer_off[layer] * 4, (tr.height >> layer) * 32,
                                 hipMemcpyDeviceToHost, tr.ctx->stream));
    P3G_CHECK_HIP(hipStreamSynchronize(tr.ctx->stream));
  });
}

int p3g_tree_matrix(p3g_tree* t, uint64_t i, const uint64_t** dptr, uint64_t* height, uint64_t* width) {
  if (!t || i >= t->t->mats.size()) return P3G_E_INVAL;
  if (dptr) *dptr = t->t->mats[i].ptr;
  if (height) *height = t->t->mats[i].height;
  if (width) *width = t->t->mats[i].width;
  return P3G_OK;
}

int p3g_tree_prove_batch(p3g_tree* t, const uint64_t* indices, uint64_t n, p3g_transcript* ch) {
  if (!t || !ch || (n && !indices)) return P3G_E_INVAL;
  return guard([&] {
    P3G_CHECK_HIP(hipSetDevice(t->t->ctx->device));
    std::vector<uint64_t> idx(indices, indices + n);
    tree_prove_batch(*t->t, idx, ch->tx, t->t->ctx->stream);
  });
}

void p3g_tree_free(p3g_tree* t) {
  if (!t) return;
  (void)hipSetDevice(t->t->ctx->device);
  (void)hipStreamSynchronize(t->t->ctx->stream);
  delete t->t;
  delete t;
}

int p3g_commit_traces(p3g_ctx* ctx, uint32_t hash_kind, const p3g_matrix* traces, uint64_t k, uint32_t log_blowup,
                      uint32_t flags, p3g_tree** out) {
  if (!ctx || !out || (!traces && k)) return P3G_E_INVAL;
  *out = nullptr;
  return guard([&] {
    Ctx& c = ctx->c;
    check_hash(hash_kind);
    P3G_CHECK_HIP(hipSetDevice(c.device));
    if (k == 0) throw ApiError(P3G_E_EMPTY, "at least one trace required");
    std::vector<DevMat> dm = import_mats(c, traces, k, flags & P3G_DEVICE_PTRS);
    Tree* t = nullptr;
    try {
      t = commit_traces(c, hash_kind, dm, log_blowup, c.stream);
    } catch (...) {
      if (!(flags & P3G_DEVICE_PTRS))
        for (auto& m : dm) c.release(m.ptr);
      throw;
    }
    if (!(flags & P3G_DEVICE_PTRS))
      for (auto& m : dm) c.release(m.ptr);
    *out = new p3g_tree{t};
  });
}

int p3g_commit_quotient(p3g_ctx* ctx, uint32_t hash_kind, const uint64_t* q_evals, uint64_t n, uint64_t d,
                        uint32_t log_blowup, uint32_t flags, p3g_tree** out) {
  if (!ctx || !out || !q_evals) return P3G_E_INVAL;
  *out = nullptr;
  return guard([&] {
    Ctx& c = ctx->c;
    check_hash(hash_kind);
    P3G_CHECK_HIP(hipSetDevice(c.device));
    if (n == 0 || d == 0) throw ApiError(P3G_E_EMPTY, "empty quotient");
    // shape checks before any upload (commit_quotient repeats them for its other callers)
    if (n & (n - 1)) throw ApiError(P3G_E_NOT_POW2, "quotient trace height must be a power of two");
    if (d & (d - 1)) throw ApiError(P3G_E_NOT_POW2, "constraint degree must be a power of two");
    if (log_blowup > 32 || d > (1ull << log_blowup)) throw ApiError(P3G_E_INVAL, "blowup must be at least the constraint degree");
    if (log2_exact(n) + log_blowup > 32) throw ApiError(P3G_E_RANGE, "LDE height exceeds the field two-adicity");
    const bool dev = flags & P3G_DEVICE_PTRS;
    const uint64_t* dq = dev ? q_evals : to_device(c, q_evals, n * 2 * d);
    Tree* t = nullptr;
    try {
      t = commit_quotient(c, hash_kind, dq, n, d, log_blowup, c.stream);
    } catch (...) {
      if (!dev) c.release(const_cast<uint64_t*>(dq));
      throw;
    }
    if (!dev) c.release(const_cast<uint64_t*>(dq));
    *out = new p3g_tree{t};
  });
}

int p3g_quotient_extend_scale(p3g_ctx* ctx, uint64_t* acc, uint64_t m, uint64_t target, const uint64_t beta[2]) {
  if (!ctx || !acc || !beta || beta[0] >= P || beta[1] >= P) return P3G_E_INVAL;
  return guard([&] {
    Ctx& c = ctx->c;
    P3G_CHECK_HIP(hipSetDevice(c.device));
    quotient_extend_scale(c, reinterpret_cast<ef2*>(acc), m, target, ef2{beta[0], beta[1]}, c.stream);
  });
}

int p3g_quotient_divide_vanishing(p3g_ctx* ctx, uint64_t* num, uint64_t n, uint64_t d, uint64_t shift) {
  if (!ctx || !num || shift >= P) return P3G_E_INVAL;
  return guard([&] {
    Ctx& c = ctx->c;
    P3G_CHECK_HIP(hipSetDevice(c.device));
    quotient_divide_vanishing(c, reinterpret_cast<ef2*>(num), n, d, shift, c.stream);
  });
}

int p3g_transcript_create(uint32_t kind, p3g_transcript** out) {
  if (!out) return P3G_E_INVAL;
  if (kind != CH_DUPLEX_P2 && kind != CH_KECCAK) return P3G_E_INVAL;
  *out = new p3g_transcript();
  (*out)->tx.kind = kind;
  return P3G_OK;
}

void p3g_transcript_free(p3g_transcript* ch) { delete ch; }

int p3g_transcript_clone(const p3g_transcript* ch, p3g_transcript** out) {
  if (!ch || !out) return P3G_E_INVAL;
  *out = new p3g_transcript(*ch);
  return P3G_OK;
}

static bool canonical(const uint64_t* v, uint64_t n) {
  for (uint64_t i = 0; i < n; i++)
    if (v[i] >= P) return false;
  return true;
}

int p3g_transcript_observe(p3g_transcript* ch, const uint64_t* values, uint64_t n) {
  if (!ch || (n && !values) || !canonical(values, n)) return P3G_E_INVAL;
  for (uint64_t i = 0; i < n; i++) ch->tx.observe(values[i]);
  return P3G_OK;
}

int p3g_transcript_observe_commitment(p3g_transcript* ch, const uint64_t c[4]) {
  if (!ch || !c) return P3G_E_INVAL;
  ch->tx.observe_commitment(c);
  return P3G_OK;
}

int p3g_transcript_sample(p3g_transcript* ch, uint64_t* out) {
  if (!ch || !out) return P3G_E_INVAL;
  *out = ch->tx.sample();
  return P3G_OK;
}

int p3g_transcript_sample_ext(p3g_transcript* ch, uint64_t out[2]) {
  if (!ch || !out) return P3G_E_INVAL;
  ef2 e = ch->tx.sample_ef();
  out[0] = e.c0;
  out[1] = e.c1;
  return P3G_OK;
}

int p3g_transcript_sample_bits(p3g_transcript* ch, uint32_t bits, uint64_t* out) {
  if (!ch || !out || bits >= 32) return P3G_E_INVAL;
  *out = ch->tx.sample_bits(bits);
  return P3G_OK;
}

int p3g_transcript_send_fields(p3g_transcript* ch, const uint64_t* values, uint64_t n) {
  if (!ch || (n && !values) || !canonical(values, n)) return P3G_E_INVAL;
  ch->tx.send_fields(values, n);
  return P3G_OK;
}

int p3g_transcript_send_commitment(p3g_transcript* ch, const uint64_t c[4]) {
  if (!ch || !c) return P3G_E_INVAL;
  ch->tx.send_commitment(c);
  return P3G_OK;
}

int p3g_transcript_hint_fields(p3g_transcript* ch, const uint64_t* values, uint64_t n) {
  if (!ch || (n && !values)) return P3G_E_INVAL;
  ch->tx.hint_fields(values, n);
  return P3G_OK;
}

int p3g_transcript_hint_commitment(p3g_transcript* ch, const uint64_t c[4]) {
  if (!ch || !c) return P3G_E_INVAL;
  ch->tx.hint_commitment(c);
  return P3G_OK;
}

int p3g_transcript_grind(p3g_transcript* ch, uint32_t bits, uint64_t* witness) {
  if (!ch || bits >= 32) return P3G_E_INVAL;
  uint64_t w = ch->tx.grind(bits);
  if (witness) *witness = w;
  return P3G_OK;
}

uint64_t p3g_transcript_num_fields(const p3g_transcript* ch) { return ch ? ch->tx.fields.size() : 0; }
uint64_t p3g_transcript_num_commitments(const p3g_transcript* ch) { return ch ? ch->tx.comms.size() / 4 : 0; }

int p3g_transcript_data(const p3g_transcript* ch, uint64_t* fields, uint64_t* commitments) {
  if (!ch) return P3G_E_INVAL;
  if (fields && !ch->tx.fields.empty()) memcpy(fields, ch->tx.fields.data(), ch->tx.fields.size() * 8);
  if (commitments && !ch->tx.comms.empty()) memcpy(commitments, ch->tx.comms.data(), ch->tx.comms.size() * 8);
  return P3G_OK;
}

static void put_le64(uint8_t* p, uint64_t v) {
  for (int i = 0; i < 8; i++) p[i] = (uint8_t)(v >> (8 * i));
}
static uint64_t get_le64(const uint8_t* p) {
  uint64_t v = 0;
  for (int i = 0; i < 8; i++) v |= (uint64_t)p[i] << (8 * i);
  return v;
}
static const uint8_t PROOF_MAGIC[4] = {'P', '3', 'G', 'T'};
constexpr uint32_t PROOF_VERSION = 1;
constexpr uint64_t PROOF_HEADER = 24;

int p3g_proof_encode(const uint64_t* fields, uint64_t nf, const uint64_t* commitments, uint64_t nc, uint8_t* out,
                     uint64_t cap, uint64_t* len) {
  if (!len || (nf && !fields) || (nc && !commitments)) return P3G_E_INVAL;
  if (nf > (1ull << 56) || nc > (1ull << 54)) return P3G_E_RANGE;
  const uint64_t need = PROOF_HEADER + 8 * nf + 32 * nc;
  *len = need;
  if (!out) return P3G_OK;
  if (cap < need) return P3G_E_RANGE;
  if (!canonical(fields, nf) || !canonical(commitments, 4 * nc)) return P3G_E_INVAL;
  memcpy(out, PROOF_MAGIC, 4);
  for (int i = 0; i < 4; i++) out[4 + i] = (uint8_t)(PROOF_VERSION >> (8 * i));
  put_le64(out + 8, nf);
  put_le64(out + 16, nc);
  uint8_t* p = out + PROOF_HEADER;
  for (uint64_t i = 0; i < nf; i++, p += 8) put_le64(p, fields[i]);
  for (uint64_t i = 0; i < 4 * nc; i++, p += 8) put_le64(p, commitments[i]);
  return P3G_OK;
}

int p3g_proof_decode(const uint8_t* in, uint64_t len, uint64_t* fields, uint64_t* nf, uint64_t* commitments,
                     uint64_t* nc) {
  if (!in || !nf || !nc || len < PROOF_HEADER) return P3G_E_INVAL;
  if (memcmp(in, PROOF_MAGIC, 4) != 0) return P3G_E_INVAL;
  uint32_t ver = 0;
  for (int i = 0; i < 4; i++) ver |= (uint32_t)in[4 + i] << (8 * i);
  if (ver != PROOF_VERSION) return P3G_E_INVAL;
  const uint64_t f = get_le64(in + 8), c = get_le64(in + 16);
  if (f > (1ull << 56) || c > (1ull << 54) || len != PROOF_HEADER + 8 * f + 32 * c) return P3G_E_INVAL;
  const uint8_t* p = in + PROOF_HEADER;
  for (uint64_t i = 0; i < f + 4 * c; i++)
    if (get_le64(p + 8 * i) >= P) return P3G_E_INVAL;
  *nf = f;
  *nc = c;
  if (fields)
    for (uint64_t i = 0; i < f; i++) fields[i] = get_le64(p + 8 * i);
  if (commitments)
    for (uint64_t i = 0; i < 4 * c; i++) commitments[i] = get_le64(p + 8 * (f + i));
  return P3G_OK;
}

int p3g_transcript_encode(const p3g_transcript* ch, uint8_t* out, uint64_t cap, uint64_t* len) {
  if (!ch) return P3G_E_INVAL;
  return p3g_proof_encode(ch->tx.fields.data(), ch->tx.fields.size(), ch->tx.comms.data(), ch->tx.comms.size() / 4,
                          out, cap, len);
}

int p3g_open_with_channel(p3g_ctx* ctx, const p3g_pcs_params* params, uint32_t log_lde_height,
                          const uint64_t* eval_points, uint32_t n_points, p3g_tree* const* trees, uint64_t ntrees,
                          p3g_transcript* ch) {
  if (!ctx || !params || !eval_points || !ch || (!trees && ntrees)) return P3G_E_INVAL;
  return guard([&] {
    Ctx& c = ctx->c;
    P3G_CHECK_HIP(hipSetDevice(c.device));
    const PcsParams p = pcs_params(params);
    if (n_points < 1 || n_points > 4) throw ApiError(P3G_E_INVAL, "1 to 4 evaluation points supported");
    if (!canonical(eval_points, 2ull * n_points)) throw ApiError(P3G_E_INVAL, "non-canonical evaluation point");
    std::vector<ef2> z(n_points);
    for (uint32_t j = 0; j < n_points; j++) z[j] = ef2{eval_points[2 * j], eval_points[2 * j + 1]};
    std::vector<Tree*> tv;
    for (uint64_t i = 0; i < ntrees; i++) {
      if (!trees[i]) throw ApiError(P3G_E_INVAL, "null tree");
      tv.push_back(trees[i]->t);
    }
    open_with_channel(c, p, log_lde_height, z.data(), (int)n_points, tv, ch->tx, c.stream);
  });
}

// ---- DEEP and FRI steps
namespace {
std::vector<ef2> ext_points(const uint64_t* v, uint32_t n) {
  // the count first: the caller's buffer holds 2 n words only when n is valid
  if (n < 1 || n > 4) throw ApiError(P3G_E_INVAL, "1 to 4 evaluation points supported");
  if (!canonical(v, 2ull * n)) throw ApiError(P3G_E_INVAL, "non-canonical extension element");
  std::vector<ef2> z(n);
  for (uint32_t j = 0; j < n; j++) z[j] = ef2{v[2 * j], v[2 * j + 1]};
  return z;
}
std::vector<Tree*> tree_list(p3g_tree* const* trees, uint64_t ntrees) {
  if (ntrees == 0) throw ApiError(P3G_E_EMPTY, "at least one trace tree required");
  std::vector<Tree*> tv;
  for (uint64_t i = 0; i < ntrees; i++) {
    if (!trees[i]) throw ApiError(P3G_E_INVAL, "null tree");
    tv.push_back(trees[i]->t);
  }
  return tv;
}
uint64_t eval_words(const std::vector<Tree*>& tv, uint32_t n_points) {
  if (n_points < 1 || n_points > 4) throw ApiError(P3G_E_INVAL, "1 to 4 evaluation points supported");
  uint64_t cols = 0;
  for (auto* t : tv)
    for (auto& m : t->mats) cols += m.width;
  return cols * n_points * 2;
}
}  // namespace

int p3g_deep_batch_eval(p3g_ctx* ctx, uint32_t log_blowup, const uint64_t* eval_points, uint32_t n_points,
                        p3g_tree* const* trees, uint64_t ntrees, uint64_t* evals, uint64_t cap, uint64_t* len) {
  if (!ctx || !eval_points || !len || (!trees && ntrees)) return P3G_E_INVAL;
  return guard([&] {
    Ctx& c = ctx->c;
    const std::vector<Tree*> tv = tree_list(trees, ntrees);
    const std::vector<ef2> z = ext_points(eval_points, n_points);
    *len = eval_words(tv, n_points);
    // the size query makes the same checks of the inputs as the call
    deep_check_inputs(tv, log_blowup, z.data(), (int)n_points);
    if (!evals) return;
    if (cap < *len) throw ApiError(P3G_E_RANGE, "evaluation buffer too small");
    P3G_CHECK_HIP(hipSetDevice(c.device));
    std::vector<ef2> ev;
    deep_batch_eval(c, log_blowup, tv, z.data(), (int)n_points, ev, c.stream);
    memcpy(evals, ev.data(), ev.size() * sizeof(ef2));
  });
}

int p3g_deep_quotient(p3g_ctx* ctx, uint32_t log_blowup, const uint64_t* eval_points, uint32_t n_points,
                      p3g_tree* const* trees, uint64_t ntrees, const uint64_t* evals, uint64_t len,
                      const uint64_t alpha[2], const uint64_t beta[2], uint64_t* out) {
  if (!ctx || !eval_points || !evals || !alpha || !beta || !out || (!trees && ntrees)) return P3G_E_INVAL;
  return guard([&] {
    Ctx& c = ctx->c;
    const std::vector<Tree*> tv = tree_list(trees, ntrees);
    const std::vector<ef2> z = ext_points(eval_points, n_points);
    const std::vector<ef2> ab = ext_points(alpha, 1), bb = ext_points(beta, 1);
    if (len != eval_words(tv, n_points)) throw ApiError(P3G_E_INVAL, "evaluation count does not match the trees");
    if (!canonical(evals, len)) throw ApiError(P3G_E_INVAL, "non-canonical evaluation");
    P3G_CHECK_HIP(hipSetDevice(c.device));
    deep_quotient(c, log_blowup, tv, z.data(), (int)n_points, reinterpret_cast<const ef2*>(evals), ab[0], bb[0],
                  reinterpret_cast<ef2*>(out), c.stream);
  });
}

int p3g_fri_fold(p3g_ctx* ctx, const uint64_t* in, uint64_t n, uint32_t log_arity, const uint64_t beta[2],
                 uint64_t* out) {
  if (!ctx || !in || !out || !beta) return P3G_E_INVAL;
  return guard([&] {
    Ctx& c = ctx->c;
    const std::vector<ef2> b = ext_points(beta, 1);
    P3G_CHECK_HIP(hipSetDevice(c.device));
    fri_fold(c, reinterpret_cast<const ef2*>(in), n, log_arity, b[0], reinterpret_cast<ef2*>(out), c.stream);
  });
}

int p3g_fri_commit(p3g_ctx* ctx, uint32_t hash_kind, const uint64_t* evals, uint64_t n, uint32_t log_arity,
                   p3g_tree** out, uint64_t root[4]) {
  if (!ctx || !evals || !out) return P3G_E_INVAL;
  *out = nullptr;
  return guard([&] {
    Ctx& c = ctx->c;
    check_hash(hash_kind);
    if (log_arity < 1 || log_arity > 3) throw ApiError(P3G_E_INVAL, "folding arity must be 2, 4 or 8");
    if (n == 0 || (n & (n - 1))) throw ApiError(P3G_E_NOT_POW2, "domain size must be a power of two");
    if ((n >> log_arity) == 0) throw ApiError(P3G_E_INVAL, "domain smaller than the folding arity");
    P3G_CHECK_HIP(hipSetDevice(c.device));
    DevMat m;
    m.ptr = const_cast<uint64_t*>(evals);
    m.height = n >> log_arity;
    m.width = 2ull << log_arity;
    m.owned = false;
    Tree* t = lmcs_build(c, hash_kind, {m}, 1, c.stream);
    if (root) memcpy(root, t->root, 32);
    *out = new p3g_tree{t};
  });
}

int p3g_fri_final_poly(const uint64_t* evals, uint64_t n, uint64_t* coeffs) {
  if (!evals || !coeffs) return P3G_E_INVAL;
  if (n == 0 || (n & (n - 1))) return P3G_E_NOT_POW2;
  if (n > (1ull << 24)) return P3G_E_RANGE;
  return guard([&] {
    if (!canonical(evals, 2 * n)) throw ApiError(P3G_E_INVAL, "non-canonical evaluation");
    std::vector<ef2> out(n);
    fri_final_poly(reinterpret_cast<const ef2*>(evals), n, out.data());
    memcpy(coeffs, out.data(), n * sizeof(ef2));
  });
}

namespace {
// ProverChannel implemented by the caller through function pointers.
struct CallbackChannel : Channel {
  const p3g_channel* cb;
  explicit CallbackChannel(const p3g_channel* c) : cb(c) {}
  static void ok(int rc) {
    if (rc != 0) throw ApiError(P3G_E_TRANSCRIPT, "channel callback failed");
  }
  void send_fields(const uint64_t* v, size_t n) override { ok(cb->send_fields(cb->user, v, n)); }
  void send_commitment(const uint64_t* c4) override { ok(cb->send_commitment(cb->user, c4)); }
  void hint_fields(const uint64_t* v, size_t n) override { ok(cb->hint_fields(cb->user, v, n)); }
  void hint_commitment(const uint64_t* c4) override { ok(cb->hint_commitment(cb->user, c4)); }
  uint64_t grind(unsigned bits) override {
    uint64_t w = 0;
    ok(cb->grind(cb->user, bits, &w));
    return w;
  }
  uint64_t sample() override {
    uint64_t v = 0;
    ok(cb->sample(cb->user, &v));
    return v;
  }
  uint64_t sample_bits(unsigned bits) override {
    uint64_t v = 0;
    ok(cb->sample_bits(cb->user, bits, &v));
    return v;
  }
};

bool channel_complete(const p3g_channel* c) {
  return c && c->send_fields && c->send_commitment && c->hint_fields && c->hint_commitment && c->grind &&
         c->sample && c->sample_bits;
}
}  // namespace

int p3g_tree_prove_batch_cb(p3g_tree* t, const uint64_t* indices, uint64_t n, const p3g_channel* ch) {
  if (!t || !channel_complete(ch) || (n && !indices)) return P3G_E_INVAL;
  return guard([&] {
    P3G_CHECK_HIP(hipSetDevice(t->t->ctx->device));
    std::vector<uint64_t> idx(indices, indices + n);
    CallbackChannel cc(ch);
    tree_prove_batch(*t->t, idx, cc, t->t->ctx->stream);
  });
}

int p3g_open_with_callbacks(p3g_ctx* ctx, const p3g_pcs_params* params, uint32_t log_lde_height,
                            const uint64_t* eval_points, uint32_t n_points, p3g_tree* const* trees, uint64_t ntrees,
                            const p3g_channel* ch) {
  if (!ctx || !params || !eval_points || !channel_complete(ch) || (!trees && ntrees)) return P3G_E_INVAL;
  return guard([&] {
    Ctx& c = ctx->c;
    P3G_CHECK_HIP(hipSetDevice(c.device));
    const PcsParams p = pcs_params(params);
    if (n_points < 1 || n_points > 4) throw ApiError(P3G_E_INVAL, "1 to 4 evaluation points supported");
    if (!canonical(eval_points, 2ull * n_points)) throw ApiError(P3G_E_INVAL, "non-canonical evaluation point");
    std::vector<ef2> z(n_points);
    for (uint32_t j = 0; j < n_points; j++) z[j] = ef2{eval_points[2 * j], eval_points[2 * j + 1]};
    std::vector<Tree*> tv;
    for (uint64_t i = 0; i < ntrees; i++) {
      if (!trees[i]) throw ApiError(P3G_E_INVAL, "null tree");
      tv.push_back(trees[i]->t);
    }
    CallbackChannel cc(ch);
    open_with_channel(c, p, log_lde_height, z.data(), (int)n_points, tv, cc, c.stream);
  });
}

int p3g_copy_device(p3g_ctx* ctx, void* dst, const void* src, uint64_t bytes) {
  if (!ctx || ((!dst || !src) && bytes)) return P3G_E_INVAL;
  return guard([&] {
    if (!bytes) return;
    P3G_CHECK_HIP(hipSetDevice(ctx->c.device));
    P3G_CHECK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->c.stream));
    P3G_CHECK_HIP(hipStreamSynchronize(ctx->c.stream));
  });
}

int p3g_fill_random_device(p3g_ctx* ctx, uint64_t seed, uint64_t n, uint64_t* dptr) {
  if (!ctx || (!dptr && n)) return P3G_E_INVAL;
  return guard([&] {
    P3G_CHECK_HIP(hipSetDevice(ctx->c.device));
    fill_random_device(ctx->c, seed, n, dptr, ctx->c.stream);
    P3G_CHECK_HIP(hipStreamSynchronize(ctx->c.stream));
  });
}

int p3g_commit_plan_create(p3g_ctx* ctx, uint32_t hash_kind, const uint64_t* heights, const uint64_t* widths,
                           uint64_t k, uint32_t log_blowup, uint32_t flags, p3g_commit_plan** out) {
  if (!ctx || !out || (k && (!heights || !widths))) return P3G_E_INVAL;
  *out = nullptr;
  return guard([&] {
    check_hash(hash_kind);
    P3G_CHECK_HIP(hipSetDevice(ctx->c.device));
    std::vector<uint64_t> h(heights, heights + k), w(widths, widths + k);
    CommitPlan* p = make_commit_plan(ctx->c, hash_kind, h, w, log_blowup, (flags & P3G_PLAN_GRAPH) != 0,
                                     (flags & P3G_PLAN_AUTOTUNE) != 0, (flags & P3G_PLAN_PER_TRACE) == 0);
    *out = new p3g_commit_plan{p};
  });
}

int p3g_commit_plan_input(p3g_commit_plan* plan, uint64_t i, uint64_t** dptr) {
  if (!plan || !dptr || i >= plan->p->traces.size()) return P3G_E_INVAL;
  *dptr = plan->p->traces[i].ptr;
  return P3G_OK;
}

int p3g_commit_plan_info(const p3g_commit_plan* plan, uint32_t* kmax, uint32_t* leaf_threads, uint32_t* launches,
                         uint32_t* grouped) {
  if (!plan) return P3G_E_INVAL;
  if (grouped) *grouped = plan->p->grouped ? 1u : 0u;
  if (kmax) *kmax = plan->p->kmax;
  if (leaf_threads) *leaf_threads = (uint32_t)plan->p->leaf_threads;
  if (launches) *launches = (uint32_t)plan->p->launches.size();
  return P3G_OK;
}

int p3g_commit_plan_run(p3g_commit_plan* plan, p3g_tree** view) {
  if (!plan || !view) return P3G_E_INVAL;
  *view = nullptr;
  return guard([&] {
    CommitPlan& p = *plan->p;
    P3G_CHECK_HIP(hipSetDevice(p.ctx->device));
    {
      ProfScope ps(*p.ctx, "batch_commit", p.ctx->stream);
      p.run(p.ctx->stream);
    }
    *view = new p3g_tree{p.tree_view()};
  });
}

void p3g_commit_plan_destroy(p3g_commit_plan* plan) {
  if (!plan) return;
  (void)hipSetDevice(plan->p->ctx->device);
  (void)hipStreamSynchronize(plan->p->ctx->stream);
  delete plan->p;
  delete plan;
}

int p3g_prove_planned(p3g_ctx* ctx, p3g_commit_plan* plan, const p3g_pcs_params* params, p3g_transcript* ch,
                      uint64_t root[4], uint64_t eval_points_out[4]) {
  if (!ctx || !plan || !params || !ch) return P3G_E_INVAL;
  if (params->log_blowup != plan->p->log_blowup) return P3G_E_INVAL;
  // reject bad parameters before the commitment touches the transcript
  if (int rc0 = guard([&] { (void)pcs_params(params); })) return rc0;
  p3g_tree* tree = nullptr;
  int rc = p3g_commit_plan_run(plan, &tree);
  if (rc != P3G_OK) return rc;
  uint64_t r[4], z[4];
  p3g_tree_root(tree, r);
  ch->tx.observe_commitment(r);
  ef2 z0 = ch->tx.sample_ef();
  ef2 z1 = ch->tx.sample_ef();
  z[0] = z0.c0;
  z[1] = z0.c1;
  z[2] = z1.c0;
  z[3] = z1.c1;
  uint32_t logL = log2_exact(p3g_tree_height(tree));
  rc = p3g_open_with_channel(ctx, params, logL, z, 2, &tree, 1, ch);
  p3g_tree_free(tree);
  if (root) memcpy(root, r, 32);
  if (eval_points_out) memcpy(eval_points_out, z, 32);
  return rc;
}

int p3g_prove_many(p3g_ctx* ctx, uint32_t hash_kind, const p3g_pcs_params* params, const p3g_matrix* traces,
                   uint64_t k, uint64_t t, uint32_t flags, p3g_transcript* const* chs, uint64_t* roots,
                   uint64_t* eval_points) {
  if (!ctx || !params || !chs || (k && (!traces || t == 0)) || k > 65535) return P3G_E_INVAL;
  if (k == 0) return P3G_OK;
  for (uint64_t j = 0; j < k; j++)
    if (!chs[j]) return P3G_E_INVAL;
  return guard([&] {
    check_hash(hash_kind);
    P3G_CHECK_HIP(hipSetDevice(ctx->c.device));
    Ctx& c = ctx->c;
    const PcsParams p = pcs_params(params);
    std::vector<std::vector<DevMat>> tr(k);
    std::vector<uint64_t*> uploaded;
    try {
      for (uint64_t j = 0; j < k; j++)
        for (uint64_t i = 0; i < t; i++) {
          const p3g_matrix& m = traces[j * t + i];
          if (!m.data && m.height * m.width) throw ApiError(E_INVAL, "null matrix");
          DevMat d;
          d.height = m.height;
          d.width = m.width;
          d.owned = false;
          if (flags & P3G_DEVICE_PTRS) {
            d.ptr = const_cast<uint64_t*>(m.data);
          } else {
            d.ptr = to_device(c, m.data, m.height * m.width);
            uploaded.push_back(d.ptr);
          }
          tr[j].push_back(d);
        }
      std::vector<Transcript*> txs(k);
      for (uint64_t j = 0; j < k; j++) txs[j] = &chs[j]->tx;
      std::vector<std::array<uint64_t, 4>> r;
      std::vector<std::array<ef2, 2>> z;
      prove_many(c, hash_kind, p, tr, txs, r, z, c.stream);
      for (uint64_t j = 0; j < k; j++) {
        if (roots) memcpy(roots + 4 * j, r[j].data(), 32);
        if (eval_points) {
          eval_points[4 * j] = z[j][0].c0;
          eval_points[4 * j + 1] = z[j][0].c1;
          eval_points[4 * j + 2] = z[j][1].c0;
          eval_points[4 * j + 3] = z[j][1].c1;
        }
      }
    } catch (...) {
      for (auto* u : uploaded) c.release(u);
      throw;
    }
    for (auto* u : uploaded) c.release(u);
  });
}

int p3g_prove(p3g_ctx* ctx, uint32_t hash_kind, const p3g_pcs_params* params, const p3g_matrix* traces, uint64_t k,
              uint32_t flags, p3g_transcript* ch, uint64_t root[4], uint64_t eval_points_out[4]) {
  if (!ctx || !params || !ch || (!traces && k)) return P3G_E_INVAL;
  // reject bad parameters before the commitment touches the transcript
  if (int rc0 = guard([&] { (void)pcs_params(params); })) return rc0;
  p3g_tree* tree = nullptr;
  int rc = p3g_commit_traces(ctx, hash_kind, traces, k, params->log_blowup, flags, &tree);
  if (rc != P3G_OK) return rc;
  uint64_t r[4], z[4];
  p3g_tree_root(tree, r);
  ch->tx.observe_commitment(r);
  ef2 z0 = ch->tx.sample_ef();
  ef2 z1 = ch->tx.sample_ef();
  z[0] = z0.c0;
  z[1] = z0.c1;
  z[2] = z1.c0;
  z[3] = z1.c1;
  uint32_t logL = log2_exact(p3g_tree_height(tree));
  rc = p3g_open_with_channel(ctx, params, logL, z, 2, &tree, 1, ch);
  p3g_tree_free(tree);
  if (root) memcpy(root, r, 32);
  if (eval_points_out) memcpy(eval_points_out, z, 32);
  return rc;
}

}  // extern "C"
