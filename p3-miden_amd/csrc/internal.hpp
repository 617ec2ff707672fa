// Internal declarations shared by the HIP library's translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <array>
#include <atomic>
#include <deque>
#include <functional>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "gl.hpp"
#include "hash.hpp"

#define P3G_CHECK_HIP(expr)                                  \
  do {                                                       \
    hipError_t e_ = (expr);                                  \
    if (e_ != hipSuccess) throw p3g::DeviceError(e_, #expr); \
  } while (0)

namespace p3g {

enum Status : int {
  OK = 0,
  E_INVAL = 1,
  E_NOT_POW2 = 2,
  E_UNSORTED = 3,
  E_NOMEM = 4,
  E_DEVICE = 5,
  E_EMPTY = 6,
  E_TRANSCRIPT = 7,
  E_RANGE = 8,
  E_CANCELED = 9,
};

struct DeviceError {
  hipError_t err;
  const char* what;
  DeviceError(hipError_t e, const char* w) : err(e), what(w) {}
};

struct ApiError {
  int code;
  std::string msg;
  ApiError(int c, std::string m) : code(c), msg(std::move(m)) {}
};

enum HashKind : uint32_t { HASH_P2 = 0, HASH_KECCAK = 1 };
enum ChallengerKind : uint32_t { CH_DUPLEX_P2 = 0, CH_KECCAK = 1 };

inline uint32_t hash_rate(uint32_t h) { return h == HASH_P2 ? 8 : 17; }
inline uint32_t hash_width(uint32_t h) { return h == HASH_P2 ? 12 : 25; }
// Alignable::ALIGNMENT: RATE for the field sponge; lcm(8, 8*17)/8 = 17 for the
// u64-serialising Keccak sponge (serializing_sponge.rs:166-203).
inline uint32_t hash_alignment(uint32_t h) { return h == HASH_P2 ? 8 : 17; }
inline uint64_t aligned_len(uint64_t n, uint64_t a) { return a <= 1 ? n : (n + a - 1) / a * a; }

// ------------------------------------------------------------------ shared block pool
// Device blocks shared by several contexts (one stream each).  A released
// block remembers the stream it was released on, so that it never changes
// hands while its previous owner's kernels still use it:
//  - taken again on that stream, stream order puts the new work after the old
//    (no HIP call; a lock-step batch releases and retakes ~1,700 small blocks);
//  - blocks of at least EXACT_EVENT_BYTES also get an event recorded at
//    release, which another stream waits for;
//  - a small block is taken on another stream only once its stream has
//    drained (hipStreamQuery); while that stream is busy a new block is
//    allocated instead, and only at the limit does the taker wait for an
//    event recorded then on the releasing stream (covering the block's last
//    uses and whatever that stream queued since).
// alloc prefers same-stream blocks, then drained or evented ones.  live +
// cached bytes stay within `limit`: beyond it cached blocks are reused even
// when oversized, and only then freed.
struct SharedPool {
  static constexpr size_t EXACT_EVENT_BYTES = size_t(4) << 20;
  static constexpr size_t LARGE_BLOCK_BYTES = size_t(64) << 20;
  struct Block {
    void* p;
    size_t sz;
    hipStream_t s;   // nullptr: no pending work (its stream was synchronized)
    hipEvent_t ev;   // recorded at release (large blocks), or nullptr
  };
  std::mutex mu;
  size_t limit = 0;
  std::multimap<size_t, Block> free_blocks;
  std::vector<hipEvent_t> spare_events;
  std::map<void*, size_t> live;
  size_t live_bytes = 0, cached_bytes = 0;
  uint64_t n_malloc = 0, n_free = 0, n_cross_stream = 0;
  double malloc_ms = 0, free_ms = 0;

  void* alloc(size_t bytes, hipStream_t s);
  void release(void* p, hipStream_t s);
  // the stream has been synchronized and is about to be destroyed
  void forget_stream(hipStream_t s);
  void trim();
  ~SharedPool();

 private:
  void* take(std::multimap<size_t, Block>::iterator it, hipStream_t s);
  std::multimap<size_t, Block>::iterator pick(std::multimap<size_t, Block>::iterator lo,
                                              std::multimap<size_t, Block>::iterator hi, hipStream_t s);
  void free_one(std::multimap<size_t, Block>::iterator it);
};

// ------------------------------------------------------------------ context
struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool owns_stream = true;  // false for a stream the caller created and keeps
  P2Consts host_p2{};
  P2Consts* dev_p2 = nullptr;
  std::mutex mu;
  // two-level power tables keyed by (base, exponent bits, lo multiplier)
  std::map<std::tuple<uint64_t, uint32_t, uint64_t>, uint64_t*> pow_tables;
  // size-bucketed cache of device allocations
  std::multimap<size_t, void*> free_blocks;
  std::map<void*, size_t> live_blocks;
  size_t cached_bytes = 0;
  size_t live_bytes = 0, peak_live_bytes = 0;
  // device allocator calls made for this context (hipMalloc / hipFree) and their host time
  uint64_t n_malloc = 0, n_free = 0;
  double malloc_ms = 0, free_ms = 0;

  // per-kernel timing with HIP events on this context's stream (off by default)
  bool profiling = false;
  std::map<std::string, std::vector<std::pair<hipEvent_t, hipEvent_t>>> prof;

  // when set, device blocks come from this pool (shared with other contexts)
  SharedPool* pool = nullptr;
  // Set by the prove-request server for the request it runs: a nonzero value
  // makes the proof stop at its next phase boundary with E_CANCELED.
  const std::atomic<int>* cancel = nullptr;
  void check_cancel() const {
    if (cancel && cancel->load(std::memory_order_relaxed)) throw ApiError(E_CANCELED, "request canceled");
  }
  // Fault injection for failure-handling tests: when >= 0, the allocation after
  // that many more fails with E_NOMEM (as if HBM were exhausted), then it resets to -1.
  int fail_alloc_after = -1;
  // pinned host staging for gathers, grown on demand; contents valid until the
  // next host_stage call on this context
  uint64_t* pinned = nullptr;
  size_t pinned_words = 0;
  uint64_t* host_stage(size_t words);

  // Pinned upload ring.  upload_async copies the caller's bytes into pinned
  // memory and enqueues the device copy from there, so the caller's buffer may
  // die at once and nothing waits for the stream.  A ring region is reused
  // only after the event recorded behind its copy has completed.
  struct UpSeg {
    size_t start, end;
    hipEvent_t ev;
  };
  uint8_t* up_ring = nullptr;
  size_t up_cap = 0, up_head = 0;
  std::deque<UpSeg> up_inflight;
  std::vector<hipEvent_t> up_spare;
  std::mutex up_mu;
  void upload_async(void* dst, const void* src, size_t bytes, hipStream_t s);
  // alloc + upload_async; the block is released by the caller as usual
  void* upload(const void* src, size_t bytes, hipStream_t s);
  void free_upload_ring();

  // Small device-to-host reads (roots, evaluations, final polynomials) go
  // through their own pinned buffer: the copy is a plain async DMA and the
  // round trip is shorter than through pageable memory.  Synchronises s.
  uint8_t* rb_pinned = nullptr;
  size_t rb_bytes = 0;
  void read_back(void* dst, const void* dsrc, size_t bytes, hipStream_t s);

  void* alloc(size_t bytes);
  void release(void* p);
  void trim();
  const uint64_t* pow_table(uint64_t base, uint32_t bits, uint64_t lo_mul = 1);
};

// Records a start/stop event pair around the launches in its scope when the
// context has profiling enabled.
struct ProfScope {
  Ctx& ctx;
  hipStream_t s;
  hipEvent_t a = nullptr, b = nullptr;
  const char* name;
  ProfScope(Ctx& c, const char* n, hipStream_t st) : ctx(c), s(st), name(n) {
    if (ctx.profiling) {
      (void)hipEventCreate(&a);
      (void)hipEventCreate(&b);
      (void)hipEventRecord(a, s);
    }
  }
  ~ProfScope() {
    if (a) {
      (void)hipEventRecord(b, s);
      std::lock_guard<std::mutex> g(ctx.mu);
      ctx.prof[name].emplace_back(a, b);
    }
  }
};

// Table layout: [lo: 2048 entries base^i * mul][hi: base^(2048 i)], so that
// base^e * mul = lo[e & 2047] * hi[e >> 11].
constexpr uint32_t POW_LO = 2048;
inline uint64_t pow_table_hi_len(uint32_t bits) { return bits > 11 ? (1ull << (bits - 11)) : 1; }

// ------------------------------------------------------------------ matrices and trees
struct DevMat {
  uint64_t* ptr = nullptr;
  uint64_t height = 0, width = 0;
  bool owned = false;
};

struct Tree {
  Ctx* ctx = nullptr;
  uint32_t hash = HASH_P2;
  uint32_t alignment = 1;
  std::vector<DevMat> mats;
  uint64_t height = 0;
  unsigned depth = 0;
  uint64_t* digests = nullptr;  // all layers, layer l at layer_off[l] (in 4-word digests)
  bool owns_digests = true;
  // hiding trees: per-leaf salt (height x salt width), absorbed after the rows
  // and hinted after them; ptr is null for a tree without salt
  DevMat salt;
  std::vector<uint64_t> layer_off;
  uint64_t root[4] = {0, 0, 0, 0};
  ~Tree();
};

// ------------------------------------------------------------------ transcript
// The prover-channel operations the PCS driver performs, in the reference's
// ProverChannel vocabulary (p3-miden-transcript/src/prover.rs:53-95).
struct Channel {
  virtual ~Channel() = default;
  virtual void send_fields(const uint64_t* v, size_t n) = 0;
  virtual void send_commitment(const uint64_t* c4) = 0;
  virtual void hint_fields(const uint64_t* v, size_t n) = 0;
  virtual void hint_commitment(const uint64_t* c4) = 0;
  virtual uint64_t grind(unsigned bits) = 0;
  virtual uint64_t sample() = 0;
  virtual uint64_t sample_bits(unsigned bits) = 0;
  ef2 sample_ef() {
    const uint64_t a = sample();
    const uint64_t b = sample();
    return ef2{a, b};
  }
};

struct Transcript : Channel {
  uint32_t kind = CH_DUPLEX_P2;
  uint64_t st[12] = {};
  uint64_t in[8] = {};
  int in_len = 0;
  uint64_t out[8] = {};
  int out_len = 0;
  std::vector<uint8_t> kin;
  uint8_t kout[32] = {};
  int kout_len = 0;
  std::vector<uint64_t> fields;
  std::vector<uint64_t> comms;  // 4 words each

  void observe(uint64_t v);
  void observe_commitment(const uint64_t* c4);
  uint64_t sample() override;
  uint64_t sample_bits(unsigned bits) override;
  void send_fields(const uint64_t* v, size_t n) override;
  void send_commitment(const uint64_t* c4) override;
  void hint_fields(const uint64_t* v, size_t n) override;
  void hint_commitment(const uint64_t* c4) override;
  uint64_t grind(unsigned bits) override;
  bool check_witness(unsigned bits, uint64_t w);
  // grind with a witness found elsewhere (the device search): observes and
  // checks it as grind would, and records it; throws if it fails the check
  void accept_witness(unsigned bits, uint64_t w);

 private:
  void duplex();
  void k_observe_u64(uint64_t v);
  uint8_t k_sample_byte();
};

const P2Consts& host_p2_consts();
void keccak256_host(const uint8_t* in, size_t len, uint8_t* out32);

// ------------------------------------------------------------------ PCS parameters
struct PcsParams {
  uint32_t log_blowup, log_arity, log_final_degree, folding_pow_bits, deep_pow_bits, num_queries, query_pow_bits;
};
// PcsParams::new's checks (p3-miden-lifted-fri/src/params.rs:54-88), plus this
// library's bound on grinding: proof-of-work bits below MAX_POW_BITS.
constexpr uint32_t MAX_POW_BITS = 32;
void check_pcs_params(const PcsParams& p);
size_t fri_num_rounds(const PcsParams& p, unsigned log_domain);
size_t fri_final_degree(const PcsParams& p, unsigned log_domain);

// ------------------------------------------------------------------ launchers
// Optional twist of the coefficients between the inverse and the forward
// transforms: coefficient k of column c is multiplied by root^(k * (c / group)),
// with root of order 2^log_order >= n * (w / group).
struct ColTwist {
  uint64_t root;
  uint32_t log_order;
  uint32_t group;
};
void lde_coset_batch(Ctx& ctx, const uint64_t* d_trace, uint64_t n, uint64_t w, unsigned log_blowup,
                     uint64_t shift, uint64_t* d_out, hipStream_t s, const ColTwist* twist = nullptr);
void bit_reverse_rows(Ctx& ctx, const uint64_t* src, uint64_t* dst, uint64_t h, uint64_t w, hipStream_t s);
// The same LDE (bit-reversed rows out) with its tables in a static per-GPU
// scratch array: no allocation, nothing waited for.  lde_scratch_init fetches
// the array's device address once per device (that device current).
void lde_scratch_init(int device);
uint64_t lde_scratch_words(unsigned logn, unsigned log_blowup);
uint64_t lde_scratch_capacity();
void lde_coset_batch_noalloc(Ctx& ctx, const uint64_t* d_trace, uint64_t n, uint64_t w, unsigned log_blowup,
                             uint64_t shift, uint64_t* d_out, hipStream_t s);
// one transform of a batch of same-shaped coset LDEs (bit-reversed rows out)
struct LdeJob {
  const uint64_t* src;
  uint64_t* dst;
  uint64_t n, w, shift;
};
void coset_lde_many(Ctx& ctx, const std::vector<LdeJob>& jobs, unsigned log_blowup, hipStream_t s);

// trees of k jobs with identical matrix shapes, built with shared launches; the
// trees take ownership of the matrices whose DevMat::owned is set
std::vector<Tree*> lmcs_build_many(Ctx& ctx, uint32_t hash, const std::vector<std::vector<DevMat>>& mats,
                                   uint32_t alignment, hipStream_t s, bool read_roots = true);
size_t prove_batch_pieces(const Tree& t, const std::vector<uint64_t>& indices,
                          std::vector<std::pair<const uint64_t*, uint64_t>>& pieces);
size_t prove_batch_emit(const Tree& t, const std::vector<std::pair<const uint64_t*, uint64_t>>& pieces, size_t first,
                        size_t last, size_t nrows, const uint64_t* raw, Channel& tx);
// read_root = false: no stream synchronisation; the root stays on the device
// salt (optional): a tree-height x SALT matrix for a hiding tree; the tree
// takes it over like the matrices (releasing it when owned)
Tree* lmcs_build(Ctx& ctx, uint32_t hash, const std::vector<DevMat>& mats, uint32_t alignment, hipStream_t s,
                 bool read_root = true, const DevMat* salt = nullptr);
void tree_rows(const Tree& t, uint64_t index, std::vector<uint64_t>& out);
void tree_prove_batch(const Tree& t, const std::vector<uint64_t>& indices, Channel& tx, hipStream_t s);
// gathers the pieces into the context's pinned staging buffer; the returned
// words are valid until the next gather on this context
const uint64_t* gather_pinned(Ctx& ctx, const std::vector<std::pair<const uint64_t*, uint64_t>>& pieces,
                              hipStream_t s);
void gather_to_host(Ctx& ctx, const std::vector<std::pair<const uint64_t*, uint64_t>>& pieces,
                    std::vector<uint64_t>& out, hipStream_t s);

void host_sponge_absorb(uint32_t hash, uint64_t* state, const uint64_t* in, size_t len);
void host_compress(uint32_t hash, const uint64_t* l, const uint64_t* r, uint64_t* out);

Tree* commit_traces(Ctx& ctx, uint32_t hash, const std::vector<DevMat>& traces, unsigned log_blowup, hipStream_t s);
// commit_quotient: q holds n*d extension elements (c0, c1), natural order on gJ,
// read as the n x 2d base matrix it already is
Tree* commit_quotient(Ctx& ctx, uint32_t hash, const uint64_t* d_q, uint64_t n, uint64_t d, unsigned log_blowup,
                      hipStream_t s);
// cyclic_extend_and_scale / divide_by_vanishing_in_place on device buffers
void quotient_extend_scale(Ctx& ctx, ef2* acc, uint64_t m, uint64_t target, ef2 beta, hipStream_t s);
void quotient_divide_vanishing(Ctx& ctx, ef2* num, uint64_t n, uint64_t d, uint64_t shift, hipStream_t s);

void open_with_channel(Ctx& ctx, const PcsParams& p, unsigned log_lde_height, const ef2* z, int npts,
                       const std::vector<Tree*>& trees, Channel& tx, hipStream_t s);

// The steps open_with_channel composes.  deep_batch_eval: evals[column][point]
// (host, unaligned widths in tree and matrix order); deep_quotient: the DEEP
// polynomial's L evaluations into the device buffer out from such evals and
// the two challenges; fri_fold: one round, n -> n >> la device elements;
// fri_final_poly: host, bit-reversed evaluations -> descending coefficients.
// the checks deep_batch_eval and deep_quotient make of their inputs, alone
void deep_check_inputs(const std::vector<Tree*>& trees, unsigned log_blowup, const ef2* z, int npts);
void deep_batch_eval(Ctx& ctx, unsigned log_blowup, const std::vector<Tree*>& trees, const ef2* z, int npts,
                     std::vector<ef2>& evals, hipStream_t s);
void deep_quotient(Ctx& ctx, unsigned log_blowup, const std::vector<Tree*>& trees, const ef2* z, int npts,
                   const ef2* evals, ef2 alpha, ef2 beta, ef2* out, hipStream_t s);
void fri_fold(Ctx& ctx, const ef2* in, uint64_t n, unsigned la, ef2 beta, ef2* out, hipStream_t s);
void fri_final_poly(const ef2* bitrev_evals, size_t n, ef2* coeffs_desc);

// f(0..n-1) on the library's host worker threads (serially when the pool is busy)
void parallel_for(size_t n, const std::function<void(size_t)>& f);

// Constraint evaluation on the quotient domain (evaluate_constraints_into): one
// straight-line program op, laid out as the C ABI's p3g_air_op
struct AirOpIn {
  uint32_t op, a, b, reserved;
  uint64_t imm;
};
void eval_constraints(Ctx& ctx, const AirOpIn* ops, uint64_t nops, const DevMat& main, const DevMat* aux,
                      const uint64_t* periodic, uint64_t prow, uint64_t pcols, const uint64_t* publics, uint64_t npub,
                      const ef2* ext_inputs, uint64_t next, uint64_t n, unsigned log_d, uint64_t shift, ef2 alpha,
                      ef2* acc, hipStream_t s);
// checks a program and compiles its kernel into the process cache; the code object's size
uint64_t air_compile(const AirOpIn* ops, uint64_t nops);

// k same-shaped proofs in lock-step (the p3g_prove flow per job)
void prove_many(Ctx& ctx, uint32_t hash, const PcsParams& p, const std::vector<std::vector<DevMat>>& traces,
                const std::vector<Transcript*>& txs, std::vector<std::array<uint64_t, 4>>& roots,
                std::vector<std::array<ef2, 2>>& zs, hipStream_t s);

}  // namespace p3g
