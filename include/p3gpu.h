/*
 * p3gpu.h — C ABI of the MI355X (gfx950) lifted-PCS library, libp3gpu.so.
 *
 * Drop-in boundary for the data-parallel core of the p3-miden lifted-STARK
 * prover over Goldilocks: coset LDE (TwoAdicSubgroupDft), LMCS commit/open
 * (Lmcs / LmcsTree), DEEP + FRI opening (open_with_channel) and the transcript
 * (ProverChannel).  Plain pointers and sizes only; every call returns a status
 * code (P3G_OK = 0) and never throws across the boundary.  Field elements are
 * canonical little-endian u64 Goldilocks values; extension elements are two
 * consecutive words (c0, c1) of F[x]/(x^2 - 7).  Citations are to files of
 * the reference repository (adr1anh/p3-miden).
 */
#ifndef P3GPU_H
#define P3GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define P3G_OK 0
#define P3G_E_INVAL 1      /* bad argument */
#define P3G_E_NOT_POW2 2   /* height not a power of two */
#define P3G_E_UNSORTED 3   /* matrices not sorted by ascending height */
#define P3G_E_NOMEM 4      /* device allocation failed */
#define P3G_E_DEVICE 5     /* HIP runtime error */
#define P3G_E_EMPTY 6      /* empty batch */
#define P3G_E_TRANSCRIPT 7 /* transcript misuse */
#define P3G_E_RANGE 8      /* index / size out of range */
#define P3G_E_CANCELED 9   /* prove request canceled (p3g_server_cancel) */

/* ---- hash and challenger kinds ---- */
#define P3G_HASH_POSEIDON2_GL12 0 /* StatefulSponge<Poseidon2Goldilocks<12>,12,8,4> + TruncatedPermutation */
#define P3G_HASH_KECCAK_GL 1      /* SerializingStatefulSponge<StatefulSponge<KeccakF,25,17,4>> + PaddingFreeSponge compress */
#define P3G_CHALLENGER_DUPLEX_P2 0 /* DuplexChallenger<Goldilocks, Poseidon2GL12, 12, 8> */
#define P3G_CHALLENGER_KECCAK 1    /* SerializingChallenger64<Goldilocks, HashChallenger<u8, Keccak256, 32>> */

/* ---- flags ---- */
#define P3G_DEVICE_PTRS 1u /* pointers are device (HBM) addresses */
#define P3G_BITREV_ROWS 2u /* coset LDE: permute output rows by bit reversal */
#define P3G_TAKE_OWNERSHIP 4u /* lmcs_build with device pointers: the tree frees them */

typedef struct p3g_ctx p3g_ctx;
typedef struct p3g_tree p3g_tree;
typedef struct p3g_transcript p3g_transcript;

/* Row-major height x width matrix of canonical Goldilocks words. */
typedef struct {
  const uint64_t* data;
  uint64_t height;
  uint64_t width;
} p3g_matrix;

/* PcsParams::new(log_blowup, log_folding_arity, log_final_degree, folding_pow_bits,
 * deep_pow_bits, num_queries, query_pow_bits)  (p3-miden-lifted-fri/src/params.rs:170-204) */
typedef struct {
  uint32_t log_blowup;
  uint32_t log_folding_arity;
  uint32_t log_final_degree;
  uint32_t folding_pow_bits;
  uint32_t deep_pow_bits;
  uint32_t num_queries;
  uint32_t query_pow_bits;
} p3g_pcs_params;

const char* p3g_status_string(int status);
const char* p3g_last_error(void); /* thread-local detail of the last failure */

/* ---- context (one per device; all work runs on its stream) ---- */
int p3g_ctx_create(int device, p3g_ctx** out);
/* the same on a stream the caller created (a hipStream_t of that device from the HIP
 * runtime this library links, not the null stream; no reference counterpart): the
 * library's work is ordered with the caller's own work on it, and p3g_ctx_destroy
 * synchronizes the stream but leaves it alive */
int p3g_ctx_create_on_stream(int device, void* stream, p3g_ctx** out);
/* frees all device memory of the context: free its trees and plans first (they refer to it) */
void p3g_ctx_destroy(p3g_ctx* ctx);
int p3g_ctx_synchronize(p3g_ctx* ctx);
int p3g_ctx_trim(p3g_ctx* ctx); /* release cached device blocks */
/* device bytes held by live allocations, by the block cache, and the peak of live bytes
 * (optionally reset to the current value) */
int p3g_ctx_memory(p3g_ctx* ctx, uint64_t* live, uint64_t* cached, uint64_t* peak_live, int reset_peak);

/* ---- per-kernel timing (instrumentation; no reference counterpart) ----
 * When enabled, HIP events are recorded on the context's stream around each
 * launch of a named kernel family ("lde_inverse_pass", "lde_forward_pass",
 * "p2_leaf", "keccak_leaf", "compress_layer", "compress_top", "deep_weights",
 * "deep_colsum", "deep_quotient", "fri_fold", "grind"). */
int p3g_profile_enable(p3g_ctx* ctx, int on);
int p3g_profile_reset(p3g_ctx* ctx);
int p3g_profile_read(p3g_ctx* ctx, const char* name, double* total_ms, uint64_t* launches);

/* ---- peak calibration (instrumentation; no reference counterpart) ----
 * The measured ceilings the roofline figures are quoted against.
 * calibrate_valu: one launch of blocks x 256 threads in which every lane runs
 *   independent chains of one VALU instruction (kind 0: v_add_u32, 1: v_mad_u64_u32,
 *   2: v_fma_f64, 3: v_mul_hi_u32, 4: a 64-bit add on a carry pair, v_add_co_u32 +
 *   v_addc_co_u32, 5: v_mov_b32, 6: v_cndmask_b32, 7: v_add_co_u32 with a carry-out SGPR
 *   pair, 8: v_lshl_add_u64, 9: v_add3_u32, 10: v_xor_b32, 11: v_alignbit_b32,
 *   12: v_cndmask_b32_e32 (VOP2, mask in vcc), 13: v_mul_u32_u24_e32, and two mixed
 *   streams: 14: v_mad_u64_u32 alternating with v_add_u32_e32, 15: v_mad_u64_u32
 *   alternating with v_cndmask_b32_e32, each counting both) for iters
 *   iterations (a positive multiple of 4); *ms is the launch
 *   time (HIP events on the context's stream), *insts_per_wave the instructions of
 *   that class each wave issues.
 * calibrate_copy: a device-to-device copy of bytes (a multiple of 16) with a
 *   16-bytes-per-lane copy kernel and with hipMemcpyAsync, reps times each after a
 *   warm-up; each *ms is the average time of one copy. */
int p3g_calibrate_valu(p3g_ctx* ctx, uint32_t kind, uint32_t iters, uint32_t blocks, double* ms,
                       uint64_t* insts_per_wave);
int p3g_calibrate_copy(p3g_ctx* ctx, uint64_t bytes, uint32_t reps, double* kernel_ms, double* memcpy_ms);

/* ---- device field operations on caller operands (instrumentation; no reference counterpart) ----
 * For tests of the kernels' arithmetic on chosen operands, including the carry
 * and borrow cases random data does not reach.  n operands (1..2^20) in host
 * arrays a, b (and c, or NULL for zeros), one wave per 64; out[i] receives:
 *   op 0: a*b (mod p, some value < 2^64) as the products of the proof's kernels
 *         compute it in full;
 *   op 1: the same with the hashing kernels' flagged shortcut: each wave's
 *         64-bit lane mask of operands whose product took the rare borrow goes
 *         to wave_flags[i / 64] (those lanes' out is a*b + (2^32 - 1) mod p);
 *   op 2, 3: a*b + c, as op 0 and op 1;
 *   op 4: a*b canonical (< p), for canonical a, b;
 *   op 5: a * 2^e canonical, for canonical a and 0 <= e < 96;
 *   op 6: (a + 2^32 b) mod p, some value < 2^64, for a, b < 2^40;
 *   op 7, 8: a + b and a - b canonical, for canonical a, b (the NTT passes'
 *         butterflies);
 *   op 9: a mod p for any 64-bit a.
 * wave_flags (ceil(n / 64) words, or NULL) is written for every op (zero
 * except for ops 1 and 3). */
int p3g_field_check(p3g_ctx* ctx, uint32_t op, uint32_t e, const uint64_t* a, const uint64_t* b, const uint64_t* c,
                    uint64_t n, uint64_t* out, uint64_t* wave_flags);

/* ---- device memory (resident inputs, benchmarks, serving) ---- */
int p3g_alloc(p3g_ctx* ctx, uint64_t bytes, void** dptr);
int p3g_free(p3g_ctx* ctx, void* dptr);
int p3g_upload(p3g_ctx* ctx, void* dst, const void* src, uint64_t bytes);
int p3g_download(p3g_ctx* ctx, void* dst, const void* src, uint64_t bytes);
int p3g_copy_device(p3g_ctx* ctx, void* dst, const void* src, uint64_t bytes); /* device to device */
/* RowMajorMatrix::rand(SmallRng::seed_from_u64(seed), n/w, w) as n canonical words, on the host. */
int p3g_fill_random(uint64_t seed, uint64_t n, uint64_t* host_out);

/* ---- TwoAdicSubgroupDft ----
 * coset_lde_batch(mat, added_bits, shift) (+ bit_reverse_rows with P3G_BITREV_ROWS);
 * replaces the call in p3-miden-lifted-stark/src/prover/commit.rs:217-221.
 * out holds (n << added_bits) x w words.  Without P3G_BITREV_ROWS the natural
 * order is produced on the host side of the copy (host pointers only). */
int p3g_coset_lde_batch(p3g_ctx* ctx, const uint64_t* trace, uint64_t n, uint64_t w, uint32_t added_bits,
                        uint64_t shift, uint32_t flags, uint64_t* out);

/* coset_lde_batch + bit_reverse_rows on device buffers, allocating nothing.  No counterpart
 * in the reference: it is for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The same values as
 * p3g_coset_lde_batch with P3G_DEVICE_PTRS | P3G_BITREV_ROWS: d_trace (n x w) and d_out
 * ((n << added_bits) x w) are device addresses.  The twiddle and coset tables live in one
 * static 768 KiB array per GPU, held until the process ends; every call on a GPU shares it,
 * so calls that overlap on one GPU race.  The work is enqueued on the context's stream and
 * the call returns without waiting (p3g_ctx_synchronize waits).  Shapes whose tables need
 * more than the array -- 2 (2048 + 2^(log n - 12)) + 2^added_bits (2048 + 2^(log n - 11))
 * words > 98,304 -- or added_bits > 3 give P3G_E_RANGE; every trace up to 2^24 rows with
 * added_bits <= 3 fits. */
int p3g_coset_lde_batch_noalloc(p3g_ctx* ctx, const uint64_t* d_trace, uint64_t n, uint64_t w, uint32_t added_bits,
                                uint64_t shift, uint64_t* d_out);

/* ---- Lmcs ----
 * LmcsConfig::build_tree (alignment = 1) / build_aligned_tree (alignment = p3g_lmcs_alignment)
 * (p3-miden-lmcs/src/lmcs.rs:102-130, lifted_tree.rs:244-310).  Matrices sorted by
 * ascending power-of-two height.  Host matrices are copied to the device; device
 * matrices are referenced (and freed by the tree with P3G_TAKE_OWNERSHIP). */
uint32_t p3g_lmcs_alignment(uint32_t hash_kind);
int p3g_lmcs_build(p3g_ctx* ctx, uint32_t hash_kind, const p3g_matrix* mats, uint64_t k, uint32_t alignment,
                   uint32_t flags, p3g_tree** out);
/* Lmcs::hash / Lmcs::compress (lib.rs:203-209), host side. */
/* Hiding LMCS (HidingLmcsConfig::build_tree / build_aligned_tree, p3-miden-lmcs/src/hiding_lmcs.rs:104-140;
 * salt absorption lifted_tree.rs:271-279): as p3g_lmcs_build, plus a salt matrix of the tree's height and
 * SALT columns that the caller draws from its RNG.  Each leaf absorbs its salt row after its rows; the
 * batch opening hints the salt (unpadded) after each leaf's rows (lifted_tree.rs:187-192).  flags apply
 * to the salt as to the matrices. */
int p3g_lmcs_build_salted(p3g_ctx* ctx, uint32_t hash_kind, const p3g_matrix* mats, uint64_t k, uint32_t alignment,
                          uint32_t flags, const p3g_matrix* salt, p3g_tree** out);
int p3g_lmcs_hash(uint32_t hash_kind, const uint64_t* elems, const uint64_t* row_widths, uint64_t nrows,
                  uint64_t out[4]);
int p3g_lmcs_compress(uint32_t hash_kind, const uint64_t left[4], const uint64_t right[4], uint64_t out[4]);

/* ---- LmcsTree (lib.rs:262-295) ---- */
int p3g_tree_root(const p3g_tree* t, uint64_t out[4]);
uint64_t p3g_tree_height(const p3g_tree* t);
uint32_t p3g_tree_alignment(const p3g_tree* t);
uint64_t p3g_tree_num_matrices(const p3g_tree* t);
/* widths(): aligned widths, one per matrix */
int p3g_tree_widths(const p3g_tree* t, uint64_t* out);
/* rows(index): upsampled rows padded to alignment (lifted_tree.rs:127-138); *len receives the count */
int p3g_tree_rows(p3g_tree* t, uint64_t index, uint64_t* out, uint64_t cap, uint64_t* len);
/* digest layer l (0 = leaves): (height >> l) digests of 4 words */
/* LiftedMerkleTree::salt(index) (lifted_tree.rs:340-360): len = SALT, 0 for a tree without salt */
int p3g_tree_salt(const p3g_tree* t, uint64_t index, uint64_t* out, uint64_t cap, uint64_t* len);
int p3g_tree_layer(p3g_tree* t, uint32_t layer, uint64_t* out);
/* device address of committed matrix i (the bit-reversed LDE for commit_traces trees) */
int p3g_tree_matrix(p3g_tree* t, uint64_t i, const uint64_t** dptr, uint64_t* height, uint64_t* width);
/* prove_batch(indices, channel) (lifted_tree.rs:157-224) */
int p3g_tree_prove_batch(p3g_tree* t, const uint64_t* indices, uint64_t n, p3g_transcript* ch);
void p3g_tree_free(p3g_tree* t);

/* ---- commit_traces (p3-miden-lifted-stark/src/prover/commit.rs:175-228) ----
 * Per trace: coset LDE on the nested shift GENERATOR^(2^(log n_max - log n_j)),
 * bit-reversed rows, then one aligned LMCS tree over all LDEs. */
int p3g_commit_traces(p3g_ctx* ctx, uint32_t hash_kind, const p3g_matrix* traces, uint64_t k, uint32_t log_blowup,
                      uint32_t flags, p3g_tree** out);

/* ---- commit_quotient (p3-miden-lifted-stark/src/prover/quotient.rs:133-238) ----
 * q_evals: the quotient's n*d extension elements in natural order on gJ, each as
 * (c0, c1) -- the n x 2d base matrix, row r = [Q(g w_J^(r d + t))]_t.  Commits the
 * bit-reversed (n*B) x 2d LDE of the chunk polynomials on gK in one aligned tree.
 * n and d powers of two, d <= 2^log_blowup (P3G_E_INVAL otherwise).
 * flags: P3G_DEVICE_PTRS when q_evals is a device address. */
int p3g_commit_quotient(p3g_ctx* ctx, uint32_t hash_kind, const uint64_t* q_evals, uint64_t n, uint64_t d,
                        uint32_t log_blowup, uint32_t flags, p3g_tree** out);

/* ---- quotient accumulation (p3-miden-lifted-stark/src/prover/quotient.rs:28-101) ----
 * On device buffers of extension elements (c0, c1 words), on the context's stream.
 * extend_scale: cyclic_extend_and_scale in place on a buffer of `target` elements whose
 *   first m hold the accumulator (m == 0 zero-fills; else scale by beta, repeat to target).
 *   m and target powers of two, m <= target.
 * divide_vanishing: divide_by_vanishing_in_place on the n*d quotient values in natural
 *   order on the coset shift*J (|J| = n*d; shift = LiftedCoset::lde_shift()).
 *   P3G_E_INVAL when the vanishing polynomial is zero there (shift^n a d-th root of unity). */
int p3g_quotient_extend_scale(p3g_ctx* ctx, uint64_t* acc, uint64_t m, uint64_t target, const uint64_t beta[2]);
int p3g_quotient_divide_vanishing(p3g_ctx* ctx, uint64_t* num, uint64_t n, uint64_t d, uint64_t shift);

/* ---- constraint evaluation on the quotient domain (evaluate_constraints_into,
 * p3-miden-lifted-stark/src/prover/constraints/mod.rs:63-179) ----
 * An AIR's constraints as a straight-line program: what a symbolic AIR builder records
 * when the AIR's eval runs (p3-air's SymbolicAirBuilder).  Op i's value is slot i; operands
 * name earlier slots.  Values are extension elements (base values have c1 = 0).
 *   CONST imm | MAIN col=a, row=b (0 current, 1 next) | AUX ext column a, row b |
 *   PERIODIC col a | PUBLIC index a | EXT_INPUT index a (extension inputs: the permutation
 *   randomness, then the permutation values) | IS_FIRST_ROW | IS_LAST_ROW | IS_TRANSITION |
 *   ADD a b | SUB a b | MUL a b | NEG a | ASSERT_ZERO a | ASSERT_ZERO_EXT a
 * The constraints are the ASSERT ops in program order, C_0 .. C_{K-1}; the folded value is
 * sum_k alpha^(K-1-k) C_k (constraints/mod.rs:53-58). */
#define P3G_AIR_CONST 0u
#define P3G_AIR_MAIN 1u
#define P3G_AIR_AUX 2u
#define P3G_AIR_PERIODIC 3u
#define P3G_AIR_PUBLIC 4u
#define P3G_AIR_EXT_INPUT 5u
#define P3G_AIR_IS_FIRST_ROW 6u
#define P3G_AIR_IS_LAST_ROW 7u
#define P3G_AIR_IS_TRANSITION 8u
#define P3G_AIR_ADD 9u
#define P3G_AIR_SUB 10u
#define P3G_AIR_MUL 11u
#define P3G_AIR_NEG 12u
#define P3G_AIR_ASSERT_ZERO 13u
#define P3G_AIR_ASSERT_ZERO_EXT 14u
typedef struct {
  uint32_t op;
  uint32_t a, b;
  uint32_t reserved; /* 0 */
  uint64_t imm;      /* P3G_AIR_CONST: a canonical value */
} p3g_air_op;
/* accumulator[i] += sum_k alpha^(K-1-k) C_k(x_i) at every point x_i = shift w^i of the
 * quotient domain gJ (natural order i < n d, |gJ| = n d, d = 2^log_d), on the context's stream.
 *  main, aux: device matrices holding a committed LDE in bit-reversed row order, at least n d
 *    rows (the first n d rows are gJ, as Committed::evals_on_quotient_domain views them,
 *    commit.rs:134-144); the next row of point i is point (i + d) mod n d.  aux holds 2 base
 *    words per extension column (c0, c1); NULL when the AIR has none.
 *  periodic: host table of the periodic columns' LDE on gJ in natural order (PeriodicLde,
 *    periodic.rs:29-99), rows x columns, read at row i mod rows; NULL when none.
 *  publics, ext_inputs: host values (ext_inputs as c0, c1 pairs).
 *  Selectors are the reference's unnormalised ones at x: Z_H(x)/(x - 1), Z_H(x)/(x - w_n^-1)
 *    and x - w_n^-1 with Z_H(x) = x^n - 1 (coset.rs:210-247).
 *  shift: outside the subgroup of order n d (P3G_E_INVAL otherwise: Z_H would vanish).
 *  accumulator: device, n d extension elements.
 * Each program is compiled once into its own gfx950 kernel (values typed base or extension,
 * kept in registers) and cached for the process; a malformed program or an index out of
 * range gives P3G_E_INVAL, an empty or assertion-free one P3G_E_EMPTY, more than 2^20 ops,
 * log_d > 20 or n d > 2^31 P3G_E_RANGE, all before any launch. */
int p3g_eval_constraints(p3g_ctx* ctx, const p3g_air_op* ops, uint64_t nops, const p3g_matrix* main,
                         const p3g_matrix* aux, const uint64_t* periodic, uint64_t periodic_rows,
                         uint64_t periodic_cols, const uint64_t* publics, uint64_t num_publics,
                         const uint64_t* ext_inputs, uint64_t num_ext_inputs, uint64_t n, uint32_t log_d,
                         uint64_t shift, const uint64_t alpha[2], uint64_t* accumulator);
/* Check a constraint program and compile its kernel ahead of the first p3g_eval_constraints
 * (the first call otherwise pays the compile, ~0.3-2 s); needs no device.  *code_bytes is
 * the size of the gfx950 code object.  No counterpart in the reference (its AIRs are
 * compiled with the prover). */
int p3g_air_compile(const p3g_air_op* ops, uint64_t nops, uint64_t* code_bytes);

/* ---- transcript (ProverChannel, p3-miden-transcript/src/prover.rs:53-145) ---- */
int p3g_transcript_create(uint32_t challenger_kind, p3g_transcript** out);
void p3g_transcript_free(p3g_transcript* ch);
int p3g_transcript_clone(const p3g_transcript* ch, p3g_transcript** out);
int p3g_transcript_observe(p3g_transcript* ch, const uint64_t* values, uint64_t n);       /* CanObserve<F> */
int p3g_transcript_observe_commitment(p3g_transcript* ch, const uint64_t c[4]);          /* CanObserve<Hash> */
int p3g_transcript_sample(p3g_transcript* ch, uint64_t* out);                             /* sample F */
int p3g_transcript_sample_ext(p3g_transcript* ch, uint64_t out[2]);                       /* sample_algebra_element */
int p3g_transcript_sample_bits(p3g_transcript* ch, uint32_t bits, uint64_t* out);
int p3g_transcript_send_fields(p3g_transcript* ch, const uint64_t* values, uint64_t n);
int p3g_transcript_send_commitment(p3g_transcript* ch, const uint64_t c[4]);
int p3g_transcript_hint_fields(p3g_transcript* ch, const uint64_t* values, uint64_t n);
int p3g_transcript_hint_commitment(p3g_transcript* ch, const uint64_t c[4]);
int p3g_transcript_grind(p3g_transcript* ch, uint32_t bits, uint64_t* witness);
/* TranscriptData { fields, commitments } (data.rs:11-14) */
uint64_t p3g_transcript_num_fields(const p3g_transcript* ch);
uint64_t p3g_transcript_num_commitments(const p3g_transcript* ch);
int p3g_transcript_data(const p3g_transcript* ch, uint64_t* fields, uint64_t* commitments);

/* ---- proof wire format (TranscriptData, data.rs:8-14, has serde only; this byte
 * encoding is the library's own) ----
 *   bytes 0..3   "P3GT"          bytes 4..7   version 1 (u32 LE)
 *   bytes 8..15  nf (u64 LE)     bytes 16..23 nc (u64 LE)
 *   then nf field words and nc commitments of 4 words, each a canonical u64 LE:
 *   24 + 8 nf + 32 nc bytes in all.
 * encode: out == NULL only reports *len; a buffer shorter than *len gives P3G_E_RANGE;
 * a non-canonical word gives P3G_E_INVAL.  decode: fields / commitments == NULL only
 * report the counts; a bad header, a length that does not match, or a non-canonical
 * word gives P3G_E_INVAL. */
int p3g_proof_encode(const uint64_t* fields, uint64_t nf, const uint64_t* commitments, uint64_t nc, uint8_t* out,
                     uint64_t cap, uint64_t* len);
int p3g_proof_decode(const uint8_t* in, uint64_t len, uint64_t* fields, uint64_t* nf, uint64_t* commitments,
                     uint64_t* nc);
int p3g_transcript_encode(const p3g_transcript* ch, uint8_t* out, uint64_t cap, uint64_t* len);

/* ---- PCS opening ----
 * open_with_channel(params, lmcs, log_lde_height, eval_points, trace_trees, channel)
 * (p3-miden-lifted-fri/src/prover.rs:35-116): DEEP quotient, FRI commit phase,
 * query grinding and sampling, batch openings of every trace tree and FRI round.
 * eval_points: n_points extension elements (2 words each), 1 <= n_points <= 4.
 * Grinding on this library's Poseidon2 duplex transcript searches witnesses on the
 * GPU (every candidate below the returned witness is tested, so it is the serial
 * search's witness); the parameters are checked as PcsParams::new does, and
 * proof-of-work bits of 32 or more give P3G_E_INVAL (here and in p3g_prove,
 * p3g_prove_many, p3g_prove_planned and p3g_server_submit). */
int p3g_open_with_channel(p3g_ctx* ctx, const p3g_pcs_params* params, uint32_t log_lde_height,
                          const uint64_t* eval_points, uint32_t n_points, p3g_tree* const* trees, uint64_t ntrees,
                          p3g_transcript* ch);

/* ---- DEEP and FRI steps ----
 * The pieces open_with_channel composes, for a caller that runs the protocol
 * around them itself (its own channel order, grinding or query phase).  Each
 * runs on the context's stream; device buffers hold extension elements as
 * (c0, c1) words.  trees: one LDE height, one alignment (as open_with_channel).
 *
 * deep_batch_eval: PointQuotients::batch_eval_lifted (p3-miden-lifted-fri/src/
 *   deep/interpolate.rs:123-201) of every committed column at the n_points
 *   points, into host evals[column][point] (unaligned widths, tree then matrix
 *   order): 2 * n_points * columns words, reported in *len; evals == NULL only
 *   reports *len; cap < *len gives P3G_E_RANGE.  A point on the LDE coset gives
 *   P3G_E_INVAL.
 * deep_quotient: DeepPoly::from_evals after its challenges (deep/prover.rs:
 *   106-285) -- pads evals to the aligned widths, f_red(z_j) by Horner in alpha,
 *   column coefficients -alpha^(W-1-i), and writes the DEEP polynomial's L
 *   evaluations (bit-reversed LDE order) to the device buffer out (2L words).
 *   evals / len as deep_batch_eval returns them.
 * fri_fold: FriFold::fold_matrix (fri/fold/mod.rs:119-144) for one round of
 *   FriPolys::new (fri/prover.rs:90-223): the n bit-reversed evaluations at in
 *   (device) fold to n >> log_arity at out (device, not overlapping in), with
 *   s_k^-1 = w_n^-bitrev(k) for row k.  log_arity 1..3.
 * fri_commit: the round's commitment ahead of its fold (fri/prover.rs:136-150):
 *   the unaligned tree over the n device evaluations read as the
 *   (n >> log_arity) x (2 << log_arity) base matrix (p3g_lmcs_build, alignment 1,
 *   of that view).  The tree refers to evals, which must outlive it; root
 *   (optional) receives its root.
 * fri_final_poly: the final polynomial (fri/prover.rs:208-221) from the first n
 *   evaluations (host, bit-reversed order) of the last domain to n coefficients
 *   in descending degree (host).  n a power of two up to 2^24. */
int p3g_deep_batch_eval(p3g_ctx* ctx, uint32_t log_blowup, const uint64_t* eval_points, uint32_t n_points,
                        p3g_tree* const* trees, uint64_t ntrees, uint64_t* evals, uint64_t cap, uint64_t* len);
int p3g_deep_quotient(p3g_ctx* ctx, uint32_t log_blowup, const uint64_t* eval_points, uint32_t n_points,
                      p3g_tree* const* trees, uint64_t ntrees, const uint64_t* evals, uint64_t len,
                      const uint64_t alpha[2], const uint64_t beta[2], uint64_t* out);
int p3g_fri_fold(p3g_ctx* ctx, const uint64_t* in, uint64_t n, uint32_t log_arity, const uint64_t beta[2],
                 uint64_t* out);
int p3g_fri_commit(p3g_ctx* ctx, uint32_t hash_kind, const uint64_t* evals, uint64_t n, uint32_t log_arity,
                   p3g_tree** out, uint64_t root[4]);
int p3g_fri_final_poly(const uint64_t* evals, uint64_t n, uint64_t* coeffs);

/* ---- caller-owned channel ----
 * The same operations driven through the caller's ProverChannel (for example a
 * Rust ProverTranscript over its own DuplexChallenger): the library calls back
 * for every send / hint / grind / sample, in exactly the order the reference's
 * open_with_channel and LmcsTree::prove_batch perform them.  Callbacks return 0
 * on success; any other value aborts the call with P3G_E_TRANSCRIPT. */
typedef struct {
  void* user;
  int (*send_fields)(void* user, const uint64_t* values, uint64_t n);  /* send_field_slice (observed) */
  int (*send_commitment)(void* user, const uint64_t c[4]);            /* send_commitment (observed) */
  int (*hint_fields)(void* user, const uint64_t* values, uint64_t n);  /* hint_field_slice */
  int (*hint_commitment)(void* user, const uint64_t c[4]);            /* hint_commitment */
  int (*grind)(void* user, uint32_t bits, uint64_t* witness);          /* grind */
  int (*sample)(void* user, uint64_t* out);                           /* Channel::sample */
  int (*sample_bits)(void* user, uint32_t bits, uint64_t* out);       /* Channel::sample_bits */
} p3g_channel;

/* prove_batch(indices, channel) (lifted_tree.rs:157-224) into a caller channel */
int p3g_tree_prove_batch_cb(p3g_tree* t, const uint64_t* indices, uint64_t n, const p3g_channel* ch);
/* open_with_channel (p3-miden-lifted-fri/src/prover.rs:35-116) into a caller channel */
int p3g_open_with_callbacks(p3g_ctx* ctx, const p3g_pcs_params* params, uint32_t log_lde_height,
                            const uint64_t* eval_points, uint32_t n_points, p3g_tree* const* trees, uint64_t ntrees,
                            const p3g_channel* ch);

/* ---- batched commit plans (no reference counterpart: the planned, graph-captured
 * form of commit_traces for many traces of uneven heights) ----
 * A plan fixes the shapes (heights ascending, powers of two), owns the device
 * input buffers (fill them through p3g_commit_plan_input), and on every run
 * recomputes the LDEs and the lifted LMCS tree exactly as p3g_commit_traces
 * would.  Every NTT pass index is one grouped launch over all traces; with
 * P3G_PLAN_GRAPH the launch sequence is captured once into a hipGraph and
 * replayed; with P3G_PLAN_AUTOTUNE the NTT stage grouping and the leaf-hash
 * workgroup size are picked by timing candidates on this shape.  The tree
 * returned by run is a view valid until the next run or destroy. */
#define P3G_PLAN_GRAPH 1u
#define P3G_PLAN_AUTOTUNE 2u
#define P3G_PLAN_PER_TRACE 4u /* without autotuning: one NTT launch per trace and pass */
typedef struct p3g_commit_plan p3g_commit_plan;
int p3g_commit_plan_create(p3g_ctx* ctx, uint32_t hash_kind, const uint64_t* heights, const uint64_t* widths,
                           uint64_t k, uint32_t log_blowup, uint32_t flags, p3g_commit_plan** out);
int p3g_commit_plan_input(p3g_commit_plan* plan, uint64_t i, uint64_t** dptr);
/* the launch shapes in use: NTT stages per pass, leaf workgroup size, launch count, and whether NTT passes
 * are grouped over traces (1) or launched per trace (0) */
int p3g_commit_plan_info(const p3g_commit_plan* plan, uint32_t* kmax, uint32_t* leaf_threads, uint32_t* launches,
                         uint32_t* grouped);
int p3g_commit_plan_run(p3g_commit_plan* plan, p3g_tree** view);
void p3g_commit_plan_destroy(p3g_commit_plan* plan);
/* run the plan, observe the root, sample two points, open (the pcs_trace flow of p3g_prove) */
int p3g_prove_planned(p3g_ctx* ctx, p3g_commit_plan* plan, const p3g_pcs_params* params, p3g_transcript* ch,
                      uint64_t root[4], uint64_t eval_points_out[4]);
/* synthetic canonical words generated on the device (counter-based SplitMix64; not the SmallRng stream) */
int p3g_fill_random_device(p3g_ctx* ctx, uint64_t seed, uint64_t n, uint64_t* dptr);

/* ---- prove-request server (no reference counterpart: continuous batching of
 * independent proofs on one GPU) ----
 * W worker contexts (own stream and allocation cache each) pull requests in
 * arrival order; a request is admitted when a worker is free and its device
 * footprint (p3g_request_footprint) fits in the HBM budget beside the requests
 * in flight.  The workers share one device block pool; the budget is clamped to
 * what the device has free at creation, less 4 GiB.  A worker that admits a
 * request also takes the later queued requests with the same hash, parameters
 * and trace shapes (up to max_batch, within the budget) and proves them
 * together with p3g_prove_many.  A request that does not fit lets later ones that do go first
 * until it has waited backfill_ms.  Each request runs the p3g_prove flow on a
 * fresh transcript of its hash's challenger kind; its result is independent of
 * the other requests.  Trace buffers belong to the caller and must stay valid
 * until the request is done (pinned host memory from p3g_host_alloc makes the
 * copies asynchronous-speed; device pointers with P3G_DEVICE_PTRS skip them). */
typedef struct p3g_server p3g_server;
typedef struct {
  uint32_t hash_kind;
  p3g_pcs_params params;
  const p3g_matrix* traces; /* ascending power-of-two heights */
  uint64_t k;
  uint32_t flags; /* P3G_DEVICE_PTRS, P3G_INJECT_ALLOC_FAILURE */
} p3g_request;
/* Failure-handling test hook (request flag): the request's third device allocation fails
 * with P3G_E_NOMEM, as if HBM ran out in mid-proof.  In a lock-step batch this fails the
 * batch, whose members are then proven alone, so that only this request reports the error. */
#define P3G_INJECT_ALLOC_FAILURE 8u
typedef struct {
  int status;
  uint64_t root[4];
  uint64_t eval_points[4];
  uint64_t num_fields;
  uint64_t num_commitments;
  double submit_ms, admit_ms, done_ms; /* since server creation */
  uint32_t worker;
  uint64_t footprint_bytes;
} p3g_result_info;
typedef struct {
  uint64_t admitted, completed, queued, running;
  uint64_t batches, batched_requests; /* lock-step batches run, and the requests they held */
  uint64_t inflight_bytes, peak_inflight_bytes, peak_running, budget_bytes;
  uint64_t n_malloc, n_free; /* device allocator calls of the shared pool */
  double malloc_ms, free_ms; /* host time spent in them */
} p3g_server_counters;
uint64_t p3g_request_footprint(const p3g_request* req); /* device bytes reserved at admission */
/* max_batch: requests of one shape signature proven together in lock-step (1 disables batching) */
int p3g_server_create(int device, uint32_t workers, uint64_t hbm_budget_bytes, double backfill_ms, uint32_t max_batch,
                      p3g_server** out);
int p3g_server_submit(p3g_server* s, const p3g_request* req, uint64_t* ticket); /* P3G_E_RANGE: above the budget */
/* Batch window (default 0): a worker that picks a request with fewer than max_batch
 * queued requests of its shape holds it until it has been queued batch_window_ms, or
 * the batch fills, so that requests arriving together are proven together. */
int p3g_server_set_batch_window(p3g_server* s, double batch_window_ms);
/* Upload ahead: a stage with a stream of its own copies the host traces of up to `requests`
 * queued requests (0 = off, the default; at most 16) into device memory, in arrival order,
 * while the workers prove earlier ones, so that PCIe copies overlap proving.  A request's
 * whole footprint is reserved in the budget when its copies start, and a worker then proves
 * from the copies; results are unchanged.  No counterpart in the reference (its prover
 * reads host matrices); for callers that hand over host traces. */
int p3g_server_set_upload_ahead(p3g_server* s, uint32_t requests);
int p3g_server_wait(p3g_server* s, uint64_t ticket, p3g_result_info* info);     /* blocks; returns the proof status */
int p3g_server_poll(p3g_server* s, uint64_t ticket, int* done);
int p3g_server_proof(p3g_server* s, uint64_t ticket, uint64_t* fields, uint64_t* commitments);
int p3g_server_release(p3g_server* s, uint64_t ticket);
/* Cancels a request: a queued one completes at once with P3G_E_CANCELED (one whose copies
 * the upload stage is enqueuing, as soon as they are enqueued); a running one
 * stops at its next phase boundary (after its commitment, its DEEP quotient, each FRI
 * round, before its queries) and completes with P3G_E_CANCELED, its device memory
 * released.  A member of a lock-step batch does not stop its batch-mates: its result is
 * dropped when the batch ends.  A request already done keeps its result (P3G_OK either way). */
int p3g_server_cancel(p3g_server* s, uint64_t ticket);
int p3g_server_stats(p3g_server* s, p3g_server_counters* out);
int p3g_server_trim(p3g_server* s);     /* frees the pool's cached (unused) device blocks */
void p3g_server_destroy(p3g_server* s); /* finishes running requests, drops queued ones */
int p3g_host_alloc(uint64_t bytes, void** out); /* pinned host memory */
int p3g_host_free(void* p);

/* ---- batched proofs (no reference counterpart: many independent proofs in lock-step) ----
 * k jobs of t traces each; traces[j*t + i] is trace i of job j, and every job
 * has the same heights and widths.  Each job is the p3g_prove flow on its own
 * transcript chs[j]; every device phase (LDE passes, leaf/node hashing, DEEP,
 * FRI rounds, query gathers) is one launch over all k jobs.  Outputs: roots
 * (k*4 words) and eval_points (k*4: two extension points per job). */
int p3g_prove_many(p3g_ctx* ctx, uint32_t hash_kind, const p3g_pcs_params* params, const p3g_matrix* traces,
                   uint64_t k, uint64_t t, uint32_t flags, p3g_transcript* const* chs, uint64_t* roots,
                   uint64_t* eval_points);

/* ---- one-shot proof (the pcs_trace harness flow, p3-miden-lifted-fri/benches/pcs_trace.rs:60-100,
 * with commit_traces semantics): commit the traces, observe the root, sample two
 * extension points, open.  root receives the commitment, eval_points_out (4 words)
 * the sampled points; the transcript holds the proof. */
int p3g_prove(p3g_ctx* ctx, uint32_t hash_kind, const p3g_pcs_params* params, const p3g_matrix* traces, uint64_t k,
              uint32_t flags, p3g_transcript* ch, uint64_t root[4], uint64_t eval_points_out[4]);

#ifdef __cplusplus
}
#endif
#endif
