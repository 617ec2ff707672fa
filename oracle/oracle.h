/*
 * oracle/oracle.h — C API of the CPU restatement of the reference's lifted
 * PCS path (coset LDE, LMCS commit/open, DEEP, FRI, transcript).
 *
 * TEST INFRASTRUCTURE ONLY.  Only the tests, the smoke check and the CPU
 * baseline of the benchmark load liboracle.so, and only as the checker.
 *
 * Pins (tests/test_oracle_pins.py, tests/test_golden.py): field arithmetic and
 * two-adic generators by closed form; the xoshiro256++ generator by its
 * published reference vector; NTT / coset LDE / bit reversal by naive DFT and
 * interpolation; the stateful sponge by the reference's mock-permutation KAT
 * and padding-equivalence tests; alignment by the reference's table; Keccak-f
 * by SHA3 (hashlib); LMCS structure by lifted == upsampled == concatenated;
 * DEEP by polynomial evaluation and the reference's negative-coefficient /
 * Horner identities; FRI folding by evaluation of low-degree polynomials and
 * the reference's final-polynomial KAT; prover/verifier round trips.
 *
 * Parity unpinned (the algorithm lives in un-vendored Plonky3 0.5.0 / rand
 * 0.10, which cannot be built or imported here, and no reference test holds
 * an output vector for it):
 *   - Poseidon2Goldilocks<12> round constants drawn from SmallRng(2025), the
 *     MATRIX_DIAG_12_GOLDILOCKS diagonal and the exact round layout
 *     (orc_p2_init_constants, orc_p2_permute);
 *   - SmallRng seeding (SplitMix64) and StandardUniform<Goldilocks> rejection
 *     sampling order (orc_rng_seed, orc_rng_gl, orc_rand_fill);
 *   - TruncatedPermutation / PaddingFreeSponge compression layouts (orc_lmcs_*);
 *   - DuplexChallenger observe/sample/grind semantics and the Keccak
 *     transcript, which the reference leaves build-defined (orc_tx_*);
 *   - Goldilocks u64 serialisation into the Keccak sponge.
 */
#ifndef ORACLE_H
#define ORACLE_H

#include "gl.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- seeded generator (rand SmallRng restatement) ---- */
typedef struct { uint64_t s[4]; } orc_rng;
void orc_rng_seed(orc_rng* r, uint64_t seed);
uint64_t orc_rng_next(orc_rng* r);
gl orc_rng_gl(orc_rng* r);
void orc_rand_fill(uint64_t seed, size_t n, uint64_t* out);
void orc_rand_fill_rng(orc_rng* r, size_t n, uint64_t* out);

/* ---- permutations ---- */
void orc_p2_init_constants(uint64_t seed);
void orc_p2_constants(uint64_t* out /* 48 + 48 + 22 */);
void orc_p2_diag(uint64_t* out /* 12 */);
void orc_p2_m4(gl* x);
void orc_p2_external_linear(gl* s);
void orc_p2_internal_linear(gl* s);
void orc_p2_permute(uint64_t* s /* 12 */);
void orc_keccakf(uint64_t* a /* 25 */);
void orc_keccak256(const uint8_t* in, size_t len, uint8_t* out32);

/* ---- field helpers exported for tests ---- */
uint64_t orc_gl_mul(uint64_t a, uint64_t b);
uint64_t orc_gl_add(uint64_t a, uint64_t b);
uint64_t orc_gl_sub(uint64_t a, uint64_t b);
uint64_t orc_gl_inv(uint64_t a);
uint64_t orc_gl_pow(uint64_t a, uint64_t e);
uint64_t orc_two_adic_generator(unsigned bits);
void orc_ef_mul(const uint64_t* a, const uint64_t* b, uint64_t* out);
void orc_ef_inv(const uint64_t* a, uint64_t* out);

/* ---- DFT ---- */
void orc_ntt(gl* a, size_t n, int inverse);
void orc_coset_lde_batch(const gl* trace, size_t n, size_t w, unsigned added_bits, gl shift, gl* out);
void orc_bit_reverse_rows(gl* m, size_t h, size_t w);
void orc_naive_coset_dft(const gl* coeffs, size_t n, size_t w, gl shift, gl* out);
void orc_idft_ext(ef2* v, size_t n);

/* ---- sponge / compression ---- */
#define ORC_HASH_P2 0
#define ORC_HASH_KECCAK 1
#define ORC_HASH_MOCK 2
void orc_mock_config(size_t width, size_t rate);
size_t orc_hash_width(int kind);
size_t orc_hash_rate(int kind);
size_t orc_hash_alignment(int kind);
size_t orc_field_alignment(size_t field_bytes, size_t word_bytes, size_t rate);
void orc_sponge_absorb(int kind, uint64_t* state, const uint64_t* in, size_t len);
void orc_compress(int kind, const uint64_t* left, const uint64_t* right, uint64_t* out);
void orc_hash_rows(int kind, const uint64_t* elems, size_t n, uint64_t* out4);

/* ---- LMCS tree ---- */
typedef struct orc_tx orc_tx;
typedef struct {
    int kind;
    size_t k;
    size_t* heights;
    size_t* widths;
    uint64_t** mats;
    size_t alignment;
    size_t height;
    unsigned depth;
    uint64_t* salt;    /* hiding trees: height x salt_w, absorbed after the rows; NULL without salt */
    size_t salt_w;
    uint64_t** layers; /* depth+1 layers, layer l holds (height >> l) digests of 4 words */
} orc_tree;

orc_tree* orc_lmcs_build(int kind, const uint64_t* const* mats, const size_t* heights,
                         const size_t* widths, size_t k, size_t alignment);
/* HidingLmcsConfig build (hiding_lmcs.rs:104-140, lifted_tree.rs:271-279) with a caller-given salt
 * matrix (height x salt_w) absorbed after each leaf's rows */
orc_tree* orc_lmcs_build_salted(int kind, const uint64_t* const* mats, const size_t* heights,
                                const size_t* widths, size_t k, size_t alignment, const uint64_t* salt,
                                size_t salt_w);
size_t orc_tree_salt(const orc_tree* t, size_t index, uint64_t* out);
void orc_tree_free(orc_tree* t);
void orc_tree_root(const orc_tree* t, uint64_t* out4);
size_t orc_tree_height(const orc_tree* t);
void orc_tree_layer(const orc_tree* t, unsigned layer, uint64_t* out);
size_t orc_tree_rows(const orc_tree* t, size_t index, uint64_t* out);
void orc_tree_prove_batch(const orc_tree* t, const size_t* indices, size_t n, orc_tx* tx);
/* Verifier side (Lmcs::open_batch).  Returns 0 ok, 1 invalid proof, 2 root mismatch, 3 transcript. */
/* open_batch of a hiding tree: every leaf's rows are followed by salt_w hinted salt elements */
int orc_lmcs_open_batch_salted(int kind, const uint64_t* root, const size_t* widths, size_t nw,
                               unsigned log_max_height, const size_t* indices, size_t n, size_t salt_w, orc_tx* tx,
                               uint64_t* rows_out);
int orc_lmcs_open_batch(int kind, const uint64_t* root, const size_t* widths, size_t nw,
                        unsigned log_max_height, const size_t* indices, size_t n, orc_tx* tx,
                        uint64_t* rows_out /* n_unique * sum(widths), sorted */);

/* ---- transcript / challenger ---- */
#define ORC_CH_DUPLEX_P2 0
#define ORC_CH_KECCAK 1
orc_tx* orc_tx_new(int kind);
orc_tx* orc_tx_clone(const orc_tx* tx);
void orc_tx_free(orc_tx* tx);
void orc_tx_observe(orc_tx* tx, uint64_t v);
void orc_tx_observe_commitment(orc_tx* tx, const uint64_t* c4);
uint64_t orc_tx_sample(orc_tx* tx);
void orc_tx_sample_ef(orc_tx* tx, uint64_t* out2);
size_t orc_tx_sample_bits(orc_tx* tx, unsigned bits);
void orc_tx_send_fields(orc_tx* tx, const uint64_t* v, size_t n);
void orc_tx_send_commitment(orc_tx* tx, const uint64_t* c4);
void orc_tx_hint_fields(orc_tx* tx, const uint64_t* v, size_t n);
void orc_tx_hint_commitment(orc_tx* tx, const uint64_t* c4);
uint64_t orc_tx_grind(orc_tx* tx, unsigned bits);
size_t orc_tx_nfields(const orc_tx* tx);
size_t orc_tx_ncomms(const orc_tx* tx);
void orc_tx_get(const orc_tx* tx, uint64_t* fields, uint64_t* comms);
/* verifier mode: load recorded data, then receive_* consume it */
void orc_tx_load(orc_tx* tx, const uint64_t* fields, size_t nf, const uint64_t* comms, size_t nc);
int orc_tx_receive_fields(orc_tx* tx, size_t n, uint64_t* out);      /* observed */
int orc_tx_receive_commitment(orc_tx* tx, uint64_t* out4);          /* observed */
int orc_tx_receive_hint_fields(orc_tx* tx, size_t n, uint64_t* out);
int orc_tx_receive_hint_commitment(orc_tx* tx, uint64_t* out4);
int orc_tx_check_grind(orc_tx* tx, unsigned bits);
int orc_tx_is_empty(const orc_tx* tx);
int orc_tx_check_witness(orc_tx* tx, unsigned bits, uint64_t witness);

/* ---- PCS ---- */
typedef struct {
    uint32_t log_blowup;
    uint32_t log_arity;
    uint32_t log_final_degree;
    uint32_t folding_pow_bits;
    uint32_t deep_pow_bits;
    uint32_t num_queries;
    uint32_t query_pow_bits;
} orc_pcs_params;

size_t orc_fri_num_rounds(const orc_pcs_params* p, unsigned log_domain);
size_t orc_fri_final_degree(const orc_pcs_params* p, unsigned log_domain);
void orc_bit_reversed_coset_points(unsigned log_n, gl* out);

/* DEEP pieces (exported for tests) */
void orc_point_quotients(const uint64_t* z, size_t npts, const gl* coset_points, size_t n, uint64_t* q_out);
void orc_batch_eval_lifted(const orc_tree* const* trees, size_t ntrees, const uint64_t* z, size_t npts,
                           unsigned log_blowup, uint64_t* evals_out /* sum widths * npts * 2 */);
void orc_fri_fold_row(unsigned log_arity, const uint64_t* row /* arity EF */, uint64_t s_inv,
                      const uint64_t* beta, uint64_t* out2);

/* open_with_channel (p3-miden-lifted-fri/src/prover.rs:35-116); deep_evals_out optional (L*2) */
/* FriPolys::new on caller evals (L EF, bit-reversed): rounds; final poly (descending) to final_out */
size_t orc_fri_prove(const orc_pcs_params* p, int hash_kind, const uint64_t* evals, unsigned log_lde_height,
                     orc_tx* tx, uint64_t* final_out);
void orc_open_with_channel(const orc_pcs_params* p, int hash_kind, orc_tree* const* trees, size_t ntrees,
                           unsigned log_lde_height, const uint64_t* z, size_t npts, orc_tx* tx,
                           uint64_t* deep_evals_out);
/* verify (p3-miden-lifted-fri/src/verifier.rs:50-102); widths are the aligned widths. 0 = accept. */
int orc_verify(const orc_pcs_params* p, int hash_kind, const uint64_t* roots, const size_t* widths_flat,
               const size_t* nwidths, size_t ntrees, unsigned log_lde_height, const uint64_t* z,
               size_t npts, orc_tx* tx, uint64_t* evals_out);

/* commit_traces (p3-miden-lifted-stark/src/prover/commit.rs:175-228) */
orc_tree* orc_commit_traces(int hash_kind, const uint64_t* const* traces, const size_t* heights,
                            const size_t* widths, size_t k, unsigned log_blowup);

/* commit_quotient (p3-miden-lifted-stark/src/prover/quotient.rs:133-238): q_evals holds N*D
 * extension elements (c0, c1) in natural order on gJ; out is the (N*B) x 2D bit-reversed LDE.
 * Returns -1 (NULL) unless N and D are powers of two with D <= B. */
int orc_quotient_lde(const uint64_t* q_evals, size_t n, size_t d, unsigned log_blowup, uint64_t* out);
orc_tree* orc_commit_quotient(int hash_kind, const uint64_t* q_evals, size_t n, size_t d, unsigned log_blowup);
/* cyclic_extend_and_scale (quotient.rs:40-55), out of place (m == 0 zero-fills) */
int orc_cyclic_extend_scale(const uint64_t* acc, size_t m, const uint64_t* beta, size_t target, uint64_t* out);
/* divide_by_vanishing_in_place (quotient.rs:72-101): n*d EF, coset shift s; -1 if Z_H vanishes */
int orc_divide_by_vanishing(uint64_t* num, size_t n, size_t d, uint64_t shift);

/* evaluate_constraints_into (p3-miden-lifted-stark/src/prover/constraints/mod.rs:63-179) for an
 * AIR given as a straight-line program (the op list of include/p3gpu.h); main / aux are committed
 * LDEs with bit-reversed rows (at least n*2^log_d rows), periodic a natural-order table on gJ;
 * acc (n*2^log_d EF) += the folded constraints.  -1 on a malformed program or shape. */
typedef struct { uint32_t op, a, b, reserved; uint64_t imm; } orc_air_op;
int orc_eval_constraints(const orc_air_op* ops, size_t nops, const uint64_t* main, size_t main_h, size_t main_w,
                         const uint64_t* aux, size_t aux_h, size_t aux_w, const uint64_t* periodic, size_t prow,
                         size_t pcols, const uint64_t* publics, size_t npub, const uint64_t* ext, size_t next,
                         size_t n, unsigned log_d, uint64_t shift, const uint64_t* alpha, uint64_t* acc);

void orc_set_threads(int n);
int orc_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
