/*
 * oracle/pcs.c — DEEP quotient, FRI commit/fold/final polynomial, query phase,
 * open_with_channel, the verifier, and commit_traces; plain C + OpenMP.
 *
 * TEST INFRASTRUCTURE ONLY (the checker; never linked by product code).
 *
 * Restated reference functions (paths relative to the reference checkout):
 *  - bit_reversed_coset_points: p3-miden-lifted-fri/src/utils.rs:324-329
 *  - PointQuotients::new: deep/interpolate.rs:81-109 (q[i][j] = 1/(z_j - x_i))
 *  - batch_eval_lifted: deep/interpolate.rs:123-201 (barycentric weights over
 *    the first L/B coset points, folded by chunk sums for lifted heights,
 *    scaled by s_j = ((z_j/g)^d - 1)/d)
 *  - DeepPoly::from_evals: deep/prover.rs:230-409 (aligned evals sent per point,
 *    grind, alpha then beta, Horner f_red(z_j), coefficients -alpha^(W-1-i),
 *    accumulate_matrices :420-488 with upsampling, Q = sum_j beta^j q_j (f_red(z_j) + acc))
 *  - FriParams::num_rounds / final_poly_degree: fri/mod.rs:66-96
 *  - FriPolys::new: fri/prover.rs:190-323; prove_queries :335-356
 *  - fold_evals arity 2/4/8: fri/fold/arity2.rs:59-79, arity4.rs:46-121,
 *    arity8.rs:194-288 (restated as the unique interpolate-and-evaluate value
 *    (1/a) sum_j c_j (beta/s)^j of the bit-reversed coset row)
 *  - open_with_channel: p3-miden-lifted-fri/src/prover.rs:35-116
 *  - verify: verifier.rs:50-102, DeepOracle deep/verifier.rs:73-211,
 *    FriOracle fri/verifier.rs:72-223
 *  - commit_traces: p3-miden-lifted-stark/src/prover/commit.rs:175-228 with
 *    LiftedCoset::lde_shift coset.rs:132-134
 */
#include <stdlib.h>
#include <string.h>
#include "oracle.h"
#include "tx_internal.h"

static size_t aligned_len(size_t len, size_t alignment) {
    if (alignment <= 1) return len;
    return (len + alignment - 1) / alignment * alignment;
}

size_t orc_fri_num_rounds(const orc_pcs_params* p, unsigned log_domain) {
    unsigned lmf = p->log_final_degree + p->log_blowup;
    unsigned diff = log_domain > lmf ? log_domain - lmf : 0;
    return (diff + p->log_arity - 1) / p->log_arity;
}

size_t orc_fri_final_degree(const orc_pcs_params* p, unsigned log_domain) {
    size_t rounds = orc_fri_num_rounds(p, log_domain);
    unsigned lfs = log_domain - (unsigned)(rounds * p->log_arity);
    unsigned e = lfs > p->log_blowup ? lfs - p->log_blowup : 0;
    return (size_t)1 << e;
}

void orc_bit_reversed_coset_points(unsigned log_n, gl* out) {
    size_t n = (size_t)1 << log_n;
    gl g = gl_two_adic_generator(log_n);
    gl x = GL_GENERATOR;
    for (size_t i = 0; i < n; i++) { out[bitrev_idx(i, log_n)] = x; x = gl_mul(x, g); }
}

/* q_out[(i * npts + j) * 2] = 1 / (z_j - x_i) */
void orc_point_quotients(const uint64_t* z, size_t npts, const gl* xs, size_t n, uint64_t* q_out) {
    ef2* q = (ef2*)q_out;
    size_t total = n * npts;
    ef2* diffs = (ef2*)malloc(total * sizeof(ef2));
    #pragma omp parallel for
    for (size_t i = 0; i < n; i++)
        for (size_t j = 0; j < npts; j++)
            diffs[i * npts + j] = ef_sub(ef_make(z[2 * j], z[2 * j + 1]), ef_from(xs[i]));
    /* Montgomery batch inversion in chunks */
    const size_t CH = 4096;
    #pragma omp parallel for schedule(dynamic, 1)
    for (size_t c0 = 0; c0 < total; c0 += CH) {
        size_t c1 = c0 + CH < total ? c0 + CH : total;
        ef2 acc = ef_from(1);
        for (size_t i = c0; i < c1; i++) { q[i] = acc; acc = ef_mul(acc, diffs[i]); }
        ef2 inv = ef_inv(acc);
        for (size_t i = c1; i-- > c0;) {
            ef2 t = ef_mul(inv, q[i]);
            inv = ef_mul(inv, diffs[i]);
            q[i] = t;
        }
    }
    free(diffs);
}

typedef struct {
    size_t count;      /* matrices over all trees */
    size_t* heights;
    size_t* widths;
    const uint64_t** mats;
    size_t* awidths;
} mat_list;

static mat_list collect(orc_tree* const* trees, size_t ntrees) {
    mat_list ml;
    ml.count = 0;
    for (size_t t = 0; t < ntrees; t++) ml.count += trees[t]->k;
    ml.heights = (size_t*)malloc(ml.count * sizeof(size_t));
    ml.widths = (size_t*)malloc(ml.count * sizeof(size_t));
    ml.awidths = (size_t*)malloc(ml.count * sizeof(size_t));
    ml.mats = (const uint64_t**)malloc(ml.count * sizeof(uint64_t*));
    size_t m = 0;
    for (size_t t = 0; t < ntrees; t++)
        for (size_t i = 0; i < trees[t]->k; i++, m++) {
            ml.heights[m] = trees[t]->heights[i];
            ml.widths[m] = trees[t]->widths[i];
            ml.awidths[m] = aligned_len(trees[t]->widths[i], trees[t]->alignment);
            ml.mats[m] = trees[t]->mats[i];
        }
    return ml;
}

static void free_ml(mat_list* ml) {
    free(ml->heights); free(ml->widths); free(ml->awidths); free((void*)ml->mats);
}

/* evals_out: for each matrix (tree order), for each column, for each point: EF (2 words). */
static void batch_eval_lifted_q(const mat_list* ml, const ef2* q, const gl* xs, size_t L,
                                const uint64_t* z, size_t npts, unsigned log_blowup, uint64_t* evals_out) {
    size_t d = L >> log_blowup;
    unsigned log_d = log2_exact(d);
    gl shift = xs[0];
    gl shift_inv = gl_inv(shift);
    gl dinv = gl_inv_pow2(log_d);
    ef2 scal[8];
    for (size_t j = 0; j < npts; j++) {
        ef2 zz = ef_mul_base(ef_make(z[2 * j], z[2 * j + 1]), shift_inv);
        ef2 t = ef_sub(ef_exp_power_of_2(zz, log_d), ef_from(1));
        scal[j] = ef_mul_base(t, dinv);
    }
    /* top weights, then folded weights per used degree (largest first) */
    ef2* top = (ef2*)malloc(d * npts * sizeof(ef2));
    #pragma omp parallel for
    for (size_t i = 0; i < d; i++)
        for (size_t j = 0; j < npts; j++) top[i * npts + j] = ef_mul_base(q[i * npts + j], xs[i]);
    unsigned maxlog = log_d;
    ef2** wts = (ef2**)calloc(maxlog + 1, sizeof(ef2*));
    wts[log_d] = top;
    /* Fold down through every degree below d that some matrix uses, in descending order. */
    int* used = (int*)calloc(maxlog + 1, sizeof(int));
    for (size_t m = 0; m < ml->count; m++) used[log2_exact(ml->heights[m] >> log_blowup)] = 1;
    ef2* prev = top;
    size_t prev_len = d;
    for (int lg = (int)log_d - 1; lg >= 0; lg--) {
        if (!used[lg]) continue;
        size_t len = (size_t)1 << lg;
        size_t cs = prev_len / len;
        ef2* nw = (ef2*)malloc(len * npts * sizeof(ef2));
        for (size_t i = 0; i < len; i++)
            for (size_t j = 0; j < npts; j++) {
                ef2 s = ef_from(0);
                for (size_t c = 0; c < cs; c++) s = ef_add(s, prev[(i * cs + c) * npts + j]);
                nw[i * npts + j] = s;
            }
        wts[lg] = nw;
        prev = nw;
        prev_len = len;
    }
    size_t off = 0;
    for (size_t m = 0; m < ml->count; m++) {
        size_t h = ml->heights[m], w = ml->widths[m];
        size_t dm = h >> log_blowup;
        ef2* wm = wts[log2_exact(dm)];
        const uint64_t* M = ml->mats[m];
        #pragma omp parallel for
        for (size_t c = 0; c < w; c++)
            for (size_t j = 0; j < npts; j++) {
                ef2 acc = ef_from(0);
                for (size_t i = 0; i < dm; i++) acc = ef_add(acc, ef_mul_base(wm[i * npts + j], M[i * w + c]));
                acc = ef_mul(acc, scal[j]);
                evals_out[((off + c) * npts + j) * 2] = acc.c0;
                evals_out[((off + c) * npts + j) * 2 + 1] = acc.c1;
            }
        off += w;
    }
    for (unsigned lg = 0; lg <= maxlog; lg++) free(wts[lg]);
    free(wts);
    free(used);
}

void orc_batch_eval_lifted(const orc_tree* const* trees, size_t ntrees, const uint64_t* z, size_t npts,
                           unsigned log_blowup, uint64_t* evals_out) {
    mat_list ml = collect((orc_tree* const*)trees, ntrees);
    size_t L = trees[0]->height;
    unsigned logL = log2_exact(L);
    gl* xs = (gl*)malloc(L * sizeof(gl));
    orc_bit_reversed_coset_points(logL, xs);
    uint64_t* q = (uint64_t*)malloc(L * npts * 2 * sizeof(uint64_t));
    orc_point_quotients(z, npts, xs, L, q);
    batch_eval_lifted_q(&ml, (const ef2*)q, xs, L, z, npts, log_blowup, evals_out);
    free(q);
    free(xs);
    free_ml(&ml);
}

/* Unique interpolate-and-evaluate fold of one bit-reversed coset row. */
void orc_fri_fold_row(unsigned log_arity, const uint64_t* row, uint64_t s_inv, const uint64_t* beta, uint64_t* out2) {
    size_t a = (size_t)1 << log_arity;
    gl om_inv = gl_inv(gl_two_adic_generator(log_arity));
    ef2 x = ef_mul_base(ef_make(beta[0], beta[1]), s_inv);
    ef2 res = ef_from(0), xp = ef_from(1);
    for (size_t j = 0; j < a; j++) {
        /* C_j = sum_k f(s w^k) w^{-jk};  f(s w^k) sits at position bitrev(k) */
        ef2 cj = ef_from(0);
        for (size_t k = 0; k < a; k++) {
            size_t p = bitrev_idx(k, log_arity);
            gl tw = gl_pow(om_inv, (j * k) % a);
            cj = ef_add(cj, ef_mul_base(ef_make(row[2 * p], row[2 * p + 1]), tw));
        }
        res = ef_add(res, ef_mul(cj, xp));
        xp = ef_mul(xp, x);
    }
    res = ef_mul_base(res, gl_inv_pow2(log_arity));
    out2[0] = res.c0;
    out2[1] = res.c1;
}

static void send_ef_slice(orc_tx* tx, const ef2* v, size_t n) {
    orc_tx_send_fields(tx, (const uint64_t*)v, 2 * n);
}

/* FriPolys::new (p3-miden-lifted-fri/src/fri/prover.rs:90-223): per round
 * commit the (m/a) x a EF matrix (flattened c0,c1) as an unaligned tree, send
 * its root, grind, sample beta, fold; then the final polynomial: truncate to the
 * final degree, bit-reverse, inverse DFT, reverse to descending order, send.
 * Consumes `cur`; returns the number of rounds; trees go to ftrees[r]. */
static size_t fri_commit(const orc_pcs_params* p, int hash_kind, ef2* cur, unsigned log_lde_height, orc_tx* tx,
                         orc_tree** ftrees, uint64_t* final_out) {
    size_t L = (size_t)1 << log_lde_height;
    unsigned la = p->log_arity;
    size_t arity = (size_t)1 << la;
    size_t domain = L;
    size_t fdeg = orc_fri_final_degree(p, log_lde_height);
    size_t fdomain = fdeg << p->log_blowup;
    size_t nrounds = 0;
    gl ginv = gl_inv(gl_two_adic_generator(log_lde_height));
    size_t ns = domain >> la;
    gl* sinv = (gl*)malloc((ns ? ns : 1) * sizeof(gl));
    {
        gl x = 1;
        unsigned lns = log2_exact(ns);
        for (size_t k = 0; k < ns; k++) { sinv[bitrev_idx(k, lns)] = x; x = gl_mul(x, ginv); }
    }
    while (domain > fdomain) {
        size_t fd = domain >> la;
        const uint64_t* mp = (const uint64_t*)cur;
        size_t hh = fd, ww = 2 * arity;
        orc_tree* t = orc_lmcs_build(hash_kind, &mp, &hh, &ww, 1, 1);
        uint64_t root[4];
        orc_tree_root(t, root);
        orc_tx_send_commitment(tx, root);
        orc_tx_grind(tx, p->folding_pow_bits);
        uint64_t bb[2];
        orc_tx_sample_ef(tx, bb);
        ef2* nxt = (ef2*)malloc(fd * sizeof(ef2));
        #pragma omp parallel for schedule(static)
        for (size_t k = 0; k < fd; k++) {
            uint64_t o[2];
            orc_fri_fold_row(la, (const uint64_t*)(cur + k * arity), sinv[k], bb, o);
            nxt[k] = ef_make(o[0], o[1]);
        }
        ftrees[nrounds++] = t;
        free(cur);
        cur = nxt;
        domain = fd;
        size_t nn = domain >> la;
        gl* s2 = (gl*)malloc((nn ? nn : 1) * sizeof(gl));
        for (size_t k = 0; k < nn; k++) s2[k] = gl_pow(sinv[k * arity], arity);
        free(sinv);
        sinv = s2;
    }
    free(sinv);
    {
        ef2* fp = (ef2*)malloc(fdeg * sizeof(ef2));
        unsigned lf = log2_exact(fdeg);
        for (size_t i = 0; i < fdeg; i++) fp[bitrev_idx(i, lf)] = cur[i];
        orc_idft_ext(fp, fdeg);
        for (size_t i = 0; i < fdeg / 2; i++) { ef2 t = fp[i]; fp[i] = fp[fdeg - 1 - i]; fp[fdeg - 1 - i] = t; }
        send_ef_slice(tx, fp, fdeg);
        if (final_out) memcpy(final_out, fp, fdeg * sizeof(ef2));
        free(fp);
    }
    free(cur);
    return nrounds;
}

size_t orc_fri_prove(const orc_pcs_params* p, int hash_kind, const uint64_t* evals, unsigned log_lde_height,
                     orc_tx* tx, uint64_t* final_out) {
    size_t L = (size_t)1 << log_lde_height;
    ef2* cur = (ef2*)malloc(L * sizeof(ef2));
    memcpy(cur, evals, L * sizeof(ef2));
    size_t maxr = orc_fri_num_rounds(p, log_lde_height) + 1;
    orc_tree** ftrees = (orc_tree**)calloc(maxr + 1, sizeof(orc_tree*));
    size_t n = fri_commit(p, hash_kind, cur, log_lde_height, tx, ftrees, final_out);
    for (size_t r = 0; r < n; r++) orc_tree_free(ftrees[r]);
    free(ftrees);
    return n;
}

void orc_open_with_channel(const orc_pcs_params* p, int hash_kind, orc_tree* const* trees, size_t ntrees,
                           unsigned log_lde_height, const uint64_t* z, size_t npts, orc_tx* tx,
                           uint64_t* deep_evals_out) {
    size_t L = (size_t)1 << log_lde_height;
    mat_list ml = collect(trees, ntrees);

    /* ---- DEEP (DeepPoly::from_trees) ---- */
    gl* xs = (gl*)malloc(L * sizeof(gl));
    orc_bit_reversed_coset_points(log_lde_height, xs);
    ef2* q = (ef2*)malloc(L * npts * sizeof(ef2));
    orc_point_quotients(z, npts, xs, L, (uint64_t*)q);
    size_t totw = 0, tota = 0;
    for (size_t m = 0; m < ml.count; m++) { totw += ml.widths[m]; tota += ml.awidths[m]; }
    uint64_t* ev = (uint64_t*)malloc((totw ? totw : 1) * npts * 2 * sizeof(uint64_t));
    batch_eval_lifted_q(&ml, q, xs, L, z, npts, p->log_blowup, ev);
    /* aligned evals per point: aev[j][col] */
    ef2* aev = (ef2*)calloc(tota * npts, sizeof(ef2));
    {
        size_t off = 0, aoff = 0;
        for (size_t m = 0; m < ml.count; m++) {
            for (size_t c = 0; c < ml.widths[m]; c++)
                for (size_t j = 0; j < npts; j++)
                    aev[j * tota + aoff + c] = ef_make(ev[((off + c) * npts + j) * 2], ev[((off + c) * npts + j) * 2 + 1]);
            off += ml.widths[m];
            aoff += ml.awidths[m];
        }
    }
    for (size_t j = 0; j < npts; j++) send_ef_slice(tx, aev + j * tota, tota);
    orc_tx_grind(tx, p->deep_pow_bits);
    uint64_t a2[2], b2[2];
    orc_tx_sample_ef(tx, a2);
    orc_tx_sample_ef(tx, b2);
    ef2 alpha = ef_make(a2[0], a2[1]), beta = ef_make(b2[0], b2[1]);
    ef2 fred[8];
    for (size_t j = 0; j < npts; j++) {
        ef2 acc = ef_from(0);
        for (size_t c = 0; c < tota; c++) acc = ef_add(ef_mul(acc, alpha), aev[j * tota + c]);
        fred[j] = acc;
    }
    /* coefficients -alpha^(W-1-i) */
    ef2* coeff = (ef2*)malloc((tota ? tota : 1) * sizeof(ef2));
    {
        ef2 pw = ef_neg(ef_from(1));
        for (size_t i = tota; i-- > 0;) { coeff[i] = pw; pw = ef_mul(pw, alpha); }
    }
    ef2 bpow[8];
    bpow[0] = ef_from(1);
    for (size_t j = 1; j < npts; j++) bpow[j] = ef_mul(bpow[j - 1], beta);
    ef2* deep = (ef2*)malloc(L * sizeof(ef2));
    #pragma omp parallel for schedule(static)
    for (size_t i = 0; i < L; i++) {
        ef2 acc = ef_from(0);
        size_t aoff = 0;
        for (size_t m = 0; m < ml.count; m++) {
            size_t h = ml.heights[m], w = ml.widths[m];
            size_t row = i >> log2_exact(L / h);
            const uint64_t* R = ml.mats[m] + row * w;
            for (size_t c = 0; c < w; c++) acc = ef_add(acc, ef_mul_base(coeff[aoff + c], R[c]));
            aoff += ml.awidths[m];
        }
        ef2 res = ef_from(0);
        for (size_t j = 0; j < npts; j++)
            res = ef_add(res, ef_mul(bpow[j], ef_mul(q[i * npts + j], ef_add(fred[j], acc))));
        deep[i] = res;
    }
    if (deep_evals_out) memcpy(deep_evals_out, deep, L * sizeof(ef2));
    free(coeff); free(aev); free(ev); free(q); free(xs);

    /* ---- FRI commit phase (FriPolys::new) ---- */
    unsigned la = p->log_arity;
    size_t maxr = orc_fri_num_rounds(p, log_lde_height) + 1;
    orc_tree** ftrees = (orc_tree**)calloc(maxr + 1, sizeof(orc_tree*));
    size_t nrounds = fri_commit(p, hash_kind, deep, log_lde_height, tx, ftrees, NULL);

    /* ---- query phase ---- */
    orc_tx_grind(tx, p->query_pow_bits);
    size_t nq = p->num_queries;
    size_t* idx = (size_t*)malloc((nq ? nq : 1) * sizeof(size_t));
    for (size_t i = 0; i < nq; i++) idx[i] = bitrev_idx(orc_tx_sample_bits(tx, log_lde_height), log_lde_height);
    for (size_t t = 0; t < ntrees; t++) orc_tree_prove_batch(trees[t], idx, nq, tx);
    size_t* ri = (size_t*)malloc((nq ? nq : 1) * sizeof(size_t));
    for (size_t r = 0; r < nrounds; r++) {
        unsigned sh = la * (unsigned)(r + 1);
        for (size_t i = 0; i < nq; i++) ri[i] = idx[i] >> sh;
        orc_tree_prove_batch(ftrees[r], ri, nq, tx);
        orc_tree_free(ftrees[r]);
    }
    free(ri);
    free(idx);
    free(ftrees);
    free_ml(&ml);
}

/* ---------------- verifier ---------------- */

typedef struct { size_t idx; ef2 v; } kv;

static int cmp_kv(const void* a, const void* b) {
    size_t x = ((const kv*)a)->idx, y = ((const kv*)b)->idx;
    return x < y ? -1 : x > y;
}

int orc_verify(const orc_pcs_params* p, int hash_kind, const uint64_t* roots, const size_t* widths_flat,
               const size_t* nwidths, size_t ntrees, unsigned log_lde_height, const uint64_t* z,
               size_t npts, orc_tx* tx, uint64_t* evals_out) {
    if (ntrees == 0) return 1;
    size_t tota = 0, nmat = 0;
    for (size_t t = 0; t < ntrees; t++) for (size_t i = 0; i < nwidths[t]; i++) tota += widths_flat[nmat++];
    int rc = 0;
    ef2* evals = (ef2*)malloc((tota ? tota : 1) * npts * sizeof(ef2));
    for (size_t j = 0; j < npts; j++)
        if (orc_tx_receive_fields(tx, 2 * tota, (uint64_t*)(evals + j * tota))) return 20;
    if (evals_out) memcpy(evals_out, evals, tota * npts * sizeof(ef2));
    if (orc_tx_check_grind(tx, p->deep_pow_bits)) { free(evals); return 21; }
    uint64_t a2[2], b2[2];
    orc_tx_sample_ef(tx, a2);
    orc_tx_sample_ef(tx, b2);
    ef2 alpha = ef_make(a2[0], a2[1]), beta = ef_make(b2[0], b2[1]);
    ef2 fred[8];
    for (size_t j = 0; j < npts; j++) {
        ef2 acc = ef_from(0);
        for (size_t c = 0; c < tota; c++) acc = ef_add(ef_mul(acc, alpha), evals[j * tota + c]);
        fred[j] = acc;
    }
    free(evals);
    /* FRI oracle */
    size_t nr = orc_fri_num_rounds(p, log_lde_height);
    uint64_t* fcom = (uint64_t*)malloc((nr ? nr : 1) * 32);
    ef2* fbeta = (ef2*)malloc((nr ? nr : 1) * sizeof(ef2));
    for (size_t r = 0; r < nr; r++) {
        if (orc_tx_receive_commitment(tx, fcom + 4 * r)) { rc = 30; goto out0; }
        if (orc_tx_check_grind(tx, p->folding_pow_bits)) { rc = 31; goto out0; }
        uint64_t bb[2];
        orc_tx_sample_ef(tx, bb);
        fbeta[r] = ef_make(bb[0], bb[1]);
    }
    size_t fdeg = orc_fri_final_degree(p, log_lde_height);
    ef2* fpoly = (ef2*)malloc(fdeg * sizeof(ef2));
    if (orc_tx_receive_fields(tx, 2 * fdeg, (uint64_t*)fpoly)) { rc = 32; goto out1; }
    if (orc_tx_check_grind(tx, p->query_pow_bits)) { rc = 33; goto out1; }
    {
        size_t nq = p->num_queries;
        size_t* idx = (size_t*)malloc(nq * sizeof(size_t));
        for (size_t i = 0; i < nq; i++) idx[i] = bitrev_idx(orc_tx_sample_bits(tx, log_lde_height), log_lde_height);
        /* unique sorted */
        kv* cur = (kv*)malloc(nq * sizeof(kv));
        size_t nu = 0;
        for (size_t i = 0; i < nq; i++) { cur[nu].idx = idx[i]; cur[nu].v = ef_from(0); nu++; }
        qsort(cur, nu, sizeof(kv), cmp_kv);
        { size_t u = 1; for (size_t i = 1; i < nu; i++) if (cur[i].idx != cur[u - 1].idx) cur[u++] = cur[i]; nu = u; }
        size_t* uidx = (size_t*)malloc(nu * sizeof(size_t));
        for (size_t i = 0; i < nu; i++) uidx[i] = cur[i].idx;
        /* DEEP open_batch over each tree */
        size_t moff = 0;
        for (size_t t = 0; t < ntrees; t++) {
            size_t tw = 0;
            for (size_t i = 0; i < nwidths[t]; i++) tw += widths_flat[moff + i];
            uint64_t* rows = (uint64_t*)malloc((tw ? tw : 1) * nu * sizeof(uint64_t));
            int e = orc_lmcs_open_batch(hash_kind, roots + 4 * t, widths_flat + moff, nwidths[t], log_lde_height,
                                        uidx, nu, tx, rows);
            if (e) { free(rows); rc = 40 + e; free(uidx); free(cur); free(idx); goto out1; }
            for (size_t q = 0; q < nu; q++)
                for (size_t c = 0; c < tw; c++)
                    cur[q].v = ef_add(ef_mul(cur[q].v, alpha), ef_from(rows[q * tw + c]));
            free(rows);
            moff += nwidths[t];
        }
        {
            gl gen = gl_two_adic_generator(log_lde_height);
            for (size_t q = 0; q < nu; q++) {
                size_t ex = bitrev_idx(cur[q].idx, log_lde_height);
                gl x = gl_mul(GL_GENERATOR, gl_pow(gen, ex));
                ef2 red = cur[q].v, acc = ef_from(0), bp = ef_from(1);
                for (size_t j = 0; j < npts; j++) {
                    ef2 den = ef_sub(ef_make(z[2 * j], z[2 * j + 1]), ef_from(x));
                    if (ef_is_zero(den)) { rc = 50; break; }
                    acc = ef_add(acc, ef_mul(bp, ef_mul(ef_sub(fred[j], red), ef_inv(den))));
                    bp = ef_mul(bp, beta);
                }
                cur[q].v = acc;
            }
        }
        /* FRI test_low_degree */
        unsigned la = p->log_arity;
        size_t arity = (size_t)1 << la;
        unsigned logd = log_lde_height;
        gl ginv = gl_inv(gl_two_adic_generator(logd));
        for (size_t r = 0; r < nr && !rc; r++) {
            unsigned lfd = logd - la;
            size_t* ri = (size_t*)malloc(nu * sizeof(size_t));
            for (size_t q = 0; q < nu; q++) ri[q] = cur[q].idx >> la;
            /* unique row indices (cur sorted => ri sorted) */
            size_t nrow = 0;
            for (size_t q = 0; q < nu; q++) if (!nrow || ri[nrow - 1] != ri[q]) ri[nrow++] = ri[q];
            size_t bw = 2 * arity;
            uint64_t* rows = (uint64_t*)malloc(nrow * bw * sizeof(uint64_t));
            int e = orc_lmcs_open_batch(hash_kind, fcom + 4 * r, &bw, 1, lfd, ri, nrow, tx, rows);
            if (e) { rc = 60 + e; free(rows); free(ri); break; }
            size_t nn = 0;
            for (size_t q = 0; q < nu; q++) {
                size_t row_idx = cur[q].idx >> la, pos = cur[q].idx & (arity - 1);
                size_t rr = 0;
                while (ri[rr] != row_idx) rr++;
                const uint64_t* R = rows + rr * bw;
                if (!ef_eq(ef_make(R[2 * pos], R[2 * pos + 1]), cur[q].v)) { rc = 70; break; }
                gl s_inv = gl_pow(ginv, bitrev_idx(row_idx, lfd));
                uint64_t o[2], bb[2] = {fbeta[r].c0, fbeta[r].c1};
                orc_fri_fold_row(la, R, s_inv, bb, o);
                if (nn && cur[nn - 1].idx == row_idx) continue;
                cur[nn].idx = row_idx;
                cur[nn].v = ef_make(o[0], o[1]);
                nn++;
            }
            nu = nn;
            free(rows);
            free(ri);
            logd = lfd;
            ginv = gl_pow(ginv, arity);
        }
        if (!rc) {
            gl gen = gl_two_adic_generator(logd);
            for (size_t q = 0; q < nu; q++) {
                gl x = gl_pow(gen, bitrev_idx(cur[q].idx, logd));
                ef2 acc = ef_from(0);
                for (size_t i = 0; i < fdeg; i++) acc = ef_add(ef_mul_base(acc, x), fpoly[i]);
                if (!ef_eq(acc, cur[q].v)) { rc = 80; break; }
            }
        }
        free(uidx);
        free(cur);
        free(idx);
    }
out1:
    free(fpoly);
out0:
    free(fcom);
    free(fbeta);
    return rc;
}

orc_tree* orc_commit_traces(int hash_kind, const uint64_t* const* traces, const size_t* heights,
                            const size_t* widths, size_t k, unsigned log_blowup) {
    if (k == 0) return NULL;
    for (size_t i = 1; i < k; i++) if (heights[i] < heights[i - 1]) return NULL;
    unsigned lmax = log2_exact(heights[k - 1]);
    uint64_t** ldes = (uint64_t**)malloc(k * sizeof(uint64_t*));
    size_t* lh = (size_t*)malloc(k * sizeof(size_t));
    for (size_t i = 0; i < k; i++) {
        unsigned lt = log2_exact(heights[i]);
        gl shift = GL_GENERATOR;
        for (unsigned s = lt; s < lmax; s++) shift = gl_mul(shift, shift);
        lh[i] = heights[i] << log_blowup;
        ldes[i] = (uint64_t*)malloc(lh[i] * widths[i] * sizeof(uint64_t) + 8);
        orc_coset_lde_batch(traces[i], heights[i], widths[i], log_blowup, shift, ldes[i]);
        orc_bit_reverse_rows(ldes[i], lh[i], widths[i]);
    }
    orc_tree* t = orc_lmcs_build(hash_kind, (const uint64_t* const*)ldes, lh, widths, k, orc_hash_alignment(hash_kind));
    for (size_t i = 0; i < k; i++) free(ldes[i]);
    free(ldes);
    free(lh);
    return t;
}
