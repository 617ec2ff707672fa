/*
 * oracle/quotient.c — quotient commitment, plain C + OpenMP.
 *
 * TEST INFRASTRUCTURE ONLY (the checker; never linked by product code).
 *
 * Restates `commit_quotient` (p3-miden-lifted-stark/src/prover/quotient.rs:133-238)
 * step by step:
 *   0. q_evals (N*D extension elements, natural order on gJ) viewed as an N x D
 *      matrix: entry (r, t) = Q(g * omega_J^t * omega_H^r)            (:156-160)
 *   1. iDFT over H of every column (`idft_algebra_batch`; the extension iDFT
 *      acts on the two base coordinates independently)                (:162-169)
 *   2. coefficient row k, column t multiplied by (omega_J^-k)^t,
 *      omega_J = two_adic_generator(log N + log D)                    (:171-187)
 *   3. flattened to base (c0 then c1 per element: N x 2D), zero-padded to N*B
 *      rows                                                            (:189-221)
 *   4. plain (unshifted) DFT of size N*B on every column               (:223-227)
 *   5. rows bit-reversed, then `build_aligned_tree` of the single matrix (:229-235)
 * B >= D is required (:154).
 *
 * Also the accumulation helpers ahead of it (quotient.rs:28-101):
 *   cyclic_extend_and_scale: an empty accumulator becomes target_len zeros;
 *     otherwise every element is multiplied by beta, then the buffer doubles
 *     by repetition up to target_len                                   (:40-55)
 *   divide_by_vanishing_in_place: numerator[i] / Z_H(x_i), Z_H(X) = X^N - 1,
 *     x_i = s * omega_J^i on the quotient coset sJ (|J| = N*D); Z_H takes the
 *     D values s^N * omega_D^t - 1, indexed by i mod D              (:72-101)
 */
#include <stdlib.h>
#include <string.h>
#include <omp.h>
#include "oracle.h"

int orc_quotient_lde(const uint64_t* q_evals, size_t n, size_t d, unsigned log_blowup, uint64_t* out) {
    if (n == 0 || (n & (n - 1)) || d == 0 || (d & (d - 1))) return -1;
    const unsigned logn = log2_exact(n), logd = log2_exact(d);
    const size_t B = (size_t)1 << log_blowup;
    if (B < d) return -1;
    const size_t N = n * B, w = 2 * d;
    const gl om_j_inv = gl_inv(gl_two_adic_generator(logn + logd));
    #pragma omp parallel
    {
        gl* col = (gl*)malloc(N * sizeof(gl));
        #pragma omp for schedule(dynamic, 1)
        for (size_t c = 0; c < w; c++) {
            const size_t t = c / 2, comp = c % 2;
            for (size_t r = 0; r < n; r++) col[r] = q_evals[2 * (r * d + t) + comp];
            orc_ntt(col, n, 1);                                  /* step 1 */
            const gl step = gl_pow(om_j_inv, t);                 /* step 2: (omega_J^-k)^t = (omega_J^-t)^k */
            gl f = 1;
            for (size_t k = 0; k < n; k++) { col[k] = gl_mul(col[k], f); f = gl_mul(f, step); }
            for (size_t k = n; k < N; k++) col[k] = 0;           /* step 3 */
            orc_ntt(col, N, 0);                                  /* step 4 */
            for (size_t r = 0; r < N; r++) out[r * w + c] = col[r];
        }
        free(col);
    }
    orc_bit_reverse_rows(out, N, w);                             /* step 5 */
    return 0;
}

orc_tree* orc_commit_quotient(int hash_kind, const uint64_t* q_evals, size_t n, size_t d, unsigned log_blowup) {
    const size_t N = n << log_blowup, w = 2 * d;
    uint64_t* lde = (uint64_t*)malloc(N * w * sizeof(uint64_t) + 8);
    if (orc_quotient_lde(q_evals, n, d, log_blowup, lde) != 0) {
        free(lde);
        return NULL;
    }
    const uint64_t* mats[1] = {lde};
    size_t h = N;
    orc_tree* t = orc_lmcs_build(hash_kind, mats, &h, &w, 1, orc_hash_alignment(hash_kind));
    free(lde);
    return t;
}

/* cyclic_extend_and_scale (quotient.rs:40-55), out of place: out has target
 * elements; m == 0 zero-fills.  EF elements are (c0, c1) word pairs. */
int orc_cyclic_extend_scale(const uint64_t* acc, size_t m, const uint64_t* beta, size_t target, uint64_t* out) {
    if (target == 0 || (target & (target - 1)) || (m && ((m & (m - 1)) || m > target))) return -1;
    if (m == 0) {
        memset(out, 0, target * 2 * sizeof(uint64_t));
        return 0;
    }
    const ef2 b = ef_make(beta[0], beta[1]);
    for (size_t i = 0; i < m; i++) {
        ef2 v = ef_mul(ef_make(acc[2 * i], acc[2 * i + 1]), b);
        out[2 * i] = v.c0;
        out[2 * i + 1] = v.c1;
    }
    for (size_t len = m; len < target; len *= 2) memcpy(out + 2 * len, out, len * 2 * sizeof(uint64_t));
    return 0;
}

/* divide_by_vanishing_in_place (quotient.rs:72-101) on n*d EF elements; the
 * quotient coset shift s is coset.lde_shift().  -1 when a Z_H value is zero. */
int orc_divide_by_vanishing(uint64_t* num, size_t n, size_t d, uint64_t shift) {
    if (n == 0 || (n & (n - 1)) || d == 0 || (d & (d - 1))) return -1;
    const unsigned logn = log2_exact(n), logd = log2_exact(d);
    gl s_pow_n = shift;
    for (unsigned i = 0; i < logn; i++) s_pow_n = gl_mul(s_pow_n, s_pow_n);
    const gl wd = gl_two_adic_generator(logd);
    gl* inv = (gl*)malloc(d * sizeof(gl));
    gl x = 1;
    for (size_t t = 0; t < d; t++) {
        const gl z = gl_sub(gl_mul(s_pow_n, x), 1);
        if (z == 0) {
            free(inv);
            return -1;
        }
        inv[t] = gl_inv(z);
        x = gl_mul(x, wd);
    }
    const size_t total = n * d;
    #pragma omp parallel for schedule(static)
    for (size_t i = 0; i < total; i++) {
        const gl f = inv[i & (d - 1)];
        num[2 * i] = gl_mul(num[2 * i], f);
        num[2 * i + 1] = gl_mul(num[2 * i + 1], f);
    }
    free(inv);
    return 0;
}
