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
