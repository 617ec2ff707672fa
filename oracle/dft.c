/*
 * oracle/dft.c — radix-2 NTT, coset LDE and bit reversal, plain C + OpenMP.
 *
 * TEST INFRASTRUCTURE ONLY (the checker; never linked by product code).
 *
 * Semantics restated: `TwoAdicSubgroupDft::coset_lde_batch(mat, added_bits,
 * shift)` followed by `.bit_reverse_rows()` as called in
 * p3-miden-lifted-stark/src/prover/commit.rs:217-221 and
 * p3-miden-lifted-fri/benches/pcs_trace.rs:73-74: per column, interpolate the
 * natural-order values over the subgroup H_n, evaluate the polynomial at
 * shift * omega_{n 2^added_bits}^i (i natural), then permute rows by
 * bit reversal.  These values are mathematically unique, so any exact
 * algorithm reproduces p3-dft's Radix2DitParallel output.
 * `idft_algebra` of the final FRI polynomial (fri/prover.rs:309-313) is
 * orc_idft_ext.
 */
#include <stdlib.h>
#include <string.h>
#include <omp.h>
#include "oracle.h"

void orc_ntt(gl* a, size_t n, int inverse) {
    unsigned logn = log2_exact(n);
    for (size_t i = 0; i < n; i++) {
        size_t j = bitrev_idx(i, logn);
        if (i < j) { gl t = a[i]; a[i] = a[j]; a[j] = t; }
    }
    for (unsigned s = 1; s <= logn; s++) {
        size_t m = (size_t)1 << s, half = m >> 1;
        gl wm = gl_two_adic_generator(s);
        if (inverse) wm = gl_inv(wm);
        /* twiddles for this stage */
        gl* tw = (gl*)malloc(half * sizeof(gl));
        tw[0] = 1;
        for (size_t j = 1; j < half; j++) tw[j] = gl_mul(tw[j - 1], wm);
        for (size_t k = 0; k < n; k += m)
            for (size_t j = 0; j < half; j++) {
                gl u = a[k + j], v = gl_mul(a[k + j + half], tw[j]);
                a[k + j] = gl_add(u, v);
                a[k + j + half] = gl_sub(u, v);
            }
        free(tw);
    }
    if (inverse && n > 1) {
        gl ninv = gl_inv_pow2(logn);
        for (size_t i = 0; i < n; i++) a[i] = gl_mul(a[i], ninv);
    }
}

void orc_coset_lde_batch(const gl* trace, size_t n, size_t w, unsigned added_bits, gl shift, gl* out) {
    size_t N = n << added_bits;
    #pragma omp parallel
    {
        gl* col = (gl*)malloc(N * sizeof(gl));
        #pragma omp for schedule(dynamic, 1)
        for (size_t c = 0; c < w; c++) {
            for (size_t r = 0; r < n; r++) col[r] = trace[r * w + c];
            orc_ntt(col, n, 1);
            gl sp = 1;
            for (size_t k = 0; k < n; k++) { col[k] = gl_mul(col[k], sp); sp = gl_mul(sp, shift); }
            for (size_t k = n; k < N; k++) col[k] = 0;
            orc_ntt(col, N, 0);
            for (size_t r = 0; r < N; r++) out[r * w + c] = col[r];
        }
        free(col);
    }
}

void orc_bit_reverse_rows(gl* m, size_t h, size_t w) {
    unsigned lh = log2_exact(h);
    gl* tmp = (gl*)malloc(w * sizeof(gl));
    for (size_t i = 0; i < h; i++) {
        size_t j = bitrev_idx(i, lh);
        if (i < j) {
            memcpy(tmp, m + i * w, w * sizeof(gl));
            memcpy(m + i * w, m + j * w, w * sizeof(gl));
            memcpy(m + j * w, tmp, w * sizeof(gl));
        }
    }
    free(tmp);
}

/* out[i][c] = sum_k coeffs[k][c] * (shift * omega_n^i)^k — the O(n^2) brute force
 * the reference uses as its own oracle (NaiveDft, deep/interpolate.rs:259). */
void orc_naive_coset_dft(const gl* coeffs, size_t n, size_t w, gl shift, gl* out) {
    gl om = gl_two_adic_generator(log2_exact(n));
    gl x = shift;
    for (size_t i = 0; i < n; i++) {
        for (size_t c = 0; c < w; c++) {
            gl acc = 0, xp = 1;
            for (size_t k = 0; k < n; k++) { acc = gl_add(acc, gl_mul(coeffs[k * w + c], xp)); xp = gl_mul(xp, x); }
            out[i * w + c] = acc;
        }
        x = gl_mul(x, om);
    }
}

/* Inverse DFT over the subgroup of size n, EF-valued (natural in, natural out). */
void orc_idft_ext(ef2* v, size_t n) {
    gl* a = (gl*)malloc(n * sizeof(gl));
    gl* b = (gl*)malloc(n * sizeof(gl));
    for (size_t i = 0; i < n; i++) { a[i] = v[i].c0; b[i] = v[i].c1; }
    orc_ntt(a, n, 1);
    orc_ntt(b, n, 1);
    for (size_t i = 0; i < n; i++) v[i] = ef_make(a[i], b[i]);
    free(a);
    free(b);
}

uint64_t orc_gl_mul(uint64_t a, uint64_t b) { return gl_mul(a, b); }
uint64_t orc_gl_add(uint64_t a, uint64_t b) { return gl_add(a, b); }
uint64_t orc_gl_sub(uint64_t a, uint64_t b) { return gl_sub(a, b); }
uint64_t orc_gl_inv(uint64_t a) { return gl_inv(a); }
uint64_t orc_gl_pow(uint64_t a, uint64_t e) { return gl_pow(a, e); }
uint64_t orc_two_adic_generator(unsigned bits) { return gl_two_adic_generator(bits); }
void orc_ef_mul(const uint64_t* a, const uint64_t* b, uint64_t* out) {
    ef2 r = ef_mul(ef_make(a[0], a[1]), ef_make(b[0], b[1]));
    out[0] = r.c0; out[1] = r.c1;
}
void orc_ef_inv(const uint64_t* a, uint64_t* out) {
    ef2 r = ef_inv(ef_make(a[0], a[1]));
    out[0] = r.c0; out[1] = r.c1;
}

void orc_set_threads(int n) { omp_set_num_threads(n); }
int orc_get_threads(void) { return omp_get_max_threads(); }
