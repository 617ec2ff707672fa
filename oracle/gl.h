/*
 * oracle/gl.h — Goldilocks field and its quadratic extension, plain C.
 *
 * TEST INFRASTRUCTURE ONLY.  Everything under oracle/ is the CPU restatement
 * of the reference prover's algorithm that the tests, the smoke check and the
 * benchmark's CPU-baseline leg use as the checker.  Product code never links,
 * loads or calls it.
 *
 * Field: p = 2^64 - 2^32 + 1, canonical representatives in [0, p).
 * Extension: EF = F[x]/(x^2 - 7) (the degree-2 BinomialExtensionField the
 * reference instantiates in p3-miden-dev-utils/src/configs.rs:184-190,
 * goldilocks_poseidon2 at :394-407).
 * Two-adic generators: omega_32 = 1753635133440165772 = 7^((p-1)/2^32),
 * omega_k = omega_32^(2^(32-k)); multiplicative generator 7
 * (F::GENERATOR, used as coset shift in p3-miden-lifted-fri/src/utils.rs:101-106
 * and p3-miden-lifted-stark/src/coset.rs:132-134).
 */
#ifndef ORACLE_GL_H
#define ORACLE_GL_H

#include <stdint.h>
#include <stddef.h>

typedef uint64_t gl;
typedef struct { gl c0, c1; } ef2;

#define GL_P   0xFFFFFFFF00000001ULL
#define GL_EPS 0xFFFFFFFFULL
#define GL_GENERATOR 7ULL
#define GL_W 7ULL
#define GL_OMEGA32 1753635133440165772ULL
#define GL_TWO_ADICITY 32

static inline gl gl_canon(uint64_t x) { return x >= GL_P ? x - GL_P : x; }

static inline gl gl_add(gl a, gl b) {
    gl s = a + b;
    if (s < a || s >= GL_P) s -= GL_P;
    return s;
}

static inline gl gl_sub(gl a, gl b) {
    gl d = a - b;
    if (a < b) d += GL_P;
    return d;
}

static inline gl gl_neg(gl a) { return a ? GL_P - a : 0; }

/* Reduce a 128-bit value: x = hh*2^96 + hl*2^64 + lo with 2^64 = eps, 2^96 = -1. */
static inline gl gl_reduce128(unsigned __int128 x) {
    uint64_t lo = (uint64_t)x, hi = (uint64_t)(x >> 64);
    uint64_t hh = hi >> 32, hl = hi & GL_EPS;
    uint64_t t0 = lo - hh;
    if (lo < hh) t0 -= GL_EPS;
    uint64_t t1 = hl * GL_EPS;
    uint64_t r = t0 + t1;
    if (r < t1) r += GL_EPS;
    return gl_canon(r);
}

static inline gl gl_mul(gl a, gl b) { return gl_reduce128((unsigned __int128)a * b); }

static inline gl gl_sqr(gl a) { return gl_mul(a, a); }

static inline gl gl_pow(gl a, uint64_t e) {
    gl r = 1;
    while (e) {
        if (e & 1) r = gl_mul(r, a);
        a = gl_mul(a, a);
        e >>= 1;
    }
    return r;
}

static inline gl gl_inv(gl a) { return gl_pow(a, GL_P - 2); }

/* 2^-k mod p. */
static inline gl gl_inv_pow2(unsigned k) { return gl_inv(gl_pow(2, k)); }

static inline gl gl_two_adic_generator(unsigned bits) {
    gl g = GL_OMEGA32;
    for (unsigned i = bits; i < GL_TWO_ADICITY; i++) g = gl_mul(g, g);
    return g;
}

/* ---- EF2 ---- */
static inline ef2 ef_make(gl a, gl b) { ef2 r = {a, b}; return r; }
static inline ef2 ef_from(gl a) { ef2 r = {a, 0}; return r; }
static inline ef2 ef_add(ef2 a, ef2 b) { return ef_make(gl_add(a.c0, b.c0), gl_add(a.c1, b.c1)); }
static inline ef2 ef_sub(ef2 a, ef2 b) { return ef_make(gl_sub(a.c0, b.c0), gl_sub(a.c1, b.c1)); }
static inline ef2 ef_neg(ef2 a) { return ef_make(gl_neg(a.c0), gl_neg(a.c1)); }
static inline ef2 ef_mul(ef2 a, ef2 b) {
    gl t = gl_mul(gl_mul(a.c1, b.c1), GL_W);
    return ef_make(gl_add(gl_mul(a.c0, b.c0), t),
                   gl_add(gl_mul(a.c0, b.c1), gl_mul(a.c1, b.c0)));
}
static inline ef2 ef_mul_base(ef2 a, gl b) { return ef_make(gl_mul(a.c0, b), gl_mul(a.c1, b)); }
static inline ef2 ef_sqr(ef2 a) { return ef_mul(a, a); }
static inline int ef_eq(ef2 a, ef2 b) { return a.c0 == b.c0 && a.c1 == b.c1; }
static inline int ef_is_zero(ef2 a) { return a.c0 == 0 && a.c1 == 0; }
static inline ef2 ef_inv(ef2 a) {
    /* (a0 + a1 x)^-1 = (a0 - a1 x) / (a0^2 - 7 a1^2) */
    gl norm = gl_sub(gl_sqr(a.c0), gl_mul(GL_W, gl_sqr(a.c1)));
    gl ni = gl_inv(norm);
    return ef_make(gl_mul(a.c0, ni), gl_neg(gl_mul(a.c1, ni)));
}
static inline ef2 ef_pow(ef2 a, uint64_t e) {
    ef2 r = ef_from(1);
    while (e) {
        if (e & 1) r = ef_mul(r, a);
        a = ef_sqr(a);
        e >>= 1;
    }
    return r;
}
static inline ef2 ef_exp_power_of_2(ef2 a, unsigned k) {
    for (unsigned i = 0; i < k; i++) a = ef_sqr(a);
    return a;
}

static inline size_t bitrev_idx(size_t x, unsigned bits) {
    size_t r = 0;
    for (unsigned i = 0; i < bits; i++) { r = (r << 1) | (x & 1); x >>= 1; }
    return r;
}

static inline unsigned log2_exact(size_t n) {
    unsigned l = 0;
    while (((size_t)1 << l) < n) l++;
    return l;
}

#endif
