/*
 * oracle/prims.c — permutations and the seeded generator, plain C.
 *
 * TEST INFRASTRUCTURE ONLY (the checker; never linked by product code).
 *
 *  - SmallRng restatement: rand 0.10 `SmallRng` on 64-bit targets is
 *    Xoshiro256++ seeded through SplitMix64 (`seed_from_u64`), and the
 *    Goldilocks `StandardUniform` distribution rejects u64 >= p.  Used by
 *    `RowMajorMatrix::rand` (p3-miden-dev-utils/src/matrix.rs:98-116,
 *    generate_matrices_from_specs :109-134) and by Poseidon2 constant
 *    generation (p3-miden-dev-utils/src/configs.rs:215-218, seed 2025 from
 *    p3-miden-dev-utils/src/fixtures.rs:15).  Parity unpinned: rand 0.10 is
 *    not vendored in the reference, so this follows the published algorithm.
 *  - Poseidon2-Goldilocks width 12 as `Poseidon2Goldilocks::<12>::new_from_rng_128`
 *    (configs.rs:126-130,394-407): x^7 S-box, 8 external rounds (4+4),
 *    22 internal rounds, external layer = M4 blocks + column sums, internal
 *    layer = diag(d) + all-ones with the published MATRIX_DIAG_12_GOLDILOCKS.
 *    Constants: 4 x [F;12] initial, 4 x [F;12] terminal, 22 x F internal,
 *    in that sampling order.  Parity unpinned (p3-poseidon2 / p3-goldilocks
 *    0.5.0 not vendored); structure pinned by tests.
 *  - Keccak-f[1600] on 25 u64 lanes (p3-keccak `KeccakF`, used by the Keccak
 *    LMCS in p3-miden-lmcs/benches/utils.rs:27,46-53); pinned against
 *    hashlib SHA3 in the tests.
 *  - Keccak-256 (original 0x01 padding, rate 136) for the build-defined
 *    Keccak transcript challenger.
 */
#include <string.h>
#include "oracle.h"

/* ---------------- SmallRng (Xoshiro256++ / SplitMix64) ---------------- */

static inline uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

void orc_rng_seed(orc_rng* r, uint64_t seed) {
    uint64_t st = seed;
    for (int i = 0; i < 4; i++) {
        st += 0x9e3779b97f4a7c15ULL;
        uint64_t z = st;
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
        r->s[i] = z ^ (z >> 31);
    }
}

uint64_t orc_rng_next(orc_rng* r) {
    uint64_t* s = r->s;
    uint64_t result = rotl64(s[0] + s[3], 23) + s[0];
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl64(s[3], 45);
    return result;
}

gl orc_rng_gl(orc_rng* r) {
    for (;;) {
        uint64_t x = orc_rng_next(r);
        if (x < GL_P) return x;
    }
}

void orc_rand_fill(uint64_t seed, size_t n, uint64_t* out) {
    orc_rng r;
    orc_rng_seed(&r, seed);
    for (size_t i = 0; i < n; i++) out[i] = orc_rng_gl(&r);
}

void orc_rand_fill_rng(orc_rng* r, size_t n, uint64_t* out) {
    for (size_t i = 0; i < n; i++) out[i] = orc_rng_gl(r);
}

/* ---------------- Poseidon2-Goldilocks-12 ---------------- */

static const gl P2_DIAG[12] = {
    0xc3b6c08e23ba9300ULL, 0xd84b5de94a324fb6ULL, 0x0d0c371c5b35b84fULL,
    0x7964f570e7188037ULL, 0x5daf18bbd996604bULL, 0x6743bc47b9595257ULL,
    0x5528b9362c59bb70ULL, 0xac45e25b7127b68bULL, 0xa2077d7dfbb606b5ULL,
    0xf3faac6faee378aeULL, 0x0c6388b51545e883ULL, 0xd27dbb6944917b60ULL,
};

static gl P2_RC_INIT[4][12];
static gl P2_RC_TERM[4][12];
static gl P2_RC_INT[22];
static int p2_ready = 0;

void orc_p2_init_constants(uint64_t seed) {
    orc_rng r;
    orc_rng_seed(&r, seed);
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 12; j++) P2_RC_INIT[i][j] = orc_rng_gl(&r);
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 12; j++) P2_RC_TERM[i][j] = orc_rng_gl(&r);
    for (int i = 0; i < 22; i++) P2_RC_INT[i] = orc_rng_gl(&r);
    p2_ready = 1;
}

static void p2_ensure(void) {
    if (!p2_ready) orc_p2_init_constants(2025);
}

void orc_p2_constants(uint64_t* out) {
    p2_ensure();
    memcpy(out, P2_RC_INIT, sizeof(P2_RC_INIT));
    memcpy(out + 48, P2_RC_TERM, sizeof(P2_RC_TERM));
    memcpy(out + 96, P2_RC_INT, sizeof(P2_RC_INT));
}

void orc_p2_diag(uint64_t* out) { memcpy(out, P2_DIAG, sizeof(P2_DIAG)); }

static inline gl sbox7(gl x) {
    gl x2 = gl_sqr(x);
    gl x3 = gl_mul(x2, x);
    gl x4 = gl_sqr(x2);
    return gl_mul(x3, x4);
}

/* M4 = [[2,3,1,1],[1,2,3,1],[1,1,2,3],[3,1,1,2]] */
void orc_p2_m4(gl* x) {
    gl a = x[0], b = x[1], c = x[2], d = x[3];
    gl ab = gl_add(a, b), cd = gl_add(c, d);
    gl abcd = gl_add(ab, cd);
    gl y0 = gl_add(gl_add(abcd, a), b);          /* 2a+2b+c+d */
    y0 = gl_add(y0, b);                            /* 2a+3b+c+d */
    gl y1 = gl_add(gl_add(abcd, b), gl_add(c, c)); /* a+2b+3c+d */
    gl y2 = gl_add(gl_add(abcd, c), gl_add(d, d)); /* a+b+2c+3d */
    gl y3 = gl_add(gl_add(abcd, d), gl_add(a, a)); /* 3a+b+c+2d */
    x[0] = y0; x[1] = y1; x[2] = y2; x[3] = y3;
}

void orc_p2_external_linear(gl* s) {
    for (int c = 0; c < 12; c += 4) orc_p2_m4(s + c);
    gl sums[4];
    for (int k = 0; k < 4; k++) sums[k] = gl_add(gl_add(s[k], s[4 + k]), s[8 + k]);
    for (int i = 0; i < 12; i++) s[i] = gl_add(s[i], sums[i & 3]);
}

void orc_p2_internal_linear(gl* s) {
    gl sum = 0;
    for (int i = 0; i < 12; i++) sum = gl_add(sum, s[i]);
    for (int i = 0; i < 12; i++) s[i] = gl_add(gl_mul(s[i], P2_DIAG[i]), sum);
}

void orc_p2_permute(uint64_t* s) {
    p2_ensure();
    orc_p2_external_linear(s);
    for (int r = 0; r < 4; r++) {
        for (int i = 0; i < 12; i++) s[i] = sbox7(gl_add(s[i], P2_RC_INIT[r][i]));
        orc_p2_external_linear(s);
    }
    for (int r = 0; r < 22; r++) {
        s[0] = sbox7(gl_add(s[0], P2_RC_INT[r]));
        orc_p2_internal_linear(s);
    }
    for (int r = 0; r < 4; r++) {
        for (int i = 0; i < 12; i++) s[i] = sbox7(gl_add(s[i], P2_RC_TERM[r][i]));
        orc_p2_external_linear(s);
    }
}

/* ---------------- Keccak-f[1600] ---------------- */

static const uint64_t KECCAK_RC[24] = {
    0x0000000000000001ULL, 0x0000000000008082ULL, 0x800000000000808aULL,
    0x8000000080008000ULL, 0x000000000000808bULL, 0x0000000080000001ULL,
    0x8000000080008081ULL, 0x8000000000008009ULL, 0x000000000000008aULL,
    0x0000000000000088ULL, 0x0000000080008009ULL, 0x000000008000000aULL,
    0x000000008000808bULL, 0x800000000000008bULL, 0x8000000000008089ULL,
    0x8000000000008003ULL, 0x8000000000008002ULL, 0x8000000000000080ULL,
    0x000000000000800aULL, 0x800000008000000aULL, 0x8000000080008081ULL,
    0x8000000000008080ULL, 0x0000000080000001ULL, 0x8000000080008008ULL,
};

/* rotation offsets r[x][y] indexed by lane x + 5y */
static const int KECCAK_ROT[25] = {
     0,  1, 62, 28, 27,
    36, 44,  6, 55, 20,
     3, 10, 43, 25, 39,
    41, 45, 15, 21,  8,
    18,  2, 61, 56, 14,
};

void orc_keccakf(uint64_t* a) {
    for (int round = 0; round < 24; round++) {
        uint64_t c[5], d[5], b[25];
        for (int x = 0; x < 5; x++) c[x] = a[x] ^ a[x + 5] ^ a[x + 10] ^ a[x + 15] ^ a[x + 20];
        for (int x = 0; x < 5; x++) d[x] = c[(x + 4) % 5] ^ rotl64(c[(x + 1) % 5], 1);
        for (int i = 0; i < 25; i++) a[i] ^= d[i % 5];
        /* rho + pi: B[y, 2x+3y] = rot(A[x,y], r[x,y]) */
        for (int x = 0; x < 5; x++)
            for (int y = 0; y < 5; y++) {
                int r = KECCAK_ROT[x + 5 * y];
                uint64_t v = a[x + 5 * y];
                b[y + 5 * ((2 * x + 3 * y) % 5)] = r ? rotl64(v, r) : v;
            }
        for (int y = 0; y < 5; y++)
            for (int x = 0; x < 5; x++)
                a[x + 5 * y] = b[x + 5 * y] ^ (~b[(x + 1) % 5 + 5 * y] & b[(x + 2) % 5 + 5 * y]);
        a[0] ^= KECCAK_RC[round];
    }
}

/* Keccak-256 with the original multi-rate padding (0x01 ... 0x80), rate 136. */
void orc_keccak256(const uint8_t* in, size_t len, uint8_t* out32) {
    uint64_t st[25];
    memset(st, 0, sizeof(st));
    const size_t rate = 136;
    uint8_t block[136];
    size_t off = 0;
    for (;;) {
        size_t take = len - off < rate ? len - off : rate;
        memset(block, 0, rate);
        memcpy(block, in + off, take);
        int last = take < rate;
        if (last) {
            block[take] ^= 0x01;
            block[rate - 1] ^= 0x80;
        }
        for (size_t i = 0; i < rate / 8; i++) {
            uint64_t lane = 0;
            for (int k = 0; k < 8; k++) lane |= (uint64_t)block[8 * i + k] << (8 * k);
            st[i] ^= lane;
        }
        orc_keccakf(st);
        off += take;
        if (last) break;
    }
    for (int i = 0; i < 32; i++) out32[i] = (uint8_t)(st[i / 8] >> (8 * (i % 8)));
}
