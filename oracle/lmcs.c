/*
 * oracle/lmcs.c — stateful sponges, 2-to-1 compression and the lifted Merkle
 * tree (build, rows, prove_batch, open_batch), plain C + OpenMP.
 *
 * TEST INFRASTRUCTURE ONLY (the checker; never linked by product code).
 *
 * Restated reference functions:
 *  - StatefulSponge::absorb_into / squeeze, overwrite mode with zero-filled
 *    partial chunk: p3-miden-stateful-hasher/src/field_sponge.rs:164-187.
 *  - SerializingStatefulSponge (field -> canonical u64 lanes):
 *    serializing_sponge.rs:72-86; alignment lcm rule :166-203.
 *  - TruncatedPermutation<Perm,2,4,12> compression (parent = perm(l||r||0)[0..4])
 *    and CompressionFunctionFromHasher<PaddingFreeSponge<KeccakF,25,17,4>,2,4>
 *    (configs.rs:199-202,313-316; p3-symmetric, parity unpinned).
 *  - LiftedMerkleTree::build_with_alignment :244-310 with
 *    build_leaf_states_upsampled :385-437 (duplicate states when the height
 *    grows), absorb_matrix :447-483, compress_uniform :494-539.
 *  - rows :127-138, prove_batch :157-224 (sorted unique indices, rows padded to
 *    alignment, then missing siblings level by level, left to right).
 *  - LmcsConfig::open_batch: p3-miden-lmcs/src/lmcs.rs:167-288.
 */
#include <stdlib.h>
#include <string.h>
#include "oracle.h"
#include "tx_internal.h"

/* Mock permutation of the reference's sponge tests
 * (p3-miden-stateful-hasher/src/testing.rs:21-52): every lane becomes
 * sum_i state[i] * (i + 1) with wrapping u64 arithmetic. */
static size_t mock_width = 4, mock_rate = 2;
void orc_mock_config(size_t width, size_t rate) { mock_width = width; mock_rate = rate; }

size_t orc_hash_width(int kind) { return kind == ORC_HASH_P2 ? 12 : kind == ORC_HASH_KECCAK ? 25 : mock_width; }
size_t orc_hash_rate(int kind) { return kind == ORC_HASH_P2 ? 8 : kind == ORC_HASH_KECCAK ? 17 : mock_rate; }
/* compute_field_alignment (p3-miden-stateful-hasher/src/serializing_sponge.rs:166-203):
 * the fewest field elements that fill whole sponge chunks of rate words of
 * word_bytes each: inner = word_bytes * rate, alignment = inner / gcd(field_bytes, inner). */
size_t orc_field_alignment(size_t field_bytes, size_t word_bytes, size_t rate) {
    size_t inner = word_bytes * rate, a = field_bytes, b = inner;
    while (b) { size_t t = a % b; a = b; b = t; }
    return inner / a;
}

size_t orc_hash_alignment(int kind) {
    /* P2: RATE (field sponge).  Keccak: 8-byte Goldilocks serialised to u64 lanes, rate 17.  Mock: RATE. */
    return kind == ORC_HASH_P2 ? 8 : kind == ORC_HASH_KECCAK ? orc_field_alignment(8, 8, 17) : mock_rate;
}

static void permute(int kind, uint64_t* s) {
    if (kind == ORC_HASH_P2) orc_p2_permute(s);
    else if (kind == ORC_HASH_KECCAK) orc_keccakf(s);
    else {
        uint64_t acc = 0;
        for (size_t i = 0; i < mock_width; i++) acc += s[i] * (uint64_t)(i + 1);
        for (size_t i = 0; i < mock_width; i++) s[i] = acc;
    }
}

void orc_sponge_absorb(int kind, uint64_t* state, const uint64_t* in, size_t len) {
    size_t rate = orc_hash_rate(kind);
    size_t pos = 0;
    for (;;) {
        size_t i;
        for (i = 0; i < rate; i++) {
            if (pos < len) {
                state[i] = in[pos++];
            } else {
                if (i != 0) {
                    for (size_t j = i; j < rate; j++) state[j] = 0;
                    permute(kind, state);
                }
                return;
            }
        }
        permute(kind, state);
    }
}

void orc_compress(int kind, const uint64_t* left, const uint64_t* right, uint64_t* out) {
    uint64_t s[64];
    memset(s, 0, sizeof(s));
    memcpy(s, left, 4 * sizeof(uint64_t));
    memcpy(s + 4, right, 4 * sizeof(uint64_t));
    permute(kind, s);
    memcpy(out, s, 4 * sizeof(uint64_t));
}

void orc_hash_rows(int kind, const uint64_t* elems, size_t n, uint64_t* out4) {
    uint64_t s[64];
    memset(s, 0, sizeof(s));
    orc_sponge_absorb(kind, s, elems, n);
    memcpy(out4, s, 4 * sizeof(uint64_t));
}

static size_t aligned_len(size_t len, size_t alignment) {
    if (alignment <= 1) return len;
    return (len + alignment - 1) / alignment * alignment;
}

orc_tree* orc_lmcs_build(int kind, const uint64_t* const* mats, const size_t* heights,
                         const size_t* widths, size_t k, size_t alignment) {
    return orc_lmcs_build_salted(kind, mats, heights, widths, k, alignment, NULL, 0);
}

orc_tree* orc_lmcs_build_salted(int kind, const uint64_t* const* mats, const size_t* heights,
                                const size_t* widths, size_t k, size_t alignment, const uint64_t* salt,
                                size_t salt_w) {
    if (k == 0) return NULL;
    if (salt && salt_w == 0) return NULL;
    for (size_t i = 0; i < k; i++) {
        if (heights[i] == 0 || (heights[i] & (heights[i] - 1))) return NULL;
        if (i && heights[i] < heights[i - 1]) return NULL;
    }
    orc_tree* t = (orc_tree*)calloc(1, sizeof(orc_tree));
    t->kind = kind;
    t->k = k;
    t->heights = (size_t*)malloc(k * sizeof(size_t));
    t->widths = (size_t*)malloc(k * sizeof(size_t));
    t->mats = (uint64_t**)malloc(k * sizeof(uint64_t*));
    for (size_t i = 0; i < k; i++) {
        t->heights[i] = heights[i];
        t->widths[i] = widths[i];
        size_t bytes = heights[i] * widths[i] * sizeof(uint64_t);
        t->mats[i] = (uint64_t*)malloc(bytes ? bytes : 1);
        if (bytes) memcpy(t->mats[i], mats[i], bytes);
    }
    t->alignment = alignment < 1 ? 1 : alignment;
    size_t H = heights[k - 1];
    t->height = H;
    t->depth = log2_exact(H);
    size_t W = orc_hash_width(kind);

    /* build_leaf_states_upsampled */
    uint64_t* states = (uint64_t*)calloc(H * W, sizeof(uint64_t));
    uint64_t* scratch = (uint64_t*)calloc(H * W, sizeof(uint64_t));
    size_t active = heights[0];
    for (size_t m = 0; m < k; m++) {
        size_t h = heights[m];
        if (h > active) {
            size_t sf = h / active;
            #pragma omp parallel for
            for (size_t i = 0; i < active; i++)
                for (size_t j = 0; j < sf; j++)
                    memcpy(scratch + (i * sf + j) * W, states + i * W, W * sizeof(uint64_t));
            uint64_t* tmp = states; states = scratch; scratch = tmp;
        }
        const uint64_t* mat = t->mats[m];
        size_t w = widths[m];
        #pragma omp parallel for schedule(static)
        for (size_t r = 0; r < h; r++) orc_sponge_absorb(kind, states + r * W, mat + r * w, w);
        active = h;
    }
    /* hiding: absorb_matrix of the salt after the rows (lifted_tree.rs:271-279) */
    if (salt) {
        t->salt_w = salt_w;
        t->salt = (uint64_t*)malloc(H * salt_w * sizeof(uint64_t));
        memcpy(t->salt, salt, H * salt_w * sizeof(uint64_t));
        #pragma omp parallel for schedule(static)
        for (size_t r = 0; r < H; r++) orc_sponge_absorb(kind, states + r * W, t->salt + r * salt_w, salt_w);
    }
    t->layers = (uint64_t**)malloc((t->depth + 1) * sizeof(uint64_t*));
    t->layers[0] = (uint64_t*)malloc(H * 4 * sizeof(uint64_t));
    #pragma omp parallel for
    for (size_t r = 0; r < H; r++) memcpy(t->layers[0] + 4 * r, states + r * W, 4 * sizeof(uint64_t));
    free(states);
    free(scratch);
    /* compress_uniform until one digest remains */
    for (unsigned l = 1; l <= t->depth; l++) {
        size_t len = H >> l;
        uint64_t* prev = t->layers[l - 1];
        uint64_t* next = (uint64_t*)malloc(len * 4 * sizeof(uint64_t));
        #pragma omp parallel for if (len > 1024)
        for (size_t i = 0; i < len; i++) orc_compress(kind, prev + 8 * i, prev + 8 * i + 4, next + 4 * i);
        t->layers[l] = next;
    }
    return t;
}

size_t orc_tree_salt(const orc_tree* t, size_t index, uint64_t* out) {
    if (!t->salt) return 0;
    if (out) memcpy(out, t->salt + index * t->salt_w, t->salt_w * sizeof(uint64_t));
    return t->salt_w;
}

void orc_tree_free(orc_tree* t) {
    if (!t) return;
    free(t->salt);
    for (size_t i = 0; i < t->k; i++) free(t->mats[i]);
    free(t->mats);
    free(t->heights);
    free(t->widths);
    for (unsigned l = 0; l <= t->depth; l++) free(t->layers[l]);
    free(t->layers);
    free(t);
}

void orc_tree_root(const orc_tree* t, uint64_t* out4) { memcpy(out4, t->layers[t->depth], 32); }
size_t orc_tree_height(const orc_tree* t) { return t->height; }
void orc_tree_layer(const orc_tree* t, unsigned layer, uint64_t* out) {
    memcpy(out, t->layers[layer], (t->height >> layer) * 32);
}

/* Upsampled rows for leaf `index`, each padded to the tree alignment. Returns count. */
size_t orc_tree_rows(const orc_tree* t, size_t index, uint64_t* out) {
    size_t pos = 0;
    for (size_t m = 0; m < t->k; m++) {
        unsigned ls = log2_exact(t->height / t->heights[m]);
        size_t row = index >> ls;
        size_t w = t->widths[m];
        size_t aw = aligned_len(w, t->alignment);
        if (out) {
            memcpy(out + pos, t->mats[m] + row * w, w * sizeof(uint64_t));
            for (size_t j = w; j < aw; j++) out[pos + j] = 0;
        }
        pos += aw;
    }
    return pos;
}

static int cmp_size(const void* a, const void* b) {
    size_t x = *(const size_t*)a, y = *(const size_t*)b;
    return x < y ? -1 : x > y;
}

static size_t sort_unique(size_t* v, size_t n) {
    if (!n) return 0;
    qsort(v, n, sizeof(size_t), cmp_size);
    size_t u = 1;
    for (size_t i = 1; i < n; i++) if (v[i] != v[u - 1]) v[u++] = v[i];
    return u;
}

static int contains_sorted(const size_t* v, size_t n, size_t x) {
    size_t lo = 0, hi = n;
    while (lo < hi) {
        size_t mid = (lo + hi) / 2;
        if (v[mid] < x) lo = mid + 1; else hi = mid;
    }
    return lo < n && v[lo] == x;
}

void orc_tree_prove_batch(const orc_tree* t, const size_t* indices, size_t n, orc_tx* tx) {
    size_t* known = (size_t*)malloc((n ? n : 1) * sizeof(size_t));
    memcpy(known, indices, n * sizeof(size_t));
    size_t nk = sort_unique(known, n);
    size_t rowlen = orc_tree_rows(t, 0, NULL);
    uint64_t* row = (uint64_t*)malloc((rowlen ? rowlen : 1) * sizeof(uint64_t));
    for (size_t q = 0; q < nk; q++) {
        orc_tree_rows(t, known[q], row);
        /* one hint per matrix row, in leaf order */
        size_t pos = 0;
        for (size_t m = 0; m < t->k; m++) {
            size_t aw = aligned_len(t->widths[m], t->alignment);
            orc_tx_hint_fields(tx, row + pos, aw);
            pos += aw;
        }
        /* then the salt, as it is (lifted_tree.rs:187-192) */
        if (t->salt) orc_tx_hint_fields(tx, t->salt + known[q] * t->salt_w, t->salt_w);
    }
    size_t* parents = (size_t*)malloc((nk ? nk : 1) * sizeof(size_t));
    for (unsigned layer = 0; layer < t->depth; layer++) {
        size_t np = 0;
        for (size_t q = 0; q < nk; q++) {
            size_t pos = known[q], parent = pos / 2;
            if (np && parents[np - 1] == parent) continue;
            parents[np++] = parent;
            size_t left = parent * 2, right = left + 1;
            int hl = contains_sorted(known, nk, left), hr = contains_sorted(known, nk, right);
            if (hl && !hr) orc_tx_hint_commitment(tx, t->layers[layer] + 4 * right);
            else if (!hl && hr) orc_tx_hint_commitment(tx, t->layers[layer] + 4 * left);
        }
        memcpy(known, parents, np * sizeof(size_t));
        nk = np;
    }
    free(parents);
    free(known);
    free(row);
}

int orc_lmcs_open_batch(int kind, const uint64_t* root, const size_t* widths, size_t nw,
                        unsigned log_max_height, const size_t* indices, size_t n, orc_tx* tx,
                        uint64_t* rows_out) {
    return orc_lmcs_open_batch_salted(kind, root, widths, nw, log_max_height, indices, n, 0, tx, rows_out);
}

int orc_lmcs_open_batch_salted(int kind, const uint64_t* root, const size_t* widths, size_t nw,
                               unsigned log_max_height, const size_t* indices, size_t n, size_t salt_w, orc_tx* tx,
                               uint64_t* rows_out) {
    size_t max_h = (size_t)1 << log_max_height;
    size_t* uniq = (size_t*)malloc((n ? n : 1) * sizeof(size_t));
    memcpy(uniq, indices, n * sizeof(size_t));
    size_t nu = sort_unique(uniq, n);
    if (nu == 0) { free(uniq); return 1; }
    size_t total = 0;
    for (size_t i = 0; i < nw; i++) total += widths[i];
    size_t* pos = (size_t*)malloc(nu * sizeof(size_t));
    uint64_t* hashes = (uint64_t*)malloc(nu * 4 * sizeof(uint64_t));
    uint64_t* elems = (uint64_t*)malloc((total ? total : 1) * sizeof(uint64_t));
    int rc = 0;
    for (size_t q = 0; q < nu; q++) {
        if (uniq[q] >= max_h) { rc = 1; goto done; }
        if (orc_tx_receive_hint_fields(tx, total, elems)) { rc = 3; goto done; }
        if (rows_out) memcpy(rows_out + q * total, elems, total * sizeof(uint64_t));
        uint64_t st[64];
        memset(st, 0, sizeof(st));
        size_t off = 0;
        for (size_t i = 0; i < nw; i++) { orc_sponge_absorb(kind, st, elems + off, widths[i]); off += widths[i]; }
        if (salt_w) {
            /* hash(rows ... salt) (lmcs.rs:198-208) */
            uint64_t sl[64];
            if (salt_w > 64 || orc_tx_receive_hint_fields(tx, salt_w, sl)) { rc = 3; goto done; }
            orc_sponge_absorb(kind, st, sl, salt_w);
        }
        memcpy(hashes + 4 * q, st, 32);
        pos[q] = uniq[q];
    }
    {
        size_t nc = nu;
        for (unsigned level = 0; level < log_max_height; level++) {
            size_t np = 0;
            size_t i = 0;
            while (i < nc) {
                size_t p = pos[i];
                uint64_t sib[4];
                size_t sp = p ^ 1;
                const uint64_t* mine = hashes + 4 * i;
                int have = (i + 1 < nc && pos[i + 1] == sp);
                if (have) {
                    memcpy(sib, hashes + 4 * (i + 1), 32);
                } else if (orc_tx_receive_hint_commitment(tx, sib)) {
                    rc = 1;
                    goto done;
                }
                uint64_t par[4];
                if ((p & 1) == 0) orc_compress(kind, mine, sib, par);
                else orc_compress(kind, sib, mine, par);
                pos[np] = p >> 1;
                memcpy(hashes + 4 * np, par, 32);
                np++;
                i += have ? 2 : 1;
            }
            nc = np;
        }
        if (!(nc == 1 && pos[0] == 0)) { rc = 1; goto done; }
        if (memcmp(hashes, root, 32) != 0) rc = 2;
    }
done:
    free(uniq);
    free(pos);
    free(hashes);
    free(elems);
    return rc;
}
