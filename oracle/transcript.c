/*
 * oracle/transcript.c — Fiat-Shamir channel: challenger + recorded streams.
 *
 * TEST INFRASTRUCTURE ONLY (the checker; never linked by product code).
 *
 * Restated reference behaviour:
 *  - ProverTranscript (p3-miden-transcript/src/prover.rs:15-145): send_* append
 *    to the fields / commitments stream and observe; hint_* append only;
 *    grind pushes the witness to the fields stream without observing it.
 *  - VerifierTranscript (verifier.rs:16-59 and the channel impl): receive_*
 *    consume the streams in order (observed variants observe); grind reads the
 *    witness and runs check_witness; finalize rejects trailing data.
 *  - Channel::sample_algebra_element (channel.rs:229-231): coefficient 0 then 1.
 *  - DuplexChallenger<Goldilocks, Poseidon2, 12, 8> (configs.rs:208; p3-challenger
 *    0.5.0, parity unpinned): observe overwrites an input buffer that is
 *    duplexed at RATE; sample duplexes when input is pending or output is
 *    exhausted and pops from the END of the 8-element output buffer;
 *    sample_bits masks the canonical value; grind with 0 bits returns witness 0
 *    without observing, otherwise the smallest witness i with
 *    observe(i); sample_bits(bits) == 0 (the serial search order).
 *  - Keccak configuration (no transcript exists in the reference for it,
 *    build-defined): SerializingChallenger64<Goldilocks, HashChallenger<u8,
 *    Keccak256, 32>> — fields observed as 8 little-endian bytes, u64 digest
 *    words likewise; a sample pops bytes from a 32-byte Keccak-256 output that
 *    is recomputed over the pending input (chained) when exhausted.
 */
#include <stdlib.h>
#include <string.h>
#include "oracle.h"
#include "tx_internal.h"

/* ---------- challenger ---------- */

static void duplex(orc_challenger* c) {
    for (int i = 0; i < c->in_len; i++) c->st[i] = c->in[i];
    c->in_len = 0;
    orc_p2_permute(c->st);
    for (int i = 0; i < 8; i++) c->out[i] = c->st[i];
    c->out_len = 8;
}

static void kin_push(orc_challenger* c, uint8_t b) {
    if (c->kin_len == c->kin_cap) {
        c->kin_cap = c->kin_cap ? 2 * c->kin_cap : 256;
        c->kin = (uint8_t*)realloc(c->kin, c->kin_cap);
    }
    c->kin[c->kin_len++] = b;
}

static void k_observe_u64(orc_challenger* c, uint64_t v) {
    c->kout_len = 0;
    for (int i = 0; i < 8; i++) kin_push(c, (uint8_t)(v >> (8 * i)));
}

static uint8_t k_sample_byte(orc_challenger* c) {
    if (c->kout_len == 0) {
        uint8_t h[32];
        orc_keccak256(c->kin, c->kin_len, h);
        memcpy(c->kout, h, 32);
        c->kout_len = 32;
        c->kin_len = 0;
        for (int i = 0; i < 32; i++) kin_push(c, h[i]);
    }
    return c->kout[--c->kout_len];
}

static void ch_observe(orc_challenger* c, gl v) {
    if (c->kind == ORC_CH_DUPLEX_P2) {
        c->out_len = 0;
        c->in[c->in_len++] = v;
        if (c->in_len == 8) duplex(c);
    } else {
        k_observe_u64(c, v);
    }
}

static void ch_observe_commitment(orc_challenger* c, const uint64_t* c4) {
    for (int i = 0; i < 4; i++) {
        if (c->kind == ORC_CH_DUPLEX_P2) ch_observe(c, c4[i]);
        else k_observe_u64(c, c4[i]);
    }
}

static gl ch_sample(orc_challenger* c) {
    if (c->kind == ORC_CH_DUPLEX_P2) {
        if (c->in_len > 0 || c->out_len == 0) duplex(c);
        return c->out[--c->out_len];
    }
    for (;;) {
        uint64_t v = 0;
        for (int i = 0; i < 8; i++) v |= (uint64_t)k_sample_byte(c) << (8 * i);
        if (v < GL_P) return v;
    }
}

static size_t ch_sample_bits(orc_challenger* c, unsigned bits) {
    gl v = ch_sample(c);
    return (size_t)(v & ((((uint64_t)1) << bits) - 1));
}

static void ch_copy(orc_challenger* dst, const orc_challenger* src) {
    *dst = *src;
    dst->kin = NULL;
    dst->kin_cap = 0;
    dst->kin_len = 0;
    if (src->kin_len) {
        dst->kin = (uint8_t*)malloc(src->kin_cap);
        dst->kin_cap = src->kin_cap;
        memcpy(dst->kin, src->kin, src->kin_len);
        dst->kin_len = src->kin_len;
    }
}

static int ch_check_witness(orc_challenger* c, unsigned bits, gl w) {
    if (bits == 0) return 1;
    ch_observe(c, w);
    return ch_sample_bits(c, bits) == 0;
}

/* ---------- transcript ---------- */

orc_tx* orc_tx_new(int kind) {
    orc_tx* t = (orc_tx*)calloc(1, sizeof(orc_tx));
    t->ch.kind = kind;
    return t;
}

orc_tx* orc_tx_clone(const orc_tx* src) {
    orc_tx* t = (orc_tx*)calloc(1, sizeof(orc_tx));
    ch_copy(&t->ch, &src->ch);
    t->nf = src->nf; t->capf = src->nf;
    t->nc = src->nc; t->capc = src->nc;
    t->fields = (uint64_t*)malloc((src->nf ? src->nf : 1) * 8);
    t->comms = (uint64_t*)malloc((src->nc ? src->nc : 1) * 32);
    memcpy(t->fields, src->fields, src->nf * 8);
    memcpy(t->comms, src->comms, src->nc * 32);
    t->rf = src->rf; t->rc = src->rc;
    return t;
}

void orc_tx_free(orc_tx* t) {
    if (!t) return;
    free(t->ch.kin);
    free(t->fields);
    free(t->comms);
    free(t);
}

static void push_fields(orc_tx* t, const uint64_t* v, size_t n) {
    if (t->nf + n > t->capf) {
        size_t cap = t->capf ? t->capf : 1024;
        while (cap < t->nf + n) cap *= 2;
        t->fields = (uint64_t*)realloc(t->fields, cap * 8);
        t->capf = cap;
    }
    memcpy(t->fields + t->nf, v, n * 8);
    t->nf += n;
}

static void push_comm(orc_tx* t, const uint64_t* c4) {
    if (t->nc + 1 > t->capc) {
        size_t cap = t->capc ? 2 * t->capc : 256;
        t->comms = (uint64_t*)realloc(t->comms, cap * 32);
        t->capc = cap;
    }
    memcpy(t->comms + 4 * t->nc, c4, 32);
    t->nc++;
}

void orc_tx_observe(orc_tx* t, uint64_t v) { ch_observe(&t->ch, v); }
void orc_tx_observe_commitment(orc_tx* t, const uint64_t* c4) { ch_observe_commitment(&t->ch, c4); }
uint64_t orc_tx_sample(orc_tx* t) { return ch_sample(&t->ch); }
void orc_tx_sample_ef(orc_tx* t, uint64_t* out2) {
    out2[0] = ch_sample(&t->ch);
    out2[1] = ch_sample(&t->ch);
}
size_t orc_tx_sample_bits(orc_tx* t, unsigned bits) { return ch_sample_bits(&t->ch, bits); }

void orc_tx_send_fields(orc_tx* t, const uint64_t* v, size_t n) {
    push_fields(t, v, n);
    for (size_t i = 0; i < n; i++) ch_observe(&t->ch, v[i]);
}
void orc_tx_send_commitment(orc_tx* t, const uint64_t* c4) {
    push_comm(t, c4);
    ch_observe_commitment(&t->ch, c4);
}
void orc_tx_hint_fields(orc_tx* t, const uint64_t* v, size_t n) { push_fields(t, v, n); }
void orc_tx_hint_commitment(orc_tx* t, const uint64_t* c4) { push_comm(t, c4); }

uint64_t orc_tx_grind(orc_tx* t, unsigned bits) {
    gl w = 0;
    if (bits > 0) {
        for (gl i = 0;; i++) {
            orc_challenger probe;
            ch_copy(&probe, &t->ch);
            int ok = ch_check_witness(&probe, bits, i);
            free(probe.kin);
            if (ok) { w = i; break; }
        }
        ch_check_witness(&t->ch, bits, w);
    }
    push_fields(t, &w, 1);
    return w;
}

size_t orc_tx_nfields(const orc_tx* t) { return t->nf; }
size_t orc_tx_ncomms(const orc_tx* t) { return t->nc; }
void orc_tx_get(const orc_tx* t, uint64_t* fields, uint64_t* comms) {
    if (fields) memcpy(fields, t->fields, t->nf * 8);
    if (comms) memcpy(comms, t->comms, t->nc * 32);
}

void orc_tx_load(orc_tx* t, const uint64_t* fields, size_t nf, const uint64_t* comms, size_t nc) {
    t->nf = 0; t->nc = 0; t->rf = 0; t->rc = 0;
    push_fields(t, fields, nf);
    for (size_t i = 0; i < nc; i++) push_comm(t, comms + 4 * i);
}

int orc_tx_receive_hint_fields(orc_tx* t, size_t n, uint64_t* out) {
    if (t->rf + n > t->nf) return 1;
    memcpy(out, t->fields + t->rf, n * 8);
    t->rf += n;
    return 0;
}
int orc_tx_receive_fields(orc_tx* t, size_t n, uint64_t* out) {
    if (orc_tx_receive_hint_fields(t, n, out)) return 1;
    for (size_t i = 0; i < n; i++) ch_observe(&t->ch, out[i]);
    return 0;
}
int orc_tx_receive_hint_commitment(orc_tx* t, uint64_t* out4) {
    if (t->rc + 1 > t->nc) return 1;
    memcpy(out4, t->comms + 4 * t->rc, 32);
    t->rc++;
    return 0;
}
int orc_tx_receive_commitment(orc_tx* t, uint64_t* out4) {
    if (orc_tx_receive_hint_commitment(t, out4)) return 1;
    ch_observe_commitment(&t->ch, out4);
    return 0;
}
int orc_tx_check_witness(orc_tx* t, unsigned bits, uint64_t w) { return ch_check_witness(&t->ch, bits, w); }
int orc_tx_check_grind(orc_tx* t, unsigned bits) {
    uint64_t w;
    if (orc_tx_receive_hint_fields(t, 1, &w)) return 1;
    return ch_check_witness(&t->ch, bits, w) ? 0 : 2;
}
int orc_tx_is_empty(const orc_tx* t) { return t->rf == t->nf && t->rc == t->nc; }
