// Host-side Fiat-Shamir channel.
//
// ProverTranscript semantics (p3-miden-transcript/src/prover.rs:15-145):
// send_* append to the recorded stream and observe, hint_* only append,
// grind appends the witness without observing it.  The challenger is
//  - DuplexChallenger<Goldilocks, Poseidon2GL12, 12, 8> for the Poseidon2
//    configuration (p3-miden-dev-utils/src/configs.rs:208): overwrite-mode
//    input buffer duplexed at rate 8, samples popped from the end of the
//    8-word output buffer, sample_bits masks the canonical value, grind(0)
//    returns witness 0 without observing, grind(b) takes the smallest witness
//    i with observe(i); sample_bits(b) == 0;
//  - for the Keccak configuration (the reference defines no transcript for it)
//    a SerializingChallenger64 over a Keccak-256 HashChallenger: words are
//    observed as 8 little-endian bytes, samples pop bytes from a chained
//    Keccak-256 output and reject u64 >= p.
#include <cstring>

#include "internal.hpp"
#include "p2_fast.hpp"

namespace p3g {

const P2Consts& host_p2_consts() {
  static P2Consts c = [] {
    P2Consts x;
    p2_generate_constants(2025, &x);
    return x;
  }();
  return c;
}

void keccak256_host(const uint8_t* in, size_t len, uint8_t* out32) {
  uint64_t st[25] = {};
  const size_t rate = 136;
  uint8_t block[136];
  size_t off = 0;
  for (;;) {
    size_t take = len - off < rate ? len - off : rate;
    memset(block, 0, rate);
    if (take) memcpy(block, in + off, take);
    bool last = take < rate;
    if (last) {
      block[take] ^= 0x01;
      block[rate - 1] ^= 0x80;
    }
    for (size_t i = 0; i < rate / 8; i++) {
      uint64_t lane = 0;
      for (int k = 0; k < 8; k++) lane |= (uint64_t)block[8 * i + k] << (8 * k);
      st[i] ^= lane;
    }
    keccak_f(st);
    off += take;
    if (last) break;
  }
  for (int i = 0; i < 32; i++) out32[i] = (uint8_t)(st[i / 8] >> (8 * (i % 8)));
}

void Transcript::duplex() {
  for (int i = 0; i < in_len; i++) st[i] = in[i];
  in_len = 0;
  p2_permute_fast(st, &host_p2_consts());  // bit-identical to p2_permute, fewer reductions
  for (int i = 0; i < 8; i++) out[i] = st[i];
  out_len = 8;
}

void Transcript::k_observe_u64(uint64_t v) {
  kout_len = 0;
  for (int i = 0; i < 8; i++) kin.push_back((uint8_t)(v >> (8 * i)));
}

uint8_t Transcript::k_sample_byte() {
  if (kout_len == 0) {
    uint8_t h[32];
    keccak256_host(kin.data(), kin.size(), h);
    memcpy(kout, h, 32);
    kout_len = 32;
    kin.assign(h, h + 32);
  }
  return kout[--kout_len];
}

void Transcript::observe(uint64_t v) {
  if (kind == CH_DUPLEX_P2) {
    out_len = 0;
    in[in_len++] = v;
    if (in_len == 8) duplex();
  } else {
    k_observe_u64(v);
  }
}

void Transcript::observe_commitment(const uint64_t* c4) {
  for (int i = 0; i < 4; i++) observe(c4[i]);
}

uint64_t Transcript::sample() {
  if (kind == CH_DUPLEX_P2) {
    if (in_len > 0 || out_len == 0) duplex();
    return out[--out_len];
  }
  for (;;) {
    uint64_t v = 0;
    for (int i = 0; i < 8; i++) v |= (uint64_t)k_sample_byte() << (8 * i);
    if (v < P) return v;
  }
}

uint64_t Transcript::sample_bits(unsigned bits) {
  uint64_t v = sample();
  return v & ((1ull << bits) - 1);
}

void Transcript::send_fields(const uint64_t* v, size_t n) {
  fields.insert(fields.end(), v, v + n);
  for (size_t i = 0; i < n; i++) observe(v[i]);
}

void Transcript::send_commitment(const uint64_t* c4) {
  comms.insert(comms.end(), c4, c4 + 4);
  observe_commitment(c4);
}

void Transcript::hint_fields(const uint64_t* v, size_t n) { fields.insert(fields.end(), v, v + n); }
void Transcript::hint_commitment(const uint64_t* c4) { comms.insert(comms.end(), c4, c4 + 4); }

bool Transcript::check_witness(unsigned bits, uint64_t w) {
  if (bits == 0) return true;
  observe(w);
  return sample_bits(bits) == 0;
}

void Transcript::accept_witness(unsigned bits, uint64_t w) {
  if (w >= P || !check_witness(bits, w)) throw ApiError(E_TRANSCRIPT, "grinding witness fails the transcript's check");
  fields.push_back(w);
}

uint64_t Transcript::grind(unsigned bits) {
  uint64_t w = 0;
  if (bits > 0) {
    for (uint64_t i = 0;; i++) {
      Transcript probe;
      probe.kind = kind;
      memcpy(probe.st, st, sizeof(st));
      memcpy(probe.in, in, sizeof(in));
      probe.in_len = in_len;
      memcpy(probe.out, out, sizeof(out));
      probe.out_len = out_len;
      probe.kin = kin;
      memcpy(probe.kout, kout, sizeof(kout));
      probe.kout_len = kout_len;
      if (probe.check_witness(bits, i)) {
        w = i;
        break;
      }
    }
    check_witness(bits, w);
  }
  fields.push_back(w);
  return w;
}

}  // namespace p3g
