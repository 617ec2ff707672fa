// Poseidon2-Goldilocks width 12 and Keccak-f[1600], for host and gfx950 device code.
//
// Poseidon2 is the permutation the reference configures as
// `Poseidon2Goldilocks::<12>::new_from_rng_128(SmallRng::seed_from_u64(2025))`
// (p3-miden-dev-utils/src/configs.rs:126-130,215-218,394-407): x^7 S-box,
// 4 + 4 external rounds with the M4-based light MDS layer, 22 internal rounds
// with the diag(d) + ones matrix.  The round constants are generated on the
// host (p2_generate_constants) from the same seeded generator and handed to
// kernels by pointer, so every lane reads them through scalar loads.
//
// Keccak-f[1600] is the permutation behind p3-keccak's `KeccakF` used by the
// Keccak LMCS of p3-miden-lmcs/benches/utils.rs:27,46-53.
#pragma once
#include "gl.hpp"

namespace p3g {

struct P2Consts {
  uint64_t rc_init[4][12];
  uint64_t rc_term[4][12];
  uint64_t rc_int[22];
  uint64_t diag[12];
};

// Seeded generator: SplitMix64 seeding + Xoshiro256++ (rand's SmallRng on 64-bit
// targets), Goldilocks samples by rejection of u64 >= p.
struct SmallRng {
  uint64_t s[4];
  explicit SmallRng(uint64_t seed) {
    uint64_t st = seed;
    for (int i = 0; i < 4; i++) {
      st += 0x9e3779b97f4a7c15ull;
      uint64_t z = st;
      z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
      z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
      s[i] = z ^ (z >> 31);
    }
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next_u64() {
    uint64_t result = rotl(s[0] + s[3], 23) + s[0];
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return result;
  }
  uint64_t next_gl() {
    for (;;) {
      uint64_t x = next_u64();
      if (x < P) return x;
    }
  }
};

inline void p2_generate_constants(uint64_t seed, P2Consts* c) {
  static const uint64_t diag[12] = {
      0xc3b6c08e23ba9300ull, 0xd84b5de94a324fb6ull, 0x0d0c371c5b35b84full, 0x7964f570e7188037ull,
      0x5daf18bbd996604bull, 0x6743bc47b9595257ull, 0x5528b9362c59bb70ull, 0xac45e25b7127b68bull,
      0xa2077d7dfbb606b5ull, 0xf3faac6faee378aeull, 0x0c6388b51545e883ull, 0xd27dbb6944917b60ull,
  };
  SmallRng rng(seed);
  for (int r = 0; r < 4; r++)
    for (int i = 0; i < 12; i++) c->rc_init[r][i] = rng.next_gl();
  for (int r = 0; r < 4; r++)
    for (int i = 0; i < 12; i++) c->rc_term[r][i] = rng.next_gl();
  for (int r = 0; r < 22; r++) c->rc_int[r] = rng.next_gl();
  for (int i = 0; i < 12; i++) c->diag[i] = diag[i];
}

P3G_HD uint64_t p2_sbox(uint64_t x) {
  uint64_t x2 = gl_sqr(x);
  uint64_t x3 = gl_mul(x2, x);
  uint64_t x4 = gl_sqr(x2);
  return gl_mul(x3, x4);
}

// M4 = [[2,3,1,1],[1,2,3,1],[1,1,2,3],[3,1,1,2]] applied to each 4-block, then
// each element gains the sum of its residue class mod 4.
P3G_HD void p2_external_linear(uint64_t* s) {
#pragma unroll
  for (int c = 0; c < 12; c += 4) {
    uint64_t a = s[c], b = s[c + 1], x = s[c + 2], d = s[c + 3];
    uint64_t ab = gl_add(a, b), xd = gl_add(x, d);
    uint64_t t = gl_add(ab, xd);
    uint64_t t_b = gl_add(t, b);  // a+2b+c+d
    uint64_t t_d = gl_add(t, d);  // a+b+c+2d
    s[c + 0] = gl_add(t_b, ab);                  // 2a+3b+c+d
    s[c + 1] = gl_add(t_b, gl_add(x, x));        // a+2b+3c+d
    s[c + 2] = gl_add(t_d, xd);                  // a+b+2c+3d
    s[c + 3] = gl_add(t_d, gl_add(a, a));        // 3a+b+c+2d
  }
#pragma unroll
  for (int k = 0; k < 4; k++) {
    uint64_t sum = gl_add(gl_add(s[k], s[4 + k]), s[8 + k]);
    s[k] = gl_add(s[k], sum);
    s[4 + k] = gl_add(s[4 + k], sum);
    s[8 + k] = gl_add(s[8 + k], sum);
  }
}

P3G_HD void p2_permute(uint64_t* s, const P2Consts* __restrict__ c) {
  p2_external_linear(s);
#pragma unroll
  for (int r = 0; r < 4; r++) {
#pragma unroll
    for (int i = 0; i < 12; i++) s[i] = p2_sbox(gl_add(s[i], c->rc_init[r][i]));
    p2_external_linear(s);
  }
#pragma unroll
  for (int r = 0; r < 22; r++) {
    s[0] = p2_sbox(gl_add(s[0], c->rc_int[r]));
    uint64_t sum = s[0];
#pragma unroll
    for (int i = 1; i < 12; i++) sum = gl_add(sum, s[i]);
#pragma unroll
    for (int i = 0; i < 12; i++) s[i] = gl_add(gl_mul(s[i], c->diag[i]), sum);
  }
#pragma unroll
  for (int r = 0; r < 4; r++) {
#pragma unroll
    for (int i = 0; i < 12; i++) s[i] = p2_sbox(gl_add(s[i], c->rc_term[r][i]));
    p2_external_linear(s);
  }
}

// ---------------------------------------------------------------- Keccak-f[1600]

#if defined(__HIP_DEVICE_COMPILE__)
// Device: a 64-bit rotation is two 32-bit funnel shifts (v_alignbit_b32)
// instead of two 64-bit shifts and two ORs.  k is a compile-time constant at
// every call site, so the branches fold.
__device__ __forceinline__ uint64_t rotl64(uint64_t x, int k) {
  const uint32_t lo = (uint32_t)x, hi = (uint32_t)(x >> 32);
  if (k == 0) return x;
  if (k == 32) return ((uint64_t)lo << 32) | hi;
  // rotating by k >= 32 is rotating the half-swapped word by k - 32
  const uint32_t h = k < 32 ? hi : lo, l = k < 32 ? lo : hi;
  const uint32_t m = (uint32_t)(k < 32 ? k : k - 32);
  const uint32_t nh = __builtin_amdgcn_alignbit(h, l, 32 - m);
  const uint32_t nl = __builtin_amdgcn_alignbit(l, h, 32 - m);
  return ((uint64_t)nh << 32) | nl;
}
// chi's b0 ^ (~b1 & b2) as one three-input bitwise instruction per 32-bit half
// (v_bitop3_b32) instead of a bitfield insert and an XOR.  The truth table is
// the function applied to the operand patterns 0xF0, 0xCC, 0xAA:
// 0xF0 ^ (~0xCC & 0xAA) = 0xD2.
__device__ __forceinline__ uint64_t chi64(uint64_t b0, uint64_t b1, uint64_t b2) {
  const uint32_t lo = __builtin_amdgcn_bitop3_b32((uint32_t)b0, (uint32_t)b1, (uint32_t)b2, 0xD2);
  const uint32_t hi =
      __builtin_amdgcn_bitop3_b32((uint32_t)(b0 >> 32), (uint32_t)(b1 >> 32), (uint32_t)(b2 >> 32), 0xD2);
  return ((uint64_t)hi << 32) | lo;
}
#else
inline uint64_t rotl64(uint64_t x, int k) { return k ? (x << k) | (x >> (64 - k)) : x; }
inline uint64_t chi64(uint64_t b0, uint64_t b1, uint64_t b2) { return b0 ^ (~b1 & b2); }
#endif

P3G_HD void keccak_f(uint64_t* a) {
  const uint64_t RC[24] = {
      0x0000000000000001ull, 0x0000000000008082ull, 0x800000000000808aull, 0x8000000080008000ull,
      0x000000000000808bull, 0x0000000080000001ull, 0x8000000080008081ull, 0x8000000000008009ull,
      0x000000000000008aull, 0x0000000000000088ull, 0x0000000080008009ull, 0x000000008000000aull,
      0x000000008000808bull, 0x800000000000008bull, 0x8000000000008089ull, 0x8000000000008003ull,
      0x8000000000008002ull, 0x8000000000000080ull, 0x000000000000800aull, 0x800000008000000aull,
      0x8000000080008081ull, 0x8000000000008080ull, 0x0000000080000001ull, 0x8000000080008008ull,
  };
  for (int round = 0; round < 24; round++) {
    uint64_t c[5], d[5];
#pragma unroll
    for (int x = 0; x < 5; x++) c[x] = a[x] ^ a[x + 5] ^ a[x + 10] ^ a[x + 15] ^ a[x + 20];
#pragma unroll
    for (int x = 0; x < 5; x++) d[x] = c[(x + 4) % 5] ^ rotl64(c[(x + 1) % 5], 1);
#pragma unroll
    for (int i = 0; i < 25; i++) a[i] ^= d[i % 5];
    // rho and pi: B[y, 2x+3y] = rot(A[x, y], r[x, y]), walked along the pi cycle
    uint64_t b[25];
    b[0] = a[0];
    b[10] = rotl64(a[1], 1);
    b[20] = rotl64(a[2], 62);
    b[5] = rotl64(a[3], 28);
    b[15] = rotl64(a[4], 27);
    b[16] = rotl64(a[5], 36);
    b[1] = rotl64(a[6], 44);
    b[11] = rotl64(a[7], 6);
    b[21] = rotl64(a[8], 55);
    b[6] = rotl64(a[9], 20);
    b[7] = rotl64(a[10], 3);
    b[17] = rotl64(a[11], 10);
    b[2] = rotl64(a[12], 43);
    b[12] = rotl64(a[13], 25);
    b[22] = rotl64(a[14], 39);
    b[23] = rotl64(a[15], 41);
    b[8] = rotl64(a[16], 45);
    b[18] = rotl64(a[17], 15);
    b[3] = rotl64(a[18], 21);
    b[13] = rotl64(a[19], 8);
    b[14] = rotl64(a[20], 18);
    b[24] = rotl64(a[21], 2);
    b[9] = rotl64(a[22], 61);
    b[19] = rotl64(a[23], 56);
    b[4] = rotl64(a[24], 14);
#pragma unroll
    for (int y = 0; y < 25; y += 5) {
#pragma unroll
      for (int x = 0; x < 5; x++) a[y + x] = chi64(b[y + x], b[y + (x + 1) % 5], b[y + (x + 2) % 5]);
    }
    a[0] ^= RC[round];
  }
}

}  // namespace p3g
