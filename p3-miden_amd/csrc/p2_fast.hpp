// Poseidon2-Goldilocks-12 with lazy reduction, for the hashing kernels.
//
// Same permutation as p2_permute in hash.hpp (the reference configuration of
// p3-miden-dev-utils/src/configs.rs:394-407), computed on non-canonical u64
// representatives: products are reduced to some value < 2^64 without the
// final conditional subtraction, the linear layers accumulate sums as
// (64-bit low word, carry count) pairs and reduce once per output, and the
// internal layer folds "+ sum" into the 128-bit product before reducing.
// Every state word is canonicalised at the end, so the output equals the
// canonical permutation bit for bit.
#pragma once
#include "hash.hpp"

namespace p3g {

// (hi * 2^64 + lo) mod p as some value < 2^64 (not necessarily < p).
P3G_HD uint64_t red_nc(uint64_t lo, uint64_t hi) { return gl_reduce_nc(lo, hi); }

// 64-bit add on 32-bit limbs; returns the carry out.
P3G_HD uint64_t add64c(uint64_t a, uint64_t b, unsigned int& carry) {
  unsigned int c0;
  const uint32_t r0 = __builtin_addc((uint32_t)a, (uint32_t)b, 0u, &c0);
  const uint32_t r1 = __builtin_addc((uint32_t)(a >> 32), (uint32_t)(b >> 32), c0, &carry);
  return ((uint64_t)r1 << 32) | r0;
}

P3G_HD uint64_t mul_nc(uint64_t a, uint64_t b) {
  uint64_t lo, hi;
  mul128(a, b, lo, hi);
  return red_nc(lo, hi);
}

// a (any u64) + c (canonical, < p): one carry fold suffices.
P3G_HD uint64_t add_canon_nc(uint64_t a, uint64_t c) {
  unsigned int carry;
  const uint64_t s = add64c(a, c, carry);
  return s + (carry ? EPS : 0);
}

// Sum of up to 2^32 words: low 64 bits and carry count.
struct Acc {
  uint64_t lo;
  uint32_t c;
};

P3G_HD Acc acc_of(uint64_t x) { return Acc{x, 0}; }
P3G_HD Acc acc_add(Acc a, uint64_t x) {
  unsigned int carry, dummy;
  const uint64_t s = add64c(a.lo, x, carry);
  return Acc{s, __builtin_addc(a.c, 0u, carry, &dummy)};
}
P3G_HD Acc acc_add(Acc a, Acc b) {
  unsigned int carry, dummy;
  const uint64_t s = add64c(a.lo, b.lo, carry);
  return Acc{s, __builtin_addc(a.c, b.c, carry, &dummy)};
}
// lo + c * 2^64 == lo + c * eps (mod p), as some value < 2^64
P3G_HD uint64_t acc_fold(Acc a) {
  const uint64_t t = ((uint64_t)a.c << 32) - a.c;
  unsigned int carry;
  const uint64_t r = add64c(a.lo, t, carry);
  return r + (carry ? EPS : 0);
}

P3G_HD uint64_t canon(uint64_t x) { return x >= P ? x - P : x; }

P3G_HD uint64_t sbox_nc(uint64_t x) {
  const uint64_t x2 = mul_nc(x, x);
  const uint64_t x3 = mul_nc(x2, x);
  const uint64_t x4 = mul_nc(x2, x2);
  return mul_nc(x3, x4);
}

#if defined(__HIP_DEVICE_COMPILE__)
// a*b (+ s) mod p as some value < 2^64 through v_mad_u64_u32 carry-outs and
// 64-bit addends, which do the carry propagation that otherwise takes add/addc
// pairs and selects.  With a = a0 + 2^32 a1, b = b0 + 2^32 b1:
//   P + cp 2^64 = a0 b0 + s                (cp = 0 without s)
//   M           = a1 b0 + P1 + 2^32 cp     (< 2^64: P1 <= 2^32 - 2 when cp = 1)
//   N + cn 2^64 = a0 b1 + M
//   R           = a1 b1 + N1               (<= 2^64 - 2^32)
// so a b + s = lo + 2^64 R + 2^96 cn with lo = P0 + 2^32 N0, and since
// 2^64 == eps and 2^96 == -1 (mod p):  a b + s == lo + eps R0 - R1 - cn.
//   T + c 2^64 = lo + eps R0               (v_mad_u64_u32 with carry-out c)
//   U - b 2^64 = T - R1 - cn               (borrow chain, borrow b)
// and the result is U + (c - b) eps mod 2^64, which lies in [0, 2^64) in all
// four cases of (c, b) (T < 2^64 - 2^33 when c = 1, U >= 2^64 - 2^32 when b = 1).
//
// The borrow b needs T < R1 + cn < 2^32, which has probability about 2^-32, so
// the products come in two forms:
//  - EXACT: folds (c - b) eps (11 single-issue VALU instructions per product;
//    the masks c & !b and b & !c are combined on the scalar unit);
//  - TRY: folds c eps only (9 instructions) and ORs b into a wave-uniform flag.
//    Its result is right in every lane whose borrow is clear, and also where
//    c = b = 1 (U >= 2^64 - 2^32 there, so U + eps wraps past 2^64, which is
//    -eps again); the flag takes every borrow all the same.  p2_permute_fast
//    runs the permutation on TRY products and redoes it with EXACT ones when the
//    flag is set, so the output is exact for every input.
// The assembly is arithmetic only.  An SGPR that a VALU instruction writes is
// read by VALU instructions two or more wait states later (s_nop), as gfx950
// requires; VALU-to-SALU and SALU-to-VALU reads of an SGPR need none.  No SGPR
// written here is read outside these statements, except the TRY flag (SALU).
template <bool TRY>
__device__ __forceinline__ uint64_t mul_tail(uint32_t a0, uint32_t a1, uint32_t b1, uint32_t P0, uint64_t M,
                                             uint64_t& flag) {
  uint64_t N, cn;
  asm("v_mad_u64_u32 %0, %1, %2, %3, %4" : "=v"(N), "=s"(cn) : "v"(a0), "v"(b1), "v"(M));
  const uint64_t R = (uint64_t)a1 * b1 + (N >> 32);
  const uint64_t lo = (N << 32) | P0;
  uint64_t T, c;
  asm("v_mad_u64_u32 %0, %1, %2, -1, %3" : "=v"(T), "=s"(c) : "v"((uint32_t)R), "v"(lo));
  uint32_t u0, u1, al, ah;
  uint64_t sb;
  if constexpr (TRY) {
    asm("s_nop 1\n\t"
        "v_subb_co_u32 %0, %3, %5, %7, %8\n\t"
        "v_cndmask_b32 %2, 0, -1, %9\n\t"
        "s_nop 0\n\t"
        "v_subb_co_u32 %1, %3, %6, 0, %3\n\t"
        "s_or_b64 %4, %4, %3"
        : "=&v"(u0), "=&v"(u1), "=&v"(al), "=&s"(sb), "+s"(flag)
        : "v"((uint32_t)T), "v"((uint32_t)(T >> 32)), "v"((uint32_t)(R >> 32)), "s"(cn), "s"(c)
        : "scc");
    ah = 0;
  } else {
    uint64_t f, g;
    asm("s_nop 1\n\t"
        "v_subb_co_u32 %0, %4, %7, %9, %10\n\t"
        "s_nop 1\n\t"
        "v_subb_co_u32 %1, %4, %8, 0, %4\n\t"
        "s_andn2_b64 %5, %4, %11\n\t"
        "s_andn2_b64 %6, %11, %4\n\t"
        "v_cndmask_b32 %2, 0, -1, %6\n\t"
        "v_cndmask_b32 %3, 0, -1, %5\n\t"
        "v_cndmask_b32 %2, %2, 1, %5"
        : "=&v"(u0), "=&v"(u1), "=&v"(al), "=&v"(ah), "=&s"(sb), "=&s"(f), "=&s"(g)
        : "v"((uint32_t)T), "v"((uint32_t)(T >> 32)), "v"((uint32_t)(R >> 32)), "s"(cn), "s"(c)
        : "scc");
  }
  const uint64_t U = ((uint64_t)u1 << 32) | u0;
  const uint64_t adj = ((uint64_t)ah << 32) | al;
  uint64_t r;
  asm("v_lshl_add_u64 %0, %1, 0, %2" : "=v"(r) : "v"(U), "v"(adj));
  return r;
}

template <bool TRY>
__device__ __forceinline__ uint64_t mul_t(uint64_t a, uint64_t b, uint64_t& flag) {
  const uint32_t a0 = (uint32_t)a, a1 = (uint32_t)(a >> 32), b0 = (uint32_t)b, b1 = (uint32_t)(b >> 32);
  const uint64_t P = (uint64_t)a0 * b0;
  const uint64_t M = (uint64_t)a1 * b0 + (P >> 32);
  return mul_tail<TRY>(a0, a1, b1, (uint32_t)P, M, flag);
}

template <bool TRY>
__device__ __forceinline__ uint64_t mul_add_t(uint64_t a, uint64_t b, uint64_t s, uint64_t& flag) {
  const uint32_t a0 = (uint32_t)a, a1 = (uint32_t)(a >> 32), b0 = (uint32_t)b, b1 = (uint32_t)(b >> 32);
  uint64_t P, cp;
  uint32_t cpv;
  asm("v_mad_u64_u32 %0, %1, %3, %4, %5\n\t"
      "s_nop 1\n\t"
      "v_cndmask_b32 %2, 0, 1, %1"
      : "=&v"(P), "=&s"(cp), "=v"(cpv)
      : "v"(a0), "v"(b0), "v"(s));
  const uint64_t M = (uint64_t)a1 * b0 + (((uint64_t)cpv << 32) | (P >> 32));
  return mul_tail<TRY>(a0, a1, b1, (uint32_t)P, M, flag);
}

template <bool TRY>
__device__ __forceinline__ uint64_t sbox_t(uint64_t x, uint64_t& flag) {
  const uint64_t x2 = mul_t<TRY>(x, x, flag);
  const uint64_t x3 = mul_t<TRY>(x2, x, flag);
  const uint64_t x4 = mul_t<TRY>(x2, x2, flag);
  return mul_t<TRY>(x3, x4, flag);
}

__device__ __forceinline__ uint64_t mul_fast(uint64_t a, uint64_t b) {
  uint64_t f = 0;
  return mul_t<false>(a, b, f);
}
__device__ __forceinline__ uint64_t mul_add_fast(uint64_t a, uint64_t b, uint64_t s) {
  uint64_t f = 0;
  return mul_add_t<false>(a, b, s, f);
}
#else
// host: the same values through 128-bit products
P3G_HD uint64_t mul_fast(uint64_t a, uint64_t b) { return mul_nc(a, b); }
P3G_HD uint64_t mul_add_fast(uint64_t a, uint64_t b, uint64_t s) {
  uint64_t lo, hi;
  mul128(a, b, lo, hi);
  unsigned int carry;
  lo = add64c(lo, s, carry);
  return red_nc(lo, hi + carry);
}
template <bool TRY>
P3G_HD uint64_t mul_t(uint64_t a, uint64_t b, uint64_t&) { return mul_nc(a, b); }
template <bool TRY>
P3G_HD uint64_t mul_add_t(uint64_t a, uint64_t b, uint64_t s, uint64_t&) { return mul_add_fast(a, b, s); }
template <bool TRY>
P3G_HD uint64_t sbox_t(uint64_t x, uint64_t&) { return sbox_nc(x); }
#endif

// External layer: y = circ(2 M4, M4, M4) x with M4 = [[2,3,1,1],[1,2,3,1],[1,1,2,3],[3,1,1,2]].
P3G_HD void ext_linear_nc(uint64_t* s) {
  Acc o[12];
#pragma unroll
  for (int j = 0; j < 12; j += 4) {
    const uint64_t a = s[j], b = s[j + 1], c = s[j + 2], d = s[j + 3];
    const Acc ab = acc_add(acc_of(a), b);
    const Acc cd = acc_add(acc_of(c), d);
    const Acc t = acc_add(ab, cd);       // a+b+c+d
    const Acc tb = acc_add(t, b);        // a+2b+c+d
    const Acc td = acc_add(t, d);        // a+b+c+2d
    o[j] = acc_add(tb, ab);              // 2a+3b+c+d
    o[j + 1] = acc_add(acc_add(tb, c), c);  // a+2b+3c+d
    o[j + 2] = acc_add(td, cd);             // a+b+2c+3d
    o[j + 3] = acc_add(acc_add(td, a), a);  // 3a+b+c+2d
  }
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const Acc sum = acc_add(acc_add(o[k], o[4 + k]), o[8 + k]);
    s[k] = acc_fold(acc_add(o[k], sum));
    s[4 + k] = acc_fold(acc_add(o[4 + k], sum));
    s[8 + k] = acc_fold(acc_add(o[8 + k], sum));
  }
}

#if defined(__HIP_DEVICE_COMPILE__)
// Linear layers on 32-bit halves.  A sum of words x_i = xl_i + 2^32 xh_i with
// small coefficients is kept as two 64-bit sums L = sum(k_i xl_i) and
// H = sum(k_i xh_i), each far below 2^64, so every step is one instruction
// with no carry to track: v_mad_u64_u32 adds k * (32-bit half) to a 64-bit sum,
// v_lshl_add_u64 adds two sums.  One reduction per output turns L + 2^32 H into
// a word (red_halves).  The carry-outs of these mads are never read (vcc).
__device__ __forceinline__ uint64_t mad1(uint32_t x, uint64_t y) {
  uint64_t r;
  asm("v_mad_u64_u32 %0, vcc, %1, 1, %2" : "=v"(r) : "v"(x), "v"(y) : "vcc");
  return r;
}
__device__ __forceinline__ uint64_t mad2(uint32_t x, uint64_t y) {
  uint64_t r;
  asm("v_mad_u64_u32 %0, vcc, %1, 2, %2" : "=v"(r) : "v"(x), "v"(y) : "vcc");
  return r;
}
__device__ __forceinline__ uint64_t add64(uint64_t x, uint64_t y) {
  uint64_t r;
  asm("v_lshl_add_u64 %0, %1, 0, %2" : "=v"(r) : "v"(x), "v"(y));
  return r;
}

// (L + 2^32 H) mod p as some value < 2^64, for L, H < 2^40.  With
// G = H + (L >> 32) the value is L0 + 2^32 G0 + 2^64 G1 == W + eps G1 where
// W = L0 + 2^32 G0; T + c 2^64 = W + eps G1 and T + c eps < 2^64 (T < eps G1
// < 2^40 when c = 1).
__device__ __forceinline__ uint64_t red_halves(uint64_t L, uint64_t H) {
  const uint64_t G = mad1((uint32_t)(L >> 32), H);
  const uint64_t W = (G << 32) | (uint32_t)L;
  uint64_t T, c;
  uint32_t e;
  asm("v_mad_u64_u32 %0, %1, %3, -1, %4\n\t"
      "s_nop 1\n\t"
      "v_cndmask_b32 %2, 0, -1, %1"
      : "=&v"(T), "=&s"(c), "=v"(e)
      : "v"((uint32_t)(G >> 32)), "v"(W));
  return add64(T, (uint64_t)e);
}

// Halves of the four M4 outputs of one block (a, b, c, d: one half of each word).
__device__ __forceinline__ void m4_halves(uint32_t a, uint32_t b, uint32_t c, uint32_t d, uint64_t* o) {
  const uint64_t ab = mad1(a, (uint64_t)b);
  const uint64_t cd = mad1(c, (uint64_t)d);
  const uint64_t t = add64(ab, cd);   // a+b+c+d
  const uint64_t tb = mad1(b, t);      // a+2b+c+d
  const uint64_t td = mad1(d, t);      // a+b+c+2d
  o[0] = add64(tb, ab);                // 2a+3b+c+d
  o[1] = mad2(c, tb);                  // a+2b+3c+d
  o[2] = add64(td, cd);                // a+b+2c+3d
  o[3] = mad2(a, td);                  // 3a+b+c+2d
}

// s <- circ(2 M4, M4, M4) s (+ rc when given: the next round's constants)
template <bool RC>
__device__ __forceinline__ void ext_linear_fast(uint64_t* s, const uint64_t* rc) {
  uint64_t ol[12], oh[12];
#pragma unroll
  for (int j = 0; j < 12; j += 4) {
    m4_halves((uint32_t)s[j], (uint32_t)s[j + 1], (uint32_t)s[j + 2], (uint32_t)s[j + 3], ol + j);
    m4_halves((uint32_t)(s[j] >> 32), (uint32_t)(s[j + 1] >> 32), (uint32_t)(s[j + 2] >> 32),
              (uint32_t)(s[j + 3] >> 32), oh + j);
  }
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const uint64_t sl = add64(add64(ol[k], ol[4 + k]), ol[8 + k]);
    const uint64_t sh = add64(add64(oh[k], oh[4 + k]), oh[8 + k]);
#pragma unroll
    for (int j = 0; j < 12; j += 4) {
      uint64_t L = add64(ol[j + k], sl), H = add64(oh[j + k], sh);
      if (RC) {
        L = mad1((uint32_t)rc[j + k], L);
        H = mad1((uint32_t)(rc[j + k] >> 32), H);
      }
      s[j + k] = red_halves(L, H);
    }
  }
}

// sum of the twelve state words
__device__ __forceinline__ uint64_t sum12_fast(const uint64_t* s) {
  uint64_t L = mad1((uint32_t)s[0], (uint64_t)(uint32_t)s[1]);
  uint64_t H = mad1((uint32_t)(s[0] >> 32), (uint64_t)(uint32_t)(s[1] >> 32));
#pragma unroll
  for (int i = 2; i < 12; i++) {
    L = mad1((uint32_t)s[i], L);
    H = mad1((uint32_t)(s[i] >> 32), H);
  }
  return red_halves(L, H);
}
#else
// (the host compilation pass parses kernels that call it)
__device__ uint64_t red_halves(uint64_t L, uint64_t H);
#endif

#if defined(__HIP_DEVICE_COMPILE__)
template <bool TRY>
__device__ __forceinline__ void p2_permute_dev(uint64_t* s, const P2Consts* __restrict__ c, uint64_t& flag) {
  // Each external layer that precedes a full round also adds that round's
  // constants (in its reductions), so the full rounds start at the S-boxes.
  ext_linear_fast<true>(s, c->rc_init[0]);
#pragma unroll
  for (int r = 0; r < 4; r++) {
#pragma unroll
    for (int i = 0; i < 12; i++) s[i] = sbox_t<TRY>(s[i], flag);
    if (r < 3) ext_linear_fast<true>(s, c->rc_init[r + 1]);
    else ext_linear_fast<false>(s, nullptr);
  }
#pragma unroll
  for (int r = 0; r < 22; r++) {
    s[0] = sbox_t<TRY>(add_canon_nc(s[0], c->rc_int[r]), flag);
    const uint64_t sm = sum12_fast(s);
#pragma unroll
    for (int i = 0; i < 12; i++) s[i] = mul_add_t<TRY>(s[i], c->diag[i], sm, flag);
  }
#pragma unroll
  for (int i = 0; i < 12; i++) s[i] = add_canon_nc(s[i], c->rc_term[0][i]);
#pragma unroll
  for (int r = 0; r < 4; r++) {
#pragma unroll
    for (int i = 0; i < 12; i++) s[i] = sbox_t<TRY>(s[i], flag);
    if (r < 3) ext_linear_fast<true>(s, c->rc_term[r + 1]);
    else ext_linear_fast<false>(s, nullptr);
  }
}
#endif

P3G_HD void p2_permute_fast(uint64_t* s, const P2Consts* __restrict__ c) {
#if defined(__HIP_DEVICE_COMPILE__)
  uint64_t flag = 0;
  p2_permute_dev<false>(s, c, flag);
#else
  ext_linear_nc(s);
  for (int r = 0; r < 4; r++) {
    for (int i = 0; i < 12; i++) s[i] = sbox_nc(add_canon_nc(s[i], c->rc_init[r][i]));
    ext_linear_nc(s);
  }
  for (int r = 0; r < 22; r++) {
    s[0] = sbox_nc(add_canon_nc(s[0], c->rc_int[r]));
    Acc sum = acc_of(s[0]);
    for (int i = 1; i < 12; i++) sum = acc_add(sum, s[i]);
    const uint64_t sm = acc_fold(sum);
    for (int i = 0; i < 12; i++) s[i] = mul_add_fast(s[i], c->diag[i], sm);
  }
  for (int r = 0; r < 4; r++) {
    for (int i = 0; i < 12; i++) s[i] = sbox_nc(add_canon_nc(s[i], c->rc_term[r][i]));
    ext_linear_nc(s);
  }
#endif
#pragma unroll
  for (int i = 0; i < 12; i++) s[i] = canon(s[i]);
}

#if defined(__HIP_DEVICE_COMPILE__)
// The canonical permutation of hash.hpp (p2_permute: canonical field
// operations throughout) on a state held in LDS (v[k * STRIDE], k < 12), one
// word at a time with every loop rolled: little code and few registers, for
// the rare redo below, where its speed does not matter.
template <int STRIDE>
__device__ __forceinline__ void p2_ext_lds(uint64_t* v) {
  uint64_t s[12];
#pragma unroll
  for (int i = 0; i < 12; i++) s[i] = v[i * STRIDE];
  p2_external_linear(s);
#pragma unroll
  for (int i = 0; i < 12; i++) v[i * STRIDE] = s[i];
}
template <int STRIDE>
__device__ __noinline__ void p2_permute_lds(uint64_t* v, const P2Consts* __restrict__ c) {
  p2_ext_lds<STRIDE>(v);
#pragma unroll 1
  for (int r = 0; r < 8; r++) {
    const uint64_t* rc = r < 4 ? c->rc_init[r] : c->rc_term[r - 4];
#pragma unroll 1
    for (int i = 0; i < 12; i++) v[i * STRIDE] = p2_sbox(gl_add(v[i * STRIDE], rc[i]));
    p2_ext_lds<STRIDE>(v);
    if (r == 3) {
#pragma unroll 1
      for (int q = 0; q < 22; q++) {
        v[0] = p2_sbox(gl_add(v[0], c->rc_int[q]));
        uint64_t sum = 0;
#pragma unroll 1
        for (int i = 0; i < 12; i++) sum = gl_add(sum, v[i * STRIDE]);
#pragma unroll 1
        for (int i = 0; i < 12; i++) v[i * STRIDE] = gl_add(gl_mul(v[i * STRIDE], c->diag[i]), sum);
      }
    }
  }
}

// The permutation on TRY products, for the hashing kernels: the input is kept
// in LDS (save[k * STRIDE], k < 12: this thread's slots in a block-wide
// array), the permutation runs on TRY products, and the whole wave redoes it
// from the saved input with the canonical permutation when any lane met a
// product borrow (about 2^-32 per product).  The output equals p2_permute's
// (canonical) for every input.
template <int STRIDE>
__device__ __forceinline__ void p2_permute_saved(uint64_t* s, const P2Consts* __restrict__ c, uint64_t* save) {
#pragma unroll
  for (int i = 0; i < 12; i++) save[i * STRIDE] = s[i];
  uint64_t flag = 0;
  p2_permute_dev<true>(s, c, flag);
  if (__builtin_expect(flag != 0, 0)) {
    p2_permute_lds<STRIDE>(save, c);
#pragma unroll
    for (int i = 0; i < 12; i++) s[i] = save[i * STRIDE];
  } else {
#pragma unroll
    for (int i = 0; i < 12; i++) s[i] = canon(s[i]);
  }
}
#else
// (the host compilation pass parses the kernels that call them)
template <int STRIDE>
__device__ void p2_permute_saved(uint64_t* s, const P2Consts* __restrict__ c, uint64_t* save);
template <bool TRY>
__device__ void p2_permute_dev(uint64_t* s, const P2Consts* __restrict__ c, uint64_t& flag);
template <int STRIDE>
__device__ void p2_permute_lds(uint64_t* v, const P2Consts* __restrict__ c);
#endif

}  // namespace p3g
