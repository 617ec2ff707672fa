// Field operations for the NTT passes: canonical sums, differences, products
// and products by powers of two built on v_mad_u64_u32 carry-outs (the EXACT
// products of p2_fast.hpp) and on carry and borrow masks used as carry-ins, in
// fewer VALU issue slots than the generic carry-chain forms of gl.hpp.  Inputs
// and outputs are canonical (< p), as every NTT pass stores canonical words.
//
// Adding +-eps (eps = 2^32 - 1) to a 64-bit word w = w0 + 2^32 w1 under a
// lane mask m takes two carry instructions and no select:
//   w + eps = (w0 - 1) + 2^32 (w1 + 1 - bb), bb the borrow of w0 - 1:
//     r0 = w0 - m (borrow bb), r1 = w1 + (m & ~bb);
//   w - eps = (w0 + 1) + 2^32 (w1 - 1 + cc), cc the carry of w0 + 1:
//     r0 = w0 + m (carry cc), r1 = w1 - (m & ~cc).
// The masks are combined on the scalar unit.  A VALU instruction that reads an
// SGPR written by the VALU two or fewer instructions before waits
// (s_nop 1); SGPRs written by the scalar unit are read at once.
#pragma once
#include "p2_fast.hpp"

namespace p3g {

#if defined(__HIP_DEVICE_COMPILE__)
// x mod p for any 64-bit x: x >= p only when x1 = 2^32 - 1 and x0 >= 1, and
// then x - p = x0 - 1 (one compare, one borrow, one select)
__device__ __forceinline__ uint64_t canon3(uint64_t x) {
  const bool ge = x >= P;
  const uint32_t lo = (uint32_t)x - (uint32_t)ge;
  const uint32_t hi = ge ? 0u : (uint32_t)(x >> 32);
  return ((uint64_t)hi << 32) | lo;
}

// canonical x + y: t = x + y - p (mod 2^64) is below x exactly when x + y >= p
// (x + y < p: t = x + y + eps > x; p <= x + y: t = x - (p - y) < x since
// y < p); two 64-bit adds, one compare, two selects
__device__ __forceinline__ uint64_t add_c(uint64_t x, uint64_t y) {
  const uint64_t s = x + y, t = s + EPS;
  return t < x ? t : s;
}

// canonical x - y: d = x - y with borrow b, then d + p = d - eps (mod 2^64)
// under b; four carry instructions
__device__ __forceinline__ uint64_t sub_c(uint64_t x, uint64_t y) {
  uint32_t r0, r1;
  uint64_t b, cc;
  asm("v_sub_co_u32 %0, %2, %4, %6\n\t"
      "s_nop 1\n\t"
      "v_subb_co_u32 %1, %2, %5, %7, %2\n\t"
      "s_nop 1\n\t"
      "v_addc_co_u32 %0, %3, %0, 0, %2\n\t"
      "s_andn2_b64 %2, %2, %3\n\t"
      "v_subb_co_u32 %1, %3, %1, 0, %2"
      : "=&v"(r0), "=v"(r1), "=s"(b), "=s"(cc)
      : "v"((uint32_t)x), "v"((uint32_t)(x >> 32)), "v"((uint32_t)y), "v"((uint32_t)(y >> 32))
      : "scc");
  return ((uint64_t)r1 << 32) | r0;
}

// canonical a * b for canonical a, b
__device__ __forceinline__ uint64_t mulc(uint64_t a, uint64_t b) {
  uint64_t f = 0;
  return canon3(mul_t<false>(a, b, f));
}

// canonical v * 2^E for canonical v and 0 <= E < 96, written for E known at
// compile time (the branches fold away after unrolling).
//  - E <= 32: v 2^E = lo + 2^64 hi with hi < 2^32, and 2^64 == eps, so
//    T + c 2^64 = lo + eps hi (one v_mad_u64_u32 with carry-out).  With c the
//    value is T + eps (< p: T < eps hi); without, T reduced once.  Both are
//    T + eps (mod 2^64) under the mask c | (T >= p).
//  - 32 < E < 64: lo = v0 2^E (low word zero), hi = v >> (64 - E) = h0 + 2^32 h1
//    with h1 < 2^(E - 32); v 2^E == lo + eps h0 - h1 (2^96 == -1).  A mad with
//    carry-out c gives T, X = T - h1 with borrow b; the value is X + eps when c
//    (canonical, also after a borrow), X - eps when b and not c (canonical), and
//    X reduced once otherwise: X - eps under b & ~c first, then + eps under
//    c | (X' >= p), the comparison read off the carry of X' + eps.
//  - 64 <= E < 96: with k = 96 - E (1..32), 2^E = -2^-k and
//    2^-k == -2^(32 - k) eps, so for v = q 2^k + r (r < 2^k)
//    v 2^E == (r 2^(32 - k)) eps - q: a product of 32-bit words (< p) minus
//    q < p, one canonical subtraction.
__device__ __forceinline__ uint64_t mul_pow2c(uint64_t v, uint32_t E) {
  if (E == 0) return v;
  if (E <= 32) {
    const uint64_t lo = v << (E & 63);
    const uint32_t hi = E == 32 ? (uint32_t)(v >> 32) : (uint32_t)(v >> (64 - E));
    uint64_t T, c;
    asm("v_mad_u64_u32 %0, %1, %2, -1, %3" : "=v"(T), "=s"(c) : "v"(hi), "v"(lo));
    uint64_t m, bb;
    asm("v_cmp_lt_u64_e64 %0, %1, %2" : "=s"(m) : "s"(P - 1), "v"(T));
    uint32_t r0 = (uint32_t)T, r1 = (uint32_t)(T >> 32);
    asm("s_or_b64 %2, %2, %4\n\t"
        "v_subb_co_u32 %0, %3, %0, 0, %2\n\t"
        "s_andn2_b64 %3, %2, %3\n\t"
        "v_addc_co_u32 %1, %3, %1, 0, %3"
        : "+v"(r0), "+v"(r1), "+s"(m), "=s"(bb)
        : "s"(c)
        : "scc");
    return ((uint64_t)r1 << 32) | r0;
  }
  if (E < 64) {
    const uint32_t v0 = (uint32_t)v;
    const uint64_t lo = (uint64_t)(v0 << (E - 32)) << 32;
    const uint64_t hi = v >> (64 - E);
    uint64_t T, c;
    asm("v_mad_u64_u32 %0, %1, %2, -1, %3" : "=v"(T), "=s"(c) : "v"((uint32_t)hi), "v"(lo));
    uint32_t r0 = (uint32_t)T, r1 = (uint32_t)(T >> 32), u0, u1;
    uint64_t b, k;
    asm("v_sub_co_u32 %0, %2, %0, %6\n\t"
        "s_nop 1\n\t"
        "v_subb_co_u32 %1, %2, %1, 0, %2\n\t"
        "s_andn2_b64 %2, %2, %7\n\t"
        "v_addc_co_u32 %0, %3, %0, 0, %2\n\t"
        "s_andn2_b64 %2, %2, %3\n\t"
        "v_subb_co_u32 %1, %3, %1, 0, %2\n\t"
        "v_add_co_u32 %4, %3, %0, -1\n\t"
        "s_nop 1\n\t"
        "v_addc_co_u32 %5, %3, %1, 0, %3\n\t"
        "s_or_b64 %3, %3, %7\n\t"
        "v_cndmask_b32 %0, %0, %4, %3\n\t"
        "v_cndmask_b32 %1, %1, %5, %3"
        : "+v"(r0), "+v"(r1), "=&s"(b), "=&s"(k), "=v"(u0), "=v"(u1)
        : "v"((uint32_t)(hi >> 32)), "s"(c)
        : "scc");
    return ((uint64_t)r1 << 32) | r0;
  }
  const uint32_t k = 96 - E;
  const uint32_t rp = k == 32 ? (uint32_t)v : (uint32_t)v << ((32 - k) & 31);
  return sub_c((uint64_t)rp * EPS, v >> k);
}

__device__ __forceinline__ uint64_t mulc_cc(uint64_t a, uint64_t b) {
  uint64_t f = 0;
  return canon(mul_t<false>(a, b, f));
}
// The same products in the carry-chain forms (a select per word, canon after
// the fold): more issue slots, fewer live registers.  The inverse first pass
// uses these: its bit-reversed row gather is latency-bound, and with them it
// compiles to 80 VGPRs (6 waves per SIMD) instead of 99 (4 waves).
//  v * 2^E:
//  - E <= 32: v 2^E = lo + 2^64 hi with hi < 2^32, and 2^64 == eps, so
//    T + c 2^64 = lo + eps hi (one v_mad_u64_u32 with carry-out) and the value
//    is T + c eps (< 2^64: T < eps hi when c = 1).
//  - 32 < E < 64: lo = v0 2^E (low word zero), hi = v >> (64 - E) = h0 + 2^32 h1
//    with h1 < 2^(E - 32); v 2^E == lo + eps h0 - h1 (2^96 == -1): a mad with
//    carry-out c, a borrow chain with borrow b, and the fold (c - b) eps as in
//    the products.
//  - 64 <= E < 96: 2^E == eps 2^(E - 64) = 2^(E - 32) - 2^(E - 64), a canonical
//    constant; one product.
__device__ __forceinline__ uint64_t mul_pow2c_cc(uint64_t v, uint32_t E) {
  if (E == 0) return v;
  if (E <= 32) {
    const uint64_t lo = v << (E & 63);
    const uint32_t hi = E == 32 ? (uint32_t)(v >> 32) : (uint32_t)(v >> (64 - E));
    uint64_t T, c;
    uint32_t e;
    asm("v_mad_u64_u32 %0, %1, %3, -1, %4\n\t"
        "s_nop 1\n\t"
        "v_cndmask_b32 %2, 0, -1, %1"
        : "=&v"(T), "=&s"(c), "=v"(e)
        : "v"(hi), "v"(lo));
    return canon(add64(T, (uint64_t)e));
  }
  if (E < 64) {
    const uint32_t v0 = (uint32_t)v, v1 = (uint32_t)(v >> 32);
    const uint64_t lo = (uint64_t)(v0 << (E - 32)) << 32;
    const uint64_t hi = v >> (64 - E);
    uint64_t T, c;
    asm("v_mad_u64_u32 %0, %1, %2, -1, %3" : "=v"(T), "=s"(c) : "v"((uint32_t)hi), "v"(lo));
    uint32_t u0, u1, al, ah;
    uint64_t sb, f, g;
    asm("s_nop 1\n\t"
        "v_sub_co_u32 %0, %4, %7, %9\n\t"
        "s_nop 1\n\t"
        "v_subb_co_u32 %1, %4, %8, 0, %4\n\t"
        "s_andn2_b64 %5, %4, %10\n\t"
        "s_andn2_b64 %6, %10, %4\n\t"
        "v_cndmask_b32 %2, 0, -1, %6\n\t"
        "v_cndmask_b32 %3, 0, -1, %5\n\t"
        "v_cndmask_b32 %2, %2, 1, %5"
        : "=&v"(u0), "=&v"(u1), "=&v"(al), "=&v"(ah), "=&s"(sb), "=&s"(f), "=&s"(g)
        : "v"((uint32_t)T), "v"((uint32_t)(T >> 32)), "v"((uint32_t)(hi >> 32)), "s"(c)
        : "scc");
    (void)v1;
    return canon(add64(((uint64_t)u1 << 32) | u0, ((uint64_t)ah << 32) | al));
  }
  return mulc_cc(v, (1ull << (E - 32)) - (1ull << (E - 64)));
}

// omega_16^j (omega_16^-j when inverse) = (-1)^neg 2^E for the butterflies: the
// sign is applied by the caller, which swaps the operands of its add and
// subtract instead of negating.
__device__ __forceinline__ void w16_pow2(uint32_t j, bool inverse, uint32_t& E, bool& neg) {
  uint32_t e = (OMEGA16_LOG2 * j) % 192;
  if (inverse) e = (192 - e) % 192;
  neg = e >= 96;
  E = e % 96;
}
#else
// (the host compilation pass parses the kernels that call them)
__device__ uint64_t canon3(uint64_t x);
__device__ uint64_t add_c(uint64_t x, uint64_t y);
__device__ uint64_t sub_c(uint64_t x, uint64_t y);
__device__ uint64_t mulc(uint64_t a, uint64_t b);
__device__ uint64_t mul_pow2c(uint64_t v, uint32_t E);
__device__ uint64_t mulc_cc(uint64_t a, uint64_t b);
__device__ uint64_t mul_pow2c_cc(uint64_t v, uint32_t E);
__device__ void w16_pow2(uint32_t j, bool inverse, uint32_t& E, bool& neg);
#endif

}  // namespace p3g
