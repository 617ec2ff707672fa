// Field operations for the NTT passes: canonical products and products by
// powers of two built on v_mad_u64_u32 carry-outs (the EXACT products of
// p2_fast.hpp), in fewer VALU issue slots than the generic carry-chain forms
// of gl.hpp.  Inputs and outputs are canonical (< p), as every NTT pass
// stores canonical words.
#pragma once
#include "p2_fast.hpp"

namespace p3g {

#if defined(__HIP_DEVICE_COMPILE__)
// canonical a * b for canonical a, b
__device__ __forceinline__ uint64_t mulc(uint64_t a, uint64_t b) {
  uint64_t f = 0;
  return canon(mul_t<false>(a, b, f));
}

// canonical v * 2^E for canonical v and 0 <= E < 96, written for E known at
// compile time (the branches fold away after unrolling).
//  - E <= 32: v 2^E = lo + 2^64 hi with hi < 2^32, and 2^64 == eps, so
//    T + c 2^64 = lo + eps hi (one v_mad_u64_u32 with carry-out) and the value
//    is T + c eps (< 2^64: T < eps hi when c = 1).
//  - 32 < E < 64: lo = v0 2^E (low word zero), hi = v >> (64 - E) = h0 + 2^32 h1
//    with h1 < 2^(E - 32); v 2^E == lo + eps h0 - h1 (2^96 == -1): a mad with
//    carry-out c, a borrow chain with borrow b, and the fold (c - b) eps as in
//    the products.
//  - 64 <= E < 96: 2^E == eps 2^(E - 64) = 2^(E - 32) - 2^(E - 64), a canonical
//    constant; one product.
__device__ __forceinline__ uint64_t mul_pow2c(uint64_t v, uint32_t E) {
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
  return mulc(v, (1ull << (E - 32)) - (1ull << (E - 64)));
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
__device__ uint64_t mulc(uint64_t a, uint64_t b);
__device__ uint64_t mul_pow2c(uint64_t v, uint32_t E);
__device__ void w16_pow2(uint32_t j, bool inverse, uint32_t& E, bool& neg);
#endif

}  // namespace p3g
