// Goldilocks field (p = 2^64 - 2^32 + 1) and its quadratic extension
// EF = F[x]/(x^2 - 7), for host and gfx950 device code.
//
// Matches the field the reference instantiates: Goldilocks with
// BinomialExtensionField<Goldilocks, 2> (p3-miden-dev-utils/src/configs.rs:394-407).
// All values are kept canonical (< p) so that every stored word equals the
// reference's canonical serialisation.
#pragma once
#if defined(__HIPCC_RTC__)
// compiled at run time (hiprtc): no system headers; the types as <stdint.h> has them
typedef unsigned long uint64_t;
typedef unsigned int uint32_t;
typedef long int64_t;
typedef int int32_t;
#else
#include <hip/hip_runtime.h>
#include <stdint.h>
#endif

#define P3G_HD __host__ __device__ __forceinline__

namespace p3g {

constexpr uint64_t P = 0xFFFFFFFF00000001ull;
constexpr uint64_t EPS = 0xFFFFFFFFull;  // 2^64 mod p
constexpr uint64_t GENERATOR = 7;
constexpr uint64_t OMEGA32 = 1753635133440165772ull;  // 7^((p-1)/2^32)
constexpr uint64_t EXT_W = 7;

P3G_HD uint64_t mulhi64(uint64_t a, uint64_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __umul64hi(a, b);
#else
  return (uint64_t)(((unsigned __int128)a * b) >> 64);
#endif
}

// Canonical add of canonical a, b: s = a + b and t = s + eps (= s - p mod
// 2^64); the result is t when either sum carries out of 64 bits, else s.
// Both candidates are computed and one is selected (no divergent branch).
P3G_HD uint64_t gl_add(uint64_t a, uint64_t b) {
  unsigned int c0, c1, d0, d1;
  const uint32_t s0 = __builtin_addc((uint32_t)a, (uint32_t)b, 0u, &c0);
  const uint32_t s1 = __builtin_addc((uint32_t)(a >> 32), (uint32_t)(b >> 32), c0, &c1);
  const uint32_t t0 = __builtin_addc(s0, 0xFFFFFFFFu, 0u, &d0);
  const uint32_t t1 = __builtin_addc(s1, 0u, d0, &d1);
  const bool over = (c1 | d1) != 0;
  return ((uint64_t)(over ? t1 : s1) << 32) | (over ? t0 : s0);
}

P3G_HD uint64_t gl_sub(uint64_t a, uint64_t b) {
  unsigned int c0, c1;
  const uint32_t d0 = __builtin_subc((uint32_t)a, (uint32_t)b, 0u, &c0);
  const uint32_t d1 = __builtin_subc((uint32_t)(a >> 32), (uint32_t)(b >> 32), c0, &c1);
  const uint64_t d = ((uint64_t)d1 << 32) | d0;
  return c1 ? d - EPS : d;  // borrow: -2^64 == -EPS
}

P3G_HD uint64_t gl_neg(uint64_t a) { return a ? P - a : 0; }

// x = hi*2^64 + lo with hi = h1*2^32 + h0: since 2^64 == 2^32 - 1 and
// 2^96 == -1, x == lo + h0*2^32 - (h0 + h1).  Computed on 32-bit limbs with
// carry chains; the final carry (+2^64) and borrow (-2^64) fold as +-eps.
// Returns a value < 2^64 that is not necessarily < p.
P3G_HD uint64_t gl_reduce_nc(uint64_t lo, uint64_t hi) {
#if !defined(__HIP_DEVICE_COMPILE__)
  // host: the same identity on 64-bit words (each fold cannot wrap twice)
  const uint64_t hh = hi >> 32, hl = hi & EPS;
  uint64_t t = lo - hh;
  if (lo < hh) t -= EPS;
  const uint64_t t2 = (hl << 32) - hl;
  uint64_t r = t + t2;
  if (r < t2) r += EPS;
  return r;
#endif
  const uint32_t l0 = (uint32_t)lo, l1 = (uint32_t)(lo >> 32);
  const uint32_t h0 = (uint32_t)hi, h1 = (uint32_t)(hi >> 32);
  unsigned int sc, b0, b1, c1;
  const uint32_t s0 = __builtin_addc(h0, h1, 0u, &sc);
  const uint32_t t0 = __builtin_subc(l0, s0, 0u, &b0);
  const uint32_t t1 = __builtin_subc(l1, sc, b0, &b1);
  const uint32_t u1 = __builtin_addc(t1, h0, 0u, &c1);
  uint64_t u = ((uint64_t)u1 << 32) | t0;
  u += c1 ? EPS : 0;
  u -= b1 ? EPS : 0;
  return u;
}

P3G_HD uint64_t gl_reduce(uint64_t lo, uint64_t hi) {
  const uint64_t r = gl_reduce_nc(lo, hi);
  return r >= P ? r - P : r;
}

// 64 x 64 -> 128-bit product: on the device from four 32 x 32 -> 64
// multiply-adds (v_mad_u64_u32), on the host one widening multiply.
P3G_HD void mul128(uint64_t a, uint64_t b, uint64_t& lo, uint64_t& hi) {
#if defined(__HIP_DEVICE_COMPILE__)
  const uint32_t a0 = (uint32_t)a, a1 = (uint32_t)(a >> 32), b0 = (uint32_t)b, b1 = (uint32_t)(b >> 32);
  const uint64_t p00 = (uint64_t)a0 * b0;
  const uint64_t p01 = (uint64_t)a0 * b1 + (p00 >> 32);
  const uint64_t p10 = (uint64_t)a1 * b0 + (uint32_t)p01;
  hi = (uint64_t)a1 * b1 + (p01 >> 32) + (p10 >> 32);
  lo = (p10 << 32) | (uint32_t)p00;
#else
  const unsigned __int128 p = (unsigned __int128)a * b;
  lo = (uint64_t)p;
  hi = (uint64_t)(p >> 64);
#endif
}

P3G_HD uint64_t gl_mul(uint64_t a, uint64_t b) {
  uint64_t lo, hi;
  mul128(a, b, lo, hi);
  return gl_reduce(lo, hi);
}

// x * 2^s for canonical x, 0 <= s < 192 (2^96 == -1, 2^192 == 1): shifts and
// one reduction, no multiplies.  Written for s known at compile time (the
// branches fold away).  For 64 < s < 96 with u = s - 64:
// x 2^s = (lo_u + hi_u 2^64) 2^64 == lo_u 2^64 - hi_u 2^32, since 2^128 == -2^32.
P3G_HD uint64_t gl_mul_pow2(uint64_t x, uint32_t s) {
  s %= 192;
  const bool neg = s >= 96;
  s %= 96;
  uint64_t r;
  if (s == 0) r = x;
  else if (s < 64) r = gl_reduce(x << s, x >> (64 - s));
  else if (s == 64) r = gl_reduce(0, x);
  else {
    const uint32_t u = s - 64;
    const uint64_t lo_u = x << u, hi_u = x >> (64 - u);
    const uint64_t a = gl_reduce(0, lo_u), b = hi_u << 32;  // b < 2^64 - 2^32 < p
    r = gl_sub(a, b);
  }
  return neg ? gl_neg(r) : r;
}

// Lazy sum of 128-bit products: value = lo + hi 2^64 + c 2^128, where c counts
// the carries out of 128 bits.  A dot product pays one reduction at the end
// instead of one per term.
struct Acc192 {
  uint64_t lo, hi;
  uint32_t c;
};
P3G_HD Acc192 acc192_zero() { return Acc192{0, 0, 0}; }
// a += x * y
P3G_HD void acc192_mad(Acc192& a, uint64_t x, uint64_t y) {
  uint64_t plo, phi;
  mul128(x, y, plo, phi);
  unsigned int c0, c1, c2, c3;
  const uint32_t r0 = __builtin_addc((uint32_t)a.lo, (uint32_t)plo, 0u, &c0);
  const uint32_t r1 = __builtin_addc((uint32_t)(a.lo >> 32), (uint32_t)(plo >> 32), c0, &c1);
  const uint32_t r2 = __builtin_addc((uint32_t)a.hi, (uint32_t)phi, c1, &c2);
  const uint32_t r3 = __builtin_addc((uint32_t)(a.hi >> 32), (uint32_t)(phi >> 32), c2, &c3);
  a.lo = ((uint64_t)r1 << 32) | r0;
  a.hi = ((uint64_t)r3 << 32) | r2;
  a.c += c3;
}
// canonical value mod p; 2^128 == -2^32 and c 2^32 < p
P3G_HD uint64_t acc192_reduce(const Acc192& a) { return gl_sub(gl_reduce(a.lo, a.hi), (uint64_t)a.c << 32); }

// omega_16 = gl_two_adic_generator(4) = 2^156 (checked at context creation):
// powers of the 16th roots of unity are shifts.
constexpr uint32_t OMEGA16_LOG2 = 156;
// x * omega_16^j (inverse: omega_16^-j)
P3G_HD uint64_t gl_mul_w16(uint64_t x, uint32_t j, bool inverse) {
  const uint32_t e = (OMEGA16_LOG2 * j) % 192;
  return gl_mul_pow2(x, inverse ? (192 - e) % 192 : e);
}
P3G_HD uint64_t gl_sqr(uint64_t a) { return gl_mul(a, a); }

P3G_HD uint64_t gl_pow(uint64_t a, uint64_t e) {
  uint64_t r = 1;
  while (e) {
    if (e & 1) r = gl_mul(r, a);
    a = gl_mul(a, a);
    e >>= 1;
  }
  return r;
}

// a^k squared s times, times b
P3G_HD uint64_t gl_sqr_n_mul(uint64_t a, unsigned s, uint64_t b) {
  for (unsigned i = 0; i < s; i++) a = gl_mul(a, a);
  return gl_mul(a, b);
}

// a^(p-2) (0 for 0) by an addition chain: with t_k = a^(2^k - 1),
// p - 2 = (2^31 - 1) 2^33 + (2^32 - 1), so a^(p-2) = t_31^(2^33) t_32.
// 64 squarings and 9 products (square-and-multiply takes 63 and 62).
P3G_HD uint64_t gl_inv(uint64_t a) {
  const uint64_t t2 = gl_sqr_n_mul(a, 1, a);
  const uint64_t t3 = gl_sqr_n_mul(t2, 1, a);
  const uint64_t t6 = gl_sqr_n_mul(t3, 3, t3);
  const uint64_t t12 = gl_sqr_n_mul(t6, 6, t6);
  const uint64_t t24 = gl_sqr_n_mul(t12, 12, t12);
  const uint64_t t30 = gl_sqr_n_mul(t24, 6, t6);
  const uint64_t t31 = gl_sqr_n_mul(t30, 1, a);
  const uint64_t t32 = gl_sqr_n_mul(t31, 1, a);
  return gl_sqr_n_mul(t31, 33, t32);
}

P3G_HD uint64_t gl_two_adic_generator(unsigned bits) {
  uint64_t g = OMEGA32;
  for (unsigned i = bits; i < 32; i++) g = gl_mul(g, g);
  return g;
}

P3G_HD uint64_t gl_inv_pow2(unsigned k) { return gl_inv(gl_pow(2, k)); }

struct ef2 {
  uint64_t c0, c1;
};

P3G_HD ef2 ef_make(uint64_t a, uint64_t b) { return ef2{a, b}; }
P3G_HD ef2 ef_from(uint64_t a) { return ef2{a, 0}; }
P3G_HD ef2 ef_add(ef2 a, ef2 b) { return ef2{gl_add(a.c0, b.c0), gl_add(a.c1, b.c1)}; }
P3G_HD ef2 ef_sub(ef2 a, ef2 b) { return ef2{gl_sub(a.c0, b.c0), gl_sub(a.c1, b.c1)}; }
P3G_HD ef2 ef_neg(ef2 a) { return ef2{gl_neg(a.c0), gl_neg(a.c1)}; }
P3G_HD ef2 ef_mul(ef2 a, ef2 b) {
  uint64_t t = gl_mul(gl_mul(a.c1, b.c1), EXT_W);
  return ef2{gl_add(gl_mul(a.c0, b.c0), t), gl_add(gl_mul(a.c0, b.c1), gl_mul(a.c1, b.c0))};
}
P3G_HD ef2 ef_mul_base(ef2 a, uint64_t b) { return ef2{gl_mul(a.c0, b), gl_mul(a.c1, b)}; }
P3G_HD ef2 ef_sqr(ef2 a) { return ef_mul(a, a); }
P3G_HD bool ef_eq(ef2 a, ef2 b) { return a.c0 == b.c0 && a.c1 == b.c1; }
P3G_HD ef2 ef_inv(ef2 a) {
  uint64_t norm = gl_sub(gl_sqr(a.c0), gl_mul(EXT_W, gl_sqr(a.c1)));
  uint64_t ni = gl_inv(norm);
  return ef2{gl_mul(a.c0, ni), gl_neg(gl_mul(a.c1, ni))};
}
P3G_HD ef2 ef_pow(ef2 a, uint64_t e) {
  ef2 r = ef_from(1);
  while (e) {
    if (e & 1) r = ef_mul(r, a);
    a = ef_sqr(a);
    e >>= 1;
  }
  return r;
}

P3G_HD uint32_t bitrev32(uint32_t x, unsigned bits) {
  if (bits == 0) return 0;
#if defined(__HIP_DEVICE_COMPILE__)
  return __brev(x) >> (32 - bits);
#else
  uint32_t r = 0;
  for (unsigned i = 0; i < bits; i++) {
    r = (r << 1) | (x & 1);
    x >>= 1;
  }
  return r;
#endif
}

inline unsigned log2_exact(uint64_t n) {
  unsigned l = 0;
  while ((1ull << l) < n) l++;
  return l;
}

}  // namespace p3g
