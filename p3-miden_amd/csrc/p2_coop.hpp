// Lane-cooperative Poseidon2-Goldilocks-12 for latency-bound tree tops.
//
// Sixteen lanes (one DPP row) compute one permutation: lane l < 12 holds state
// word l and lanes 12..15 hold zero.  S-boxes run in every lane at once; the
// external layer forms M4 inside each quad from quad broadcasts and adds the
// three blocks with two row rotations; the internal layer's sum is a four-step
// row-rotation reduction, after which every lane multiplies its own word by
// its diagonal entry.  The dependent chain is about a fifth of the
// one-lane-per-permutation form, which is what the top levels of a tree (a
// handful of permutations per level) are bound by.  The products are the
// TRY ones of p2_fast.hpp, with an EXACT redo of the wave when one flags its
// rare borrow; every value a DPP instruction reads comes from a
// compiler-visible select, not straight from their assembly.  The output
// equals the canonical permutation bit for bit.
#pragma once
#include "p2_fast.hpp"

namespace p3g {

template <int CTRL>
__device__ __forceinline__ uint32_t dpp32(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
__device__ __forceinline__ uint64_t dpp64(uint64_t v) {
  return ((uint64_t)dpp32<CTRL>((uint32_t)(v >> 32)) << 32) | dpp32<CTRL>((uint32_t)v);
}
template <int CTRL>
__device__ __forceinline__ Acc dpp_acc(Acc a) {
  return Acc{dpp64<CTRL>(a.lo), dpp32<CTRL>(a.c)};
}

// DPP controls: quad_perm broadcast of quad lane k, and rotations within a 16-lane row
constexpr int DPP_QUAD_BCAST0 = 0x00, DPP_QUAD_BCAST1 = 0x55, DPP_QUAD_BCAST2 = 0xAA, DPP_QUAD_BCAST3 = 0xFF;
constexpr int DPP_ROW_ROR1 = 0x121, DPP_ROW_ROR2 = 0x122, DPP_ROW_ROR4 = 0x124, DPP_ROW_ROR8 = 0x128;

// External layer y = circ(2 M4, M4, M4) x on the row; dummy lanes come out zero.
__device__ __forceinline__ uint64_t coop_ext_linear(uint64_t x, uint32_t l) {
  const uint64_t a = dpp64<DPP_QUAD_BCAST0>(x), b = dpp64<DPP_QUAD_BCAST1>(x);
  const uint64_t c = dpp64<DPP_QUAD_BCAST2>(x), d = dpp64<DPP_QUAD_BCAST3>(x);
  const Acc ab = acc_add(acc_of(a), b), cd = acc_add(acc_of(c), d);
  const Acc t = acc_add(ab, cd);
  const Acc tb = acc_add(t, b), td = acc_add(t, d);
  const uint32_t j = l & 3;
  // M4 = [[2,3,1,1],[1,2,3,1],[1,1,2,3],[3,1,1,2]], row j
  const Acc o = j == 0 ? acc_add(tb, ab) : j == 1 ? acc_add(acc_add(tb, c), c) : j == 2 ? acc_add(td, cd)
                                                                                       : acc_add(acc_add(td, a), a);
  // sum of the block outputs at this position (the dummy quad contributes zero)
  Acc s = acc_add(o, dpp_acc<DPP_ROW_ROR4>(o));
  s = acc_add(s, dpp_acc<DPP_ROW_ROR8>(s));
  const uint64_t y = acc_fold(acc_add(o, s));
  return l < 12 ? y : 0;
}

// x: this lane's state word (zero in lanes 12..15); l: lane index within the row.
// TRY / EXACT products as in p2_fast.hpp; the result is not canonicalised.
template <bool TRY>
__device__ __forceinline__ uint64_t p2_permute_coop_body(uint64_t x, uint32_t l, const P2Consts* __restrict__ c,
                                                         uint64_t& flag) {
  const uint32_t li = l < 12 ? l : 0;
  x = coop_ext_linear(x, l);
#pragma unroll 1
  for (int r = 0; r < 4; r++) {
    x = sbox_t<TRY>(add_canon_nc(x, c->rc_init[r][li]), flag);
    x = coop_ext_linear(l < 12 ? x : 0, l);
  }
  const uint64_t dg = c->diag[li];
#pragma unroll 1
  for (int r = 0; r < 22; r++) {
    const uint64_t y0 = sbox_t<TRY>(add_canon_nc(x, c->rc_int[r]), flag);
    x = l == 0 ? y0 : x;
    Acc s = acc_of(l < 12 ? x : 0);
    s = acc_add(s, dpp_acc<DPP_ROW_ROR8>(s));
    s = acc_add(s, dpp_acc<DPP_ROW_ROR4>(s));
    s = acc_add(s, dpp_acc<DPP_ROW_ROR2>(s));
    s = acc_add(s, dpp_acc<DPP_ROW_ROR1>(s));
    const uint64_t sm = acc_fold(s);
    const uint64_t y = mul_add_t<TRY>(x, dg, sm, flag);
    x = l < 12 ? y : 0;
  }
#pragma unroll 1
  for (int r = 0; r < 4; r++) {
    x = sbox_t<TRY>(add_canon_nc(x, c->rc_term[r][li]), flag);
    x = coop_ext_linear(l < 12 ? x : 0, l);
  }
  return x;
}

// TRY products first; the whole wave (four rows) redoes the permutation with
// EXACT products when any lane flagged a product borrow (about 2^-32 per
// product): the one input word per lane is all that has to be kept.
__device__ __forceinline__ uint64_t p2_permute_coop(uint64_t x, uint32_t l, const P2Consts* __restrict__ c) {
  uint64_t flag = 0;
  uint64_t y = p2_permute_coop_body<true>(x, l, c, flag);
  if (__builtin_expect(flag != 0, 0)) y = p2_permute_coop_body<false>(x, l, c, flag);
  return canon(y);
}

}  // namespace p3g
