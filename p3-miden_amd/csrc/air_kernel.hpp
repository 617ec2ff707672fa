// Fixed part of a compiled AIR kernel: constraint evaluation on the quotient domain
// (evaluate_constraints_into, p3-miden-lifted-stark/src/prover/constraints/mod.rs:63-179).
//
// For each AIR program the library writes a source that defines
//   AIR_SELECTORS, AIR_AUX, AIR_PERIODIC   1 when the program reads them, else 0
// includes gl.hpp and this file, and then defines
//   air_point(const AirKArgs&, const AirRow&) -> ef2
//                                          the folded constraints at one point;
// hiprtc compiles it for gfx950 (constraints.cpp).  Host code includes this file too,
// for the argument structures only.
//
// Point i of the quotient domain gJ (natural order, M = n d = 2^L points) is stored
// row bitrev_L(i) of the committed LDE.  Workgroup `mid` takes the 2^(2k) points
//   i = hi 2^(L-k) + mid 2^k + lo        (hi, lo < 2^k)
// whose stored rows are rev_k(lo) 2^(L-k) + rev(mid) 2^k + rev_k(hi): runs of 2^k
// consecutive rows.  Threads evaluate the points in stored-row order (consecutive
// threads read consecutive rows), write the folded values into an LDS tile, and after
// a barrier add them to the accumulator in natural order (consecutive threads update
// consecutive points).
#pragma once

namespace p3g {

constexpr int AIR_T = 256;        // threads per workgroup
constexpr int AIR_PPT = 4;        // points per thread
constexpr unsigned AIR_KMAX = 5;  // tiles of up to 2^5 x 2^5 points (2^10 = AIR_T * AIR_PPT)

struct AirKArgs {
  const uint64_t* main;
  const uint64_t* aux;
  const uint64_t* periodic;
  const uint64_t* pub;    // public values
  const ef2* ext;         // extension inputs
  const uint64_t* apow;   // alpha^e as (c0, c1, 7 c1) for e = 0 .. K-1
  const uint64_t* coset;  // shift w^l (l < 2^k) | w^(h 2^(L-k)) (h < 2^k) | Z_H at i mod d (d values)
  ef2* acc;
  uint64_t main_w, aux_w;  // words per row
  uint64_t prow, pcols;    // periodic table (prow < 2^32)
  uint64_t omega_tile;     // w^(2^k): x at tile index mid carries w^(mid 2^k)
  uint64_t omega_n_inv;    // w_n^-1, the last row of H
  uint32_t logM, k, log_d;
  uint32_t blocks_on;      // all ones: the guard of every basic block of a compiled program
};

// What air_point reads at one point.
struct AirRow {
  const uint64_t* mc;   // main: current row
  const uint64_t* mn;   //       next row
  const uint64_t* ac;   // aux (c0, c1 pairs): current row
  const uint64_t* an;   //       next row
  const uint64_t* per;  // periodic table row
  uint64_t first, last, trans;
};

#if defined(__HIPCC_RTC__)

// the folded constraints at one point: generated per program, after this file
__device__ __forceinline__ ef2 air_point(const AirKArgs& a, const AirRow& r);

// Per-slot values live in small arrays read only at constant indices (slot 0) and
// rotated after each point, so they stay in registers while the point loop stays
// rolled (one copy of air_point); a loop-indexed read would move them to private memory.
template <typename T, int N>
__device__ __forceinline__ void air_rotate(T (&v)[N]) {
#pragma unroll
  for (int t = 0; t + 1 < N; t++) v[t] = v[t + 1];
}

extern "C" __global__ void __launch_bounds__(AIR_T) p3g_air_kernel(AirKArgs a) {
  __shared__ ef2 tile[(1u << AIR_KMAX) * ((1u << AIR_KMAX) + 1)];
  const uint32_t k = a.k, L = a.logM;
  const uint32_t side = 1u << k, mask = side - 1, npts = side << k;
  const uint32_t mid = blockIdx.x;
  const uint32_t Mmask = (uint32_t)((1ull << L) - 1);  // L <= 31
  const uint32_t d = 1u << a.log_d;
  uint32_t pt[AIR_PPT];  // point per slot; idle slots (small domains) hold ~0
#pragma unroll
  for (int j = 0; j < AIR_PPT; j++) {
    const uint32_t u = threadIdx.x + j * AIR_T;
    pt[j] = u < npts ? (bitrev32(u & mask, k) << (L - k)) | (mid << k) | bitrev32(u >> k, k) : ~0u;
  }
#if AIR_SELECTORS
  // x = shift w^lo * w^(hi 2^(L-k)) * w^(mid 2^k); 1/(x - 1) and 1/(x - w_n^-1) for this
  // thread's points with one inversion (Montgomery's trick); then the selectors
  // Z_H(x)/(x - 1), Z_H(x)/(x - w_n^-1) and x - w_n^-1
  const uint64_t wmid = gl_pow(a.omega_tile, mid);
  uint64_t xs[AIR_PPT], den[2 * AIR_PPT];
#pragma unroll
  for (int j = 0; j < AIR_PPT; j++) {
    const uint32_t u = threadIdx.x + j * AIR_T;
    xs[j] = 1;
    den[2 * j] = den[2 * j + 1] = 1;
    if (pt[j] != ~0u) {
      xs[j] = gl_mul(gl_mul(a.coset[bitrev32(u >> k, k)], a.coset[side + bitrev32(u & mask, k)]), wmid);
      den[2 * j] = gl_sub(xs[j], 1);
      den[2 * j + 1] = gl_sub(xs[j], a.omega_n_inv);
    }
  }
  {
    uint64_t pre[2 * AIR_PPT], run = 1;
#pragma unroll
    for (int t = 0; t < 2 * AIR_PPT; t++) {
      pre[t] = run;
      run = gl_mul(run, den[t]);
    }
    uint64_t inv = gl_inv(run);
#pragma unroll
    for (int t = 2 * AIR_PPT - 1; t >= 0; t--) {
      const uint64_t it = gl_mul(inv, pre[t]);
      inv = gl_mul(inv, den[t]);
      den[t] = it;
    }
  }
  uint64_t sel_first[AIR_PPT], sel_last[AIR_PPT], sel_trans[AIR_PPT];
#pragma unroll
  for (int j = 0; j < AIR_PPT; j++) {
    const uint64_t zh = pt[j] != ~0u ? a.coset[2 * side + (pt[j] & (d - 1))] : 0;
    sel_first[j] = gl_mul(zh, den[2 * j]);
    sel_last[j] = gl_mul(zh, den[2 * j + 1]);
    sel_trans[j] = gl_sub(xs[j], a.omega_n_inv);
  }
#endif
#pragma unroll 1
  for (int j = 0; j < AIR_PPT; j++) {
    const uint32_t i = pt[0];
    if (i != ~0u) {
      const uint64_t cur = bitrev32(i, L), nxt = bitrev32((i + d) & Mmask, L);
      AirRow r{};
      r.mc = a.main + cur * a.main_w;
      r.mn = a.main + nxt * a.main_w;
#if AIR_AUX
      r.ac = a.aux + cur * a.aux_w;
      r.an = a.aux + nxt * a.aux_w;
#endif
#if AIR_PERIODIC
      r.per = a.periodic + (uint64_t)(i % (uint32_t)a.prow) * a.pcols;
#endif
#if AIR_SELECTORS
      r.first = sel_first[0];
      r.last = sel_last[0];
      r.trans = sel_trans[0];
#endif
      const uint32_t u = threadIdx.x + j * AIR_T;
      tile[bitrev32(u & mask, k) * (side + 1) + bitrev32(u >> k, k)] = air_point(a, r);
    }
    air_rotate(pt);
#if AIR_SELECTORS
    air_rotate(sel_first);
    air_rotate(sel_last);
    air_rotate(sel_trans);
#endif
  }
  __syncthreads();
  // the accumulator in natural order: all loads first, then the sums
  ef2 old[AIR_PPT];
#pragma unroll
  for (int j = 0; j < AIR_PPT; j++) {
    const uint32_t v = threadIdx.x + j * AIR_T;
    if (v < npts) old[j] = a.acc[((v >> k) << (L - k)) | (mid << k) | (v & mask)];
  }
#pragma unroll
  for (int j = 0; j < AIR_PPT; j++) {
    const uint32_t v = threadIdx.x + j * AIR_T;
    if (v < npts) {
      const uint32_t hi = v >> k, lo = v & mask;
      a.acc[(hi << (L - k)) | (mid << k) | lo] = ef_add(old[j], tile[hi * (side + 1) + lo]);
    }
  }
}

#endif  // __HIPCC_RTC__

}  // namespace p3g
