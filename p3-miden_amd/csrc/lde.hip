// Coset low-degree extension with bit-reversed output rows, for gfx950.
//
// Replaces `TwoAdicSubgroupDft::coset_lde_batch(trace, added_bits, shift)` followed
// by `.bit_reverse_rows()` (p3-miden-lifted-stark/src/prover/commit.rs:217-221).
//
// Layout: the trace is row-major n x w (u64 words); the output is row-major
// (n*B) x w.  Output row r holds f(shift * omega_{nB}^{bitrev(r)}) per column.
//
// Method (one permutation, no zero padding):
//   1. inverse NTT of size n, decimation in time: rows are gathered in
//      bit-reversed order straight from the trace (each row is a contiguous
//      w-word segment, so the gather stays coalesced) and the natural-order
//      coefficients land in output block 0;
//   2. the LDE over nB points splits into B cosets c_t * H_n with
//      c_t = shift * omega_{nB}^t.  For each t the coefficients are scaled by
//      c_t^k / n and a size-n decimation-in-frequency NTT produces the
//      bit-reversed evaluations, which are exactly output rows
//      [bitrev_b(t)*n, (bitrev_b(t)+1)*n).  The first forward pass reads the
//      coefficient tile once and emits all B cosets.
// Each pass runs up to NTT_K butterfly stages on a 2^K x C tile staged in LDS;
// twiddles come from two-level power tables (one multiply per twiddle),
// computed once per workgroup into LDS.
#include <cstring>

#include "internal.hpp"
#include "ntt_ops.hpp"

namespace p3g {

constexpr int NTT_K = 8;
constexpr int NTT_C = 16;
constexpr int NTT_T = 256;

enum NttMode : uint32_t { INV_FIRST = 0, INV = 1, FWD_FIRST = 2, FWD = 3 };

struct NttArgs {
  const uint64_t* src;
  uint64_t* dst;
  const uint64_t* tw;     // power table of omega_n^{-1} (inverse) or omega_n (forward)
  const uint64_t* scale;  // FWD_FIRST: B power tables of c_t with 1/n folded into lo
  uint64_t scale_stride;  // words per scale table
  uint64_t n;
  uint32_t logn, s0, K, w, C, ncoltiles, logB, mode;
  uint32_t logC;  // C = 2^logC
};

__device__ __forceinline__ uint64_t pw(const uint64_t* __restrict__ tab, uint64_t e) {
  return gl_mul(tab[e & (POW_LO - 1)], tab[POW_LO + (e >> 11)]);
}

__device__ __forceinline__ void ntt_pass_body(const NttArgs& a, uint64_t bid) {
  __shared__ uint64_t tile[(1 << NTT_K) * NTT_C];
  __shared__ uint64_t tws[1 << NTT_K];
  const uint32_t K = a.K, C = a.C, tid = threadIdx.x;
  const bool inverse = a.mode <= INV;
  const uint32_t ct = (uint32_t)(bid % a.ncoltiles);
  bid /= a.ncoltiles;
  // unit decomposition
  uint32_t lb;  // number of low index bits below the unit's butterfly bits
  if (inverse) lb = a.s0;
  else lb = a.logn - a.s0 - K;
  const uint64_t nunits = a.n >> K;
  const uint64_t u = bid % nunits;
  const uint64_t blk = bid / nunits;  // FWD: which of the B transforms
  const uint64_t lo = u & ((1ull << lb) - 1);
  const uint64_t hi = u >> lb;
  const uint64_t base = hi << (lb + K) | lo;  // index of x = 0
  const uint64_t xstride = 1ull << lb;
  const uint32_t col0 = ct * C;
  const uint32_t nel = (1u << K) * C;

  // twiddles for this unit
  for (uint32_t e = tid; e + 1 < (1u << K); e += NTT_T) {
    // e = (2^t - 1) + m
    uint32_t t = 31 - __clz(e + 1);
    uint32_t m = e + 1 - (1u << t);
    uint64_t ex;
    if (inverse) {
      // DIT stage t: j = m*2^s0 + lo, exponent j * 2^(logn - s0 - t - 1)
      ex = ((uint64_t)m << a.s0 | lo) << (a.logn - a.s0 - t - 1);
    } else {
      // DIF stage with q = t here (table slot indexed by q): j = m*2^lb + lo,
      // exponent j * 2^(s0 + K - 1 - q)
      ex = ((uint64_t)m << lb | lo) << (a.s0 + K - 1 - t);
    }
    tws[e] = pw(a.tw, ex);
  }

  constexpr int PER = (1 << NTT_K) * NTT_C / NTT_T;
  uint64_t regs[PER];
#pragma unroll
  for (int r = 0; r < PER; r++) {
    uint32_t e = tid + r * NTT_T;
    if (e < nel) {
      uint32_t x = e >> a.logC, c = e & (C - 1);
      uint32_t col = col0 + c;
      uint64_t idx = base + (uint64_t)x * xstride;
      uint64_t v = 0;
      if (col < a.w) {
        if (a.mode == INV_FIRST) {
          uint64_t srow = bitrev32((uint32_t)idx, a.logn);
          v = a.src[srow * a.w + col];
        } else {
          uint64_t row = (a.mode == FWD ? blk * a.n : 0) + idx;
          v = a.dst[row * a.w + col];
        }
      }
      regs[r] = v;
    }
  }

  const uint32_t nbf = (1u << (K ? K - 1 : 0)) * C;
  const uint32_t nrep = a.mode == FWD_FIRST ? (1u << a.logB) : 1;
  for (uint32_t t_rep = 0; t_rep < nrep; t_rep++) {
    __syncthreads();
#pragma unroll
    for (int r = 0; r < PER; r++) {
      uint32_t e = tid + r * NTT_T;
      if (e < nel) {
        uint64_t v = regs[r];
        if (a.mode == FWD_FIRST) {
          uint32_t x = e >> a.logC;
          uint64_t k = base + (uint64_t)x * xstride;
          v = gl_mul(v, pw(a.scale + t_rep * a.scale_stride, k));
        }
        tile[e] = v;
      }
    }
    __syncthreads();
    for (uint32_t t = 0; t < K; t++) {
      for (uint32_t b = tid; b < nbf; b += NTT_T) {
        uint32_t p = b >> a.logC, c = b & (C - 1);
        if (inverse) {
          uint32_t mlo = p & ((1u << t) - 1);
          uint32_t x0 = ((p >> t) << (t + 1)) | mlo;
          uint32_t x1 = x0 | (1u << t);
          uint64_t twv = tws[(1u << t) - 1 + mlo];
          uint64_t uu = tile[x0 * C + c];
          uint64_t vv = gl_mul(tile[x1 * C + c], twv);
          tile[x0 * C + c] = gl_add(uu, vv);
          tile[x1 * C + c] = gl_sub(uu, vv);
        } else {
          uint32_t q = K - 1 - t;
          uint32_t mlo = p & ((1u << q) - 1);
          uint32_t x0 = ((p >> q) << (q + 1)) | mlo;
          uint32_t x1 = x0 | (1u << q);
          uint64_t twv = tws[(1u << q) - 1 + mlo];
          uint64_t uu = tile[x0 * C + c];
          uint64_t vv = tile[x1 * C + c];
          tile[x0 * C + c] = gl_add(uu, vv);
          tile[x1 * C + c] = gl_mul(gl_sub(uu, vv), twv);
        }
      }
      __syncthreads();
    }
    uint64_t out_block;
    if (a.mode == FWD_FIRST) out_block = bitrev32(t_rep, a.logB);
    else if (a.mode == FWD) out_block = blk;
    else out_block = 0;
    for (uint32_t e = tid; e < nel; e += NTT_T) {
      uint32_t x = e >> a.logC, c = e & (C - 1);
      uint32_t col = col0 + c;
      if (col < a.w) {
        uint64_t row = out_block * a.n + base + (uint64_t)x * xstride;
        a.dst[row * a.w + col] = tile[e];
      }
    }
  }
}


// Register-blocked pass of LR stages (5 <= LR <= 8) on a 2^LR-row x 16-column
// tile with 2^LR threads: each thread owns one column and 16 rows, performs the
// stages on x bits 0..3 in registers (layout A: x = 16 g + k), exchanges through
// an LDS tile once, and performs the stages on x bits 4..LR-1 in registers
// (layout B: x = g + G k with G = 2^(LR-4)).  Inverse passes run A then B,
// forward passes B then A.
//
// The exchange moves the low 32-bit halves, then the high halves, through a
// tile of one 32-bit word per element: half the LDS of a 64-bit tile, so more
// workgroups fit on a CU (a 128-thread pass: 4 -> up to 7 waves per SIMD).
// Word of element x, column c: rows of 16 words, XOR-swizzled by the row
// group, with 16 words of spacing after every 16 rows so that the row groups
// of one access land in different banks.
constexpr uint32_t xchg_words(uint32_t rows) { return rows * 16 + (rows / 16) * 16; }
__device__ __forceinline__ uint32_t swz(uint32_t x, uint32_t c) {
  return x * 16 + (c ^ ((x >> 4) & 15)) + (x >> 4) * 16;
}

// Every thread writes its 16 values at wpos(k) and reads back the 16 at
// rpos(k), low halves first.  The tile must be free on entry; it is in use
// until the caller's next barrier.
template <typename WP, typename RP>
__device__ __forceinline__ void xchg_halves(uint32_t* tile, uint64_t (&r)[16], WP wpos, RP rpos) {
  uint32_t hi[16], lo[16];
#pragma unroll
  for (int k = 0; k < 16; k++) {
    tile[wpos(k)] = (uint32_t)r[k];
    hi[k] = (uint32_t)(r[k] >> 32);
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 16; k++) lo[k] = tile[rpos(k)];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 16; k++) tile[wpos(k)] = hi[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 16; k++) r[k] = ((uint64_t)tile[rpos(k)] << 32) | lo[k];
}

// Field operations of a pass.  add_c / sub_c / mulc / mul_pow2c take the
// fewest issue slots; two passes keep some carry-chain forms for their
// register budget:
//  - the coset-expanding first forward pass holds its input tile across the
//    coset loop and spills with add_c's wider live ranges (gl_add there);
//  - the inverse first pass gathers bit-reversed rows and is latency-bound:
//    the carry-chain forms compile it to 80 VGPRs (6 waves per SIMD) where the
//    others take 99 (4 waves), which measured 13% slower.
template <uint32_t MODE>
struct NttOps {
  static __device__ __forceinline__ uint64_t add(uint64_t x, uint64_t y) {
    if constexpr (MODE == FWD_FIRST || MODE == INV_FIRST) return gl_add(x, y);
    else return add_c(x, y);
  }
  static __device__ __forceinline__ uint64_t sub(uint64_t x, uint64_t y) {
    if constexpr (MODE == INV_FIRST) return gl_sub(x, y);
    else return sub_c(x, y);
  }
  static __device__ __forceinline__ uint64_t mul(uint64_t x, uint64_t y) {
    if constexpr (MODE == INV_FIRST) return mulc_cc(x, y);
    else return mulc(x, y);
  }
  static __device__ __forceinline__ uint64_t pow2(uint64_t x, uint32_t E) {
    if constexpr (MODE == INV_FIRST) return mul_pow2c_cc(x, E);
    else return mul_pow2c(x, E);
  }
};

template <uint32_t MODE, int LR>
__device__ __forceinline__ void ntt_reg_body(const NttArgs& a, uint64_t bid) {
  using Ops = NttOps<MODE>;
  constexpr uint32_t T = 1u << LR;      // threads = rows per tile
  constexpr uint32_t G = 1u << (LR - 4);  // row groups in layout A / row stride in layout B
  constexpr int KB0 = 8 - LR;             // layout B: k bits below KB0 belong to layout-A stages
  __shared__ uint32_t tile[xchg_words(T)];
  __shared__ uint64_t tws[T];
  __shared__ uint64_t tfa[16];  // layout-A twist factors (the same for every thread)
  __shared__ uint64_t tfb[T];   // layout-B twist factors, entry 16 g + k
  __shared__ uint64_t scl[MODE == FWD_FIRST ? G * 17 : 1];  // coset scale factors, entry 17 g + k
  const uint32_t tid = threadIdx.x;
  const uint32_t c = tid & 15, g = tid >> 4;
  constexpr bool inverse = MODE <= INV;
  const uint32_t ct = (uint32_t)(bid % a.ncoltiles);
  bid /= a.ncoltiles;
  const uint32_t lb = inverse ? a.s0 : a.logn - a.s0 - LR;
  const uint64_t nunits = a.n >> LR;
  const uint64_t u = bid % nunits;
  const uint64_t blk = bid / nunits;
  const uint64_t lo = u & ((1ull << lb) - 1);
  const uint64_t hi = u >> lb;
  const uint64_t base = hi << (lb + LR) | lo;
  const uint64_t xstride = 1ull << lb;
  const uint32_t col = ct * 16 + c;
  const bool live = col < a.w;

  for (uint32_t e = tid; e < T - 1; e += T) {
    uint32_t t = 31 - __clz(e + 1);
    uint32_t m = e + 1 - (1u << t);
    uint64_t ex;
    if (inverse) ex = ((uint64_t)m << a.s0 | lo) << (a.logn - a.s0 - t - 1);
    else ex = ((uint64_t)m << lb | lo) << (a.s0 + LR - 1 - t);
    tws[e] = pw(a.tw, ex);
  }

  // one branch around all 16 loads (columns past the width read as zero)
  uint64_t in[16];
  if (live) {
#pragma unroll
    for (int k = 0; k < 16; k++) {
      const uint32_t x = inverse ? 16 * g + k : g + G * k;
      const uint64_t idx = base + (uint64_t)x * xstride;
      if (MODE == INV_FIRST) in[k] = a.src[(uint64_t)bitrev32((uint32_t)idx, a.logn) * a.w + col];
      else in[k] = a.dst[((MODE == FWD ? blk * a.n : 0) + idx) * a.w + col];
    }
  } else {
#pragma unroll
    for (int k = 0; k < 16; k++) in[k] = 0;
  }
  __syncthreads();  // twiddles ready

  // Each group of stages on the 4 k bits is a twist by per-element factors
  // followed by butterflies whose twiddles are 16th roots of unity (shifts):
  // the stage-t twiddle tws[2^t - 1 + m] factors into omega_{2^(t+1)}^(high
  // bits of m) times an entry that does not depend on the butterfly's active k
  // bits, and those entries multiply out to one factor per element.  Inverse
  // (DIT) groups twist before their butterflies, forward (DIF) groups after.
  // Layout A: stages t = 0..3 are k bits 0..3, factor entries tws[2^t - 1].
  // Layout B: stages t = 4..LR-1 are k bits KB0..3; k bits below KB0 (s) and
  // the row group g select the entries tws[2^t - 1 + g + G s].
  for (uint32_t e = tid; e < 16; e += T) {
    uint64_t f = 1;
#pragma unroll
    for (int kb = 0; kb < 4; kb++)
      if ((e >> kb) & 1) f = gl_mul(f, tws[(1u << kb) - 1]);
    tfa[e] = f;
  }
  {
    const uint32_t gq = tid >> 4, k = tid & 15;  // T = 16 G entries, one per thread
    const uint32_t sb = k & ((1u << KB0) - 1);
    uint64_t f = 1;
#pragma unroll
    for (int kb = KB0; kb < 4; kb++)
      if ((k >> kb) & 1) f = gl_mul(f, tws[(1u << (kb + LR - 4)) - 1 + gq + G * sb]);
    tfb[tid] = f;
  }
  __syncthreads();  // twist factors ready

  const int nrep = MODE == FWD_FIRST ? (1 << a.logB) : 1;
  // cosets in descending order (block 0 last, as in every pass order)
  for (int rep = nrep - 1; rep >= 0; rep--) {
    uint64_t r[16];
    if (MODE == FWD_FIRST) {
      // coefficient k_glob = base + (g + G k) xstride is scaled by c_t^k_glob / n
      // (the power table's lo half carries 1/n): each thread computes one
      // factor of the tile into LDS, then multiplies its 16 elements
      const uint64_t* sc = a.scale + rep * a.scale_stride;
      {
        const uint32_t gq = tid >> 4, kq = tid & 15;
        scl[17 * gq + kq] = pw(sc, base + (uint64_t)(gq + G * kq) * xstride);
      }
      __syncthreads();  // the previous coset's readers passed the tile barriers below
#pragma unroll
      for (int k = 0; k < 16; k++) r[k] = Ops::mul(in[k], scl[17 * g + k]);
    } else {
#pragma unroll
      for (int k = 0; k < 16; k++) r[k] = in[k];
    }
    if (inverse) {
      // stages on x bits 0..3 (k bits), layout A
#pragma unroll
      for (int k = 1; k < 16; k++) r[k] = Ops::mul(r[k], tfa[k]);
#pragma unroll
      for (int t = 0; t < 4; t++) {
#pragma unroll
        for (int k0 = 0; k0 < 16; k0++) {
          if (k0 & (1 << t)) continue;
          const int k1 = k0 | (1 << t);
          const uint32_t jw = (uint32_t)(k0 & ((1 << t) - 1)) << (3 - t);
          uint32_t E;
          bool neg;
          w16_pow2(jw, true, E, neg);  // omega^-jw = (-1)^neg 2^E
          const uint64_t x = r[k0], y = Ops::pow2(r[k1], E);
          r[k0] = neg ? Ops::sub(x, y) : Ops::add(x, y);
          r[k1] = neg ? Ops::add(x, y) : Ops::sub(x, y);
        }
      }
      xchg_halves(tile, r, [&](int k) { return swz(16 * g + k, c); }, [&](int k) { return swz(g + G * k, c); });
      // stages on x bits 4..LR-1, layout B (x bit t is k bit t - (LR - 4))
#pragma unroll
      for (int k = 0; k < 16; k++)
        if (k >> KB0) r[k] = Ops::mul(r[k], tfb[16 * g + k]);
#pragma unroll
      for (int t = 4; t < LR; t++) {
        const int kb = t - (LR - 4);
#pragma unroll
        for (int k0 = 0; k0 < 16; k0++) {
          if (k0 & (1 << kb)) continue;
          const int k1 = k0 | (1 << kb);
          const uint32_t jw = (uint32_t)((k0 >> KB0) & ((1 << (kb - KB0)) - 1)) << (3 - kb + KB0);
          uint32_t E;
          bool neg;
          w16_pow2(jw, true, E, neg);  // omega^-jw = (-1)^neg 2^E
          const uint64_t x = r[k0], y = Ops::pow2(r[k1], E);
          r[k0] = neg ? Ops::sub(x, y) : Ops::add(x, y);
          r[k1] = neg ? Ops::add(x, y) : Ops::sub(x, y);
        }
      }
      if (live) {
#pragma unroll
        for (int k = 0; k < 16; k++) a.dst[(base + (uint64_t)(g + G * k) * xstride) * a.w + col] = r[k];
      }
    } else {
      // DIF: x bits LR-1..4 first (layout B), then bits 3..0 (layout A)
#pragma unroll
      for (int t = 0; t < LR - 4; t++) {
        const int q = LR - 1 - t;
        const int kb = q - (LR - 4);
#pragma unroll
        for (int k0 = 0; k0 < 16; k0++) {
          if (k0 & (1 << kb)) continue;
          const int k1 = k0 | (1 << kb);
          const uint32_t jw = (uint32_t)((k0 >> KB0) & ((1 << (kb - KB0)) - 1)) << (3 - kb + KB0);
          uint32_t E;
          bool neg;
          w16_pow2(jw, false, E, neg);  // omega^jw = (-1)^neg 2^E
          const uint64_t x = r[k0], y = r[k1];
          r[k0] = Ops::add(x, y);
          r[k1] = Ops::pow2(neg ? Ops::sub(y, x) : Ops::sub(x, y), E);
        }
      }
#pragma unroll
      for (int k = 0; k < 16; k++)
        if (k >> KB0) r[k] = Ops::mul(r[k], tfb[16 * g + k]);
      __syncthreads();  // previous repetition's readers are done with the tile
      xchg_halves(tile, r, [&](int k) { return swz(g + G * k, c); }, [&](int k) { return swz(16 * g + k, c); });
#pragma unroll
      for (int q = 3; q >= 0; q--) {
#pragma unroll
        for (int k0 = 0; k0 < 16; k0++) {
          if (k0 & (1 << q)) continue;
          const int k1 = k0 | (1 << q);
          const uint32_t jw = (uint32_t)(k0 & ((1 << q) - 1)) << (3 - q);
          uint32_t E;
          bool neg;
          w16_pow2(jw, false, E, neg);  // omega^jw = (-1)^neg 2^E
          const uint64_t x = r[k0], y = r[k1];
          r[k0] = Ops::add(x, y);
          r[k1] = Ops::pow2(neg ? Ops::sub(y, x) : Ops::sub(x, y), E);
        }
      }
#pragma unroll
      for (int k = 1; k < 16; k++) r[k] = Ops::mul(r[k], tfa[k]);
      const uint64_t ob = MODE == FWD_FIRST ? bitrev32((uint32_t)rep, a.logB) : (MODE == FWD ? blk : 0);
      // the row stride is made opaque inside the coset loop so that the 16 store
      // addresses are formed per coset instead of being held across it
      uint64_t xs = xstride;
      if (MODE == FWD_FIRST) asm volatile("" : "+v"(xs));
      if (live) {
        uint64_t* out = a.dst + ((ob * a.n + base + (uint64_t)(16 * g) * xs) * a.w + col);
        const uint64_t rs = xs * a.w;
#pragma unroll
        for (int k = 0; k < 16; k++) out[(uint64_t)k * rs] = r[k];
      }
    }
  }
}

__global__ void __launch_bounds__(NTT_T) __attribute__((amdgpu_waves_per_eu(4))) ntt_pass_kernel(NttArgs a) {
  ntt_pass_body(a, blockIdx.x);
}

// Register-blocked kernels: 2^LR threads per workgroup, at least 4 waves per
// EU (3 for FWD_FIRST, whose coset loop keeps the coefficient tile live).
template <uint32_t MODE, int LR>
__global__ void __launch_bounds__(1 << LR) __attribute__((amdgpu_waves_per_eu(MODE == FWD_FIRST ? 3 : 4)))
ntt_reg_kernel(NttArgs a) {
  ntt_reg_body<MODE, LR>(a, blockIdx.x);
}

// One launch for the same pass of several transforms (the traces of a batch):
// block b belongs to descriptor k with starts[k] <= b < starts[k + 1].
__device__ __forceinline__ uint32_t find_desc(const uint64_t* __restrict__ starts, uint32_t nd, uint64_t b) {
  uint32_t k = 0;
  while (k + 1 < nd && starts[k + 1] <= b) k++;
  return k;
}

__global__ void __launch_bounds__(NTT_T) __attribute__((amdgpu_waves_per_eu(4)))
ntt_pass_grouped_kernel(const NttArgs* __restrict__ d, const uint64_t* __restrict__ starts, uint32_t nd) {
  const uint32_t k = find_desc(starts, nd, blockIdx.x);
  const NttArgs a = d[k];
  ntt_pass_body(a, blockIdx.x - starts[k]);
}

template <uint32_t MODE, int LR>
__global__ void __launch_bounds__(1 << LR) __attribute__((amdgpu_waves_per_eu(MODE == FWD_FIRST ? 3 : 4)))
ntt_reg_grouped_kernel(const NttArgs* __restrict__ d, const uint64_t* __restrict__ starts, uint32_t nd) {
  const uint32_t k = find_desc(starts, nd, blockIdx.x);
  const NttArgs a = d[k];
  ntt_reg_body<MODE, LR>(a, blockIdx.x - starts[k]);
}

// host-side dispatch over (mode, LR) for the register-blocked kernels
template <int LR>
static void launch_reg_lr(uint32_t mode, const NttArgs* a, const NttArgs* d, const uint64_t* starts, uint32_t nd,
                          uint64_t blocks, hipStream_t s) {
  const dim3 grid((uint32_t)blocks), block(1u << LR);
  if (a) {
    switch (mode) {
      case INV_FIRST: hipLaunchKernelGGL((ntt_reg_kernel<INV_FIRST, LR>), grid, block, 0, s, *a); break;
      case INV: hipLaunchKernelGGL((ntt_reg_kernel<INV, LR>), grid, block, 0, s, *a); break;
      case FWD_FIRST: hipLaunchKernelGGL((ntt_reg_kernel<FWD_FIRST, LR>), grid, block, 0, s, *a); break;
      default: hipLaunchKernelGGL((ntt_reg_kernel<FWD, LR>), grid, block, 0, s, *a); break;
    }
  } else {
    switch (mode) {
      case INV_FIRST: hipLaunchKernelGGL((ntt_reg_grouped_kernel<INV_FIRST, LR>), grid, block, 0, s, d, starts, nd); break;
      case INV: hipLaunchKernelGGL((ntt_reg_grouped_kernel<INV, LR>), grid, block, 0, s, d, starts, nd); break;
      case FWD_FIRST: hipLaunchKernelGGL((ntt_reg_grouped_kernel<FWD_FIRST, LR>), grid, block, 0, s, d, starts, nd); break;
      default: hipLaunchKernelGGL((ntt_reg_grouped_kernel<FWD, LR>), grid, block, 0, s, d, starts, nd); break;
    }
  }
}

// Launches a register-blocked pass of K stages (5 <= K <= 8): either one
// transform (a != nullptr) or a grouped launch over descriptors d/starts.
static void launch_reg(uint32_t K, uint32_t mode, const NttArgs* a, const NttArgs* d, const uint64_t* starts,
                       uint32_t nd, uint64_t blocks, hipStream_t s) {
  switch (K) {
    case 5: launch_reg_lr<5>(mode, a, d, starts, nd, blocks, s); break;
    case 6: launch_reg_lr<6>(mode, a, d, starts, nd, blocks, s); break;
    case 7: launch_reg_lr<7>(mode, a, d, starts, nd, blocks, s); break;
    default: launch_reg_lr<8>(mode, a, d, starts, nd, blocks, s); break;
  }
  P3G_CHECK_HIP(hipGetLastError());
}

bool ntt_reg_ok(uint32_t K, uint32_t logC) { return logC == 4 && K >= 5 && K <= 8; }

std::vector<uint32_t> ntt_pass_split(uint32_t logn, uint32_t kmax);
void launch_ntt_grouped(uint32_t mode, bool reg, uint32_t K, const void* descs, const uint64_t* starts, uint32_t nd,
                        uint64_t blocks, hipStream_t s);

// coefficient k of column c times root^(k * (c / group)) (root's powers in tab)
__global__ void coef_twist_kernel(uint64_t* __restrict__ coef, uint64_t n, uint32_t w, const uint64_t* __restrict__ tab,
                                  uint32_t group) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * w) return;
  const uint64_t k = i / w;
  const uint32_t c = (uint32_t)(i - k * w);
  const uint64_t e = k * (c / group);
  if (e) coef[i] = gl_mul(coef[i], pw(tab, e));
}

// Twiddle and coset-scale tables of one LDE shape, wherever they live.
struct LdeTables {
  const uint64_t* tw_inv;  // power table of omega_n^-1
  const uint64_t* tw_fwd;  // power table of omega_n
  const uint64_t* scale;   // B power tables of c_t = shift * omega_nB^t with 1/n folded into lo
};

static void lde_check_shape(uint64_t n, unsigned log_blowup) {
  if (n == 0 || (n & (n - 1))) throw ApiError(E_NOT_POW2, "trace height must be a power of two");
  if (log2_exact(n) + log_blowup > 32) throw ApiError(E_RANGE, "LDE height exceeds the field two-adicity");
}

static void lde_coset_launches(Ctx& ctx, const uint64_t* d_trace, uint64_t n, uint64_t w, unsigned log_blowup,
                               uint64_t* d_out, hipStream_t s, const ColTwist* twist, const LdeTables& tabs);

void lde_coset_batch(Ctx& ctx, const uint64_t* d_trace, uint64_t n, uint64_t w, unsigned log_blowup,
                     uint64_t shift, uint64_t* d_out, hipStream_t s, const ColTwist* twist) {
  lde_check_shape(n, log_blowup);
  if (w == 0) return;
  const unsigned logn = log2_exact(n);
  const uint32_t B = 1u << log_blowup;
  const uint32_t twbits = logn ? logn - 1 : 0;
  const uint64_t* tw_inv = ctx.pow_table(gl_inv(gl_two_adic_generator(logn)), twbits);
  const uint64_t* tw_fwd = ctx.pow_table(gl_two_adic_generator(logn), twbits);
  // scale tables for the B cosets, stacked in one allocation
  const uint64_t ninv = gl_inv_pow2(logn);
  const uint64_t stride = POW_LO + pow_table_hi_len(logn);
  uint64_t* scale = (uint64_t*)ctx.alloc(stride * B * sizeof(uint64_t));
  try {
    std::vector<uint64_t> host(stride * B);
    const uint64_t om = gl_two_adic_generator(logn + log_blowup);
    uint64_t ct = shift;
    for (uint32_t t = 0; t < B; t++) {
      uint64_t* tab = host.data() + t * stride;
      uint64_t v = ninv;
      for (uint32_t i = 0; i < POW_LO; i++) {
        tab[i] = v;
        v = gl_mul(v, ct);
      }
      uint64_t step = gl_pow(ct, POW_LO), h = 1;
      for (uint64_t i = 0; i < pow_table_hi_len(logn); i++) {
        tab[POW_LO + i] = h;
        h = gl_mul(h, step);
      }
      ct = gl_mul(ct, om);
    }
    ctx.upload_async(scale, host.data(), host.size() * 8, s);
    lde_coset_launches(ctx, d_trace, n, w, log_blowup, d_out, s, twist, LdeTables{tw_inv, tw_fwd, scale});
  } catch (...) {
    ctx.release(scale);
    throw;
  }
  ctx.release(scale);
}

static void lde_coset_launches(Ctx& ctx, const uint64_t* d_trace, uint64_t n, uint64_t w, unsigned log_blowup,
                               uint64_t* d_out, hipStream_t s, const ColTwist* twist, const LdeTables& tabs) {
  const unsigned logn = log2_exact(n);
  uint32_t logC = 0;
  while ((2u << logC) <= NTT_C && (2ull << logC) <= w) logC++;
  if (logC == 3) {
    // widths 8..15: register tiles of 16 columns, those past the width masked,
    // beat the per-stage LDS kernel on 8-column tiles (a 2^22 x 8 -> 2^24 LDE
    // and its tree: 18.0 -> 17.6 ms); at widths 4..7 three quarters of the lanes
    // would idle and the LDS kernel stays faster.  Every pass must be 5..8 stages.
    bool reg = true;
    for (uint32_t k : ntt_pass_split(logn, NTT_K)) reg = reg && ntt_reg_ok(k, 4);
    if (reg) logC = 4;
  }
  const uint32_t C = 1u << logC;
  const uint32_t ncol = (uint32_t)((w + C - 1) / C);
  const uint32_t B = 1u << log_blowup;
  const uint64_t* tw_inv = tabs.tw_inv;
  const uint64_t* tw_fwd = tabs.tw_fwd;
  const uint64_t* scale = tabs.scale;
  const uint64_t stride = POW_LO + pow_table_hi_len(logn);

  auto launch = [&](uint32_t mode, uint32_t s0, uint32_t K, uint64_t nblk) {
    NttArgs a{};
    a.src = d_trace;
    a.dst = d_out;
    a.tw = (mode <= INV) ? tw_inv : tw_fwd;
    a.scale = scale;
    a.scale_stride = stride;
    a.n = n;
    a.logn = logn;
    a.s0 = s0;
    a.K = K;
    a.w = (uint32_t)w;
    a.C = C;
    a.logC = logC;
    a.ncoltiles = ncol;
    a.logB = log_blowup;
    a.mode = mode;
    uint64_t grid = (n >> K) * ncol * nblk;
    ProfScope ps(ctx, mode <= INV ? "lde_inverse_pass" : "lde_forward_pass", s);
    if (ntt_reg_ok(K, logC)) {
      launch_reg(K, mode, &a, nullptr, nullptr, 0, grid, s);
    } else {
      hipLaunchKernelGGL(ntt_pass_kernel, dim3((uint32_t)grid), dim3(NTT_T), 0, s, a);
      P3G_CHECK_HIP(hipGetLastError());
    }
  };

  // inverse (DIT), first pass gathers bit-reversed rows from the trace;
  // then forward (DIF) per coset, the first pass emitting all B cosets
  const std::vector<uint32_t> split = ntt_pass_split(logn, NTT_K);
  {
    uint32_t s0 = 0;
    for (size_t i = 0; i < split.size(); i++) {
      launch(i == 0 ? INV_FIRST : INV, s0, split[i], 1);
      s0 += split[i];
    }
  }
  if (twist) {
    // exponents k * (c / group) < n * ceil(w / group) must stay below the root's order
    const uint64_t groups = (w + twist->group - 1) / twist->group;
    if (twist->group == 0 || twist->log_order > 32 || n * groups > (1ull << twist->log_order))
      throw ApiError(E_RANGE, "coefficient twist exponent exceeds the root's order");
    const uint64_t* tab = ctx.pow_table(twist->root, twist->log_order);
    const uint64_t total = n * w;
    hipLaunchKernelGGL(coef_twist_kernel, dim3((uint32_t)((total + 255) / 256)), dim3(256), 0, s, d_out, n,
                       (uint32_t)w, tab, twist->group);
    P3G_CHECK_HIP(hipGetLastError());
  }
  {
    uint32_t s0 = 0;
    for (size_t i = 0; i < split.size(); i++) {
      launch(i == 0 ? FWD_FIRST : FWD, s0, split[i], i == 0 ? 1 : B);
      s0 += split[i];
    }
  }
}

// ------------------------------------------------------------------ the LDE without allocation
// One array per GPU for the twiddle and coset-scale tables of the no-alloc
// entry point, so that it makes no allocation of its own.  In host code the
// name `scratch` is only a stand-in: its device address is fetched once per
// device (lde_scratch_init, called when a context is created with that device
// current) and only that address is handed to kernels.  Every call on one GPU
// shares the array, so calls that overlap on one GPU race.
constexpr uint64_t LDE_SCRATCH_WORDS = 96 * 1024;  // 768 KiB
static __device__ __align__(256) uint64_t scratch[LDE_SCRATCH_WORDS];

static std::mutex g_scratch_mu;
static std::map<int, uint64_t*> g_scratch_ptr;

void lde_scratch_init(int device) {
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (g_scratch_ptr.count(device)) return;
  void* ptr = nullptr;
  P3G_CHECK_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  g_scratch_ptr[device] = static_cast<uint64_t*>(ptr);
}

static uint64_t* lde_scratch(int device) {
  std::lock_guard<std::mutex> g(g_scratch_mu);
  auto it = g_scratch_ptr.find(device);
  if (it == g_scratch_ptr.end()) throw ApiError(E_INVAL, "no scratch address for this device (context not set up)");
  return it->second;
}

// Words of table space an LDE of 2^logn rows and blowup 2^log_blowup needs.
uint64_t lde_scratch_words(unsigned logn, unsigned log_blowup) {
  const unsigned twbits = logn ? logn - 1 : 0;
  return 2 * (POW_LO + pow_table_hi_len(twbits)) + (uint64_t(1) << log_blowup) * (POW_LO + pow_table_hi_len(logn));
}

uint64_t lde_scratch_capacity() { return LDE_SCRATCH_WORDS; }

// Up to 8 power tables of `len` words at dst + t * stride, in the layout of
// Ctx::pow_table: [lo_mul * base^i, i < POW_LO][base^(POW_LO j)].  Each entry
// is one exponentiation, so the tables come out of one launch with no host data.
struct PowFill {
  uint64_t base[8];
  uint64_t lo_mul;
  uint64_t* dst;
  uint64_t stride, len;
};

__global__ void pow_fill_kernel(PowFill f) {
  const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const uint32_t t = blockIdx.y;
  if (i >= f.len) return;
  const uint64_t b = f.base[t];
  f.dst[t * f.stride + i] = i < POW_LO ? gl_mul(f.lo_mul, gl_pow(b, i)) : gl_pow(b, uint64_t(POW_LO) * (i - POW_LO));
}

static void pow_fill(const PowFill& f, uint32_t ntab, hipStream_t s) {
  hipLaunchKernelGGL(pow_fill_kernel, dim3((uint32_t)((f.len + 255) / 256), ntab), dim3(256), 0, s, f);
  P3G_CHECK_HIP(hipGetLastError());
}

void lde_coset_batch_noalloc(Ctx& ctx, const uint64_t* d_trace, uint64_t n, uint64_t w, unsigned log_blowup,
                             uint64_t shift, uint64_t* d_out, hipStream_t s) {
  lde_check_shape(n, log_blowup);
  const unsigned logn = log2_exact(n);
  const uint64_t need = lde_scratch_words(logn, log_blowup);
  if (log_blowup > 3 || need > LDE_SCRATCH_WORDS)
    throw ApiError(E_RANGE, "coset_lde_batch_noalloc: this shape needs " + std::to_string(need) +
                                " table words (added_bits <= 3); the scratch array holds " +
                                std::to_string(LDE_SCRATCH_WORDS));
  if (w == 0) return;
  uint64_t* base = lde_scratch(ctx.device);
  const unsigned twbits = logn ? logn - 1 : 0;
  const uint64_t twlen = POW_LO + pow_table_hi_len(twbits);
  const uint64_t stride = POW_LO + pow_table_hi_len(logn);
  const uint32_t B = 1u << log_blowup;
  uint64_t* tw_inv = base;
  uint64_t* tw_fwd = base + twlen;
  uint64_t* scale = base + 2 * twlen;
  PowFill f{};
  f.lo_mul = 1;
  f.dst = tw_inv;
  f.stride = twlen;  // tw_fwd follows tw_inv
  f.len = twlen;
  f.base[0] = gl_inv(gl_two_adic_generator(logn));
  f.base[1] = gl_two_adic_generator(logn);
  pow_fill(f, 2, s);
  PowFill g{};
  g.lo_mul = gl_inv_pow2(logn);
  g.dst = scale;
  g.stride = stride;
  g.len = stride;
  const uint64_t om = gl_two_adic_generator(logn + log_blowup);
  uint64_t ct = shift;
  for (uint32_t t = 0; t < B; t++) {
    g.base[t] = ct;
    ct = gl_mul(ct, om);
  }
  pow_fill(g, B, s);
  lde_coset_launches(ctx, d_trace, n, w, log_blowup, d_out, s, nullptr, LdeTables{tw_inv, tw_fwd, scale});
}

}  // namespace p3g

namespace p3g {

__global__ void bitrev_rows_kernel(const uint64_t* __restrict__ src, uint64_t* __restrict__ dst, uint64_t h,
                                   unsigned logh, uint64_t w) {
  const uint64_t r = blockIdx.x;
  const uint64_t sr = bitrev32((uint32_t)r, logh);
  for (uint64_t c = threadIdx.x; c < w; c += blockDim.x) dst[r * w + c] = src[sr * w + c];
  (void)h;
}

// Out-of-place row bit reversal (bit_reverse_rows of p3-matrix).
void bit_reverse_rows(Ctx& ctx, const uint64_t* src, uint64_t* dst, uint64_t h, uint64_t w, hipStream_t s) {
  (void)ctx;
  if (h == 0 || w == 0) return;
  hipLaunchKernelGGL(bitrev_rows_kernel, dim3((uint32_t)h), dim3(64), 0, s, src, dst, h, log2_exact(h), w);
  P3G_CHECK_HIP(hipGetLastError());
}

}  // namespace p3g

namespace p3g {

size_t ntt_args_size() { return sizeof(NttArgs); }

void make_ntt_args(void* out, const uint64_t* src, uint64_t* dst, const uint64_t* tw, const uint64_t* scale,
                   uint64_t scale_stride, uint64_t n, uint32_t logn, uint32_t s0, uint32_t K, uint32_t w, uint32_t logC,
                   uint32_t ncol, uint32_t logB, uint32_t mode) {
  NttArgs a{};
  a.src = src;
  a.dst = dst;
  a.tw = tw;
  a.scale = scale;
  a.scale_stride = scale_stride;
  a.n = n;
  a.logn = logn;
  a.s0 = s0;
  a.K = K;
  a.w = w;
  a.C = 1u << logC;
  a.logC = logC;
  a.ncoltiles = ncol;
  a.logB = logB;
  a.mode = mode;
  memcpy(out, &a, sizeof(a));
}

// One pass of one transform with its descriptor passed by value (no host
// interaction: safe inside stream capture).
void launch_ntt_single(uint32_t mode, bool reg, const void* args, uint64_t blocks, hipStream_t s) {
  NttArgs a;
  memcpy(&a, args, sizeof(a));
  if (reg) {
    launch_reg(a.K, mode, &a, nullptr, nullptr, 0, blocks, s);
  } else {
    hipLaunchKernelGGL(ntt_pass_kernel, dim3((uint32_t)blocks), dim3(NTT_T), 0, s, a);
    P3G_CHECK_HIP(hipGetLastError());
  }
}

// Grouped launch of one pass over several transforms (no host interaction:
// safe inside stream capture); register-blocked groups share one K.
void launch_ntt_grouped(uint32_t mode, bool reg, uint32_t K, const void* descs, const uint64_t* starts, uint32_t nd,
                        uint64_t blocks, hipStream_t s) {
  const NttArgs* d = reinterpret_cast<const NttArgs*>(descs);
  if (reg) {
    launch_reg(K, mode, nullptr, d, starts, nd, blocks, s);
  } else {
    hipLaunchKernelGGL(ntt_pass_grouped_kernel, dim3((uint32_t)blocks), dim3(NTT_T), 0, s, d, starts, nd);
    P3G_CHECK_HIP(hipGetLastError());
  }
}

// Coset LDEs of many traces of one shape (n x w) at once: every pass is one
// grouped launch over all transforms, scale tables shared per distinct shift.
void coset_lde_many(Ctx& ctx, const std::vector<LdeJob>& jobs, unsigned log_blowup, hipStream_t s) {
  if (jobs.empty()) return;
  const uint64_t n = jobs[0].n, w = jobs[0].w;
  for (auto& j : jobs)
    if (j.n != n || j.w != w) throw ApiError(E_INVAL, "batched LDE needs one trace shape");
  if (n == 0 || (n & (n - 1))) throw ApiError(E_NOT_POW2, "trace height must be a power of two");
  if (w == 0) return;
  const unsigned logn = log2_exact(n);
  if (logn + log_blowup > 32) throw ApiError(E_RANGE, "LDE height exceeds the field two-adicity");
  uint32_t logC = 0;
  while ((2u << logC) <= NTT_C && (2ull << logC) <= w) logC++;
  const uint32_t ncol = (uint32_t)((w + (1u << logC) - 1) >> logC);
  const uint32_t B = 1u << log_blowup;
  const uint32_t twbits = logn ? logn - 1 : 0;
  const uint64_t* tw_inv = ctx.pow_table(gl_inv(gl_two_adic_generator(logn)), twbits);
  const uint64_t* tw_fwd = ctx.pow_table(gl_two_adic_generator(logn), twbits);
  const uint64_t stride = POW_LO + pow_table_hi_len(logn);
  std::map<uint64_t, uint64_t*> scales;
  std::vector<uint64_t*> owned;
  std::vector<NttArgs> descs;
  std::vector<uint64_t> starts;
  auto cleanup = [&] {
    (void)hipStreamSynchronize(s);
    for (auto* p : owned) ctx.release(p);
  };
  try {
    for (auto& j : jobs) {
      if (scales.count(j.shift)) continue;
      std::vector<uint64_t> host(stride * B);
      const uint64_t ninv = gl_inv_pow2(logn);
      const uint64_t om = gl_two_adic_generator(logn + log_blowup);
      uint64_t ct = j.shift;
      for (uint32_t t = 0; t < B; t++) {
        uint64_t* tab = host.data() + t * stride;
        uint64_t v = ninv;
        for (uint32_t i = 0; i < POW_LO; i++) {
          tab[i] = v;
          v = gl_mul(v, ct);
        }
        uint64_t step = gl_pow(ct, POW_LO), h = 1;
        for (uint64_t i = 0; i < pow_table_hi_len(logn); i++) {
          tab[POW_LO + i] = h;
          h = gl_mul(h, step);
        }
        ct = gl_mul(ct, om);
      }
      uint64_t* sc = (uint64_t*)ctx.upload(host.data(), host.size() * 8, s);
      owned.push_back(sc);
      scales[j.shift] = sc;
    }
    // descriptors of every pass for every job, uploaded once: [pass][job]
    const std::vector<uint32_t> split = ntt_pass_split(logn, NTT_K);
    const size_t np = 2 * split.size(), k = jobs.size();
    descs.assign(np * k, NttArgs{});
    starts.assign(np * (k + 1), 0);
    std::vector<uint32_t> modes(np), Ks(np);
    std::vector<uint64_t> nblks(np);
    uint32_t s0 = 0;
    for (size_t i = 0; i < split.size(); i++) {
      modes[i] = i == 0 ? INV_FIRST : INV;
      modes[split.size() + i] = i == 0 ? FWD_FIRST : FWD;
      Ks[i] = Ks[split.size() + i] = split[i];
      nblks[i] = 1;
      nblks[split.size() + i] = i == 0 ? 1 : B;
      for (size_t q = 0; q < 2; q++) {
        const size_t pi = q * split.size() + i;
        for (size_t jj = 0; jj < k; jj++) {
          const LdeJob& jb = jobs[jj];
          make_ntt_args(&descs[pi * k + jj], jb.src, jb.dst, q == 0 ? tw_inv : tw_fwd, scales[jb.shift], stride, n,
                        logn, s0, split[i], (uint32_t)w, logC, ncol, log_blowup, modes[pi]);
          starts[pi * (k + 1) + jj] = jj * ((n >> split[i]) * ncol * nblks[pi]);
        }
        starts[pi * (k + 1) + k] = k * ((n >> split[i]) * ncol * nblks[pi]);
      }
      s0 += split[i];
    }
    NttArgs* ddesc = (NttArgs*)ctx.upload(descs.data(), descs.size() * sizeof(NttArgs), s);
    owned.push_back((uint64_t*)ddesc);
    uint64_t* dstarts = (uint64_t*)ctx.upload(starts.data(), starts.size() * 8, s);
    owned.push_back(dstarts);
    for (size_t pi = 0; pi < np; pi++) {
      ProfScope ps(ctx, modes[pi] <= INV ? "lde_inverse_pass" : "lde_forward_pass", s);
      launch_ntt_grouped(modes[pi], ntt_reg_ok(Ks[pi], logC), Ks[pi], ddesc + pi * k, dstarts + pi * (k + 1),
                         (uint32_t)k, starts[pi * (k + 1) + k], s);
    }
  } catch (...) {
    cleanup();
    throw;
  }
  // stream-ordered release: nothing here waits for the transforms
  for (auto* p : owned) ctx.release(p);
}

// Stages per pass for a size-2^logn transform: as few passes as kmax allows,
// stage counts balanced so that no pass falls far below the others.
std::vector<uint32_t> ntt_pass_split(uint32_t logn, uint32_t kmax) {
  std::vector<uint32_t> out;
  if (logn == 0) {
    out.push_back(0);
    return out;
  }
  const uint32_t np = (logn + kmax - 1) / kmax;
  for (uint32_t i = 0; i < np; i++) out.push_back(logn / np + (i < logn % np ? 1 : 0));
  return out;
}

}  // namespace p3g
