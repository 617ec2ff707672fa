// DEEP quotient, FRI commit/fold and the query phase on gfx950, sequenced by
// the host exactly as `open_with_channel` (p3-miden-lifted-fri/src/prover.rs:35-116).
//
// DEEP (deep/interpolate.rs:81-201, deep/prover.rs:230-409):
//   deep_weights_kernel   barycentric weights x_i / (z_j - x_i) over the first
//                         L/B coset points x_i = g * omega_L^{bitrev(i)};
//                         1/(z - x) = (z0 - x, -z1) / N(z - x) with the base-field
//                         norm inverted by a per-thread Montgomery batch;
//   fold_weights_kernel   chunk sums for lifted (shorter) matrices;
//   colsum kernels        per-column dot products with the weights (lanes over
//                         columns, coalesced row reads), then a block reduce;
//   deep_quotient_kernel  Q(x_i) = sum_j beta^j (f_red(z_j) + acc_i) / (z_j - x_i)
//                         with acc_i = sum_c -alpha^(W-1-c) f_c(x_i) over the
//                         lifted rows (row i >> log2(L/h) of each matrix).
// FRI (fri/prover.rs:190-356, fold/arity{2,4,8}.rs): each round commits the
//   current EF vector viewed as (m/a) x 2a base-field rows with an unaligned
//   LMCS tree, then fri_fold_kernel interpolates each bit-reversed coset row
//   and evaluates at beta / s (s^-1 = omega_L^{-bitrev(k) * L/m} from a
//   two-level table).  The final polynomial is recovered on the host.
#include <algorithm>
#include <array>
#include <cstring>
#include <map>

#include "internal.hpp"
#include "p2_coop.hpp"

namespace p3g {

constexpr int MAXPTS = 4;

struct Points {
  ef2 z[MAXPTS];
};

__device__ __forceinline__ uint64_t pwt(const uint64_t* __restrict__ tab, uint64_t e) {
  return gl_mul(tab[e & (POW_LO - 1)], tab[POW_LO + (e >> 11)]);
}

constexpr int DW_T = 256;
constexpr int DW_CH = 8;

template <int N>
__device__ __forceinline__ void deep_weights_body(const uint64_t* __restrict__ powL, unsigned logL, const Points& zp,
                                                  uint64_t d, ef2* __restrict__ W, uint64_t i0) {
  if (i0 >= d) return;
  const int cnt = (int)((d - i0) < (uint64_t)DW_CH ? (d - i0) : DW_CH);
  uint64_t xs[DW_CH], d0[DW_CH][N], nrm[DW_CH][N], pre[DW_CH][N];
  uint64_t acc = 1;
#pragma unroll
  for (int r = 0; r < DW_CH; r++) {
    if (r < cnt) {
      uint64_t i = i0 + r;
      uint64_t x = gl_mul(GENERATOR, pwt(powL, bitrev32((uint32_t)i, logL)));
      xs[r] = x;
#pragma unroll
      for (int j = 0; j < N; j++) {
        uint64_t a0 = gl_sub(zp.z[j].c0, x);
        d0[r][j] = a0;
        uint64_t n = gl_sub(gl_sqr(a0), gl_mul(EXT_W, gl_sqr(zp.z[j].c1)));
        nrm[r][j] = n;
        pre[r][j] = acc;
        acc = gl_mul(acc, n);
      }
    }
  }
  uint64_t inv = gl_inv(acc);
#pragma unroll
  for (int r = DW_CH - 1; r >= 0; r--) {
    if (r < cnt) {
#pragma unroll
      for (int j = N - 1; j >= 0; j--) {
        uint64_t ni = gl_mul(inv, pre[r][j]);
        inv = gl_mul(inv, nrm[r][j]);
        // q = (d0 * ni, -z1 * ni); weight = x * q
        uint64_t q0 = gl_mul(d0[r][j], ni);
        uint64_t q1 = gl_neg(gl_mul(zp.z[j].c1, ni));
        W[(i0 + r) * N + j] = ef2{gl_mul(q0, xs[r]), gl_mul(q1, xs[r])};
      }
    }
  }
}

template <int N>
__global__ void __launch_bounds__(DW_T) deep_weights_kernel(const uint64_t* __restrict__ powL, unsigned logL,
                                                            Points zp, uint64_t d, ef2* __restrict__ W) {
  deep_weights_body<N>(powL, logL, zp, d, W, ((uint64_t)blockIdx.x * DW_T + threadIdx.x) * DW_CH);
}

// ---- job-indexed forms for batches of same-shaped proofs (job = blockIdx.y, or .z)
template <int N>
__global__ void __launch_bounds__(DW_T) deep_weights_many_kernel(const uint64_t* __restrict__ powL, unsigned logL,
                                                                 const Points* __restrict__ zps, uint64_t d,
                                                                 ef2* const* __restrict__ Ws) {
  const Points zp = zps[blockIdx.y];
  deep_weights_body<N>(powL, logL, zp, d, Ws[blockIdx.y], ((uint64_t)blockIdx.x * DW_T + threadIdx.x) * DW_CH);
}

template <int N>
__device__ __forceinline__ void fold_weights_body(const ef2* __restrict__ in, ef2* __restrict__ out, uint64_t len,
                                                  uint64_t cs, uint64_t i) {
  if (i >= len) return;
#pragma unroll
  for (int j = 0; j < N; j++) {
    ef2 s = ef_from(0);
    for (uint64_t c = 0; c < cs; c++) s = ef_add(s, in[(i * cs + c) * N + j]);
    out[i * N + j] = s;
  }
}
template <int N>
__global__ void fold_weights_kernel(const ef2* __restrict__ in, ef2* __restrict__ out, uint64_t len, uint64_t cs) {
  fold_weights_body<N>(in, out, len, cs, (uint64_t)blockIdx.x * blockDim.x + threadIdx.x);
}
template <int N>
__global__ void fold_weights_many_kernel(const ef2* const* __restrict__ ins, ef2* const* __restrict__ outs,
                                         uint64_t len, uint64_t cs) {
  fold_weights_body<N>(ins[blockIdx.y], outs[blockIdx.y], len, cs, (uint64_t)blockIdx.x * blockDim.x + threadIdx.x);
}

constexpr int CS_COLS = 64;
constexpr int CS_SLICES = 4;
constexpr int CS_T = CS_COLS * CS_SLICES;

template <int N>
__device__ __forceinline__ void deep_colsum_body(const uint64_t* __restrict__ M, uint64_t w, uint64_t rows,
                                                 const ef2* __restrict__ W, uint64_t rows_per_block,
                                                 ef2* __restrict__ partial, uint32_t bx, uint32_t by) {
  __shared__ ef2 red[CS_SLICES][CS_COLS][N];
  const uint32_t cl = threadIdx.x % CS_COLS, rs = threadIdx.x / CS_COLS;
  const uint64_t col = (uint64_t)bx * CS_COLS + cl;
  const uint64_t r0 = (uint64_t)by * rows_per_block;
  const uint64_t r1 = r0 + rows_per_block < rows ? r0 + rows_per_block : rows;
  // unreduced sums of the two extension coordinates per point, reduced once
  Acc192 a0[N], a1[N];
#pragma unroll
  for (int j = 0; j < N; j++) a0[j] = a1[j] = acc192_zero();
  if (col < w) {
    // 8 rows per step, their loads issued together before the reductions
    constexpr int U = 8;
    uint64_t r = r0 + rs;
    for (; r + (U - 1) * CS_SLICES < r1; r += U * CS_SLICES) {
      uint64_t v[U];
#pragma unroll
      for (int u = 0; u < U; u++) v[u] = M[(r + (uint64_t)u * CS_SLICES) * w + col];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const uint64_t ru = r + (uint64_t)u * CS_SLICES;
#pragma unroll
        for (int j = 0; j < N; j++) {
          const ef2 wt = W[ru * N + j];
          acc192_mad(a0[j], wt.c0, v[u]);
          acc192_mad(a1[j], wt.c1, v[u]);
        }
      }
    }
    for (; r < r1; r += CS_SLICES) {
      const uint64_t v = M[r * w + col];
#pragma unroll
      for (int j = 0; j < N; j++) {
        const ef2 wt = W[r * N + j];
        acc192_mad(a0[j], wt.c0, v);
        acc192_mad(a1[j], wt.c1, v);
      }
    }
  }
#pragma unroll
  for (int j = 0; j < N; j++) red[rs][cl][j] = ef2{acc192_reduce(a0[j]), acc192_reduce(a1[j])};
  __syncthreads();
  if (rs == 0 && col < w) {
#pragma unroll
    for (int j = 0; j < N; j++) {
      ef2 s = red[0][cl][j];
      for (int k = 1; k < CS_SLICES; k++) s = ef_add(s, red[k][cl][j]);
      partial[((uint64_t)by * w + col) * N + j] = s;
    }
  }
}
template <int N>
__global__ void __launch_bounds__(CS_T) deep_colsum_kernel(const uint64_t* __restrict__ M, uint64_t w, uint64_t rows,
                                                           const ef2* __restrict__ W, uint64_t rows_per_block,
                                                           ef2* __restrict__ partial) {
  deep_colsum_body<N>(M, w, rows, W, rows_per_block, partial, blockIdx.x, blockIdx.y);
}
template <int N>
__global__ void __launch_bounds__(CS_T) deep_colsum_many_kernel(const uint64_t* const* __restrict__ Ms, uint64_t w,
                                                                uint64_t rows, const ef2* const* __restrict__ Ws,
                                                                uint64_t rows_per_block, ef2* const* __restrict__ partials) {
  deep_colsum_body<N>(Ms[blockIdx.z], w, rows, Ws[blockIdx.z], rows_per_block, partials[blockIdx.z], blockIdx.x,
                      blockIdx.y);
}

// one workgroup per output t (column x point): strided partial sums, then a
// tree in LDS
constexpr int CSR_T = 256;
__device__ __forceinline__ void colsum_reduce_body(const ef2* __restrict__ partial, uint64_t nblocks, uint64_t wn,
                                                   ef2* __restrict__ out, uint64_t t) {
  __shared__ ef2 red[CSR_T];
  const uint32_t tid = threadIdx.x;
  ef2 s = ef_from(0);
  for (uint64_t b = tid; b < nblocks; b += CSR_T) s = ef_add(s, partial[b * wn + t]);
  red[tid] = s;
  __syncthreads();
#pragma unroll
  for (uint32_t off = CSR_T / 2; off > 0; off >>= 1) {
    if (tid < off) red[tid] = ef_add(red[tid], red[tid + off]);
    __syncthreads();
  }
  if (tid == 0) out[t] = red[0];
}
template <int N>
__global__ void __launch_bounds__(CSR_T) colsum_reduce_kernel(const ef2* __restrict__ partial, uint64_t nblocks,
                                                            uint64_t w, ef2* __restrict__ out) {
  colsum_reduce_body(partial, nblocks, w * N, out, blockIdx.x);
}
template <int N>
__global__ void __launch_bounds__(CSR_T) colsum_reduce_many_kernel(const ef2* const* __restrict__ partials,
                                                                 uint64_t nblocks, uint64_t w,
                                                                 ef2* const* __restrict__ outs) {
  colsum_reduce_body(partials[blockIdx.y], nblocks, w * N, outs[blockIdx.y], blockIdx.x);
}

struct QMat {
  const uint64_t* ptr;
  uint64_t width;
  uint32_t shift;
  uint32_t coeff_off;
};

struct QuotArgs {
  const QMat* mats;
  uint32_t nmats;
  const ef2* coeffs;
  uint64_t L;
  unsigned logL;
  const uint64_t* powL;
  Points z;
  ef2 fred[MAXPTS];
  ef2 bpow[MAXPTS];
  ef2* out;
  uint32_t nsplit;  // column tiles dealt round-robin over nsplit workgroups per row block
  ef2* partial;     // nsplit > 1: per row and split, the partial column sum (L x nsplit)
};

constexpr int DQ_T = 256;
constexpr int DQ_CH = 4;  // rows per thread: tid + 256 r, r < DQ_CH
constexpr int DQ_ROWS = DQ_T * DQ_CH;

// Each workgroup owns DQ_ROWS consecutive LDE rows.  Matrix rows are staged
// through LDS in 256-row x 16-column tiles with coalesced loads (16 lanes per
// 128-byte row segment); each thread then reduces its own row from LDS with
// the (uniform, scalar-loaded) -alpha^(W-1-c) coefficients.
// acc[r] = sum over this split's column tiles of -alpha^(W-1-c) f_c(x_i) for
// rows i = wbase + tid + 256 r (tiles g with g % nsplit == split)
__device__ __forceinline__ void deep_quotient_acc(const QuotArgs& a, uint64_t wbase, uint32_t split, uint32_t nsplit,
                                                  ef2* acc) {
  __shared__ uint64_t tile[256 * 16];
  const uint32_t tid = threadIdx.x;
  // this split's column tiles, in matrix order: tile g = (matrix m, first column c0)
  auto next_tile = [&](uint32_t& m, uint64_t& c0, uint32_t& g) {
    for (;;) {
      if (m >= a.nmats) return false;
      if (c0 >= a.mats[m].width) {
        m++;
        c0 = 0;
        continue;
      }
      if (g % nsplit == split) return true;
      c0 += 16;
      g++;
    }
  };
  // the 16 words this thread stages for tile (m, c0) of row chunk r
  auto load_tile = [&](uint32_t m, uint64_t c0, int r, uint64_t* v) {
    const QMat mm = a.mats[m];
    const uint32_t nc = (uint32_t)((mm.width - c0) < 16 ? (mm.width - c0) : 16);
#pragma unroll
    for (int k = 0; k < 16; k++) {
      const uint32_t e = tid + 256 * k;
      const uint32_t rl = e >> 4, cc = e & 15;
      const uint64_t row = wbase + 256 * r + rl;
      v[k] = (cc < nc && row < a.L) ? mm.ptr[(row >> mm.shift) * mm.width + c0 + cc] : 0;
    }
  };
#pragma unroll
  for (int r = 0; r < DQ_CH; r++) {
    // unreduced sums of the two extension coordinates, reduced once per row
    Acc192 s0 = acc192_zero(), s1 = acc192_zero();
    uint32_t m = 0, g = 0;
    uint64_t c0 = 0;
    bool have = next_tile(m, c0, g);
    uint64_t v[16];
    if (have) load_tile(m, c0, r, v);
    while (have) {
      const uint32_t cm = m;
      const uint64_t cc0 = c0;
      __syncthreads();  // the previous tile's readers are done
#pragma unroll
      for (int k = 0; k < 16; k++) {
        const uint32_t e = tid + 256 * k;
        const uint32_t rl = e >> 4, cc = e & 15;
        tile[rl * 16 + (cc ^ (rl & 15))] = v[k];
      }
      __syncthreads();
      // issue the next tile's loads before reducing this one
      c0 += 16;
      g++;
      have = next_tile(m, c0, g);
      if (have) load_tile(m, c0, r, v);
      const QMat mm = a.mats[cm];
      const ef2* cf = a.coeffs + mm.coeff_off;
      const uint32_t nc = (uint32_t)((mm.width - cc0) < 16 ? (mm.width - cc0) : 16);
#pragma unroll 4
      for (uint32_t cc = 0; cc < nc; cc++) {
        const uint64_t x = tile[tid * 16 + (cc ^ (tid & 15))];
        const ef2 k = cf[cc0 + cc];
        acc192_mad(s0, k.c0, x);
        acc192_mad(s1, k.c1, x);
      }
    }
    acc[r] = ef2{acc192_reduce(s0), acc192_reduce(s1)};
  }
}

template <int N>
__device__ __forceinline__ void deep_quotient_finish(const QuotArgs& a, uint64_t wbase, const ef2* acc) {
  const uint32_t tid = threadIdx.x;
  uint64_t d0[DQ_CH][N], nrm[DQ_CH][N], pre[DQ_CH][N];
  uint64_t run = 1;
#pragma unroll
  for (int r = 0; r < DQ_CH; r++) {
    const uint64_t i = wbase + tid + 256 * r;
    const uint64_t x = gl_mul(GENERATOR, pwt(a.powL, bitrev32((uint32_t)(i < a.L ? i : 0), a.logL)));
#pragma unroll
    for (int j = 0; j < N; j++) {
      uint64_t a0 = gl_sub(a.z.z[j].c0, x);
      d0[r][j] = a0;
      uint64_t n = gl_sub(gl_sqr(a0), gl_mul(EXT_W, gl_sqr(a.z.z[j].c1)));
      nrm[r][j] = n;
      pre[r][j] = run;
      run = gl_mul(run, n);
    }
  }
  uint64_t inv = gl_inv(run);
  ef2 res[DQ_CH];
#pragma unroll
  for (int r = 0; r < DQ_CH; r++) res[r] = ef_from(0);
#pragma unroll
  for (int r = DQ_CH - 1; r >= 0; r--) {
#pragma unroll
    for (int j = N - 1; j >= 0; j--) {
      uint64_t ni = gl_mul(inv, pre[r][j]);
      inv = gl_mul(inv, nrm[r][j]);
      ef2 q = ef2{gl_mul(d0[r][j], ni), gl_neg(gl_mul(a.z.z[j].c1, ni))};
      ef2 t = ef_mul(q, ef_add(a.fred[j], acc[r]));
      if (j > 0) t = ef_mul(a.bpow[j], t);
      res[r] = ef_add(res[r], t);
    }
  }
#pragma unroll
  for (int r = 0; r < DQ_CH; r++) {
    const uint64_t i = wbase + tid + 256 * r;
    if (i < a.L) a.out[i] = res[r];
  }
}
template <int N>
__device__ __forceinline__ void deep_quotient_body(const QuotArgs& a, uint32_t bx) {
  const uint64_t wbase = (uint64_t)bx * DQ_ROWS;
  ef2 acc[DQ_CH];
  deep_quotient_acc(a, wbase, 0, 1, acc);
  deep_quotient_finish<N>(a, wbase, acc);
}
template <int N>
// 3 waves per EU: room for the row tile's staged loads and the batch inversion
// of the finishing step without spilling (4 waves capped it at 128 VGPRs and
// spilled 92 bytes per lane; 2^24 rows x 64 columns: 2.30 -> 2.19 ms)
__global__ void __launch_bounds__(DQ_T) __attribute__((amdgpu_waves_per_eu(3))) deep_quotient_kernel(QuotArgs a) {
  deep_quotient_body<N>(a, blockIdx.x);
}
// column-split form for short, wide LDEs: partial sums, then a finishing pass
// 3 waves per EU, as deep_quotient_kernel (no spill of the staged tile loads)
__global__ void __launch_bounds__(DQ_T) __attribute__((amdgpu_waves_per_eu(3))) deep_quotient_partial_kernel(QuotArgs a) {
  const uint64_t wbase = (uint64_t)blockIdx.x * DQ_ROWS;
  ef2 acc[DQ_CH];
  deep_quotient_acc(a, wbase, blockIdx.y, a.nsplit, acc);
#pragma unroll
  for (int r = 0; r < DQ_CH; r++) {
    const uint64_t i = wbase + threadIdx.x + 256 * r;
    if (i < a.L) a.partial[i * a.nsplit + blockIdx.y] = acc[r];
  }
}
template <int N>
__global__ void __launch_bounds__(DQ_T) __attribute__((amdgpu_waves_per_eu(4))) deep_quotient_finish_kernel(QuotArgs a) {
  const uint64_t wbase = (uint64_t)blockIdx.x * DQ_ROWS;
  ef2 acc[DQ_CH];
#pragma unroll
  for (int r = 0; r < DQ_CH; r++) {
    const uint64_t i = wbase + threadIdx.x + 256 * r;
    ef2 sum = ef_from(0);
    if (i < a.L)
      for (uint32_t y = 0; y < a.nsplit; y++) sum = ef_add(sum, a.partial[i * a.nsplit + y]);
    acc[r] = sum;
  }
  deep_quotient_finish<N>(a, wbase, acc);
}
template <int N>
__global__ void __launch_bounds__(DQ_T) __attribute__((amdgpu_waves_per_eu(3)))
deep_quotient_many_kernel(const QuotArgs* __restrict__ args) {
  const QuotArgs a = args[blockIdx.y];
  deep_quotient_body<N>(a, blockIdx.x);
}

struct FoldTw {
  uint64_t w[8];  // inverse twiddles for the in-row inverse DFT (stage-major)
};

template <int LA>
__device__ __forceinline__ void fri_fold_body(const ef2* __restrict__ in, ef2* __restrict__ out, uint64_t nrows,
                                              unsigned log_rows, const uint64_t* __restrict__ powLinv,
                                              unsigned exp_shift, ef2 beta, const FoldTw& tw, uint64_t ainv,
                                              uint64_t k) {
  constexpr int A = 1 << LA;
  if (k >= nrows) return;
  ef2 v[A];
#pragma unroll
  for (int p = 0; p < A; p++) v[p] = in[k * A + p];
  // bit-reversed input + decimation in time = natural-order unscaled inverse DFT
  int twi = 0;
#pragma unroll
  for (int s = 1; s <= LA; s++) {
    const int m = 1 << s, h = m >> 1;
#pragma unroll
    for (int j = 0; j < h; j++) {
      const uint64_t t = tw.w[twi + j];
#pragma unroll
      for (int b = 0; b < A; b += m) {
        ef2 uu = v[b + j];
        ef2 vv = ef_mul_base(v[b + j + h], t);
        v[b + j] = ef_add(uu, vv);
        v[b + j + h] = ef_sub(uu, vv);
      }
    }
    twi += h;
  }
  const uint64_t e = (uint64_t)bitrev32((uint32_t)k, log_rows) << exp_shift;
  const uint64_t s_inv = pwt(powLinv, e);
  const ef2 x = ef_mul_base(beta, s_inv);
  ef2 r = v[A - 1];
#pragma unroll
  for (int j = A - 2; j >= 0; j--) r = ef_add(ef_mul(r, x), v[j]);
  out[k] = ef_mul_base(r, ainv);
}
template <int LA>
__global__ void fri_fold_kernel(const ef2* __restrict__ in, ef2* __restrict__ out, uint64_t nrows,
                                unsigned log_rows, const uint64_t* __restrict__ powLinv, unsigned exp_shift,
                                ef2 beta, FoldTw tw, uint64_t ainv) {
  fri_fold_body<LA>(in, out, nrows, log_rows, powLinv, exp_shift, beta, tw, ainv,
                    (uint64_t)blockIdx.x * blockDim.x + threadIdx.x);
}
// beta read from device memory (written there by fs_root_beta_kernel)
template <int LA>
__global__ void fri_fold_dev_kernel(const ef2* __restrict__ in, ef2* __restrict__ out, uint64_t nrows,
                                    unsigned log_rows, const uint64_t* __restrict__ powLinv, unsigned exp_shift,
                                    const ef2* __restrict__ beta, FoldTw tw, uint64_t ainv) {
  fri_fold_body<LA>(in, out, nrows, log_rows, powLinv, exp_shift, *beta, tw, ainv,
                    (uint64_t)blockIdx.x * blockDim.x + threadIdx.x);
}

// The Poseidon2 duplex challenger (Transcript with CH_DUPLEX_P2) as device
// words: rate-8 input buffer, 8-word output buffer, 12-word sponge state.
struct DevSponge {
  uint64_t st[12];
  uint64_t in[8];
  uint64_t out[8];
  uint64_t in_len, out_len;
};

// Transcript steps on the device (16 threads, one DPP row), with exactly the
// host's Transcript::observe / sample / duplex steps: observe the nw words at
// `words`, draw `discard` samples (the sample_bits check of a grind), then, if
// beta is set, beta = sample_ef().  Lane l holds sponge word l; the buffers and
// their lengths are uniform.  An FRI commit round is send_commitment(root),
// grind, sample beta (p3-miden-lifted-fri/src/fri/prover.rs:149-156): without
// grinding one call (root, no discard, beta); with grinding one call for the
// root, the witness search, then one call for the witness (one discard, beta).
__device__ __forceinline__ void fs_steps(DevSponge* __restrict__ sp, const uint64_t* __restrict__ words, int nw,
                                         int discard, ef2* __restrict__ beta, const P2Consts* __restrict__ p2) {
  __shared__ uint64_t IN[8], OUT[8];
  const uint32_t l = threadIdx.x;
  uint64_t x = l < 12 ? sp->st[l] : 0;
  if (l < 8) {
    IN[l] = sp->in[l];
    OUT[l] = sp->out[l];
  }
  int inl = (int)sp->in_len, outl = (int)sp->out_len;
  __syncthreads();
  auto duplex = [&]() {
    if (l < 8 && (int)l < inl) x = IN[l];  // overwrite the rate with the pending input
    inl = 0;
    x = p2_permute_coop(x, l, p2);
    if (l < 8) OUT[l] = x;
    outl = 8;
    __syncthreads();
  };
  for (int k = 0; k < nw; k++) {
    const uint64_t v = words[k];
    outl = 0;
    if (l == 0) IN[inl] = v;
    __syncthreads();
    inl++;
    if (inl == 8) duplex();
  }
  auto sample = [&]() {
    if (inl > 0 || outl == 0) duplex();
    return OUT[--outl];
  };
  for (int q = 0; q < discard; q++) (void)sample();
  if (beta) {
    uint64_t b[2];
    for (int q = 0; q < 2; q++) b[q] = sample();
    if (l == 0) *beta = ef2{b[0], b[1]};
  }
  __syncthreads();
  if (l < 12) sp->st[l] = x;
  if (l < 8) {
    sp->in[l] = IN[l];
    sp->out[l] = OUT[l];
  }
  if (l == 0) {
    sp->in_len = (uint64_t)inl;
    sp->out_len = (uint64_t)outl;
  }
}
__global__ void __launch_bounds__(16) fs_root_beta_kernel(DevSponge* __restrict__ sp,
                                                          const uint64_t* __restrict__ root, ef2* __restrict__ beta,
                                                          const P2Consts* __restrict__ p2) {
  fs_steps(sp, root, 4, 0, beta, p2);
}
// one 16-lane workgroup per job: sponge j, root roots[j], beta betas[j]
__global__ void __launch_bounds__(16) fs_root_beta_many_kernel(DevSponge* __restrict__ sps,
                                                               const uint64_t* const* __restrict__ roots,
                                                               ef2* __restrict__ betas,
                                                               const P2Consts* __restrict__ p2) {
  fs_steps(sps + blockIdx.x, roots[blockIdx.x], 4, 0, betas + blockIdx.x, p2);
}
// a grinding round: observe the root, and reset the witness slot the search fills
__global__ void __launch_bounds__(16) fs_root_kernel(DevSponge* __restrict__ sp, const uint64_t* __restrict__ root,
                                                     unsigned long long* __restrict__ wit,
                                                     const P2Consts* __restrict__ p2) {
  fs_steps(sp, root, 4, 0, nullptr, p2);
  if (threadIdx.x == 0) *wit = ~0ull;
}
__global__ void __launch_bounds__(16) fs_root_many_kernel(DevSponge* __restrict__ sps,
                                                          const uint64_t* const* __restrict__ roots,
                                                          unsigned long long* __restrict__ wits,
                                                          const P2Consts* __restrict__ p2) {
  fs_steps(sps + blockIdx.x, roots[blockIdx.x], 4, 0, nullptr, p2);
  if (threadIdx.x == 0) wits[blockIdx.x] = ~0ull;
}
// after the search: observe the witness, draw the sample it was checked against, sample beta
__global__ void __launch_bounds__(16) fs_witness_beta_kernel(DevSponge* __restrict__ sp,
                                                             const unsigned long long* __restrict__ wit,
                                                             ef2* __restrict__ beta, const P2Consts* __restrict__ p2) {
  fs_steps(sp, reinterpret_cast<const uint64_t*>(wit), 1, 1, beta, p2);
}
__global__ void __launch_bounds__(16) fs_witness_beta_many_kernel(DevSponge* __restrict__ sps,
                                                                  const unsigned long long* __restrict__ wits,
                                                                  ef2* __restrict__ betas,
                                                                  const P2Consts* __restrict__ p2) {
  fs_steps(sps + blockIdx.x, reinterpret_cast<const uint64_t*>(wits + blockIdx.x), 1, 1, betas + blockIdx.x, p2);
}

// ------------------------------------------------------------------ proof-of-work grinding
// DuplexChallenger::grind(bits) (called at p3-miden-transcript/src/prover.rs:140-144):
// the smallest witness i with observe(i); sample_bits(bits) == 0.  With k < 8
// input words pending, observing i and sampling is one permutation of the state
// whose rate words 0..k-1 are the pending input and word k is i; the sample is
// word 7 of the result (observing the 8th word duplexes at once, and the sample
// then pops the same word).  Each thread tests one candidate per sweep, and a
// sweep of the grid covers gridDim.x * GRIND_T consecutive candidates.  A
// workgroup stops before a sweep that starts above the smallest witness found
// so far, so every candidate below the final minimum has been tested: the
// result is the witness of the reference's serial search.  *wit keeps its
// value (~0: none) when [start, start + count) holds no witness.
constexpr int GRIND_T = 256;
constexpr uint32_t GRIND_BLOCKS = 1024;

__global__ void __launch_bounds__(GRIND_T) grind_p2_kernel(const DevSponge* __restrict__ sps, uint32_t bits,
                                                           uint64_t start, uint64_t count,
                                                           unsigned long long* __restrict__ wits,
                                                           const P2Consts* __restrict__ p2) {
  // job blockIdx.y: sponge sps[y], witness wits[y]
  const DevSponge* __restrict__ sp = sps + blockIdx.y;
  unsigned long long* __restrict__ wit = wits + blockIdx.y;
  __shared__ uint64_t S[12];
  __shared__ unsigned long long seen;
  __shared__ uint64_t save[12 * GRIND_T];  // the permutation inputs (p2_permute_saved)
  const uint32_t t = threadIdx.x;
  const uint32_t k = (uint32_t)sp->in_len;
  if (t < 12) S[t] = t < k ? sp->in[t] : sp->st[t];
  __syncthreads();
  const uint64_t mask = (1ull << bits) - 1;
  const uint64_t sweep = uint64_t(gridDim.x) * GRIND_T;
  for (uint64_t off = 0; off < count; off += sweep) {
    if (t == 0) seen = __hip_atomic_load(wit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    const bool stop = seen < start + off;
    __syncthreads();
    if (stop) break;
    const uint64_t j = off + uint64_t(blockIdx.x) * GRIND_T + t;
    if (j < count) {
      const uint64_t i = start + j;
      uint64_t st[12];
#pragma unroll
      for (int q = 0; q < 12; q++) st[q] = (uint32_t)q == k ? i : S[q];
      p2_permute_saved<GRIND_T>(st, p2, save + t);
      if ((st[7] & mask) == 0) atomicMin(wit, (unsigned long long)i);
    }
  }
}

__global__ void grind_reset_kernel(unsigned long long* __restrict__ wits, uint32_t k) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < k) wits[j] = ~0ull;
}

// Candidates per search launch: a witness of `bits` bits is missing from
// 2^(bits + 8) consecutive candidates with probability about e^-256.
static uint64_t grind_span(unsigned bits) { return 1ull << (bits + 8); }

// k searches side by side (sponge j, witness j), about GRIND_BLOCKS workgroups in all
static void launch_grind(Ctx& ctx, const DevSponge* dsp, unsigned bits, uint64_t start, unsigned long long* wit,
                         hipStream_t s, uint32_t k = 1) {
  ProfScope ps(ctx, "grind", s);
  const uint32_t per_job = std::max<uint32_t>(16, GRIND_BLOCKS / k);
  hipLaunchKernelGGL(grind_p2_kernel, dim3(per_job, k), dim3(GRIND_T), 0, s, dsp, bits, start, grind_span(bits), wit,
                     ctx.dev_p2);
  P3G_CHECK_HIP(hipGetLastError());
}

static DevSponge device_image(const Transcript& t) {
  DevSponge hs{};
  memcpy(hs.st, t.st, sizeof(hs.st));
  memcpy(hs.in, t.in, sizeof(hs.in));
  memcpy(hs.out, t.out, sizeof(hs.out));
  hs.in_len = (uint64_t)t.in_len;
  hs.out_len = (uint64_t)t.out_len;
  return hs;
}

// Channel::grind(bits) for k transcripts at once, with the witness searches on
// the device when every channel is this library's Poseidon2 duplex transcript;
// the host then observes each witness and checks it, as its own grind would.
// Other channels grind themselves.
static void grind_all(Ctx& ctx, const std::vector<Channel*>& txs, unsigned bits, hipStream_t s) {
  const size_t k = txs.size();
  std::vector<Transcript*> ts(k);
  bool dev = bits > 0;
  for (size_t j = 0; j < k; j++) {
    ts[j] = dynamic_cast<Transcript*>(txs[j]);
    dev = dev && ts[j] && ts[j]->kind == CH_DUPLEX_P2;
  }
  if (!dev) {
    parallel_for(k, [&](size_t j) { txs[j]->grind(bits); });
    return;
  }
  if (bits >= MAX_POW_BITS) throw ApiError(E_INVAL, "grinding: bits must be below 32");
  std::vector<DevSponge> hs(k);
  for (size_t j = 0; j < k; j++) {
    if (ts[j]->in_len >= 8) throw ApiError(E_TRANSCRIPT, "grinding: transcript input buffer is full");
    hs[j] = device_image(*ts[j]);
  }
  DevSponge* dsp = (DevSponge*)ctx.upload(hs.data(), k * sizeof(DevSponge), s);
  unsigned long long* wit = (unsigned long long*)ctx.alloc(k * sizeof(unsigned long long));
  std::vector<unsigned long long> w(k, ~0ull), found(k, ~0ull);
  try {
    // a bounded number of spans: each misses a witness with probability about e^-256
    bool all = false;
    for (uint64_t start = 0, span = 0; !all && span < 16; start += grind_span(bits), span++) {
      hipLaunchKernelGGL(grind_reset_kernel, dim3((uint32_t)((k + 63) / 64)), dim3(64), 0, s, wit, (uint32_t)k);
      P3G_CHECK_HIP(hipGetLastError());
      launch_grind(ctx, dsp, bits, start, wit, s, (uint32_t)k);
      ctx.read_back(w.data(), wit, k * sizeof(unsigned long long), s);
      all = true;
      for (size_t j = 0; j < k; j++) {
        if (found[j] == ~0ull) found[j] = w[j];  // the first span with a witness holds the smallest
        all = all && found[j] != ~0ull;
      }
    }
  } catch (...) {
    ctx.release(dsp);
    ctx.release(wit);
    throw;
  }
  ctx.release(dsp);
  ctx.release(wit);
  for (size_t j = 0; j < k; j++)
    if (found[j] == ~0ull) throw ApiError(E_TRANSCRIPT, "grinding found no witness");
  parallel_for(k, [&](size_t j) { ts[j]->accept_witness(bits, found[j]); });
}

static void grind(Ctx& ctx, Channel& tx, unsigned bits, hipStream_t s) { grind_all(ctx, {&tx}, bits, s); }

struct FoldJob {
  const ef2* in;
  ef2* out;
  ef2 beta;
};
// as fri_fold_many_kernel, with job j's beta read from betas[j]
template <int LA>
__global__ void fri_fold_many_dev_kernel(const FoldJob* __restrict__ jobs, const ef2* __restrict__ betas, uint64_t nrows,
                                         unsigned log_rows, const uint64_t* __restrict__ powLinv, unsigned exp_shift,
                                         FoldTw tw, uint64_t ainv) {
  const FoldJob j = jobs[blockIdx.y];
  fri_fold_body<LA>(j.in, j.out, nrows, log_rows, powLinv, exp_shift, betas[blockIdx.y], tw, ainv,
                    (uint64_t)blockIdx.x * blockDim.x + threadIdx.x);
}
template <int LA>
__global__ void fri_fold_many_kernel(const FoldJob* __restrict__ jobs, uint64_t nrows, unsigned log_rows,
                                     const uint64_t* __restrict__ powLinv, unsigned exp_shift, FoldTw tw,
                                     uint64_t ainv) {
  const FoldJob j = jobs[blockIdx.y];
  fri_fold_body<LA>(j.in, j.out, nrows, log_rows, powLinv, exp_shift, j.beta, tw, ainv,
                    (uint64_t)blockIdx.x * blockDim.x + threadIdx.x);
}

// ------------------------------------------------------------------ host driver

template <int N>
static void run_deep_kernels(Ctx& ctx, const PcsParams& p, unsigned logL, const ef2* z,
                             const std::vector<DevMat>& mats, std::vector<ef2>& evals_host, hipStream_t s) {
  const uint64_t L = 1ull << logL;
  const unsigned logd = logL - p.log_blowup;
  const uint64_t d = 1ull << logd;
  Points zp{};
  for (int j = 0; j < N; j++) zp.z[j] = z[j];
  const uint64_t* powL = ctx.pow_table(gl_two_adic_generator(logL), logL);
  std::map<uint64_t, ef2*> wts;
  ef2* top = (ef2*)ctx.alloc(d * N * sizeof(ef2));
  {
    uint64_t threads = (d + DW_CH - 1) / DW_CH;
    ProfScope ps(ctx, "deep_weights", s);
    hipLaunchKernelGGL(deep_weights_kernel<N>, dim3((uint32_t)((threads + DW_T - 1) / DW_T)), dim3(DW_T), 0, s,
                       powL, logL, zp, d, top);
    P3G_CHECK_HIP(hipGetLastError());
  }
  wts[d] = top;
  std::vector<uint64_t> degs;
  for (auto& m : mats) degs.push_back(m.height >> p.log_blowup);
  std::sort(degs.begin(), degs.end());
  degs.erase(std::unique(degs.begin(), degs.end()), degs.end());
  if (degs.back() != d) throw ApiError(E_INVAL, "no matrix spans the full LDE height");
  uint64_t prev_len = d;
  ef2* prev = top;
  for (auto it = degs.rbegin(); it != degs.rend(); ++it) {
    uint64_t len = *it;
    if (len == d) continue;
    ef2* nw = (ef2*)ctx.alloc(len * N * sizeof(ef2));
    hipLaunchKernelGGL(fold_weights_kernel<N>, dim3((uint32_t)((len + 127) / 128)), dim3(128), 0, s, prev, nw, len,
                       prev_len / len);
    P3G_CHECK_HIP(hipGetLastError());
    wts[len] = nw;
    prev = nw;
    prev_len = len;
  }
  uint64_t totw = 0;
  for (auto& m : mats) totw += m.width;
  ef2* evals_dev = (ef2*)ctx.alloc((totw ? totw : 1) * N * sizeof(ef2));
  uint64_t off = 0;
  for (auto& m : mats) {
    if (m.width == 0) continue;
    uint64_t rows = m.height >> p.log_blowup;
    // at most 1024 row blocks: enough blocks to stream the matrix, few enough
    // partials for the reduction kernel
    uint64_t rpb = rows > 1024ull * 1024 ? (rows + 1023) / 1024 : 1024;
    uint64_t nblk = (rows + rpb - 1) / rpb;
    uint64_t ncb = (m.width + CS_COLS - 1) / CS_COLS;
    ef2* partial = (ef2*)ctx.alloc(nblk * m.width * N * sizeof(ef2));
    {
    ProfScope ps(ctx, "deep_colsum", s);
    hipLaunchKernelGGL(deep_colsum_kernel<N>, dim3((uint32_t)ncb, (uint32_t)nblk), dim3(CS_T), 0, s, m.ptr, m.width,
                       rows, wts[rows], rpb, partial);
    P3G_CHECK_HIP(hipGetLastError());
    }
    uint64_t nt = m.width * N;
    hipLaunchKernelGGL(colsum_reduce_kernel<N>, dim3((uint32_t)nt), dim3(CSR_T), 0, s, partial, nblk, m.width,
                       evals_dev + off * N);
    P3G_CHECK_HIP(hipGetLastError());
    ctx.release(partial);
    off += m.width;
  }
  evals_host.resize(totw * N);
  ctx.read_back(evals_host.data(), evals_dev, totw * N * sizeof(ef2), s);
  ctx.release(evals_dev);
  for (auto& kv : wts) ctx.release(kv.second);
  // barycentric scaling s_j = ((z_j / g)^d - 1) / d
  const uint64_t ginv = gl_inv(GENERATOR);
  const uint64_t dinv = gl_inv_pow2(logd);
  ef2 sc[N];
  for (int j = 0; j < N; j++) {
    ef2 t = ef_mul_base(z[j], ginv);
    for (unsigned k = 0; k < logd; k++) t = ef_sqr(t);
    sc[j] = ef_mul_base(ef_sub(t, ef_from(1)), dinv);
  }
  for (uint64_t c = 0; c < totw; c++)
    for (int j = 0; j < N; j++) evals_host[c * N + j] = ef_mul(evals_host[c * N + j], sc[j]);
  (void)L;
}

template <int N>
static void run_quotient(Ctx& ctx, unsigned logL, const ef2* z, const std::vector<DevMat>& mats,
                         const std::vector<uint64_t>& awidths, const std::vector<ef2>& coeffs, const ef2* fred,
                         ef2 beta, ef2* out, hipStream_t s) {
  const uint64_t L = 1ull << logL;
  std::vector<QMat> qm;
  uint64_t aoff = 0;
  for (size_t i = 0; i < mats.size(); i++) {
    qm.push_back(QMat{mats[i].ptr, mats[i].width, log2_exact(L / mats[i].height), (uint32_t)aoff});
    aoff += awidths[i];
  }
  QMat* dq = (QMat*)ctx.upload(qm.data(), qm.size() * sizeof(QMat), s);
  ef2* dc = (ef2*)ctx.upload(coeffs.data(), coeffs.size() * sizeof(ef2), s);
  QuotArgs a{};
  a.mats = dq;
  a.nmats = (uint32_t)qm.size();
  a.coeffs = dc;
  a.L = L;
  a.logL = logL;
  a.powL = ctx.pow_table(gl_two_adic_generator(logL), logL);
  for (int j = 0; j < N; j++) {
    a.z.z[j] = z[j];
    a.fred[j] = fred[j];
  }
  a.bpow[0] = ef_from(1);
  for (int j = 1; j < N; j++) a.bpow[j] = ef_mul(a.bpow[j - 1], beta);
  a.out = out;
  // short, wide LDEs: deal the column tiles over more workgroups than there are
  // row blocks, so the launch fills the device
  const uint64_t nwg = (L + DQ_ROWS - 1) / DQ_ROWS;
  uint64_t tiles = 0;
  for (auto& q : qm) tiles += (q.width + 15) / 16;
  a.nsplit = 1;
  a.partial = nullptr;
  if (nwg < 1024 && tiles > 1) {
    uint64_t want = 2048 / nwg;
    a.nsplit = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(tiles, want));
  }
  {
    ProfScope ps(ctx, "deep_quotient", s);
    if (a.nsplit == 1) {
      hipLaunchKernelGGL(deep_quotient_kernel<N>, dim3((uint32_t)nwg), dim3(DQ_T), 0, s, a);
    } else {
      a.partial = (ef2*)ctx.alloc(L * a.nsplit * sizeof(ef2));
      hipLaunchKernelGGL(deep_quotient_partial_kernel, dim3((uint32_t)nwg, a.nsplit), dim3(DQ_T), 0, s, a);
      P3G_CHECK_HIP(hipGetLastError());
      hipLaunchKernelGGL(deep_quotient_finish_kernel<N>, dim3((uint32_t)nwg), dim3(DQ_T), 0, s, a);
    }
    P3G_CHECK_HIP(hipGetLastError());
  }
  // stream-ordered releases: the quotient's consumers queue behind it
  ctx.release(a.partial);
  ctx.release(dq);
  ctx.release(dc);
}

// The final polynomial (fri/prover.rs:208-221): n evaluations of the last FRI
// domain in bit-reversed order to the coefficients in descending degree.  A
// decimation-in-time inverse DFT takes the bit-reversed input as it stands and
// yields natural-order coefficients, which are then scaled by 1/n and reversed.
void fri_final_poly(const ef2* br, size_t n, ef2* out) {
  std::vector<ef2> v(br, br + n);
  for (size_t len = 2; len <= n; len <<= 1) {
    const uint64_t w = gl_inv(gl_two_adic_generator(log2_exact(len)));
    const size_t h = len >> 1;
    std::vector<uint64_t> tw(h);
    uint64_t x = 1;
    for (size_t j = 0; j < h; j++) {
      tw[j] = x;
      x = gl_mul(x, w);
    }
    for (size_t i = 0; i < n; i += len)
      for (size_t j = 0; j < h; j++) {
        const ef2 u = v[i + j], t = ef_mul_base(v[i + j + h], tw[j]);
        v[i + j] = ef_add(u, t);
        v[i + j + h] = ef_sub(u, t);
      }
  }
  const uint64_t ninv = n > 1 ? gl_inv_pow2(log2_exact(n)) : 1;
  for (size_t k = 0; k < n; k++) out[n - 1 - k] = ef_mul_base(v[k], ninv);
}

// In-row inverse DFT twiddles of the fold kernels, stage-major: stage s holds
// w_{2^s}^{-j} for j < 2^(s-1).
static FoldTw fold_twiddles(unsigned la) {
  FoldTw tw{};
  int twi = 0;
  for (unsigned st = 1; st <= la; st++) {
    const uint64_t wm = gl_inv(gl_two_adic_generator(st));
    uint64_t x = 1;
    for (uint64_t j = 0; j < (1ull << (st - 1)); j++) {
      tw.w[twi + j] = x;
      x = gl_mul(x, wm);
    }
    twi += 1 << (st - 1);
  }
  return tw;
}

template <int LA>
static void launch_fold(const ef2* in, ef2* out, uint64_t fd, unsigned log_rows, const uint64_t* powinv,
                        unsigned exp_shift, ef2 beta, const FoldTw& tw, uint64_t ainv, hipStream_t s) {
  hipLaunchKernelGGL(fri_fold_kernel<LA>, dim3((uint32_t)((fd + 255) / 256)), dim3(256), 0, s, in, out, fd, log_rows,
                     powinv, exp_shift, beta, tw, ainv);
}

// Aligned OOD evaluations [point][aligned column] from evals[column][point]:
// the zero padding the prover sends for the columns the alignment adds
// (deep/prover.rs:120-160).
static std::vector<ef2> align_evals(const std::vector<DevMat>& mats, const std::vector<uint64_t>& awidths,
                                    const ef2* evals, int npts, uint64_t tota) {
  std::vector<ef2> aligned((size_t)tota * npts, ef_from(0));
  uint64_t off = 0, aoff = 0;
  for (size_t m = 0; m < mats.size(); m++) {
    for (uint64_t c = 0; c < mats[m].width; c++)
      for (int j = 0; j < npts; j++) aligned[j * tota + aoff + c] = evals[(off + c) * npts + j];
    off += mats[m].width;
    aoff += awidths[m];
  }
  return aligned;
}

// DEEP quotient on the LDE domain from the aligned evaluations and the two
// challenges (deep/prover.rs:106-285): f_red(z_j) by Horner in alpha, column
// coefficients -alpha^(W-1-i), then the accumulation kernels.
static void deep_from_aligned(Ctx& ctx, unsigned logL, const ef2* z, int npts, const std::vector<DevMat>& mats,
                              const std::vector<uint64_t>& awidths, const std::vector<ef2>& aligned, uint64_t tota,
                              ef2 alpha, ef2 beta, ef2* out, hipStream_t s) {
  ef2 fred[MAXPTS];
  for (int j = 0; j < npts; j++) {
    ef2 acc = ef_from(0);
    for (uint64_t c = 0; c < tota; c++) acc = ef_add(ef_mul(acc, alpha), aligned[j * tota + c]);
    fred[j] = acc;
  }
  std::vector<ef2> coeffs(tota);
  {
    ef2 pw = ef_neg(ef_from(1));
    for (uint64_t i = tota; i-- > 0;) {
      coeffs[i] = pw;
      pw = ef_mul(pw, alpha);
    }
  }
  switch (npts) {
    case 1: run_quotient<1>(ctx, logL, z, mats, awidths, coeffs, fred, beta, out, s); break;
    case 2: run_quotient<2>(ctx, logL, z, mats, awidths, coeffs, fred, beta, out, s); break;
    case 3: run_quotient<3>(ctx, logL, z, mats, awidths, coeffs, fred, beta, out, s); break;
    default: run_quotient<4>(ctx, logL, z, mats, awidths, coeffs, fred, beta, out, s); break;
  }
}

static void deep_evals(Ctx& ctx, const PcsParams& p, unsigned logL, const ef2* z, int npts,
                       const std::vector<DevMat>& mats, std::vector<ef2>& evals, hipStream_t s) {
  switch (npts) {
    case 1: run_deep_kernels<1>(ctx, p, logL, z, mats, evals, s); break;
    case 2: run_deep_kernels<2>(ctx, p, logL, z, mats, evals, s); break;
    case 3: run_deep_kernels<3>(ctx, p, logL, z, mats, evals, s); break;
    default: run_deep_kernels<4>(ctx, p, logL, z, mats, evals, s); break;
  }
}

// Checks shared by the step-level DEEP calls: the trees' common LDE height,
// their matrices and aligned widths; every matrix at least 2^log_blowup tall
// and no point on the LDE coset (where 1/(z - x) does not exist).
static unsigned deep_inputs(const std::vector<Tree*>& trees, unsigned log_blowup, const ef2* z, int npts,
                            std::vector<DevMat>& mats, std::vector<uint64_t>& awidths, uint64_t& tota) {
  if (trees.empty()) throw ApiError(E_EMPTY, "at least one trace tree required");
  if (npts < 1 || npts > MAXPTS) throw ApiError(E_INVAL, "1 to 4 evaluation points supported");
  const uint64_t L = trees[0]->height;
  const uint32_t alignment = trees[0]->alignment;
  for (auto* t : trees) {
    if (t->height != L) throw ApiError(E_INVAL, "mixed LDE heights are not supported");
    if (t->alignment != alignment) throw ApiError(E_INVAL, "mixed trace tree alignments are not supported");
  }
  const unsigned logL = log2_exact(L);
  if (log_blowup > logL) throw ApiError(E_INVAL, "LDE height below blowup");
  tota = 0;
  for (auto* t : trees)
    for (auto& m : t->mats) {
      if ((m.height >> log_blowup) == 0) throw ApiError(E_INVAL, "matrix shorter than the blowup");
      mats.push_back(m);
      awidths.push_back(aligned_len(m.width, alignment));
      tota += awidths.back();
    }
  for (int j = 0; j < npts; j++) {
    if (z[j].c1 != 0) continue;
    // x in g*H_L  <=>  (x / g)^L == 1
    uint64_t t = gl_mul(z[j].c0, gl_inv(GENERATOR));
    for (unsigned k = 0; k < logL; k++) t = gl_mul(t, t);
    if (t == 1) throw ApiError(E_INVAL, "evaluation point lies on the LDE coset");
  }
  return logL;
}

void deep_check_inputs(const std::vector<Tree*>& trees, unsigned log_blowup, const ef2* z, int npts) {
  std::vector<DevMat> mats;
  std::vector<uint64_t> awidths;
  uint64_t tota = 0;
  (void)deep_inputs(trees, log_blowup, z, npts, mats, awidths, tota);
}

void deep_batch_eval(Ctx& ctx, unsigned log_blowup, const std::vector<Tree*>& trees, const ef2* z, int npts,
                     std::vector<ef2>& evals, hipStream_t s) {
  std::vector<DevMat> mats;
  std::vector<uint64_t> awidths;
  uint64_t tota = 0;
  const unsigned logL = deep_inputs(trees, log_blowup, z, npts, mats, awidths, tota);
  PcsParams p{};
  p.log_blowup = log_blowup;
  deep_evals(ctx, p, logL, z, npts, mats, evals, s);
}

void deep_quotient(Ctx& ctx, unsigned log_blowup, const std::vector<Tree*>& trees, const ef2* z, int npts,
                   const ef2* evals, ef2 alpha, ef2 beta, ef2* out, hipStream_t s) {
  std::vector<DevMat> mats;
  std::vector<uint64_t> awidths;
  uint64_t tota = 0;
  const unsigned logL = deep_inputs(trees, log_blowup, z, npts, mats, awidths, tota);
  const std::vector<ef2> aligned = align_evals(mats, awidths, evals, npts, tota);
  deep_from_aligned(ctx, logL, z, npts, mats, awidths, aligned, tota, alpha, beta, out, s);
}

// One FRI round's fold (FriFold::fold_matrix, fri/fold/mod.rs:119-144) of the
// n evaluations of round domain D (|D| = n, bit-reversed) into n / 2^la: row k
// of arity 2^la lies on the coset s_k <w_a> with s_k^-1 = w_n^-bitrev(k).
void fri_fold(Ctx& ctx, const ef2* in, uint64_t n, unsigned la, ef2 beta, ef2* out, hipStream_t s) {
  if (la < 1 || la > 3) throw ApiError(E_INVAL, "folding arity must be 2, 4 or 8");
  if (n == 0 || (n & (n - 1))) throw ApiError(E_NOT_POW2, "domain size must be a power of two");
  const unsigned logn = log2_exact(n);
  if (logn < la) throw ApiError(E_INVAL, "domain smaller than the folding arity");
  if (logn > 32) throw ApiError(E_RANGE, "domain exceeds the field two-adicity");
  const uint64_t fd = n >> la;
  if (out < in + n && in < out + fd) throw ApiError(E_INVAL, "fold output overlaps its input");
  const unsigned log_rows = logn - la;
  const uint64_t* powinv = ctx.pow_table(gl_inv(gl_two_adic_generator(logn)), log_rows);
  const FoldTw tw = fold_twiddles(la);
  const uint64_t ainv = gl_inv_pow2(la);
  ProfScope ps(ctx, "fri_fold", s);
  switch (la) {
    case 1: launch_fold<1>(in, out, fd, log_rows, powinv, 0, beta, tw, ainv, s); break;
    case 2: launch_fold<2>(in, out, fd, log_rows, powinv, 0, beta, tw, ainv, s); break;
    default: launch_fold<3>(in, out, fd, log_rows, powinv, 0, beta, tw, ainv, s); break;
  }
  P3G_CHECK_HIP(hipGetLastError());
}

void open_with_channel(Ctx& ctx, const PcsParams& p, unsigned logL, const ef2* z, int npts,
                       const std::vector<Tree*>& trees, Channel& tx, hipStream_t s) {
  if (trees.empty()) throw ApiError(E_EMPTY, "at least one trace tree required");
  if (npts < 1 || npts > MAXPTS) throw ApiError(E_INVAL, "1 to 4 evaluation points supported");
  if (p.log_arity < 1 || p.log_arity > 3) throw ApiError(E_INVAL, "invalid folding arity");
  if (p.log_blowup == 0 || p.num_queries == 0) throw ApiError(E_INVAL, "invalid PCS parameters");
  const uint64_t L = 1ull << logL;
  for (auto* t : trees)
    if (t->height != L) throw ApiError(E_INVAL, "mixed LDE heights are not supported");
  const uint32_t alignment = trees[0]->alignment;
  for (auto* t : trees)
    if (t->alignment != alignment) throw ApiError(E_INVAL, "mixed trace tree alignments are not supported");
  if (logL < p.log_blowup) throw ApiError(E_INVAL, "LDE height below blowup");
  const uint32_t hash = trees[0]->hash;

  std::vector<DevMat> mats;
  std::vector<uint64_t> awidths;
  for (auto* t : trees)
    for (auto& m : t->mats) {
      mats.push_back(m);
      awidths.push_back(aligned_len(m.width, alignment));
    }

  // ---- DEEP: evaluations at the OOD points
  ctx.check_cancel();
  std::vector<ef2> evals;  // [column][point]
  deep_evals(ctx, p, logL, z, npts, mats, evals, s);
  uint64_t tota = 0;
  for (auto a : awidths) tota += a;
  const std::vector<ef2> aligned = align_evals(mats, awidths, evals.data(), npts, tota);  // [point][aligned column]
  for (int j = 0; j < npts; j++) tx.send_fields(reinterpret_cast<const uint64_t*>(aligned.data() + j * tota), 2 * tota);
  grind(ctx, tx, p.deep_pow_bits, s);
  const ef2 alpha = tx.sample_ef();
  const ef2 beta = tx.sample_ef();
  ef2* Q = (ef2*)ctx.alloc(L * sizeof(ef2));
  deep_from_aligned(ctx, logL, z, npts, mats, awidths, aligned, tota, alpha, beta, Q, s);
  ctx.check_cancel();

  // ---- FRI commit phase
  const unsigned la = p.log_arity;
  const uint64_t arity = 1ull << la;
  const size_t fdeg = fri_final_degree(p, logL);
  const uint64_t fdomain = (uint64_t)fdeg << p.log_blowup;
  const unsigned tbits = logL > la ? logL - la : 0;
  const uint64_t* powLinv = ctx.pow_table(gl_inv(gl_two_adic_generator(logL)), tbits);
  const FoldTw tw = fold_twiddles(la);
  const uint64_t ainv = gl_inv_pow2(la);
  std::vector<Tree*> ftrees;
  std::vector<ef2*> bufs;
  ef2* cur = Q;
  uint64_t domain = L;
  // Fiat-Shamir of the commit rounds on the device, for this library's own
  // Poseidon2 duplex transcript: no host round trip per round.  With folding
  // grinding each round observes its root, searches the witness on the device
  // and observes it before sampling beta.  The host replays the rounds
  // afterwards from the read-back roots and witnesses (checking each witness)
  // and checks every beta against the device's.
  Transcript* dtx = dynamic_cast<Transcript*>(&tx);
  const bool dev_fs = dtx && dtx->kind == CH_DUPLEX_P2;
  const unsigned fbits = p.folding_pow_bits;
  DevSponge* dsp = nullptr;
  ef2* dbeta = nullptr;
  unsigned long long* dwit = nullptr;
  try {
    if (dev_fs) {
      size_t nr = 0;
      for (uint64_t d = L; d > fdomain; d >>= la) nr++;
      const DevSponge hs = device_image(*dtx);
      dsp = (DevSponge*)ctx.upload(&hs, sizeof(hs), s);
      dbeta = (ef2*)ctx.alloc((nr ? nr : 1) * sizeof(ef2));
      if (fbits) dwit = (unsigned long long*)ctx.alloc((nr ? nr : 1) * sizeof(unsigned long long));
    }
    size_t r = 0;
    while (domain > fdomain) {
      ctx.check_cancel();
      const uint64_t fd = domain >> la;
      DevMat m;
      m.ptr = reinterpret_cast<uint64_t*>(cur);
      m.height = fd;
      m.width = 2 * arity;
      m.owned = false;
      Tree* t = lmcs_build(ctx, hash, {m}, 1, s, !dev_fs);
      ftrees.push_back(t);
      bufs.push_back(cur);
      ef2 b{0, 0};
      if (dev_fs) {
        const uint64_t* droot = t->digests + t->layer_off[t->depth] * 4;
        if (fbits == 0) {
          hipLaunchKernelGGL(fs_root_beta_kernel, dim3(1), dim3(16), 0, s, dsp, droot, dbeta + r, ctx.dev_p2);
        } else {
          hipLaunchKernelGGL(fs_root_kernel, dim3(1), dim3(16), 0, s, dsp, droot, dwit + r, ctx.dev_p2);
          P3G_CHECK_HIP(hipGetLastError());
          launch_grind(ctx, dsp, fbits, 0, dwit + r, s);
          hipLaunchKernelGGL(fs_witness_beta_kernel, dim3(1), dim3(16), 0, s, dsp, dwit + r, dbeta + r, ctx.dev_p2);
        }
        P3G_CHECK_HIP(hipGetLastError());
      } else {
        tx.send_commitment(t->root);
        grind(ctx, tx, fbits, s);
        b = tx.sample_ef();
      }
      ef2* nxt = (ef2*)ctx.alloc(fd * sizeof(ef2));
      const unsigned log_rows = log2_exact(fd);
      const unsigned exp_shift = logL - log2_exact(domain);
      const uint32_t grid = (uint32_t)((fd + 255) / 256);
      ProfScope ps(ctx, "fri_fold", s);
      const ef2* db = dbeta + r;
      switch (la) {
        case 1:
          if (dev_fs)
            hipLaunchKernelGGL(fri_fold_dev_kernel<1>, dim3(grid), dim3(256), 0, s, cur, nxt, fd, log_rows, powLinv,
                               exp_shift, db, tw, ainv);
          else
            hipLaunchKernelGGL(fri_fold_kernel<1>, dim3(grid), dim3(256), 0, s, cur, nxt, fd, log_rows, powLinv,
                               exp_shift, b, tw, ainv);
          break;
        case 2:
          if (dev_fs)
            hipLaunchKernelGGL(fri_fold_dev_kernel<2>, dim3(grid), dim3(256), 0, s, cur, nxt, fd, log_rows, powLinv,
                               exp_shift, db, tw, ainv);
          else
            hipLaunchKernelGGL(fri_fold_kernel<2>, dim3(grid), dim3(256), 0, s, cur, nxt, fd, log_rows, powLinv,
                               exp_shift, b, tw, ainv);
          break;
        default:
          if (dev_fs)
            hipLaunchKernelGGL(fri_fold_dev_kernel<3>, dim3(grid), dim3(256), 0, s, cur, nxt, fd, log_rows, powLinv,
                               exp_shift, db, tw, ainv);
          else
            hipLaunchKernelGGL(fri_fold_kernel<3>, dim3(grid), dim3(256), 0, s, cur, nxt, fd, log_rows, powLinv,
                               exp_shift, b, tw, ainv);
          break;
      }
      P3G_CHECK_HIP(hipGetLastError());
      cur = nxt;
      domain = fd;
      r++;
    }
    // final polynomial: first fdeg evaluations, natural order, inverse DFT, descending
    std::vector<ef2> fin(fdeg);
    if (dev_fs) {
      // one read-back for every round's root and beta and the final evaluations
      const size_t nr = ftrees.size();
      std::vector<std::pair<const uint64_t*, uint64_t>> pieces;
      for (auto* t : ftrees) pieces.emplace_back(t->digests + t->layer_off[t->depth] * 4, 4);
      pieces.emplace_back(reinterpret_cast<const uint64_t*>(dbeta), 2 * nr);
      pieces.emplace_back(reinterpret_cast<const uint64_t*>(cur), 2 * fdeg);
      if (fbits) pieces.emplace_back(reinterpret_cast<const uint64_t*>(dwit), nr);
      const uint64_t* raw = gather_pinned(ctx, pieces, s);
      for (size_t q = 0; q < nr; q++) {
        memcpy(ftrees[q]->root, raw + 4 * q, 32);
        tx.send_commitment(ftrees[q]->root);
        if (fbits) {
          // the device searched every candidate below its witness (the serial search's answer)
          const uint64_t w = raw[6 * nr + 2 * fdeg + q];
          if (w == ~0ull) throw ApiError(E_TRANSCRIPT, "device grinding found no witness");
          dtx->accept_witness(fbits, w);
        } else {
          tx.grind(0);
        }
        const ef2 hb = tx.sample_ef();
        const uint64_t* dbq = raw + 4 * nr + 2 * q;
        if (hb.c0 != dbq[0] || hb.c1 != dbq[1])
          throw ApiError(E_TRANSCRIPT, "device transcript diverged from the host replay");
      }
      memcpy(fin.data(), raw + 6 * nr, fdeg * sizeof(ef2));
    } else {
      ctx.read_back(fin.data(), cur, fdeg * sizeof(ef2), s);
    }
    {
      std::vector<ef2> coeffs(fdeg);
      fri_final_poly(fin.data(), fdeg, coeffs.data());
      tx.send_fields(reinterpret_cast<const uint64_t*>(coeffs.data()), 2 * fdeg);
    }
    ctx.release(cur);
    cur = nullptr;

    // ---- query phase
    ctx.check_cancel();
    grind(ctx, tx, p.query_pow_bits, s);
    std::vector<uint64_t> idx(p.num_queries);
    for (uint32_t q = 0; q < p.num_queries; q++) idx[q] = bitrev32((uint32_t)tx.sample_bits(logL), logL);
    // every opening of every tree in one gather, then the hints tree by tree
    struct Seg {
      const Tree* t;
      size_t first, last, nrows;
    };
    std::vector<std::pair<const uint64_t*, uint64_t>> pieces;
    std::vector<Seg> segs;
    auto collect = [&](const Tree& t, const std::vector<uint64_t>& ix) {
      const size_t first = pieces.size();
      const size_t nr = prove_batch_pieces(t, ix, pieces);
      segs.push_back(Seg{&t, first, pieces.size(), nr});
    };
    for (auto* t : trees) collect(*t, idx);
    std::vector<uint64_t> ri(idx.size());
    for (size_t r = 0; r < ftrees.size(); r++) {
      unsigned sh = la * (unsigned)(r + 1);
      for (size_t q = 0; q < idx.size(); q++) ri[q] = idx[q] >> sh;
      collect(*ftrees[r], ri);
    }
    const uint64_t* raw = gather_pinned(ctx, pieces, s);
    for (auto& sg : segs) raw += prove_batch_emit(*sg.t, pieces, sg.first, sg.last, sg.nrows, raw, tx);
  } catch (...) {
    for (auto* t : ftrees) delete t;
    for (auto* b : bufs) ctx.release(b);
    if (cur) ctx.release(cur);
    ctx.release(dsp);
    ctx.release(dbeta);
    ctx.release(dwit);
    throw;
  }
  for (auto* t : ftrees) delete t;
  for (auto* b : bufs) ctx.release(b);
  ctx.release(dsp);
  ctx.release(dbeta);
  ctx.release(dwit);
}


// ------------------------------------------------------------------ batched proofs
// k independent proofs in lock-step: the p3g_prove flow (commit_traces, observe
// the root, sample two extension points, open_with_channel) for jobs whose
// trace lists have one shape.  Every device phase is a single launch over all
// jobs (job-indexed kernels), followed by one synchronisation; only the
// transcripts run per job on the host, each in exactly the order of the
// single-proof flow, so every proof is the one p3g_prove would produce.
void prove_many(Ctx& ctx, uint32_t hash, const PcsParams& p, const std::vector<std::vector<DevMat>>& traces,
                const std::vector<Transcript*>& txs, std::vector<std::array<uint64_t, 4>>& roots,
                std::vector<std::array<ef2, 2>>& zs, hipStream_t s) {
  constexpr int N = 2;
  const size_t k = traces.size();
  if (k == 0 || txs.size() != k) throw ApiError(E_INVAL, "one transcript per job required");
  const std::vector<DevMat>& t0 = traces[0];
  if (t0.empty()) throw ApiError(E_EMPTY, "at least one trace required");
  for (size_t i = 0; i < t0.size(); i++) {
    if (t0[i].height == 0 || (t0[i].height & (t0[i].height - 1)))
      throw ApiError(E_NOT_POW2, "trace height must be a power of two");
    if (i && t0[i].height < t0[i - 1].height) throw ApiError(E_UNSORTED, "traces must be sorted by height");
  }
  for (auto& tj : traces) {
    if (tj.size() != t0.size()) throw ApiError(E_INVAL, "batched proofs need one trace shape");
    for (size_t i = 0; i < tj.size(); i++)
      if (tj[i].height != t0[i].height || tj[i].width != t0[i].width)
        throw ApiError(E_INVAL, "batched proofs need one trace shape");
  }
  if (p.log_arity < 1 || p.log_arity > 3) throw ApiError(E_INVAL, "invalid folding arity");
  if (p.log_blowup == 0 || p.num_queries == 0) throw ApiError(E_INVAL, "invalid PCS parameters");
  const unsigned lmax = log2_exact(t0.back().height);
  const unsigned logL = lmax + p.log_blowup;
  if (logL > 32) throw ApiError(E_RANGE, "LDE height exceeds the field two-adicity");
  const uint64_t L = 1ull << logL;
  const uint32_t B = 1u << p.log_blowup;

  std::vector<Tree*> trees;
  std::vector<std::vector<Tree*>> ftrees;  // [round][job]
  std::vector<ef2*> bufs;                  // every FRI input buffer (Q and folds)
  std::vector<void*> tmp;                  // released at the end
  auto cleanup = [&] {
    (void)hipStreamSynchronize(s);
    for (auto& r : ftrees)
      for (auto* t : r) delete t;
    for (auto* t : trees) delete t;
    for (auto* b : bufs) ctx.release(b);
    for (auto* q : tmp) ctx.release(q);
  };
  try {
    // ---- commit: batched LDEs, then the k trees (which take the LDEs)
    std::vector<std::vector<DevMat>> ldes(k);
    try {
      for (size_t j = 0; j < k; j++)
        for (auto& tr : traces[j]) {
          DevMat m;
          m.height = tr.height << p.log_blowup;
          m.width = tr.width;
          m.owned = true;
          m.ptr = (uint64_t*)ctx.alloc(m.height * m.width * 8);
          ldes[j].push_back(m);
        }
      for (size_t i = 0; i < t0.size(); i++) {
        uint64_t shift = GENERATOR;
        for (unsigned q = log2_exact(t0[i].height); q < lmax; q++) shift = gl_mul(shift, shift);
        std::vector<LdeJob> jobs;
        for (size_t j = 0; j < k; j++)
          jobs.push_back(LdeJob{traces[j][i].ptr, ldes[j][i].ptr, t0[i].height, t0[i].width, shift});
        coset_lde_many(ctx, jobs, p.log_blowup, s);
      }
    } catch (...) {
      (void)hipStreamSynchronize(s);
      for (auto& lj : ldes)
        for (auto& m : lj) ctx.release(m.ptr);
      throw;
    }
    trees = lmcs_build_many(ctx, hash, ldes, hash_alignment(hash), s);
    roots.resize(k);
    zs.resize(k);
    std::vector<Points> zp(k);
    for (size_t j = 0; j < k; j++) {
      memcpy(roots[j].data(), trees[j]->root, 32);
      txs[j]->observe_commitment(trees[j]->root);
      zs[j][0] = txs[j]->sample_ef();
      zs[j][1] = txs[j]->sample_ef();
      zp[j] = Points{};
      zp[j].z[0] = zs[j][0];
      zp[j].z[1] = zs[j][1];
    }

    // ---- DEEP evaluations at the two points, all jobs per launch
    const std::vector<DevMat>& m0 = trees[0]->mats;
    const size_t nm = m0.size();
    const uint32_t alignment = trees[0]->alignment;
    std::vector<uint64_t> awidths;
    uint64_t totw = 0, tota = 0;
    for (auto& m : m0) {
      awidths.push_back(aligned_len(m.width, alignment));
      totw += m.width;
      tota += awidths.back();
    }
    const unsigned logd = logL - p.log_blowup;
    const uint64_t d = 1ull << logd;
    const uint64_t* powL = ctx.pow_table(gl_two_adic_generator(logL), logL);
    // argument arrays go up through the context's pinned upload ring: the host
    // images may die at once and nothing waits for the stream
    auto upload = [&](const void* src, size_t bytes) {
      void* dst = ctx.upload(src, bytes, s);
      tmp.push_back(dst);
      return dst;
    };
    std::map<uint64_t, std::vector<ef2*>> wts;  // weight vectors per trace height, per job
    {
      std::vector<ef2*> top(k);
      for (size_t j = 0; j < k; j++) {
        top[j] = (ef2*)ctx.alloc(d * N * sizeof(ef2));
        tmp.push_back(top[j]);
      }
      const Points* dz = (const Points*)upload(zp.data(), k * sizeof(Points));
      ef2* const* dW = (ef2* const*)upload(top.data(), k * sizeof(ef2*));
      const uint64_t threads = (d + DW_CH - 1) / DW_CH;
      ProfScope ps(ctx, "deep_weights", s);
      hipLaunchKernelGGL(deep_weights_many_kernel<N>, dim3((uint32_t)((threads + DW_T - 1) / DW_T), (uint32_t)k),
                         dim3(DW_T), 0, s, powL, logL, dz, d, dW);
      P3G_CHECK_HIP(hipGetLastError());
      wts[d] = top;
    }
    {
      std::vector<uint64_t> degs;
      for (auto& m : m0) degs.push_back(m.height >> p.log_blowup);
      std::sort(degs.begin(), degs.end());
      degs.erase(std::unique(degs.begin(), degs.end()), degs.end());
      uint64_t prev_len = d;
      for (auto it = degs.rbegin(); it != degs.rend(); ++it) {
        const uint64_t len = *it;
        if (len == d) continue;
        std::vector<ef2*> nw(k);
        for (size_t j = 0; j < k; j++) {
          nw[j] = (ef2*)ctx.alloc(len * N * sizeof(ef2));
          tmp.push_back(nw[j]);
        }
        const ef2* const* din = (const ef2* const*)upload(wts[prev_len].data(), k * sizeof(ef2*));
        ef2* const* dout = (ef2* const*)upload(nw.data(), k * sizeof(ef2*));
        hipLaunchKernelGGL(fold_weights_many_kernel<N>, dim3((uint32_t)((len + 127) / 128), (uint32_t)k), dim3(128), 0,
                           s, din, dout, len, prev_len / len);
        P3G_CHECK_HIP(hipGetLastError());
        wts[len] = nw;
        prev_len = len;
      }
    }
    ef2* evals_dev = (ef2*)ctx.alloc(k * (totw ? totw : 1) * N * sizeof(ef2));
    tmp.push_back(evals_dev);
    {
      uint64_t off = 0;
      for (size_t mi = 0; mi < nm; mi++) {
        const uint64_t w = m0[mi].width, rows = m0[mi].height >> p.log_blowup;
        const uint64_t rpb = rows > 1024ull * 1024 ? (rows + 1023) / 1024 : 1024;
        const uint64_t nblk = (rows + rpb - 1) / rpb;
        const uint64_t ncb = (w + CS_COLS - 1) / CS_COLS;
        std::vector<const uint64_t*> Ms(k);
        std::vector<ef2*> parts(k), outs(k);
        for (size_t j = 0; j < k; j++) {
          Ms[j] = trees[j]->mats[mi].ptr;
          parts[j] = (ef2*)ctx.alloc(nblk * w * N * sizeof(ef2));
          tmp.push_back(parts[j]);
          outs[j] = evals_dev + (j * totw + off) * N;
        }
        const uint64_t* const* dM = (const uint64_t* const*)upload(Ms.data(), k * sizeof(uint64_t*));
        const ef2* const* dW = (const ef2* const*)upload(wts[rows].data(), k * sizeof(ef2*));
        ef2* const* dP = (ef2* const*)upload(parts.data(), k * sizeof(ef2*));
        ef2* const* dO = (ef2* const*)upload(outs.data(), k * sizeof(ef2*));
        {
          ProfScope ps(ctx, "deep_colsum", s);
          hipLaunchKernelGGL(deep_colsum_many_kernel<N>, dim3((uint32_t)ncb, (uint32_t)nblk, (uint32_t)k), dim3(CS_T),
                             0, s, dM, w, rows, dW, rpb, dP);
          P3G_CHECK_HIP(hipGetLastError());
        }
        const uint64_t nt = w * N;
        hipLaunchKernelGGL(colsum_reduce_many_kernel<N>, dim3((uint32_t)nt, (uint32_t)k), dim3(CSR_T), 0, s,
                           (const ef2* const*)dP, nblk, w, dO);
        P3G_CHECK_HIP(hipGetLastError());
        off += w;
      }
    }
    std::vector<ef2> evals_host(k * totw * N);
    ctx.read_back(evals_host.data(), evals_dev, evals_host.size() * sizeof(ef2), s);

    // ---- per job: scale, send, grind, alpha and beta; quotient arguments
    const uint64_t ginv = gl_inv(GENERATOR), dinv = gl_inv_pow2(logd);
    std::vector<QMat> qm_all(k * nm);
    std::vector<ef2> coeffs_all(k * (tota ? tota : 1));
    std::vector<QuotArgs> qa(k);
    std::vector<ef2*> Q(k);
    for (size_t j = 0; j < k; j++) {
      Q[j] = (ef2*)ctx.alloc(L * sizeof(ef2));
      bufs.push_back(Q[j]);
    }
    QMat* dqm = (QMat*)ctx.alloc(qm_all.size() * sizeof(QMat));
    tmp.push_back(dqm);
    ef2* dcoef = (ef2*)ctx.alloc(coeffs_all.size() * sizeof(ef2));
    tmp.push_back(dcoef);
    const uint64_t* powQ = ctx.pow_table(gl_two_adic_generator(logL), logL);
    // per job: scale the evaluations, align and send them; then the DEEP grinds of
    // all jobs (one device search); then per job the challenges and quotient arguments
    std::vector<std::vector<ef2>> aligned_all(k);
    parallel_for(k, [&](size_t j) {
      ef2* ev = evals_host.data() + j * totw * N;
      ef2 sc[N];
      for (int q = 0; q < N; q++) {
        ef2 t = ef_mul_base(zs[j][q], ginv);
        for (unsigned b = 0; b < logd; b++) t = ef_sqr(t);
        sc[q] = ef_mul_base(ef_sub(t, ef_from(1)), dinv);
      }
      for (uint64_t c = 0; c < totw; c++)
        for (int q = 0; q < N; q++) ev[c * N + q] = ef_mul(ev[c * N + q], sc[q]);
      std::vector<ef2> aligned((size_t)tota * N, ef_from(0));
      {
        uint64_t off = 0, aoff = 0;
        for (size_t m = 0; m < nm; m++) {
          for (uint64_t c = 0; c < m0[m].width; c++)
            for (int q = 0; q < N; q++) aligned[q * tota + aoff + c] = ev[(off + c) * N + q];
          off += m0[m].width;
          aoff += awidths[m];
        }
      }
      Transcript& tx = *txs[j];
      for (int q = 0; q < N; q++) tx.send_fields(reinterpret_cast<const uint64_t*>(aligned.data() + q * tota), 2 * tota);
      aligned_all[j] = std::move(aligned);
    });
    grind_all(ctx, std::vector<Channel*>(txs.begin(), txs.end()), p.deep_pow_bits, s);
    parallel_for(k, [&](size_t j) {
      Transcript& tx = *txs[j];
      const std::vector<ef2>& aligned = aligned_all[j];
      const ef2 alpha = tx.sample_ef();
      const ef2 beta = tx.sample_ef();
      QuotArgs a{};
      for (int q = 0; q < N; q++) {
        ef2 acc = ef_from(0);
        for (uint64_t c = 0; c < tota; c++) acc = ef_add(ef_mul(acc, alpha), aligned[q * tota + c]);
        a.fred[q] = acc;
        a.z.z[q] = zs[j][q];
      }
      {
        ef2 pw = ef_neg(ef_from(1));
        ef2* cf = coeffs_all.data() + j * tota;
        for (uint64_t i = tota; i-- > 0;) {
          cf[i] = pw;
          pw = ef_mul(pw, alpha);
        }
      }
      uint64_t aoff = 0;
      for (size_t m = 0; m < nm; m++) {
        qm_all[j * nm + m] = QMat{trees[j]->mats[m].ptr, m0[m].width, log2_exact(L / m0[m].height), (uint32_t)aoff};
        aoff += awidths[m];
      }
      a.mats = dqm + j * nm;
      a.nmats = (uint32_t)nm;
      a.coeffs = dcoef + j * tota;
      a.L = L;
      a.logL = logL;
      a.powL = powQ;
      a.bpow[0] = ef_from(1);
      a.bpow[1] = beta;
      a.out = Q[j];
      a.nsplit = 1;
      a.partial = nullptr;
      qa[j] = a;
    });
    ctx.upload_async(dqm, qm_all.data(), qm_all.size() * sizeof(QMat), s);
    ctx.upload_async(dcoef, coeffs_all.data(), coeffs_all.size() * sizeof(ef2), s);
    const QuotArgs* dqa = (const QuotArgs*)upload(qa.data(), k * sizeof(QuotArgs));
    {
      ProfScope ps(ctx, "deep_quotient", s);
      hipLaunchKernelGGL(deep_quotient_many_kernel<N>, dim3((uint32_t)((L + DQ_ROWS - 1) / DQ_ROWS), (uint32_t)k),
                         dim3(DQ_T), 0, s, dqa);
      P3G_CHECK_HIP(hipGetLastError());
    }

    // ---- FRI commit phase, all jobs per round
    const unsigned la = p.log_arity;
    const uint64_t arity = 1ull << la;
    const size_t fdeg = fri_final_degree(p, logL);
    const uint64_t fdomain = (uint64_t)fdeg << p.log_blowup;
    const unsigned tbits = logL > la ? logL - la : 0;
    const uint64_t* powLinv = ctx.pow_table(gl_inv(gl_two_adic_generator(logL)), tbits);
    const FoldTw tw = fold_twiddles(la);
    const uint64_t ainv = gl_inv_pow2(la);
    std::vector<ef2*> cur = Q;
    uint64_t domain = L;
    // Fiat-Shamir of the commit rounds on the device (one sponge and one
    // 16-lane row per job, and with folding grinding one witness search per
    // job and round), as in the single-proof path; the hosts replay every
    // job's rounds afterwards, checking every witness and beta.
    bool dev_fs = true;
    for (size_t j = 0; j < k; j++) dev_fs = dev_fs && txs[j]->kind == CH_DUPLEX_P2;
    const unsigned fbits = p.folding_pow_bits;
    unsigned long long* dwit = nullptr;
    size_t nr = 0;
    for (uint64_t d = L; d > fdomain; d >>= la) nr++;
    DevSponge* dsp = nullptr;
    ef2* dbeta = nullptr;
    if (dev_fs) {
      std::vector<DevSponge> hs(k);
      for (size_t j = 0; j < k; j++) {
        const Transcript& tx = *txs[j];
        memcpy(hs[j].st, tx.st, sizeof(hs[j].st));
        memcpy(hs[j].in, tx.in, sizeof(hs[j].in));
        memcpy(hs[j].out, tx.out, sizeof(hs[j].out));
        hs[j].in_len = (uint64_t)tx.in_len;
        hs[j].out_len = (uint64_t)tx.out_len;
      }
      dsp = (DevSponge*)upload(hs.data(), k * sizeof(DevSponge));
      dbeta = (ef2*)ctx.alloc((nr ? nr : 1) * k * sizeof(ef2));
      tmp.push_back(dbeta);
      if (fbits) {
        dwit = (unsigned long long*)ctx.alloc((nr ? nr : 1) * k * sizeof(unsigned long long));
        tmp.push_back(dwit);
      }
    }
    size_t r = 0;
    while (domain > fdomain) {
      const uint64_t fd = domain >> la;
      std::vector<std::vector<DevMat>> fm(k);
      for (size_t j = 0; j < k; j++) {
        DevMat m;
        m.ptr = reinterpret_cast<uint64_t*>(cur[j]);
        m.height = fd;
        m.width = 2 * arity;
        m.owned = false;
        fm[j].push_back(m);
      }
      ftrees.push_back(lmcs_build_many(ctx, hash, fm, 1, s, !dev_fs));
      std::vector<FoldJob> fj(k);
      std::vector<ef2*> nxt(k);
      for (size_t j = 0; j < k; j++) {
        nxt[j] = (ef2*)ctx.alloc(fd * sizeof(ef2));
        bufs.push_back(nxt[j]);
      }
      const std::vector<Tree*>& ft = ftrees.back();
      ef2* db = dbeta + r * k;
      if (dev_fs) {
        std::vector<const uint64_t*> rp(k);
        for (size_t j = 0; j < k; j++) {
          rp[j] = ft[j]->digests + ft[j]->layer_off[ft[j]->depth] * 4;
          fj[j] = FoldJob{cur[j], nxt[j], ef2{0, 0}};
        }
        const uint64_t* const* drp = (const uint64_t* const*)upload(rp.data(), k * sizeof(uint64_t*));
        if (fbits == 0) {
          hipLaunchKernelGGL(fs_root_beta_many_kernel, dim3((uint32_t)k), dim3(16), 0, s, dsp, drp, db, ctx.dev_p2);
        } else {
          unsigned long long* wr = dwit + r * k;
          hipLaunchKernelGGL(fs_root_many_kernel, dim3((uint32_t)k), dim3(16), 0, s, dsp, drp, wr, ctx.dev_p2);
          P3G_CHECK_HIP(hipGetLastError());
          launch_grind(ctx, dsp, fbits, 0, wr, s, (uint32_t)k);
          hipLaunchKernelGGL(fs_witness_beta_many_kernel, dim3((uint32_t)k), dim3(16), 0, s, dsp, wr, db,
                             ctx.dev_p2);
        }
        P3G_CHECK_HIP(hipGetLastError());
      } else {
        parallel_for(k, [&](size_t j) { txs[j]->send_commitment(ft[j]->root); });
        grind_all(ctx, std::vector<Channel*>(txs.begin(), txs.end()), fbits, s);
        parallel_for(k, [&](size_t j) { fj[j] = FoldJob{cur[j], nxt[j], txs[j]->sample_ef()}; });
      }
      const FoldJob* dfj = (const FoldJob*)upload(fj.data(), k * sizeof(FoldJob));
      const unsigned log_rows = log2_exact(fd);
      const unsigned exp_shift = logL - log2_exact(domain);
      const dim3 grid((uint32_t)((fd + 255) / 256), (uint32_t)k);
      {
        ProfScope ps(ctx, "fri_fold", s);
        switch (la) {
          case 1:
            if (dev_fs)
              hipLaunchKernelGGL(fri_fold_many_dev_kernel<1>, grid, dim3(256), 0, s, dfj, db, fd, log_rows, powLinv,
                                 exp_shift, tw, ainv);
            else
              hipLaunchKernelGGL(fri_fold_many_kernel<1>, grid, dim3(256), 0, s, dfj, fd, log_rows, powLinv, exp_shift, tw, ainv);
            break;
          case 2:
            if (dev_fs)
              hipLaunchKernelGGL(fri_fold_many_dev_kernel<2>, grid, dim3(256), 0, s, dfj, db, fd, log_rows, powLinv,
                                 exp_shift, tw, ainv);
            else
              hipLaunchKernelGGL(fri_fold_many_kernel<2>, grid, dim3(256), 0, s, dfj, fd, log_rows, powLinv, exp_shift, tw, ainv);
            break;
          default:
            if (dev_fs)
              hipLaunchKernelGGL(fri_fold_many_dev_kernel<3>, grid, dim3(256), 0, s, dfj, db, fd, log_rows, powLinv,
                                 exp_shift, tw, ainv);
            else
              hipLaunchKernelGGL(fri_fold_many_kernel<3>, grid, dim3(256), 0, s, dfj, fd, log_rows, powLinv, exp_shift, tw, ainv);
            break;
        }
        P3G_CHECK_HIP(hipGetLastError());
      }
      cur = nxt;
      domain = fd;
      r++;
    }
    // ---- final polynomials (with the device path: also every round's roots and betas)
    {
      std::vector<std::pair<const uint64_t*, uint64_t>> pieces;
      for (size_t j = 0; j < k; j++) pieces.emplace_back(reinterpret_cast<const uint64_t*>(cur[j]), 2 * fdeg);
      const size_t rounds = ftrees.size();  // ftrees[q][j]: round q, job j
      if (dev_fs) {
        for (size_t q = 0; q < rounds; q++)
          for (size_t j = 0; j < k; j++) {
            const Tree* t = ftrees[q][j];
            pieces.emplace_back(t->digests + t->layer_off[t->depth] * 4, 4);
          }
        pieces.emplace_back(reinterpret_cast<const uint64_t*>(dbeta), 2 * rounds * k);
        if (fbits) pieces.emplace_back(reinterpret_cast<const uint64_t*>(dwit), rounds * k);
      }
      std::vector<uint64_t> raw;
      gather_to_host(ctx, pieces, raw, s);
      if (dev_fs) {
        const uint64_t* rr = raw.data() + 2 * fdeg * k;  // roots [round][job], then betas [round][job]
        const uint64_t* bb = rr + 4 * rounds * k;
        const uint64_t* ww = bb + 2 * rounds * k;  // witnesses [round][job] (folding grinding)
        std::vector<int> ok(k, 1);
        parallel_for(k, [&](size_t j) {
          Transcript& tx = *txs[j];
          for (size_t q = 0; q < rounds; q++) {
            Tree* t = ftrees[q][j];
            memcpy(t->root, rr + 4 * (q * k + j), 32);
            tx.send_commitment(t->root);
            if (fbits) {
              const uint64_t w = ww[q * k + j];
              if (w == ~0ull || w >= P || !tx.check_witness(fbits, w)) {
                ok[j] = 0;
                return;
              }
              tx.fields.push_back(w);
            } else {
              tx.grind(0);
            }
            const ef2 hb = tx.sample_ef();
            const uint64_t* dbq = bb + 2 * (q * k + j);
            if (hb.c0 != dbq[0] || hb.c1 != dbq[1]) ok[j] = 0;
          }
        });
        for (size_t j = 0; j < k; j++)
          if (!ok[j]) throw ApiError(E_TRANSCRIPT, "device transcript diverged from the host replay");
      }
      parallel_for(k, [&](size_t j) {
        const ef2* fin = reinterpret_cast<const ef2*>(raw.data() + j * 2 * fdeg);
        std::vector<ef2> coeffs(fdeg);
        fri_final_poly(fin, fdeg, coeffs.data());
        txs[j]->send_fields(reinterpret_cast<const uint64_t*>(coeffs.data()), 2 * fdeg);
      });
    }
    // ---- query phase: indices per job, then one gather for every opening of every job
    struct Seg {
      const Tree* t;
      size_t first, last, nrows;
    };
    // per job (in parallel): indices and the pieces of every opening, then one
    // gather for all jobs, then the hints per job (in parallel)
    std::vector<std::vector<std::pair<const uint64_t*, uint64_t>>> jp(k);
    std::vector<std::vector<Seg>> segs(k);
    grind_all(ctx, std::vector<Channel*>(txs.begin(), txs.end()), p.query_pow_bits, s);
    parallel_for(k, [&](size_t j) {
      Transcript& tx = *txs[j];
      std::vector<uint64_t> idx(p.num_queries);
      for (uint32_t q = 0; q < p.num_queries; q++) idx[q] = bitrev32((uint32_t)tx.sample_bits(logL), logL);
      size_t first = 0;
      size_t nr = prove_batch_pieces(*trees[j], idx, jp[j]);
      segs[j].push_back(Seg{trees[j], first, jp[j].size(), nr});
      std::vector<uint64_t> ri(idx.size());
      for (size_t r = 0; r < ftrees.size(); r++) {
        const unsigned sh = la * (unsigned)(r + 1);
        for (size_t q = 0; q < idx.size(); q++) ri[q] = idx[q] >> sh;
        first = jp[j].size();
        nr = prove_batch_pieces(*ftrees[r][j], ri, jp[j]);
        segs[j].push_back(Seg{ftrees[r][j], first, jp[j].size(), nr});
      }
    });
    // offsets per job, then every job's pieces copied into place in parallel
    std::vector<size_t> job_piece0(k + 1, 0), job_word0(k + 1, 0), job_words(k, 0);
    parallel_for(k, [&](size_t j) {
      size_t wj = 0;
      for (auto& pc : jp[j]) wj += pc.second;
      job_words[j] = wj;
    });
    for (size_t j = 0; j < k; j++) {
      job_piece0[j + 1] = job_piece0[j] + jp[j].size();
      job_word0[j + 1] = job_word0[j] + job_words[j];
    }
    std::vector<std::pair<const uint64_t*, uint64_t>> pieces(job_piece0[k]);
    parallel_for(k, [&](size_t j) { std::copy(jp[j].begin(), jp[j].end(), pieces.begin() + job_piece0[j]); });
    const uint64_t* raw = gather_pinned(ctx, pieces, s);
    parallel_for(k, [&](size_t j) {
      const uint64_t* base = raw + job_word0[j];
      for (auto& sg : segs[j]) base += prove_batch_emit(*sg.t, jp[j], sg.first, sg.last, sg.nrows, base, *txs[j]);
    });
    (void)B;
  } catch (...) {
    cleanup();
    throw;
  }
  cleanup();
}

}  // namespace p3g
