// Quotient accumulation helpers on a device-resident accumulator, for gfx950.
//
// Replaces `cyclic_extend_and_scale` and `divide_by_vanishing_in_place`
// (p3-miden-lifted-stark/src/prover/quotient.rs:28-101), the steps between
// constraint evaluation and `commit_quotient`.  Both are elementwise over
// extension elements (16 bytes each) and bound by HBM.
#include "internal.hpp"

namespace p3g {

constexpr int QH_T = 256;

// acc[i] *= beta for i < m
__global__ void __launch_bounds__(QH_T) acc_scale_kernel(ef2* __restrict__ acc, uint64_t m, ef2 beta) {
  const uint64_t i = (uint64_t)blockIdx.x * QH_T + threadIdx.x;
  if (i < m) acc[i] = ef_mul(acc[i], beta);
}

// acc[j] = acc[j mod m] for m <= j < target: reads only [0, m), which no
// thread writes, so the extension is one pass whatever target / m is
__global__ void __launch_bounds__(QH_T) acc_repeat_kernel(ef2* __restrict__ acc, uint64_t m, uint64_t target) {
  const uint64_t j = m + (uint64_t)blockIdx.x * QH_T + threadIdx.x;
  if (j < target) acc[j] = acc[j & (m - 1)];
}

// acc[i] = 0 for i < n (a kernel rather than hipMemsetAsync, so that the call
// can be recorded into a captured graph)
__global__ void __launch_bounds__(QH_T) acc_zero_kernel(ef2* __restrict__ acc, uint64_t n) {
  const uint64_t i = (uint64_t)blockIdx.x * QH_T + threadIdx.x;
  if (i < n) acc[i] = ef2{0, 0};
}

// num[i] *= inv[i mod d]: Z_H takes d values on the quotient coset
__global__ void __launch_bounds__(QH_T) vanishing_div_kernel(ef2* __restrict__ num, uint64_t total,
                                                             const uint64_t* __restrict__ inv, uint64_t dmask) {
  const uint64_t i = (uint64_t)blockIdx.x * QH_T + threadIdx.x;
  if (i < total) num[i] = ef_mul_base(num[i], inv[i & dmask]);
}

static uint32_t qh_grid(uint64_t n) {
  const uint64_t g = (n + QH_T - 1) / QH_T;
  if (g > 0x7fffffffull) throw ApiError(E_RANGE, "quotient buffer too large for one launch");
  return (uint32_t)g;
}

// cyclic_extend_and_scale (quotient.rs:40-55) in place on a buffer of target
// elements whose first m hold the accumulator: m == 0 zero-fills; otherwise
// the m elements are scaled by beta and repeated up to target.
void quotient_extend_scale(Ctx& ctx, ef2* acc, uint64_t m, uint64_t target, ef2 beta, hipStream_t s) {
  if (target == 0 || (target & (target - 1))) throw ApiError(E_NOT_POW2, "target length must be a power of two");
  if (m & (m - 1)) throw ApiError(E_NOT_POW2, "accumulator length must be a power of two");
  if (m > target) throw ApiError(E_INVAL, "target length below the accumulator length");
  (void)ctx;
  if (m == 0) {
    hipLaunchKernelGGL(acc_zero_kernel, dim3(qh_grid(target)), dim3(QH_T), 0, s, acc, target);
    P3G_CHECK_HIP(hipGetLastError());
    return;
  }
  hipLaunchKernelGGL(acc_scale_kernel, dim3(qh_grid(m)), dim3(QH_T), 0, s, acc, m, beta);
  P3G_CHECK_HIP(hipGetLastError());
  if (target > m) {
    hipLaunchKernelGGL(acc_repeat_kernel, dim3(qh_grid(target - m)), dim3(QH_T), 0, s, acc, m, target);
    P3G_CHECK_HIP(hipGetLastError());
  }
}

// divide_by_vanishing_in_place (quotient.rs:72-101) on n*d elements of the
// quotient coset s*J (|J| = n*d, natural order): Z_H(s w_J^i) = s^n w_d^(i mod d) - 1.
void quotient_divide_vanishing(Ctx& ctx, ef2* num, uint64_t n, uint64_t d, uint64_t shift, hipStream_t s) {
  if (n == 0 || (n & (n - 1))) throw ApiError(E_NOT_POW2, "trace height must be a power of two");
  if (d == 0 || (d & (d - 1))) throw ApiError(E_NOT_POW2, "constraint degree must be a power of two");
  if (d > (1u << 20)) throw ApiError(E_RANGE, "constraint degree above 2^20");
  const unsigned logn = log2_exact(n), logd = log2_exact(d);
  if (logn + logd > 32) throw ApiError(E_RANGE, "quotient domain exceeds the field two-adicity");
  uint64_t s_pow_n = shift;
  for (unsigned i = 0; i < logn; i++) s_pow_n = gl_mul(s_pow_n, s_pow_n);
  std::vector<uint64_t> z(d), inv(d);
  const uint64_t wd = gl_two_adic_generator(logd);
  uint64_t x = 1;
  for (uint64_t t = 0; t < d; t++) {
    z[t] = gl_sub(gl_mul(s_pow_n, x), 1);
    if (z[t] == 0) throw ApiError(E_INVAL, "the vanishing polynomial is zero on this coset (shift in H)");
    x = gl_mul(x, wd);
  }
  // batch inversion: prefix products, one inverse, then back
  uint64_t run = 1;
  for (uint64_t t = 0; t < d; t++) {
    inv[t] = run;
    run = gl_mul(run, z[t]);
  }
  uint64_t r = gl_inv(run);
  for (uint64_t t = d; t-- > 0;) {
    inv[t] = gl_mul(inv[t], r);
    r = gl_mul(r, z[t]);
  }
  uint64_t* dinv = (uint64_t*)ctx.upload(inv.data(), d * sizeof(uint64_t), s);
  const uint64_t total = n * d;
  hipLaunchKernelGGL(vanishing_div_kernel, dim3(qh_grid(total)), dim3(QH_T), 0, s, num, total, dinv, d - 1);
  const hipError_t e = hipGetLastError();
  ctx.release(dinv);  // stream-ordered: the block is reused only after the launch
  P3G_CHECK_HIP(e);
}

}  // namespace p3g
