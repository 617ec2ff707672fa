// The device field operations of the hashing and NTT kernels run on caller
// operands, one wave per 64 operands: instrumentation for tests of the
// operations' rare branches (the carry and borrow cases of the v_mad_u64_u32
// products that random data reaches with probability ~2^-32, non-canonical
// inputs, every power-of-two twiddle exponent).  Nothing here is on the
// proving path.
#include <hip/hip_runtime.h>

#include "internal.hpp"
#include "ntt_ops.hpp"

namespace p3g {

namespace {

// op 0: EXACT a*b        1: TRY a*b (wave flag)     2: EXACT a*b + c
//    3: TRY a*b + c      4: canonical a*b (mulc)    5: canonical a*2^e (mul_pow2c)
//    6: (a + 2^32 b) mod p from 64-bit sums a, b < 2^40 (red_halves)
//    7: canonical a + b (add_c)   8: canonical a - b (sub_c)   9: a mod p (canon3)
__global__ void __launch_bounds__(64) field_check_kernel(uint32_t op, uint32_t e, const uint64_t* __restrict__ a,
                                                         const uint64_t* __restrict__ b,
                                                         const uint64_t* __restrict__ c, uint64_t n,
                                                         uint64_t* __restrict__ out, uint64_t* __restrict__ flags) {
  const uint64_t i = (uint64_t)blockIdx.x * 64 + threadIdx.x;
  // every lane computes (lanes past n on zeros), so each wave's flag covers its 64 operands
  const uint64_t x = i < n ? a[i] : 0, y = i < n ? b[i] : 0, z = i < n ? c[i] : 0;
  uint64_t f = 0, r = 0;
  switch (op) {
    case 0: r = mul_t<false>(x, y, f); break;
    case 1: r = mul_t<true>(x, y, f); break;
    case 2: r = mul_add_t<false>(x, y, z, f); break;
    case 3: r = mul_add_t<true>(x, y, z, f); break;
    case 4: r = mulc(x, y); break;
    case 5: r = mul_pow2c(x, e); break;
    case 6: r = red_halves(x, y); break;
    case 7: r = add_c(x, y); break;
    case 8: r = sub_c(x, y); break;
    default: r = canon3(x); break;
  }
  if (i < n) out[i] = r;
  if (threadIdx.x == 0) flags[blockIdx.x] = f;
}

}  // namespace

void field_check(Ctx& ctx, uint32_t op, uint32_t e, const uint64_t* a, const uint64_t* b, const uint64_t* c,
                 uint64_t n, uint64_t* out, uint64_t* wave_flags) {
  if (op > 9) throw ApiError(E_INVAL, "field_check: op is 0..9");
  if (op == 5 && e >= 96) throw ApiError(E_INVAL, "field_check: e is 0..95");
  if (n == 0 || n > (1u << 20)) throw ApiError(E_INVAL, "field_check: n is 1..2^20");
  hipStream_t s = ctx.stream;
  const uint64_t waves = (n + 63) / 64;
  uint64_t* d = static_cast<uint64_t*>(ctx.alloc(size_t(4 * n + waves) * 8));
  uint64_t *da = d, *db = d + n, *dc = d + 2 * n, *dout = d + 3 * n, *dfl = d + 4 * n;
  P3G_CHECK_HIP(hipMemcpyAsync(da, a, n * 8, hipMemcpyHostToDevice, s));
  P3G_CHECK_HIP(hipMemcpyAsync(db, b, n * 8, hipMemcpyHostToDevice, s));
  if (c) P3G_CHECK_HIP(hipMemcpyAsync(dc, c, n * 8, hipMemcpyHostToDevice, s));
  else P3G_CHECK_HIP(hipMemsetAsync(dc, 0, n * 8, s));
  hipLaunchKernelGGL(field_check_kernel, dim3((uint32_t)waves), dim3(64), 0, s, op, e, da, db, dc, n, dout, dfl);
  P3G_CHECK_HIP(hipGetLastError());
  P3G_CHECK_HIP(hipMemcpyAsync(out, dout, n * 8, hipMemcpyDeviceToHost, s));
  if (wave_flags) P3G_CHECK_HIP(hipMemcpyAsync(wave_flags, dfl, waves * 8, hipMemcpyDeviceToHost, s));
  P3G_CHECK_HIP(hipStreamSynchronize(s));
  ctx.release(d);
}

}  // namespace p3g
