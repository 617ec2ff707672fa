// Peak calibration kernels: the measured ceilings that the roofline figures of
// the hashing and transform kernels are quoted against.
//
//  - VALU issue rate, per instruction class: each lane runs independent chains
//    of one instruction (written with inline asm, so the count per wave is
//    exact and the compiler cannot fold the chains); the launch covers every
//    CU many times over, so the rate is the chip's sustained issue rate for
//    that class.
//  - Device copy bandwidth: a 16-bytes-per-lane copy kernel and the runtime's
//    own device-to-device copy over a buffer far larger than the L2 and MALL.
//
// Nothing here is on the proving path.
#include <hip/hip_runtime.h>

#include "internal.hpp"

namespace p3g {

namespace {

constexpr int CAL_THREADS = 256;
constexpr int CAL_UNROLL = 4;

// kind 0: v_add_u32, 16 chains
__global__ __launch_bounds__(CAL_THREADS) void cal_add_kernel(uint32_t iters, uint32_t k, uint64_t* out) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t a[16];
#pragma unroll
  for (int i = 0; i < 16; i++) a[i] = t ^ (0x9e3779b9u * (i + 1));
  for (uint32_t it = 0; it < iters; it += CAL_UNROLL) {
#pragma unroll
    for (int u = 0; u < CAL_UNROLL; u++) {
#pragma unroll
      for (int i = 0; i < 16; i++) asm volatile("v_add_u32 %0, %0, %1" : "+v"(a[i]) : "v"(k));
    }
  }
  uint32_t x = 0;
#pragma unroll
  for (int i = 0; i < 16; i++) x ^= a[i];
  out[t] = x;
}

// kind 1: v_mad_u64_u32 (32 x 32 + 64 -> 64, the partial product of a Goldilocks multiply), 8 chains
__global__ __launch_bounds__(CAL_THREADS) void cal_mad64_kernel(uint32_t iters, uint32_t k, uint64_t* out) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t acc[8];
  uint32_t x[8];
#pragma unroll
  for (int i = 0; i < 8; i++) {
    acc[i] = t + i;
    x[i] = t * 0x01000193u + i;
  }
  for (uint32_t it = 0; it < iters; it += CAL_UNROLL) {
#pragma unroll
    for (int u = 0; u < CAL_UNROLL; u++) {
      // a carry pair of its own per chain: back-to-back VALU writes of one SGPR pair need a wait state
      asm volatile(
          "v_mad_u64_u32 %0, s[40:41], %8, %16, %0\n\t"
          "v_mad_u64_u32 %1, s[42:43], %9, %16, %1\n\t"
          "v_mad_u64_u32 %2, s[44:45], %10, %16, %2\n\t"
          "v_mad_u64_u32 %3, s[46:47], %11, %16, %3\n\t"
          "v_mad_u64_u32 %4, s[48:49], %12, %16, %4\n\t"
          "v_mad_u64_u32 %5, s[50:51], %13, %16, %5\n\t"
          "v_mad_u64_u32 %6, s[52:53], %14, %16, %6\n\t"
          "v_mad_u64_u32 %7, s[54:55], %15, %16, %7"
          : "+v"(acc[0]), "+v"(acc[1]), "+v"(acc[2]), "+v"(acc[3]), "+v"(acc[4]), "+v"(acc[5]), "+v"(acc[6]),
            "+v"(acc[7])
          : "v"(x[0]), "v"(x[1]), "v"(x[2]), "v"(x[3]), "v"(x[4]), "v"(x[5]), "v"(x[6]), "v"(x[7]), "v"(k)
          : "s40", "s41", "s42", "s43", "s44", "s45", "s46", "s47", "s48", "s49", "s50", "s51", "s52", "s53", "s54",
            "s55");
    }
  }
  uint64_t r = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) r ^= acc[i];
  out[t] = r;
}

// kind 2: v_fma_f64, 8 chains
__global__ __launch_bounds__(CAL_THREADS) void cal_fma64_kernel(uint32_t iters, uint32_t k, uint64_t* out) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  double d[8];
  const double m = 1.0 + 1e-12 * (k & 7), c = 1e-9;
#pragma unroll
  for (int i = 0; i < 8; i++) d[i] = 1.0 + t * 1e-9 + i;
  for (uint32_t it = 0; it < iters; it += CAL_UNROLL) {
#pragma unroll
    for (int u = 0; u < CAL_UNROLL; u++) {
#pragma unroll
      for (int i = 0; i < 8; i++) asm volatile("v_fma_f64 %0, %1, %2, %0" : "+v"(d[i]) : "v"(m), "v"(c));
    }
  }
  double s = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) s += d[i];
  out[t] = __double_as_longlong(s);
}

// kind 3: v_mul_hi_u32, 16 chains
__global__ __launch_bounds__(CAL_THREADS) void cal_mulhi_kernel(uint32_t iters, uint32_t k, uint64_t* out) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t a[16];
#pragma unroll
  for (int i = 0; i < 16; i++) a[i] = t ^ (0x9e3779b9u * (i + 1));
  for (uint32_t it = 0; it < iters; it += CAL_UNROLL) {
#pragma unroll
    for (int u = 0; u < CAL_UNROLL; u++) {
#pragma unroll
      for (int i = 0; i < 16; i++) asm volatile("v_mul_hi_u32 %0, %0, %1" : "+v"(a[i]) : "v"(k));
    }
  }
  uint32_t x = 0;
#pragma unroll
  for (int i = 0; i < 16; i++) x ^= a[i];
  out[t] = x;
}

// kind 4: 64-bit add on a carry pair (v_add_co_u32 then v_addc_co_u32 through vcc), 8 chains
__global__ __launch_bounds__(CAL_THREADS) void cal_add64_kernel(uint32_t iters, uint32_t k, uint64_t* out) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t lo[8], hi[8];
  const uint32_t k1 = k * 3u + 1u;
#pragma unroll
  for (int i = 0; i < 8; i++) {
    lo[i] = t + i;
    hi[i] = t ^ i;
  }
  for (uint32_t it = 0; it < iters; it += CAL_UNROLL) {
#pragma unroll
    for (int u = 0; u < CAL_UNROLL; u++) {
#pragma unroll
      for (int i = 0; i < 8; i++)
        asm volatile(
            "v_add_co_u32 %0, vcc, %0, %2\n\t"
            "v_addc_co_u32 %1, vcc, %1, %3, vcc"
            : "+v"(lo[i]), "+v"(hi[i])
            : "v"(k), "v"(k1)
            : "vcc");
    }
  }
  uint64_t r = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) r ^= (uint64_t(hi[i]) << 32) | lo[i];
  out[t] = r;
}

// kinds 5..11: one more instruction class each, 16 independent chains per lane
//   5 v_mov_b32 (a rotation of the chains), 6 v_cndmask_b32 on an SGPR mask, 7 v_add_co_u32
//   (VOP3, a carry-out SGPR pair per chain), 8 v_lshl_add_u64 (64-bit add, 8 chains),
//   9 v_add3_u32, 10 v_xor_b32, 11 v_alignbit_b32
template <int KIND>
__global__ __launch_bounds__(CAL_THREADS) void cal_misc_kernel(uint32_t iters, uint32_t k, uint64_t* out) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t a[16];
#pragma unroll
  for (int i = 0; i < 16; i++) a[i] = t ^ (0x9e3779b9u * (i + 1));
  uint64_t q[8];
#pragma unroll
  for (int i = 0; i < 8; i++) q[i] = (uint64_t(a[2 * i]) << 32) | a[2 * i + 1];
  const uint64_t k64 = (uint64_t(k) << 32) | (k * 3u);
  for (uint32_t it = 0; it < iters; it += CAL_UNROLL) {
#pragma unroll
    for (int u = 0; u < CAL_UNROLL; u++) {
      if constexpr (KIND == 5) {
#pragma unroll
        for (int i = 0; i < 16; i++) asm volatile("v_mov_b32 %0, %1" : "=v"(a[i]) : "v"(a[(i + 1) & 15]));
      } else if constexpr (KIND == 6) {
#pragma unroll
        for (int i = 0; i < 16; i++) asm volatile("v_cndmask_b32_e64 %0, %0, %1, %2" : "+v"(a[i]) : "v"(k), "s"(k64));
      } else if constexpr (KIND == 7) {
        asm volatile(
            "v_add_co_u32_e64 %0, s[40:41], %0, %16\n\t"
            "v_add_co_u32_e64 %1, s[42:43], %1, %16\n\t"
            "v_add_co_u32_e64 %2, s[44:45], %2, %16\n\t"
            "v_add_co_u32_e64 %3, s[46:47], %3, %16\n\t"
            "v_add_co_u32_e64 %4, s[48:49], %4, %16\n\t"
            "v_add_co_u32_e64 %5, s[50:51], %5, %16\n\t"
            "v_add_co_u32_e64 %6, s[52:53], %6, %16\n\t"
            "v_add_co_u32_e64 %7, s[54:55], %7, %16\n\t"
            "v_add_co_u32_e64 %8, s[56:57], %8, %16\n\t"
            "v_add_co_u32_e64 %9, s[58:59], %9, %16\n\t"
            "v_add_co_u32_e64 %10, s[60:61], %10, %16\n\t"
            "v_add_co_u32_e64 %11, s[62:63], %11, %16\n\t"
            "v_add_co_u32_e64 %12, s[64:65], %12, %16\n\t"
            "v_add_co_u32_e64 %13, s[66:67], %13, %16\n\t"
            "v_add_co_u32_e64 %14, s[68:69], %14, %16\n\t"
            "v_add_co_u32_e64 %15, s[70:71], %15, %16"
            : "+v"(a[0]), "+v"(a[1]), "+v"(a[2]), "+v"(a[3]), "+v"(a[4]), "+v"(a[5]), "+v"(a[6]), "+v"(a[7]),
              "+v"(a[8]), "+v"(a[9]), "+v"(a[10]), "+v"(a[11]), "+v"(a[12]), "+v"(a[13]), "+v"(a[14]), "+v"(a[15])
            : "v"(k)
            : "s40", "s41", "s42", "s43", "s44", "s45", "s46", "s47", "s48", "s49", "s50", "s51", "s52", "s53",
              "s54", "s55", "s56", "s57", "s58", "s59", "s60", "s61", "s62", "s63", "s64", "s65", "s66", "s67",
              "s68", "s69", "s70", "s71");
      } else if constexpr (KIND == 8) {
#pragma unroll
        for (int i = 0; i < 8; i++) asm volatile("v_lshl_add_u64 %0, %0, 0, %1" : "+v"(q[i]) : "v"(k64));
      } else if constexpr (KIND == 9) {
#pragma unroll
        for (int i = 0; i < 16; i++) asm volatile("v_add3_u32 %0, %0, %1, %1" : "+v"(a[i]) : "v"(k));
      } else if constexpr (KIND == 10) {
#pragma unroll
        for (int i = 0; i < 16; i++) asm volatile("v_xor_b32 %0, %0, %1" : "+v"(a[i]) : "v"(k));
      } else {
#pragma unroll
        for (int i = 0; i < 16; i++) asm volatile("v_alignbit_b32 %0, %0, %1, 7" : "+v"(a[i]) : "v"(k));
      }
    }
  }
  uint64_t x = 0;
#pragma unroll
  for (int i = 0; i < 16; i++) x ^= a[i];
#pragma unroll
  for (int i = 0; i < 8; i++) x ^= q[i];
  out[t] = x;
}

// kinds 12..15: the pairing of VOP2 forms, alone and beside 64-bit multiply-adds
//   12 v_cndmask_b32_e32 (VOP2, mask in vcc), 16 chains
//   13 v_mul_u32_u24_e32 (VOP2), 16 chains
//   14 v_mad_u64_u32 and v_add_u32_e32 alternating, 8 chains of each
//   15 v_mad_u64_u32 and v_cndmask_b32_e32 alternating, 8 chains of each
template <int KIND>
__global__ __launch_bounds__(CAL_THREADS) void cal_pair_kernel(uint32_t iters, uint32_t k, uint64_t* out) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t a[16];
#pragma unroll
  for (int i = 0; i < 16; i++) a[i] = t ^ (0x9e3779b9u * (i + 1));
  uint64_t q[8];
#pragma unroll
  for (int i = 0; i < 8; i++) q[i] = (uint64_t(a[2 * i]) << 32) | a[2 * i + 1];
  const uint64_t k64 = (uint64_t(k) << 32) | (k * 3u);
  for (uint32_t it = 0; it < iters; it += CAL_UNROLL) {
#pragma unroll
    for (int u = 0; u < CAL_UNROLL; u++) {
      if constexpr (KIND == 12) {
#pragma unroll
        for (int i = 0; i < 16; i++)
          asm volatile("s_mov_b64 vcc, %2\n\tv_cndmask_b32_e32 %0, %0, %1, vcc" : "+v"(a[i]) : "v"(k), "s"(k64) : "vcc");
      } else if constexpr (KIND == 13) {
#pragma unroll
        for (int i = 0; i < 16; i++) asm volatile("v_mul_u32_u24_e32 %0, %0, %1" : "+v"(a[i]) : "v"(k));
      } else if constexpr (KIND == 14) {
#pragma unroll
        for (int i = 0; i < 8; i++)
          asm volatile("v_mad_u64_u32 %0, vcc, %2, %3, %0\n\tv_add_u32_e32 %1, %1, %3"
                       : "+v"(q[i]), "+v"(a[i])
                       : "v"(a[8 + i]), "v"(k)
                       : "vcc");
      } else {
#pragma unroll
        for (int i = 0; i < 8; i++)
          asm volatile("v_mad_u64_u32 %0, s[40:41], %2, %3, %0\n\tv_cndmask_b32_e32 %1, %1, %3, vcc"
                       : "+v"(q[i]), "+v"(a[i])
                       : "v"(a[8 + i]), "v"(k)
                       : "s40", "s41");
      }
    }
  }
  uint64_t x = 0;
#pragma unroll
  for (int i = 0; i < 16; i++) x ^= a[i];
#pragma unroll
  for (int i = 0; i < 8; i++) x ^= q[i];
  out[t] = x;
}

__global__ __launch_bounds__(CAL_THREADS) void cal_copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                              uint64_t n16) {
  const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
  uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  for (; i + 3 * stride < n16; i += 4 * stride) {
    const uint4 a = src[i], b = src[i + stride], c = src[i + 2 * stride], d = src[i + 3 * stride];
    dst[i] = a;
    dst[i + stride] = b;
    dst[i + 2 * stride] = c;
    dst[i + 3 * stride] = d;
  }
  for (; i < n16; i += stride) dst[i] = src[i];
}

}  // namespace

// Instructions of the measured class that one wave issues for `iters` (a
// multiple of CAL_UNROLL) iterations.
uint64_t calib_valu_insts_per_wave(unsigned kind, uint32_t iters) {
  // kind 12 counts the selects only (each follows an s_mov_b64 to vcc)
  static const uint32_t per_iter[16] = {16, 8, 8, 16, 16, 16, 16, 16, 8, 16, 16, 16, 16, 16, 16, 16};
  return uint64_t(iters) * per_iter[kind];
}

double calib_valu(Ctx& ctx, unsigned kind, uint32_t iters, uint32_t blocks) {
  if (kind > 15) throw ApiError(E_INVAL, "calibrate_valu: kind is 0..15");
  if (iters == 0 || iters % CAL_UNROLL) throw ApiError(E_INVAL, "calibrate_valu: iters must be a positive multiple of 4");
  if (blocks == 0 || blocks > (1u << 20)) throw ApiError(E_INVAL, "calibrate_valu: blocks must be 1..2^20");
  hipStream_t s = ctx.stream;
  uint64_t* out = static_cast<uint64_t*>(ctx.alloc(size_t(blocks) * CAL_THREADS * 8));
  auto launch = [&](uint32_t n) {
    switch (kind) {
      case 0: cal_add_kernel<<<blocks, CAL_THREADS, 0, s>>>(n, 0x3u, out); break;
      case 1: cal_mad64_kernel<<<blocks, CAL_THREADS, 0, s>>>(n, 0x9e3779b9u, out); break;
      case 2: cal_fma64_kernel<<<blocks, CAL_THREADS, 0, s>>>(n, 5u, out); break;
      case 3: cal_mulhi_kernel<<<blocks, CAL_THREADS, 0, s>>>(n, 0xfffffff1u, out); break;
      case 4: cal_add64_kernel<<<blocks, CAL_THREADS, 0, s>>>(n, 0x7u, out); break;
      case 5: cal_misc_kernel<5><<<blocks, CAL_THREADS, 0, s>>>(n, 0x7u, out); break;
      case 6: cal_misc_kernel<6><<<blocks, CAL_THREADS, 0, s>>>(n, 0x7u, out); break;
      case 7: cal_misc_kernel<7><<<blocks, CAL_THREADS, 0, s>>>(n, 0x7u, out); break;
      case 8: cal_misc_kernel<8><<<blocks, CAL_THREADS, 0, s>>>(n, 0x7u, out); break;
      case 9: cal_misc_kernel<9><<<blocks, CAL_THREADS, 0, s>>>(n, 0x7u, out); break;
      case 10: cal_misc_kernel<10><<<blocks, CAL_THREADS, 0, s>>>(n, 0x7u, out); break;
      case 11: cal_misc_kernel<11><<<blocks, CAL_THREADS, 0, s>>>(n, 0x7u, out); break;
      case 12: cal_pair_kernel<12><<<blocks, CAL_THREADS, 0, s>>>(n, 0x7u, out); break;
      case 13: cal_pair_kernel<13><<<blocks, CAL_THREADS, 0, s>>>(n, 0x7u, out); break;
      case 14: cal_pair_kernel<14><<<blocks, CAL_THREADS, 0, s>>>(n, 0x7u, out); break;
      default: cal_pair_kernel<15><<<blocks, CAL_THREADS, 0, s>>>(n, 0x7u, out); break;
    }
    P3G_CHECK_HIP(hipGetLastError());
  };
  hipEvent_t a, b;
  P3G_CHECK_HIP(hipEventCreate(&a));
  P3G_CHECK_HIP(hipEventCreate(&b));
  launch(CAL_UNROLL);  // warm-up (code object load, clocks)
  P3G_CHECK_HIP(hipEventRecord(a, s));
  launch(iters);
  P3G_CHECK_HIP(hipEventRecord(b, s));
  P3G_CHECK_HIP(hipEventSynchronize(b));
  float ms = 0;
  P3G_CHECK_HIP(hipEventElapsedTime(&ms, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  ctx.release(out);
  return ms;
}

void calib_copy(Ctx& ctx, uint64_t bytes, uint32_t reps, double* kernel_ms, double* memcpy_ms) {
  if (bytes < 16 || bytes % 16) throw ApiError(E_INVAL, "calibrate_copy: bytes must be a positive multiple of 16");
  if (reps == 0) throw ApiError(E_INVAL, "calibrate_copy: reps must be positive");
  hipStream_t s = ctx.stream;
  void* src = ctx.alloc(bytes);
  void* dst = ctx.alloc(bytes);
  P3G_CHECK_HIP(hipMemsetAsync(src, 0x5a, bytes, s));
  const uint64_t n16 = bytes / 16;
  const unsigned blocks = 8192;
  hipEvent_t a, b;
  P3G_CHECK_HIP(hipEventCreate(&a));
  P3G_CHECK_HIP(hipEventCreate(&b));
  auto kcopy = [&] {
    cal_copy_kernel<<<blocks, CAL_THREADS, 0, s>>>(static_cast<const uint4*>(src), static_cast<uint4*>(dst), n16);
    P3G_CHECK_HIP(hipGetLastError());
  };
  auto time_it = [&](auto&& f) {
    f();  // warm-up
    P3G_CHECK_HIP(hipEventRecord(a, s));
    for (uint32_t r = 0; r < reps; r++) f();
    P3G_CHECK_HIP(hipEventRecord(b, s));
    P3G_CHECK_HIP(hipEventSynchronize(b));
    float ms = 0;
    P3G_CHECK_HIP(hipEventElapsedTime(&ms, a, b));
    return double(ms) / reps;
  };
  *kernel_ms = time_it(kcopy);
  *memcpy_ms = time_it([&] { P3G_CHECK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, s)); });
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  ctx.release(src);
  ctx.release(dst);
}

}  // namespace p3g
