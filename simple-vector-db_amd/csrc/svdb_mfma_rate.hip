// svdb_mfma_rate.hip — the sustained matrix-core rate of the device
// (gfx950): a loop of independent MFMAs per data type, the practical
// ceiling beside the published peaks.
#include "svdb_kernels.h"

namespace svdb {

typedef double f64x4 __attribute__((ext_vector_type(4)));

// ------------------------------------------------ MFMA throughput loop
//
// Independent MFMAs of one data type with no memory traffic in the loop
// (four accumulators per wavefront, operands fixed in registers): the
// sustained matrix throughput of this device, the practical ceiling the
// scans are measured against beside the published peak.  Each thread
// writes one float (out holds blocks · 256).
//
// This file is compiled with the MFMA accumulators forced into ordinary
// VGPRs, the form the scans use: with the accumulators in AGPRs (the
// compiler's choice for a loop this small) the same loops ran 30-60 %
// slower on MI355X (FP64 32-47 against 68-73 TF/s, BF16 16x16x32 0.9-1.4
// against 2.0-2.3 PF/s, FP8 16x16x128 at one wavefront per SIMD 2.9
// against 4.3 PF/s), which would understate the ceiling.

typedef float f32x4_t __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef int i32x8_t __attribute__((ext_vector_type(8)));

template <int KIND>
__global__ __launch_bounds__(256) void k_mfma_loop(float* __restrict__ out, int iters) {
  float s = 0.f;
  if constexpr (KIND == 0) {
    f64x4 acc[4];
#pragma unroll
    for (int t = 0; t < 4; t++) acc[t] = (f64x4){0.0, 0.0, 0.0, 0.0};
    const double a = 1e-3 * threadIdx.x, b = 0.5;
    for (int i = 0; i < iters; i++)
#pragma unroll
      for (int t = 0; t < 4; t++) acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[t], 0, 0, 0);
#pragma unroll
    for (int t = 0; t < 4; t++) s += (float)(acc[t][0] + acc[t][3]);
  } else if constexpr (KIND == 1) {
    f32x4_t acc[4];
#pragma unroll
    for (int t = 0; t < 4; t++) acc[t] = (f32x4_t){0.f, 0.f, 0.f, 0.f};
    bf16x8_t a, b;
#pragma unroll
    for (int i = 0; i < 8; i++) {
      a[i] = (__bf16)(1e-3f * threadIdx.x + i);
      b[i] = (__bf16)(0.5f * i);
    }
    for (int i = 0; i < iters; i++)
#pragma unroll
      for (int t = 0; t < 4; t++) acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[t], 0, 0, 0);
#pragma unroll
    for (int t = 0; t < 4; t++) s += acc[t][0] + acc[t][3];
  } else {
    f32x4_t acc[4];
#pragma unroll
    for (int t = 0; t < 4; t++) acc[t] = (f32x4_t){0.f, 0.f, 0.f, 0.f};
    i32x8_t a, b;
#pragma unroll
    for (int i = 0; i < 8; i++) {
      a[i] = 0x38383838 + (int)threadIdx.x;
      b[i] = 0x30303030 + i;
    }
    for (int i = 0; i < iters; i++)
#pragma unroll
      for (int t = 0; t < 4; t++)
        acc[t] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(a, b, acc[t], 0, 0, 0, 127, 0,
                                                                  127);
#pragma unroll
    for (int t = 0; t < 4; t++) s += acc[t][0] + acc[t][3];
  }
  out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

double mfma_loop_flop(int kind, uint32_t blocks, int iters) {
  const double per = kind == 0 ? 16.0 * 16 * 4 * 2 : kind == 1 ? 16.0 * 16 * 32 * 2
                                                                : 16.0 * 16 * 128 * 2;
  return per * 4.0 * iters * (double)blocks * 4.0;  // 4 MFMAs an iteration, 4 wavefronts a block
}

hipError_t launch_mfma_loop(int kind, uint32_t blocks, int iters, float* out, hipStream_t s) {
  switch (kind) {
    case 0: hipLaunchKernelGGL(k_mfma_loop<0>, dim3(blocks), dim3(256), 0, s, out, iters); break;
    case 1: hipLaunchKernelGGL(k_mfma_loop<1>, dim3(blocks), dim3(256), 0, s, out, iters); break;
    case 2: hipLaunchKernelGGL(k_mfma_loop<2>, dim3(blocks), dim3(256), 0, s, out, iters); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace svdb
