// BF16 GEMM on the matrix cores with FP32 accumulation, for the learned
// Newton initial guess (the only MFMA work in the project; the FP64 path has
// none).  C (fp32) = A (bf16, M x K) * B (bf16, K x N) [+ C], every operand
// addressed through (row, column) strides so that the forward products and
// both backward products of a dense layer (X W^T, dY^T X, dY W) are the same
// kernel without transposed copies.
//
// One wave computes one 16 x 16 tile of C with v_mfma_f32_16x16x16bf16_1k:
// lane l holds A(m0 + l % 16, k0 + 4 (l / 16) + t) and
// B(k0 + 4 (l / 16) + t, n0 + l % 16) for t = 0..3 and receives
// C(m0 + 4 (l / 16) + t, n0 + l % 16).  Operands outside M, N, K read as 0.
// Split-K: blockIdx.y takes every ksplit-th block of 16 in K and adds its
// partial tile to C with float atomics (C prepared by the caller).
#pragma once
#include <hip/hip_runtime.h>

namespace bca {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(256) k_gemm_bf16(int M, int N, int K, const unsigned short* __restrict__ A,
                                                   long sam, long sak, const unsigned short* __restrict__ B,
                                                   long sbk, long sbn, float* __restrict__ C, long scm, long scn,
                                                   int accumulate) {
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int tiles_n = (N + 15) / 16;
  const long tiles = (long)((M + 15) / 16) * tiles_n;
  const long tile = (long)blockIdx.x * 4 + wave;
  if (tile >= tiles) return;  // whole waves leave; the kernel has no barrier
  const int m0 = (int)(tile / tiles_n) * 16, n0 = (int)(tile % tiles_n) * 16;
  const int r = lane % 16, g = lane / 16;
  const int m = m0 + r, n = n0 + r;
  const int ksteps = (K + 15) / 16, ksplit = gridDim.y;
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  for (int ks = blockIdx.y; ks < ksteps; ks += ksplit) {
    bf16x4 a, b;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int k = ks * 16 + 4 * g + t;
      a[t] = (m < M && k < K) ? (short)A[m * sam + k * sak] : (short)0;
      b[t] = (n < N && k < K) ? (short)B[k * sbk + n * sbn] : (short)0;
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
  }
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int mo = m0 + 4 * g + t;
    if (mo < M && n < N) {
      float* p = C + mo * scm + n * scn;
      if (ksplit > 1) atomicAdd(p, acc[t]);
      else *p = accumulate ? *p + acc[t] : acc[t];
    }
  }
}

}  // namespace bca
