// LU without pivoting of the dense diagonal blocks of the Newton matrix, and
// the CSR product used to refine its solves.
//
// The Newton driver factors its block-banded matrix block by block (block
// Gaussian elimination in j- or i-line blocks); the diagonal block of each
// step is dense (B = 5 im or 5 jm unknowns, ~1000-2000) and factored here in
// place, A = L U with L unit lower, row-major with leading dimension ld:
//
//   for each 32-column panel k:
//     k_lu_panel_np   the 32 x 32 diagonal tile factored in LDS, then the
//                     row panel to its right (L^-1 A, a column per thread)
//                     and the column panel below it (A U^-1, a row per
//                     thread);
//     k_gemm_f64_sub  the trailing matrix minus the product of the two
//                     panels, 64 x 64 output tiles on the FP64 matrix cores
//                     (v_mfma_f64_16x16x4f64), panels staged in LDS.
//
// No pivoting: the relaxed (or pure) Newton Jacobians of these flows factor
// stably without it (element growth below 5 on the oracle's Jacobians, solves
// as accurate as the pivoted ones); a pivot below tiny * the tile's largest
// entry is reported so the caller can fall back, and the Newton driver
// refines every solve with the CSR residual.
#pragma once
#include <hip/hip_runtime.h>

namespace bca {

constexpr int LU_NB = 32;

typedef double f64x4 __attribute__((ext_vector_type(4)));

// One panel step at the diagonal tile A (ld), nb = min(LU_NB, rows left),
// m rows/columns past the tile; one wave per workgroup, so the barriers of
// the tile's 32 elimination steps cost next to nothing.  Every workgroup
// factors the tile in LDS (nb^3/3 flops, cheaper than a launch); workgroup 0
// reports a pivot below tiny times the tile's largest |entry| in *bad and
// stores the factored tile; workgroups 1..nrb solve row-panel columns c
// (A(0:nb, nb + c) = L^-1 A(0:nb, nb + c), unit L), the next nrb
// column-panel rows r (A(nb + r, 0:nb) = A(nb + r, 0:nb) U^-1), one per
// thread.
constexpr int LU_THREADS = 64;
__global__ void __launch_bounds__(LU_THREADS) k_lu_panel_np(double* __restrict__ A, long ld, int nb, int m, int nrb,
                                                            double tiny, int* __restrict__ bad,
                                                            double* __restrict__ work) {
  __shared__ double T[LU_NB][LU_NB + 1];
  __shared__ double X[LU_THREADS][LU_NB + 1];
  const int tid = threadIdx.x;
  double amax = 0.0;
  for (int x = tid; x < LU_NB * LU_NB; x += LU_THREADS) {
    const int r = x / LU_NB, c = x % LU_NB;
    const double v = (r < nb && c < nb) ? A[r * ld + c] : (r == c ? 1.0 : 0.0);
    T[r][c] = v;
    amax = fmax(amax, (r < nb && c < nb) ? fabs(v) : 0.0);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) amax = fmax(amax, __shfl_xor(amax, off));
  const double thr = tiny * amax;
  __syncthreads();
  for (int j = 0; j < nb; ++j) {
    const double p = T[j][j];
    if (blockIdx.x == 0 && tid == 0 && !(fabs(p) > thr)) *bad = 1;
    const double pinv = 1.0 / p;
    if (tid > j && tid < nb) T[tid][j] *= pinv;
    __syncthreads();
    const int q = nb - j - 1;
    for (int x = tid; x < q * q; x += LU_THREADS) {
      const int r = j + 1 + x / q, c = j + 1 + x % q;
      T[r][c] -= T[r][j] * T[j][c];
    }
    __syncthreads();
  }
  if (blockIdx.x == 0) {
    // the other workgroups may still be loading the tile from A: unless it
    // is alone (no panels) workgroup 0 leaves the factored tile in `work`,
    // and the trailing update's first workgroup puts it in place
    double* dst = m == 0 ? A : work;
    const long ldd = m == 0 ? ld : LU_NB;
    for (int x = tid; x < nb * nb; x += LU_THREADS) dst[(x / nb) * ldd + x % nb] = T[x / nb][x % nb];
    return;
  }
  // each thread solves one column (row panel) or row (column panel) of
  // length nb, kept in LDS
  double* x = X[tid];
  if ((int)blockIdx.x <= nrb) {
    const int c = (blockIdx.x - 1) * LU_THREADS + tid;
    if (c >= m) return;
    double* col = A + nb + c;
    for (int r = 0; r < nb; ++r) x[r] = col[r * ld];
    for (int r = 1; r < nb; ++r) {
      double sum = x[r];
      for (int i = 0; i < r; ++i) sum -= T[r][i] * x[i];
      x[r] = sum;
    }
    for (int r = 0; r < nb; ++r) col[r * ld] = x[r];
  } else {
    const int r = (blockIdx.x - 1 - nrb) * LU_THREADS + tid;
    if (r >= m) return;
    double* row = A + (long)(nb + r) * ld;
    for (int c = 0; c < nb; ++c) x[c] = row[c];
    for (int c = 0; c < nb; ++c) {
      double sum = x[c];
      for (int i = 0; i < c; ++i) sum -= x[i] * T[i][c];
      x[c] = sum / T[c][c];
    }
    for (int c = 0; c < nb; ++c) row[c] = x[c];
  }
}

// C (M x N, ldc) -= A (M x K, lda) B (K x N, ldb), row-major, K <= LU_NB.
// Workgroup of 4 waves per 64 x 64 tile of C, each wave a 32 x 32 quarter
// as 2 x 2 MFMA tiles.  v_mfma_f64_16x16x4f64: lane l holds A(l % 16, l / 16)
// and B(l / 16, l % 16) of the 16 x 4 / 4 x 16 operands and receives
// D(l / 16 + 4 t, l % 16) in its t-th result, t = 0..3 (checked against
// torch's FP64 product by the GPU tests).  Operands past M, N, K read as 0.
// With tile != nullptr, workgroup (0, 0) also copies the factored nb x nb
// tile from `work` (ld LU_NB) to `tile` (ldt): the panel step's write-back.
__global__ void __launch_bounds__(256) k_gemm_f64_sub(int M, int N, int K, const double* __restrict__ A, long lda,
                                                      const double* __restrict__ B, long ldb,
                                                      double* __restrict__ C, long ldc,
                                                      double* __restrict__ tile = nullptr, long ldt = 0,
                                                      const double* __restrict__ work = nullptr, int nb = 0) {
  __shared__ double As[64][LU_NB + 1];
  __shared__ double Bs[LU_NB][64 + 1];
  const int tid = threadIdx.x, wave = tid / 64, lane = tid % 64;
  if (tile && blockIdx.x == 0 && blockIdx.y == 0)
    for (int x = tid; x < nb * nb; x += 256) tile[(x / nb) * ldt + x % nb] = work[(x / nb) * LU_NB + x % nb];
  const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
  for (int x = tid; x < 64 * LU_NB; x += 256) {
    const int r = x / LU_NB, k = x % LU_NB;
    As[r][k] = (m0 + r < M && k < K) ? A[(long)(m0 + r) * lda + k] : 0.0;
    const int kb = x / 64, c = x % 64;
    Bs[kb][c] = (kb < K && n0 + c < N) ? B[(long)kb * ldb + n0 + c] : 0.0;
  }
  __syncthreads();
  const int wm = (wave / 2) * 32, wn = (wave % 2) * 32;
  const int lr = lane % 16, lk = lane / 16;
  f64x4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = f64x4{0.0, 0.0, 0.0, 0.0};
  const int kend = (K + 3) / 4 * 4;
  for (int k0 = 0; k0 < kend; k0 += 4) {
    double av[2], bv[2];
#pragma unroll
    for (int a = 0; a < 2; ++a) av[a] = As[wm + 16 * a + lr][k0 + lk];
#pragma unroll
    for (int b = 0; b < 2; ++b) bv[b] = Bs[k0 + lk][wn + 16 * b + lr];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a], bv[b], acc[a][b], 0, 0, 0);
  }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int r = m0 + wm + 16 * a + lk + 4 * t, c = n0 + wn + 16 * b + lr;
        if (r < M && c < N) C[(long)r * ldc + c] -= acc[a][b][t];
      }
}

// y = A x for a CSR matrix (one row per thread).
__global__ void k_csr_spmv(long n, const int* __restrict__ indptr, const int* __restrict__ indices,
                           const double* __restrict__ data, const double* __restrict__ x, double* __restrict__ y) {
  const long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  double s = 0.0;
  for (int p = indptr[r]; p < indptr[r + 1]; ++p) s += data[p] * x[indices[p]];
  y[r] = s;
}

}  // namespace bca
