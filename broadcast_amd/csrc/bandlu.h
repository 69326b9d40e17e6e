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

// ------------------------------------------------------------------ inverse
// Inverse without pivoting of a dense n x n block by Gauss-Jordan steps on
// GJ_NB-column panels, each step one launch from Ain to Aout (every entry
// written, so the caller alternates two buffers).  For the panel kk =
// k0..k0+nb-1 with pivot tile P = A(kk, kk):
//
//   Aout(kk, kk) = P^-1                      Aout(kk, j) = P^-1 A(kk, j)
//   Aout(i, kk)  = -A(i, kk) P^-1            Aout(i, j)  = A(i, j) - A(i, kk) P^-1 A(kk, j)
//
// for i, j outside kk.  Every entry is one product out = A' - T U of a
// GJ_T x GJ_T output tile, with T = A(i, kk) P^-1 (rows of kk: T = -P^-1)
// and U = A(kk, j) (columns of kk: U = the identity), A' = A off the panel
// rows and columns and 0 on them.  A workgroup (4 waves, one 64 x 64
// quadrant each, v_mfma_f64_16x16x4f64) first issues the loads of its A'
// tile into registers and of A(i, kk), A(kk, j) into LDS; meanwhile wave 0
// inverts P in registers (lane r + 32 h holds row r, columns 16 h .. 16 h +
// 15; each elimination step broadcasts the pivot row by lane shuffles).
// Workgroup (0, 0) reports a pivot below tiny times the tile's largest
// |entry| in *bad.  After n / GJ_NB steps Aout holds A^-1 (2 n^3 flops).
// The Newton driver multiplies the off-diagonal blocks of its block row by
// this inverse (one GEMM instead of two triangular solves) and applies it to
// the right-hand sides of its solves.
constexpr int GJ_NB = 32;
constexpr int GJ_T = 128;
__global__ void __launch_bounds__(256) k_gj_step(const double* __restrict__ Ain, long lda, double* __restrict__ Aout,
                                                 long ldo, int n, int k0, int nb, double tiny, int* __restrict__ bad) {
  __shared__ double P[GJ_NB][GJ_NB + 1];
  // A(i, kk) of the tile's rows, overwritten by T (each wave reads and then
  // writes only its own 32 rows)
  __shared__ double AT[GJ_T][GJ_NB + 1];
  __shared__ double U[GJ_NB][GJ_T + 1];
  const int tid = threadIdx.x, wave = tid / 64, lane = tid % 64;
  const int lr = lane % 16, lk = lane / 16;
  const int i0 = blockIdx.y * GJ_T, j0 = blockIdx.x * GJ_T;
  const int wm = (wave / 2) * 64, wn = (wave % 2) * 64;
  auto in_panel = [&](int g) { return g >= k0 && g < k0 + nb; };
  // A' of this thread's output entries: D(lk + 4 t, lr) of MFMA tile (a, b)
  double a0[4][4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int gi = i0 + wm + 16 * a + lk + 4 * t, gj = j0 + wn + 16 * b + lr;
        a0[a][b][t] = (gi < n && gj < n && !in_panel(gi) && !in_panel(gj)) ? Ain[(long)gi * lda + gj] : 0.0;
      }
  for (int x = tid; x < GJ_T * GJ_NB; x += 256) {
    const int r = x / GJ_NB, c = x % GJ_NB;
    const int gi = i0 + r;
    AT[r][c] = (gi < n && c < nb) ? Ain[(long)gi * lda + k0 + c] : 0.0;
    const int q = x / GJ_T, cc = x % GJ_T;
    const int gj = j0 + cc;
    double u = 0.0;
    if (in_panel(gj)) u = (q == gj - k0) ? 1.0 : 0.0;
    else if (gj < n && q < nb) u = Ain[(long)(k0 + q) * lda + gj];
    U[q][cc] = u;
  }
  if (wave == 0) {
    const int r = lane % 32, h = lane / 32;
    double p[16];
    double amax = 0.0;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int c = 16 * h + q;
      const bool in = r < nb && c < nb;
      p[q] = in ? Ain[(long)(k0 + r) * lda + k0 + c] : (r == c ? 1.0 : 0.0);
      amax = fmax(amax, in ? fabs(p[q]) : 0.0);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) amax = fmax(amax, __shfl_xor(amax, off));
    const double thr = tiny * amax;
#pragma unroll
    for (int k = 0; k < GJ_NB; ++k) {
      const int kh = k / 16, kq = k % 16;
      double rowk[16];
#pragma unroll
      for (int q = 0; q < 16; ++q) rowk[q] = __shfl(p[q], k + 32 * h);
      const double piv = __shfl(p[kq], k + 32 * kh);
      const double f = __shfl(p[kq], r + 32 * kh);
      if (k < nb && blockIdx.x == 0 && blockIdx.y == 0 && lane == 0 && !(fabs(piv) > thr)) *bad = 1;
      const double pinv = 1.0 / piv;
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int c = 16 * h + q;
        const double nk = rowk[q] * pinv;
        if (r == k) p[q] = (c == k) ? pinv : nk;
        else p[q] = (c == k) ? -f * pinv : p[q] - f * nk;
      }
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) P[r][16 * h + q] = p[q];
  }
  __syncthreads();
  // T = A(i, kk) P^-1: wave w rows 32 w .. 32 w + 31
  {
    f64x4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) acc[a][b] = f64x4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < GJ_NB; k += 4) {
      double av[2], bv[2];
#pragma unroll
      for (int a = 0; a < 2; ++a) av[a] = AT[32 * wave + 16 * a + lr][k + lk];
#pragma unroll
      for (int b = 0; b < 2; ++b) bv[b] = P[k + lk][16 * b + lr];
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
          const int r = 32 * wave + 16 * a + lk + 4 * t, c = 16 * b + lr;
          const int gi = i0 + r;
          AT[r][c] = in_panel(gi) ? -P[gi - k0][c] : acc[a][b][t];
        }
  }
  __syncthreads();
  // out = A' - T U over the wave's 64 x 64 quadrant, 4 x 4 MFMA tiles
  f64x4 acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = f64x4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int k = 0; k < GJ_NB; k += 4) {
    double av[4], bv[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) av[a] = AT[wm + 16 * a + lr][k + lk];
#pragma unroll
    for (int b = 0; b < 4; ++b) bv[b] = U[k + lk][wn + 16 * b + lr];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a], bv[b], acc[a][b], 0, 0, 0);
  }
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int gi = i0 + wm + 16 * a + lk + 4 * t, gj = j0 + wn + 16 * b + lr;
        if (gi < n && gj < n) Aout[(long)gi * ldo + gj] = a0[a][b][t] - acc[a][b][t];
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
