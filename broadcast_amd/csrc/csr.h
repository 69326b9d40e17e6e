// Pruned CSR of the colored COO, on the device.
//
// The reference keeps the COO entries with |J| > mini (remove_zero_jac,
// BROADCAST_npz.py:129-135) and hands them to scipy.sparse.csr_matrix
// (misc/PETSc_func.py:85), which sums duplicate (row, column) pairs and sorts
// each row by column.  The colored COO has a fixed structure: the slots of
// row (i, j, e) are (m, l, k[, zone]) at known strides, and the column of a
// slot is the window cell (vali(i, l[, zone]), valj(j, k)) and variable m.
// So each row is produced directly in column order -- distinct vali
// ascending, then valj ascending, then m -- summing the slots that share a
// column (at most one per zone) in slot order, with no sort of the COO.
//
// Two passes over the values: count (row lengths), exclusive scan into
// indptr, fill.  Reads are coalesced (a wave holds consecutive i of one row
// kind); each lane writes its row's run.
#pragma once
#include "jacobian.h"

namespace bca {

struct RowWindow {
  int nvi;             // distinct column indices vali (ascending)
  int vi[14];
  unsigned mask[14];   // slots (l + 7 zone) that map to vi[a]
  int nvj;             // valid valj (ascending) with their k
  int vj[7], kk[7];
};

__device__ inline void row_window(const Mesh& m, const Zones& z, int i, int j, RowWindow& rw) {
  const int gh = m.gh, step = 2 * gh + 1;
  rw.nvi = 0;
  for (int n = 0; n < z.n; ++n)
    for (int l = 0; l < step; ++l) {
      int vali;
      if (z.n == 1) {
        if (!seed_vali(i, l, gh, m.im, z.withjn != 0, vali)) continue;
      } else {
        bool wrapped;
        seed_vali_zone(i, l, gh, m.im, n, vali, wrapped);
      }
      const unsigned bit = 1u << (l + step * n);
      int a = 0;
      while (a < rw.nvi && rw.vi[a] < vali) ++a;
      if (a < rw.nvi && rw.vi[a] == vali) {
        rw.mask[a] |= bit;
        continue;
      }
      for (int b = rw.nvi; b > a; --b) {
        rw.vi[b] = rw.vi[b - 1];
        rw.mask[b] = rw.mask[b - 1];
      }
      rw.vi[a] = vali;
      rw.mask[a] = bit;
      ++rw.nvi;
    }
  rw.nvj = 0;
  for (int k = 0; k < step; ++k) {
    const int valj = seed_valj(j, k, gh);
    if (valj >= m.jm) continue;
    int a = rw.nvj;
    while (a > 0 && rw.vj[a - 1] > valj) {
      rw.vj[a] = rw.vj[a - 1];
      rw.kk[a] = rw.kk[a - 1];
      --a;
    }
    rw.vj[a] = valj;
    rw.kk[a] = k;
    ++rw.nvj;
  }
}

// Grid over (i, j) x e; FILL = false writes the row length to cnt[row],
// FILL = true writes the row at indptr[row].  A kept column outside the
// matrix (possible only for values the reference would also have to
// reject) raises *bad.
template <bool FILL>
__global__ void __launch_bounds__(256) k_csr_rows(Mesh m, Zones z, const double* __restrict__ jac, double mini,
                                                  int* __restrict__ cnt, const int* __restrict__ indptr,
                                                  int* __restrict__ indices, double* __restrict__ data,
                                                  int* __restrict__ bad) {
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  const int e = blockIdx.z;
  if (i > m.im || j > m.jm) return;
  const int im = m.im, jm = m.jm, step = 2 * m.gh + 1;
  const long plane = (long)im * jm;
  const long base = (long)(i - 1) + (long)(j - 1) * im + (long)e * plane;
  const long ncols = 5L * im * jm;
  const int row = e + (j - 1) * 5 + (i - 1) * jm * 5;
  RowWindow rw;
  row_window(m, z, i, j, rw);
  int out = FILL ? indptr[row] : 0;
  for (int a = 0; a < rw.nvi; ++a)
    for (int b = 0; b < rw.nvj; ++b)
      for (int mm = 0; mm < 5; ++mm) {
        double sum = 0.0;
        bool any = false;
        for (int n = 0; n < z.n; ++n)
          for (int l = 0; l < step; ++l) {
            if (!(rw.mask[a] >> (l + step * n) & 1u)) continue;
            const long slot = base + (long)rw.kk[b] * plane * 5 + (long)l * plane * 5 * step +
                              (long)mm * plane * 5 * step * step + (long)n * plane * 25 * step * step;
            const double v = jac[slot];
            if (fabs(v) > mini) {
              sum = any ? sum + v : v;
              any = true;
            }
          }
        if (!any) continue;
        const long col = mm + 5L * rw.vj[b] + 5L * jm * rw.vi[a];
        if (col < 0 || col >= ncols) {
          *bad = 1;
          continue;
        }
        if (FILL) {
          indices[out] = (int)col;
          data[out] = sum;
        }
        ++out;
      }
  if (!FILL) cnt[row] = out;
}

// ------------------------------------------------------------------ scan
// Exclusive prefix sum of n ints into out[0..n] (out[n] = total), three
// passes: per-block scans of SCAN_B elements, a single-block scan of the
// block totals, then the block offsets added.
constexpr int SCAN_B = 1024;

__device__ inline int block_exclusive_scan(int v, int* sh) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int d = 1; d < SCAN_B; d <<= 1) {
    const int add = t >= d ? sh[t - d] : 0;
    __syncthreads();
    sh[t] += add;
    __syncthreads();
  }
  return sh[t] - v;
}

__global__ void __launch_bounds__(SCAN_B) k_scan_blocks(const int* __restrict__ in, long n, int* __restrict__ out,
                                                        int* __restrict__ sums) {
  __shared__ int sh[SCAN_B];
  const long g = (long)blockIdx.x * SCAN_B + threadIdx.x;
  const int v = g < n ? in[g] : 0;
  const int ex = block_exclusive_scan(v, sh);
  if (g < n) out[g] = ex;
  if (threadIdx.x == SCAN_B - 1) sums[blockIdx.x] = ex + v;
}

// one block; nb <= SCAN_B * SCAN_B / SCAN_B blocks handled in chunks
__global__ void __launch_bounds__(SCAN_B) k_scan_sums(int* __restrict__ sums, int nb, int* __restrict__ total) {
  __shared__ int sh[SCAN_B];
  int carry = 0;
  for (int c0 = 0; c0 < nb; c0 += SCAN_B) {
    const int g = c0 + threadIdx.x;
    const int v = g < nb ? sums[g] : 0;
    const int ex = block_exclusive_scan(v, sh);
    if (g < nb) sums[g] = carry + ex;
    __syncthreads();
    carry += sh[SCAN_B - 1];
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = carry;
}

__global__ void __launch_bounds__(SCAN_B) k_scan_add(int* __restrict__ out, long n, const int* __restrict__ sums,
                                                     const int* __restrict__ total) {
  const long g = (long)blockIdx.x * SCAN_B + threadIdx.x;
  if (g < n) out[g] += sums[blockIdx.x];
  if (g == 0) out[n] = *total;
}

}  // namespace bca
