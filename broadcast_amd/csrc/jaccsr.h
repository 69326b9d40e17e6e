// Colored Jacobian straight into the pruned CSR, without the COO.
//
// The values are those the reference's seed loop leaves in its COO
// (computejacobianfromjv_relaxed[_withjn][andcheck], misc/ComputeJacobian.f90)
// and the matrix is what remove_zero_jac + scipy csr_matrix make of it
// (BROADCAST_npz.py:129-135, misc/PETSc_func.py:85): per row, the slots with
// |value| > mini, columns ascending, the slots that share a column summed.
//
// The slot values are formed from the face-local blocks -- the sum
// k_assemble_colored writes to a slot, + coefdiag on the diagonal, + the
// ghost part (compact array of k_assemble_ghosts<true>) on rows near the
// edge, in that order, so every value is the COO's bit for bit -- in two
// passes over tiles of JC_TI consecutive rows in i:
//   count: marks the kept column-order indices t = (a * 7 + b) * 5 + m of each
//          row (a, b: ranks of the slot's vali / valj among the row's
//          distinct values), one ballot per 64 of them, and stores the kept
//          mask and the row length;
//   fill:  forms the same values into an LDS tile and writes each row's run
//          in CSR order (contiguous stores).
// The COO's 8 B per slot write and its two reads are gone, at the price of
// forming every value twice: on the 512 x 384 O-mesh this path is ~20 %
// slower than the COO path, and it needs none of the COO's memory.
#pragma once
#include "csr.h"
#include "jacobian.h"

namespace bca {

constexpr int JC_TI = 16;        // rows (consecutive i) per block
constexpr int JC_THREADS = 576;  // 9 waves: 16 x 7 x 5 (i, k, m) value items, rounded up to waves

// Window columns of the tile's rows, shared by all their slots: vil(i, l,
// zone) (-1: dummy), its rank among the row's distinct values, the inverse
// maps of the fill; valj(j, k) and its rank; the color row offsets dj(k).
// Built in two short phases (values, then ranks) by ~200 threads.
template <int NZ, bool FILL>
struct JcTables {
  static constexpr int NA = 7 * NZ;
  int vil[JC_TI][NA];
  signed char posl[JC_TI][NA];
  int vjk[7];
  signed char posk[7];
  int dj[7];
  signed char inv_l[FILL ? JC_TI : 1][NZ][NA];
  int col_a[FILL ? JC_TI : 1][NA];
  signed char inv_k[7];
  int col_b[7];
};

template <int NZ, bool FILL>
__device__ inline void jc_tables(const Mesh& m, const Zones& z, int i0, int j, JcTables<NZ, FILL>& T) {
  constexpr int NA = 7 * NZ;
  constexpr int GH = 3;
  const int im = m.im, jm = m.jm;
  for (int tid = threadIdx.x; tid < JC_TI * NA + 7; tid += blockDim.x) {
  if (tid < JC_TI * NA) {
    const int il = tid / NA, t = tid % NA, i = i0 + il;
    int v = -1;
    if (i <= im) {
      if (NZ == 1) {
        if (!seed_vali(i, t, GH, im, z.withjn != 0, v)) v = -1;
      } else {
        bool wrapped;
        seed_vali_zone(i, t % 7, GH, im, t / 7, v, wrapped);
      }
    }
    T.vil[il][t] = v;
    if constexpr (FILL) {
#pragma unroll
      for (int n = 0; n < NZ; ++n) T.inv_l[il][n][t] = -1;
    }
  } else if (tid < JC_TI * NA + 7) {
    const int k = tid - JC_TI * NA;
    const int v = seed_valj(j, k, GH);
    T.vjk[k] = v < jm ? v : -1;
    T.dj[k] = color_offset(k, j, jm, GH);
  }
  }
  __syncthreads();
  for (int tid = threadIdx.x; tid < JC_TI * NA + 7; tid += blockDim.x) {
  if (tid < JC_TI * NA) {
    const int il = tid / NA, t = tid % NA;
    const int v = T.vil[il][t];
    int r = 0;
#pragma unroll
    for (int u = 0; u < NA; ++u) {
      const int x = T.vil[il][u];
      bool first = x >= 0;
#pragma unroll
      for (int q = 0; q < u; ++q) first = first && T.vil[il][q] != x;
      r += (first && x < v) ? 1 : 0;
    }
    const int a = v >= 0 ? r : -1;
    T.posl[il][t] = (signed char)a;
    if constexpr (FILL) {
      if (a >= 0) {
        T.inv_l[il][t / 7][a] = (signed char)(t % 7);
        T.col_a[il][a] = 5 * jm * v;
      }
    }
  } else if (tid < JC_TI * NA + 7) {
    const int k = tid - JC_TI * NA;
    const int v = T.vjk[k];
    int r = 0;
#pragma unroll
    for (int u = 0; u < 7; ++u) r += (T.vjk[u] >= 0 && T.vjk[u] < v) ? 1 : 0;
    const int b = v >= 0 ? r : -1;
    T.posk[k] = (signed char)b;
    if constexpr (FILL) {
      if (b >= 0) {
        T.inv_k[b] = (signed char)k;
        T.col_b[b] = 5 * v;
      }
    }
  }
  }
}

// Count pass.  Block = JC_TI rows of one (j, e), 128 threads: thread (il, k)
// forms the slot values of seed color row k for the 5 variables m of row il
// (5 x 28 face-block loads in flight per thread, no LDS tile), marks the kept
// column-order indices, then one ballot per mask word and row.
constexpr int JCC_THREADS = 128;
template <int NZ>
__global__ void __launch_bounds__(JCC_THREADS) k_jac_csr_count(Mesh m, Zones z, const double* __restrict__ bi,
                                                               const double* __restrict__ bj,
                                                               const double* __restrict__ coefdiag,
                                                               const double* __restrict__ G, double mini,
                                                               unsigned long long* __restrict__ masks,
                                                               int* __restrict__ cnt, int* __restrict__ bad) {
  constexpr int STEP = 7;  // 2 gh + 1, gh = 3 (the contexts accept no other width)
  __shared__ JcTables<NZ, false> T;
  __shared__ __align__(16) unsigned char flags[JC_TI][64 * CSR_MASK_WORDS];
  const int tid = threadIdx.x;
  const int i0 = 1 + blockIdx.x * JC_TI, j = 1 + blockIdx.y, e = blockIdx.z;
  const int im = m.im, jm = m.jm;
  const long nt = 5L * im * jm;
  for (int x = tid; x < JC_TI * 64 * CSR_MASK_WORDS / 4; x += JCC_THREADS) reinterpret_cast<unsigned*>(&flags[0][0])[x] = 0u;
  jc_tables<NZ, false>(m, z, i0, j, T);
  __syncthreads();
  const int il = tid % JC_TI, k = tid / JC_TI, i = i0 + il;
  if (k < 7 && i <= im) {
    const int nfi = (im + 1) * jm, nfj = im * (jm + 1);
    const int fi0 = (i - 1) + (im + 1) * (j - 1), fj0 = (i - 1) + im * (j - 1);
    const int dj = T.dj[k];
    const bool jin = j + dj >= 1 && j + dj <= jm;
    int b0[7], b1[7], b2[7], b3[7], zc[7];
#pragma unroll
    for (int D = 0; D < 7; ++D) {
      const int di = D - 3;
      const bool in = jin && i + di >= 1 && i + di <= im;
      int o0 = ioff_index(di, dj), o1 = ioff_index(di - 1, dj);
      int o2 = joff_index(di, dj), o3 = joff_index(di, dj - 1);
      // base offsets of the four faces' blocks at q = 0; -1: no block
      b0[D] = (in && o0 >= 0 && ioff_live(m, o0, i)) ? o0 * 25 * nfi + fi0 : -1;
      b1[D] = (in && o1 >= 0 && ioff_live(m, o1, i + 1)) ? o1 * 25 * nfi + fi0 + 1 : -1;
      b2[D] = (in && o2 >= 0 && joff_live(m, o2, j)) ? o2 * 25 * nfj + fj0 : -1;
      b3[D] = (in && o3 >= 0 && joff_live(m, o3, j + 1)) ? o3 * 25 * nfj + fj0 + im : -1;
      zc[D] = (NZ > 1 && i + di > z.hi[0]) ? 1 : 0;
    }
    const int rot = (i - 1) % STEP;
    const double cd = coefdiag ? coefdiag[(i - 1) + (long)im * (j - 1)] : 0.0;
    const long tn = near_index(m, i, j);
    const int vj = T.vjk[k], b = T.posk[k];
    const int row = e + (j - 1) * 5 + (i - 1) * jm * 5;
    bool oob = false;
#pragma unroll 1
    for (int mm = 0; mm < 5; ++mm) {
      const int q = e * 5 + mm;
      double v[7];
#pragma unroll
      for (int D = 0; D < 7; ++D) {
        // the sum of k_assemble_colored, unguarded loads (a missing block
        // reads offset 0) and selects
        const double x0 = bi[(b0[D] >= 0 ? b0[D] : 0) + q * nfi], x1 = bi[(b1[D] >= 0 ? b1[D] : 0) + q * nfi];
        const double x2 = bj[(b2[D] >= 0 ? b2[D] : 0) + q * nfj], x3 = bj[(b3[D] >= 0 ? b3[D] : 0) + q * nfj];
        double acc = 0.0;
        acc = b0[D] >= 0 ? acc - x0 : acc;
        acc = b1[D] >= 0 ? acc + x1 : acc;
        acc = b2[D] >= 0 ? acc - x2 : acc;
        acc = b3[D] >= 0 ? acc + x3 : acc;
        v[D] = acc;
      }
#pragma unroll
      for (int l = 0; l < 7; ++l) {
        int r = l - rot;
        r += r < -3 ? 7 : 0;
        r -= r > 3 ? 7 : 0;
        double x = v[0];
        int xz = zc[0];
#pragma unroll
        for (int D = 1; D < 7; ++D) {
          x = (r == D - 3) ? v[D] : x;
          xz = (r == D - 3) ? zc[D] : xz;
        }
#pragma unroll
        for (int zn = 0; zn < NZ; ++zn) {
          const int vi = T.vil[il][l + 7 * zn];
          const bool ok = vi >= 0 && vj >= 0;
          const int c = ok ? vj * 5 + vi * jm * 5 + mm : 0;
          double val = (ok && xz == zn) ? x : 0.0;
          if (ok && coefdiag && row == c) val += cd;
          if (ok && tn >= 0) val += G[(tn * (NZ * 49) + zn * 49 + l * 7 + k) * 25 + q];
          if (ok && fabs(val) > mini) {
            flags[il][(T.posl[il][l + 7 * zn] * 7 + b) * 5 + mm] = 1;
            oob = oob || c >= nt;
          }
        }
      }
    }
    if (oob) *bad = 1;
  }
  __syncthreads();
  const int wave = tid / 64, lane = tid % 64;
  for (int r = wave; r < JC_TI; r += JCC_THREADS / 64) {
    const int ir = i0 + r;
    if (ir > im) break;
    int c = 0;
    const long tr = csr_tid(m, ir, j, e);
#pragma unroll
    for (int w = 0; w < CSR_MASK_WORDS; ++w) {
      const unsigned long long word = __ballot(flags[r][w * 64 + lane] != 0);
      c += __popcll(word);
      if (lane == 0) masks[w * nt + tr] = word;
    }
    if (lane == 0) cnt[e + (j - 1) * 5 + (ir - 1) * jm * 5] = c;
  }
}

// Fill pass.  Block = JC_TI rows of one j, all 5 variables e in turn:
// thread (il, mm, k) keeps the face-block offsets of its row and color in
// registers, forms the slot values of variable row e into the LDS tile, then
// each wave writes whole rows in CSR order (lane L of pass w takes
// column-order index t = 64 w + L; position indptr[row] + popcount of the
// mask below t).
template <int NZ>
__global__ void __launch_bounds__(JC_THREADS) k_jac_csr_fill(Mesh m, Zones z, const double* __restrict__ bi,
                                                             const double* __restrict__ bj,
                                                             const double* __restrict__ coefdiag,
                                                             const double* __restrict__ G, double mini,
                                                             const unsigned long long* __restrict__ masks,
                                                             const int* __restrict__ indptr,
                                                             int* __restrict__ indices, double* __restrict__ data) {
  constexpr int NS = 245 * NZ;
  constexpr int PITCH = JC_TI + 1;
  constexpr int NWAVE = JC_THREADS / 64;
  constexpr int STEP = 7;  // 2 gh + 1, gh = 3 (the contexts accept no other width)
  __shared__ double tile[NS * PITCH];
  __shared__ JcTables<NZ, true> T;
  __shared__ unsigned long long mk_s[CSR_MASK_WORDS][JC_TI];
  __shared__ int start_s[JC_TI];
  const int tid = threadIdx.x;
  const int i0 = 1 + blockIdx.x * JC_TI, j = 1 + blockIdx.y;
  const int im = m.im, jm = m.jm;
  const long nt = 5L * im * jm;
  const int wave = tid / 64, lane = tid % 64;
  jc_tables<NZ, true>(m, z, i0, j, T);
  __syncthreads();
  // item (row il, variable mm, seed color row k); its face-block offsets, the
  // column cells' zones and the row constants serve the 5 variables e
  const bool item = tid < JC_TI * 35;
  const int il = tid % JC_TI, rest = tid / JC_TI, mm = rest % 5, k = item ? rest / 5 : 0;
  const int i = i0 + il;
  const bool live = item && i <= im;
  const int nfi = (im + 1) * jm, nfj = im * (jm + 1);
  // per column offset D the offset indices o + 1 (0: no block) of the four
  // faces' blocks, 5 bits each, and the column cell's zone in bit 20
  int code[7];
  double cd = 0.0;
  long tn = -1;
  int rot = 0, vj = -1;
  const int fi0 = (i - 1) + (im + 1) * (j - 1), fj0 = (i - 1) + im * (j - 1);
  if (live) {
    const int dj = T.dj[k];
    const bool jin = j + dj >= 1 && j + dj <= jm;
#pragma unroll
    for (int D = 0; D < 7; ++D) {
      const int di = D - 3;
      const bool in = jin && i + di >= 1 && i + di <= im;
      int o0 = ioff_index(di, dj), o1 = ioff_index(di - 1, dj);
      int o2 = joff_index(di, dj), o3 = joff_index(di, dj - 1);
      o0 = (in && o0 >= 0 && ioff_live(m, o0, i)) ? o0 + 1 : 0;
      o1 = (in && o1 >= 0 && ioff_live(m, o1, i + 1)) ? o1 + 1 : 0;
      o2 = (in && o2 >= 0 && joff_live(m, o2, j)) ? o2 + 1 : 0;
      o3 = (in && o3 >= 0 && joff_live(m, o3, j + 1)) ? o3 + 1 : 0;
      code[D] = o0 | (o1 << 5) | (o2 << 10) | (o3 << 15) | ((NZ > 1 && i + di > z.hi[0]) ? 1 << 20 : 0);
    }
    rot = (i - 1) % STEP;
    cd = coefdiag ? coefdiag[(i - 1) + (long)im * (j - 1)] : 0.0;
    tn = near_index(m, i, j);
    vj = T.vjk[k];
  }
  for (int e = 0; e < 5; ++e) {
    if (tid >= JC_TI * 35 && tid < JC_TI * 35 + JC_TI) {
      const int r = tid - JC_TI * 35;
      if (i0 + r <= im) {
        start_s[r] = indptr[e + (j - 1) * 5 + (i0 + r - 1) * jm * 5];
#pragma unroll
        for (int w = 0; w < CSR_MASK_WORDS; ++w) mk_s[w][r] = masks[w * nt + csr_tid(m, i0 + r, j, e)];
      }
    }
    if (live) {
      const int q = e * 5 + mm;
      double v[7];
#pragma unroll
      for (int D = 0; D < 7; ++D) {
        // jac = -dR/dw with R(i,j) = hi(i) - hi(i+1) + hj(j) - hj(j+1), the
        // order of k_assemble_colored; unguarded loads (a missing block reads
        // offset 0) keep all 28 in flight, the selects keep the sums exact
        const int o0 = code[D] & 31, o1 = (code[D] >> 5) & 31, o2 = (code[D] >> 10) & 31, o3 = (code[D] >> 15) & 31;
        const double x0 = bi[(o0 ? (o0 - 1) * 25 * nfi + fi0 : 0) + q * nfi];
        const double x1 = bi[(o1 ? (o1 - 1) * 25 * nfi + fi0 + 1 : 0) + q * nfi];
        const double x2 = bj[(o2 ? (o2 - 1) * 25 * nfj + fj0 : 0) + q * nfj];
        const double x3 = bj[(o3 ? (o3 - 1) * 25 * nfj + fj0 + im : 0) + q * nfj];
        double acc = 0.0;
        acc = o0 ? acc - x0 : acc;
        acc = o1 ? acc + x1 : acc;
        acc = o2 ? acc - x2 : acc;
        acc = o3 ? acc + x3 : acc;
        v[D] = acc;
      }
      const int row = e + (j - 1) * 5 + (i - 1) * jm * 5;
#pragma unroll
      for (int l = 0; l < 7; ++l) {
        int r = l - rot;
        r += r < -3 ? 7 : 0;
        r -= r > 3 ? 7 : 0;
        double x = v[0];
        int xz = code[0] >> 20;
#pragma unroll
        for (int D = 1; D < 7; ++D) {
          x = (r == D - 3) ? v[D] : x;
          xz = (r == D - 3) ? code[D] >> 20 : xz;
        }
#pragma unroll
        for (int zn = 0; zn < NZ; ++zn) {
          const int vi = T.vil[il][l + 7 * zn];
          const bool ok = vi >= 0 && vj >= 0;
          const int c = ok ? vj * 5 + vi * jm * 5 + mm : 0;
          double val = (ok && xz == zn) ? x : 0.0;
          if (ok && coefdiag && row == c) val += cd;
          if (ok && tn >= 0) val += G[(tn * (NZ * 49) + zn * 49 + l * 7 + k) * 25 + q];
          tile[(k + 7 * l + 49 * mm + 245 * zn) * PITCH + il] = val;
        }
      }
    }
    __syncthreads();
    // each row's run in CSR order
    for (int r = wave; r < JC_TI; r += NWAVE) {
      if (i0 + r > im) break;
      int pos = start_s[r];
#pragma unroll
      for (int w = 0; w < CSR_MASK_WORDS; ++w) {
        const unsigned long long mk = mk_s[w][r];
        if (mk == 0) continue;
        if ((mk >> lane) & 1ull) {
          const int t = w * 64 + lane;
          const int a = t / 35, b = (t / 5) % 7, m5 = t % 5;
          const int kk = T.inv_k[b];
          double val = 0.0;
#pragma unroll
          for (int n = 0; n < NZ; ++n) {
            const int l = T.inv_l[r][n][a];
            if (l >= 0) {
              const double y = tile[(kk + 7 * l + 49 * m5 + 245 * n) * PITCH + r];
              if (fabs(y) > mini) val += y;
            }
          }
          const int p = pos + __popcll(mk & ((1ull << lane) - 1ull));
          indices[p] = m5 + T.col_b[b] + T.col_a[r][a];
          data[p] = val;
        }
        pos += __popcll(mk);
      }
    }
    __syncthreads();
  }
}

}  // namespace bca
