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
// Two passes over the values: count (row lengths and kept masks), exclusive
// scan into indptr, fill.  Reads are coalesced (a wave holds consecutive i of
// one row kind and reads the same slot kind); each lane writes its row's run.
#pragma once
#include "jacobian.h"

namespace bca {

// The CSR passes read the COO in slot order -- (k, l, zone, m) the same for
// every lane, so each load is a coalesced row of consecutive i -- and place
// each kept slot at its column-order index t = (a * 7 + b) * 5 + m, where a
// is the rank of the slot's vali among the row's distinct values and b that
// of its valj.  Count: the row's kept set as a bitmask over t (8 words, kept
// for the fill) and its popcount.  Fill: the slot's CSR position is
// indptr[row] + popcount of the mask below t; a second slot on the same
// column (the other zone) is added to the first.
constexpr int CSR_MASK_WORDS = 8;  // 7 * 2 * 7 * 5 = 490 column-order bits

template <int NZ>
struct RowOrder {
  int posl[7 * NZ];  // rank a of the vali of slot (l, zone) among distinct values, -1: dummy
  int vil[7 * NZ];
  int posk[7];       // rank b of valj(k), -1: dummy
  int vjk[7];
};

template <int NZ>
__device__ inline void row_order(const Mesh& m, int withjn, int i, int j, RowOrder<NZ>& o) {
  constexpr int NI = 7 * NZ;
  const int gh = m.gh, im = m.im, jm = m.jm;
#pragma unroll
  for (int t = 0; t < NI; ++t) {
    int v;
    bool ok = true;
    if (NZ == 1) {
      ok = seed_vali(i, t % 7, gh, im, withjn != 0, v);
    } else {
      bool wrapped;
      seed_vali_zone(i, t % 7, gh, im, t / 7, v, wrapped);
    }
    o.vil[t] = ok ? v : -1;
  }
  // rank among distinct valid values: count of valid values strictly smaller
  // counted once each (a value shared by two slots counts once)
#pragma unroll
  for (int t = 0; t < NI; ++t) {
    int r = 0;
#pragma unroll
    for (int u = 0; u < NI; ++u) {
      bool first = o.vil[u] >= 0;
#pragma unroll
      for (int w = 0; w < u; ++w) first = first && o.vil[w] != o.vil[u];
      r += (first && o.vil[u] < o.vil[t]) ? 1 : 0;
    }
    o.posl[t] = o.vil[t] >= 0 ? r : -1;
  }
#pragma unroll
  for (int k = 0; k < 7; ++k) {
    const int v = seed_valj(j, k, gh);
    o.vjk[k] = v < jm ? v : -1;
  }
#pragma unroll
  for (int k = 0; k < 7; ++k) {
    int r = 0;
#pragma unroll
    for (int u = 0; u < 7; ++u) r += (o.vjk[u] >= 0 && o.vjk[u] < o.vjk[k]) ? 1 : 0;
    o.posk[k] = o.vjk[k] >= 0 ? r : -1;
  }
}

__device__ inline void mask_set(unsigned long long (&mk)[CSR_MASK_WORDS], int t) {
#pragma unroll
  for (int w = 0; w < CSR_MASK_WORDS; ++w)
    if (w == (t >> 6)) mk[w] |= 1ull << (t & 63);
}
__device__ inline bool mask_test(const unsigned long long (&mk)[CSR_MASK_WORDS], int t) {
  bool r = false;
#pragma unroll
  for (int w = 0; w < CSR_MASK_WORDS; ++w)
    if (w == (t >> 6)) r = (mk[w] >> (t & 63)) & 1ull;
  return r;
}
__device__ inline int mask_rank(const unsigned long long (&mk)[CSR_MASK_WORDS], int t) {
  int r = 0;
#pragma unroll
  for (int w = 0; w < CSR_MASK_WORDS; ++w) {
    const int lo = w * 64;
    if (t >= lo + 64) r += __popcll(mk[w]);
    else if (t > lo) r += __popcll(mk[w] & ((1ull << (t - lo)) - 1ull));
  }
  return r;
}

// Thread index of row (i, j, e) in the mask array: the launch order.
__device__ inline long csr_tid(const Mesh& m, int i, int j, int e) {
  return (long)(i - 1) + (long)m.im * ((j - 1) + (long)m.jm * e);
}

// Visits every non-dummy slot of row (i, j, e) in slot order (zone, l, k, m
// the same for every lane, so each load is coalesced).
template <int NZ, class A, class K>
__device__ inline void row_scan_slots(const Mesh& m, const RowOrder<NZ>& o, const A& jac, int i, int j, int e,
                                      K&& slot) {
  const int im = m.im, jm = m.jm;
  const long plane = (long)im * jm;
  const long base = (long)(i - 1) + (long)(j - 1) * im + (long)e * plane;
#pragma unroll
  for (int n = 0; n < NZ; ++n)
#pragma unroll
    for (int l = 0; l < 7; ++l) {
      const int a = o.posl[l + 7 * n];
#pragma unroll
      for (int k = 0; k < 7; ++k) {
        const int b = o.posk[k];
#pragma unroll
        for (int mm = 0; mm < 5; ++mm) {
          const long s = base + (long)k * plane * 5 + (long)l * plane * 35 + (long)mm * plane * 245 +
                         (long)n * plane * 1225;
          const double v = jac[s];
          if (a < 0 || b < 0) continue;
          slot((a * 7 + b) * 5 + mm, mm + 5L * o.vjk[k] + 5L * jm * o.vil[l + 7 * n], v, s);
        }
      }
    }
}

// Count: kept mask (column order) of every row into masks[tid], length into
// cnt[row].
template <int NZ, class A>
__global__ void __launch_bounds__(256) k_csr_count(Mesh m, int withjn, A jac, double mini, int* __restrict__ cnt,
                                                   unsigned long long* __restrict__ masks, int* __restrict__ bad) {
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  const int e = blockIdx.z;
  if (i > m.im || j > m.jm) return;
  RowOrder<NZ> o;
  row_order<NZ>(m, withjn, i, j, o);
  unsigned long long mk[CSR_MASK_WORDS];
#pragma unroll
  for (int w = 0; w < CSR_MASK_WORDS; ++w) mk[w] = 0;
  const long ncols = 5L * m.im * m.jm;
  bool oob = false;
  row_scan_slots<NZ>(m, o, jac, i, j, e, [&](int t, long col, double v, long) {
    if (fabs(v) > mini) {
      mask_set(mk, t);
      oob = oob || col >= ncols;
    }
  });
  if (oob) *bad = 1;
  int c = 0;
  const long tid = csr_tid(m, i, j, e), nt = 5L * m.im * m.jm;
#pragma unroll
  for (int w = 0; w < CSR_MASK_WORDS; ++w) {
    c += __popcll(mk[w]);
    masks[w * nt + tid] = mk[w];
  }
  cnt[e + (j - 1) * 5 + (i - 1) * m.jm * 5] = c;
}

// Fill.  A block owns FILL_TI consecutive i of one (j, e): it first copies
// every slot value of those rows into LDS (a 128-byte run of consecutive i
// per slot plane, so the COO is read once, in whole lines), then each wave
// writes whole rows in CSR order: lane L of pass w takes column-order index
// t = 64 w + L, tests the kept mask and stores at indptr[row] + the popcount
// of the mask below t, so a pass writes one contiguous run of the row.  The
// value of column t sums the kept slots on it (at most one per zone: within
// a zone the 7 window columns of a row are distinct mod 7), zone 0 first --
// the order scipy's sum of duplicates sees them in the COO.
constexpr int FILL_TI = 16;        // rows (consecutive i) per block
constexpr int FILL_THREADS = 512;  // 8 waves, FILL_TI / 8 rows each

template <int NZ, bool JB = false>
__global__ void __launch_bounds__(FILL_THREADS) k_csr_fill(Mesh m, int withjn, const double* __restrict__ jac,
                                                          double mini, const unsigned long long* __restrict__ masks,
                                                          const int* __restrict__ indptr, int* __restrict__ indices,
                                                          double* __restrict__ data,
                                                          const double* __restrict__ jdz = nullptr,
                                                          const double* __restrict__ jdz2 = nullptr,
                                                          const double* __restrict__ vol = nullptr, long nnz = 0) {
  constexpr int NS = 245 * NZ;          // slots per row: (k, l, m, zone)
  constexpr int NA = 7 * NZ;            // distinct window columns in i, at most
  constexpr int PITCH = FILL_TI + 1;    // odd pitch: lanes reading different slots spread over the banks
  __shared__ double tile[NS * PITCH];
  __shared__ signed char inv_l[FILL_TI][NZ][NA];  // slot l of zone n on column rank a, -1: none
  __shared__ int col_a[FILL_TI][NA];              // 5 jm vali of column rank a
  __shared__ signed char inv_k[7];                // slot k on column rank b
  __shared__ int col_b[7];                        // 5 valj of column rank b
  __shared__ unsigned long long mk_s[CSR_MASK_WORDS][FILL_TI];
  __shared__ int start_s[FILL_TI];
  const int tid = threadIdx.x;
  const int i0 = 1 + blockIdx.x * FILL_TI, j = 1 + blockIdx.y, e = blockIdx.z;
  const int im = m.im, jm = m.jm;
  const long plane = (long)im * jm;
  const long nt = 5L * im * jm;
#pragma unroll 1
  for (int c = 0; c < (JB ? 3 : 1); ++c) {
  {
    // slot values: thread (s0, il) loads planes s0, s0 + 32, ...
    const int il = tid % FILL_TI, s0 = tid / FILL_TI;
    const bool in = i0 + il <= im;
    const double* arr = c == 0 ? jac : c == 1 ? jdz : jdz2;
    const double* src = arr + (i0 - 1 + il) + (long)(j - 1) * im + (long)e * plane;
    constexpr int STEP = FILL_THREADS / FILL_TI;
    constexpr int NL = (NS + STEP - 1) / STEP;
    double v[NL];
#pragma unroll
    for (int q = 0; q < NL; ++q) {
      const int s = s0 + q * STEP;
      v[q] = (in && s < NS) ? src[5L * s * plane] : 0.0;
    }
    if (c == 0 && tid < FILL_TI) {
      const int i = i0 + tid;
      if (i <= im) {
        RowOrder<NZ> o;
        row_order<NZ>(m, withjn, i, j, o);
#pragma unroll
        for (int n = 0; n < NZ; ++n)
#pragma unroll
          for (int a = 0; a < NA; ++a) inv_l[tid][n][a] = -1;
#pragma unroll
        for (int t = 0; t < NA; ++t)
          if (o.posl[t] >= 0) {
            inv_l[tid][t / 7][o.posl[t]] = (signed char)(t % 7);
            col_a[tid][o.posl[t]] = 5 * jm * o.vil[t];
          }
        if (tid == 0) {
#pragma unroll
          for (int k = 0; k < 7; ++k)
            if (o.posk[k] >= 0) {
              inv_k[o.posk[k]] = (signed char)k;
              col_b[o.posk[k]] = 5 * o.vjk[k];
            }
        }
        start_s[tid] = indptr[e + (j - 1) * 5 + (i - 1) * jm * 5];
      }
    }
    if (c == 0 && tid < CSR_MASK_WORDS * FILL_TI) {
      const int w = tid / FILL_TI, r = tid % FILL_TI;
      if (i0 + r <= im) mk_s[w][r] = masks[w * nt + csr_tid(m, i0 + r, j, e)];
    }
#pragma unroll
    for (int q = 0; q < NL; ++q) {
      const int s = s0 + q * STEP;
      if (s < NS) tile[s * PITCH + il] = v[q];
    }
  }
  __syncthreads();
  const int wave = tid / 64, lane = tid % 64;
  for (int r = wave; r < FILL_TI; r += FILL_THREADS / 64) {
    if (i0 + r > im) break;
    int pos = start_s[r];
    const double v0 = JB ? vol[cidx(m, i0 + r, j)] : 1.0;
#pragma unroll
    for (int w = 0; w < CSR_MASK_WORDS; ++w) {
      const unsigned long long mk = mk_s[w][r];
      if (mk == 0) continue;
      if ((mk >> lane) & 1ull) {
        const int t = w * 64 + lane;
        const int a = t / 35, b = (t / 5) % 7, mm = t % 5;
        const int k = inv_k[b];
        double val = 0.0;
#pragma unroll
        for (int n = 0; n < NZ; ++n) {
          const int l = inv_l[r][n][a];
          if (l >= 0) {
            const double x = tile[(k + 7 * l + 49 * mm + 245 * n) * PITCH + r];
            if (fabs(x) > mini) val += x;
          }
        }
        const int p = pos + __popcll(mk & ((1ull << lane) - 1ull));
        if (c == 0) indices[p] = mm + col_b[b] + col_a[r][a];
        if constexpr (JB) data[c * nnz + p] = c == 0 ? (val != 0.0 ? val / v0 : 0.0) : val;
        else data[p] = val;
      }
      pos += __popcll(mk);
    }
  }
  // the next array's tile overwrites this one
  if constexpr (JB) __syncthreads();
  }
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
