// Colored Jacobian straight into the pruned CSR, without the COO.
//
// The values are those the reference's seed loop leaves in its COO
// (computejacobianfromjv_relaxed[_withjn][andcheck], misc/ComputeJacobian.f90)
// and the matrix is what remove_zero_jac + scipy csr_matrix make of it
// (BROADCAST_npz.py:129-135, misc/PETSc_func.py:85): per row, the slots with
// |value| > mini, columns ascending, the slots that share a column summed.
// Every value is the COO's bit for bit: the sum k_assemble_colored writes to
// a slot (0 - left i-face + right i-face - lower j-face + upper j-face block
// entries), + coefdiag on the diagonal, + the ghost part (compact array of
// k_ghost_part) on rows near the edge, in that order.
//
// k_jac_csr_rows works on tiles of JC_TI consecutive i of one j: the tile's
// Jacobian rows of one variable e are summed in LDS from the face blocks
// (each block entry read once per row cell that uses it, coalesced along
// i), then each wave turns whole rows into their CSR runs.  Most rows are
// plain (every zone's window is the 7 x 7 neighbourhood of the row cell and
// no ghost part is added): their CSR row is the 29 stencil cells in a fixed
// column order, handled with a lane map shared by the tile.  The other rows
// (near the mesh edge or a zone's wrapped columns) go through the general
// slot logic in a second launch over the edge tiles.
#pragma once
#include "csr.h"
#include "jacobian.h"

namespace bca {

constexpr int JC_TI = 16;  // rows (consecutive i) per tile

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

// ---------------------------------------------------------------- row pass
// One kernel for the three uses of the pruned CSR without the COO; each slot
// value is formed once per launch:
//
//   for each variable row e of a tile of JC_TI consecutive i (one j):
//     J(il, k, D, m) = d(-R)/dw of the tile's row cells over their 7 x 7
//     windows, each entry summed in registers from its (at most four) face
//     block entries in the order of the COO path, the e-independent block
//     offsets kept in registers from one e to the next, stored in LDS;
//     then each wave takes whole rows and forms the kept masks by ballot
//     (plain rows: 145 stencil candidates in three words; the others: the
//     column-order index t = (a 7 + b) 5 + m of every slot in eight words),
//     and
//   JR_COUNT    stores the kept masks and row lengths;
//   JR_FILL     writes indices and data of each kept entry at indptr[row] +
//               popcount of the mask below it;
//   JR_REFRESH  compares the masks with the stored ones: a row whose kept set
//               is unchanged gets its data rewritten in place (its indptr and
//               indices are those of the earlier call); a changed row stores
//               its new mask and length and raises *changed (the caller then
//               rescans and refills).
// The mask format of a row (three words in candidate order for a plain row,
// eight in column order otherwise) depends only on the mesh and the zones,
// so all three modes agree on it.
enum { JR_COUNT = 0, JR_FILL = 1, JR_REFRESH = 2 };
constexpr int JR_THREADS = 512;       // plain launch
constexpr int JR_EDGE_THREADS = 256;  // edge launch
constexpr int JR_PITCH = JC_TI + 1;
// The cells (di, dj) relative to a row cell that its four faces' stencils
// reach: the 29-cell stencil of the interior plus (0, 4), which the row j = 1
// reaches through the off-centred Euler flux of the face j = 2.
constexpr int JR_NCELL = 30;
__host__ __device__ constexpr int jr_cell_di(int c) {
  // rows of the table: di = -3 .. 3 with the dj values each reaches
  return c < 1 ? -3 : c < 6 ? -2 : c < 11 ? -1 : c < 19 ? 0 : c < 24 ? 1 : c < 29 ? 2 : 3;
}
__host__ __device__ constexpr int jr_cell_dj(int c) {
  return c < 1 ? 0 : c < 6 ? c - 3 : c < 11 ? c - 8 : c < 19 ? c - 14 : c < 24 ? c - 21 : c < 29 ? c - 26 : 0;
}
constexpr int JR_NELEM = JC_TI * JR_NCELL * 5;                   // (il, m, cell) value items of one e

// The row phase's loads of a wave's rows -- the stored mask words and the
// row start of each variable row, coefdiag of each row cell -- issued once
// before the variable loop, one item per lane (items past 64 in a second
// register), and read back with readlane: the row phase then waits on no
// global load.  Item q: row slot qr = q / PER_ROW, then per variable e the
// W mask words (refresh) and indptr (fill, refresh), last coefdiag.
__device__ inline unsigned long long readlane64(unsigned long long x, int l) {
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)x, l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(x >> 32), l);
  return ((unsigned long long)hi << 32) | lo;
}
template <int MODE, int W, int RPW>
struct RowPrefetch {
  static constexpr int NMASK = MODE == JR_REFRESH ? W : 0;
  static constexpr int PER_E = NMASK + (MODE != JR_COUNT ? 1 : 0);
  static constexpr int PER_ROW = 5 * PER_E + 1;
  static constexpr int NREG = (RPW * PER_ROW + 63) / 64;
  unsigned long long v[NREG];
  __device__ inline void load(const Mesh& m, int i0, int j, int wave, int nwave,
                              const unsigned long long* masks, const int* indptr, const double* coefdiag) {
    const int lane = threadIdx.x % 64;
    const long nt = 5L * m.im * m.jm;
#pragma unroll
    for (int g = 0; g < NREG; ++g) {
      const int q = g * 64 + lane;
      const int qr = q / PER_ROW, x = q % PER_ROW;
      const int i = i0 + wave + nwave * qr;
      unsigned long long val = 0ull;
      if (qr < RPW && i <= m.im) {
        if (x == PER_ROW - 1) {
          if (coefdiag) val = (unsigned long long)__double_as_longlong(coefdiag[(i - 1) + (long)m.im * (j - 1)]);
        } else {
          constexpr int DIV = PER_E > 0 ? PER_E : 1;  // PER_E = 0 (count) leaves only coefdiag
          const int e = x / DIV, w = x % DIV;
          if (w < NMASK)
            val = masks[w * nt + csr_tid(m, i, j, e)];
          else
            val = (unsigned long long)(unsigned)indptr[e + (j - 1) * 5 + (i - 1) * m.jm * 5];
        }
      }
      v[g] = val;
    }
  }
  __device__ inline unsigned long long get(int q) const {
    unsigned long long x = v[0];
#pragma unroll
    for (int g = 1; g < NREG; ++g) x = q >= g * 64 ? v[g] : x;
    return readlane64(x, q % 64);
  }
  __device__ inline unsigned long long mask(int qr, int e, int w) const { return get(qr * PER_ROW + e * PER_E + w); }
  __device__ inline int start(int qr, int e) const { return (int)get(qr * PER_ROW + e * PER_E + NMASK); }
  __device__ inline double cd(int qr) const { return __longlong_as_double((long long)get(qr * PER_ROW + PER_ROW - 1)); }
};

// A row whose window is not the plain 7 x 7 neighbourhood (near the mesh
// edge, where the ghost part is added from G when tn >= 0, or near a zone's
// wrapped columns), in the edge launch: every slot from the tables, all
// mask words in flight at once (this launch has a register budget of its
// own).
template <int NZ, int MODE, class PF>
__device__ inline void jr_row_edge(const double* J, const JcTables<NZ, true>& T, const Zones& z, int r, int i,
                                   int jm, int row, long tr, long nt, int e, const PF& pf, int qr, bool has_cd,
                                   const double* G, long tn, double mini, unsigned long long* masks, int* cnt,
                                   int* indices, double* data, int* changed, int* bad) {
  constexpr int NA = 7 * NZ;
  const int lane = threadIdx.x % 64;
  const int rot = (i - 1) % 7;
  double v[CSR_MASK_WORDS];
  unsigned long long mk[CSR_MASK_WORDS];
  bool oob = false;
  const double cd = pf.cd(qr);
  // the lane's slots (word w, zone n): window indices first, then every
  // ghost-part load issued at once (from entry 0 where a slot has none, the
  // branch on tn being row-uniform), so that they are in flight together
  int kks[CSR_MASK_WORDS], ls[CSR_MASK_WORDS][NZ];
  double gv[CSR_MASK_WORDS][NZ];
#pragma unroll
  for (int w = 0; w < CSR_MASK_WORDS; ++w) {
    const int t = w * 64 + lane;
    const int a = t / 35, b = (t / 5) % 7;
    kks[w] = t < NA * 35 ? (int)T.inv_k[b] : -1;
#pragma unroll
    for (int n = 0; n < NZ; ++n) {
      ls[w][n] = kks[w] >= 0 ? (int)T.inv_l[r][n][a] : -1;
      gv[w][n] = 0.0;
    }
  }
  if (tn >= 0) {
#pragma unroll
    for (int w = 0; w < CSR_MASK_WORDS; ++w) {
      const int m5 = (w * 64 + lane) % 5;
#pragma unroll
      for (int n = 0; n < NZ; ++n) {
        const int l = ls[w][n];
        const long gi = l >= 0 ? (tn * (NZ * 49) + n * 49 + l * 7 + kks[w]) * 25 + e * 5 + m5 : 0;
        gv[w][n] = G[gi];
      }
    }
  }
#pragma unroll
  for (int w = 0; w < CSR_MASK_WORDS; ++w) {
    const int m5 = (w * 64 + lane) % 5;
    const int kk = kks[w];
    double acc = 0.0;
    bool keep = false;
    if (kk >= 0) {
#pragma unroll
      for (int n = 0; n < NZ; ++n) {
        const int l = ls[w][n];
        if (l < 0) continue;
        int dl = l - rot;
        dl += dl < -3 ? 7 : 0;
        dl -= dl > 3 ? 7 : 0;
        const int xz = (NZ > 1 && i + dl > z.hi[0]) ? 1 : 0;
        const int c = T.vjk[kk] * 5 + T.vil[r][l + 7 * n] * jm * 5 + m5;
        double y = (xz == n) ? J[((kk * 7 + dl + 3) * 5 + m5) * JR_PITCH + r] : 0.0;
        if (has_cd && row == c) y += cd;
        if (tn >= 0) y += gv[w][n];
        if (fabs(y) > mini) {
          acc += y;
          keep = true;
          oob = oob || c >= nt;
        }
      }
    }
    v[w] = acc;
    mk[w] = __ballot(keep);
  }
  if (oob) *bad = 1;
  bool same = true;
  if constexpr (MODE == JR_REFRESH) {
#pragma unroll
    for (int w = 0; w < CSR_MASK_WORDS; ++w) same = (pf.mask(qr, e, w) == mk[w]) & same;
  }
  if (MODE == JR_FILL || (MODE == JR_REFRESH && same)) {
    int pos = pf.start(qr, e);
#pragma unroll
    for (int w = 0; w < CSR_MASK_WORDS; ++w) {
      if ((mk[w] >> lane) & 1ull) {
        const int p = pos + __popcll(mk[w] & ((1ull << lane) - 1ull));
        if constexpr (MODE == JR_FILL) {
          const int t = w * 64 + lane;
          indices[p] = t % 5 + T.col_b[(t / 5) % 7] + T.col_a[r][t / 35];
        }
        data[p] = v[w];
      }
      pos += __popcll(mk[w]);
    }
  } else {
    int count = 0;
#pragma unroll
    for (int w = 0; w < CSR_MASK_WORDS; ++w) {
      count += __popcll(mk[w]);
      if (lane == w) masks[w * nt + tr] = mk[w];
    }
    if (lane == 0) {
      cnt[row] = count;
      if (MODE == JR_REFRESH) *changed = 1;
    }
  }
}

// A plain row: in every zone the window is the 7 x 7 neighbourhood of the
// row cell, and no ghost part is added.  Its CSR row is then the 29 stencil
// cells in column order -- (di, dj) ascending, i.e. the JR cell table
// without (0, 4) -- times the 5 variables: candidate c = 5 s + m of 145,
// three mask words in candidate order.  A column's value is J from the zone
// that owns its cell (the other zones' slots there are zero) and, on the
// diagonal, coefdiag from every zone's slot: (J + cd) + cd with two zones,
// the sums of the general path bit for bit (the kept terms are added in
// zone order; with two kept terms the sum does not depend on the order).
constexpr int JR_PLAIN_WORDS = 3;
// Lane map of a plain row, the same for every plain row of one j: per mask
// word the LDS offset of the candidate's J value (without the row) or -1,
// and whether it is the diagonal cell (di = dj = 0) with its variable m.
struct PlainLane {
  int off[JR_PLAIN_WORDS];
  int diagm[JR_PLAIN_WORDS];  // m of the diagonal cell's candidate, -1 otherwise
};
__device__ inline PlainLane plain_lane(int j) {
  PlainLane pl;
  const int lane = threadIdx.x % 64;
#pragma unroll
  for (int w = 0; w < JR_PLAIN_WORDS; ++w) {
    const int c = w * 64 + lane;
    const int s0 = c / 5, m5 = c % 5;
    const int s = s0 < 18 ? s0 : s0 + 1;  // skip (0, 4)
    const int di = jr_cell_di(s), dj = jr_cell_dj(s);
    const int k = (dj + j - 1 + 7) % 7;
    pl.off[w] = c < 145 ? ((k * 7 + di + 3) * 5 + m5) * JR_PITCH : -1;
    pl.diagm[w] = (c < 145 && di == 0 && dj == 0) ? m5 : -1;
  }
  return pl;
}

template <int NZ, int MODE, class PF>
__device__ inline void jr_row_plain(const PlainLane& pl, const double* J, int r, int i, int j, int jm, int row,
                                    long tr, long nt, int e, const PF& pf, int qr, bool has_cd, double mini,
                                    unsigned long long* masks, int* cnt, int* indices, double* data,
                                    int* changed) {
  const int lane = threadIdx.x % 64;
  double v[JR_PLAIN_WORDS];
  unsigned long long mk[JR_PLAIN_WORDS];
  const double cd = pf.cd(qr);
#pragma unroll
  for (int w = 0; w < JR_PLAIN_WORDS; ++w) {
    double acc = 0.0;
    bool keep = false;
    if (pl.off[w] >= 0) {
      double y = J[pl.off[w] + r];
      if (has_cd && pl.diagm[w] == e) {
        y += cd;
        if (fabs(y) > mini) {
          acc = y;
          keep = true;
        }
        if (NZ > 1 && fabs(cd) > mini) {
          acc += cd;
          keep = true;
        }
      } else if (fabs(y) > mini) {
        acc = y;
        keep = true;
      }
    }
    v[w] = acc;
    mk[w] = __ballot(keep);
  }
  bool same = true;
  if constexpr (MODE == JR_REFRESH) {
#pragma unroll
    for (int w = 0; w < JR_PLAIN_WORDS; ++w) same = (pf.mask(qr, e, w) == mk[w]) & same;
  }
  if (MODE == JR_FILL || (MODE == JR_REFRESH && same)) {
    int pos = pf.start(qr, e);
#pragma unroll
    for (int w = 0; w < JR_PLAIN_WORDS; ++w) {
      if ((mk[w] >> lane) & 1ull) {
        const int p = pos + __popcll(mk[w] & ((1ull << lane) - 1ull));
        if constexpr (MODE == JR_FILL) {
          const int c = w * 64 + lane;
          const int s0 = c / 5, s = s0 < 18 ? s0 : s0 + 1;
          indices[p] = c % 5 + (j - 1 + jr_cell_dj(s)) * 5 + (i - 1 + jr_cell_di(s)) * jm * 5;
        }
        data[p] = v[w];
      }
      pos += __popcll(mk[w]);
    }
  } else {
    int count = 0;
#pragma unroll
    for (int w = 0; w < CSR_MASK_WORDS; ++w) {
      const unsigned long long word = w < JR_PLAIN_WORDS ? mk[w < JR_PLAIN_WORDS ? w : 0] : 0ull;
      count += __popcll(word);
      if (lane == w) masks[w * nt + tr] = word;
    }
    if (lane == 0) {
      cnt[row] = count;
      if (MODE == JR_REFRESH) *changed = 1;
    }
  }
}

// The edge launch: tiles that may hold rows which are not plain (the x
// tiles of i <= 2 JC_TI and of the last two tiles, for every j, and the
// other x tiles of the NEAR rows at each j end).  t -> (tile x, j).
__host__ __device__ inline int jr_edge_nx(int nxt) { return nxt < 4 ? nxt : 4; }
__host__ __device__ inline int jr_edge_nj(int jm) { return jm <= 2 * NEAR ? jm : 2 * NEAR; }
__host__ __device__ inline long jr_edge_tiles(int nxt, int jm) {
  return (long)jr_edge_nx(nxt) * jm + (long)(nxt - jr_edge_nx(nxt)) * jr_edge_nj(jm);
}
__host__ __device__ inline bool jr_edge_x(int x, int nxt) { return x < 2 || x >= nxt - 2; }
__host__ __device__ inline bool jr_edge_j(int j, int jm) { return j <= NEAR || j > jm - NEAR; }
__device__ inline void jr_edge_tile(long t, int nxt, int jm, int& x, int& j) {
  const int nxe = jr_edge_nx(nxt);
  if (t < (long)nxe * jm) {
    const int k = (int)(t % nxe);
    x = (nxt <= 4 || k < 2) ? k : nxt - 4 + k;
    j = 1 + (int)(t / nxe);
    return;
  }
  t -= (long)nxe * jm;
  const int nmid = nxt - nxe;
  x = 2 + (int)(t % nmid);
  const int q = (int)(t / nmid);
  j = (jm <= 2 * NEAR || q < NEAR) ? 1 + q : jm - 2 * NEAR + 1 + q;
}

template <int NZ, int MODE, bool EDGE>
__global__ void __launch_bounds__(EDGE ? JR_EDGE_THREADS : JR_THREADS) k_jac_csr_rows(Mesh m, Zones z, const double* __restrict__ bi,
                                                             const double* __restrict__ bj,
                                                             const double* __restrict__ coefdiag,
                                                             const double* __restrict__ G, double mini,
                                                             unsigned long long* __restrict__ masks,
                                                             int* __restrict__ cnt, const int* __restrict__ indptr,
                                                             int* __restrict__ indices, double* __restrict__ data,
                                                             int* __restrict__ changed, int* __restrict__ bad) {
  constexpr int NA = 7 * NZ;
  constexpr int TH = EDGE ? JR_EDGE_THREADS : JR_THREADS;
  constexpr int JR_NIT = (JR_NELEM + TH - 1) / TH;
  constexpr int NWAVE = TH / 64;
  constexpr int RPW = JC_TI / NWAVE;                             // rows per wave
  constexpr int STEP = 7;                                        // 2 gh + 1, gh = 3
  __shared__ double J[245 * JR_PITCH];                           // (k, D, m) x il
  __shared__ JcTables<NZ, true> T;
  __shared__ bool plain[JC_TI];
  const int tid = threadIdx.x;
  const int im = m.im, jm = m.jm;
  const int nxt = (im + JC_TI - 1) / JC_TI;
  int tx = blockIdx.x, j = 1 + blockIdx.y;
  if constexpr (EDGE) jr_edge_tile(blockIdx.x, nxt, jm, tx, j);
  const int i0 = 1 + tx * JC_TI;
  const bool edge_tile = jr_edge_x(tx, nxt) || jr_edge_j(j, jm);
  const long nt = 5L * im * jm;
  const int nfi = (im + 1) * jm, nfj = im * (jm + 1);
  if (tid < 7) T.inv_k[tid] = -1;
  jc_tables<NZ, true>(m, z, i0, j, T);
  // window entries no face reaches stay zero for every e
  for (int x = tid; x < 245 * JR_PITCH; x += TH) J[x] = 0.0;
  __syncthreads();
  const int wave = tid / 64, lane = tid % 64;

  // value items (il, m, cell): the four faces' block offsets at e = 0 (-1:
  // no block) and the LDS slot, e-independent, kept in registers
  int vofs[JR_NIT][4];
  int vdst[JR_NIT];
#pragma unroll
  for (int it = 0; it < JR_NIT; ++it) {
    const int item = tid + it * TH;
    const int il = item % JC_TI, rest = item / JC_TI, mm = rest % 5, cell = rest / 5;
    const int i = i0 + il;
    vdst[it] = -1;
#pragma unroll
    for (int f = 0; f < 4; ++f) vofs[it][f] = -1;
    if (item >= JR_NELEM || i > im) continue;
    const int di = jr_cell_di(cell), dj = jr_cell_dj(cell);
    const int ci = i + di, cj = j + dj;
    const int k = ((dj + j - 1) % STEP + STEP) % STEP;
    if (ci < 1 || ci > im || cj < 1 || cj > jm || T.dj[k] != dj) continue;
    const int o0 = ioff_index(di, dj), o1 = ioff_index(di - 1, dj);
    const int o2 = joff_index(di, dj), o3 = joff_index(di, dj - 1);
    const int fi = (i - 1) + (im + 1) * (j - 1), fj = (i - 1) + im * (j - 1);
    if (o0 >= 0 && ioff_live(m, o0, i)) vofs[it][0] = (o0 * 25 + mm) * nfi + fi;
    if (o1 >= 0 && ioff_live(m, o1, i + 1)) vofs[it][1] = (o1 * 25 + mm) * nfi + fi + 1;
    if (o2 >= 0 && joff_live(m, o2, j)) vofs[it][2] = (o2 * 25 + mm) * nfj + fj;
    if (o3 >= 0 && joff_live(m, o3, j + 1)) vofs[it][3] = (o3 * 25 + mm) * nfj + fj + im;
    vdst[it] = ((k * 7 + di + 3) * 5 + mm) * JR_PITCH + il;
  }

  // plain rows: every zone's window is the 7 x 7 neighbourhood, no ghost part
  if (tid < JC_TI) {
    const int r = tid, i = i0 + r;
    bool pl = i <= im && near_index(m, i, j) < 0;
    const int rot = (i - 1) % STEP;
    for (int k = 0; k < 7 && pl; ++k) pl = T.dj[k] >= -3 && T.dj[k] <= 3 && T.vjk[k] == j - 1 + T.dj[k];
    for (int t = 0; t < NA && pl; ++t) {
      int dl = t % 7 - rot;
      dl += dl < -3 ? 7 : 0;
      dl -= dl > 3 ? 7 : 0;
      pl = T.vil[r][t] == i - 1 + dl;
    }
    plain[r] = pl;
    // every row that is not plain lies in a tile of the edge launch
    if (!EDGE && !pl && i <= im && !edge_tile) *bad = 2;
  }
  __syncthreads();
  {
    // a tile with no row for this launch leaves (the decision is the same
    // for every thread of the block)
    bool any = false;
    for (int r = 0; r < JC_TI; ++r) any = any || (i0 + r <= im && (EDGE ? !plain[r] : plain[r]));
    if (!any) return;
  }
  const PlainLane pln = plain_lane(j);
  RowPrefetch<MODE, EDGE ? CSR_MASK_WORDS : JR_PLAIN_WORDS, RPW> pf;
  pf.load(m, i0, j, wave, NWAVE, masks, indptr, coefdiag);

#pragma unroll 1
  for (int e = 0; e < 5; ++e) {
    // J(il, cell, m) of variable row e, each summed in registers in the face
    // order of the COO path; loads from a stand-in offset 0 where a face has
    // no block keep them all in flight, selects keep the sums exact
#pragma unroll
    for (int it = 0; it < JR_NIT; ++it) {
      // only the rows this launch turns into CSR rows (plain rows here, the
      // others in the edge launch)
      if (vdst[it] < 0 || (EDGE ? plain[vdst[it] % JR_PITCH] : !plain[vdst[it] % JR_PITCH])) continue;
      const int h0 = vofs[it][0] >= 0, h1 = vofs[it][1] >= 0, h2 = vofs[it][2] >= 0, h3 = vofs[it][3] >= 0;
      const double x0 = bi[h0 ? vofs[it][0] + e * 5 * nfi : 0];
      const double x1 = bi[h1 ? vofs[it][1] + e * 5 * nfi : 0];
      const double x2 = bj[h2 ? vofs[it][2] + e * 5 * nfj : 0];
      const double x3 = bj[h3 ? vofs[it][3] + e * 5 * nfj : 0];
      double acc = 0.0;
      acc = h0 ? acc - x0 : acc;
      acc = h1 ? acc + x1 : acc;
      acc = h2 ? acc - x2 : acc;
      acc = h3 ? acc + x3 : acc;
      J[vdst[it]] = acc;
    }
    __syncthreads();
    // rows in column order
#pragma unroll
    for (int qr = 0; qr < RPW; ++qr) {
      const int r = wave + NWAVE * qr, i = i0 + r;
      if (i > im) break;
      const int row = e + (j - 1) * 5 + (i - 1) * jm * 5;
      const long tr = csr_tid(m, i, j, e);
      const long tn = near_index(m, i, j);
      if constexpr (!EDGE) {
        if (plain[r])
          jr_row_plain<NZ, MODE>(pln, J, r, i, j, jm, row, tr, nt, e, pf, qr, coefdiag != nullptr, mini, masks,
                                 cnt, indices, data, changed);
      } else {
        if (!plain[r])
          jr_row_edge<NZ, MODE>(J, T, z, r, i, jm, row, tr, nt, e, pf, qr, coefdiag != nullptr, G, tn, mini,
                                masks, cnt, indices, data, changed, bad);
      }
    }
    __syncthreads();
  }
}

}  // namespace bca
