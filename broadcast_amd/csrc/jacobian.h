// Colored sparse-Jacobian assembly from face-local Jacobian blocks.
//
// The reference builds -dR/dw column-group by column-group: 245 one-hot seeds
// (5 variables x 7 x 7 spatial colors), each a full tangent evaluation of BCs
// + residual, scattered into COO (BROADCAST_npz.py:1068-1127,
// misc/ComputeJacobian.f90:357-374, 503-570).  Here the same numbers come from
// two passes that never evaluate a dense tangent:
//
//  phase 1  (k_iface_blocks / k_jface_blocks) every face computes
//           B(face, o) = d hn(face) / d w(face + o) for the 24 stencil offsets
//           o of its family by compile-time sparse forward mode (SeededAcc),
//           ghost cells included as independent inputs, one plane per offset;
//  ghosts   (k_seed_frames + linearised BCs + k_ghost_tangents, all 49 colors
//           per launch) the tangent the BC sequence gives every ghost cell,
//           T(color, g), computed on the boundary frame only;
//  phase 2  (k_assemble_colored, then k_assemble_ghosts on the rows near the
//           mesh edge) for every row cell and color, the tangent residual of
//           that color -- the sum over the cell's 4 faces of the block of the
//           color's interior cell in the face stencil, plus B(face, ghost)
//           T(color, ghost) -- written into the reference's COO slots with its
//           window rules.
//
// For every seed this is exactly the reference's JVP (the tangent of a face is
// the sum over its active inputs), so the COO values agree to rounding.
#pragma once
#include "bcs.h"
#include "blockskern.h"
#include "misc.h"

namespace bca {

// Seed zones in i.  One zone spanning the mesh is the reference's plain
// coloring (testvector + computejacobianfromjv[_relaxed][_withjn]); two zones
// are the periodic two-zone coloring (testvector_partial +
// computejacobianfromjv_relaxed_withjnandcheck, cylinder.py:895-935), each
// zone filling its own block of the doubled COO.  lo/hi are 1-based inclusive.
struct Zones {
  int n;
  int lo[2], hi[2];
  int withjn;
};

// the row's seed column under the zone's window rules (empty-slot semantics)
__device__ inline bool zone_column(const Mesh& m, const Zones& z, int zn, int i, int j, int l, int k, int& vali,
                                   int& valj) {
  if (z.n == 1) return seed_column(i, j, l, k, m.gh, m.im, m.jm, z.withjn != 0, vali, valj);
  bool wrapped;
  return seed_column_zone(i, j, l, k, m.gh, m.im, m.jm, zn, vali, valj, wrapped);
}

// ------------------------------------------------------------------ ghosts
// Seeded states of all zones x 49 colors at once (blockIdx.y = zone * 49 +
// color), stored as the compact frame the boundary conditions read or write
// (the ghost frame plus 4 interior layers; StateView<T, true> layout, planes
// of frame_count(m, BC_FRAME) cells).  State values from w, unit tangent
// d[m] on variable m of the cells the color seeds in the zone, zero elsewhere.
__global__ void k_seed_frames(Mesh m, Zones z, const double* __restrict__ w, Dual<5>* __restrict__ ws,
                              long cstride) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= frame_count(m, BC_FRAME)) return;
  const int step = 2 * m.gh + 1, ncol = step * step;
  const int color = blockIdx.y % ncol, zn = blockIdx.y / ncol;
  const int l = color / step, k = color % step;
  int a, b;
  frame_cell(m, BC_FRAME, t, a, b);
  const int i = a - m.gh + 1, j = b - m.gh + 1;
  const int i0 = z.lo[zn];
  const bool on = i >= i0 + l && i <= z.hi[zn] && j >= k + 1 && j <= m.jm && (i - i0 - l) % step == 0 &&
                  (j - k - 1) % step == 0;
  const long c = a + (long)m.ni * b, fc = frame_count(m, BC_FRAME);
  Dual<5>* out = ws + (long)blockIdx.y * cstride;
#pragma unroll
  for (int e = 0; e < 5; ++e) {
    Dual<5> x(w[c + e * m.ncell]);
    if (on) x.d[e] = 1.0;
    out[t + e * fc] = x;
  }
}

// T(color, g)[q][m]: tangent of ghost cell g, variable q, under seed (m, color);
// nz(color, g) = 1 where T(color, g) has a nonzero entry.  A ghost state
// depends on a few interior cells only (mirror, extrapolation and join
// sources), so most (color, ghost) pairs are zero and the fold skips them.
__global__ void k_ghost_tangents(Mesh m, const Dual<5>* __restrict__ ws, long cstride, double* __restrict__ T,
                                 unsigned char* __restrict__ nz) {
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long ng = ghost_count(m);
  if (g >= ng) return;
  const int color = blockIdx.y;
  int a, b;
  frame_cell(m, m.gh, g, a, b);  // the ghost frame, in ghost_index order
  const long c = frame_index(m, BC_FRAME, a, b), fc = frame_count(m, BC_FRAME);
  const Dual<5>* src = ws + (long)color * cstride;
  double* out = T + ((long)color * ng + g) * 25;
  bool any = false;
#pragma unroll
  for (int q = 0; q < 5; ++q) {
    const Dual<5> x = src[c + q * fc];
#pragma unroll
    for (int mm = 0; mm < 5; ++mm) {
      out[q * 5 + mm] = x.d[mm];
      any = any || x.d[mm] != 0.0;
    }
  }
  nz[(long)color * ng + g] = any ? 1 : 0;
}

// ------------------------------------------------------------------ phase 2
__device__ inline bool interior(const Mesh& m, int i, int j) {
  return i >= 1 && i <= m.im && j >= 1 && j <= m.jm;
}

// Row offset dj (relative to row j) of the cell of seed color k that lies in
// the row's stencil: the residue of k - (j - 1) taken in [-3, 3], moved by one
// period when that cell is outside the mesh (the near-wall Euler flux of the
// face j = 2 reaches the cell j = 5 from the row j = 1).
__device__ inline int color_offset(int k, int j, int n, int gh) {
  const int step = 2 * gh + 1;
  int d = ((k - (j - 1)) % step + step) % step;
  if (d > gh) d -= step;
  if (j + d < 1) d += step;
  else if (j + d > n) d -= step;
  return d;
}

// Interior part.  Grid (x: i, y: j, z: color row k), one thread per (row cell,
// k).  Within a wave j and k are uniform, so is the column row offset dj; the
// row's blocks for the 7 column offsets di = -3..3 are sums of the blocks of
// its 4 faces read at wave-uniform offset planes (coalesced), and color l of
// the lane picks di = residue of l - (i - 1) by a select chain.  A stencil
// holds at most one interior cell per color (its widest extent, 8 cells,
// occurs only on the first/last face where one end is a ghost), so this is the
// whole interior contribution of each seed.  Values -dR/dw (+ coefdiag on the
// diagonal) go to the reference's COO slots (misc/ComputeJacobian.f90:503-570).
template <int NZ>
__global__ void __launch_bounds__(256, NZ > 1 ? 4 : 3) k_assemble_colored(Mesh m, Zones z, const double* __restrict__ bi,
                                                          const double* __restrict__ bj,
                                                          const double* __restrict__ coefdiag,
                                                          double* __restrict__ jac, int* __restrict__ ia,
                                                          int* __restrict__ ja, int write_index) {
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  const int k = blockIdx.z;
  if (i > m.im || j > m.jm) return;
  const int im = m.im, jm = m.jm, gh = m.gh, step = 2 * gh + 1;
  // 32-bit face-block offsets (the host checks that the block arrays fit)
  const int nfi = (im + 1) * jm, nfj = im * (jm + 1);
  const int fi0 = (i - 1) + (im + 1) * (j - 1), fj0 = (i - 1) + im * (j - 1);
  const int dj = color_offset(k, j, jm, gh);
  const bool jin = j + dj >= 1 && j + dj <= jm;
  // per column offset di = D - 3 (wave-uniform): the offset indices o + 1 of
  // the four faces' blocks (0: no block), 5 bits each, and the zone of the
  // column cell i + di in bit 20 (each cell belongs to one zone)
  int code[7];
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
  // seed columns of the 7 colors (l, k) of this row, per zone (-1: dummy slot)
  int col[NZ][7];
  const int rot = (i - 1) % step;
#pragma unroll
  for (int zn = 0; zn < NZ; ++zn)
#pragma unroll
    for (int l = 0; l < 7; ++l) {
      int vali, valj;
      col[zn][l] = zone_column(m, z, zn, i, j, l, k, vali, valj) ? valj * 5 + vali * jm * 5 : -1;
    }
  const double cd = coefdiag ? coefdiag[(i - 1) + (long)im * (j - 1)] : 0.0;
  const long plane = (long)im * jm;
  const long base = (long)(i - 1) + (long)(j - 1) * im + (long)k * plane * 5;
  const int dummy = (z.withjn || NZ > 1) ? 5 * im * jm - 2 : 5 * im * jm - 1;
#pragma unroll
  for (int e = 0; e < 5; ++e) {
    const int row = e + (j - 1) * 5 + (i - 1) * jm * 5;
#pragma unroll
    for (int mm = 0; mm < 5; ++mm) {
      const int q = e * 5 + mm;
      double v[7];
#pragma unroll
      for (int D = 0; D < 7; ++D) {
        // jac = -dR/dw with R(i,j) = hi(i) - hi(i+1) + hj(j) - hj(j+1); the
        // loads are unguarded (a missing block reads offset 0) so that they
        // are all in flight at once, the selects keep the sum exact
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
#pragma unroll
      for (int l = 0; l < 7; ++l) {
        // column offset di of color l: residue of l - (i-1) in [-3, 3]
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
          const long cur = base + (long)e * plane + (long)l * plane * 5 * step +
                           (long)mm * plane * 5 * step * step + (long)zn * plane * 25 * step * step;
          const bool ok = col[zn][l] >= 0;
          const int c = ok ? col[zn][l] + mm : 0;
          double val = (ok && xz == zn) ? x : 0.0;
          if (ok && coefdiag && row == c) val += cd;
          jac[cur] = val;
          if (write_index) {
            ia[cur] = ok ? row : dummy;
            ja[cur] = c;
          }
        }
      }
    }
  }
}

// Ghost part, rows within NEAR cells of the mesh edge only: for each seed color
// c, sum over the row's faces and the ghost cells g of their stencils of
// B(face, g) T(c, g), added to the values k_assemble_colored wrote.
constexpr int NEAR = 5;
__host__ __device__ inline bool near_whole(const Mesh& m) { return 2 * NEAR >= m.im || 2 * NEAR >= m.jm; }
__host__ __device__ inline long near_count(const Mesh& m) {
  if (near_whole(m)) return (long)m.im * m.jm;
  return 2L * NEAR * m.im + 2L * NEAR * (m.jm - 2 * NEAR);
}
__device__ inline void near_cell(const Mesh& m, long t, int& i, int& j) {
  const int W = NEAR;
  if (near_whole(m) || t < (long)W * m.im) { j = 1 + (int)(t / m.im); i = 1 + (int)(t % m.im); return; }
  t -= (long)W * m.im;
  if (t < (long)W * m.im) { j = m.jm - W + 1 + (int)(t / m.im); i = 1 + (int)(t % m.im); return; }
  t -= (long)W * m.im;
  j = W + 1 + (int)(t / (2 * W));
  const int s = (int)(t % (2 * W));
  i = 1 + (s < W ? s : m.im - 2 * W + s);
}

// index t of near_cell(m, t) = (i, j), or -1 when (i, j) is not a near row
__device__ inline long near_index(const Mesh& m, int i, int j) {
  const int W = NEAR;
  if (near_whole(m)) return (long)(j - 1) * m.im + (i - 1);
  if (j <= W) return (long)(j - 1) * m.im + (i - 1);
  if (j > m.jm - W) return (long)W * m.im + (long)(j - (m.jm - W + 1)) * m.im + (i - 1);
  long t = 2L * W * m.im + (long)(j - W - 1) * 2 * W;
  if (i <= W) return t + (i - 1);
  if (i > m.im - W) return t + (i - 1 - m.im + 2 * W);
  return -1;
}

// Row e of the ghost part of one face: acc[m] += sg sum_q B(face, o)[e][q]
// T(color, g)[q][m] over the face's stencil offsets o whose cell g is a ghost
// with a nonzero tangent.
template <bool JFAM>
__device__ inline void ghost_face(const Mesh& m, const double* blk, long nf, long f, int fi, int fj, double sg,
                                  const double* Tc, const unsigned char* nzc, int e, double acc[5]) {
  for (int o = 0; o < NOFF; ++o) {
    if (JFAM ? !joff_live(m, o, fj) : !ioff_live(m, o, fi)) continue;
    const int ci = fi + (JFAM ? joff_di(o) : ioff_di(o));
    const int cj = fj + (JFAM ? joff_dj(o) : ioff_dj(o));
    if (interior(m, ci, cj)) continue;
    if (ci < 1 - m.gh || ci > m.im + m.gh || cj < 1 - m.gh || cj > m.jm + m.gh) continue;
    const long g = ghost_index(m, ci + m.gh - 1, cj + m.gh - 1);
    if (!nzc[g]) continue;
    const double* tg = Tc + g * 25;
    const double* b = blk + ((long)o * 25 + e * 5) * nf + f;
    double bb[5];
#pragma unroll
    for (int q = 0; q < 5; ++q) bb[q] = b[q * nf];
#pragma unroll
    for (int mm = 0; mm < 5; ++mm) {
      double s = 0.0;
#pragma unroll
      for (int q = 0; q < 5; ++q) s += bb[q] * tg[q * 5 + mm];
      acc[mm] += sg * s;
    }
  }
}

// Grid (x: near rows, y: (zone * 49 + color) * 5 + e), one thread per near
// row cell, slot color and variable row e: adds the ghost part to the COO
// slots.
__global__ void __launch_bounds__(256) k_assemble_ghosts(Mesh m, Zones z, const double* __restrict__ bi,
                                                         const double* __restrict__ bj,
                                                         const double* __restrict__ T,
                                                         const unsigned char* __restrict__ nz,
                                                         double* __restrict__ jac) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= near_count(m)) return;
  const int im = m.im, jm = m.jm, gh = m.gh, step = 2 * gh + 1;
  const int slot = blockIdx.y / 5, e = blockIdx.y % 5;
  const int color = slot % (step * step), zn = slot / (step * step);
  const int l = color / step, k = color % step;
  int i, j;
  near_cell(m, t, i, j);
  int vali, valj;
  if (!zone_column(m, z, zn, i, j, l, k, vali, valj)) return;
  const long nfi = (long)(im + 1) * jm, nfj = (long)im * (jm + 1);
  const long fi0 = (i - 1) + (long)(im + 1) * (j - 1), fi1 = fi0 + 1;
  const long fj0 = (i - 1) + (long)im * (j - 1), fj1 = fj0 + im;
  const long ng = ghost_count(m);
  const double* Tc = T + (long)slot * ng * 25;
  const unsigned char* nzc = nz + (long)slot * ng;
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  ghost_face<false>(m, bi, nfi, fi0, i, j, -1.0, Tc, nzc, e, acc);
  ghost_face<false>(m, bi, nfi, fi1, i + 1, j, +1.0, Tc, nzc, e, acc);
  ghost_face<true>(m, bj, nfj, fj0, i, j, -1.0, Tc, nzc, e, acc);
  ghost_face<true>(m, bj, nfj, fj1, i, j + 1, +1.0, Tc, nzc, e, acc);
  const long plane = (long)im * jm;
  const long base = (long)(i - 1) + (long)(j - 1) * im + (long)k * plane * 5 + (long)l * plane * 5 * step +
                    (long)zn * plane * 25 * step * step + (long)e * plane;
#pragma unroll
  for (int mm = 0; mm < 5; ++mm) jac[base + (long)mm * plane * 5 * step * step] += acc[mm];
}

// The ghost cell that face f of the row cell (i, j) -- 0, 1: left and right
// i-face, 2, 3: lower and upper j-face -- reads through its stencil offset
// o, or -1 when that cell is not a ghost (or the offset is not live).
__device__ inline long ghost_of(const Mesh& m, int f, int o, int i, int j) {
  const bool jf = f >= 2;
  const int fi = i + (f == 1 ? 1 : 0), fj = j + (f == 3 ? 1 : 0);
  if (jf ? !joff_live(m, o, fj) : !ioff_live(m, o, fi)) return -1;
  const int ci = fi + (jf ? joff_di(o) : ioff_di(o));
  const int cj = fj + (jf ? joff_dj(o) : ioff_dj(o));
  if (interior(m, ci, cj)) return -1;
  if (ci < 1 - m.gh || ci > m.im + m.gh || cj < 1 - m.gh || cj > m.jm + m.gh) return -1;
  return ghost_index(m, ci + m.gh - 1, cj + m.gh - 1);
}

// The ghost part in the compact array of the fused CSR path:
// G[(t * nslot + zone * 49 + color) * 25 + e * 5 + m], zeros for a slot
// outside the zone's window.  Grid (x: near rows, y: slot); one thread per
// near row cell and slot forms its five variable rows (the face and ghost
// search done once for them, its nonzero-tangent probes all in flight
// together) and stores its 25 values as one run.
__global__ void __launch_bounds__(256) k_ghost_part(Mesh m, Zones z, const double* __restrict__ bi,
                                                    const double* __restrict__ bj, const double* __restrict__ T,
                                                    const unsigned char* __restrict__ nz, double* __restrict__ G) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= near_count(m)) return;
  const int im = m.im, jm = m.jm, gh = m.gh, step = 2 * gh + 1;
  const int slot = blockIdx.y, nslot = gridDim.y;
  const int color = slot % (step * step), zn = slot / (step * step);
  const int l = color / step, k = color % step;
  int i, j;
  near_cell(m, t, i, j);
  int vali, valj;
  double* gout = G + (t * nslot + slot) * 25;
  double acc[25];
#pragma unroll
  for (int x = 0; x < 25; ++x) acc[x] = 0.0;
  if (zone_column(m, z, zn, i, j, l, k, vali, valj)) {
    const long nfi = (long)(im + 1) * jm, nfj = (long)im * (jm + 1);
    const long fi0 = (i - 1) + (long)(im + 1) * (j - 1);
    const long fj0 = (i - 1) + (long)im * (j - 1);
    const long ng = ghost_count(m);
    const double* Tc = T + (long)slot * ng * 25;
    const unsigned char* nzc = nz + (long)slot * ng;
    // (face, offset) pairs whose ghost has a nonzero tangent, bit f NOFF + o:
    // every probe issued at once (from entry 0 where there is no ghost)
    unsigned long long mk[2] = {0ull, 0ull};
#pragma unroll
    for (int f = 0; f < 4; ++f) {
#pragma unroll
      for (int o = 0; o < NOFF; ++o) {
        const long g = ghost_of(m, f, o, i, j);
        const bool live = (g >= 0) & (nzc[g >= 0 ? g : 0] != 0);
        const int b = f * NOFF + o;
        mk[b / 64] |= (unsigned long long)live << (b % 64);
      }
    }
    // the terms in the order of ghost_face: faces left, right, lower,
    // upper, offsets ascending within each
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      unsigned long long bits = mk[h];
      while (bits) {
        const int b = h * 64 + __ffsll((long long)bits) - 1;
        bits &= bits - 1;
        const int f = b / NOFF, o = b % NOFF;
        const long g = ghost_of(m, f, o, i, j);
        const double sg = (f == 0 || f == 2) ? -1.0 : 1.0;
        const double* blk = f < 2 ? bi : bj;
        const long nf = f < 2 ? nfi : nfj;
        const long fc = f == 0 ? fi0 : f == 1 ? fi0 + 1 : f == 2 ? fj0 : fj0 + im;
        const double* tg = Tc + g * 25;
        double tt[25];
#pragma unroll
        for (int x = 0; x < 25; ++x) tt[x] = tg[x];
#pragma unroll
        for (int e = 0; e < 5; ++e) {
          const double* bp = blk + ((long)o * 25 + e * 5) * nf + fc;
          double bb[5];
#pragma unroll
          for (int q = 0; q < 5; ++q) bb[q] = bp[q * nf];
#pragma unroll
          for (int mm = 0; mm < 5; ++mm) {
            double sum = 0.0;
#pragma unroll
            for (int q = 0; q < 5; ++q) sum += bb[q] * tt[q * 5 + mm];
            acc[e * 5 + mm] += sg * sum;
          }
        }
      }
    }
  }
#pragma unroll
  for (int x = 0; x < 25; ++x) gout[x] = acc[x];
}

}  // namespace bca
