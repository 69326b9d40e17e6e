// Adjoint (reverse mode) of the ghost-filling boundary conditions, as the
// TAPENADE _b routines of the reference (srcfv/adjoint/bc_supandsubinlet_b.f90,
// bc_no_reflexion_b.f90, bc_extrapolateo2_b.f90, bc_wall_viscous_b.f90,
// jn_match_2d_b.f90): on exit wb(ghosts of the BC) = 0 and every cell the BC
// reads has received (d ghosts / d cell)^T wb(ghosts); w is left as given.
//
// A BC line reads the cells at offsets 0..3 along its inward normal and writes
// its gh ghost layers (deeper layers may read the ones before, which the
// forward-mode line function composes).  So the transpose is formed from the
// tangent line functions: for each offset d, a unit Dual<5> tangent on the
// offset-d cell of every line (lines are disjoint), the tangent BC applied,
// and the ghost tangents contracted with wb(ghosts).  The four offsets are
// four slots of one seeded array (slot d at ws + d cs, blockIdx.y = d), so a
// BC takes one seed, one tangent BC and one accumulate launch.
#pragma once
#include "bcs.h"
#include "jacobian.h"

namespace bca {

constexpr int ADJ_REACH = 4;  // offsets 0..3 along the inward normal

__device__ inline void line_cell(int loc, Box b, int l, int& i, int& j, int& i0, int& j0) {
  const Side s = decode_side(loc, b.imin, b.jmin, b.imax, b.jmax);
  i = b.imin + (l - 1) * s.j0 * s.j0;
  j = b.jmin + (l - 1) * s.i0 * s.i0;
  i0 = s.i0;
  j0 = s.j0;
}

// ws = w as Dual<5> with unit tangents d[e] on variable e of the offset-d
// cell of every line of the interface, stored as the compact frame the
// boundary conditions read or write (the ghost frame plus 4 interior layers,
// StateView<T, true> layout; the offset-d cells, d <= 3, lie in it).
__global__ void k_adj_seed(Mesh m, const double* __restrict__ w, Dual<5>* __restrict__ wsd, int loc, Box b,
                           long cs) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int d = blockIdx.y;
  Dual<5>* __restrict__ ws = wsd + d * cs;
  if (t >= frame_count(m, BC_FRAME)) return;
  int a, bb;
  frame_cell(m, BC_FRAME, t, a, bb);
  const long c = a + (long)m.ni * bb, fc = frame_count(m, BC_FRAME);
  const int i = a - m.gh + 1, j = bb - m.gh + 1;
  const Side s = decode_side(loc, b.imin, b.jmin, b.imax, b.jmax);
  const bool on = s.kdir == 1 ? (i == b.imin + d * s.i0 && j >= b.jmin && j <= b.jmax)
                              : (j == b.jmin + d * s.j0 && i >= b.imin && i <= b.imax);
#pragma unroll
  for (int e = 0; e < 5; ++e) {
    Dual<5> x(w[c + e * m.ncell]);
    if (on) x.d[e] = 1.0;
    ws[t + e * fc] = x;
  }
}

// wb(offset-d cell, q) += sum over the line's ghosts g and variables e of
// d ws(g, e) / d w(offset-d cell, q) * wb(g, e)
__global__ void k_adj_accumulate(Mesh m, const Dual<5>* __restrict__ wsd, double* __restrict__ wb, int loc, Box b,
                                 long cs, int lmax) {
  const int l = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int d = blockIdx.y;
  const Dual<5>* __restrict__ ws = wsd + d * cs;
  if (l > lmax) return;
  int i, j, i0, j0;
  line_cell(loc, b, l, i, j, i0, j0);
  const long n = m.ncell, fc = frame_count(m, BC_FRAME);
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (int de = 1; de <= m.gh; ++de) {
    const int ig = i - de * i0, jg = j - de * j0;
    const long g = cidx(m, ig, jg), gf = frame_index(m, BC_FRAME, ig + m.gh - 1, jg + m.gh - 1);
#pragma unroll
    for (int e = 0; e < 5; ++e) {
      const Dual<5> x = ws[gf + e * fc];
      const double y = wb[g + e * n];
#pragma unroll
      for (int q = 0; q < 5; ++q) acc[q] += x.d[q] * y;
    }
  }
  const long s = cidx(m, i + d * i0, j + d * j0);
#pragma unroll
  for (int q = 0; q < 5; ++q) wb[s + q * n] += acc[q];
}

__global__ void k_adj_clear_ghosts(Mesh m, double* __restrict__ wb, int loc, Box b, int lmax) {
  const int l = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  if (l > lmax) return;
  int i, j, i0, j0;
  line_cell(loc, b, l, i, j, i0, j0);
  for (int de = 1; de <= m.gh; ++de) {
    const long g = cidx(m, i - de * i0, j - de * j0);
#pragma unroll
    for (int e = 0; e < 5; ++e) wb[g + e * m.ncell] = 0.0;
  }
}

// In-block join (identity transform, disjoint boxes): the copy wb(prr) <-
// wb(prd) transposed, wb(prd cell) += wb(prr cell), then wb(prr cell) = 0.
__global__ void k_join_adjoint(Mesh m, double* __restrict__ wb, Box prr, Box prd, int em) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int ni_ = prd.imax - prd.imin + 1, nj_ = prd.jmax - prd.jmin + 1;
  if (t >= ni_ * nj_) return;
  const int a = t % ni_, q = t / ni_;
  const long dst = cidx(m, prr.imin + a, prr.jmin + q), src = cidx(m, prd.imin + a, prd.jmin + q);
  for (int e = 0; e < em; ++e) {
    wb[src + e * m.ncell] += wb[dst + e * m.ncell];
    wb[dst + e * m.ncell] = 0.0;
  }
}

}  // namespace bca
