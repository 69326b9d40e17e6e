// Face-local Jacobian block kernels: B(face, o) = d hn(face) / d w(face + o)
// for the 24 stencil offsets o of the face family, by compile-time sparse
// forward mode (SeededAcc).  Two outputs of the same passes:
//   MODE_BLOCKS  the blocks themselves (phase 1 of the colored assembly);
//   MODE_VJP     hnb(face)^T B(face, o), 5 values per offset, with hnb the
//                face adjoint of residub (the VJP, gathered per cell later).
// Templates only: each variant is instantiated in its own object (blocks.hip).
#pragma once
#include "faces.h"

namespace bca {

constexpr int NOFF = 24;  // stencil offsets per face family

// i-face (between cells (i-1, j) and (i, j)) stencil, offsets from (i, j):
// row 0: -3..2, rows -2, -1, 1, 2: -2..1, plus (3, 0) and (-4, 0) reached only
// through the extrapolated ghost gradients of the faces i = 1 and i = im+1.
__host__ __device__ constexpr int ioff_di(int o) {
  return o < 6 ? o - 3 : o < 22 ? ((o - 6) % 4) - 2 : (o == 22 ? 3 : -4);
}
__host__ __device__ constexpr int ioff_dj(int o) {
  return o < 6 ? 0 : o < 22 ? ((o - 6) / 4 < 2 ? (o - 6) / 4 - 2 : (o - 6) / 4 - 1) : 0;
}
// j-face (between (i, j-1) and (i, j)): transpose of the above, plus (0, 3)
// (the off-centred Euler flux of row 2) and (0, -4) (ghost gradients of row
// jm+1).
__host__ __device__ constexpr int joff_di(int o) { return ioff_dj(o); }
__host__ __device__ constexpr int joff_dj(int o) { return ioff_di(o); }

// offset -> index lookups (-1: not in the stencil); i-face table over
// di in [-4, 3] x dj in [-2, 2], j-face table over di in [-2, 2] x dj in [-4, 3]
__host__ __device__ inline int ioff_index(int di, int dj) {
  if (dj == 0) {
    if (di >= -3 && di <= 2) return di + 3;
    if (di == 3) return 22;
    if (di == -4) return 23;
    return -1;
  }
  if (dj < -2 || dj > 2 || di < -2 || di > 1) return -1;
  const int r = dj < 0 ? dj + 2 : dj + 1;  // -2,-1,1,2 -> 0,1,2,3
  return 6 + r * 4 + (di + 2);
}
__host__ __device__ inline int joff_index(int di, int dj) { return ioff_index(dj, di); }

// is offset index o of a face at (fi, fj) a stored block?  (the two extra
// offsets exist only on the faces listed in the phase-1 kernels)
__host__ __device__ inline bool ioff_live(const Mesh& m, int o, int fi) {
  return o < 22 || (o == 22 && fi == 1) || (o == 23 && fi == m.im + 1);
}
__host__ __device__ inline bool joff_live(const Mesh& m, int o, int fj) {
  return o < 22 || (o == 22 && fj == 2) || (o == 23 && fj == m.jm + 1);
}

// compact index of a ghost-frame cell given storage coordinates (a, b)
__host__ __device__ inline long ghost_index(const Mesh& m, int a, int b) {
  const int gh = m.gh;
  if (b < gh) return (long)b * m.ni + a;
  if (b >= gh + m.jm) return (long)gh * m.ni + (long)(b - gh - m.jm) * m.ni + a;
  return 2L * gh * m.ni + (long)(b - gh) * 2 * gh + (a < gh ? a : a - m.im);
}
__host__ __device__ inline long ghost_count(const Mesh& m) {
  return 2L * m.gh * m.ni + 2L * m.gh * m.jm;
}

// ------------------------------------------------------------------ phase 1
enum { MODE_BLOCKS = 0, MODE_VJP = 1 };

template <int SI, int SJ>
__device__ inline SeededAcc<SI, SJ> seeded(const Mesh& m, const Phys& ph, const Metric& g,
                                           const double* w, const double* prim, const double* grad,
                                           int i, int j) {
  SeededAcc<SI, SJ> a;
  a.m = m; a.ph = ph; a.g = g; a.i = i; a.j = j;
  a.w = w; a.prim = prim; a.grad = grad;
  a.seed();
  return a;
}

// Output of the seeded pass of offset o at face f.  MODE_BLOCKS: B(face, o)
// in plane o, out[(o * 25 + e * 5 + q) * nf + f]; MODE_VJP: sum_e hb[e]
// B(face, o)[e][q] in out[(o * 5 + q) * nf + f].  Either way a wave
// (consecutive faces, one offset) stores coalesced rows.
template <int MODE>
struct FaceSink {
  double* out;
  long nf, f;
  double hb[5];
  __device__ void put(int o, const Dual<5> hn[5]) const {
    if constexpr (MODE == MODE_BLOCKS) {
      double* p = out + (long)o * 25 * nf + f;
#pragma unroll
      for (int e = 0; e < 5; ++e)
#pragma unroll
        for (int q = 0; q < 5; ++q) p[(e * 5 + q) * nf] = hn[e].d[q];
    } else {
      double* p = out + (long)o * 5 * nf + f;
#pragma unroll
      for (int q = 0; q < 5; ++q) {
        double t = 0.0;
#pragma unroll
        for (int e = 0; e < 5; ++e) t += hb[e] * hn[e].d[q];
        p[q * nf] = t;
      }
    }
  }
};

// Face adjoints of the balance R(i,j) = hn(i) - hn(i+1) + hn(j) - hn(j+1) over
// the interior rows (srcfv/adjoint/flux_num_dnc5_b.f90:800-829).
__device__ inline void face_adjoint(const Mesh& m, const double* resb, int i, int j, int di, int dj,
                                    double hb[5]) {
  const bool lo = (i - di) >= 1 && (j - dj) >= 1, hi = i <= m.im && j <= m.jm;
#pragma unroll
  for (int e = 0; e < 5; ++e) {
    double v = 0.0;
    if (hi) v += resb[cidx(m, i, j) + (long)e * m.ncell];
    if (lo) v -= resb[cidx(m, i - di, j - dj) + (long)e * m.ncell];
    hb[e] = v;
  }
}

// The 24 seeded passes of an interior face are split into groups of
// OFF_GROUP offsets, one kernel instantiation (and launch) per group: one
// thread evaluating all 24 lets the compiler interleave them far beyond the
// register file (hundreds of bytes of spills per lane).  Offsets [O0, O1) per
// instantiation.  At 512 x 384, groups of 4 (one of six spills) take
// 0.66 + 0.52 ms for the i- and j-face blocks, groups of 2 0.62 + 0.48 ms and
// single offsets (no spills, 24 launches) 0.52 + 0.43 ms.
constexpr int OFF_GROUP = 1;

// i-faces (i, j), i = i0, i0 + istep, ... <= im+1, rows j in [j0, j1]; O4 selects the
// compact viscous flux (rows j >= 3) over the 2nd-order one (rows 1, 2).
// The two extra offsets (3, 0), (-4, 0) only exist for the faces i = 1 and
// i = im+1 (extrapolated ghost gradients); elsewhere nothing is stored for
// them and the assembly does not read them.
template <bool O4, int MODE, int O0 = 0, int O1 = NOFF>
__global__ void __launch_bounds__(256) k_iface_blocks(Mesh m, Phys ph, Metric g, const double* __restrict__ w,
                                                      const double* __restrict__ prim,
                                                      const double* __restrict__ grad,
                                                      double* __restrict__ out, const double* __restrict__ resb,
                                                      int j0, int j1, int i0, int istep) {
  int bx, by;
  xcd_block(bx, by);
  const int i = i0 + (bx * (int)blockDim.x + (int)threadIdx.x) * istep;
  const int j = j0 + by * blockDim.y + threadIdx.y;
  if (i > m.im + 1 || j > j1) return;
  const long nf = (long)(m.im + 1) * m.jm;
  FaceSink<MODE> sink{out, nf, (i - 1) + (long)(m.im + 1) * (j - 1), {}};
  if constexpr (MODE == MODE_VJP) face_adjoint(m, resb, i, j, 1, 0, sink.hb);
  sfor<O0, O1>([&](auto O) {
    constexpr int o = decltype(O)::value;
    constexpr int SI = ioff_di(o), SJ = ioff_dj(o);
    if (o == 22 && i != 1) return;
    if (o == 23 && i != m.im + 1) return;
    auto a = seeded<SI, SJ>(m, ph, g, w, prim, grad, i, j);
    Dual<5> hn[5];
    iface<O4>(a, hn);
    sink.put(o, hn);
  });
}

// j-faces (i, j), i = i0, i0 + istep, ... <= im, one row variant per launch: ROW 0 = centred
// (rows j >= 4), 3, 2 = off-centred near-wall rows, 1 = the wall face.
// Extra offsets: (0, 3) only for row 2, (0, -4) only for row jm+1; the wall
// face depends on (0, 0) and (0, 1) only.
template <int ROW, int MODE, int O0 = 0, int O1 = NOFF>
__global__ void __launch_bounds__(256) k_jface_blocks(Mesh m, Phys ph, Metric g, const double* __restrict__ w,
                                                      const double* __restrict__ prim,
                                                      const double* __restrict__ grad,
                                                      double* __restrict__ out, const double* __restrict__ resb,
                                                      int j0, int j1, int i0, int istep) {
  int bx, by;
  xcd_block(bx, by);
  const int i = i0 + (bx * (int)blockDim.x + (int)threadIdx.x) * istep;
  const int j = j0 + by * blockDim.y + threadIdx.y;
  if (i > m.im || j > j1) return;
  const long nf = (long)m.im * (m.jm + 1);
  FaceSink<MODE> sink{out, nf, (i - 1) + (long)m.im * (j - 1), {}};
  if constexpr (MODE == MODE_VJP) face_adjoint(m, resb, i, j, 0, 1, sink.hb);
  sfor<O0, O1>([&](auto O) {
    constexpr int o = decltype(O)::value;
    constexpr int SI = joff_di(o), SJ = joff_dj(o);
    if constexpr (o == 22 && ROW != 2) return;
    else if constexpr (ROW == 1 && !(SI == 0 && (SJ == 0 || SJ == 1))) {
      // the wall face does not depend on this cell: store an explicit zero
      // block where the assembly may look for one
      Dual<5> z[5];
      sink.put(o, z);
      return;
    } else {
      if (o == 23 && !(ROW == 0 && j == m.jm + 1)) return;
      auto a = seeded<SI, SJ>(m, ph, g, w, prim, grad, i, j);
      Dual<5> hn[5];
      if constexpr (ROW == 0) jface<0, true>(a, hn);
      else if constexpr (ROW == 3) jface<3, true>(a, hn);
      else if constexpr (ROW == 2) jface<2, false>(a, hn);
      else jface_wall(a, hn);
      sink.put(o, hn);
    }
  });
}

}  // namespace bca
