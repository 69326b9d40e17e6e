// DNC5 residual of BROADCAST (srcfv/rhs/flux_num_dnc5.F90:7-226) as device
// code templated on the scalar type: T = double gives the residual, T = Dual<N>
// gives N tangent directions at once.
//
// Storage is the reference's: Fortran order with ghost width gh on every side.
// Cell planes are ni x nj (ni = im+2gh), face/node planes (ni+1) x (nj+1);
// every multi-component field is a stack of such planes (SoA), so a wavefront
// walking i reads 64 consecutive doubles per plane.
//
// The pipeline is four passes over HBM-resident planes:
//   prim   : every cell (ghosts too)  w -> u, v, w, T, p, mu, H
//   grad   : interior cells           4th-order velocity gradients (u, v)
//   ghost  : one extrapolated layer of gradients where faces read it
//   face   : every face               hn = Euler - dissipation - viscous
//   balance: interior cells           residu = -(d_i hn_i + d_j hn_j)
#pragma once
#include <hip/hip_runtime.h>
#include "dual.h"

namespace bca {

struct Mesh {
  int im, jm, gh, ni, nj;
  long ncell, nnode;
};

__host__ inline Mesh make_mesh(int im, int jm, int gh) {
  Mesh m;
  m.im = im; m.jm = jm; m.gh = gh;
  m.ni = im + 2 * gh; m.nj = jm + 2 * gh;
  m.ncell = (long)m.ni * m.nj;
  m.nnode = (long)(m.ni + 1) * (m.nj + 1);
  return m;
}

// XCD-aware block order for the stencil kernels' 2-D grids.  MI355X hands
// workgroups to its 8 XCDs in turn (workgroup b to XCD b % 8), each XCD with
// its own L2; in launch order, blocks that share stencil cells (the rows
// above and below) land on different XCDs and every L2 fetches its own
// copy.  This gives XCD x one contiguous run of blocks in row-major order,
// so the rows a block shares with its neighbours were fetched into the same
// L2 a moment before.  Returns the logical (bx, by) of this workgroup.
constexpr int NXCD = 8;
__device__ inline void xcd_block(int& bx, int& by) {
  const int nbx = gridDim.x, nblk = gridDim.x * gridDim.y;
  const int b = blockIdx.x + blockIdx.y * nbx;
  const int q = nblk / NXCD, r = nblk % NXCD;
  const int x = b % NXCD, s = b / NXCD;
  const int t = x * q + (x < r ? x : r) + s;
  bx = t % nbx;
  by = t / nbx;
}

// Fortran (i, j), lower bound 1-gh, to the flat storage offset
__host__ __device__ inline long cidx(const Mesh& m, int i, int j) {
  return (long)(i + m.gh - 1) + (long)m.ni * (j + m.gh - 1);
}
__host__ __device__ inline long nidx(const Mesh& m, int i, int j) {
  return (long)(i + m.gh - 1) + (long)(m.ni + 1) * (j + m.gh - 1);
}

struct Phys {
  double cp, cv, prandtl, gam, rgaz, cs, muref, tref, s_suth, k2, k4;
};

enum { P_U = 0, P_V, P_W, P_T, P_P, P_MU, P_H, NPRIM };
enum { G_UX = 0, G_UY, G_VX, G_VY, NGRAD };

// scheme coefficients (srcfv/rhs/flux_num_dnc5.F90:65-118, coefnearbnd_7p.F)
// of the 5th-order scheme; Coef3 holds those of the 3rd-order one
// (srcfv/rhs/flux_num_dnc3.F90:103-119): 2nd-order gradient, 4-point centred
// Euler flux and predictor.
struct Coef3 {
  static constexpr double b1 = 0.5;
  static constexpr double c1 = 7.0 / 12.0, c2 = -1.0 / 12.0;
  static constexpr double d1 = 3.0 / 12.0, d2 = 1.0 / 12.0;
};
// Orders 7 and 9 (srcfv/rhs/flux_num_dnc7.F90:74-90, coefnearbnd_9p.F;
// flux_num_dnc9.F90:106-128, coefnearbnd_11p.F), formed as the reference
// forms them: b(n) the centred gradient, c(n) the centred Euler flux and d(n)
// the predictor (n = 0 the innermost pair), w(row, r) the off-centred Euler
// flux of the near-wall face row `row` on cell rows 1 .. NW.
template <int ORD>
struct CoefHi;
template <>
struct CoefHi<7> {
  static constexpr int NB = 3, NC = 4, NW = 9, NROW = 3;
  static constexpr double b(int n) {
    return (n == 0 ? 45.0 : n == 1 ? -9.0 : 1.0) * (1.0 / 60.0);
  }
  static constexpr double c(int n) {
    return (n == 0 ? 533.0 : n == 1 ? -139.0 : n == 2 ? 29.0 : -3.0) * (1.0 / 840.0);
  }
  static constexpr double d(int n) { return (n == 0 ? 35.0 : n == 1 ? 21.0 : n == 2 ? 7.0 : 1.0) * (1.0 / 280.0); }
  static constexpr double w(int row, int r) {
    constexpr double k3[9] = {92.0, 1547.0, -1861.0, 2171.0, -1917.0, 1191.0, -489.0, 119.0, -13.0};
    constexpr double k5[9] = {-13.0, 209.0, 1079.0, -769.0, 533.0, -279.0, 99.0, -21.0, 2.0};
    constexpr double k7[9] = {2.0, -31.0, 281.0, 911.0, -517.0, 281.0, -111.0, 27.0, -3.0};
    return (row == 2 ? k3[r] : row == 3 ? k5[r] : k7[r]) * (1.0 / 840.0);
  }
};
template <>
struct CoefHi<9> {
  static constexpr int NB = 4, NC = 5, NW = 11, NROW = 4;
  static constexpr double b(int n) {
    return (n == 0 ? 672.0 : n == 1 ? -168.0 : n == 2 ? 32.0 : -3.0) * (1.0 / 840.0);
  }
  static constexpr double c(int n) {
    return (n == 0 ? 1627.0 : n == 1 ? -473.0 : n == 2 ? 127.0 : n == 3 ? -23.0 : 2.0) * (1.0 / 2520.0);
  }
  static constexpr double d(int n) {
    return n == 4 ? 1.0 / 1260.0 : (n == 0 ? 126.0 : n == 1 ? 84.0 : n == 2 ? 36.0 : 9.0) * (1.0 / 1260.0);
  }
  // coefnearbnd_11p.F: the 9/2 face's coefficients from the centred ones
  // minus a correction, each row below from the one above minus another
  static constexpr double w(int row, int r) {
    constexpr double a9[11] = {1.0 / 840.0, -1.0 / 63.0, 3.0 / 28.0, -4.0 / 7.0, -11.0 / 30.0, 6.0 / 5.0,
                               -0.5,        4.0 / 21.0,  -3.0 / 56.0, 1.0 / 105.0, -1.0 / 1260.0};
    constexpr double a7[11] = {-1.0 / 360.0, 1.0 / 24.0, -3.0 / 8.0, -319.0 / 420.0, 1.75, -21.0 / 20.0,
                               7.0 / 12.0,   -0.25,      0.075,      -1.0 / 72.0,    1.0 / 840.0};
    constexpr double a5[11] = {1.0 / 90.0, -2.0 / 9.0,  -341.0 / 280.0, 8.0 / 3.0,  -7.0 / 3.0, 28.0 / 15.0,
                               -7.0 / 6.0, 8.0 / 15.0,  -1.0 / 6.0,     2.0 / 63.0, -1.0 / 360.0};
    constexpr double a3[11] = {-0.1, -4609.0 / 2520.0, 4.5, -6.0, 7.0, -6.3, 21.0 * 0.2, -2.0, 9.0 / 14.0,
                               -0.125, 1.0 / 90.0};
    const double cc = r == 0 ? c(4) : r == 1 ? c(3) : r == 2 ? c(2) : r == 3 ? c(1) : r == 4 ? c(0)
                    : r == 5 ? c(0) : r == 6 ? c(1) : r == 7 ? c(2) : r == 8 ? c(3) : r == 9 ? c(4) : 0.0;
    const double w9 = r == 10 ? -a9[10] : cc - a9[r];
    const double w7 = w9 - a7[r], w5 = w7 - a5[r], w3 = w5 - a3[r];
    return row == 5 ? w9 : row == 4 ? w7 : row == 3 ? w5 : w3;
  }
};

struct Coef {
  static constexpr double b1 = 8.0 / 12.0, b2 = -1.0 / 12.0;            // 4th-order gradient
  static constexpr double c1 = 37.0 / 60.0, c2 = -8.0 / 60.0, c3 = 1.0 / 60.0;  // centred Euler
  static constexpr double d1 = 10.0 / 60.0, d2 = 5.0 / 60.0, d3 = 1.0 / 60.0;   // predictor
  static constexpr double r24 = 1.0 / 24.0;                               // face value, o4
  static constexpr double cross = (0.25 / 3.0) * 0.0625;                  // cross value, o4
  static constexpr double twothird = 2.0 / 3.0;
};

template <class T>
struct Planes {
  const T* w;      // 5 cell planes
  T* prim;         // NPRIM cell planes
  T* grad;         // NGRAD cell planes
  T* hn;           // 10 node planes: e + 5*dir
};

struct Metric {
  const double* nx;   // 2 node planes
  const double* ny;
  const double* vol;  // 1 cell plane
  const double* volf; // 2 cell planes
};

// The state (w, wd) of a tangent read straight from its two double arrays as
// Dual<1> values, in place of a packed Dual<1> copy: w[k] = (w[k], wd[k]).
struct SplitState {
  const double* v;
  const double* d;
  __device__ Dual<1> operator[](long k) const {
    Dual<1> x(v[k]);
    x.d[0] = d[k];
    return x;
  }
};

// ---------------------------------------------------------------- primitives
// WP: a pointer to planes of T, or a SplitState for T = Dual<1>
template <class T, class WP = const T*>
__device__ inline void primitives(const Mesh& m, const Phys& ph, WP w, T* prim, long c) {
  const double cvm1 = 1.0 / ph.cv;
  const double betas = ph.muref * (ph.tref + ph.cs) / (sqrt(ph.tref) * ph.tref);
  const long n = m.ncell;
  T ro = w[c];
  T rom1 = 1.0 / ro;
  T u = w[c + n] * rom1;
  T v = w[c + 2 * n] * rom1;
  T ww = w[c + 3 * n] * rom1;
  T ec = 0.5 * (u * u + v * v + ww * ww);
  T e = (w[c + 4 * n] - ec * ro) * rom1;
  T t = e * cvm1;
  T p = (ph.gam - 1.0) * ro * e;
  prim[c + P_U * n] = u;
  prim[c + P_V * n] = v;
  prim[c + P_W * n] = ww;
  prim[c + P_T * n] = t;
  prim[c + P_P * n] = p;
  prim[c + P_H * n] = (w[c + 4 * n] + p) * rom1;
  prim[c + P_MU * n] = betas / (t + ph.s_suth) * dsqrt(t) * t;
}

// wcopy (optional): the state as planes of T as well, for the kernels that
// read it by stencil (one load per value there, where a SplitState takes two)
template <class T, class WP = const T*>
__global__ void k_prim(Mesh m, Phys ph, WP w, T* __restrict__ prim, T* __restrict__ wcopy = nullptr) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= m.ncell) return;
  primitives<T, WP>(m, ph, w, prim, c);
  if (wcopy)
#pragma unroll
    for (int e = 0; e < 5; ++e) wcopy[c + e * m.ncell] = w[c + e * m.ncell];
}

// ---------------------------------------------------------------- gradients
// gradop_5pi/j.F (ORD 5), gradop_3pi/j.F (ORD 3), gradop_7pi/j.F, gradop_9pi/j.F
// + gradient.F + geom/dxdy.F
template <int ORD = 5, class T>
__device__ inline void gradient_at(const Mesh& m, const Metric& g, const T* prim, int i, int j,
                                   T out[4]) {
  const long n = m.ncell;
  const T* U = prim + P_U * n;
  const T* V = prim + P_V * n;
  auto di = [&](const T* f) -> T {
    if constexpr (ORD == 3) return Coef3::b1 * (f[cidx(m, i + 1, j)] - f[cidx(m, i - 1, j)]);
    else if constexpr (ORD == 5)
      return Coef::b1 * (f[cidx(m, i + 1, j)] - f[cidx(m, i - 1, j)]) +
             Coef::b2 * (f[cidx(m, i + 2, j)] - f[cidx(m, i - 2, j)]);
    else {
      T acc = CoefHi<ORD>::b(0) * (f[cidx(m, i + 1, j)] - f[cidx(m, i - 1, j)]);
#pragma unroll
      for (int n = 1; n < CoefHi<ORD>::NB; ++n)
        acc = acc + CoefHi<ORD>::b(n) * (f[cidx(m, i + n + 1, j)] - f[cidx(m, i - n - 1, j)]);
      return acc;
    }
  };
  auto dj = [&](const T* f) -> T {
    if constexpr (ORD == 3) return Coef3::b1 * (f[cidx(m, i, j + 1)] - f[cidx(m, i, j - 1)]);
    else if constexpr (ORD == 5)
      return Coef::b1 * (f[cidx(m, i, j + 1)] - f[cidx(m, i, j - 1)]) +
             Coef::b2 * (f[cidx(m, i, j + 2)] - f[cidx(m, i, j - 2)]);
    else {
      T acc = CoefHi<ORD>::b(0) * (f[cidx(m, i, j + 1)] - f[cidx(m, i, j - 1)]);
#pragma unroll
      for (int n = 1; n < CoefHi<ORD>::NB; ++n)
        acc = acc + CoefHi<ORD>::b(n) * (f[cidx(m, i, j + n + 1)] - f[cidx(m, i, j - n - 1)]);
      return acc;
    }
  };
  T gui = di(U), gvi = di(V), guj = dj(U), gvj = dj(V);
  const long nn = m.nnode;
  const double volm1 = 1.0 / g.vol[cidx(m, i, j)];
  const double dxm1 = 0.5 * (g.nx[nidx(m, i, j)] + g.nx[nidx(m, i + 1, j)]) * volm1;
  const double dxm2 = 0.5 * (g.nx[nidx(m, i, j) + nn] + g.nx[nidx(m, i, j + 1) + nn]) * volm1;
  const double dym1 = 0.5 * (g.ny[nidx(m, i, j)] + g.ny[nidx(m, i + 1, j)]) * volm1;
  const double dym2 = 0.5 * (g.ny[nidx(m, i, j) + nn] + g.ny[nidx(m, i, j + 1) + nn]) * volm1;
  out[G_UX] = dxm1 * gui + dxm2 * guj;
  out[G_UY] = dym1 * gui + dym2 * guj;
  out[G_VX] = dxm1 * gvi + dxm2 * gvj;
  out[G_VY] = dym1 * gvi + dym2 * gvj;
}

template <class T, int ORD = 5>
__global__ void k_grad(Mesh m, Metric g, const T* __restrict__ prim, T* __restrict__ grad) {
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  if (i > m.im || j > m.jm) return;
  T out[4];
  gradient_at<ORD>(m, g, prim, i, j, out);
  const long c = cidx(m, i, j);
#pragma unroll
  for (int k = 0; k < 4; ++k) grad[c + k * m.ncell] = out[k];
}

// gradveloingh.F: the faces that reach the residual read ghost gradients only
// in columns 0 and im+1 (rows 1..jm), in row jm+1 and -- the variant without
// wall rows, whose first j-faces are centred -- row 0 (columns 1..im); there
// the reference's two-pass linear extrapolation reduces to 2 g(1) - g(2).
template <class T>
__global__ void k_grad_ghost(Mesh m, T* __restrict__ grad) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const long n = m.ncell;
  if (t >= m.jm + m.im && t < m.jm + 2 * m.im) {
    const int i = 1 + t - m.jm - m.im;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      T* G = grad + k * n;
      G[cidx(m, i, 0)] = 2.0 * G[cidx(m, i, 1)] - G[cidx(m, i, 2)];
    }
    return;
  }
  if (t < m.jm) {
    const int j = 1 + t;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      T* G = grad + k * n;
      G[cidx(m, 0, j)] = 2.0 * G[cidx(m, 1, j)] - G[cidx(m, 2, j)];
      G[cidx(m, m.im + 1, j)] = 2.0 * G[cidx(m, m.im, j)] - G[cidx(m, m.im - 1, j)];
    }
  } else if (t < m.jm + m.im) {
    const int i = 1 + t - m.jm;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      T* G = grad + k * n;
      G[cidx(m, i, m.jm + 1)] = 2.0 * G[cidx(m, i, m.jm)] - G[cidx(m, i, m.jm - 1)];
    }
  }
}

// balanceloop.F
template <class T>
__device__ inline T balance_at(const Mesh& m, const T* hn, int i, int j, int e) {
  const long nn = m.nnode;
  const T* hi = hn + (long)e * nn;
  const T* hj = hn + (long)(5 + e) * nn;
  return -(hi[nidx(m, i + 1, j)] - hi[nidx(m, i, j)]) - (hj[nidx(m, i, j + 1)] - hj[nidx(m, i, j)]);
}

template <class T>
__global__ void k_balance(Mesh m, const T* __restrict__ hn, T* __restrict__ residu) {
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  if (i > m.im || j > m.jm) return;
  const long c = cidx(m, i, j);
#pragma unroll
  for (int e = 0; e < 5; ++e) residu[c + (long)e * m.ncell] = balance_at(m, hn, i, j, e);
}

// The balance of Dual<N> fluxes written as residu (values) and residud (the
// first tangent) on the interior (either may be null; ghosts untouched).
template <int N>
__global__ void k_balance_split(Mesh m, const Dual<N>* __restrict__ hn, double* __restrict__ residu,
                                double* __restrict__ residud) {
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  if (i > m.im || j > m.jm) return;
  const long c = cidx(m, i, j);
#pragma unroll
  for (int e = 0; e < 5; ++e) {
    const Dual<N> v = balance_at(m, hn, i, j, e);
    if (residu) residu[c + (long)e * m.ncell] = v.v;
    if (residud) residud[c + (long)e * m.ncell] = v.d[0];
  }
}

}  // namespace bca
