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

// ---------------------------------------------------------------- primitives
template <class T>
__device__ inline void primitives(const Mesh& m, const Phys& ph, const T* w, T* prim, long c) {
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

template <class T>
__global__ void k_prim(Mesh m, Phys ph, const T* __restrict__ w, T* __restrict__ prim) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c < m.ncell) primitives(m, ph, w, prim, c);
}

// ---------------------------------------------------------------- gradients
// gradop_5pi/j.F + gradient.F + geom/dxdy.F
template <class T>
__device__ inline void gradient_at(const Mesh& m, const Metric& g, const T* prim, int i, int j,
                                   T out[4]) {
  const long n = m.ncell;
  const T* U = prim + P_U * n;
  const T* V = prim + P_V * n;
  auto di = [&](const T* f) {
    return Coef::b1 * (f[cidx(m, i + 1, j)] - f[cidx(m, i - 1, j)]) +
           Coef::b2 * (f[cidx(m, i + 2, j)] - f[cidx(m, i - 2, j)]);
  };
  auto dj = [&](const T* f) {
    return Coef::b1 * (f[cidx(m, i, j + 1)] - f[cidx(m, i, j - 1)]) +
           Coef::b2 * (f[cidx(m, i, j + 2)] - f[cidx(m, i, j - 2)]);
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

template <class T>
__global__ void k_grad(Mesh m, Metric g, const T* __restrict__ prim, T* __restrict__ grad) {
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  if (i > m.im || j > m.jm) return;
  T out[4];
  gradient_at(m, g, prim, i, j, out);
  const long c = cidx(m, i, j);
#pragma unroll
  for (int k = 0; k < 4; ++k) grad[c + k * m.ncell] = out[k];
}

// gradveloingh.F: the faces that reach the residual read ghost gradients only
// in columns 0 and im+1 (rows 1..jm) and in row jm+1 (columns 1..im); there
// the reference's two-pass linear extrapolation reduces to 2 g(1) - g(2).
template <class T>
__global__ void k_grad_ghost(Mesh m, T* __restrict__ grad) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const long n = m.ncell;
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

// ---------------------------------------------------------------- face flux
// Everything one face needs, gathered from the planes.
template <class T>
struct FaceCtx {
  Mesh m;
  Phys ph;
  Metric g;
  const T* w;
  const T* prim;
  const T* grad;
  __device__ T W(int i, int j, int e) const { return w[cidx(m, i, j) + (long)e * m.ncell]; }
  __device__ T P(int k, int i, int j) const { return prim[cidx(m, i, j) + (long)k * m.ncell]; }
  __device__ T G(int k, int i, int j) const { return grad[cidx(m, i, j) + (long)k * m.ncell]; }
  __device__ double NX(int i, int j, int d) const { return g.nx[nidx(m, i, j) + (long)d * m.nnode]; }
  __device__ double NY(int i, int j, int d) const { return g.ny[nidx(m, i, j) + (long)d * m.nnode]; }
  __device__ double VOL(int i, int j) const { return g.vol[cidx(m, i, j)]; }
  __device__ double VOLF(int i, int j, int d) const { return g.volf[cidx(m, i, j) + (long)d * m.ncell]; }
  // Euler flux components of cell (i, j): F = (rho u, rho u^2 + p, rho u v, rho u w, rho u H)
  __device__ void fg(int i, int j, int e, T& f, T& gg) const {
    const T ru = W(i, j, 1), rv = W(i, j, 2);
    switch (e) {
      case 0: f = ru; gg = rv; break;
      case 1: { T u = P(P_U, i, j), p = P(P_P, i, j); f = ru * u + p; gg = rv * u; } break;
      case 2: { T v = P(P_V, i, j), p = P(P_P, i, j); f = ru * v; gg = rv * v + p; } break;
      case 3: { T ww = P(P_W, i, j); f = ru * ww; gg = rv * ww; } break;
      default: { T h = P(P_H, i, j); f = ru * h; gg = rv * h; } break;
    }
  }
};

// Newtonian stress and heat flux (tail of flux_visqueux_o*.F)
template <class T>
__device__ inline void stress(T mmu, T lambda, T uu, T vv, T ww, const T gx[4], const T gy[4],
                              T fv[5], T gv[5]) {
  // gx/gy: d/dx, d/dy of (u, v, w, T)
  fv[0] = T(0.0);
  gv[0] = T(0.0);
  fv[1] = Coef::twothird * mmu * (2.0 * gx[0] - gy[1]);
  fv[2] = mmu * (gy[0] + gx[1]);
  fv[3] = mmu * gx[2];
  fv[4] = lambda * gx[3] + uu * fv[1] + vv * fv[2] + ww * fv[3];
  gv[1] = mmu * (gy[0] + gx[1]);
  gv[2] = Coef::twothird * mmu * (-gx[0] + 2.0 * gy[1]);
  gv[3] = mmu * gy[2];
  gv[4] = lambda * gy[3] + uu * gv[1] + vv * gv[2] + ww * gv[3];
}

// Scalar Roe-weighted dissipation with the Jameson-Ducros sensor
// (spectralradius_i.F, ducrosfordnc_i.F, dissipation_ducros_i.F).
// Right cell (i, j), left cell (i - di, j - dj), face normal family d.
template <class T>
__device__ inline void dissipation(const FaceCtx<T>& c, int i, int j, int di, int dj, int d,
                                   const T pred[5], T diss[5]) {
  const int il = i - di, jl = j - dj;
  const double gr = c.ph.gam * c.ph.rgaz;
  T rr_ = c.W(i, j, 0), rl = c.W(il, jl, 0);
  T ur = c.W(i, j, 1) / rr_, vr = c.W(i, j, 2) / rr_;
  T ul = c.W(il, jl, 1) / rl, vl = c.W(il, jl, 2) / rl;
  T c2r = gr * c.P(P_T, i, j), c2l = gr * c.P(P_T, il, jl);
  T r = dsqrt(rr_ / rl);
  T wl = 1.0 / (1.0 + r);
  T wr = 1.0 - wl;
  T u = ul * wl + ur * wr;
  T v = vl * wl + vr * wr;
  T c2 = c2l * wl + c2r * wr;
  const double nxf = c.NX(i, j, d), nyf = c.NY(i, j, d);
  const double n2 = nxf * nxf + nyf * nyf;
  T rspec = dabs(nxf * u + nyf * v) + dsqrt(c2 * n2);
  // pressure sensor
  T pm2 = c.P(P_P, il - di, jl - dj), pm1 = c.P(P_P, il, jl), p0 = c.P(P_P, i, j),
    pp1 = c.P(P_P, i + di, j + dj);
  T ks1 = dabs(pm1 - 2.0 * p0 + pp1) / dabs(pm1 + 2.0 * p0 + pp1);
  T ks2 = dabs(pm2 - 2.0 * pm1 + p0) / dabs(pm2 + 2.0 * pm1 + p0);
  // Ducros and compression switches on both sides
  auto side = [&](int ii, int jj, T c2s, T& duc, T& sw) {
    T divu = c.G(G_UX, ii, jj) + c.G(G_VY, ii, jj);
    T rot = c.G(G_VX, ii, jj) - c.G(G_UY, ii, jj);
    T divu2 = divu * divu;
    duc = divu2 / (divu2 + rot * rot + 1e-15);
    sw = 0.5 * (1.0 - dtanh(2.5 + 10.0 * c.VOL(ii, jj) / (dsqrt(c2s * n2) + 1e-15) * divu));
  };
  T duc1, sw1, duc2, sw2;
  side(i, j, c2r, duc1, sw1);
  side(il, jl, c2l, duc2, sw2);
  T coef = dmax(ks1, ks2) * dmax(duc1, duc2) * dmax(sw1, sw2);
  T eps2 = c.ph.k2 * coef;
  T eps4 = dmax(T(0.0), c.ph.k4 - eps2 * 12.0);
#pragma unroll
  for (int e = 0; e < 5; ++e) {
    T half_jump = 0.5 * (c.W(i, j, e) - c.W(il, jl, e));
    diss[e] = rspec * (eps2 * half_jump + eps4 * pred[e]);
  }
}

// 6-point predictor difference across the face (predictor_7p_i.F)
template <class T>
__device__ inline void predictor(const FaceCtx<T>& c, int i, int j, int di, int dj, T pred[5]) {
#pragma unroll
  for (int e = 0; e < 5; ++e)
    pred[e] = -(Coef::d3 * c.W(i - 3 * di, j - 3 * dj, e)) + Coef::d2 * c.W(i - 2 * di, j - 2 * dj, e) -
              Coef::d1 * c.W(i - di, j - dj, e) + Coef::d1 * c.W(i, j, e) -
              Coef::d2 * c.W(i + di, j + dj, e) + Coef::d3 * c.W(i + 2 * di, j + 2 * dj, e);
}

template <class T>
__device__ inline void assemble(const FaceCtx<T>& c, int i, int j, int d, const T fx[5],
                                const T diss[5], const T fv[5], const T gv[5], T hn[5]) {
  const double s1 = c.NX(i, j, d), s2 = c.NY(i, j, d);
  const double sn = sqrt(s1 * s1 + s2 * s2);
  const double inv = 1.0 / sn;
  const double nxl = s1 * inv, nyl = s2 * inv;
  hn[0] = fx[0] - diss[0];
#pragma unroll
  for (int e = 1; e < 5; ++e) hn[e] = fx[e] - diss[e] - (fv[e] * nxl + gv[e] * nyl) * sn;
}

// i-face (i, j): between cells (i-1, j) and (i, j).  o4 selects the compact
// 4th-order viscous flux (flux_visqueux_o4_i.F) over the 2nd-order one
// (flux_visqueux_o2_i.F) used on rows j = 1, 2.
template <class T>
__device__ inline void iface_flux(const FaceCtx<T>& c, int i, int j, bool o4, T hn[5]) {
  T fx[5], pred[5], diss[5], fv[5], gv[5];
#pragma unroll
  for (int e = 0; e < 5; ++e) {
    T f0, g0, f1, g1, f2, g2, f3, g3, f4, g4, f5, g5;
    c.fg(i, j, e, f0, g0);
    c.fg(i - 1, j, e, f1, g1);
    c.fg(i + 1, j, e, f2, g2);
    c.fg(i - 2, j, e, f3, g3);
    c.fg(i + 2, j, e, f4, g4);
    c.fg(i - 3, j, e, f5, g5);
    fx[e] = (Coef::c1 * (f0 + f1) + Coef::c2 * (f2 + f3) + Coef::c3 * (f4 + f5)) * c.NX(i, j, 0) +
            (Coef::c1 * (g0 + g1) + Coef::c2 * (g2 + g3) + Coef::c3 * (g4 + g5)) * c.NY(i, j, 0);
  }
  predictor(c, i, j, 1, 0, pred);
  // contour of the staggered control volume around the face
  const double vf = c.VOLF(i, j, 0);
  const double nxN = 0.5 * (c.NX(i + 1, j, 0) + c.NX(i, j, 0));
  const double nxS = -(0.5 * (c.NX(i - 1, j, 0) + c.NX(i, j, 0)));
  const double nxO = 0.5 * (c.NX(i - 1, j + 1, 1) + c.NX(i, j + 1, 1));
  const double nxE = -(0.5 * (c.NX(i - 1, j, 1) + c.NX(i, j, 1)));
  const double nyN = 0.5 * (c.NY(i + 1, j, 0) + c.NY(i, j, 0));
  const double nyS = -(0.5 * (c.NY(i - 1, j, 0) + c.NY(i, j, 0)));
  const double nyO = 0.5 * (c.NY(i - 1, j + 1, 1) + c.NY(i, j + 1, 1));
  const double nyE = -(0.5 * (c.NY(i - 1, j, 1) + c.NY(i, j, 1)));
  T gx[4], gy[4];
  const int q[4] = {P_U, P_V, P_W, P_T};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int pk = q[k];
    T vN, vS, vE, vO;
    if (o4) {
      vN = Coef::r24 * (-c.P(pk, i + 1, j) + 26.0 * c.P(pk, i, j) - c.P(pk, i - 1, j));
      vS = Coef::r24 * (-c.P(pk, i, j) + 26.0 * c.P(pk, i - 1, j) - c.P(pk, i - 2, j));
      T rw[5];
#pragma unroll
      for (int s = 0; s < 5; ++s) {
        const int jj = j - 2 + s;
        rw[s] = -c.P(pk, i - 2, jj) + 9.0 * c.P(pk, i - 1, jj) + 9.0 * c.P(pk, i, jj) - c.P(pk, i + 1, jj);
      }
      vE = Coef::cross * (-rw[0] + 7.0 * rw[1] + 7.0 * rw[2] - rw[3]);
      vO = Coef::cross * (-rw[1] + 7.0 * rw[2] + 7.0 * rw[3] - rw[4]);
    } else {
      vN = c.P(pk, i, j);
      vS = c.P(pk, i - 1, j);
      vE = 0.25 * (c.P(pk, i, j) + c.P(pk, i, j - 1) + c.P(pk, i - 1, j) + c.P(pk, i - 1, j - 1));
      vO = 0.25 * (c.P(pk, i, j) + c.P(pk, i, j + 1) + c.P(pk, i - 1, j) + c.P(pk, i - 1, j + 1));
    }
    gx[k] = (vN * nxN + vS * nxS + vO * nxO + vE * nxE) * vf;
    gy[k] = (vN * nyN + vS * nyS + vO * nyO + vE * nyE) * vf;
  }
  T uu, vv, ww, mmu, lambda;
  const double cpp = c.ph.cp / c.ph.prandtl;
  if (o4) {
    auto mid = [&](int pk) {
      return 0.0625 * (-c.P(pk, i - 2, j) + 9.0 * c.P(pk, i - 1, j) + 9.0 * c.P(pk, i, j) - c.P(pk, i + 1, j));
    };
    uu = mid(P_U); vv = mid(P_V); ww = mid(P_W); mmu = mid(P_MU);
    lambda = mmu * cpp;
  } else {
    uu = 0.5 * (c.P(P_U, i, j) + c.P(P_U, i - 1, j));
    vv = 0.5 * (c.P(P_V, i, j) + c.P(P_V, i - 1, j));
    ww = 0.5 * (c.P(P_W, i, j) + c.P(P_W, i - 1, j));
    mmu = 0.5 * (c.P(P_MU, i, j) + c.P(P_MU, i - 1, j));
    lambda = mmu * cpp;
  }
  stress(mmu, lambda, uu, vv, ww, gx, gy, fv, gv);
  dissipation(c, i, j, 1, 0, 0, pred, diss);
  assemble(c, i, j, 0, fx, diss, fv, gv, hn);
}

// j-face (i, j): between cells (i, j-1) and (i, j).  Rows 1..3 carry the
// wall treatment of flux_num_dnc5.F90:161-220.
template <class T>
__device__ inline void jface_flux(const FaceCtx<T>& c, int i, int j, T hn[5]) {
  const double n2x = c.NX(i, j, 1), n2y = c.NY(i, j, 1);
  if (j == 1) {
    // adiabatic no-slip wall (fluxwall.F): extrapolated pressure, one-sided
    // velocity gradient against the zero wall velocity, no heat flux
    T pw = 1.125 * c.P(P_P, i, 1) - 0.125 * c.P(P_P, i, 2);
    T mmu = c.P(P_MU, i, 1);
    const double vf = c.VOLF(i, 1, 1);
    T u = c.P(P_U, i, 1), v = c.P(P_V, i, 1), ww = c.P(P_W, i, 1);
    T ux = 2.0 * u * n2x * vf, vx = 2.0 * v * n2x * vf, wx = 2.0 * ww * n2x * vf;
    T uy = 2.0 * u * n2y * vf, vy = 2.0 * v * n2y * vf, wy = 2.0 * ww * n2y * vf;
    T fu = Coef::twothird * mmu * (2.0 * ux - vy);
    T fvv = mmu * (uy + vx);
    T fw = mmu * wx;
    T gu = mmu * (uy + vx);
    T gvv = Coef::twothird * mmu * (-ux + 2.0 * vy);
    T gw = mmu * wy;
    hn[0] = T(0.0);
    hn[1] = pw * n2x - (fu * n2x + gu * n2y);
    hn[2] = pw * n2y - (fvv * n2x + gvv * n2y);
    hn[3] = -(fw * n2x + gw * n2y);
    hn[4] = T(0.0);
    return;
  }
  T fx[5], pred[5], diss[5], fv[5], gv[5];
  if (j <= 3) {
    // off-centred near-wall Euler fluxes (nearbndfluxes{3,5}demi_7p.F)
    const bool j2 = (j == 2);
    const double kk[5] = {j2 ? 12.0 / 60.0 : -3.0 / 60.0, j2 ? 77.0 / 60.0 : 27.0 / 60.0,
                          j2 ? -43.0 / 60.0 : 47.0 / 60.0, j2 ? 17.0 / 60.0 : -13.0 / 60.0,
                          j2 ? -3.0 / 60.0 : 2.0 / 60.0};
#pragma unroll
    for (int e = 0; e < 5; ++e) {
      T f[5], g[5];
#pragma unroll
      for (int s = 0; s < 5; ++s) c.fg(i, 1 + s, e, f[s], g[s]);
      fx[e] = (kk[0] * f[0] + kk[1] * f[1] + kk[2] * f[2] + kk[3] * f[3] + kk[4] * f[4]) * n2x +
              (kk[0] * g[0] + kk[1] * g[1] + kk[2] * g[2] + kk[3] * g[3] + kk[4] * g[4]) * n2y;
    }
  } else {
#pragma unroll
    for (int e = 0; e < 5; ++e) {
      T f0, g0, f1, g1, f2, g2, f3, g3, f4, g4, f5, g5;
      c.fg(i, j, e, f0, g0);
      c.fg(i, j - 1, e, f1, g1);
      c.fg(i, j + 1, e, f2, g2);
      c.fg(i, j - 2, e, f3, g3);
      c.fg(i, j + 2, e, f4, g4);
      c.fg(i, j - 3, e, f5, g5);
      fx[e] = (Coef::c1 * (f0 + f1) + Coef::c2 * (f2 + f3) + Coef::c3 * (f4 + f5)) * n2x +
              (Coef::c1 * (g0 + g1) + Coef::c2 * (g2 + g3) + Coef::c3 * (g4 + g5)) * n2y;
    }
  }
  predictor(c, i, j, 0, 1, pred);
  const bool o4 = (j >= 3);
  const double vf = c.VOLF(i, j, 1);
  const double nxN = 0.5 * (c.NX(i + 1, j - 1, 0) + c.NX(i + 1, j, 0));
  const double nxS = -(0.5 * (c.NX(i, j - 1, 0) + c.NX(i, j, 0)));
  const double nxO = 0.5 * (c.NX(i, j + 1, 1) + c.NX(i, j, 1));
  const double nxE = -(0.5 * (c.NX(i, j - 1, 1) + c.NX(i, j, 1)));
  const double nyN = 0.5 * (c.NY(i + 1, j - 1, 0) + c.NY(i + 1, j, 0));
  const double nyS = -(0.5 * (c.NY(i, j - 1, 0) + c.NY(i, j, 0)));
  const double nyO = 0.5 * (c.NY(i, j + 1, 1) + c.NY(i, j, 1));
  const double nyE = -(0.5 * (c.NY(i, j - 1, 1) + c.NY(i, j, 1)));
  T gx[4], gy[4];
  const int q[4] = {P_U, P_V, P_W, P_T};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int pk = q[k];
    T vN, vS, vE, vO;
    if (o4) {
      T cl[5];
#pragma unroll
      for (int s = 0; s < 5; ++s) {
        const int ii = i - 2 + s;
        cl[s] = -c.P(pk, ii, j - 2) + 9.0 * c.P(pk, ii, j - 1) + 9.0 * c.P(pk, ii, j) - c.P(pk, ii, j + 1);
      }
      vN = Coef::cross * (-cl[1] + 7.0 * cl[2] + 7.0 * cl[3] - cl[4]);
      vS = Coef::cross * (-cl[0] + 7.0 * cl[1] + 7.0 * cl[2] - cl[3]);
      vE = Coef::r24 * (-c.P(pk, i, j) + 26.0 * c.P(pk, i, j - 1) - c.P(pk, i, j - 2));
      vO = Coef::r24 * (-c.P(pk, i, j + 1) + 26.0 * c.P(pk, i, j) - c.P(pk, i, j - 1));
    } else {
      vN = 0.25 * (c.P(pk, i, j) + c.P(pk, i + 1, j) + c.P(pk, i, j - 1) + c.P(pk, i + 1, j - 1));
      vS = 0.25 * (c.P(pk, i, j) + c.P(pk, i - 1, j) + c.P(pk, i, j - 1) + c.P(pk, i - 1, j - 1));
      vO = c.P(pk, i, j);
      vE = c.P(pk, i, j - 1);
    }
    gx[k] = (vN * nxN + vS * nxS + vO * nxO + vE * nxE) * vf;
    gy[k] = (vN * nyN + vS * nyS + vO * nyO + vE * nyE) * vf;
  }
  T uu, vv, ww, mmu, lambda;
  const double cpp = c.ph.cp / c.ph.prandtl;
  if (o4) {
    auto mid = [&](int pk) {
      return 0.0625 * (-c.P(pk, i, j - 2) + 9.0 * c.P(pk, i, j - 1) + 9.0 * c.P(pk, i, j) - c.P(pk, i, j + 1));
    };
    uu = mid(P_U); vv = mid(P_V); ww = mid(P_W); mmu = mid(P_MU);
    lambda = mmu * cpp;
  } else {
    uu = 0.5 * (c.P(P_U, i, j) + c.P(P_U, i, j - 1));
    vv = 0.5 * (c.P(P_V, i, j) + c.P(P_V, i, j - 1));
    ww = 0.5 * (c.P(P_W, i, j) + c.P(P_W, i, j - 1));
    mmu = 0.5 * (c.P(P_MU, i, j) + c.P(P_MU, i, j - 1));
    lambda = mmu * cpp;
  }
  stress(mmu, lambda, uu, vv, ww, gx, gy, fv, gv);
  dissipation(c, i, j, 0, 1, 1, pred, diss);
  assemble(c, i, j, 1, fx, diss, fv, gv, hn);
}

// One thread per (i, j) in [1, im+1] x [1, jm+1]: the i-face (i, j) when
// j <= jm and the j-face (i, j) when i <= im (the faces the balance reads).
template <class T>
__global__ void __launch_bounds__(256) k_face(Mesh m, Phys ph, Metric g, const T* __restrict__ w,
                       const T* __restrict__ prim, const T* __restrict__ grad, T* __restrict__ hn) {
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  if (i > m.im + 1 || j > m.jm + 1) return;
  FaceCtx<T> c{m, ph, g, w, prim, grad};
  const long f = nidx(m, i, j);
  T h[5];
  if (j <= m.jm) {
    iface_flux(c, i, j, j >= 3, h);
#pragma unroll
    for (int e = 0; e < 5; ++e) hn[f + (long)e * m.nnode] = h[e];
  }
  if (i <= m.im) {
    jface_flux(c, i, j, h);
#pragma unroll
    for (int e = 0; e < 5; ++e) hn[f + (long)(5 + e) * m.nnode] = h[e];
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

}  // namespace bca
