// Face fluxes of the DNC5 scheme written against a stencil accessor whose
// cell offsets are compile-time constants.  One source serves three
// instantiations:
//
//   PlainAcc<double>    the residual                (every value a double)
//   PlainAcc<Dual<N>>   dense tangents (JVP)        (every value a Dual<N>)
//   SeededAcc<SI, SJ>   face-local Jacobian block   (only the cell at offset
//                       d hn(face) / d w(cell)       (SI, SJ) carries a 5-way
//                                                     tangent; all else double)
//
// With SeededAcc, C++ typing carries the tangent exactly along the data flow
// of the one seeded cell: a sum of six cell values becomes a Dual only at the
// seeded term and everything that does not depend on it stays plain FP64, so
// the 24 blocks of a face cost a fraction of 24 dense 5-way tangents.
//
// Formulas: srcfv/rhs/flux_num_dnc5.F90 (expanded srcfv/prepro/flux_num_dnc5.f90)
//   Euler        euler_o6_i.F / nearbndfluxes{3,5}demi_7p.F (coefnearbnd_7p.F)
//   dissipation  predictor_7p_i.F, spectralradius_i.F, ducrosfordnc_i.F
//   viscous      flux_visqueux_o4_i.F, flux_visqueux_o2_i.F, fluxwall.F
//   assembly     fluxnumassembly_i.F
#pragma once
#include <type_traits>

#include "scheme.h"

namespace bca {

// ------------------------------------------------------------------ typing
template <class A, class B> struct Prom;
template <> struct Prom<double, double> { using type = double; };
template <int N> struct Prom<Dual<N>, Dual<N>> { using type = Dual<N>; };
template <int N> struct Prom<double, Dual<N>> { using type = Dual<N>; };
template <int N> struct Prom<Dual<N>, double> { using type = Dual<N>; };
template <class A, class B> using prom_t = typename Prom<std::decay_t<A>, std::decay_t<B>>::type;

template <class A, class B>
__device__ inline prom_t<A, B> pmax(const A& a, const B& b) {
  using R = prom_t<A, B>;
  return val(a) < val(b) ? R(b) : R(a);
}

template <int B, int E, class F>
__device__ inline void sfor(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    sfor<B + 1, E>(f);
  }
}

// primitives of one cell from its 5 conservative values (primvisc.F)
template <class T>
__device__ inline void prim_of(const Phys& ph, const T& ro, const T& ru, const T& rv, const T& rw,
                               const T& re, T out[NPRIM]) {
  const double cvm1 = 1.0 / ph.cv;
  const double betas = ph.muref * (ph.tref + ph.cs) / (sqrt(ph.tref) * ph.tref);
  T rom1 = 1.0 / ro;
  T u = ru * rom1, v = rv * rom1, ww = rw * rom1;
  T ec = 0.5 * (u * u + v * v + ww * ww);
  T e = (re - ec * ro) * rom1;
  T t = e * cvm1;
  T p = (ph.gam - 1.0) * ro * e;
  out[P_U] = u;
  out[P_V] = v;
  out[P_W] = ww;
  out[P_T] = t;
  out[P_P] = p;
  out[P_H] = (re + p) * rom1;
  out[P_MU] = betas / (t + ph.s_suth) * dsqrt(t) * t;
}

// ------------------------------------------------------------ accessors
// Common part: position of the face's right cell (i, j) and metrics.
struct AccBase {
  Mesh m;
  Phys ph;
  Metric g;
  int i, j;
  __device__ double NX(int di, int dj, int d) const { return g.nx[nidx(m, i + di, j + dj) + (long)d * m.nnode]; }
  __device__ double NY(int di, int dj, int d) const { return g.ny[nidx(m, i + di, j + dj) + (long)d * m.nnode]; }
  __device__ double VOL(int di, int dj) const { return g.vol[cidx(m, i + di, j + dj)]; }
  __device__ double VOLF(int di, int dj, int d) const { return g.volf[cidx(m, i + di, j + dj) + (long)d * m.ncell]; }
};

// Every cell value read from planes of T (T = double or Dual<N>).
template <class T>
struct PlainAcc : AccBase {
  using Out = T;
  template <int DI, int DJ> using CT = T;
  template <int DI, int DJ> using GT = T;
  const T* w;
  const T* prim;
  const T* grad;
  template <int DI, int DJ> __device__ T W(int e) const { return w[cidx(m, i + DI, j + DJ) + (long)e * m.ncell]; }
  template <int DI, int DJ> __device__ T P(int k) const { return prim[cidx(m, i + DI, j + DJ) + (long)k * m.ncell]; }
  template <int DI, int DJ> __device__ T G(int k) const { return grad[cidx(m, i + DI, j + DJ) + (long)k * m.ncell]; }
};

// 4th-order gradient (gradop_5pi/j.F + gradient.F) at the cell at offset
// (GI, GJ), computed from the accessor's primitives so that it is typed by
// whether the seeded cell is in its stencil.  k: G_UX, G_UY, G_VX, G_VY.
template <int GI, int GJ, class A>
__device__ inline auto grad_direct(const A& a, int k) {
  const int q = (k < 2) ? P_U : P_V;
  auto di = Coef::b1 * (a.template P<GI + 1, GJ>(q) - a.template P<GI - 1, GJ>(q)) +
            Coef::b2 * (a.template P<GI + 2, GJ>(q) - a.template P<GI - 2, GJ>(q));
  auto dj = Coef::b1 * (a.template P<GI, GJ + 1>(q) - a.template P<GI, GJ - 1>(q)) +
            Coef::b2 * (a.template P<GI, GJ + 2>(q) - a.template P<GI, GJ - 2>(q));
  const double volm1 = 1.0 / a.VOL(GI, GJ);
  const bool xd = (k == G_UX || k == G_VX);
  const double m1 = xd ? 0.5 * (a.NX(GI, GJ, 0) + a.NX(GI + 1, GJ, 0)) * volm1
                       : 0.5 * (a.NY(GI, GJ, 0) + a.NY(GI + 1, GJ, 0)) * volm1;
  const double m2 = xd ? 0.5 * (a.NX(GI, GJ, 1) + a.NX(GI, GJ + 1, 1)) * volm1
                       : 0.5 * (a.NY(GI, GJ, 1) + a.NY(GI, GJ + 1, 1)) * volm1;
  return m1 * di + m2 * dj;
}

__host__ __device__ constexpr bool in_grad_stencil(int dx, int dy) {
  return (dy == 0 && (dx == 1 || dx == -1 || dx == 2 || dx == -2)) ||
         (dx == 0 && (dy == 1 || dy == -1 || dy == 2 || dy == -2));
}
// can the gradient at offset (GI, GJ) -- directly, or through the one-layer
// ghost extrapolation 2 g(a) - g(b) of gradveloingh.F -- depend on (SI, SJ)?
__host__ __device__ constexpr bool grad_reach(int SI, int SJ, int GI, int GJ) {
  return in_grad_stencil(SI - GI, SJ - GJ) || in_grad_stencil(SI - GI - 1, SJ - GJ) ||
         in_grad_stencil(SI - GI - 2, SJ - GJ) || in_grad_stencil(SI - GI + 1, SJ - GJ) ||
         in_grad_stencil(SI - GI + 2, SJ - GJ) || in_grad_stencil(SI - GI, SJ - GJ + 1) ||
         in_grad_stencil(SI - GI, SJ - GJ + 2);
}

// Unit 5-way tangent on the cell at offset (SI, SJ) (the 5 conservative
// variables of that cell), FP64 everywhere else.
template <int SI, int SJ>
struct SeededAcc : AccBase {
  using Out = Dual<5>;
  template <int DI, int DJ> using CT = std::conditional_t<(DI == SI && DJ == SJ), Dual<5>, double>;
  template <int DI, int DJ> using GT = std::conditional_t<grad_reach(SI, SJ, DI, DJ), Dual<5>, double>;
  const double* w;
  const double* prim;
  const double* grad;
  Dual<5> sw[5];      // seeded cell state
  Dual<5> sp[NPRIM];  // its primitives

  __device__ void seed() {
    const long c = cidx(m, i + SI, j + SJ);
#pragma unroll
    for (int e = 0; e < 5; ++e) {
      sw[e] = Dual<5>(w[c + (long)e * m.ncell]);
      sw[e].d[e] = 1.0;
    }
    prim_of(ph, sw[0], sw[1], sw[2], sw[3], sw[4], sp);
  }
  template <int DI, int DJ> __device__ CT<DI, DJ> W(int e) const {
    if constexpr (DI == SI && DJ == SJ) return sw[e];
    else return w[cidx(m, i + DI, j + DJ) + (long)e * m.ncell];
  }
  template <int DI, int DJ> __device__ CT<DI, DJ> P(int k) const {
    if constexpr (DI == SI && DJ == SJ) return sp[k];
    else return prim[cidx(m, i + DI, j + DJ) + (long)k * m.ncell];
  }
  template <int DI, int DJ> __device__ GT<DI, DJ> G(int k) const {
    using R = GT<DI, DJ>;
    if constexpr (!grad_reach(SI, SJ, DI, DJ)) {
      return grad[cidx(m, i + DI, j + DJ) + (long)k * m.ncell];
    } else {
      const int gi = i + DI, gj = j + DJ;
      if (gi == 0) return R(2.0 * grad_direct<DI + 1, DJ>(*this, k) - grad_direct<DI + 2, DJ>(*this, k));
      if (gi == m.im + 1) return R(2.0 * grad_direct<DI - 1, DJ>(*this, k) - grad_direct<DI - 2, DJ>(*this, k));
      if (gj == m.jm + 1) return R(2.0 * grad_direct<DI, DJ - 1>(*this, k) - grad_direct<DI, DJ - 2>(*this, k));
      return R(grad_direct<DI, DJ>(*this, k));
    }
  }
};

// ------------------------------------------------------------ building blocks
// Euler flux components of the cell at offset (DI, DJ): x- and y-direction
template <int DI, int DJ, class A>
__device__ inline auto efx(const A& a, int e) {
  auto ru = a.template W<DI, DJ>(1);
  switch (e) {
    case 0: return ru;
    case 1: return ru * a.template P<DI, DJ>(P_U) + a.template P<DI, DJ>(P_P);
    case 2: return ru * a.template P<DI, DJ>(P_V);
    case 3: return ru * a.template P<DI, DJ>(P_W);
    default: return ru * a.template P<DI, DJ>(P_H);
  }
}
template <int DI, int DJ, class A>
__device__ inline auto efy(const A& a, int e) {
  auto rv = a.template W<DI, DJ>(2);
  switch (e) {
    case 0: return rv;
    case 1: return rv * a.template P<DI, DJ>(P_U);
    case 2: return rv * a.template P<DI, DJ>(P_V) + a.template P<DI, DJ>(P_P);
    case 3: return rv * a.template P<DI, DJ>(P_W);
    default: return rv * a.template P<DI, DJ>(P_H);
  }
}

// Dissipation scalars of a face whose right cell is offset (0, 0) and left
// cell (-DX, -DY): Roe-weighted spectral radius and eps2/eps4 of the sensor.
template <int DX, int DY, class A>
__device__ inline auto diss_scalars(const A& a, int d) {
  const double gr = a.ph.gam * a.ph.rgaz;
  auto rr_ = a.template W<0, 0>(0);
  auto rl = a.template W<-DX, -DY>(0);
  auto ur = a.template W<0, 0>(1) / rr_, vr = a.template W<0, 0>(2) / rr_;
  auto ul = a.template W<-DX, -DY>(1) / rl, vl = a.template W<-DX, -DY>(2) / rl;
  auto c2r = gr * a.template P<0, 0>(P_T);
  auto c2l = gr * a.template P<-DX, -DY>(P_T);
  auto r = dsqrt(rr_ / rl);
  auto wl = 1.0 / (1.0 + r);
  auto wr = 1.0 - wl;
  auto u = ul * wl + ur * wr;
  auto v = vl * wl + vr * wr;
  auto c2 = c2l * wl + c2r * wr;
  const double nxf = a.NX(0, 0, d), nyf = a.NY(0, 0, d);
  const double n2 = nxf * nxf + nyf * nyf;
  auto rspec = dabs(nxf * u + nyf * v) + dsqrt(c2 * n2);
  using R = prom_t<decltype(rspec), prom_t<decltype(a.template P<-2 * DX, -2 * DY>(P_P)),
                                           decltype(a.template P<DX, DY>(P_P))>>;
  using RG = prom_t<decltype(a.template G<0, 0>(0)), decltype(a.template G<-DX, -DY>(0))>;
  using RS = prom_t<R, RG>;
  RS eps2(0.0);
  if (a.ph.k2 != 0.0) {
    auto pm2 = a.template P<-2 * DX, -2 * DY>(P_P);
    auto pm1 = a.template P<-DX, -DY>(P_P);
    auto p0 = a.template P<0, 0>(P_P);
    auto pp1 = a.template P<DX, DY>(P_P);
    auto ks1 = dabs(pm1 - 2.0 * p0 + pp1) / dabs(pm1 + 2.0 * p0 + pp1);
    auto ks2 = dabs(pm2 - 2.0 * pm1 + p0) / dabs(pm2 + 2.0 * pm1 + p0);
    auto divr = a.template G<0, 0>(G_UX) + a.template G<0, 0>(G_VY);
    auto rotr = a.template G<0, 0>(G_VX) - a.template G<0, 0>(G_UY);
    auto div2r = divr * divr;
    auto duc1 = div2r / (div2r + rotr * rotr + 1e-15);
    auto sw1 = 0.5 * (1.0 - dtanh(2.5 + 10.0 * a.VOL(0, 0) / (dsqrt(c2r * n2) + 1e-15) * divr));
    auto divl = a.template G<-DX, -DY>(G_UX) + a.template G<-DX, -DY>(G_VY);
    auto rotl = a.template G<-DX, -DY>(G_VX) - a.template G<-DX, -DY>(G_UY);
    auto div2l = divl * divl;
    auto duc2 = div2l / (div2l + rotl * rotl + 1e-15);
    auto sw2 = 0.5 * (1.0 - dtanh(2.5 + 10.0 * a.VOL(-DX, -DY) / (dsqrt(c2l * n2) + 1e-15) * divl));
    eps2 = RS(a.ph.k2 * (pmax(ks1, ks2) * pmax(duc1, duc2) * pmax(sw1, sw2)));
  }
  auto eps4 = pmax(0.0, a.ph.k4 - eps2 * 12.0);
  struct S {
    decltype(rspec) rspec;
    RS eps2;
    decltype(eps4) eps4;
  };
  return S{rspec, eps2, eps4};
}

// Newtonian stress and heat flux from the face values and face gradients
// gx/gy of (u, v, w, T); returned as (fv1..fv4, gv1..gv4).
template <class MU, class LA, class UU, class VV, class WW, class GX0, class GX1, class GX2, class GX3,
          class GY0, class GY1, class GY2, class GY3>
__device__ inline auto stress_of(const MU& mmu, const LA& lambda, const UU& uu, const VV& vv,
                                 const WW& ww, const GX0& ux, const GX1& vx, const GX2& wx,
                                 const GX3& tx, const GY0& uy, const GY1& vy, const GY2& wy,
                                 const GY3& ty) {
  auto fv1 = Coef::twothird * mmu * (2.0 * ux - vy);
  auto fv2 = mmu * (uy + vx);
  auto fv3 = mmu * wx;
  auto fv4 = lambda * tx + uu * fv1 + vv * fv2 + ww * fv3;
  auto gv1 = mmu * (uy + vx);
  auto gv2 = Coef::twothird * mmu * (-ux + 2.0 * vy);
  auto gv3 = mmu * wy;
  auto gv4 = lambda * ty + uu * gv1 + vv * gv2 + ww * gv3;
  struct V {
    decltype(fv1) f1; decltype(fv2) f2; decltype(fv3) f3; decltype(fv4) f4;
    decltype(gv1) g1; decltype(gv2) g2; decltype(gv3) g3; decltype(gv4) g4;
  };
  return V{fv1, fv2, fv3, fv4, gv1, gv2, gv3, gv4};
}

template <int E, class V>
__device__ inline auto vis_f(const V& v) {
  if constexpr (E == 1) return v.f1;
  else if constexpr (E == 2) return v.f2;
  else if constexpr (E == 3) return v.f3;
  else return v.f4;
}
template <int E, class V>
__device__ inline auto vis_g(const V& v) {
  if constexpr (E == 1) return v.g1;
  else if constexpr (E == 2) return v.g2;
  else if constexpr (E == 3) return v.g3;
  else return v.g4;
}

// contour-integral face gradient of primitive k from the four staggered values
struct Contour {
  double nxN, nxS, nxO, nxE, nyN, nyS, nyO, nyE, vf;
};
template <class VN, class VS, class VO, class VE>
__device__ inline auto cgx(const Contour& c, const VN& vN, const VS& vS, const VO& vO, const VE& vE) {
  return (vN * c.nxN + vS * c.nxS + vO * c.nxO + vE * c.nxE) * c.vf;
}
template <class VN, class VS, class VO, class VE>
__device__ inline auto cgy(const Contour& c, const VN& vN, const VS& vS, const VO& vO, const VE& vE) {
  return (vN * c.nyN + vS * c.nyS + vO * c.nyO + vE * c.nyE) * c.vf;
}

// ---------------------------------------------------------------- i-faces
template <class A>
__device__ inline Contour contour_i(const A& a) {
  Contour c;
  c.vf = a.VOLF(0, 0, 0);
  c.nxN = 0.5 * (a.NX(1, 0, 0) + a.NX(0, 0, 0));
  c.nxS = -(0.5 * (a.NX(-1, 0, 0) + a.NX(0, 0, 0)));
  c.nxO = 0.5 * (a.NX(-1, 1, 1) + a.NX(0, 1, 1));
  c.nxE = -(0.5 * (a.NX(-1, 0, 1) + a.NX(0, 0, 1)));
  c.nyN = 0.5 * (a.NY(1, 0, 0) + a.NY(0, 0, 0));
  c.nyS = -(0.5 * (a.NY(-1, 0, 0) + a.NY(0, 0, 0)));
  c.nyO = 0.5 * (a.NY(-1, 1, 1) + a.NY(0, 1, 1));
  c.nyE = -(0.5 * (a.NY(-1, 0, 1) + a.NY(0, 0, 1)));
  return c;
}

// staggered values of primitive k around the i-face, compact 4th order
template <class A>
__device__ inline auto ivals_o4(const A& a, int k) {
  auto row = [&](auto JJ) {
    constexpr int jj = decltype(JJ)::value;
    return -a.template P<-2, jj>(k) + 9.0 * a.template P<-1, jj>(k) + 9.0 * a.template P<0, jj>(k) -
           a.template P<1, jj>(k);
  };
  auto r0 = row(std::integral_constant<int, -2>{});
  auto r1 = row(std::integral_constant<int, -1>{});
  auto r2 = row(std::integral_constant<int, 0>{});
  auto r3 = row(std::integral_constant<int, 1>{});
  auto r4 = row(std::integral_constant<int, 2>{});
  auto vN = Coef::r24 * (-a.template P<1, 0>(k) + 26.0 * a.template P<0, 0>(k) - a.template P<-1, 0>(k));
  auto vS = Coef::r24 * (-a.template P<0, 0>(k) + 26.0 * a.template P<-1, 0>(k) - a.template P<-2, 0>(k));
  auto vE = Coef::cross * (-r0 + 7.0 * r1 + 7.0 * r2 - r3);
  auto vO = Coef::cross * (-r1 + 7.0 * r2 + 7.0 * r3 - r4);
  struct Q { decltype(vN) N; decltype(vS) S; decltype(vO) O; decltype(vE) E; };
  return Q{vN, vS, vO, vE};
}
template <class A>
__device__ inline auto ivals_o2(const A& a, int k) {
  auto vN = a.template P<0, 0>(k);
  auto vS = a.template P<-1, 0>(k);
  auto vE = 0.25 * (a.template P<0, 0>(k) + a.template P<0, -1>(k) + a.template P<-1, 0>(k) +
                    a.template P<-1, -1>(k));
  auto vO = 0.25 * (a.template P<0, 0>(k) + a.template P<0, 1>(k) + a.template P<-1, 0>(k) +
                    a.template P<-1, 1>(k));
  struct Q { decltype(vN) N; decltype(vS) S; decltype(vO) O; decltype(vE) E; };
  return Q{vN, vS, vO, vE};
}

template <bool O4, class A>
__device__ inline auto iface_viscous(const A& a) {
  const Contour c = contour_i(a);
  auto q = [&](int k) {
    if constexpr (O4) return ivals_o4(a, k);
    else return ivals_o2(a, k);
  };
  auto qu = q(P_U), qv = q(P_V), qw = q(P_W), qt = q(P_T);
  auto ux = cgx(c, qu.N, qu.S, qu.O, qu.E), uy = cgy(c, qu.N, qu.S, qu.O, qu.E);
  auto vx = cgx(c, qv.N, qv.S, qv.O, qv.E), vy = cgy(c, qv.N, qv.S, qv.O, qv.E);
  auto wx = cgx(c, qw.N, qw.S, qw.O, qw.E), wy = cgy(c, qw.N, qw.S, qw.O, qw.E);
  auto tx = cgx(c, qt.N, qt.S, qt.O, qt.E), ty = cgy(c, qt.N, qt.S, qt.O, qt.E);
  const double cpp = a.ph.cp / a.ph.prandtl;
  if constexpr (O4) {
    auto mid = [&](int k) {
      return 0.0625 * (-a.template P<-2, 0>(k) + 9.0 * a.template P<-1, 0>(k) + 9.0 * a.template P<0, 0>(k) -
                       a.template P<1, 0>(k));
    };
    auto mmu = mid(P_MU);
    return stress_of(mmu, mmu * cpp, mid(P_U), mid(P_V), mid(P_W), ux, vx, wx, tx, uy, vy, wy, ty);
  } else {
    auto half = [&](int k) { return 0.5 * (a.template P<0, 0>(k) + a.template P<-1, 0>(k)); };
    auto mmu = half(P_MU);
    return stress_of(mmu, mmu * cpp, half(P_U), half(P_V), half(P_W), ux, vx, wx, tx, uy, vy, wy, ty);
  }
}

// centred Euler flux and predictor of a face whose right cell is offset
// (0, 0) and left cell (-DX, -DY): 6 points for ORD 5 (euler_o6_i.F,
// predictor_7p_i.F), 4 points for ORD 3 (euler_o4_i.F, predictor_5p_i.F)
template <int ORD, int DX, int DY, class A>
__device__ inline auto euler_centred(const A& a, int e, double nxf, double nyf) {
  if constexpr (ORD == 7)
    return (CoefHi<7>::c(0) * (efx<0, 0>(a, e) + efx<-DX, -DY>(a, e)) +
            CoefHi<7>::c(1) * (efx<DX, DY>(a, e) + efx<-2 * DX, -2 * DY>(a, e)) +
            CoefHi<7>::c(2) * (efx<2 * DX, 2 * DY>(a, e) + efx<-3 * DX, -3 * DY>(a, e)) +
            CoefHi<7>::c(3) * (efx<3 * DX, 3 * DY>(a, e) + efx<-4 * DX, -4 * DY>(a, e))) * nxf +
           (CoefHi<7>::c(0) * (efy<0, 0>(a, e) + efy<-DX, -DY>(a, e)) +
            CoefHi<7>::c(1) * (efy<DX, DY>(a, e) + efy<-2 * DX, -2 * DY>(a, e)) +
            CoefHi<7>::c(2) * (efy<2 * DX, 2 * DY>(a, e) + efy<-3 * DX, -3 * DY>(a, e)) +
            CoefHi<7>::c(3) * (efy<3 * DX, 3 * DY>(a, e) + efy<-4 * DX, -4 * DY>(a, e))) * nyf;
  else if constexpr (ORD == 9)
    return (CoefHi<9>::c(0) * (efx<0, 0>(a, e) + efx<-DX, -DY>(a, e)) +
            CoefHi<9>::c(1) * (efx<DX, DY>(a, e) + efx<-2 * DX, -2 * DY>(a, e)) +
            CoefHi<9>::c(2) * (efx<2 * DX, 2 * DY>(a, e) + efx<-3 * DX, -3 * DY>(a, e)) +
            CoefHi<9>::c(3) * (efx<3 * DX, 3 * DY>(a, e) + efx<-4 * DX, -4 * DY>(a, e)) +
            CoefHi<9>::c(4) * (efx<4 * DX, 4 * DY>(a, e) + efx<-5 * DX, -5 * DY>(a, e))) * nxf +
           (CoefHi<9>::c(0) * (efy<0, 0>(a, e) + efy<-DX, -DY>(a, e)) +
            CoefHi<9>::c(1) * (efy<DX, DY>(a, e) + efy<-2 * DX, -2 * DY>(a, e)) +
            CoefHi<9>::c(2) * (efy<2 * DX, 2 * DY>(a, e) + efy<-3 * DX, -3 * DY>(a, e)) +
            CoefHi<9>::c(3) * (efy<3 * DX, 3 * DY>(a, e) + efy<-4 * DX, -4 * DY>(a, e)) +
            CoefHi<9>::c(4) * (efy<4 * DX, 4 * DY>(a, e) + efy<-5 * DX, -5 * DY>(a, e))) * nyf;
  else if constexpr (ORD == 3)
    return (Coef3::c1 * (efx<0, 0>(a, e) + efx<-DX, -DY>(a, e)) +
            Coef3::c2 * (efx<DX, DY>(a, e) + efx<-2 * DX, -2 * DY>(a, e))) * nxf +
           (Coef3::c1 * (efy<0, 0>(a, e) + efy<-DX, -DY>(a, e)) +
            Coef3::c2 * (efy<DX, DY>(a, e) + efy<-2 * DX, -2 * DY>(a, e))) * nyf;
  else
    return (Coef::c1 * (efx<0, 0>(a, e) + efx<-DX, -DY>(a, e)) +
            Coef::c2 * (efx<DX, DY>(a, e) + efx<-2 * DX, -2 * DY>(a, e)) +
            Coef::c3 * (efx<2 * DX, 2 * DY>(a, e) + efx<-3 * DX, -3 * DY>(a, e))) * nxf +
           (Coef::c1 * (efy<0, 0>(a, e) + efy<-DX, -DY>(a, e)) +
            Coef::c2 * (efy<DX, DY>(a, e) + efy<-2 * DX, -2 * DY>(a, e)) +
            Coef::c3 * (efy<2 * DX, 2 * DY>(a, e) + efy<-3 * DX, -3 * DY>(a, e))) * nyf;
}
template <int ORD, int DX, int DY, class A>
__device__ inline auto predictor(const A& a, int e) {
  // predictor_9p_i.F, predictor_11p_i.F: alternating signs from the
  // outermost left cell to the outermost right one
  if constexpr (ORD == 7)
    return CoefHi<7>::d(3) * a.template W<-4 * DX, -4 * DY>(e) - CoefHi<7>::d(2) * a.template W<-3 * DX, -3 * DY>(e) +
           CoefHi<7>::d(1) * a.template W<-2 * DX, -2 * DY>(e) - CoefHi<7>::d(0) * a.template W<-DX, -DY>(e) +
           CoefHi<7>::d(0) * a.template W<0, 0>(e) - CoefHi<7>::d(1) * a.template W<DX, DY>(e) +
           CoefHi<7>::d(2) * a.template W<2 * DX, 2 * DY>(e) - CoefHi<7>::d(3) * a.template W<3 * DX, 3 * DY>(e);
  else if constexpr (ORD == 9)
    return -(CoefHi<9>::d(4) * a.template W<-5 * DX, -5 * DY>(e)) +
           CoefHi<9>::d(3) * a.template W<-4 * DX, -4 * DY>(e) - CoefHi<9>::d(2) * a.template W<-3 * DX, -3 * DY>(e) +
           CoefHi<9>::d(1) * a.template W<-2 * DX, -2 * DY>(e) - CoefHi<9>::d(0) * a.template W<-DX, -DY>(e) +
           CoefHi<9>::d(0) * a.template W<0, 0>(e) - CoefHi<9>::d(1) * a.template W<DX, DY>(e) +
           CoefHi<9>::d(2) * a.template W<2 * DX, 2 * DY>(e) - CoefHi<9>::d(3) * a.template W<3 * DX, 3 * DY>(e) +
           CoefHi<9>::d(4) * a.template W<4 * DX, 4 * DY>(e);
  else if constexpr (ORD == 3)
    return Coef3::d2 * a.template W<-2 * DX, -2 * DY>(e) - Coef3::d1 * a.template W<-DX, -DY>(e) +
           Coef3::d1 * a.template W<0, 0>(e) - Coef3::d2 * a.template W<DX, DY>(e);
  else
    return -(Coef::d3 * a.template W<-3 * DX, -3 * DY>(e)) + Coef::d2 * a.template W<-2 * DX, -2 * DY>(e) -
           Coef::d1 * a.template W<-DX, -DY>(e) + Coef::d1 * a.template W<0, 0>(e) -
           Coef::d2 * a.template W<DX, DY>(e) + Coef::d3 * a.template W<2 * DX, 2 * DY>(e);
}

// hn = Euler - dissipation - viscous on the i-face between (i-1, j) and (i, j)
template <bool O4, int ORD = 5, class A>
__device__ inline void iface(const A& a, typename A::Out hn[5]) {
  using Out = typename A::Out;
  const double nxf = a.NX(0, 0, 0), nyf = a.NY(0, 0, 0);
  const auto ds = diss_scalars<1, 0>(a, 0);
  const auto vis = iface_viscous<O4>(a);
  const double sn = sqrt(nxf * nxf + nyf * nyf);
  const double inv = 1.0 / sn;
  const double nxl = nxf * inv, nyl = nyf * inv;
  sfor<0, 5>([&](auto E) {
    constexpr int e = decltype(E)::value;
    auto fx = euler_centred<ORD, 1, 0>(a, e, nxf, nyf);
    auto pred = predictor<ORD, 1, 0>(a, e);
    auto jump = 0.5 * (a.template W<0, 0>(e) - a.template W<-1, 0>(e));
    auto diss = ds.rspec * (ds.eps2 * jump + ds.eps4 * pred);
    if constexpr (e == 0) hn[0] = Out(fx - diss);
    else hn[e] = Out(fx - diss - (vis_f<e>(vis) * nxl + vis_g<e>(vis) * nyl) * sn);
  });
}

// ---------------------------------------------------------------- j-faces
template <class A>
__device__ inline Contour contour_j(const A& a) {
  Contour c;
  c.vf = a.VOLF(0, 0, 1);
  c.nxN = 0.5 * (a.NX(1, -1, 0) + a.NX(1, 0, 0));
  c.nxS = -(0.5 * (a.NX(0, -1, 0) + a.NX(0, 0, 0)));
  c.nxO = 0.5 * (a.NX(0, 1, 1) + a.NX(0, 0, 1));
  c.nxE = -(0.5 * (a.NX(0, -1, 1) + a.NX(0, 0, 1)));
  c.nyN = 0.5 * (a.NY(1, -1, 0) + a.NY(1, 0, 0));
  c.nyS = -(0.5 * (a.NY(0, -1, 0) + a.NY(0, 0, 0)));
  c.nyO = 0.5 * (a.NY(0, 1, 1) + a.NY(0, 0, 1));
  c.nyE = -(0.5 * (a.NY(0, -1, 1) + a.NY(0, 0, 1)));
  return c;
}

template <class A>
__device__ inline auto jvals_o4(const A& a, int k) {
  auto col = [&](auto II) {
    constexpr int ii = decltype(II)::value;
    return -a.template P<ii, -2>(k) + 9.0 * a.template P<ii, -1>(k) + 9.0 * a.template P<ii, 0>(k) -
           a.template P<ii, 1>(k);
  };
  auto c0 = col(std::integral_constant<int, -2>{});
  auto c1 = col(std::integral_constant<int, -1>{});
  auto c2 = col(std::integral_constant<int, 0>{});
  auto c3 = col(std::integral_constant<int, 1>{});
  auto c4 = col(std::integral_constant<int, 2>{});
  auto vN = Coef::cross * (-c1 + 7.0 * c2 + 7.0 * c3 - c4);
  auto vS = Coef::cross * (-c0 + 7.0 * c1 + 7.0 * c2 - c3);
  auto vE = Coef::r24 * (-a.template P<0, 0>(k) + 26.0 * a.template P<0, -1>(k) - a.template P<0, -2>(k));
  auto vO = Coef::r24 * (-a.template P<0, 1>(k) + 26.0 * a.template P<0, 0>(k) - a.template P<0, -1>(k));
  struct Q { decltype(vN) N; decltype(vS) S; decltype(vO) O; decltype(vE) E; };
  return Q{vN, vS, vO, vE};
}
template <class A>
__device__ inline auto jvals_o2(const A& a, int k) {
  auto vN = 0.25 * (a.template P<0, 0>(k) + a.template P<1, 0>(k) + a.template P<0, -1>(k) +
                    a.template P<1, -1>(k));
  auto vS = 0.25 * (a.template P<0, 0>(k) + a.template P<-1, 0>(k) + a.template P<0, -1>(k) +
                    a.template P<-1, -1>(k));
  auto vO = a.template P<0, 0>(k);
  auto vE = a.template P<0, -1>(k);
  struct Q { decltype(vN) N; decltype(vS) S; decltype(vO) O; decltype(vE) E; };
  return Q{vN, vS, vO, vE};
}

template <bool O4, class A>
__device__ inline auto jface_viscous(const A& a) {
  const Contour c = contour_j(a);
  auto q = [&](int k) {
    if constexpr (O4) return jvals_o4(a, k);
    else return jvals_o2(a, k);
  };
  auto qu = q(P_U), qv = q(P_V), qw = q(P_W), qt = q(P_T);
  auto ux = cgx(c, qu.N, qu.S, qu.O, qu.E), uy = cgy(c, qu.N, qu.S, qu.O, qu.E);
  auto vx = cgx(c, qv.N, qv.S, qv.O, qv.E), vy = cgy(c, qv.N, qv.S, qv.O, qv.E);
  auto wx = cgx(c, qw.N, qw.S, qw.O, qw.E), wy = cgy(c, qw.N, qw.S, qw.O, qw.E);
  auto tx = cgx(c, qt.N, qt.S, qt.O, qt.E), ty = cgy(c, qt.N, qt.S, qt.O, qt.E);
  const double cpp = a.ph.cp / a.ph.prandtl;
  if constexpr (O4) {
    auto mid = [&](int k) {
      return 0.0625 * (-a.template P<0, -2>(k) + 9.0 * a.template P<0, -1>(k) + 9.0 * a.template P<0, 0>(k) -
                       a.template P<0, 1>(k));
    };
    auto mmu = mid(P_MU);
    return stress_of(mmu, mmu * cpp, mid(P_U), mid(P_V), mid(P_W), ux, vx, wx, tx, uy, vy, wy, ty);
  } else {
    auto half = [&](int k) { return 0.5 * (a.template P<0, 0>(k) + a.template P<0, -1>(k)); };
    auto mmu = half(P_MU);
    return stress_of(mmu, mmu * cpp, half(P_U), half(P_V), half(P_W), ux, vx, wx, tx, uy, vy, wy, ty);
  }
}

// Euler j-flux of component e.  ROW: 0 = centred 6-point (rows j >= 4),
// 2 = off-centred 3/2 face (row j = 2), 3 = off-centred 5/2 face (row j = 3);
// the off-centred ones read cells of rows 1..5 (flux_num_dnc5.F90:164-196).
// ORD 3 has one off-centred row, ROW 2, the mean of rows 1 and 2
// (nearbndfluxes3demi_5p.F, flux_num_dnc3.F90:169-172).
// sum_r w(ROW, r) f(cell row r + 1), r = 0 .. NW-1, added left to right
// (nearbndfluxes*_9p.F, *_11p.F); R0 = 1 - ROW is cell row 1's offset from
// the face row
template <int R0, int R, int ORD, bool Y, class A, class S>
__device__ inline auto wall_acc(const A& a, int e, S acc) {
  if constexpr (R < CoefHi<ORD>::NW) {
    constexpr int ROW = 1 - R0;
    if constexpr (Y) return wall_acc<R0, R + 1, ORD, Y>(a, e, acc + CoefHi<ORD>::w(ROW, R) * efy<0, R0 + R>(a, e));
    else return wall_acc<R0, R + 1, ORD, Y>(a, e, acc + CoefHi<ORD>::w(ROW, R) * efx<0, R0 + R>(a, e));
  } else {
    return acc;
  }
}
template <int R0, int ORD, bool Y, class A>
__device__ inline auto wall_sum(const A& a, int e) {
  constexpr int ROW = 1 - R0;
  if constexpr (Y) return wall_acc<R0, 1, ORD, Y>(a, e, CoefHi<ORD>::w(ROW, 0) * efy<0, R0>(a, e));
  else return wall_acc<R0, 1, ORD, Y>(a, e, CoefHi<ORD>::w(ROW, 0) * efx<0, R0>(a, e));
}

template <int ROW, int ORD = 5, class A>
__device__ inline auto jeuler(const A& a, int e, double n2x, double n2y) {
  if constexpr (ROW == 0) {
    return euler_centred<ORD, 0, 1>(a, e, n2x, n2y);
  } else if constexpr (ORD == 7 || ORD == 9) {
    // nearbndfluxes{3,5,7,9}demi_{9,11}p.F: cell rows 1 .. NW
    return wall_sum<1 - ROW, ORD, false>(a, e) * n2x + wall_sum<1 - ROW, ORD, true>(a, e) * n2y;
  } else if constexpr (ORD == 3) {
    static_assert(ROW == 2, "the 3rd-order scheme has one off-centred row");
    return (0.5 * efx<0, -1>(a, e) + 0.5 * efx<0, 0>(a, e)) * n2x +
           (0.5 * efy<0, -1>(a, e) + 0.5 * efy<0, 0>(a, e)) * n2y;
  } else {
    constexpr int o = 1 - ROW;  // offset of row 1 relative to the face row
    constexpr double k0 = ROW == 2 ? 12.0 / 60.0 : -3.0 / 60.0;
    constexpr double k1 = ROW == 2 ? 77.0 / 60.0 : 27.0 / 60.0;
    constexpr double k2 = ROW == 2 ? -43.0 / 60.0 : 47.0 / 60.0;
    constexpr double k3 = ROW == 2 ? 17.0 / 60.0 : -13.0 / 60.0;
    constexpr double k4 = ROW == 2 ? -3.0 / 60.0 : 2.0 / 60.0;
    return (k0 * efx<0, o>(a, e) + k1 * efx<0, o + 1>(a, e) + k2 * efx<0, o + 2>(a, e) +
            k3 * efx<0, o + 3>(a, e) + k4 * efx<0, o + 4>(a, e)) * n2x +
           (k0 * efy<0, o>(a, e) + k1 * efy<0, o + 1>(a, e) + k2 * efy<0, o + 2>(a, e) +
            k3 * efy<0, o + 3>(a, e) + k4 * efy<0, o + 4>(a, e)) * n2y;
  }
}

// j-face between (i, j-1) and (i, j) for rows j >= 2.
template <int ROW, bool O4, int ORD = 5, class A>
__device__ inline void jface(const A& a, typename A::Out hn[5]) {
  using Out = typename A::Out;
  const double n2x = a.NX(0, 0, 1), n2y = a.NY(0, 0, 1);
  const auto ds = diss_scalars<0, 1>(a, 1);
  const auto vis = jface_viscous<O4>(a);
  const double sn = sqrt(n2x * n2x + n2y * n2y);
  const double inv = 1.0 / sn;
  const double nxl = n2x * inv, nyl = n2y * inv;
  sfor<0, 5>([&](auto E) {
    constexpr int e = decltype(E)::value;
    auto fx = jeuler<ROW, ORD>(a, e, n2x, n2y);
    auto pred = predictor<ORD, 0, 1>(a, e);
    auto jump = 0.5 * (a.template W<0, 0>(e) - a.template W<0, -1>(e));
    auto diss = ds.rspec * (ds.eps2 * jump + ds.eps4 * pred);
    if constexpr (e == 0) hn[0] = Out(fx - diss);
    else hn[e] = Out(fx - diss - (vis_f<e>(vis) * nxl + vis_g<e>(vis) * nyl) * sn);
  });
}

// adiabatic no-slip wall j-face at j = 1 (fluxwall.F)
template <class A>
__device__ inline void jface_wall(const A& a, typename A::Out hn[5]) {
  using Out = typename A::Out;
  const double n2x = a.NX(0, 0, 1), n2y = a.NY(0, 0, 1), vf = a.VOLF(0, 0, 1);
  auto pw = 1.125 * a.template P<0, 0>(P_P) - 0.125 * a.template P<0, 1>(P_P);
  auto mmu = a.template P<0, 0>(P_MU);
  auto u = a.template P<0, 0>(P_U), v = a.template P<0, 0>(P_V), ww = a.template P<0, 0>(P_W);
  auto ux = 2.0 * u * n2x * vf, vx = 2.0 * v * n2x * vf, wx = 2.0 * ww * n2x * vf;
  auto uy = 2.0 * u * n2y * vf, vy = 2.0 * v * n2y * vf, wy = 2.0 * ww * n2y * vf;
  auto fu = Coef::twothird * mmu * (2.0 * ux - vy);
  auto fvv = mmu * (uy + vx);
  auto fw = mmu * wx;
  auto gu = mmu * (uy + vx);
  auto gvv = Coef::twothird * mmu * (-ux + 2.0 * vy);
  auto gw = mmu * wy;
  hn[0] = Out(0.0);
  hn[1] = Out(pw * n2x - (fu * n2x + gu * n2y));
  hn[2] = Out(pw * n2y - (fvv * n2x + gvv * n2y));
  hn[3] = Out(-(fw * n2x + gw * n2y));
  hn[4] = Out(0.0);
}

// dispatch of the i-face / j-face variants on the row.  ORD 3 takes the
// 2nd-order viscous flux everywhere and one off-centred Euler row
// (flux_num_dnc3.F90:143-205).
template <int ORD = 5, class A>
__device__ inline void iface_any(const A& a, typename A::Out hn[5]) {
  if constexpr (ORD == 3) iface<false, 3>(a, hn);
  else if (a.j >= 3) iface<true, ORD>(a, hn);
  else iface<false, ORD>(a, hn);
}
template <int ORD = 5, class A>
__device__ inline void jface_any(const A& a, typename A::Out hn[5]) {
  if constexpr (ORD == 3) {
    if (a.j >= 3) jface<0, false, 3>(a, hn);
    else if (a.j == 2) jface<2, false, 3>(a, hn);
    else jface_wall(a, hn);
  } else if constexpr (ORD == 7) {
    // flux_num_dnc7.F90:117-176: off-centred Euler rows 2 .. 4
    if (a.j >= 5) jface<0, true, 7>(a, hn);
    else if (a.j == 4) jface<4, true, 7>(a, hn);
    else if (a.j == 3) jface<3, true, 7>(a, hn);
    else if (a.j == 2) jface<2, false, 7>(a, hn);
    else jface_wall(a, hn);
  } else if constexpr (ORD == 9) {
    // flux_num_dnc9.F90: off-centred Euler rows 2 .. 5
    if (a.j >= 6) jface<0, true, 9>(a, hn);
    else if (a.j == 5) jface<5, true, 9>(a, hn);
    else if (a.j == 4) jface<4, true, 9>(a, hn);
    else if (a.j == 3) jface<3, true, 9>(a, hn);
    else if (a.j == 2) jface<2, false, 9>(a, hn);
    else jface_wall(a, hn);
  } else {
    if (a.j >= 4) jface<0, true>(a, hn);
    else if (a.j == 3) jface<3, true>(a, hn);
    else if (a.j == 2) jface<2, false>(a, hn);
    else jface_wall(a, hn);
  }
}

// One thread per (i, j) in [1, im+1] x [1, jm+1]: the i-face (i, j) when
// j <= jm and the j-face (i, j) when i <= im (the faces the balance reads).
// NOWALL: the variants without wall rows (srcfv/rhs/flux_num_dnc5_nowall.F90:
// 140-154, flux_num_dnc3_nowall.F90): every face takes the interior formulas,
// the centred Euler flux and the viscous flux of the order (compact 4th-order
// for ORD 5, 7, 9, 2nd-order for ORD 3), the ghost layers supplying the stencil at
// the block edges.  ORD: 3, 5, 7, 9 = flux_num_dnc3*, _dnc5*, _dnc7*, _dnc9*.
template <class T, bool NOWALL = false, int ORD = 5>
__global__ void __launch_bounds__(256) k_face(Mesh m, Phys ph, Metric g, const T* __restrict__ w,
                                              const T* __restrict__ prim, const T* __restrict__ grad,
                                              T* __restrict__ hn) {
  int bx, by;
  xcd_block(bx, by);
  const int i = 1 + bx * blockDim.x + threadIdx.x;
  const int j = 1 + by * blockDim.y + threadIdx.y;
  if (i > m.im + 1 || j > m.jm + 1) return;
  PlainAcc<T> a;
  a.m = m; a.ph = ph; a.g = g; a.i = i; a.j = j;
  a.w = w; a.prim = prim; a.grad = grad;
  const long f = nidx(m, i, j);
  T h[5];
  if (j <= m.jm) {
    if constexpr (NOWALL) iface<ORD != 3, ORD>(a, h);
    else iface_any<ORD>(a, h);
#pragma unroll
    for (int e = 0; e < 5; ++e) hn[f + (long)e * m.nnode] = h[e];
  }
  if (i <= m.im) {
    if constexpr (NOWALL) jface<0, ORD != 3, ORD>(a, h);
    else jface_any<ORD>(a, h);
#pragma unroll
    for (int e = 0; e < 5; ++e) hn[f + (long)(5 + e) * m.nnode] = h[e];
  }
}

}  // namespace bca
