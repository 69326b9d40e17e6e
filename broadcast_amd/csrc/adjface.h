// Reverse-mode face adjoints of the interior faces (the VJP of the DNC5
// residual): for an i-face with the compact viscous flux (rows j >= 3) or a
// centred j-face (rows j >= 4), hnb(face)^T d hn(face) / d w(cell) for the 24
// stencil cells, the same 5 values per offset that the seeded passes of
// k_iface_blocks / k_jface_blocks (MODE_VJP) store, in the same planes, so
// that k_vjp_gather reads either.  The reference's adjoint is TAPENADE's
// reverse sweep of the face loop with recomputation
// (srcfv/adjoint/flux_num_dnc5_b.f90:15-6676); here one thread takes one
// face:
//
//   1. the face's primal quantities (PlainAcc<double>): the dissipation
//      scalars, the jump and predictor of each variable, the viscous
//      gradients and face values;
//   2. the face-level adjoints: of the Roe spectral radius and eps2 / eps4
//      (rspec_b = -sum_e hnb_e (eps2 jump_e + eps4 pred_e), ...), and of the
//      viscous flux by the transposed stress (stress_of reversed by hand,
//      down to the four staggered values N, S, O, E of u, v, w, T and the
//      face values of mu, u, v, w);
//   3. per stencil cell (compile-time offsets), the cell's 5 values:
//        Euler flux, jump and predictor: forward-mode Dual<5> on that cell
//          only (SeededAcc), weighted by hnb_e and the primal scalars;
//        dissipation scalars: reversed by hand at the face level (diss_reverse:
//          adjoints of the state, T and p of the cells on the normal line
//          and of the two cells' gradients), each gradient's stencil
//          coefficients found by evaluating it on the indicator accessor;
//        viscous flux: the cell's coefficient in each linear stencil (found
//          at compile time by evaluating ivals_o4 / jvals_o4 on an indicator
//          accessor) times the face-level adjoints, then the cell's
//          primitive Jacobian (prim_of on Dual<5>) transposed.
// The derivatives are those of the seeded passes (the same Dual
// conventions at abs / max / sqrt / tanh), summed in another order.
#pragma once
#include "blockskern.h"

namespace bca {

// Indicator accessor: the cell at offset (SI, SJ) reads 1, every other cell
// 0, so that a linear stencil evaluated on it is that cell's coefficient
// (constants folded at compile time).
template <int SI, int SJ>
struct IndAcc : AccBase {
  template <int DI, int DJ> __device__ double P(int) const { return (DI == SI && DJ == SJ) ? 1.0 : 0.0; }
  template <int DI, int DJ> __device__ double W(int) const { return (DI == SI && DJ == SJ) ? 1.0 : 0.0; }
};

// weight of the cell at offset (SI, SJ) in the 4-point face value mid(k) of
// the compact viscous flux (0.0625 (-1, 9, 9, -1) across the face)
template <bool J>
__host__ __device__ constexpr double mid_coef(int SI, int SJ) {
  const int along = J ? SJ : SI, across = J ? SI : SJ;
  if (across != 0) return 0.0;
  return along == -2 ? -0.0625 : along == -1 ? 0.5625 : along == 0 ? 0.5625 : along == 1 ? -0.0625 : 0.0;
}

// wb[q] += c d x / d w_q for a Dual<5>; nothing for a value that does not
// depend on the seeded cell
__device__ inline void add_d(double wb[5], const Dual<5>& x, double c) {
#pragma unroll
  for (int q = 0; q < 5; ++q) wb[q] += c * x.d[q];
}
__device__ inline void add_d(double[5], double, double) {}

// primal of the compact viscous flux: face values and gradients
struct ViscPrimal {
  Contour c;
  double mmu, lam, uu, vv, ww;
  double ux, vx, wx, tx, uy, vy, wy, ty;
};
template <bool J, class A>
__device__ inline ViscPrimal visc_primal(const A& a) {
  ViscPrimal v;
  v.c = J ? contour_j(a) : contour_i(a);
  auto q = [&](int k) {
    if constexpr (J) return jvals_o4(a, k);
    else return ivals_o4(a, k);
  };
  const auto qu = q(P_U), qv = q(P_V), qw = q(P_W), qt = q(P_T);
  v.ux = cgx(v.c, qu.N, qu.S, qu.O, qu.E), v.uy = cgy(v.c, qu.N, qu.S, qu.O, qu.E);
  v.vx = cgx(v.c, qv.N, qv.S, qv.O, qv.E), v.vy = cgy(v.c, qv.N, qv.S, qv.O, qv.E);
  v.wx = cgx(v.c, qw.N, qw.S, qw.O, qw.E), v.wy = cgy(v.c, qw.N, qw.S, qw.O, qw.E);
  v.tx = cgx(v.c, qt.N, qt.S, qt.O, qt.E), v.ty = cgy(v.c, qt.N, qt.S, qt.O, qt.E);
  auto mid = [&](int k) {
    if constexpr (J)
      return 0.0625 * (-a.template P<0, -2>(k) + 9.0 * a.template P<0, -1>(k) + 9.0 * a.template P<0, 0>(k) -
                       a.template P<0, 1>(k));
    else
      return 0.0625 * (-a.template P<-2, 0>(k) + 9.0 * a.template P<-1, 0>(k) + 9.0 * a.template P<0, 0>(k) -
                       a.template P<1, 0>(k));
  };
  v.mmu = mid(P_MU);
  v.lam = v.mmu * (a.ph.cp / a.ph.prandtl);
  v.uu = mid(P_U);
  v.vv = mid(P_V);
  v.ww = mid(P_W);
  return v;
}

// Face-level adjoints of the viscous flux: of the staggered values (N, S, O,
// E) of u, v, w, T and of the face values of u, v, w, mu, from the adjoints
// fb[e], gb[e] (e = 1..4) of the stress components vis_f<e>, vis_g<e>.
struct ViscAdj {
  double N[4], S[4], O[4], E[4];  // k = P_U, P_V, P_W, P_T
  double mid[4];                  // u, v, w, mu
};
__device__ inline ViscAdj visc_reverse(const ViscPrimal& v, const double fb[5], const double gb[5], double cpp) {
  constexpr double tt = Coef::twothird;
  // stress_of forward, for the products the reverse needs
  const double fv1 = tt * v.mmu * (2.0 * v.ux - v.vy), fv2 = v.mmu * (v.uy + v.vx), fv3 = v.mmu * v.wx;
  const double gv1 = v.mmu * (v.uy + v.vx), gv2 = tt * v.mmu * (-v.ux + 2.0 * v.vy), gv3 = v.mmu * v.wy;
  const double f1b = fb[1] + fb[4] * v.uu, f2b = fb[2] + fb[4] * v.vv, f3b = fb[3] + fb[4] * v.ww;
  const double g1b = gb[1] + gb[4] * v.uu, g2b = gb[2] + gb[4] * v.vv, g3b = gb[3] + gb[4] * v.ww;
  const double lamb = fb[4] * v.tx + gb[4] * v.ty;
  const double txb = fb[4] * v.lam, tyb = gb[4] * v.lam;
  ViscAdj r;
  r.mid[0] = fb[4] * fv1 + gb[4] * gv1;
  r.mid[1] = fb[4] * fv2 + gb[4] * gv2;
  r.mid[2] = fb[4] * fv3 + gb[4] * gv3;
  r.mid[3] = f1b * tt * (2.0 * v.ux - v.vy) + f2b * (v.uy + v.vx) + f3b * v.wx + g1b * (v.uy + v.vx) +
             g2b * tt * (-v.ux + 2.0 * v.vy) + g3b * v.wy + lamb * cpp;
  const double uxb = f1b * tt * 2.0 * v.mmu - g2b * tt * v.mmu;
  const double vyb = -f1b * tt * v.mmu + g2b * tt * 2.0 * v.mmu;
  const double uyb = (f2b + g1b) * v.mmu, vxb = (f2b + g1b) * v.mmu;
  const double wxb = f3b * v.mmu, wyb = g3b * v.mmu;
  const double xb[4] = {uxb, vxb, wxb, txb}, yb[4] = {uyb, vyb, wyb, tyb};
  const Contour& c = v.c;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    r.N[k] = (xb[k] * c.nxN + yb[k] * c.nyN) * c.vf;
    r.S[k] = (xb[k] * c.nxS + yb[k] * c.nyS) * c.vf;
    r.O[k] = (xb[k] * c.nxO + yb[k] * c.nyO) * c.vf;
    r.E[k] = (xb[k] * c.nxE + yb[k] * c.nyE) * c.vf;
  }
  return r;
}

// Coefficient of u (k = G_UX, G_UY) or v (G_VX, G_VY) of the indicator's
// cell in the gradient G(k) at offset (GI, GJ), as SeededAcc::G forms it:
// the 4th-order gradient, or at the ghost layer i = 0, i = im+1, j = jm+1 its
// linear extrapolation 2 g(a) - g(b).
template <int GI, int GJ, class A>
__device__ inline double grad_lin(const A& ind, int k) {
  const int gi = ind.i + GI, gj = ind.j + GJ;
  if (gi == 0) return 2.0 * grad_direct<GI + 1, GJ>(ind, k) - grad_direct<GI + 2, GJ>(ind, k);
  if (gi == ind.m.im + 1) return 2.0 * grad_direct<GI - 1, GJ>(ind, k) - grad_direct<GI - 2, GJ>(ind, k);
  if (gj == ind.m.jm + 1) return 2.0 * grad_direct<GI, GJ - 1>(ind, k) - grad_direct<GI, GJ - 2>(ind, k);
  return grad_direct<GI, GJ>(ind, k);
}

// Adjoints of the inputs of diss_scalars<DX, DY> (the face between cell 0
// and cell -1 = (-DX, -DY)): of W(0..2) and T of cells 0 and -1, of the
// pressure of cells -2, -1, 0, 1 along the face normal, and of the gradients
// G of cells 0 and -1; from the adjoints of rspec, eps2 and eps4.  The
// derivative conventions are Dual's: abs takes + at 0, sqrt has derivative 0
// at 0, max(a, b) = a < b ? b : a.
struct DissAdj {
  double w0[3], w1[3];  // rho, rho u, rho v of cells 0, -1
  double t0, t1;        // T of cells 0, -1
  double p[4];          // p of cells -2, -1, 0, 1
  double g0[NGRAD], g1[NGRAD];
};
template <int DX, int DY, class A>
__device__ inline DissAdj diss_reverse(const A& a, int d, double rspb, double e2b, double e4b) {
  DissAdj r;
  const double gr = a.ph.gam * a.ph.rgaz;
  const double rr = a.template W<0, 0>(0), rl = a.template W<-DX, -DY>(0);
  const double rur = a.template W<0, 0>(1), rvr = a.template W<0, 0>(2);
  const double rul = a.template W<-DX, -DY>(1), rvl = a.template W<-DX, -DY>(2);
  const double ur = rur / rr, vr = rvr / rr, ul = rul / rl, vl = rvl / rl;
  const double c2r = gr * a.template P<0, 0>(P_T), c2l = gr * a.template P<-DX, -DY>(P_T);
  const double q = rr / rl;
  const double rt = sqrt(q);
  const double wl = 1.0 / (1.0 + rt), wr = 1.0 - wl;
  const double u = ul * wl + ur * wr, v = vl * wl + vr * wr, c2 = c2l * wl + c2r * wr;
  const double nxf = a.NX(0, 0, d), nyf = a.NY(0, 0, d);
  const double n2 = nxf * nxf + nyf * nyf;
  const double un = nxf * u + nyf * v, c2n = c2 * n2;
  double c2r_b = 0.0, c2l_b = 0.0;
#pragma unroll
  for (int k = 0; k < 4; ++k) r.p[k] = 0.0;
#pragma unroll
  for (int k = 0; k < NGRAD; ++k) r.g0[k] = r.g1[k] = 0.0;
  if (a.ph.k2 != 0.0) {
    const double pm2 = a.template P<-2 * DX, -2 * DY>(P_P), pm1 = a.template P<-DX, -DY>(P_P);
    const double p0 = a.template P<0, 0>(P_P), pp1 = a.template P<DX, DY>(P_P);
    const double n1 = pm1 - 2.0 * p0 + pp1, d1 = pm1 + 2.0 * p0 + pp1;
    const double n2s = pm2 - 2.0 * pm1 + p0, d2s = pm2 + 2.0 * pm1 + p0;
    const double ks1 = fabs(n1) / fabs(d1), ks2 = fabs(n2s) / fabs(d2s);
    const double divr = a.template G<0, 0>(G_UX) + a.template G<0, 0>(G_VY);
    const double rotr = a.template G<0, 0>(G_VX) - a.template G<0, 0>(G_UY);
    const double div2r = divr * divr, den1 = div2r + rotr * rotr + 1e-15, duc1 = div2r / den1;
    const double s1 = sqrt(c2r * n2), h1 = 10.0 * a.VOL(0, 0) / (s1 + 1e-15), th1 = tanh(2.5 + h1 * divr);
    const double sw1 = 0.5 * (1.0 - th1);
    const double divl = a.template G<-DX, -DY>(G_UX) + a.template G<-DX, -DY>(G_VY);
    const double rotl = a.template G<-DX, -DY>(G_VX) - a.template G<-DX, -DY>(G_UY);
    const double div2l = divl * divl, den2 = div2l + rotl * rotl + 1e-15, duc2 = div2l / den2;
    const double s2 = sqrt(c2l * n2), h2 = 10.0 * a.VOL(-DX, -DY) / (s2 + 1e-15), th2 = tanh(2.5 + h2 * divl);
    const double sw2 = 0.5 * (1.0 - th2);
    const double Ks = ks1 < ks2 ? ks2 : ks1, Du = duc1 < duc2 ? duc2 : duc1, Sw = sw1 < sw2 ? sw2 : sw1;
    const double eps2 = a.ph.k2 * (Ks * Du * Sw);
    // eps4 = max(0, k4 - 12 eps2)
    const double eb = e2b + ((0.0 < a.ph.k4 - eps2 * 12.0) ? -12.0 * e4b : 0.0);
    const double Ksb = eb * a.ph.k2 * Du * Sw, Dub = eb * a.ph.k2 * Ks * Sw, Swb = eb * a.ph.k2 * Ks * Du;
    const double ks1b = ks1 < ks2 ? 0.0 : Ksb, ks2b = ks1 < ks2 ? Ksb : 0.0;
    const double duc1b = duc1 < duc2 ? 0.0 : Dub, duc2b = duc1 < duc2 ? Dub : 0.0;
    const double sw1b = sw1 < sw2 ? 0.0 : Swb, sw2b = sw1 < sw2 ? Swb : 0.0;
    // ks = |n| / |d|
    const double n1b = (ks1b / fabs(d1)) * (n1 >= 0.0 ? 1.0 : -1.0);
    const double d1b = (-ks1b * ks1 / fabs(d1)) * (d1 >= 0.0 ? 1.0 : -1.0);
    const double n2b = (ks2b / fabs(d2s)) * (n2s >= 0.0 ? 1.0 : -1.0);
    const double d2b = (-ks2b * ks2 / fabs(d2s)) * (d2s >= 0.0 ? 1.0 : -1.0);
    r.p[0] = n2b + d2b;
    r.p[1] = n1b + d1b - 2.0 * n2b + 2.0 * d2b;
    r.p[2] = -2.0 * n1b + 2.0 * d1b + n2b + d2b;
    r.p[3] = n1b + d1b;
    // duc = div^2 / (div^2 + rot^2 + 1e-15)
    const double dv1b = duc1b / den1 - duc1b * duc1 / den1, rt1b = -duc1b * duc1 / den1;
    const double dv2b = duc2b / den2 - duc2b * duc2 / den2, rt2b = -duc2b * duc2 / den2;
    // sw = 0.5 (1 - tanh(2.5 + h div)), h = 10 vol / (sqrt(c2 n2) + 1e-15)
    const double z1b = -0.5 * sw1b * (1.0 - th1 * th1), z2b = -0.5 * sw2b * (1.0 - th2 * th2);
    const double divrb = 2.0 * divr * dv1b + z1b * h1, rotrb = 2.0 * rotr * rt1b;
    const double divlb = 2.0 * divl * dv2b + z2b * h2, rotlb = 2.0 * rotl * rt2b;
    const double s1b = -(z1b * divr) * h1 / (s1 + 1e-15), s2b = -(z2b * divl) * h2 / (s2 + 1e-15);
    c2r_b += (c2r * n2 == 0.0 ? 0.0 : s1b * 0.5 / s1) * n2;
    c2l_b += (c2l * n2 == 0.0 ? 0.0 : s2b * 0.5 / s2) * n2;
    r.g0[G_UX] = divrb, r.g0[G_VY] = divrb, r.g0[G_VX] = rotrb, r.g0[G_UY] = -rotrb;
    r.g1[G_UX] = divlb, r.g1[G_VY] = divlb, r.g1[G_VX] = rotlb, r.g1[G_UY] = -rotlb;
  }
  // rspec = |n . (u, v)| + sqrt(c2 |n|^2) of the Roe average
  const double unb = rspb * (un >= 0.0 ? 1.0 : -1.0);
  const double c2nb = c2n == 0.0 ? 0.0 : rspb * 0.5 / sqrt(c2n);
  const double ub = unb * nxf, vb = unb * nyf, c2b = c2nb * n2;
  const double ulb = ub * wl, urb = ub * wr, vlb = vb * wl, vrb = vb * wr;
  c2l_b += c2b * wl;
  c2r_b += c2b * wr;
  const double wrb = ub * ur + vb * vr + c2b * c2r;
  const double wlb = ub * ul + vb * vl + c2b * c2l - wrb;
  const double rtb = -wlb * wl * wl;
  const double qb = q == 0.0 ? 0.0 : rtb * 0.5 / rt;
  r.w0[0] = qb / rl - urb * ur / rr - vrb * vr / rr;
  r.w1[0] = -qb * q / rl - ulb * ul / rl - vlb * vl / rl;
  r.w0[1] = urb / rr, r.w0[2] = vrb / rr;
  r.w1[1] = ulb / rl, r.w1[2] = vlb / rl;
  r.t0 = gr * c2r_b;
  r.t1 = gr * c2l_b;
  return r;
}

// The face adjoint of one face for the offsets [O0, O1): J = false an
// i-face (compact viscous flux), true a centred j-face.  hb: the face's
// adjoint hnb; out: out[(o * 5 + q) * nf + f].
template <bool J, int O0, int O1>
__device__ inline void face_reverse(const Mesh& m, const Phys& ph, const Metric& g, const double* w,
                                    const double* prim, const double* grad, int i, int j, const double hb[5],
                                    double* out, long nf, long f) {
  constexpr int DX = J ? 0 : 1, DY = J ? 1 : 0, D = J ? 1 : 0;
  PlainAcc<double> a;
  a.m = m; a.ph = ph; a.g = g; a.i = i; a.j = j;
  a.w = w; a.prim = prim; a.grad = grad;
  const double nxf = a.NX(0, 0, D), nyf = a.NY(0, 0, D);
  const auto ds = diss_scalars<DX, DY>(a, D);
  const double sn = sqrt(nxf * nxf + nyf * nyf);
  const double inv = 1.0 / sn;
  const double nxl = nxf * inv, nyl = nyf * inv;
  // dissipation: diss_e = rspec (eps2 jump_e + eps4 pred_e), hn_e -= diss_e
  double rspb = 0.0, e2b = 0.0, e4b = 0.0;
#pragma unroll
  for (int e = 0; e < 5; ++e) {
    const double jp = 0.5 * (a.template W<0, 0>(e) - a.template W<-DX, -DY>(e));
    const double pr = predictor<5, DX, DY>(a, e);
    rspb -= hb[e] * (ds.eps2 * jp + ds.eps4 * pr);
    e2b -= hb[e] * ds.rspec * jp;
    e4b -= hb[e] * ds.rspec * pr;
  }
  // viscous: hn_e -= (vis_f<e> nxl + vis_g<e> nyl) sn, e = 1..4
  const ViscPrimal vp = visc_primal<J>(a);
  double fb[5], gb[5];
#pragma unroll
  for (int e = 0; e < 5; ++e) {
    fb[e] = e == 0 ? 0.0 : -hb[e] * sn * nxl;
    gb[e] = e == 0 ? 0.0 : -hb[e] * sn * nyl;
  }
  const ViscAdj va = visc_reverse(vp, fb, gb, ph.cp / ph.prandtl);
  const DissAdj da = diss_reverse<DX, DY>(a, D, rspb, e2b, e4b);

  sfor<O0, O1>([&](auto O) {
    constexpr int o = decltype(O)::value;
    constexpr int SI = J ? joff_di(o) : ioff_di(o), SJ = J ? joff_dj(o) : ioff_dj(o);
    // the two extra offsets only exist on the faces the seeded passes store
    if constexpr (J) {
      if constexpr (o == 22) return;
      else if (o == 23 && j != m.jm + 1) return;
    } else {
      if (o == 22 && i != 1) return;
      if (o == 23 && i != m.im + 1) return;
    }
    auto s = seeded<SI, SJ>(m, ph, g, w, prim, grad, i, j);
    double wb[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    // Euler flux, jump and predictor through the cell's state
    sfor<0, 5>([&](auto E) {
      constexpr int e = decltype(E)::value;
      add_d(wb, euler_centred<5, DX, DY>(s, e, nxf, nyf), hb[e]);
      add_d(wb, 0.5 * (s.template W<0, 0>(e) - s.template W<-DX, -DY>(e)), -hb[e] * ds.rspec * ds.eps2);
      add_d(wb, predictor<5, DX, DY>(s, e), -hb[e] * ds.rspec * ds.eps4);
    });
    // the dissipation scalars and the viscous flux through the cell's state
    // and primitives
    {
      IndAcc<SI, SJ> ind;
      ind.m = m; ind.ph = ph; ind.g = g; ind.i = i; ind.j = j;
      const auto qc = [&] {
        if constexpr (J) return jvals_o4(ind, 0);
        else return ivals_o4(ind, 0);
      }();
      constexpr double cm = mid_coef<J>(SI, SJ);
      double pb[NPRIM];
#pragma unroll
      for (int k = 0; k < NPRIM; ++k) pb[k] = 0.0;
#pragma unroll
      for (int k = 0; k < 4; ++k) pb[k] = qc.N * va.N[k] + qc.S * va.S[k] + qc.O * va.O[k] + qc.E * va.E[k];
      pb[P_U] += cm * va.mid[0];
      pb[P_V] += cm * va.mid[1];
      pb[P_W] += cm * va.mid[2];
      pb[P_MU] += cm * va.mid[3];
      if constexpr (SI == 0 && SJ == 0) {
#pragma unroll
        for (int q = 0; q < 3; ++q) wb[q] += da.w0[q];
        pb[P_T] += da.t0;
        pb[P_P] += da.p[2];
      }
      if constexpr (SI == -DX && SJ == -DY) {
#pragma unroll
        for (int q = 0; q < 3; ++q) wb[q] += da.w1[q];
        pb[P_T] += da.t1;
        pb[P_P] += da.p[1];
      }
      if constexpr (SI == -2 * DX && SJ == -2 * DY) pb[P_P] += da.p[0];
      if constexpr (SI == DX && SJ == DY) pb[P_P] += da.p[3];
      if constexpr (grad_reach(SI, SJ, 0, 0)) {
        pb[P_U] += grad_lin<0, 0>(ind, G_UX) * da.g0[G_UX] + grad_lin<0, 0>(ind, G_UY) * da.g0[G_UY];
        pb[P_V] += grad_lin<0, 0>(ind, G_VX) * da.g0[G_VX] + grad_lin<0, 0>(ind, G_VY) * da.g0[G_VY];
      }
      if constexpr (grad_reach(SI, SJ, -DX, -DY)) {
        pb[P_U] += grad_lin<-DX, -DY>(ind, G_UX) * da.g1[G_UX] + grad_lin<-DX, -DY>(ind, G_UY) * da.g1[G_UY];
        pb[P_V] += grad_lin<-DX, -DY>(ind, G_VX) * da.g1[G_VX] + grad_lin<-DX, -DY>(ind, G_VY) * da.g1[G_VY];
      }
      add_d(wb, s.sp[P_U], pb[P_U]);
      add_d(wb, s.sp[P_V], pb[P_V]);
      add_d(wb, s.sp[P_W], pb[P_W]);
      add_d(wb, s.sp[P_T], pb[P_T]);
      add_d(wb, s.sp[P_P], pb[P_P]);
      add_d(wb, s.sp[P_MU], pb[P_MU]);
    }
    double* p = out + (long)o * 5 * nf + f;
#pragma unroll
    for (int q = 0; q < 5; ++q) p[q * nf] = wb[q];
  });
}

// offsets per launch of the reverse kernels (the face-level part is
// recomputed per launch): 12 keeps the i-face kernel free of scratch spills
// (24 spill 236 bytes per lane)
constexpr int REV_GROUP = 12;

// One thread per face: i-faces (i, j), i in [1, im+1], j in [j0, j1] (J =
// false, rows j >= 3), or centred j-faces (i, j), i in [1, im], j in [j0, j1]
// (J = true, rows j >= 4); offsets [O0, O1) of the face's stencil.
template <bool J, int O0 = 0, int O1 = NOFF>
__global__ void __launch_bounds__(256) k_face_reverse(Mesh m, Phys ph, Metric g, const double* __restrict__ w,
                                                      const double* __restrict__ prim,
                                                      const double* __restrict__ grad, double* __restrict__ out,
                                                      const double* __restrict__ resb, int j0, int j1) {
  int bx, by;
  xcd_block(bx, by);
  const int i = 1 + bx * blockDim.x + threadIdx.x;
  const int j = j0 + by * blockDim.y + threadIdx.y;
  if (i > (J ? m.im : m.im + 1) || j > j1) return;
  const long nf = J ? (long)m.im * (m.jm + 1) : (long)(m.im + 1) * m.jm;
  const long f = J ? (i - 1) + (long)m.im * (j - 1) : (i - 1) + (long)(m.im + 1) * (j - 1);
  double hb[5];
  face_adjoint(m, resb, i, j, J ? 0 : 1, J ? 1 : 0, hb);
  face_reverse<J, O0, O1>(m, ph, g, w, prim, grad, i, j, hb, out, nf, f);
}

}  // namespace bca
