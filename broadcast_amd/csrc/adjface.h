// Reverse-mode face adjoints (the VJP of the DNC5 residual): for every i-face
// and every j-face above the wall row, hnb(face)^T d hn(face) / d w(cell) for the 24
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
//   3. per stencil cell (compile-time offsets), the cell's 5 values, all
//      in plain FP64:
//        Euler flux: the cell's flux components reversed by hand, weighted
//          by its coefficient in the centred flux and hnb_e;
//        jump and predictor: the cell's coefficient (indicator accessor)
//          times hnb_e and the primal scalars;
//        dissipation scalars: reversed by hand at the face level (diss_adjoint:
//          adjoints of the state, T and p of the cells on the normal line
//          and of the two cells' gradients), each gradient's stencil
//          coefficients found by evaluating it on the indicator accessor;
//        viscous flux: the cell's coefficient in each linear stencil (found
//          at compile time by evaluating ivals_o4 / jvals_o4 on an indicator
//          accessor) times the face-level adjoints;
//        then the primitives' adjoints through prim_of reversed by hand.
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

// weight of the cell at offset (SI, SJ) in the face value of the viscous
// flux: 0.0625 (-1, 9, 9, -1) across the face for the compact flux (O4),
// 0.5 (1, 1) for the 2nd-order one
template <bool J, bool O4>
__host__ __device__ constexpr double mid_coef(int SI, int SJ) {
  const int along = J ? SJ : SI, across = J ? SI : SJ;
  if (across != 0) return 0.0;
  if (!O4) return (along == 0 || along == -1) ? 0.5 : 0.0;
  return along == -2 ? -0.0625 : along == -1 ? 0.5625 : along == 0 ? 0.5625 : along == 1 ? -0.0625 : 0.0;
}

// wb[q] += c d x / d w_q for a Dual<5>; nothing for a value that does not
// depend on the seeded cell
__device__ inline void add_d(double wb[5], const Dual<5>& x, double c) {
#pragma unroll
  for (int q = 0; q < 5; ++q) wb[q] += c * x.d[q];
}
__device__ inline void add_d(double[5], double, double) {}

// primal of the viscous flux (compact, O4, or 2nd-order): face values and
// gradients
struct ViscPrimal {
  Contour c;
  double mmu, lam, uu, vv, ww;
  double ux, vx, wx, tx, uy, vy, wy, ty;
};
template <bool J, bool O4, class A>
__device__ inline ViscPrimal visc_primal(const A& a) {
  ViscPrimal v;
  v.c = J ? contour_j(a) : contour_i(a);
  auto q = [&](int k) {
    if constexpr (J && O4) return jvals_o4(a, k);
    else if constexpr (J) return jvals_o2(a, k);
    else if constexpr (O4) return ivals_o4(a, k);
    else return ivals_o2(a, k);
  };
  const auto qu = q(P_U), qv = q(P_V), qw = q(P_W), qt = q(P_T);
  v.ux = cgx(v.c, qu.N, qu.S, qu.O, qu.E), v.uy = cgy(v.c, qu.N, qu.S, qu.O, qu.E);
  v.vx = cgx(v.c, qv.N, qv.S, qv.O, qv.E), v.vy = cgy(v.c, qv.N, qv.S, qv.O, qv.E);
  v.wx = cgx(v.c, qw.N, qw.S, qw.O, qw.E), v.wy = cgy(v.c, qw.N, qw.S, qw.O, qw.E);
  v.tx = cgx(v.c, qt.N, qt.S, qt.O, qt.E), v.ty = cgy(v.c, qt.N, qt.S, qt.O, qt.E);
  auto mid = [&](int k) {
    if constexpr (!O4 && J) return 0.5 * (a.template P<0, 0>(k) + a.template P<0, -1>(k));
    else if constexpr (!O4) return 0.5 * (a.template P<0, 0>(k) + a.template P<-1, 0>(k));
    else if constexpr (J)
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

// The quantities of diss_scalars<DX, DY> (the face between cell 0 and cell
// -1 = (-DX, -DY)) that its reverse needs, computed as diss_scalars does.
struct DissPrim {
  double rr, rl, ur, vr, ul, vl, c2r, c2l, q, rt, wl, wr, nxf, nyf, n2, un, c2n;
  double rspec, eps2, eps4;
  double n1, d1, n2s, d2s, ks1, ks2, divr, rotr, den1, duc1, s1, h1, th1, sw1;
  double divl, rotl, den2, duc2, s2, h2, th2, sw2;
};
template <int DX, int DY, class A>
__device__ inline DissPrim diss_primal(const A& a, int d) {
  DissPrim P;
  const double gr = a.ph.gam * a.ph.rgaz;
  P.rr = a.template W<0, 0>(0), P.rl = a.template W<-DX, -DY>(0);
  P.ur = a.template W<0, 0>(1) / P.rr, P.vr = a.template W<0, 0>(2) / P.rr;
  P.ul = a.template W<-DX, -DY>(1) / P.rl, P.vl = a.template W<-DX, -DY>(2) / P.rl;
  P.c2r = gr * a.template P<0, 0>(P_T), P.c2l = gr * a.template P<-DX, -DY>(P_T);
  P.q = P.rr / P.rl;
  P.rt = sqrt(P.q);
  P.wl = 1.0 / (1.0 + P.rt), P.wr = 1.0 - P.wl;
  const double u = P.ul * P.wl + P.ur * P.wr, v = P.vl * P.wl + P.vr * P.wr;
  const double c2 = P.c2l * P.wl + P.c2r * P.wr;
  P.nxf = a.NX(0, 0, d), P.nyf = a.NY(0, 0, d);
  P.n2 = P.nxf * P.nxf + P.nyf * P.nyf;
  P.un = P.nxf * u + P.nyf * v, P.c2n = c2 * P.n2;
  P.rspec = fabs(P.un) + sqrt(P.c2n);
  P.eps2 = 0.0;
  if (a.ph.k2 != 0.0) {
    const double pm2 = a.template P<-2 * DX, -2 * DY>(P_P), pm1 = a.template P<-DX, -DY>(P_P);
    const double p0 = a.template P<0, 0>(P_P), pp1 = a.template P<DX, DY>(P_P);
    P.n1 = pm1 - 2.0 * p0 + pp1, P.d1 = pm1 + 2.0 * p0 + pp1;
    P.n2s = pm2 - 2.0 * pm1 + p0, P.d2s = pm2 + 2.0 * pm1 + p0;
    P.ks1 = fabs(P.n1) / fabs(P.d1), P.ks2 = fabs(P.n2s) / fabs(P.d2s);
    P.divr = a.template G<0, 0>(G_UX) + a.template G<0, 0>(G_VY);
    P.rotr = a.template G<0, 0>(G_VX) - a.template G<0, 0>(G_UY);
    P.den1 = P.divr * P.divr + P.rotr * P.rotr + 1e-15, P.duc1 = P.divr * P.divr / P.den1;
    P.s1 = sqrt(P.c2r * P.n2), P.h1 = 10.0 * a.VOL(0, 0) / (P.s1 + 1e-15), P.th1 = tanh(2.5 + P.h1 * P.divr);
    P.sw1 = 0.5 * (1.0 - P.th1);
    P.divl = a.template G<-DX, -DY>(G_UX) + a.template G<-DX, -DY>(G_VY);
    P.rotl = a.template G<-DX, -DY>(G_VX) - a.template G<-DX, -DY>(G_UY);
    P.den2 = P.divl * P.divl + P.rotl * P.rotl + 1e-15, P.duc2 = P.divl * P.divl / P.den2;
    P.s2 = sqrt(P.c2l * P.n2), P.h2 = 10.0 * a.VOL(-DX, -DY) / (P.s2 + 1e-15), P.th2 = tanh(2.5 + P.h2 * P.divl);
    P.sw2 = 0.5 * (1.0 - P.th2);
    P.eps2 = a.ph.k2 * ((P.ks1 < P.ks2 ? P.ks2 : P.ks1) * (P.duc1 < P.duc2 ? P.duc2 : P.duc1) *
                        (P.sw1 < P.sw2 ? P.sw2 : P.sw1));
  }
  P.eps4 = 0.0 < a.ph.k4 - P.eps2 * 12.0 ? a.ph.k4 - P.eps2 * 12.0 : 0.0;
  return P;
}

// Adjoints of the inputs of diss_scalars: of W(0..2) and T of cells 0 and
// -1, of the pressure of cells -2, -1, 0, 1 along the face normal, and of
// the gradients G of cells 0 and -1; from the adjoints of rspec, eps2 and
// eps4.  The derivative conventions are Dual's: abs takes + at 0, sqrt has
// derivative 0 at 0, max(a, b) = a < b ? b : a.
struct DissAdj {
  double w0[3], w1[3];  // rho, rho u, rho v of cells 0, -1
  double t0, t1;        // T of cells 0, -1
  double p[4];          // p of cells -2, -1, 0, 1
  double g0[NGRAD], g1[NGRAD];
};
__device__ inline DissAdj diss_adjoint(const DissPrim& P, const Phys& ph, double rspb, double e2b, double e4b) {
  DissAdj r;
  const double gr = ph.gam * ph.rgaz;
  double c2r_b = 0.0, c2l_b = 0.0;
#pragma unroll
  for (int k = 0; k < 4; ++k) r.p[k] = 0.0;
#pragma unroll
  for (int k = 0; k < NGRAD; ++k) r.g0[k] = r.g1[k] = 0.0;
  if (ph.k2 != 0.0) {
    const double Ks = P.ks1 < P.ks2 ? P.ks2 : P.ks1, Du = P.duc1 < P.duc2 ? P.duc2 : P.duc1;
    const double Sw = P.sw1 < P.sw2 ? P.sw2 : P.sw1;
    // eps4 = max(0, k4 - 12 eps2)
    const double eb = e2b + ((0.0 < ph.k4 - P.eps2 * 12.0) ? -12.0 * e4b : 0.0);
    const double Ksb = eb * ph.k2 * Du * Sw, Dub = eb * ph.k2 * Ks * Sw, Swb = eb * ph.k2 * Ks * Du;
    const double ks1b = P.ks1 < P.ks2 ? 0.0 : Ksb, ks2b = P.ks1 < P.ks2 ? Ksb : 0.0;
    const double duc1b = P.duc1 < P.duc2 ? 0.0 : Dub, duc2b = P.duc1 < P.duc2 ? Dub : 0.0;
    const double sw1b = P.sw1 < P.sw2 ? 0.0 : Swb, sw2b = P.sw1 < P.sw2 ? Swb : 0.0;
    // ks = |n| / |d|
    const double n1b = (ks1b / fabs(P.d1)) * (P.n1 >= 0.0 ? 1.0 : -1.0);
    const double d1b = (-ks1b * P.ks1 / fabs(P.d1)) * (P.d1 >= 0.0 ? 1.0 : -1.0);
    const double n2b = (ks2b / fabs(P.d2s)) * (P.n2s >= 0.0 ? 1.0 : -1.0);
    const double d2b = (-ks2b * P.ks2 / fabs(P.d2s)) * (P.d2s >= 0.0 ? 1.0 : -1.0);
    r.p[0] = n2b + d2b;
    r.p[1] = n1b + d1b - 2.0 * n2b + 2.0 * d2b;
    r.p[2] = -2.0 * n1b + 2.0 * d1b + n2b + d2b;
    r.p[3] = n1b + d1b;
    // duc = div^2 / (div^2 + rot^2 + 1e-15)
    const double dv1b = duc1b / P.den1 - duc1b * P.duc1 / P.den1, rt1b = -duc1b * P.duc1 / P.den1;
    const double dv2b = duc2b / P.den2 - duc2b * P.duc2 / P.den2, rt2b = -duc2b * P.duc2 / P.den2;
    // sw = 0.5 (1 - tanh(2.5 + h div)), h = 10 vol / (sqrt(c2 n2) + 1e-15)
    const double z1b = -0.5 * sw1b * (1.0 - P.th1 * P.th1), z2b = -0.5 * sw2b * (1.0 - P.th2 * P.th2);
    const double divrb = 2.0 * P.divr * dv1b + z1b * P.h1, rotrb = 2.0 * P.rotr * rt1b;
    const double divlb = 2.0 * P.divl * dv2b + z2b * P.h2, rotlb = 2.0 * P.rotl * rt2b;
    const double s1b = -(z1b * P.divr) * P.h1 / (P.s1 + 1e-15), s2b = -(z2b * P.divl) * P.h2 / (P.s2 + 1e-15);
    c2r_b += (P.c2r * P.n2 == 0.0 ? 0.0 : s1b * 0.5 / P.s1) * P.n2;
    c2l_b += (P.c2l * P.n2 == 0.0 ? 0.0 : s2b * 0.5 / P.s2) * P.n2;
    r.g0[G_UX] = divrb, r.g0[G_VY] = divrb, r.g0[G_VX] = rotrb, r.g0[G_UY] = -rotrb;
    r.g1[G_UX] = divlb, r.g1[G_VY] = divlb, r.g1[G_VX] = rotlb, r.g1[G_UY] = -rotlb;
  }
  // rspec = |n . (u, v)| + sqrt(c2 |n|^2) of the Roe average
  const double unb = rspb * (P.un >= 0.0 ? 1.0 : -1.0);
  const double c2nb = P.c2n == 0.0 ? 0.0 : rspb * 0.5 / sqrt(P.c2n);
  const double ub = unb * P.nxf, vb = unb * P.nyf, c2b = c2nb * P.n2;
  const double ulb = ub * P.wl, urb = ub * P.wr, vlb = vb * P.wl, vrb = vb * P.wr;
  c2l_b += c2b * P.wl;
  c2r_b += c2b * P.wr;
  const double wrb = ub * P.ur + vb * P.vr + c2b * P.c2r;
  const double wlb = ub * P.ul + vb * P.vl + c2b * P.c2l - wrb;
  const double rtb = -wlb * P.wl * P.wl;
  const double qb = P.q == 0.0 ? 0.0 : rtb * 0.5 / P.rt;
  r.w0[0] = qb / P.rl - urb * P.ur / P.rr - vrb * P.vr / P.rr;
  r.w1[0] = -qb * P.q / P.rl - ulb * P.ul / P.rl - vlb * P.vl / P.rl;
  r.w0[1] = urb / P.rr, r.w0[2] = vrb / P.rr;
  r.w1[1] = ulb / P.rl, r.w1[2] = vlb / P.rl;
  r.t0 = gr * c2r_b;
  r.t1 = gr * c2l_b;
  return r;
}

// Metric factors of the 4th-order gradient at offset (GI, GJ) (grad_direct:
// G(k) = m1 di + m2 dj, x components of the normals for G_UX / G_VX, y for
// G_UY / G_VY), folded with the gradient adjoints gb of that cell: the
// weights cu / cv of u / v along i and j.
struct GradW {
  double cu_i, cu_j, cv_i, cv_j;
};
template <int GI, int GJ, class A>
__device__ inline GradW grad_weights(const A& a, const double gb[NGRAD]) {
  const double volm1 = 1.0 / a.VOL(GI, GJ);
  const double m1x = 0.5 * (a.NX(GI, GJ, 0) + a.NX(GI + 1, GJ, 0)) * volm1;
  const double m2x = 0.5 * (a.NX(GI, GJ, 1) + a.NX(GI, GJ + 1, 1)) * volm1;
  const double m1y = 0.5 * (a.NY(GI, GJ, 0) + a.NY(GI + 1, GJ, 0)) * volm1;
  const double m2y = 0.5 * (a.NY(GI, GJ, 1) + a.NY(GI, GJ + 1, 1)) * volm1;
  return {gb[G_UX] * m1x + gb[G_UY] * m1y, gb[G_UX] * m2x + gb[G_UY] * m2y,
          gb[G_VX] * m1x + gb[G_VY] * m1y, gb[G_VX] * m2x + gb[G_VY] * m2y};
}
// weight of the cell (dx, dy) from the gradient's cell in di (along i) and
// dj (along j) of grad_direct
__host__ __device__ constexpr double grad_wi(int dx, int dy) {
  return dy != 0 ? 0.0 : dx == 1 ? Coef::b1 : dx == -1 ? -Coef::b1 : dx == 2 ? Coef::b2 : dx == -2 ? -Coef::b2 : 0.0;
}
__host__ __device__ constexpr double grad_wj(int dx, int dy) { return grad_wi(dy, dx); }
// primitives of one cell (prim_of) with the intermediates its reverse needs;
// cvm1 = 1 / cv and betas (Sutherland) are the face's constants
struct CellPrim {
  double rom1, u, v, ww, ec, e, t, p, h;
};
__device__ inline CellPrim cell_prim(const Phys& ph, double cvm1, const double w[5]) {
  CellPrim c;
  c.rom1 = 1.0 / w[0];
  c.u = w[1] * c.rom1, c.v = w[2] * c.rom1, c.ww = w[3] * c.rom1;
  c.ec = 0.5 * (c.u * c.u + c.v * c.v + c.ww * c.ww);
  c.e = (w[4] - c.ec * w[0]) * c.rom1;
  c.t = c.e * cvm1;
  c.p = (ph.gam - 1.0) * w[0] * c.e;
  c.h = (w[4] + c.p) * c.rom1;
  return c;
}
// wb += (d prim / d w)^T pb for one cell of state w (prim_of reversed by
// hand; sqrt has derivative 0 at 0, as for Dual)
__device__ inline void prim_reverse(const Phys& ph, double cvm1, double betas, const double w[5], const CellPrim& c,
                                    const double pb[NPRIM], double wb[5]) {
  const double ro = w[0], ru = w[1], rv = w[2], rw = w[3], re = w[4];
  // mu = (betas / (t + S)) sqrt(t) t: d mu / dt = A B (1.5 - t / (t + S))
  const double inv = 1.0 / (c.t + ph.s_suth);
  const double tb = pb[P_T] + pb[P_MU] * ((betas * inv) * sqrt(c.t) * (1.5 - c.t * inv));
  const double hb = pb[P_H];
  const double pbt = pb[P_P] + hb * c.rom1;
  double reb = hb * c.rom1;
  double rom1b = hb * (re + c.p);
  const double eb = pbt * (ph.gam - 1.0) * ro + tb * cvm1;
  double rob = pbt * (ph.gam - 1.0) * c.e;
  reb += eb * c.rom1;
  const double ecb = -eb * ro * c.rom1;
  rob -= eb * c.ec * c.rom1;
  rom1b += eb * (re - c.ec * ro);
  const double ub = pb[P_U] + ecb * c.u, vb = pb[P_V] + ecb * c.v, wwb = pb[P_W] + ecb * c.ww;
  rom1b += ub * ru + vb * rv + wwb * rw;
  rob -= rom1b * c.rom1 * c.rom1;
  wb[0] += rob;
  wb[1] += ub * c.rom1;
  wb[2] += vb * c.rom1;
  wb[3] += wwb * c.rom1;
  wb[4] += reb;
}

// weight of the cell `along` cells from the face (0 and -1 adjacent) in the
// Euler flux: ROW 0 the centred 6-point flux (euler_o6_i.F), ROW 2, 3 the
// off-centred j-flux of those rows over cell rows 1..5 (jeuler)
template <int ROW>
__host__ __device__ constexpr double euler_coef(int along) {
  if constexpr (ROW == 0) {
    return (along == 0 || along == -1) ? Coef::c1 : (along == 1 || along == -2) ? Coef::c2
         : (along == 2 || along == -3) ? Coef::c3 : 0.0;
  } else {
    const int r = along - (1 - ROW);
    const double k2[5] = {12.0 / 60.0, 77.0 / 60.0, -43.0 / 60.0, 17.0 / 60.0, -3.0 / 60.0};
    const double k3[5] = {-3.0 / 60.0, 27.0 / 60.0, 47.0 / 60.0, -13.0 / 60.0, 2.0 / 60.0};
    return (r < 0 || r > 4) ? 0.0 : ROW == 2 ? k2[r] : k3[r];
  }
}

// The face adjoint of one face for the offsets [O0, O1): J = false an
// i-face (compact viscous flux), true a centred j-face.  hb: the face's
// adjoint hnb; out: out[(o * 5 + q) * nf + f].
template <bool J, bool O4, int ROW, int O0, int O1>
__device__ inline void face_reverse(const Mesh& m, const Phys& ph, const Metric& g, const double* w,
                                    const double* prim, const double* grad, int i, int j, const double hb[5],
                                    double* out, long nf, long f) {
  constexpr int DX = J ? 0 : 1, DY = J ? 1 : 0, D = J ? 1 : 0;
  PlainAcc<double> a;
  a.m = m; a.ph = ph; a.g = g; a.i = i; a.j = j;
  a.w = w; a.prim = prim; a.grad = grad;
  const double nxf = a.NX(0, 0, D), nyf = a.NY(0, 0, D);
  const DissPrim ds = diss_primal<DX, DY>(a, D);
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
  const ViscPrimal vp = visc_primal<J, O4>(a);
  double fb[5], gb[5];
#pragma unroll
  for (int e = 0; e < 5; ++e) {
    fb[e] = e == 0 ? 0.0 : -hb[e] * sn * nxl;
    gb[e] = e == 0 ? 0.0 : -hb[e] * sn * nyl;
  }
  const ViscAdj va = visc_reverse(vp, fb, gb, ph.cp / ph.prandtl);
  const DissAdj da = diss_adjoint(ds, ph, rspb, e2b, e4b);
  // The gradients' adjoints as weights of u and v along i and j, on the four
  // cells P_k = (k DX, k DY), k = -2..1, whose 4th-order gradients they
  // reach: cell 0's and cell -1's own gradients, or where one of them is on
  // the extrapolated ghost layer of SeededAcc::G (i = 0, i = im+1, j = jm+1;
  // faces i = 1, i = im+1, j = jm+1) the two gradients 2 g(a) - g(b) it is
  // extrapolated from.
  const bool ext0 = J ? j == m.jm + 1 : i == m.im + 1;  // cell 0 at i = im+1 or j = jm+1
  const bool ext1 = !J && i == 1;                         // cell -1 at i = 0
  double gbp[4][NGRAD];
#pragma unroll
  for (int n = 0; n < NGRAD; ++n) {
    gbp[0][n] = ext0 ? -da.g0[n] : 0.0;
    gbp[1][n] = (ext0 ? 2.0 * da.g0[n] : 0.0) + (ext1 ? 0.0 : da.g1[n]);
    gbp[2][n] = (ext0 ? 0.0 : da.g0[n]) + (ext1 ? 2.0 * da.g1[n] : 0.0);
    gbp[3][n] = ext1 ? -da.g1[n] : 0.0;
  }
  GradW gw[4];
  sfor<0, 4>([&](auto K) {
    constexpr int k = decltype(K)::value;
    // only the cells a gradient of this face uses (the others may lie
    // beyond the ghost frame)
    const bool used = k == 2 ? !ext0 || ext1 : k == 1 ? !ext1 || ext0 : k == 0 ? ext0 : ext1;
    gw[k] = used ? grad_weights<(k - 2) * DX, (k - 2) * DY>(a, gbp[k]) : GradW{0.0, 0.0, 0.0, 0.0};
  });
  const double cvm1 = 1.0 / ph.cv;
  const double betas = ph.muref * (ph.tref + ph.cs) / (sqrt(ph.tref) * ph.tref);

  sfor<O0, O1>([&](auto O) {
    constexpr int o = decltype(O)::value;
    constexpr int SI = J ? joff_di(o) : ioff_di(o), SJ = J ? joff_dj(o) : ioff_dj(o);
    // the two extra offsets exist only on the faces the seeded passes store
    // them for (reached through the extrapolated gradients)
    if constexpr (o >= 22) {
      if (!(J ? joff_live(m, o, j) : ioff_live(m, o, i))) return;
    }
    // the cell's state and primitives (plain FP64)
    const long cc = cidx(m, i + SI, j + SJ);
    double wc[5];
#pragma unroll
    for (int q = 0; q < 5; ++q) wc[q] = w[cc + (long)q * m.ncell];
    const CellPrim pc = cell_prim(ph, cvm1, wc);
    double wb[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    double pb[NPRIM];
#pragma unroll
    for (int k = 0; k < NPRIM; ++k) pb[k] = 0.0;
    constexpr int along = J ? SJ : SI, across = J ? SI : SJ;
    // centred Euler flux: c (efx_e nx + efy_e ny) of the cell, weighted by hnb_e
    if constexpr (across == 0 && euler_coef<ROW>(along) != 0.0) {
      constexpr double c = euler_coef<ROW>(along);
      double ae[5];
#pragma unroll
      for (int e = 0; e < 5; ++e) ae[e] = c * hb[e];
      const double sv = ae[0] + ae[1] * pc.u + ae[2] * pc.v + ae[3] * pc.ww + ae[4] * pc.h;
      const double rn = nxf * wc[1] + nyf * wc[2];
      wb[1] += nxf * sv;
      wb[2] += nyf * sv;
      pb[P_U] += ae[1] * rn;
      pb[P_V] += ae[2] * rn;
      pb[P_W] += ae[3] * rn;
      pb[P_H] += ae[4] * rn;
      pb[P_P] += ae[1] * nxf + ae[2] * nyf;
    }
    IndAcc<SI, SJ> ind;
    ind.m = m; ind.ph = ph; ind.g = g; ind.i = i; ind.j = j;
    // jump and predictor of the dissipation (linear in w_e)
#pragma unroll
    for (int e = 0; e < 5; ++e) {
      const double cj = 0.5 * (ind.template W<0, 0>(e) - ind.template W<-DX, -DY>(e));
      const double cp = predictor<5, DX, DY>(ind, e);
      wb[e] -= hb[e] * ds.rspec * (ds.eps2 * cj + ds.eps4 * cp);
    }
    // the dissipation scalars through the cell's state, T, p and gradients
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
    sfor<0, 4>([&](auto K) {
      constexpr int k = decltype(K)::value;
      constexpr double wi = grad_wi(SI - (k - 2) * DX, SJ - (k - 2) * DY);
      constexpr double wj = grad_wj(SI - (k - 2) * DX, SJ - (k - 2) * DY);
      if constexpr (wi != 0.0 || wj != 0.0) {
        pb[P_U] += gw[k].cu_i * wi + gw[k].cu_j * wj;
        pb[P_V] += gw[k].cv_i * wi + gw[k].cv_j * wj;
      }
    });
    // the viscous flux: the cell's weight in each linear stencil
    {
      const auto qc = [&] {
        if constexpr (J && O4) return jvals_o4(ind, 0);
        else if constexpr (J) return jvals_o2(ind, 0);
        else if constexpr (O4) return ivals_o4(ind, 0);
        else return ivals_o2(ind, 0);
      }();
      constexpr double cm = mid_coef<J, O4>(SI, SJ);
#pragma unroll
      for (int k = 0; k < 4; ++k) pb[k] += qc.N * va.N[k] + qc.S * va.S[k] + qc.O * va.O[k] + qc.E * va.E[k];
      pb[P_U] += cm * va.mid[0];
      pb[P_V] += cm * va.mid[1];
      pb[P_W] += cm * va.mid[2];
      pb[P_MU] += cm * va.mid[3];
    }
    prim_reverse(ph, cvm1, betas, wc, pc, pb, wb);
    double* p = out + (long)o * 5 * nf + f;
#pragma unroll
    for (int q = 0; q < 5; ++q) p[q * nf] = wb[q];
    // one offset at a time: the scheduler would otherwise hoist every
    // offset's loads to the top, far beyond the register file
    __builtin_amdgcn_sched_barrier(0);
  });
}

// offsets per launch of the reverse kernels (the face-level part is
// recomputed per launch)
constexpr int REV_GROUP = 24;

// One thread per face of rows [j0, j1], i = i0, i0 + istep, ...: J = false
// i-faces (O4: rows j >= 3, the compact viscous flux; else rows 1, 2), J =
// true j-faces (ROW 0: rows j >= 4, centred Euler flux; ROW 3, 2: the
// off-centred rows 3 (compact viscous flux) and 2 (2nd-order)); offsets
// [O0, O1) of the face's stencil.
template <bool J, bool O4, int ROW, int O0 = 0, int O1 = NOFF>
__global__ void __launch_bounds__(256) k_face_reverse(Mesh m, Phys ph, Metric g, const double* __restrict__ w,
                                                      const double* __restrict__ prim,
                                                      const double* __restrict__ grad, double* __restrict__ out,
                                                      const double* __restrict__ resb, int j0, int j1, int i0,
                                                      int istep) {
  int bx, by;
  xcd_block(bx, by);
  const int i = i0 + (bx * (int)blockDim.x + (int)threadIdx.x) * istep;
  const int j = j0 + by * blockDim.y + threadIdx.y;
  if (i > (J ? m.im : m.im + 1) || j > j1) return;
  const long nf = J ? (long)m.im * (m.jm + 1) : (long)(m.im + 1) * m.jm;
  const long f = J ? (i - 1) + (long)m.im * (j - 1) : (i - 1) + (long)(m.im + 1) * (j - 1);
  double hb[5];
  face_adjoint(m, resb, i, j, J ? 0 : 1, J ? 1 : 0, hb);
  face_reverse<J, O4, ROW, O0, O1>(m, ph, g, w, prim, grad, i, j, hb, out, nf, f);
}

}  // namespace bca
