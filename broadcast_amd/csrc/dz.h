// Spanwise operators of the 3D linearisation about a 2D state and the J(beta)
// batch.
//
//   dz  = sum_n coeffs(e, n) d func_n(e)/dw . wd   (16 terms)
//         coeffs_5p_dz   srcfv/dz/coeffs_5p_dz.F90:10-174
//   dz2 = the same with 4 pointwise terms
//         coeffs_5p_dz2  srcfv/dz/coeffs_5p_dz2.F90
//
// Both are linear in wd.  dz at a cell reads wd there and, through the
// 4th-order velocity-tangent gradients, at the 8 cells (i +- 1, 2; j +- 1, 2)
// of a cross; dz2 reads wd at the cell only.  The tangent type TD is double
// for one direction (the f2py routines) or Lin<5> for the 5 variables of a
// seed color carried together (the batched assembly).
//
// Batched assembly (the seed loop of BROADCAST_npz.py:1230-1245): for every
// row cell and seed color, the seeded interior cell of the color in the cross
// (at most one: the cross spans 5 cells, the coloring period is 7) enters as
// unit tangents, ghost cells of the cross as their BC tangents T(color, g);
// the results go to the COO slots of computejacobianfromdz
// (misc/ComputeJacobian.f90:708-779).
//
// J(beta) = Jac/vol(row) - beta^2 Dzz + i beta Dz over the union pattern of
// the three pruned matrices (BROADCAST_npz.py:1206-1209, biglobal.py:82).
#pragma once
#include "csr.h"

namespace bca {

template <int N>
struct Lin {
  double d[N];
};
template <int N>
__device__ inline Lin<N> operator+(const Lin<N>& a, const Lin<N>& b) {
  Lin<N> r;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] + b.d[k];
  return r;
}
template <int N>
__device__ inline Lin<N> operator-(const Lin<N>& a, const Lin<N>& b) {
  Lin<N> r;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] - b.d[k];
  return r;
}
template <int N>
__device__ inline Lin<N> operator-(const Lin<N>& a) {
  Lin<N> r;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = -a.d[k];
  return r;
}
template <int N>
__device__ inline Lin<N> operator*(double s, const Lin<N>& a) {
  Lin<N> r;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = s * a.d[k];
  return r;
}
template <int N>
__device__ inline Lin<N> operator*(const Lin<N>& a, double s) {
  Lin<N> r;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * s;
  return r;
}
template <int N>
__device__ inline Lin<N> operator/(const Lin<N>& a, double s) {
  Lin<N> r;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] / s;
  return r;
}
template <class TD> __device__ inline TD td_zero();
template <> __device__ inline double td_zero<double>() { return 0.0; }
template <> __device__ inline Lin<5> td_zero<Lin<5>>() {
  Lin<5> r;
#pragma unroll
  for (int k = 0; k < 5; ++k) r.d[k] = 0.0;
  return r;
}

// Primal primitives of one cell (rhs/primvisc.F with phys/viscosity.F).
struct DzPrim {
  double ro, rom1, u, v, w, ec, eloc, t, p, mu;
};
__device__ inline DzPrim dz_prim(const Mesh& m, const Phys& ph, const double* w, long c, double betas) {
  const long n = m.ncell;
  DzPrim q;
  q.ro = w[c];
  q.rom1 = 1.0 / q.ro;
  q.u = w[c + n] * q.rom1;
  q.v = w[c + 2 * n] * q.rom1;
  q.w = w[c + 3 * n] * q.rom1;
  q.ec = 0.5 * (q.u * q.u + q.v * q.v + q.w * q.w);
  q.eloc = (w[c + 4 * n] - q.ec * q.ro) * q.rom1;
  q.t = q.eloc * (1.0 / ph.cv);
  q.p = (ph.gam - 1.0) * q.ro * q.eloc;
  q.mu = betas / (q.t + ph.s_suth) * sqrt(q.t) * q.t;
  return q;
}

// Tangent primitives of one cell (dz/function_5p_dz_d.f90:157-205).
template <class TD>
struct DzTan {
  TD ud, vd, wd, td, pd, mud;
};
template <class TD>
__device__ inline DzTan<TD> dz_tan(const Phys& ph, const DzPrim& q, const double* w, long c, long n,
                                   const TD wdv[5], bool full, double betas) {
  DzTan<TD> r;
  const TD rod = wdv[0];
  const TD rom1d = -(rod / (q.ro * q.ro));
  r.ud = q.rom1 * wdv[1] + w[c + n] * rom1d;
  r.vd = q.rom1 * wdv[2] + w[c + 2 * n] * rom1d;
  r.wd = q.rom1 * wdv[3] + w[c + 3 * n] * rom1d;
  if (!full) return r;
  const TD ecd = 0.5 * ((2 * q.u) * r.ud + (2 * q.v) * r.vd + (2 * q.w) * r.wd);
  const double temp = w[c + 4 * n] - q.ec * q.ro;
  const TD elocd = q.rom1 * (wdv[4] - q.ro * ecd - q.ec * rod) + temp * rom1d;
  r.td = (1.0 / ph.cv) * elocd;
  r.pd = (ph.gam - 1.0) * (q.eloc * rod + q.ro * elocd);
  const double tl = q.t;
  const double r1 = sqrt(tl);
  const TD r1d = tl == 0.0 ? td_zero<TD>() : r.td / (2.0 * r1);
  const double t2 = r1 * tl / (ph.s_suth + tl);
  r.mud = betas * (tl * r1d + (r1 - t2) * r.td) / (ph.s_suth + tl);
  return r;
}

// dz and dz2 of interior cell (i, j); wdacc(ci, cj, wdv) fills the 5
// tangent inputs of cell (ci, cj).  Either output may be skipped.
template <class TD, class WD>
__device__ inline void dz_cell(const Mesh& m, const Phys& ph, const Metric& g, const double* w, int i, int j,
                               const WD& wdacc, TD* dz, TD* dz2) {
  const double betas = ph.muref * (ph.tref + ph.cs) / (sqrt(ph.tref) * ph.tref);
  const long n = m.ncell;
  const long c = cidx(m, i, j);
  const DzPrim q0 = dz_prim(m, ph, w, c, betas);
  TD wdv[5];
  wdacc(i, j, wdv);
  const DzTan<TD> t0 = dz_tan<TD>(ph, q0, w, c, n, wdv, true, betas);
  if (dz2) {
    // dz/coeffs_dz2.F, function_5p_dz2_d.f90:160-184, dz/sum_coeffs_dz2.F
    const double mu = q0.mu;
    const double cpprandtl = ph.cp / ph.prandtl;
    const double twothird = 2.0 / 3.0;
    dz2[0] = 0.0 * td_zero<TD>();
    dz2[1] = -mu * t0.ud;
    dz2[2] = -mu * t0.vd;
    dz2[3] = -(2.0 * twothird * mu) * t0.wd;
    dz2[4] = (-mu * cpprandtl) * t0.td + (-mu * q0.u) * t0.ud + (-mu * q0.v) * t0.vd +
             (-(2.0 * twothird * mu * q0.w)) * t0.wd;
  }
  if (!dz) return;
  // 4th-order gradients of the primal velocities and mu, and of the velocity
  // tangents (rhs/gradop_5p[ij].F, gradient.F)
  const double b1 = Coef::b1, b2 = Coef::b2;
  double pu[4], pv[4], pw[4], pm[4], qu[4], qv[4], qw[4], qm[4];  // i-1, i+1, i-2, i+2 ; j likewise
  TD tu[4], tv[4], tw[4], su[4], sv[4], sw[4];
  const int di[4] = {-1, 1, -2, 2};
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    {
      const long cc = cidx(m, i + di[s], j);
      const DzPrim q = dz_prim(m, ph, w, cc, betas);
      pu[s] = q.u; pv[s] = q.v; pw[s] = q.w; pm[s] = q.mu;
      TD x[5];
      wdacc(i + di[s], j, x);
      const DzTan<TD> t = dz_tan<TD>(ph, q, w, cc, n, x, false, betas);
      tu[s] = t.ud; tv[s] = t.vd; tw[s] = t.wd;
    }
    {
      const long cc = cidx(m, i, j + di[s]);
      const DzPrim q = dz_prim(m, ph, w, cc, betas);
      qu[s] = q.u; qv[s] = q.v; qw[s] = q.w; qm[s] = q.mu;
      TD x[5];
      wdacc(i, j + di[s], x);
      const DzTan<TD> t = dz_tan<TD>(ph, q, w, cc, n, x, false, betas);
      su[s] = t.ud; sv[s] = t.vd; sw[s] = t.wd;
    }
  }
  auto g4 = [&](const double* a) { return b1 * (a[1] - a[0]) + b2 * (a[3] - a[2]); };
  auto g4t = [&](const TD* a) { return b1 * (a[1] - a[0]) + b2 * (a[3] - a[2]); };
  const double volm1 = 1.0 / g.vol[c];
  const long f0 = nidx(m, i, j);
  const double dxm1 = 0.5 * (g.nx[f0] + g.nx[f0 + 1]) * volm1;
  const double dxm2 = 0.5 * (g.nx[f0 + m.nnode] + g.nx[f0 + m.ni + 1 + m.nnode]) * volm1;
  const double dym1 = 0.5 * (g.ny[f0] + g.ny[f0 + 1]) * volm1;
  const double dym2 = 0.5 * (g.ny[f0 + m.nnode] + g.ny[f0 + m.ni + 1 + m.nnode]) * volm1;
  const double gui = g4(pu), gvi = g4(pv), gwi = g4(pw), gmui = g4(pm);
  const double guj = g4(qu), gvj = g4(qv), gwj = g4(qw), gmuj = g4(qm);
  const double gu1 = dxm1 * gui + dxm2 * guj, gv1 = dxm1 * gvi + dxm2 * gvj;
  const double gw1 = dxm1 * gwi + dxm2 * gwj, gm1 = dxm1 * gmui + dxm2 * gmuj;
  const double gu2 = dym1 * gui + dym2 * guj, gv2 = dym1 * gvi + dym2 * gvj;
  const double gw2 = dym1 * gwi + dym2 * gwj, gm2 = dym1 * gmui + dym2 * gmuj;
  const TD guid = g4t(tu), gvid = g4t(tv), gwid = g4t(tw);
  const TD gujd = g4t(su), gvjd = g4t(sv), gwjd = g4t(sw);
  const TD gud1 = dxm1 * guid + dxm2 * gujd, gvd1 = dxm1 * gvid + dxm2 * gvjd, gwd1 = dxm1 * gwid + dxm2 * gwjd;
  const TD gud2 = dym1 * guid + dym2 * gujd, gvd2 = dym1 * gvid + dym2 * gvjd, gwd2 = dym1 * gwid + dym2 * gwjd;
  // the tangent function's own mu (betas * sqrt(T) T / (S + T))
  const double r1 = sqrt(q0.t);
  const double fmu = betas * (r1 * q0.t / (ph.s_suth + q0.t));
  const double mu = q0.mu, u = q0.u, v = q0.v, wz = q0.w;
  const double twothird = 2.0 / 3.0;
  // func*d (function_5p_dz_d.f90:281-400) and the sums of dz/sum_coeffs_dz.F
  // with the coefficients of dz/coeffs_dz.F; the terms whose coefficient or
  // tangent is identically zero are left out
  const TD f02 = wz * wdv[1] + w[c + n] * t0.wd - gw1 * t0.mud - fmu * gwd1;
  const TD f03 = wz * wdv[2] + w[c + 2 * n] * t0.wd - gw2 * t0.mud - fmu * gwd2;
  const double s12 = gu1 + gv2;
  const TD f04 = wz * wdv[3] + w[c + 3 * n] * t0.wd + t0.pd + twothird * (s12 * t0.mud + fmu * (gud1 + gvd2));
  const TD f05 = wz * (wdv[4] + t0.pd) + (w[c + 4 * n] + q0.p) * t0.wd;
  dz[0] = wdv[3];
  dz[1] = f02 + (twothird * gm1) * t0.wd + (twothird * mu) * gwd1;
  dz[2] = f03 + (twothird * gm2) * t0.wd + (twothird * mu) * gwd2;
  dz[3] = f04 + (-gm1) * t0.ud + (-gm2) * t0.vd + (-mu) * (gud1 + gvd2);
  dz[4] = f05 + (twothird * (gm1 * u + mu * gu1)) * t0.wd + (twothird * mu * u) * gwd1 +
          (-(gm1 * wz + mu * gw1)) * t0.ud + (-mu * wz) * gud1 + (twothird * (gm2 * v + mu * gv2)) * t0.wd +
          (twothird * mu * v) * gwd2 + (-(gm2 * wz + mu * gw2)) * t0.vd + (-mu * wz) * gvd2 +
          (-mu * gw1) * t0.ud + (-u) * (gw1 * t0.mud + fmu * gwd1) + (-mu * gw2) * t0.vd +
          (-v) * (gw2 * t0.mud + fmu * gwd2) + (twothird * mu * s12) * t0.wd + (twothird * wz * s12) * t0.mud +
          (twothird * mu * wz) * (gud1 + gvd2);
}

// ------------------------------------------------------------------ f2py form
struct ArrayTangent {
  Mesh m;
  const double* wd;
  __device__ void operator()(int ci, int cj, double out[5]) const {
    const long c = cidx(m, ci, cj);
#pragma unroll
    for (int q = 0; q < 5; ++q) out[q] = wd[c + q * m.ncell];
  }
};

__global__ void __launch_bounds__(256) k_dz(Mesh m, Phys ph, Metric g, const double* __restrict__ w,
                                            const double* __restrict__ wd, double* __restrict__ dz,
                                            double* __restrict__ dz2) {
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  if (i > m.im || j > m.jm) return;
  double a[5], b[5];
  dz_cell<double>(m, ph, g, w, i, j, ArrayTangent{m, wd}, dz ? a : nullptr, dz2 ? b : nullptr);
  const long c = cidx(m, i, j);
#pragma unroll
  for (int e = 0; e < 5; ++e) {
    if (dz) dz[c + e * m.ncell] = a[e];
    if (dz2) dz2[c + e * m.ncell] = b[e];
  }
}

// ------------------------------------------------------------------ scatter
// Window rules of computejacobianfromdz: symmetric in i and j, no wrap.
__device__ inline bool seed_column_dz(int i, int j, int l, int k, int gh, int im, int jm, int& vali, int& valj) {
  const int step = 2 * gh + 1;
  if (k <= gh) valj = (j <= k + 1 + gh) ? k : (j - (k + 1) + gh) / step * step + k;
  else valj = (j <= k - gh) ? jm + 1 : (j - (k + 1) + gh) / step * step + k;
  if (valj >= jm) return false;
  if (l <= gh) vali = (i <= l + 1 + gh) ? l : (i - (l + 1) + gh) / step * step + l;
  else vali = (i <= l - gh) ? im + 1 : (i - (l + 1) + gh) / step * step + l;
  return vali < im;
}

__global__ void k_scatter_dz(Mesh m, double* __restrict__ jac, int* __restrict__ ia, int* __restrict__ ja,
                             const double* __restrict__ dz, int mv, int l, int k) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long nint = (long)m.im * m.jm;
  if (q >= 5 * nint) return;
  const int e = (int)(q / nint);
  const long r = q % nint;
  const int i = 1 + (int)(r % m.im), j = 1 + (int)(r / m.im);
  const int step = 2 * m.gh + 1, im = m.im, jm = m.jm;
  const long cur = (long)(i - 1) + (long)(j - 1) * im + (long)e * im * jm + (long)k * im * jm * 5 +
                   (long)l * im * jm * 5 * step + (long)mv * im * jm * 5 * step * step;
  int vali, valj;
  if (!seed_column_dz(i, j, l, k, m.gh, im, jm, vali, valj)) {
    ia[cur] = 5 * im * jm - 1;
    ja[cur] = 0;
    jac[cur] = 0.0;
    return;
  }
  ia[cur] = e + (j - 1) * 5 + (i - 1) * jm * 5;
  ja[cur] = mv + valj * 5 + vali * jm * 5;
  jac[cur] = dz[cidx(m, i, j) + (long)e * m.ncell];
}

// ------------------------------------------------------------------ batched
// Tangent inputs of one seed color: e_q on the color's seeded interior cell
// (sc_i, sc_j) if any, the BC tangents T(color, g) on ghost cells, 0 else.
struct ColorTangent {
  Mesh m;
  int si, sj;  // seeded cell, si = 0 for none
  const double* Tc;
  const unsigned char* nzc;
  __device__ void operator()(int ci, int cj, Lin<5> out[5]) const {
#pragma unroll
    for (int q = 0; q < 5; ++q)
#pragma unroll
      for (int mm = 0; mm < 5; ++mm) out[q].d[mm] = 0.0;
    if (ci >= 1 && ci <= m.im && cj >= 1 && cj <= m.jm) {
      if (ci == si && cj == sj) {
#pragma unroll
        for (int q = 0; q < 5; ++q) out[q].d[q] = 1.0;
      }
      return;
    }
    const long gidx = ghost_index(m, ci + m.gh - 1, cj + m.gh - 1);
    if (!nzc[gidx]) return;
    const double* t = Tc + gidx * 25;
#pragma unroll
    for (int q = 0; q < 5; ++q)
#pragma unroll
      for (int mm = 0; mm < 5; ++mm) out[q].d[mm] = t[q * 5 + mm];
  }
};

// Grid (x: i, y: j, z: color).  ia/ja are those of both operators.
__global__ void __launch_bounds__(256) k_dz_colored(Mesh m, Phys ph, Metric g, const double* __restrict__ w,
                                                    const double* __restrict__ T,
                                                    const unsigned char* __restrict__ nz,
                                                    double* __restrict__ jdz, double* __restrict__ jdz2,
                                                    int* __restrict__ ia, int* __restrict__ ja, int write_index) {
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  const int color = blockIdx.z;
  if (i > m.im || j > m.jm) return;
  const int im = m.im, jm = m.jm, gh = m.gh, step = 2 * gh + 1;
  const int l = color / step, k = color % step;
  const long ng = ghost_count(m);
  const double* Tc = T + (long)color * ng * 25;
  const unsigned char* nzc = nz + (long)color * ng;
  // the color's cell in the cross, if interior
  int di = ((l - (i - 1)) % step + step) % step, dj = ((k - (j - 1)) % step + step) % step;
  if (di > gh) di -= step;
  if (dj > gh) dj -= step;
  const bool cross = (dj == 0 && di >= -2 && di <= 2) || (di == 0 && dj >= -2 && dj <= 2);
  const bool seeded = cross && i + di >= 1 && i + di <= im && j + dj >= 1 && j + dj <= jm;
  bool ghosts = false;
  const int cdi[8] = {-1, 1, -2, 2, 0, 0, 0, 0}, cdj[8] = {0, 0, 0, 0, -1, 1, -2, 2};
#pragma unroll
  for (int s = 0; s < 8; ++s) {
    const int ci = i + cdi[s], cj = j + cdj[s];
    if (ci < 1 || ci > im || cj < 1 || cj > jm) ghosts = ghosts || nzc[ghost_index(m, ci + gh - 1, cj + gh - 1)];
  }
  Lin<5> a[5], b[5];
  if (seeded || ghosts) {
    const ColorTangent acc{m, seeded ? i + di : 0, seeded ? j + dj : 0, Tc, nzc};
    dz_cell<Lin<5>>(m, ph, g, w, i, j, acc, jdz ? a : nullptr, jdz2 ? b : nullptr);
  } else {
#pragma unroll
    for (int e = 0; e < 5; ++e) a[e] = b[e] = td_zero<Lin<5>>();
  }
  int vali, valj;
  const bool ok = seed_column_dz(i, j, l, k, gh, im, jm, vali, valj);
  const long plane = (long)im * jm;
  const long base = (long)(i - 1) + (long)(j - 1) * im + (long)k * plane * 5 + (long)l * plane * 5 * step;
#pragma unroll
  for (int e = 0; e < 5; ++e) {
    const int row = e + (j - 1) * 5 + (i - 1) * jm * 5;
#pragma unroll
    for (int mm = 0; mm < 5; ++mm) {
      const long cur = base + (long)e * plane + (long)mm * plane * 5 * step * step;
      if (jdz) jdz[cur] = ok ? a[e].d[mm] : 0.0;
      if (jdz2) jdz2[cur] = ok ? b[e].d[mm] : 0.0;
      if (write_index) {
        ia[cur] = ok ? row : 5 * im * jm - 1;
        ja[cur] = ok ? mm + valj * 5 + vali * jm * 5 : 0;
      }
    }
  }
}

// ------------------------------------------------------------------ J(beta)
// J(beta) rows for nb wavenumbers from the three COOs on the Jacobian's slot
// layout (plain/relaxed single-zone coloring): a slot is kept if any of
// |jac|, |dz|, |dz2| exceeds mini (k_csr_count over MaxAbs3), each term then
// pruned on its own; values re = jac/vol(row) - beta^2 dz2, im = beta dz,
// stored [b][nnz] as (re, im).
struct MaxAbs3 {
  const double *a, *b, *c;
  __device__ double operator[](long s) const { return fmax(fabs(a[s]), fmax(fabs(b[s]), fabs(c[s]))); }
};

// Pass 1 is k_csr_fill<1, true> (csr.h): the column indices and the three
// pruned components on the pattern, comp[p] = (jac/vol(row), dz, dz2) (zero
// where pruned), from LDS tiles of the three COOs.

// Pass 2, element-wise and coalesced: vals[b][p] = (jv - beta_b^2 dz2, beta_b dz).
__global__ void __launch_bounds__(256) k_jbeta_values(const double* __restrict__ comp, long nnz,
                                                      const double* __restrict__ beta, int nb,
                                                      double* __restrict__ vals) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nnz) return;
  const double jv = comp[p], d = comp[nnz + p], d2 = comp[2 * nnz + p];
  for (int q = 0; q < nb; ++q) {
    const double bq = beta[q];
    double2 r;
    r.x = jv - (bq * bq) * d2;
    r.y = bq * d;
    reinterpret_cast<double2*>(vals)[(long)q * nnz + p] = r;
  }
}

}  // namespace bca
