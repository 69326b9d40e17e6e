// Ghost-frame boundary conditions of the DNC5 path, templated on the scalar
// type (double for the state, Dual<N> for the linearised BCs that fill the
// ghost tangents of seeds).  One thread owns one boundary line l and fills its
// gh ghost layers in order, since layer de reads layer de-1.
//
//   inlet   : bc_supandsubinlet_2d   (srcfv/borders/bc_supandsubinlet.F90:2-171)
//   noref   : bc_no_reflexion_2d     (srcfv/borders/bc_no_reflexion.F90:8-152)
//   extrap2 : bc_extrapolate_o2_2d   (srcfv/borders/bc_extrapolate.F90:39-71)
//   wall    : bc_wall_viscous_adia_2d(srcfv/borders/bc_wall_viscous.F90:2-99)
//   join    : jn_match_2d            (srcfv/borders/jn_match.F90:3-66)
//
// Interfaces are the reference's boxes {imin, jmin, imax, jmax} with a side
// code loc (0 Ilo, 1 Ihi, 2 Jlo, 3 Jhi) decoded as srcfv/borders/init_2d.F.
#pragma once
#include "scheme.h"

namespace bca {

struct Side {
  int kdir, i0, j0, high, lmax;
};

__host__ __device__ inline Side decode_side(int loc, int imin, int jmin, int imax, int jmax) {
  Side s{0, 0, 0, 0, 0};
  switch (loc) {
    case 0: s.kdir = 1; s.i0 = 1; s.lmax = jmax - jmin + 1; break;
    case 1: s.kdir = 1; s.i0 = -1; s.lmax = jmax - jmin + 1; s.high = 1; break;
    case 2: s.kdir = 2; s.j0 = 1; s.lmax = imax - imin + 1; break;
    default: s.kdir = 2; s.j0 = -1; s.lmax = imax - imin + 1; s.high = 1; break;
  }
  return s;
}

struct Box {
  int imin, jmin, imax, jmax;
};

// ------------------------------------------------------------------ frames
// Storage coordinates (a, b) of the t-th cell of the frame of width W around
// the block (all cells within W of the array edge): rows b < W, rows
// b >= nj - W, then the left/right W columns of the rows in between.  A block
// too small for a frame of that width is taken whole, row by row.
__host__ __device__ inline bool frame_whole(const Mesh& m, int W) { return 2 * W >= m.ni || 2 * W >= m.nj; }
__host__ __device__ inline long frame_count(const Mesh& m, int W) {
  if (frame_whole(m, W)) return m.ncell;
  return 2L * W * m.ni + 2L * W * (m.nj - 2 * W);
}
__device__ inline void frame_cell(const Mesh& m, int W, long t, int& a, int& b) {
  if (frame_whole(m, W) || t < (long)W * m.ni) { b = (int)(t / m.ni); a = (int)(t % m.ni); return; }
  t -= (long)W * m.ni;
  if (t < (long)W * m.ni) { b = m.nj - W + (int)(t / m.ni); a = (int)(t % m.ni); return; }
  t -= (long)W * m.ni;
  b = W + (int)(t / (2 * W));
  const int s = (int)(t % (2 * W));
  a = s < W ? s : m.ni - 2 * W + s;
}
// the inverse: frame position t of the storage cell (a, b), which must lie in
// the frame
__host__ __device__ inline long frame_index(const Mesh& m, int W, int a, int b) {
  if (frame_whole(m, W) || b < W) return a + (long)m.ni * b;
  if (b >= m.nj - W) return (long)W * m.ni + a + (long)m.ni * (b - (m.nj - W));
  return 2L * W * m.ni + 2L * W * (b - W) + (a < W ? a : a - (m.ni - 2 * W));
}

// Frame read or written by the boundary conditions of a box on the block
// edge: the ghost layers plus 4 interior layers.
constexpr int BC_FRAME = 7;

// Element (i, j, e) of a state stored as the whole array (FR false: planes
// of ncell cells) or as the compact boundary frame of width BC_FRAME (FR
// true: planes of frame_count cells; for the seeded Dual<5> states that only
// the boundary conditions touch).
template <class T, bool FR = false>
struct StateView {
  Mesh m;
  T* w;
  long plane;
  __device__ StateView(const Mesh& mm, T* ww) : m(mm), w(ww), plane(FR ? frame_count(mm, BC_FRAME) : mm.ncell) {}
  __device__ T& operator()(int i, int j, int e) const {
    const long c = FR ? frame_index(m, BC_FRAME, i + m.gh - 1, j + m.gh - 1) : cidx(m, i, j);
    return w[c + (long)e * plane];
  }
};

// Characteristic sub/supersonic inflow.  Kept quirks of the reference:
// the Roe-state pressure lacks the 1/2 (:49), velo2 uses ww unsquared (:69),
// the ghost spanwise momentum is field(j, de, 4) (:160).
template <class T, bool FR = false>
__device__ inline void bc_inlet_line(const Mesh& m, T* __restrict__ wp, int loc, Box b,
                                     const double* __restrict__ field, int lm, const double* __restrict__ nxp,
                                     const double* __restrict__ nyp, double gam, int l) {
  const Side s = decode_side(loc, b.imin, b.jmin, b.imax, b.jmax);
  StateView<T, FR> w{m, wp};
  auto F = [&](int a, int d, int e) { return field[(a - 1) + (long)lm * ((d - 1) + (long)m.gh * e)]; };
  const int i = b.imin + (l - 1) * s.j0 * s.j0;
  const int j = b.jmin + (l - 1) * s.i0 * s.i0;
  const double g1 = gam - 1.0;
  T r1 = w(i, j, 0);
  T rm1 = 1.0 / r1;
  T u1 = w(i, j, 1) * rm1, v1 = w(i, j, 2) * rm1, z1 = w(i, j, 3) * rm1;
  T pr = g1 * (w(i, j, 4) - r1 * (u1 * u1 + v1 * v1 + z1 * z1));
  T hr = (w(i, j, 4) + pr) * rm1;
  const double rl = F(j, i, 0), rlm1 = 1.0 / rl;
  const double ul = F(j, i, 1) * rlm1, vl = F(j, i, 2) * rlm1, zl = F(j, i, 3) * rlm1;
  const double pl = g1 * (F(j, i, 4) - rl * (ul * ul + vl * vl + zl * zl));
  const double hl = (F(j, i, 4) + pl) * rlm1;
  T sr = dsqrt(r1 / rl);
  T rr = dsqrt(r1 * rl);
  T wgt = 1.0 / (sr + 1.0);
  T uu = (u1 * sr + ul) * wgt, vv = (v1 * sr + vl) * wgt, zz = (z1 * sr + zl) * wgt;
  T hh = (hr * sr + hl) * wgt;
  T q2 = uu * uu + vv * vv + zz;
  T a2 = g1 * (hh - 0.5 * q2);
  T mach = dsqrt(q2 / a2);
  const bool sub = val(mach) < 1.0;
  double nxn = 0.0, nyn = 0.0, em = 0.0, e0 = 0.0, ep = 0.0;
  if (sub) {
    const long f = nidx(m, i, j) + (long)(s.kdir - 1) * m.nnode;
    const double nxl = nxp[f], nyl = nyp[f];
    const double ns = 1.0 / sqrt(nxl * nxl + nyl * nyl);
    nxn = nxl * ns;
    nyn = nyl * ns;
    const double rc = val(rr * dsqrt(a2));
    const double rvn = val(rr * (uu * nxn + vv * nyn));
    em = 0.5 + copysign(0.5, rc - rvn);
    e0 = 0.5 + copysign(0.5, -rvn);
    ep = 0.5 + copysign(0.5, -rc - rvn);
  }
  for (int de = 1; de <= m.gh; ++de) {
    const int ig = i - de * s.i0, jg = j - de * s.j0;
#pragma unroll
    for (int e = 0; e < 5; ++e) w(ig, jg, e) = T(F(l, de, e));
    if (!sub) continue;
    const int is = i - (de - 1) * s.i0, js = j - (de - 1) * s.j0;
    // outer state d (the prescribed profile), inner state s (previous layer)
    T rd = w(ig, jg, 0);
    T rdm1 = 1.0 / rd;
    T ud = w(ig, jg, 1) * rdm1, vd = w(ig, jg, 2) * rdm1, zd = w(ig, jg, 3) * rdm1;
    T pd = g1 * (w(ig, jg, 4) - rd * 0.5 * (ud * ud + vd * vd + zd * zd));
    T vnd = ud * nxn + vd * nyn;
    T utd = ud - vnd * nxn, vtd = vd - vnd * nyn;
    T hd = (w(ig, jg, 4) + pd) * rdm1;
    T rs = w(is, js, 0);
    T rsm1 = 1.0 / rs;
    T us = w(is, js, 1) * rsm1, vs = w(is, js, 2) * rsm1, zs = w(is, js, 3) * rsm1;
    T ps = g1 * (w(is, js, 4) - rs * 0.5 * (us * us + vs * vs + zs * zs));
    T vns = us * nxn + vs * nyn;
    T uts = us - vns * nxn, vts = vs - vns * nyn;
    T hs = (w(is, js, 4) + ps) * rsm1;
    T sr2 = dsqrt(rs * rdm1);
    T rr2 = dsqrt(rs * rd);
    T wg2 = 1.0 / (sr2 + 1.0);
    T ua = (us * sr2 + ud) * wg2, va = (vs * sr2 + vd) * wg2, za = (zs * sr2 + zd) * wg2;
    T ka = 0.5 * (ua * ua + va * va + za * za);
    T ha = (hs * sr2 + hd) * wg2;
    T aa2 = g1 * (ha - ka);
    T rc = rr2 * dsqrt(aa2);
    T rcm1 = 1.0 / rc;
    T ut = e0 * uts + (1.0 - e0) * utd;
    T vt = e0 * vts + (1.0 - e0) * vtd;
    T am = em * (ps - rc * vns) + (1.0 - em) * (pd - rc * vnd);
    T ap = ep * (ps + rc * vns) + (1.0 - ep) * (pd + rc * vnd);
    T vn = (ap - am) * 0.5 * rcm1;
    T p = (ap + am) * 0.5;
    T bs = (p - ps) * rr2 * rr2 * rcm1 * rcm1 + rs;
    T b0 = (p - pd) * rr2 * rr2 * rcm1 * rcm1 + rd;
    T ro = e0 * bs + (1.0 - e0) * b0;
    T rou = ro * (ut + vn * nxn);
    T rov = ro * (vt + vn * nyn);
    T row = T(F(j, de, 3));
    w(ig, jg, 0) = ro;
    w(ig, jg, 1) = rou;
    w(ig, jg, 2) = rov;
    w(ig, jg, 3) = row;
    w(ig, jg, 4) = p / g1 + 0.5 * (rou * rou + rov * rov + row * row) / ro;
  }
}

template <class T, bool FR = false>
__global__ void k_bc_inlet(Mesh m, T* __restrict__ wbase, int loc, Box b, const double* __restrict__ field,
                           int lm, const double* __restrict__ nxp, const double* __restrict__ nyp,
                           double gam, long cstride) {
  const int l = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  if (l <= decode_side(loc, b.imin, b.jmin, b.imax, b.jmax).lmax)
    bc_inlet_line<T, FR>(m, wbase + (long)blockIdx.y * cstride, loc, b, field, lm, nxp, nyp, gam, l);
}

// Non-reflecting far field against the reference state wbd(l, :).
template <class T, bool FR = false>
__device__ inline void bc_noref_line(const Mesh& m, T* __restrict__ wp, const double* __restrict__ wbd, int lm,
                                     int loc, Box b, const double* __restrict__ nxp,
                                     const double* __restrict__ nyp, double gam, int l) {
  const Side s = decode_side(loc, b.imin, b.jmin, b.imax, b.jmax);
  StateView<T, FR> w{m, wp};
  const int i1 = s.i0 * s.i0, j1 = s.j0 * s.j0;
  const int i = b.imin + (l - 1) * j1;
  const int j = b.jmin + (l - 1) * i1;
  const double g1 = gam - 1.0;
  const double sens = (double)(s.i0 + s.j0);
  const long f = nidx(m, i + s.high * i1, j + s.high * j1) + (long)(s.kdir - 1) * m.nnode;
  const double nxl = nxp[f], nyl = nyp[f];
  const double ninv = 1.0 / sqrt(nxl * nxl + nyl * nyl);
  const double nxn = nxl * ninv * sens, nyn = nyl * ninv * sens;
  T r0 = w(i, j, 0);
  T r0m1 = 1.0 / r0;
  T u0 = w(i, j, 1) * r0m1, v0 = w(i, j, 2) * r0m1, z0 = w(i, j, 3) * r0m1;
  T p0 = g1 * (w(i, j, 4) - 0.5 * r0 * (u0 * u0 + v0 * v0 + z0 * z0));
  T rc = r0 * dsqrt(gam * p0 * r0m1);
  T rcm1 = 1.0 / rc;
  T rvn = w(i, j, 1) * nxn + w(i, j, 2) * nyn;
  const double em = 0.5 + copysign(0.5, val(rc - rvn));
  const double e0 = 0.5 + copysign(0.5, val(-rvn));
  const double ep = 0.5 + copysign(0.5, val(-rc - rvn));
  auto B = [&](int e) { return wbd[(l - 1) + (long)lm * e]; };
  const double rd = B(0), rdm1 = 1.0 / rd;
  const double ud = B(1) * rdm1, vd = B(2) * rdm1, zd = B(3) * rdm1;
  const double pd = g1 * (B(4) - rd * 0.5 * (ud * ud + vd * vd + zd * zd));
  const double vnd = ud * nxn + vd * nyn;
  const double utd = ud - vnd * nxn, vtd = vd - vnd * nyn;
  T ps = g1 * (w(i, j, 4) - r0 * 0.5 * (u0 * u0 + v0 * v0 + z0 * z0));
  T vns = u0 * nxn + v0 * nyn;
  T uts = u0 - vns * nxn, vts = v0 - vns * nyn;
  T ut = e0 * uts + (1.0 - e0) * utd;
  T vt = e0 * vts + (1.0 - e0) * vtd;
  T zt = e0 * z0 + (1.0 - e0) * zd;
  T am = em * (ps - rc * vns) + (1.0 - em) * (pd - rc * vnd);
  T ap = ep * (ps + rc * vns) + (1.0 - ep) * (pd + rc * vnd);
  T vn = (ap - am) * 0.5 * rcm1;
  T p = (ap + am) * 0.5;
  T bs = (p - ps) * r0 * r0 * rcm1 * rcm1 + r0;
  T b0 = (p - pd) * r0 * r0 * rcm1 * rcm1 + rd;
  T ro = e0 * bs + (1.0 - e0) * b0;
  T rom1 = 1.0 / ro;
  T rou = ro * (ut + vn * nxn);
  T rov = ro * (vt + vn * nyn);
  T row = ro * zt;
  T roe = p / g1 + 0.5 * rom1 * (rou * rou + rov * rov + row * row);
  // first ghost layer from the characteristics, the others linearly extrapolated
  T q[5] = {ro, rou, rov, row, roe};
  for (int de = 1; de <= m.gh; ++de) {
    const int ig = i - de * s.i0, jg = j - de * s.j0;
    const int ia = i - (de - 1) * s.i0, ja = j - (de - 1) * s.j0;
#pragma unroll
    for (int e = 0; e < 5; ++e) {
      w(ig, jg, e) = q[e];
      q[e] = 2.0 * q[e] - w(ia, ja, e);
    }
  }
}

template <class T, bool FR = false>
__global__ void k_bc_noref(Mesh m, T* __restrict__ wbase, const double* __restrict__ wbd, int lm, int loc,
                           Box b, const double* __restrict__ nxp, const double* __restrict__ nyp,
                           double gam, long cstride) {
  const int l = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  if (l <= decode_side(loc, b.imin, b.jmin, b.imax, b.jmax).lmax)
    bc_noref_line<T, FR>(m, wbase + (long)blockIdx.y * cstride, wbd, lm, loc, b, nxp, nyp, gam, l);
}

template <class T, bool FR = false>
__device__ inline void bc_extrap2_line(const Mesh& m, T* __restrict__ wp, int loc, Box b, int em, int l) {
  const Side s = decode_side(loc, b.imin, b.jmin, b.imax, b.jmax);
  StateView<T, FR> w{m, wp};
  const int i = b.imin + (l - 1) * s.j0 * s.j0;
  const int j = b.jmin + (l - 1) * s.i0 * s.i0;
  for (int de = 1; de <= m.gh; ++de)
    for (int e = 0; e < em; ++e)
      w(i - s.i0 * de, j - s.j0 * de, e) =
          2.0 * w(i - s.i0 * (de - 1), j - s.j0 * (de - 1), e) - w(i - s.i0 * (de - 2), j - s.j0 * (de - 2), e);
}

template <class T, bool FR = false>
__global__ void k_bc_extrap2(Mesh m, T* __restrict__ wbase, int loc, Box b, int em, long cstride) {
  const int l = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  if (l <= decode_side(loc, b.imin, b.jmin, b.imax, b.jmax).lmax)
    bc_extrap2_line<T, FR>(m, wbase + (long)blockIdx.y * cstride, loc, b, em, l);
}

// Adiabatic no-slip wall: first ghost from the 2nd-order zero-normal-gradient
// pressure and the isentropic density, the deeper ones mirror interior cells
// with reversed velocity.
template <class T, bool FR = false>
__device__ inline void bc_wall_line(const Mesh& m, T* __restrict__ wp, int loc, Box b, double gam, int l) {
  const Side s = decode_side(loc, b.imin, b.jmin, b.imax, b.jmax);
  StateView<T, FR> w{m, wp};
  const int i = b.imin + (l - 1) * s.j0 * s.j0;
  const int j = b.jmin + (l - 1) * s.i0 * s.i0;
  const double g1 = gam - 1.0;
  T re = w(i, j, 0);
  T rem1 = 1.0 / re;
  T ue = w(i, j, 1) * rem1, ve = w(i, j, 2) * rem1, ze = w(i, j, 3) * rem1;
  T pe = g1 * (w(i, j, 4) - 0.5 * re * (ue * ue + ve * ve));
  T r1 = w(i + s.i0, j + s.j0, 0);
  T r1m1 = 1.0 / r1;
  T u1 = w(i + s.i0, j + s.j0, 1) * r1m1, v1 = w(i + s.i0, j + s.j0, 2) * r1m1;
  T pe1 = g1 * (w(i + s.i0, j + s.j0, 4) - 0.5 * r1 * (u1 * u1 + v1 * v1));
  T pw = 1.125 * pe - 0.125 * pe1;
  T pi = (1.0 / 3.0) * (4.0 * pw - pe);
  T ri = dpow(pi / pe * dpow(re, gam), 1.0 / gam);
  T ui = -ue, vi = -ve, zi = -ze;
  T ei = pi / g1 + 0.5 * ri * (ui * ui + vi * vi);
  for (int de = 1; de <= m.gh; ++de) {
    const int ig = i - de * s.i0, jg = j - de * s.j0;
    w(ig, jg, 0) = ri;
    w(ig, jg, 1) = ri * ui;
    w(ig, jg, 2) = ri * vi;
    w(ig, jg, 3) = ri * zi;
    w(ig, jg, 4) = ei;
    const int ii = i + de * s.i0, jj = j + de * s.j0;
    ri = w(ii, jj, 0);
    T rim1 = 1.0 / ri;
    ui = -(w(ii, jj, 1) * rim1);
    vi = -(w(ii, jj, 2) * rim1);
    zi = -(w(ii, jj, 3) * rim1);
    ei = w(ii, jj, 4);
  }
}

template <class T, bool FR = false>
__global__ void k_bc_wall(Mesh m, T* __restrict__ wbase, int loc, Box b, double gam, long cstride) {
  const int l = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  if (l <= decode_side(loc, b.imin, b.jmin, b.imax, b.jmax).lmax)
    bc_wall_line<T, FR>(m, wbase + (long)blockIdx.y * cstride, loc, b, gam, l);
}

// Block copy between two boxes of one block (the periodic join of the O-mesh).
// prr/prd: {imin, jmin, imax, jmax}; tr: the index transform (tr[0] axis and
// sign of the receiving i, tr[1] of j).  em components.
__host__ __device__ inline int join_count(Box prd, int tr0, int tr1) {
  const int idir = tr0 >= 0 ? tr0 : -tr0, jdir = tr1 >= 0 ? tr1 : -tr1;
  const int plo[2] = {prd.imin, prd.jmin}, phi[2] = {prd.imax, prd.jmax};
  return (phi[idir - 1] - plo[idir - 1] + 1) * (phi[jdir - 1] - plo[jdir - 1] + 1);
}

template <class T, bool FR = false>
__device__ inline void join_cell(const Mesh& m, T* __restrict__ wp, Box prr, Box prd, int tr0, int tr1, int em,
                                 int ni_arr, long plane, int t) {
  const int istep = tr0 >= 0 ? 1 : -1, jstep = tr1 >= 0 ? 1 : -1;
  const int idir = tr0 >= 0 ? tr0 : -tr0, jdir = tr1 >= 0 ? tr1 : -tr1;
  const int plo[2] = {prd.imin, prd.jmin}, phi[2] = {prd.imax, prd.jmax};
  const int ni_ = phi[idir - 1] - plo[idir - 1] + 1;
  const int nj_ = phi[jdir - 1] - plo[jdir - 1] + 1;
  const int a = t % ni_, bq = t / ni_;          // receiving offsets (idd-1, jdd-1)
  const int ii = istep == 1 ? 1 + a : ni_ - a;  // loop counters of the donor walk
  const int jj = jstep == 1 ? 1 + bq : nj_ - bq;
  const int ir = prr.imin + a, jr = prr.jmin + bq;
  const int id = plo[idir - 1] + ii - 1, jd = plo[jdir - 1] + jj - 1;
  const int off = m.gh - 1;
  const long rdst = FR ? frame_index(m, BC_FRAME, ir + off, jr + off) : (ir + off) + (long)ni_arr * (jr + off);
  const long rsrc = FR ? frame_index(m, BC_FRAME, id + off, jd + off) : (id + off) + (long)ni_arr * (jd + off);
  for (int e = 0; e < em; ++e) wp[rdst + e * plane] = wp[rsrc + e * plane];
}

template <class T, bool FR = false>
__global__ void k_join(Mesh m, T* __restrict__ wbase, Box prr, Box prd, int tr0, int tr1, int em,
                       int ni_arr, int nj_arr, long plane, long cstride) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < join_count(prd, tr0, tr1))
    join_cell<T, FR>(m, wbase + (long)blockIdx.y * cstride, prr, prd, tr0, tr1, em, ni_arr, plane, t);
}

}  // namespace bca
