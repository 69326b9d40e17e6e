// Set-up, reduction and Jacobian-scatter kernels around the DNC5 residual:
//   geometry  computegeom_2d           (srcfv/geom/computegeom.F90:3-104)
//   norms     compute_norml2[inf]      (srcfv/norm.F90:2-77)
//   seeds     testvector               (misc/ComputeJacobian.f90:357-374)
//   scatter   computejacobianfromjv*   (misc/ComputeJacobian.f90:292-355, 503-570, 847-926)
//   set_bnd   set_bndbl_2d             (set_bnd.f90:2-24)
#pragma once
#include "scheme.h"

namespace bca {

// ------------------------------------------------------------------ geometry
// One workgroup walks the reference's statement sequence; every statement is
// an array assignment over one line of the grid, so threads split the line
// and a barrier separates statements.  Runs once per mesh.
__global__ void __launch_bounds__(1024) k_geom(Mesh m, double* x0, double* y0, double* nx,
                                               double* ny, double* xc, double* yc, double* vol,
                                               double* volf) {
  const int im = m.im, jm = m.jm, gh = m.gh;
  const int nin = m.ni + 1;  // node-plane leading dimension
  const long nn = m.nnode, nc = m.ncell;
  auto N = [&](double* a, int i, int j) -> double& { return a[(i + gh - 1) + (long)nin * (j + gh - 1)]; };
  auto Cc = [&](double* a, int i, int j) -> double& { return a[(i + gh - 1) + (long)m.ni * (j + gh - 1)]; };
  const int t = threadIdx.x, nt = blockDim.x;
  const int ilo = 1 - gh, ihi = im + gh + 1, jlo = 1 - gh, jhi = jm + gh + 1;
  for (int g = 1; g <= gh; ++g)
    for (int rep = 0; rep < 2; ++rep) {
      for (int j = jlo + t; j <= jhi; j += nt) {
        N(x0, 1 - g, j) = 2.0 * N(x0, 2 - g, j) - N(x0, 3 - g, j);
        N(x0, im + 1 + g, j) = 2.0 * N(x0, im + g, j) - N(x0, im - 1 + g, j);
      }
      __syncthreads();
      for (int i = ilo + t; i <= ihi; i += nt) {
        N(x0, i, 1 - g) = 2.0 * N(x0, i, 2 - g) - N(x0, i, 3 - g);
        N(x0, i, jm + 1 + g) = 2.0 * N(x0, i, jm + g) - N(x0, i, jm - 1 + g);
      }
      __syncthreads();
      for (int i = ilo + t; i <= ihi; i += nt) {
        N(y0, i, 1 - g) = 2.0 * N(y0, i, 2 - g) - N(y0, i, 3 - g);
        N(y0, i, jm + 1 + g) = 2.0 * N(y0, i, jm + g) - N(y0, i, jm - 1 + g);
      }
      __syncthreads();
      for (int j = jlo + t; j <= jhi; j += nt) {
        N(y0, 1 - g, j) = 2.0 * N(y0, 2 - g, j) - N(y0, 3 - g, j);
        N(y0, im + 1 + g, j) = 2.0 * N(y0, im + g, j) - N(y0, im - 1 + g, j);
      }
      __syncthreads();
    }
  // centres, two-triangle volumes and unnormalised face normals
  const int ncol = im + 1;
  for (long q = t; q < (long)(im + 1) * (jm + 1); q += nt) {
    const int i = 1 + (int)(q % ncol), j = 1 + (int)(q / ncol);
    Cc(xc, i, j) = 0.25 * (N(x0, i, j) + N(x0, i + 1, j) + N(x0, i, j + 1) + N(x0, i + 1, j + 1));
    Cc(yc, i, j) = 0.25 * (N(y0, i, j) + N(y0, i + 1, j) + N(y0, i, j + 1) + N(y0, i + 1, j + 1));
    const double xa = N(x0, i, j), ya = N(y0, i, j), xb = N(x0, i + 1, j), yb = N(y0, i + 1, j);
    const double xe = N(x0, i, j + 1), ye = N(y0, i, j + 1), xd = N(x0, i + 1, j + 1), yd = N(y0, i + 1, j + 1);
    const double q1 = 0.5 * fabs((xb - xa) * (ye - ya) - (xe - xa) * (yb - ya));
    const double q2 = 0.5 * fabs((xe - xd) * (yb - yd) - (xb - xd) * (ye - yd));
    Cc(vol, i, j) = q1 + q2;
    N(nx, i, j) = N(y0, i, j + 1) - N(y0, i, j);
    N(ny, i, j) = N(x0, i, j) - N(x0, i, j + 1);
    nx[nn + (i + gh - 1) + (long)nin * (j + gh - 1)] = N(y0, i, j) - N(y0, i + 1, j);
    ny[nn + (i + gh - 1) + (long)nin * (j + gh - 1)] = N(x0, i + 1, j) - N(x0, i, j);
  }
  __syncthreads();
  auto N2 = [&](double* a, int i, int j, int d) -> double& {
    return a[(long)d * nn + (i + gh - 1) + (long)nin * (j + gh - 1)];
  };
  for (int g = 1; g <= gh; ++g)
    for (int rep = 0; rep < 2; ++rep) {
      double* cf[2] = {xc, yc};
      for (int f = 0; f < 2; ++f) {
        double* a = cf[f];
        for (int j = 1 - gh + t; j <= jm + gh; j += nt) {
          Cc(a, 1 - g, j) = 2.0 * Cc(a, 2 - g, j) - Cc(a, 3 - g, j);
          Cc(a, im + g, j) = 2.0 * Cc(a, im - 1 + g, j) - Cc(a, im - 2 + g, j);
        }
        __syncthreads();
        for (int i = 1 - gh + t; i <= im + gh; i += nt) {
          Cc(a, i, 1 - g) = 2.0 * Cc(a, i, 2 - g) - Cc(a, i, 3 - g);
          Cc(a, i, jm + g) = 2.0 * Cc(a, i, jm - 1 + g) - Cc(a, i, jm - 2 + g);
        }
        __syncthreads();
      }
      for (int j = 1 + t; j <= jm; j += nt) {
        Cc(vol, 1 - g, j) = 2.0 * Cc(vol, 2 - g, j) - Cc(vol, 3 - g, j);
        Cc(vol, im + g, j) = 2.0 * Cc(vol, im - 1 + g, j) - Cc(vol, im - 2 + g, j);
      }
      __syncthreads();
      for (int i = 1 - gh + t; i <= im + gh; i += nt) {
        Cc(vol, i, 1 - g) = 2.0 * Cc(vol, i, 2 - g) - Cc(vol, i, 3 - g);
        Cc(vol, i, jm + g) = 2.0 * Cc(vol, i, jm - 1 + g) - Cc(vol, i, jm - 2 + g);
      }
      __syncthreads();
      for (int q = t; q < 2 * (jhi - jlo + 1); q += nt) {
        const int d = q / (jhi - jlo + 1), j = jlo + q % (jhi - jlo + 1);
        N2(nx, 1 - g, j, d) = 2.0 * N2(nx, 2 - g, j, d) - N2(nx, 3 - g, j, d);
        N2(ny, 1 - g, j, d) = 2.0 * N2(ny, 2 - g, j, d) - N2(ny, 3 - g, j, d);
        N2(nx, im + 1 + g, j, d) = 2.0 * N2(nx, im + g, j, d) - N2(nx, im - 1 + g, j, d);
        N2(ny, im + 1 + g, j, d) = 2.0 * N2(ny, im + g, j, d) - N2(ny, im - 1 + g, j, d);
      }
      __syncthreads();
      for (int q = t; q < 2 * (ihi - ilo + 1); q += nt) {
        const int d = q / (ihi - ilo + 1), i = ilo + q % (ihi - ilo + 1);
        N2(nx, i, 1 - g, d) = 2.0 * N2(nx, i, 2 - g, d) - N2(nx, i, 3 - g, d);
        N2(ny, i, 1 - g, d) = 2.0 * N2(ny, i, 2 - g, d) - N2(ny, i, 3 - g, d);
        N2(nx, i, jm + 1 + g, d) = 2.0 * N2(nx, i, jm + g, d) - N2(nx, i, jm - 1 + g, d);
        N2(ny, i, jm + 1 + g, d) = 2.0 * N2(ny, i, jm + g, d) - N2(ny, i, jm - 1 + g, d);
      }
      __syncthreads();
    }
  for (long q = t; q < (long)(im + 1) * (jm + 1); q += nt) {
    const int i = 1 + (int)(q % ncol), j = 1 + (int)(q / ncol);
    volf[cidx(m, i, j)] = 2.0 / (Cc(vol, i, j) + Cc(vol, i - 1, j));
    volf[cidx(m, i, j) + nc] = 2.0 / (Cc(vol, i, j) + Cc(vol, i, j - 1));
  }
}

// volf from the final volumes (the O-mesh recomputes it after the periodic
// copy of the ghost volumes, cylinder.py:524-527)
__global__ void k_volf(Mesh m, const double* __restrict__ vol, double* __restrict__ volf) {
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  if (i > m.im + 1 || j > m.jm + 1) return;
  volf[cidx(m, i, j)] = 2.0 / (vol[cidx(m, i, j)] + vol[cidx(m, i - 1, j)]);
  volf[cidx(m, i, j) + m.ncell] = 2.0 / (vol[cidx(m, i, j)] + vol[cidx(m, i, j - 1)]);
}

// ------------------------------------------------------------------ norms
// Per-variable sums of r^2 and r^10 over the interior: one partial per
// workgroup (wavefront shuffles, then LDS across the 4 waves), then a single
// workgroup folds the partials.  Deterministic for a fixed launch shape.
constexpr int NORM_THREADS = 256;

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__global__ void __launch_bounds__(NORM_THREADS) k_norm_partial(Mesh m, const double* __restrict__ r,
                                                              double* __restrict__ part) {
  double s[10];
#pragma unroll
  for (int k = 0; k < 10; ++k) s[k] = 0.0;
  const long ncells = (long)m.im * m.jm;
  for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < ncells; q += (long)gridDim.x * blockDim.x) {
    const int i = 1 + (int)(q % m.im), j = 1 + (int)(q / m.im);
    const long c = cidx(m, i, j);
#pragma unroll
    for (int e = 0; e < 5; ++e) {
      const double x = r[c + e * m.ncell];
      const double x2 = x * x;
      const double x4 = x2 * x2;
      s[e] += x2;
      s[5 + e] += x4 * x4 * x2;
    }
  }
  __shared__ double red[NORM_THREADS / 64][10];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < 10; ++k) {
    const double v = wave_sum(s[k]);
    if (lane == 0) red[wv][k] = v;
  }
  __syncthreads();
  if (threadIdx.x < 10) {
    double v = 0.0;
    for (int q = 0; q < NORM_THREADS / 64; ++q) v += red[q][threadIdx.x];
    part[(long)blockIdx.x * 10 + threadIdx.x] = v;
  }
}

__global__ void __launch_bounds__(64) k_norm_final(const double* __restrict__ part, int nparts,
                                                   double* __restrict__ out) {
  // out: norm[5] = sqrt(sum r^2), ninf[5] = (sum r^10)^0.1, nmoy[5] = norm / (im jm) set by host
  const int k = threadIdx.x;
  if (k >= 10) return;
  double v = 0.0;
  for (int q = 0; q < nparts; ++q) v += part[(long)q * 10 + k];
  out[k] = (k < 5) ? sqrt(v) : pow(v, 0.1);
}

// ------------------------------------------------------------------ seeds
__global__ void k_testvector(Mesh m, double* __restrict__ wd, int mv, int l, int k) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= 5 * m.ncell) return;
  const int e = (int)(c / m.ncell);
  const long cc = c % m.ncell;
  const int i = (int)(cc % m.ni) - m.gh + 1, j = (int)(cc / m.ni) - m.gh + 1;
  const int step = 2 * m.gh + 1;
  const bool on = e == mv && i >= l + 1 && i <= m.im && j >= k + 1 && j <= m.jm &&
                  (i - l - 1) % step == 0 && (j - k - 1) % step == 0;
  wd[c] = on ? 1.0 : 0.0;
}

// Window rules of the reference's scatter (misc/ComputeJacobian.f90:528-552):
// the column cell (0-based vali, valj) that seed color (l, k) attributes to
// row cell (i, j).  valj depends on (j, k) only, vali on (i, l) only.
__host__ __device__ inline int seed_valj(int j, int k, int gh) {
  const int step = 2 * gh + 1;
  return (j <= k + 1 + gh) ? k : (j - gh - (k + 1) + 2 * gh) / step * step + k;
}
__host__ __device__ inline bool seed_vali(int i, int l, int gh, int im, bool withjn, int& vali) {
  const int step = 2 * gh + 1;
  if (l <= gh) vali = (i <= l + 1 + gh) ? l : (i - (l + 1) + gh) / step * step + l;
  else if (i <= l - gh) vali = withjn ? im - 1 - 2 * gh + l : im + 1;
  else vali = (i - (l + 1) + gh) / step * step + l;
  if (vali >= im) {
    if (withjn && vali - im <= gh - 1) vali = l;
    else return false;
  }
  return true;
}
// false for the dummy slot
__host__ __device__ inline bool seed_column(int i, int j, int l, int k, int gh, int im, int jm,
                                            bool withjn, int& vali, int& valj) {
  valj = seed_valj(j, k, gh);
  if (valj >= jm) return false;
  return seed_vali(i, l, gh, im, withjn, vali);
}

// Two-zone variant (computejacobianfromjv_relaxed_withjnandcheck,
// misc/ComputeJacobian.f90:1095-1204) for an empty slot: zone 0 wraps the
// columns at or past im-2gh to l, zone 1 those at or below 2gh to
// im - mod(im, 2gh+1) + l (one period lower if that passes im-1).  `wrapped`
// reports which rule applied (the check resets an occupied wrapped slot of
// zone 0 and an occupied unwrapped slot of zone 1 to the dummy).
__host__ __device__ inline void seed_vali_zone(int i, int l, int gh, int im, int n, int& vali, bool& wrapped) {
  const int step = 2 * gh + 1;
  if (l <= gh) vali = (i <= l + 1 + gh) ? l : (i - (l + 1) + gh) / step * step + l;
  else vali = (i <= l - gh) ? l : (i - (l + 1) + gh) / step * step + l;
  if (n == 0) {
    wrapped = vali >= im - 2 * gh;
    if (wrapped) vali = l;
  } else {
    wrapped = vali <= 2 * gh;
    if (wrapped) {
      vali = im - im % step + l;
      if (vali > im - 1) vali -= step;
    }
  }
}
__host__ __device__ inline bool seed_column_zone(int i, int j, int l, int k, int gh, int im, int jm, int n,
                                                 int& vali, int& valj, bool& wrapped) {
  valj = seed_valj(j, k, gh);
  wrapped = false;
  if (valj >= jm) return false;
  seed_vali_zone(i, l, gh, im, n, vali, wrapped);
  return true;
}

// testvector_partial (misc/ComputeJacobian.f90:1075-1092): seeds of (m, l, k)
// restricted to the 1-based index box [i0, i1] x [j0, j1].
__global__ void k_testvector_partial(Mesh m, double* __restrict__ wd, int mv, int l, int k, int i0, int i1,
                                     int j0, int j1) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= 5 * m.ncell) return;
  const int e = (int)(c / m.ncell);
  const long cc = c % m.ncell;
  const int i = (int)(cc % m.ni) - m.gh + 1, j = (int)(cc / m.ni) - m.gh + 1;
  const int step = 2 * m.gh + 1;
  const bool on = e == mv && i >= i0 + l && i <= i1 && j >= j0 + k && j <= j1 && (i - i0 - l) % step == 0 &&
                  (j - j0 - k) % step == 0;
  wd[c] = on ? 1.0 : 0.0;
}

// One JVP result of seed (m, l, k) of zone n into slot block n of the doubled
// COO, writing only slots that hold |jac| < mini (the reference's check).
__global__ void k_scatter_jv_zone(Mesh m, double* __restrict__ jac, int* __restrict__ ia, int* __restrict__ ja,
                                  const double* __restrict__ resd, int mv, int l, int k,
                                  const double* __restrict__ coefdiag, double mini, int n) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long nint = (long)m.im * m.jm;
  if (q >= 5 * nint) return;
  const int e = (int)(q / nint);
  const long r = q % nint;
  const int i = 1 + (int)(r % m.im), j = 1 + (int)(r / m.im);
  const int step = 2 * m.gh + 1, im = m.im, jm = m.jm;
  const long cur = (long)(i - 1) + (long)(j - 1) * im + (long)e * im * jm + (long)k * im * jm * 5 +
                   (long)l * im * jm * 5 * step + (long)mv * im * jm * 5 * step * step +
                   (long)n * im * jm * 25 * step * step;
  const int row = e + (j - 1) * 5 + (i - 1) * jm * 5;
  const int dummy = 5 * im * jm - 2;
  ia[cur] = row;
  int vali, valj;
  bool wrapped;
  if (!seed_column_zone(i, j, l, k, m.gh, im, jm, n, vali, valj, wrapped)) {
    ia[cur] = dummy;
    ja[cur] = 0;
    jac[cur] = 0.0;
    return;
  }
  if (fabs(jac[cur]) < mini) {
    const int col = mv + valj * 5 + vali * jm * 5;
    ja[cur] = col;
    const double rv = resd[cidx(m, i, j) + (long)e * m.ncell];
    jac[cur] = (coefdiag && row == col) ? coefdiag[(i - 1) + (long)im * (j - 1)] - rv : -rv;
  } else if ((n == 0) == wrapped) {
    ia[cur] = dummy;
    ja[cur] = 0;
    jac[cur] = 0.0;
  }
}

// One JVP result resd of seed (m, l, k) into the reference's COO layout.
__global__ void k_scatter_jv(Mesh m, double* __restrict__ jac, int* __restrict__ ia,
                             int* __restrict__ ja, const double* __restrict__ resd, int mv, int l,
                             int k, const double* __restrict__ coefdiag, int withjn) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long nint = (long)m.im * m.jm;
  if (q >= 5 * nint) return;
  const int e = (int)(q / nint);
  const long r = q % nint;
  const int i = 1 + (int)(r % m.im), j = 1 + (int)(r / m.im);
  const int step = 2 * m.gh + 1, im = m.im, jm = m.jm;
  const long cur = (long)(i - 1) + (long)(j - 1) * im + (long)e * im * jm + (long)k * im * jm * 5 +
                   (long)l * im * jm * 5 * step + (long)mv * im * jm * 5 * step * step;
  const int row = e + (j - 1) * 5 + (i - 1) * jm * 5;
  int vali, valj;
  if (!seed_column(i, j, l, k, m.gh, im, jm, withjn != 0, vali, valj)) {
    ia[cur] = withjn ? 5 * im * jm - 2 : 5 * im * jm - 1;
    ja[cur] = 0;
    jac[cur] = 0.0;
    return;
  }
  const int col = mv + valj * 5 + vali * jm * 5;
  ia[cur] = row;
  ja[cur] = col;
  double v = -resd[cidx(m, i, j) + (long)e * m.ncell];
  if (coefdiag && row == col) v = coefdiag[(i - 1) + (long)im * (j - 1)] - resd[cidx(m, i, j) + (long)e * m.ncell];
  jac[cur] = v;
}

__global__ void k_set_bndbl(Mesh m, const double* __restrict__ w, double* __restrict__ field,
                            double* __restrict__ wbd) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int gh = m.gh, jm = m.jm, im = m.im;
  if (t < jm * gh) {
    const int j = 1 + t % jm, d = 1 + t / jm;
    for (int e = 0; e < 5; ++e)
      field[(j - 1) + (long)jm * ((d - 1) + (long)gh * e)] = w[cidx(m, 1 - d, j) + (long)e * m.ncell];
  } else if (t < jm * gh + im + gh) {
    const int i = 1 + t - jm * gh;
    for (int e = 0; e < 5; ++e) wbd[(i - 1) + (long)(im + gh) * e] = w[cidx(m, i - gh, jm) + (long)e * m.ncell];
  }
}

// ------------------------------------------------------------- dual packing
template <int N>
__global__ void k_pack(long n, const double* __restrict__ a, const double* __restrict__ ad,
                       Dual<N>* __restrict__ out) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  Dual<N> x(a[c]);
  if (ad) x.d[0] = ad[c];
  out[c] = x;
}

template <int N>
__global__ void k_unpack(long n, const Dual<N>* __restrict__ x, double* __restrict__ a,
                         double* __restrict__ ad) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  if (a) a[c] = x[c].v;
  if (ad) ad[c] = x[c].d[0];
}

}  // namespace bca
