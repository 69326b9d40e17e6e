// Cell gather of the residual VJP.  The face kernels in VJP mode leave, for
// every face F and stencil offset o, c(F, o) = hnb(F)^T d hn(F) / d w(F + o)
// (5 values, one plane per offset).  The adjoint of cell c is the sum of the
// c(F, o) with F + o = c over both face families: a fixed-order gather, so
// the result does not depend on scheduling.
#pragma once
#include "blockskern.h"

namespace bca {

// Grid over the whole array (ghost frame included): wb is overwritten
// everywhere, as the reference zeroes it before accumulating
// (srcfv/adjoint/flux_num_dnc5_b.f90:834).
__global__ void __launch_bounds__(256) k_vjp_gather(Mesh m, const double* __restrict__ ci,
                                                    const double* __restrict__ cj, double* __restrict__ wb) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y * blockDim.y + threadIdx.y;
  if (a >= m.ni || b >= m.nj) return;
  const int i = a - m.gh + 1, j = b - m.gh + 1;
  const long nfi = (long)(m.im + 1) * m.jm, nfj = (long)m.im * (m.jm + 1);
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  sfor<0, NOFF>([&](auto O) {
    constexpr int o = decltype(O)::value;
    {
      const int fi = i - ioff_di(o), fj = j - ioff_dj(o);
      if (fi >= 1 && fi <= m.im + 1 && fj >= 1 && fj <= m.jm && ioff_live(m, o, fi)) {
        const double* p = ci + (long)o * 5 * nfi + (fi - 1) + (long)(m.im + 1) * (fj - 1);
#pragma unroll
        for (int q = 0; q < 5; ++q) acc[q] += p[q * nfi];
      }
    }
    {
      const int fi = i - joff_di(o), fj = j - joff_dj(o);
      if (fi >= 1 && fi <= m.im && fj >= 1 && fj <= m.jm + 1 && joff_live(m, o, fj)) {
        const double* p = cj + (long)o * 5 * nfj + (fi - 1) + (long)m.im * (fj - 1);
#pragma unroll
        for (int q = 0; q < 5; ++q) acc[q] += p[q * nfj];
      }
    }
  });
  const long c = a + (long)m.ni * b;
#pragma unroll
  for (int q = 0; q < 5; ++q) wb[c + q * m.ncell] = acc[q];
}

// residub is consumed: its interior is zeroed after use
// (srcfv/adjoint/flux_num_dnc5_b.f90:809-829).
__global__ void k_zero_interior(Mesh m, double* __restrict__ r) {
  const int i = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 1 + blockIdx.y * blockDim.y + threadIdx.y;
  if (i > m.im || j > m.jm) return;
  const long c = cidx(m, i, j);
#pragma unroll
  for (int e = 0; e < 5; ++e) r[c + e * m.ncell] = 0.0;
}

}  // namespace bca
