// Host launchers of the face-block kernels (each compiled in its own object).
#pragma once
#include <hip/hip_runtime.h>

#include "scheme.h"

namespace bca {

struct BlockArgs {
  Mesh m;
  Phys ph;
  Metric g;
  const double* w;
  const double* prim;
  const double* grad;
  double* out;          // blocks (25 planes per offset) or VJP contributions (5)
  const double* resb;   // residual adjoint (VJP only)
  int j0, j1;
  int i0 = 1, istep = 1;  // faces i = i0 + k istep (k = 0, 1, ...) of each row
};

// s2 (optional): every second offset group goes to s2 instead of s (the
// groups write planes of their own)
void launch_iface_blocks_o4(dim3 gr, dim3 bl, hipStream_t s, const BlockArgs& a, hipStream_t s2 = nullptr);
void launch_iface_blocks_o2(dim3 gr, dim3 bl, hipStream_t s, const BlockArgs& a, hipStream_t s2 = nullptr);
void launch_jface_blocks_c(dim3 gr, dim3 bl, hipStream_t s, const BlockArgs& a, hipStream_t s2 = nullptr);
void launch_jface_blocks_r3(dim3 gr, dim3 bl, hipStream_t s, const BlockArgs& a, hipStream_t s2 = nullptr);
void launch_jface_blocks_r2(dim3 gr, dim3 bl, hipStream_t s, const BlockArgs& a, hipStream_t s2 = nullptr);
void launch_jface_blocks_wall(dim3 gr, dim3 bl, hipStream_t s, const BlockArgs& a, hipStream_t s2 = nullptr);
void launch_jface_vjp_wall(dim3 gr, dim3 bl, hipStream_t s, const BlockArgs& a, hipStream_t s2 = nullptr);
void launch_iface_vjp_rev(dim3 gr, dim3 bl, hipStream_t s, const BlockArgs& a, hipStream_t s2 = nullptr);
void launch_iface_vjp_rev_o2(dim3 gr, dim3 bl, hipStream_t s, const BlockArgs& a, hipStream_t s2 = nullptr);
void launch_jface_vjp_rev(dim3 gr, dim3 bl, hipStream_t s, const BlockArgs& a, hipStream_t s2 = nullptr);
void launch_jface_vjp_rev_r3(dim3 gr, dim3 bl, hipStream_t s, const BlockArgs& a, hipStream_t s2 = nullptr);
void launch_jface_vjp_rev_r2(dim3 gr, dim3 bl, hipStream_t s, const BlockArgs& a, hipStream_t s2 = nullptr);

}  // namespace bca
