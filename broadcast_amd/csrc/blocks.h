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
  double* blk;
  int j0, j1;
};

void launch_iface_blocks_o4(dim3 gr, dim3 bl, hipStream_t s, const BlockArgs& a);
void launch_iface_blocks_o2(dim3 gr, dim3 bl, hipStream_t s, const BlockArgs& a);
void launch_jface_blocks_c(dim3 gr, dim3 bl, hipStream_t s, const BlockArgs& a);
void launch_jface_blocks_r3(dim3 gr, dim3 bl, hipStream_t s, const BlockArgs& a);
void launch_jface_blocks_r2(dim3 gr, dim3 bl, hipStream_t s, const BlockArgs& a);
void launch_jface_blocks_wall(dim3 gr, dim3 bl, hipStream_t s, const BlockArgs& a);

}  // namespace bca
