// One face-block kernel variant per compilation (selected by BCA_VARIANT):
// the 24 seeded face instantiations of each variant are large, so the build
// compiles the variants as separate objects in parallel.
#include <hip/hip_runtime.h>

#include "blocks.h"
#include "blockskern.h"

namespace bca {

#if BCA_VARIANT == 0
void launch_iface_blocks_o4(dim3 gr, dim3 bl, hipStream_t s, const BlockArgs& a) {
  k_iface_blocks<true><<<gr, bl, 0, s>>>(a.m, a.ph, a.g, a.w, a.prim, a.grad, a.blk, a.j0, a.j1);
}
#elif BCA_VARIANT == 1
void launch_iface_blocks_o2(dim3 gr, dim3 bl, hipStream_t s, const BlockArgs& a) {
  k_iface_blocks<false><<<gr, bl, 0, s>>>(a.m, a.ph, a.g, a.w, a.prim, a.grad, a.blk, a.j0, a.j1);
}
#elif BCA_VARIANT == 2
void launch_jface_blocks_c(dim3 gr, dim3 bl, hipStream_t s, const BlockArgs& a) {
  k_jface_blocks<0><<<gr, bl, 0, s>>>(a.m, a.ph, a.g, a.w, a.prim, a.grad, a.blk, a.j0, a.j1);
}
#elif BCA_VARIANT == 3
void launch_jface_blocks_r3(dim3 gr, dim3 bl, hipStream_t s, const BlockArgs& a) {
  k_jface_blocks<3><<<gr, bl, 0, s>>>(a.m, a.ph, a.g, a.w, a.prim, a.grad, a.blk, a.j0, a.j1);
}
#elif BCA_VARIANT == 4
void launch_jface_blocks_r2(dim3 gr, dim3 bl, hipStream_t s, const BlockArgs& a) {
  k_jface_blocks<2><<<gr, bl, 0, s>>>(a.m, a.ph, a.g, a.w, a.prim, a.grad, a.blk, a.j0, a.j1);
}
void launch_jface_blocks_wall(dim3 gr, dim3 bl, hipStream_t s, const BlockArgs& a) {
  k_jface_blocks<1><<<gr, bl, 0, s>>>(a.m, a.ph, a.g, a.w, a.prim, a.grad, a.blk, a.j0, a.j1);
}
#else
#error "BCA_VARIANT must be 0..4"
#endif

}  // namespace bca
