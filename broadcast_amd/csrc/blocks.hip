// One face-kernel variant per compilation (selected by BCA_VARIANT): the 24
// seeded face instantiations of each variant are large, so the build compiles
// the variants as separate objects in parallel.  Variants 0-4 store Jacobian
// blocks, 5-9 the VJP contributions of the same passes.
#include <hip/hip_runtime.h>

#include "blocks.h"
#include "blockskern.h"

namespace bca {

#define BCA_ARGS a.m, a.ph, a.g, a.w, a.prim, a.grad, a.out, a.resb, a.j0, a.j1

// launch(integral_constant<int, g>) for every offset group g
template <int G = 0, class L>
void for_groups(L&& launch) {
  if constexpr (G < OFF_GROUPS) {
    launch(std::integral_constant<int, G>{});
    for_groups<G + 1>(launch);
  }
}

#define BCA_LAUNCH(name, kern, ...)                                                      \
  void name(dim3 gr, dim3 bl, hipStream_t s, const BlockArgs& a) {                       \
    for_groups([&](auto G) {                                                             \
      constexpr int o0 = decltype(G)::value * OFF_GROUP;                                 \
      kern<__VA_ARGS__, o0, o0 + OFF_GROUP><<<gr, bl, 0, s>>>(BCA_ARGS);                 \
    });                                                                                  \
  }

#if BCA_VARIANT == 0
BCA_LAUNCH(launch_iface_blocks_o4, k_iface_blocks, true, MODE_BLOCKS)
#elif BCA_VARIANT == 1
BCA_LAUNCH(launch_iface_blocks_o2, k_iface_blocks, false, MODE_BLOCKS)
#elif BCA_VARIANT == 2
BCA_LAUNCH(launch_jface_blocks_c, k_jface_blocks, 0, MODE_BLOCKS)
#elif BCA_VARIANT == 3
BCA_LAUNCH(launch_jface_blocks_r3, k_jface_blocks, 3, MODE_BLOCKS)
#elif BCA_VARIANT == 4
BCA_LAUNCH(launch_jface_blocks_r2, k_jface_blocks, 2, MODE_BLOCKS)
BCA_LAUNCH(launch_jface_blocks_wall, k_jface_blocks, 1, MODE_BLOCKS)
#elif BCA_VARIANT == 5
BCA_LAUNCH(launch_iface_vjp_o4, k_iface_blocks, true, MODE_VJP)
#elif BCA_VARIANT == 6
BCA_LAUNCH(launch_iface_vjp_o2, k_iface_blocks, false, MODE_VJP)
#elif BCA_VARIANT == 7
BCA_LAUNCH(launch_jface_vjp_c, k_jface_blocks, 0, MODE_VJP)
#elif BCA_VARIANT == 8
BCA_LAUNCH(launch_jface_vjp_r3, k_jface_blocks, 3, MODE_VJP)
#elif BCA_VARIANT == 9
BCA_LAUNCH(launch_jface_vjp_r2, k_jface_blocks, 2, MODE_VJP)
BCA_LAUNCH(launch_jface_vjp_wall, k_jface_blocks, 1, MODE_VJP)
#else
#error "BCA_VARIANT must be 0..9"
#endif

}  // namespace bca
