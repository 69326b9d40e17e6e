// One face-kernel variant per compilation (selected by BCA_VARIANT): the 24
// seeded face instantiations of each variant are large, so the build compiles
// the variants as separate objects in parallel.  Variants 0-4 store Jacobian
// blocks, 5 the VJP contributions of the wall faces by the same passes, 6 the
// reverse-mode face adjoints of the other faces.
#include <hip/hip_runtime.h>

#include "adjface.h"
#include "blocks.h"
#include "blockskern.h"

namespace bca {

#define BCA_ARGS a.m, a.ph, a.g, a.w, a.prim, a.grad, a.out, a.resb, a.j0, a.j1, a.i0, a.istep

// launch(integral_constant<int, o0>) for o0 = 0, GS, 2 GS, ... < NOFF
template <int GS, int O0 = 0, class L>
void for_groups(L&& launch) {
  if constexpr (O0 < NOFF) {
    launch(std::integral_constant<int, O0>{});
    for_groups<GS, O0 + GS>(launch);
  }
}

// Interior families (hundreds of thousands of faces) run one offset per
// launch; the edge families (one or two rows of faces) keep all 24 offsets in
// one launch, where 24 launches of a few blocks each would cost more in
// launch latency than the spills do.
#define BCA_LAUNCH(name, GS, kern, ...)                                                  \
  void name(dim3 gr, dim3 bl, hipStream_t s, const BlockArgs& a, hipStream_t s2) {       \
    static_assert(NOFF % (GS) == 0, "offset groups");                                    \
    for_groups<GS>([&](auto O) {                                                         \
      constexpr int o0 = decltype(O)::value;                                             \
      hipStream_t st = (s2 && (o0 / (GS)) % 2) ? s2 : s;                                 \
      kern<__VA_ARGS__, o0, o0 + (GS)><<<gr, bl, 0, st>>>(BCA_ARGS);                     \
    });                                                                                  \
  }

#if BCA_VARIANT == 0
BCA_LAUNCH(launch_iface_blocks_o4, OFF_GROUP, k_iface_blocks, true, MODE_BLOCKS)
#elif BCA_VARIANT == 1
BCA_LAUNCH(launch_iface_blocks_o2, NOFF, k_iface_blocks, false, MODE_BLOCKS)
#elif BCA_VARIANT == 2
BCA_LAUNCH(launch_jface_blocks_c, OFF_GROUP, k_jface_blocks, 0, MODE_BLOCKS)
#elif BCA_VARIANT == 3
BCA_LAUNCH(launch_jface_blocks_r3, NOFF, k_jface_blocks, 3, MODE_BLOCKS)
#elif BCA_VARIANT == 4
BCA_LAUNCH(launch_jface_blocks_r2, NOFF, k_jface_blocks, 2, MODE_BLOCKS)
BCA_LAUNCH(launch_jface_blocks_wall, NOFF, k_jface_blocks, 1, MODE_BLOCKS)
#elif BCA_VARIANT == 5
BCA_LAUNCH(launch_jface_vjp_wall, NOFF, k_jface_blocks, 1, MODE_VJP)
#elif BCA_VARIANT == 6
BCA_LAUNCH(launch_iface_vjp_rev, REV_GROUP, k_face_reverse, false, true, 0)
BCA_LAUNCH(launch_iface_vjp_rev_o2, REV_GROUP, k_face_reverse, false, false, 0)
BCA_LAUNCH(launch_jface_vjp_rev, REV_GROUP, k_face_reverse, true, true, 0)
BCA_LAUNCH(launch_jface_vjp_rev_r3, REV_GROUP, k_face_reverse, true, true, 3)
BCA_LAUNCH(launch_jface_vjp_rev_r2, REV_GROUP, k_face_reverse, true, false, 2)
#else
#error "BCA_VARIANT must be 0..6"
#endif

}  // namespace bca
