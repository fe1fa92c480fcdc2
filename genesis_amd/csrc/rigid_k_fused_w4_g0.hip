// Batching form: the single fused kernel k_fused<4, 0> (4 resident wavefronts per SIMD, GJK/EPA deferred to k_gjk_resume).
#include "rigid_tiles.h"

namespace gsr {

const void* fused_kernel_w4_g0() { return (const void*)k_fused<4, 0>; }

void launch_fused_w4_g0(dim3 grid, unsigned lds, hipStream_t st, const char* table, uint32_t* queue, int persistent) {
  hipLaunchKernelGGL((k_fused<4, 0>), grid, dim3(kWave), lds, st, table, queue, persistent);
}

}  // namespace gsr
