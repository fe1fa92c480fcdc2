// Batching form: the single fused kernel k_fused<2, 1> (2 resident wavefronts per SIMD, GJK/EPA inside the launch).
#include "rigid_tiles.h"

namespace gsr {

const void* fused_kernel_w2_g1() { return (const void*)k_fused<2, 1>; }

void launch_fused_w2_g1(dim3 grid, unsigned lds, hipStream_t st, const char* table, uint32_t* queue, int persistent) {
  hipLaunchKernelGGL((k_fused<2, 1>), grid, dim3(kWave), lds, st, table, queue, persistent);
}

}  // namespace gsr
