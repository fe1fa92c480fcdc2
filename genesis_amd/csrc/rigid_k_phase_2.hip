// Batching form, phased launches: stage 2 (GJK/EPA pass of deferred envs) over every tile of the batch.
#include "rigid_tiles.h"

namespace gsr {

const void* phase_kernel_2() { return (const void*)k_phase<2>; }

void launch_phase_2(dim3 grid, unsigned lds, hipStream_t st, const char* table) {
  hipLaunchKernelGGL(k_phase<2>, grid, dim3(kWave), lds, st, table);
}

}  // namespace gsr
