// Batching form, phased launches: stage 0 (step_1 dynamics) over every tile of the batch.
#include "rigid_tiles.h"

namespace gsr {

const void* phase_kernel_0() { return (const void*)k_phase<0>; }

void launch_phase_0(dim3 grid, unsigned lds, hipStream_t st, const char* table) {
  hipLaunchKernelGGL(k_phase<0>, grid, dim3(kWave), lds, st, table);
}

}  // namespace gsr
