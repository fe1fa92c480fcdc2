// Batching form, phased launches: stage 4 (step_2) over every tile of the batch.
#include "rigid_tiles.h"

namespace gsr {

const void* phase_kernel_4() { return (const void*)k_phase<4>; }

void launch_phase_4(dim3 grid, unsigned lds, hipStream_t st, const char* table) {
  hipLaunchKernelGGL(k_phase<4>, grid, dim3(kWave), lds, st, table);
}

}  // namespace gsr
