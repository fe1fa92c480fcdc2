// Batching form, phased launches: stage 3 (Newton solve) over every tile of the batch.
#include "rigid_tiles.h"

namespace gsr {

const void* phase_kernel_3() { return (const void*)k_phase<3>; }

void launch_phase_3(dim3 grid, unsigned lds, hipStream_t st, const char* table) {
  hipLaunchKernelGGL(k_phase<3>, grid, dim3(kWave), lds, st, table);
}

}  // namespace gsr
