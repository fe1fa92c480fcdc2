// Batching form, phased launches: stage 1 (equality rows + collision + inequality rows) over every tile of the batch.
#include "rigid_tiles.h"

namespace gsr {

const void* phase_kernel_1() { return (const void*)k_phase<1>; }

void launch_phase_1(dim3 grid, unsigned lds, hipStream_t st, const char* table) {
  hipLaunchKernelGGL(k_phase<1>, grid, dim3(kWave), lds, st, table);
}

}  // namespace gsr
