// Batching form, GJK/EPA deferred: the rest of the substep for the envs of one scene type that the fused launch
// left at their collision: the GJK/EPA pass and inequality rows, then the model's solver and integration variants.
// One wavefront per env; envs without deferred pairs return at once.
#include "rigid_tiles.h"

namespace gsr {
namespace {

__global__ void __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(2, 8)))
k_gjk_resume(gsr_model m, gsr_state s) {
  extern __shared__ char lds[];
  if ((int)blockIdx.x >= s.n_envs) return;
  const int b = xcd_env(blockIdx.x, s.n_envs);
  Env e(m, s, b);
  if (!gjk_pending(e)) return;  // uniform across the block
  wave_gjk_fixup(e);
  wave_rows(e);
  if (m.solver_variant == 1) wave_solve(e, lds);
  else if (threadIdx.x == 0) dev_solve(e);
  __syncthreads();
  if (m.dyn_variant == 1) {
    wave_step2(e);
  } else if (threadIdx.x == 0) {
    dev_integrate(e);
    dev_kinematics(e);
  }
}

}  // namespace

void launch_gjk_resume(const gsr_model& m, const gsr_state& s, unsigned lds, hipStream_t st) {
  hipLaunchKernelGGL(k_gjk_resume, dim3(s.n_envs), dim3(kWave), lds, st, m, s);
}

}  // namespace gsr
