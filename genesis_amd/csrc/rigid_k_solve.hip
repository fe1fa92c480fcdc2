// Newton solve of the plain step: k_solve (one lane per env) or k_solve_wave (one wavefront per env, LDS vectors
// and Hessian).
#include "rigid_tiles.h"

namespace gsr {
namespace {

__global__ void __launch_bounds__(kBlock) k_solve(gsr_model m, gsr_state s) {
  int b = env_id();
  if (b >= s.n_envs) return;
  Env e(m, s, b);
  dev_solve(e);
}

__global__ void __launch_bounds__(kWave) k_solve_wave(gsr_model m, gsr_state s) {
  extern __shared__ char lds[];
  if ((int)blockIdx.x >= s.n_envs) return;  // uniform across the block
  const int b = xcd_env(blockIdx.x, s.n_envs);
  Env e(m, s, b);
  wave_solve(e, lds);
}

}  // namespace

void launch_solve(const gsr_model& m, const gsr_state& s, hipStream_t st) {
  if (m.solver_variant == 1)
    hipLaunchKernelGGL(k_solve_wave, dim3(s.n_envs), dim3(kWave), wave_lds_bytes(m), st, m, s);
  else
    hipLaunchKernelGGL(k_solve, dim3((s.n_envs + kBlock - 1) / kBlock), dim3(kBlock), 0, st, m, s);
}

}  // namespace gsr
