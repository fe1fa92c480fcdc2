// Collision stage of the plain step: k_collide (one lane per env) or k_collide_wave (one wavefront per env), then
// the GJK/EPA pass k_gjk_fixup over the pairs the narrowphase deferred.
#include "rigid_tiles.h"

namespace gsr {
namespace {

__global__ void __launch_bounds__(kBlock) k_collide(gsr_model m, gsr_state s) {
  int b = env_id();
  if (b >= s.n_envs) return;
  Env e(m, s, b);
  dev_equality_rows(e);
  dev_collision(e);
  if (!gjk_pending(e)) dev_inequality_rows(e);
}

// the GJK/EPA pass of the plain step, one wavefront per env: deferred pairs, contact insertion, then the env's
// inequality rows; envs without deferred pairs return at once
__global__ void __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(2, 8)))
k_gjk_fixup(gsr_model m, gsr_state s) {
  if ((int)blockIdx.x >= s.n_envs) return;
  const int b = xcd_env(blockIdx.x, s.n_envs);
  Env e(m, s, b);
  if (wave_gjk_fixup(e)) wave_rows(e);
}

// Register budget of two waves per SIMD: left to itself the narrowphase takes 256 VGPRs plus AGPRs (one wave per
// SIMD); held to 256 it spills more to scratch but doubles the waves that hide each other's latency (25 % faster on
// the 20-box pile)
__global__ void __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(2, 8)))
k_collide_wave(gsr_model m, gsr_state s) {
  extern __shared__ char lds[];
  if ((int)blockIdx.x >= s.n_envs) return;  // uniform across the block
  const int b = xcd_env(blockIdx.x, s.n_envs);
  Env e(m, s, b);
  wave_collide(e, lds);
}

}  // namespace

void launch_collide(const gsr_model& m, const gsr_state& s, hipStream_t st) {
  if (m.collide_variant == 1)
    hipLaunchKernelGGL(k_collide_wave, dim3(s.n_envs), dim3(kWave), wave_collide_lds_bytes(m.n_geoms), st, m, s);
  else
    hipLaunchKernelGGL(k_collide, dim3((s.n_envs + kBlock - 1) / kBlock), dim3(kBlock), 0, st, m, s);
  hipLaunchKernelGGL(k_gjk_fixup, dim3(s.n_envs), dim3(kWave), 0, st, m, s);
}

}  // namespace gsr
