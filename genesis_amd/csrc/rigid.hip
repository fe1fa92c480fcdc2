// gfx950 kernels and C-ABI entry points of the batched rigid step (libgenesis_rigid.so).
//
// Plain drop-in form: one substep = five stream-ordered launches, one lane per environment over env-minor SoA state
// (consecutive lanes touch consecutive addresses for every per-env field):
//   k_step1      kinematics refresh (only when the pose cache is stale) + smooth dynamics (kernel_step_1)
//   k_collide    equality rows, AABBs, sweep-and-prune, narrowphase, inequality rows (_func_constraint_force)
//   k_solve      Newton solve + qacc/contact-force epilogue (ConstraintSolver.resolve)
//   k_step2      integration, NaN-guarded commit, end-of-step kinematics (kernel_step_2)
#include <hip/hip_runtime.h>

#include "genesis_rigid.h"
#include "rigid_kin.h"
#include "rigid_collide.h"
#include "rigid_solver.h"
#include "rigid_solver_wave.h"

using namespace gsr;

namespace {

constexpr int kBlock = 64;

__device__ __forceinline__ int env_id() { return blockIdx.x * blockDim.x + threadIdx.x; }

__global__ void __launch_bounds__(kBlock) k_kinematics(gsr_model m, gsr_state s, const int32_t* envs_idx, int n_idx) {
  int t = env_id();
  if (t >= n_idx) return;
  int b = envs_idx ? envs_idx[t] : t;
  if (b < 0 || b >= s.n_envs) return;
  Env e(m, s, b);
  dev_kinematics(e);
  if (m.enable_collision) dev_update_aabbs(e);
}

__global__ void __launch_bounds__(kBlock) k_step1(gsr_model m, gsr_state s, int refresh_kinematics) {
  int b = env_id();
  if (b >= s.n_envs) return;
  Env e(m, s, b);
  if (refresh_kinematics) dev_kinematics(e);
  dev_smooth_dynamics(e);
}

__global__ void __launch_bounds__(kBlock) k_collide(gsr_model m, gsr_state s) {
  int b = env_id();
  if (b >= s.n_envs) return;
  Env e(m, s, b);
  dev_equality_rows(e);
  dev_collision(e);
  dev_inequality_rows(e);
}

__global__ void __launch_bounds__(kBlock) k_solve(gsr_model m, gsr_state s) {
  int b = env_id();
  if (b >= s.n_envs) return;
  Env e(m, s, b);
  dev_solve(e);
}

__global__ void __launch_bounds__(kWave) k_solve_wave(gsr_model m, gsr_state s) {
  extern __shared__ char lds[];
  int b = blockIdx.x;
  if (b >= s.n_envs) return;  // uniform across the block
  Env e(m, s, b);
  wave_solve(e, lds);
}

__global__ void __launch_bounds__(kBlock) k_step2(gsr_model m, gsr_state s) {
  int b = env_id();
  if (b >= s.n_envs) return;
  Env e(m, s, b);
  dev_integrate(e);
  dev_kinematics(e);
}

__global__ void __launch_bounds__(kBlock) k_reset(gsr_model m, gsr_state s, const int32_t* envs_idx, int n_idx,
                                                  uint32_t what) {
  int t = env_id();
  if (t >= n_idx) return;
  int b = envs_idx ? envs_idx[t] : t;
  if (b < 0 || b >= s.n_envs) return;
  Env e(m, s, b);
  if (what & GSR_RESET_COLLIDER) {
    for (int p = 0; p < m.n_pairs; p++) e.S3(s.cache_normal, p, zero3());
    e.I(s.first_time, 0) = 1;
  }
  if (what & GSR_RESET_WARMSTART) {
    e.I(s.is_warmstart, 0) = 0;
    for (int d = 0; d < m.n_dofs; d++) e.F(s.qacc_ws, d) = 0.f;
  }
  e.I((int32_t*)s.errno_env, 0) = 0;
}

__global__ void k_errno_reduce(gsr_state s, uint32_t* out) {
  // one block of 256 threads ORs all env words
  __shared__ uint32_t part[256];
  uint32_t v = 0;
  for (int b = threadIdx.x; b < s.n_envs; b += blockDim.x) v |= s.errno_env[b];
  part[threadIdx.x] = v;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) part[threadIdx.x] |= part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = part[0];
}

int check(const gsr_model* m, const gsr_state* s) {
  if (!m || !s) return GSR_E_ARG;
  if (s->n_envs <= 0) return GSR_E_ARG;
  if (m->row_width > kMaxRowWidth) return GSR_E_SHAPE;
  if (m->n_geoms >= (1 << 15)) return GSR_E_SHAPE;
  if (m->solver_variant == 1) {
    if (m->n_entities > kMaxPatternEntities) return GSR_E_SHAPE;
    if (wave_lds_bytes(m->n_dofs, m->n_entities) > 160 * 1024) return GSR_E_SHAPE;
  }
  return GSR_OK;
}

int launch_status() {
  hipError_t err = hipGetLastError();
  return err == hipSuccess ? GSR_OK : GSR_E_LAUNCH;
}

inline dim3 grid_for(int n) { return dim3((n + kBlock - 1) / kBlock); }

void launch_one(const gsr_model* m, const gsr_state* s, int k, hipStream_t st) {
  dim3 g = grid_for(s->n_envs);
  switch (k) {
    // kinematics of the current pose are refreshed at the end of every step (and by gsr_forward_kinematics after a
    // pose write), so kernel_step_1 never needs to recompute them (enable_mujoco_compatibility = False)
    case 0: hipLaunchKernelGGL(k_step1, g, dim3(kBlock), 0, st, *m, *s, 0); break;
    case 1: hipLaunchKernelGGL(k_collide, g, dim3(kBlock), 0, st, *m, *s); break;
    case 2:
      if (m->solver_variant == 1)
        hipLaunchKernelGGL(k_solve_wave, dim3(s->n_envs), dim3(kWave), wave_lds_bytes(m->n_dofs, m->n_entities), st,
                           *m, *s);
      else
        hipLaunchKernelGGL(k_solve, g, dim3(kBlock), 0, st, *m, *s);
      break;
    default: hipLaunchKernelGGL(k_step2, g, dim3(kBlock), 0, st, *m, *s); break;
  }
}

int run_substep(const gsr_model* m, const gsr_state* s, hipStream_t st) {
  for (int k = 0; k < 4; k++) launch_one(m, s, k, st);
  return launch_status();
}

}  // namespace

extern "C" {

int gsr_abi_version(void) { return GSR_ABI_VERSION; }
int gsr_model_size(void) { return (int)sizeof(gsr_model); }
int gsr_state_size(void) { return (int)sizeof(gsr_state); }

int gsr_step(const gsr_model* model, const gsr_state* state, int32_t n_substeps, hipStream_t stream) {
  int rc = check(model, state);
  if (rc) return rc;
  if (n_substeps < 1) return GSR_E_ARG;
  for (int i = 0; i < n_substeps; i++) {
    rc = run_substep(model, state, stream);
    if (rc) return rc;
  }
  return GSR_OK;
}

int gsr_forward_kinematics(const gsr_model* model, const gsr_state* state, const int32_t* envs_idx, int32_t n_idx,
                           hipStream_t stream) {
  int rc = check(model, state);
  if (rc) return rc;
  int n = envs_idx ? n_idx : state->n_envs;
  if (n <= 0) return GSR_OK;
  hipLaunchKernelGGL(k_kinematics, grid_for(n), dim3(kBlock), 0, stream, *model, *state, envs_idx, n);
  return launch_status();
}

int gsr_reset_envs(const gsr_model* model, const gsr_state* state, const int32_t* envs_idx, int32_t n_idx,
                   uint32_t what, hipStream_t stream) {
  int rc = check(model, state);
  if (rc) return rc;
  int n = envs_idx ? n_idx : state->n_envs;
  if (n <= 0) return GSR_OK;
  hipLaunchKernelGGL(k_reset, grid_for(n), dim3(kBlock), 0, stream, *model, *state, envs_idx, n, what);
  return launch_status();
}

int gsr_get_errno(const gsr_state* state, uint32_t* dev_out, hipStream_t stream) {
  if (!state || !dev_out || state->n_envs <= 0) return GSR_E_ARG;
  hipLaunchKernelGGL(k_errno_reduce, dim3(1), dim3(256), 0, stream, *state, dev_out);
  return launch_status();
}

int gsr_launch_kernel(const gsr_model* model, const gsr_state* state, int32_t kernel, hipStream_t stream) {
  int rc = check(model, state);
  if (rc) return rc;
  if (kernel < 0 || kernel > 3) return GSR_E_ARG;
  launch_one(model, state, kernel, stream);
  return launch_status();
}

int gsr_run_stage(const gsr_model* model, const gsr_state* state, int32_t stage, hipStream_t stream) {
  int rc = check(model, state);
  if (rc) return rc;
  dim3 g = grid_for(state->n_envs);
  if (stage == 0) return run_substep(model, state, stream);
  hipLaunchKernelGGL(k_kinematics, g, dim3(kBlock), 0, stream, *model, *state, (const int32_t*)nullptr, state->n_envs);
  if (stage >= 2) hipLaunchKernelGGL(k_step1, g, dim3(kBlock), 0, stream, *model, *state, 0);
  if (stage >= 3) hipLaunchKernelGGL(k_collide, g, dim3(kBlock), 0, stream, *model, *state);
  return launch_status();
}

}  // extern "C"
