// gfx950 kernels and C-ABI entry points of the batched rigid step (libgenesis_rigid.so).
//
// Plain drop-in form: one substep = four stages launched in stream order over env-minor SoA state (consecutive lanes
// or workgroups touch consecutive envs for every per-env field):
//   step_1     smooth dynamics (kernel_step_1): k_step1 (one lane per env) or k_step1_wave (one wavefront per env)
//   collision  equality rows, AABBs, sweep-and-prune, narrowphase, inequality rows (_func_constraint_force):
//              k_collide / k_collide_wave, then the GJK/EPA pass k_gjk_fixup (rigid_k_collide.hip)
//   solve      Newton solve + qacc/contact-force epilogue (ConstraintSolver.resolve): k_solve / k_solve_wave
//              (rigid_k_solve.hip)
//   step_2     integration, NaN-guarded commit, end-of-step kinematics (kernel_step_2): k_step2 / k_step2_wave
// This unit holds step_1, step_2, the kinematics refresh, resets and readback kernels and every C-ABI entry point;
// the batching form's kernels live in rigid_k_fused_*.hip, rigid_k_phase_*.hip and rigid_k_resume.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cstring>
#include <mutex>
#include <vector>

#include "rigid_tiles.h"

using namespace gsr;

namespace {

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


__global__ void __launch_bounds__(kWave) k_step1_wave(gsr_model m, gsr_state s) {
  if ((int)blockIdx.x >= s.n_envs) return;  // uniform across the block
  const int b = xcd_env(blockIdx.x, s.n_envs);
  Env e(m, s, b);
  wave_step1(e);
}

__global__ void __launch_bounds__(kWave) k_step2_wave(gsr_model m, gsr_state s) {
  if ((int)blockIdx.x >= s.n_envs) return;  // uniform across the block
  const int b = xcd_env(blockIdx.x, s.n_envs);
  Env e(m, s, b);
  wave_step2(e);
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

// collider_kernel_get_contacts (collider_decomp.py:687-724). Contact i of env b goes to output row
// `starts[b] + i` (ragged) or `b * n_max + i` (padded, starts == nullptr). Padded rows past the env's contact count
// get the values the reference's freshly allocated buffers hold there: -1 indices, zero floats. One thread per
// (slot i, env b), env fastest, so the state reads of a wavefront are consecutive envs of one slot.
__global__ void k_get_contacts(gsr_model m, gsr_state s, const int32_t* starts, int32_t n_max, int64_t out_rows,
                               int32_t* iout, float* fout) {
  const int B = s.n_envs;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)B * n_max) return;
  const int i = (int)(t / B), b = (int)(t % B);
  const Env e(m, s, b);
  const int nc = min(e.I(s.n_contacts, 0), m.max_contacts);
  if (starts && i >= nc) return;
  const int64_t row = starts ? (int64_t)starts[b] + i : (int64_t)b * n_max + i;
  if (row < 0 || row >= out_rows) return;
  int32_t* io = iout + 4 * row;
  float* fo = fout + 10 * row;
  if (i >= nc) {
    for (int k = 0; k < 4; k++) io[k] = -1;
    for (int k = 0; k < 10; k++) fo[k] = 0.f;
    return;
  }
  const int ga = e.I(s.contact_geoms, 2 * i), gb = e.I(s.contact_geoms, 2 * i + 1);
  io[0] = m.geom_link[ga];
  io[1] = m.geom_link[gb];
  io[2] = ga;
  io[3] = gb;
  fo[0] = e.F(s.contact_pen, i);
  const f3 p = e.L3(s.contact_pos, i), n = e.L3(s.contact_normal, i), f = e.L3(s.contact_force, i);
  fo[1] = p.x; fo[2] = p.y; fo[3] = p.z;
  fo[4] = n.x; fo[5] = n.y; fo[6] = n.z;
  fo[7] = f.x; fo[8] = f.y; fo[9] = f.z;
}

// links_state.contact_force of func_update_contact_force (constraint_solver_decomp.py:1350-1379): per env, zero
// every link, then for each contact in order subtract its force from link_a and add it to link_b, the reference's
// order of float additions. Output row k (env envs_idx[k], or env k when envs_idx is null) holds n_links x 3 floats.
__global__ void k_links_contact_force(gsr_model m, gsr_state s, const int32_t* envs_idx, int32_t n_idx, float* out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n_idx) return;
  const int b = envs_idx ? envs_idx[k] : k;
  if (b < 0 || b >= s.n_envs) return;
  const Env e(m, s, b);
  float* o = out + (size_t)k * m.n_links * 3;
  for (int j = 0; j < 3 * m.n_links; j++) o[j] = 0.f;
  const int nc = m.enable_collision ? min(e.I(s.n_contacts, 0), m.max_contacts) : 0;
  for (int i = 0; i < nc; i++) {
    const f3 f = e.L3(s.contact_force, i);
    float* a = o + 3 * m.geom_link[e.I(s.contact_geoms, 2 * i)];
    float* c = o + 3 * m.geom_link[e.I(s.contact_geoms, 2 * i + 1)];
    a[0] = a[0] - f.x; a[1] = a[1] - f.y; a[2] = a[2] - f.z;
    c[0] = c[0] + f.x; c[1] = c[1] + f.y; c[2] = c[2] + f.z;
  }
}

// kernel_get_equality_constraints (constraint_solver_decomp.py:2510-2558) for the JOINT equalities this solver
// builds (one row each, rows 0..n_eq-1 in model order). Every env has the model's n_equalities, so padded and ragged
// outputs coincide: row b * n_equalities + i holds {type, obj1, obj2} and the row force in column 0 of 6.
// kernel_get_equality_constraints (constraint_solver_decomp.py:2510-2558), padded: row b * (n_equalities +
// max_dyn_welds) + i; rows past an env's own count hold type -1 and zero forces
__global__ void k_get_equalities(gsr_model m, gsr_state s, int32_t* iout, float* fout) {
  const int B = s.n_envs, E = m.n_equalities, N = E + m.max_dyn_welds;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)B * N) return;
  const int i = (int)(t / B), b = (int)(t % B);
  const Env e(m, s, b);
  const int64_t row = (int64_t)b * N + i;
  const int n_dyn = m.max_dyn_welds > 0 ? e.I(s.eq_dyn_n, 0) : 0;
  if (i >= E + n_dyn) {
    for (int k = 0; k < 3; k++) iout[3 * row + k] = -1;
    for (int k = 0; k < 6; k++) fout[6 * row + k] = 0.f;
    return;
  }
  auto rows_of = [&](int j) { return j >= E ? 6 : m.eq_type[j] == EQ_CONNECT ? 3 : m.eq_type[j] == EQ_WELD ? 6 : 1; };
  iout[3 * row] = i < E ? m.eq_type[i] : (int)EQ_WELD;
  iout[3 * row + 1] = i < E ? m.eq_obj1[i] : e.I(s.eq_dyn_obj, 2 * (i - E));
  iout[3 * row + 2] = i < E ? m.eq_obj2[i] : e.I(s.eq_dyn_obj, 2 * (i - E) + 1);
  int c = 0;
  for (int j = 0; j < i; j++) c += rows_of(j);
  const int nr = rows_of(i);
  for (int k = 0; k < 6; k++) fout[6 * row + k] = k < nr ? e.EF(s.efc_force, c + k) : 0.f;
}

int check(const gsr_model* m, const gsr_state* s) {
  if (!m || !s) return GSR_E_ARG;
  if (s->n_envs <= 0) return GSR_E_ARG;
  if (m->row_width > kMaxRowWidth) return GSR_E_SHAPE;
  if (m->n_geoms >= (1 << 15)) return GSR_E_SHAPE;
  if (m->solver_variant == 1) {
    if (m->n_entities > kMaxPatternEntities) return GSR_E_SHAPE;
    if (m->wave_h_floats < 0 || wave_lds_bytes(*m) > 160 * 1024)
      return GSR_E_SHAPE;
  }
  if (m->collide_variant == 1) {
    if (m->n_geoms > kWaveMaxGeoms || m->n_contacts_per_pair > kMaxPairContacts) return GSR_E_SHAPE;
  }
  if (m->solver_variant < 0 || m->solver_variant > 1 || m->collide_variant < 0 || m->collide_variant > 1 ||
      m->dyn_variant < 0 || m->dyn_variant > 1)
    return GSR_E_ARG;
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
    case 0:
      if (m->dyn_variant == 1) hipLaunchKernelGGL(k_step1_wave, dim3(s->n_envs), dim3(kWave), 0, st, *m, *s);
      else hipLaunchKernelGGL(k_step1, g, dim3(kBlock), 0, st, *m, *s, 0);
      break;
    case 1:
      launch_collide(*m, *s, st);
      break;
    case 2:
      launch_solve(*m, *s, st);
      break;
    default:
      if (m->dyn_variant == 1) hipLaunchKernelGGL(k_step2_wave, dim3(s->n_envs), dim3(kWave), 0, st, *m, *s);
      else hipLaunchKernelGGL(k_step2, g, dim3(kBlock), 0, st, *m, *s);
      break;
  }
}

int run_substep(const gsr_model* m, const gsr_state* s, hipStream_t st) {
  for (int k = 0; k < 4; k++) launch_one(m, s, k, st);
  return launch_status();
}

}  // namespace

namespace gsr {

const void* fused_kernel(int W, int G) {
  if (W == 4) return G ? fused_kernel_w4_g1() : fused_kernel_w4_g0();
  return G ? fused_kernel_w2_g1() : fused_kernel_w2_g0();
}

void launch_fused(int W, int G, dim3 grid, unsigned lds, hipStream_t st, const char* table, uint32_t* queue,
                  int persistent) {
  if (W == 4) {
    if (G) launch_fused_w4_g1(grid, lds, st, table, queue, persistent);
    else launch_fused_w4_g0(grid, lds, st, table, queue, persistent);
  } else {
    if (G) launch_fused_w2_g1(grid, lds, st, table, queue, persistent);
    else launch_fused_w2_g0(grid, lds, st, table, queue, persistent);
  }
}

const void* phase_kernel(int P) {
  switch (P) {
    case 0: return phase_kernel_0();
    case 1: return phase_kernel_1();
    case 2: return phase_kernel_2();
    case 3: return phase_kernel_3();
    default: return phase_kernel_4();
  }
}

void launch_phase(int P, dim3 grid, unsigned lds, hipStream_t st, const char* table) {
  switch (P) {
    case 0: launch_phase_0(grid, lds, st, table); break;
    case 1: launch_phase_1(grid, lds, st, table); break;
    case 2: launch_phase_2(grid, lds, st, table); break;
    case 3: launch_phase_3(grid, lds, st, table); break;
    default: launch_phase_4(grid, lds, st, table); break;
  }
}

}  // namespace gsr

// ---- the allocation-free step (gsr_step_noalloc) -------------------------------------------------------------
// Per-substep scratch of gsr_step_noalloc: constraint rows, Newton solver vectors and Hessian, per-entity row lists,
// GJK/EPA workspace and deferral list, integration scratch. One copy per GPU, held until the process ends; the host
// name `scratch` is only a stand-in for it, so host code uses the per-device address from hipGetSymbolAddress.
static __device__ __align__(256) unsigned char scratch[GSR_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_dev[kMaxDevices] = {};

// device address of the current device's `scratch`, looked up once per device
int scratch_address(void** out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return GSR_E_LAUNCH;
  if (dev < 0 || dev >= kMaxDevices) return GSR_E_SHAPE;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return GSR_E_LAUNCH;
    g_scratch_dev[dev] = p;
  }
  *out = g_scratch_dev[dev];
  return GSR_OK;
}

// Carves the scratch arrays of n_envs envs out of `base` (each 256-byte aligned) into `s`; returns the bytes used,
// or only counts them when base is null. Sizes per env follow the state layout of each array.
size_t carve_scratch(const gsr_model& m, int n_envs, char* base, gsr_state* s) {
  const size_t B = (size_t)n_envs;
  const size_t R = (size_t)std::max(m.max_rows, 1), W = (size_t)m.row_width, D = (size_t)std::max(m.n_dofs, 1);
  const size_t Q = (size_t)std::max(m.n_qs, 1), E = (size_t)m.n_entities, MB = (size_t)std::max(m.max_broad, 1);
  size_t off = 0;
  auto take = [&](auto*& field, size_t per_env) {
    using T = std::remove_reference_t<decltype(*field)>;
    const size_t bytes = (std::max(per_env, (size_t)1) * B * sizeof(T) + 255) & ~(size_t)255;
    if (base) field = reinterpret_cast<T*>(base + off);
    off += bytes;
  };
  take(s->jac_val, R * W); take(s->jac_idx, R * W); take(s->jac_nnz, R);
  take(s->efc_diag, R); take(s->efc_aref, R); take(s->efc_D, R); take(s->efc_fl, R); take(s->efc_Jaref, R);
  take(s->efc_jv, R); take(s->efc_quad, 3 * R); take(s->efc_active, R);
  take(s->qacc, D); take(s->Ma, D); take(s->grad, D); take(s->Mgrad, D); take(s->search, D); take(s->mv, D);
  take(s->qfrc_constraint, D); take(s->nt_vec, D); take(s->nt_H, D * (D + 1) / 2); take(s->solver_iters, 1);
  take(s->ent_row_ptr, E + 1); take(s->ent_row_list, 2 * R); take(s->efc_ent, R); take(s->ent_row_run, 2 * R);
  take(s->gjk_ws, (size_t)kGjkWsWords); take(s->gjk_defer_n, 1); take(s->gjk_defer, 2 * MB);
  take(s->qpos_next, Q); take(s->vel_next, D);
  return off;
}

}  // namespace

extern "C" {

int gsr_abi_version(void) { return GSR_ABI_VERSION; }
int gsr_model_size(void) { return (int)sizeof(gsr_model); }
int gsr_state_size(void) { return (int)sizeof(gsr_state); }
int gsr_group_size(void) { return (int)sizeof(gsr_group); }
int gsr_batch_info_size(void) { return (int)sizeof(gsr_batch_info); }
int gsr_gjk_workspace_words(void) { return kGjkWsWords; }

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

int64_t gsr_noalloc_scratch_bytes(const gsr_model* model, int32_t n_envs) {
  if (!model || n_envs <= 0) return GSR_E_ARG;
  gsr_state tmp;
  std::memset(&tmp, 0, sizeof(tmp));
  return (int64_t)carve_scratch(*model, n_envs, nullptr, &tmp);
}

int gsr_step_noalloc(const gsr_model* model, const gsr_state* state, int32_t n_substeps, hipStream_t stream) {
  if (!model || !state) return GSR_E_ARG;
  if (state->n_envs <= 0 || n_substeps < 1) return GSR_E_ARG;
  gsr_state s = *state;
  if (carve_scratch(*model, s.n_envs, nullptr, &s) > (size_t)GSR_NOALLOC_SCRATCH_BYTES) return GSR_E_SHAPE;
  void* base = nullptr;
  int rc = scratch_address(&base);
  if (rc) return rc;
  carve_scratch(*model, s.n_envs, static_cast<char*>(base), &s);
  rc = check(model, &s);
  if (rc) return rc;
  for (int i = 0; i < n_substeps; i++) {
    rc = run_substep(model, &s, stream);
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

int gsr_get_contacts(const gsr_model* model, const gsr_state* state, const int32_t* dev_starts, int32_t n_max,
                     int64_t out_rows, int32_t* dev_iout, float* dev_fout, hipStream_t stream) {
  int rc = check(model, state);
  if (rc) return rc;
  if (n_max < 0 || out_rows < 0 || ((!dev_iout || !dev_fout) && out_rows > 0)) return GSR_E_ARG;
  if (!dev_starts && out_rows < (int64_t)state->n_envs * n_max) return GSR_E_ARG;
  const int64_t n = (int64_t)state->n_envs * n_max;
  if (n == 0 || out_rows == 0) return 0;
  hipLaunchKernelGGL(k_get_contacts, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, *model, *state,
                     dev_starts, n_max, out_rows, dev_iout, dev_fout);
  return launch_status();
}

int gsr_links_contact_force(const gsr_model* model, const gsr_state* state, const int32_t* dev_envs_idx,
                            int32_t n_idx, float* dev_out, hipStream_t stream) {
  int rc = check(model, state);
  if (rc) return rc;
  const int n = dev_envs_idx ? n_idx : state->n_envs;
  if (n < 0 || (n > 0 && !dev_out)) return GSR_E_ARG;
  if (n == 0 || model->n_links == 0) return 0;
  hipLaunchKernelGGL(k_links_contact_force, grid_for(n), dim3(kBlock), 0, stream, *model, *state, dev_envs_idx, n,
                     dev_out);
  return launch_status();
}

int gsr_get_equalities(const gsr_model* model, const gsr_state* state, int32_t* dev_iout, float* dev_fout,
                       hipStream_t stream) {
  int rc = check(model, state);
  if (rc) return rc;
  const int64_t n = (int64_t)state->n_envs * (model->n_equalities + model->max_dyn_welds);
  if (n == 0) return 0;
  if (!dev_iout || !dev_fout) return GSR_E_ARG;
  hipLaunchKernelGGL(k_get_equalities, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, *model, *state,
                     dev_iout, dev_fout);
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
  if (stage >= 2) launch_one(model, state, 0, stream);
  if (stage >= 3) launch_one(model, state, 1, stream);
  return launch_status();
}

int64_t gsr_batch_table_bytes(int32_t n_groups) {
  if (n_groups < 1 || n_groups > GSR_MAX_GROUPS) return GSR_E_ARG;
  return (int64_t)(kRangesOffset + sizeof(BatchRange) * (size_t)n_groups);
}

// dynamic LDS of phase 1 (wavefront collision) or 3 (wavefront solver): the largest over the batch's scene types
static int phase_lds(const gsr_batch_info& b, int phase) {
  size_t lds = 0;
  for (int k = 0; k < b.n_types; k++) {
    const gsr_model& m = b.models[k];
    if (phase == 1 && m.collide_variant == 1) lds = std::max(lds, (size_t)wave_collide_lds_bytes(m.n_geoms));
    if (phase == 3 && m.solver_variant == 1)
      lds = std::max(lds, (size_t)wave_lds_bytes(m));
  }
  return (int)lds;
}

static int finish_launch_info(gsr_batch_info& out, int n_groups, int tiles, int wave_tiles, size_t lds,
                              gsr_batch_info* info);

int gsr_batch_upload(const gsr_group* groups, int32_t n_groups, void* dev_table, int64_t table_bytes,
                     gsr_batch_info* info) {
  if (!groups || !dev_table || !info || n_groups < 1 || n_groups > GSR_MAX_GROUPS) return GSR_E_ARG;
  if (table_bytes < gsr_batch_table_bytes(n_groups)) return GSR_E_ARG;
  std::vector<BatchRange> ranges(n_groups);
  gsr_batch_info out;
  std::memset(&out, 0, sizeof(out));
  int tiles = 0, wtiles = 0;
  size_t lds = 0;
  for (int g = 0; g < n_groups; g++) {
    const gsr_group& G = groups[g];
    int rc = check(&G.model, &G.state);
    if (rc) return rc;
    if (G.n_envs < 1 || G.env_begin < 0 || G.env_begin + G.n_envs > G.state.n_envs) return GSR_E_ARG;
    if (G.envs_per_tile < 1 || G.envs_per_tile > kWave) return GSR_E_ARG;
    const int Tw = G.wave_envs_per_tile > 0 ? G.wave_envs_per_tile : G.envs_per_tile;
    if (G.wave_envs_per_tile < 0 || Tw > kWave) return GSR_E_ARG;
    int type = -1;
    for (int k = 0; k < out.n_types; k++)
      if (!std::memcmp(&out.models[k], &G.model, sizeof(gsr_model)) &&
          !std::memcmp(&out.states[k], &G.state, sizeof(gsr_state)))
        type = k;
    if (type < 0) {
      if (out.n_types == GSR_MAX_TYPES) return GSR_E_SHAPE;
      type = out.n_types++;
      out.models[type] = G.model;
      out.states[type] = G.state;
    }
    if (G.model.solver_variant == 1)
      lds = std::max(lds, wave_lds_bytes(G.model));
    if (G.model.collide_variant == 1) lds = std::max(lds, wave_collide_lds_bytes(G.model.n_geoms));
    const int n_t = (G.n_envs + G.envs_per_tile - 1) / G.envs_per_tile, n_w = (G.n_envs + Tw - 1) / Tw;
    ranges[g] = BatchRange{type, G.env_begin, G.n_envs, G.envs_per_tile, tiles, tiles + n_t, Tw, wtiles, wtiles + n_w,
                           0, 0, 0};
    tiles += n_t;
    wtiles += n_w;
  }
  // Groups over one state must agree on the execution variants (the constraint-row and Hessian layouts follow the
  // solver variant, so two variants on one state would write each other's envs) and must not share envs.
  {
    std::vector<int> order(n_groups);
    for (int g = 0; g < n_groups; g++) order[g] = g;
    std::sort(order.begin(), order.end(), [&](int a, int b) {
      const uintptr_t pa = (uintptr_t)groups[a].state.qpos, pb = (uintptr_t)groups[b].state.qpos;
      return pa != pb ? pa < pb : groups[a].env_begin < groups[b].env_begin;
    });
    for (int i = 1; i < n_groups; i++) {
      const gsr_group& A = groups[order[i - 1]];
      const gsr_group& B = groups[order[i]];
      if (A.state.qpos != B.state.qpos) continue;
      if (A.model.solver_variant != B.model.solver_variant || A.model.collide_variant != B.model.collide_variant ||
          A.model.dyn_variant != B.model.dyn_variant)
        return GSR_E_ARG;
      if (B.env_begin < A.env_begin + A.n_envs) return GSR_E_ARG;
    }
  }
  BatchHeader h{n_groups, tiles, 0u, 0u, wtiles, 0, 0, 0};
  FusedTypes ty;
  std::memset(&ty, 0, sizeof(ty));
  for (int k = 0; k < out.n_types; k++) {
    ty.m[k] = out.models[k];
    ty.s[k] = out.states[k];
  }
  // the queue words stay as they are on the device: zero-filled by the caller once, reset by every launch
  if (hipMemcpy(static_cast<char*>(dev_table) + kRangesOffset, ranges.data(), sizeof(BatchRange) * ranges.size(),
                hipMemcpyHostToDevice) != hipSuccess)
    return GSR_E_LAUNCH;
  if (hipMemcpy(static_cast<char*>(dev_table) + kTypesOffset, &ty, sizeof(ty), hipMemcpyHostToDevice) != hipSuccess)
    return GSR_E_LAUNCH;
  if (hipMemcpy(dev_table, &h, 2 * sizeof(int32_t), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(static_cast<char*>(dev_table) + offsetof(BatchHeader, n_wave_tiles), &h.n_wave_tiles, sizeof(int32_t),
                hipMemcpyHostToDevice) != hipSuccess)
    return GSR_E_LAUNCH;
  return finish_launch_info(out, n_groups, tiles, wtiles, lds, info);
}

// Launch form of a batch whose types are in `out`: kernel attributes (dynamic LDS), grid and the form's flags, read
// from `info` and written back to it.
static int finish_launch_info(gsr_batch_info& out, int n_groups, int tiles, int wave_tiles, size_t lds,
                              gsr_batch_info* info) {
  int dev = 0, n_cu = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess) return GSR_E_LAUNCH;
  if (hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return GSR_E_LAUNCH;
  if (info->phased) {
    const int lc = phase_lds(out, 1), ls = phase_lds(out, 3);
    if (lc > 0 && hipFuncSetAttribute(phase_kernel(1), hipFuncAttributeMaxDynamicSharedMemorySize, lc) !=
                      hipSuccess)
      return GSR_E_LAUNCH;
    if (ls > 0 && hipFuncSetAttribute(phase_kernel(3), hipFuncAttributeMaxDynamicSharedMemorySize, ls) !=
                      hipSuccess)
      return GSR_E_LAUNCH;
    out.n_groups = n_groups;
    out.n_tiles = tiles;
    out.n_wave_tiles = wave_tiles;
    out.persistent = 0;
    out.waves_per_simd = 0;
    out.gjk_inline = 0;
    out.phased = 1;
    out.grid = tiles;
    out.lds_bytes = (int32_t)lds;
    *info = out;
    return GSR_OK;
  }
  const int W = info->waves_per_simd == 4 ? 4 : 2;
  const int G = info->gjk_inline ? 1 : 0;
  const void* fn = fused_kernel(W, G);
  if (lds > 0 && hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
    return GSR_E_LAUNCH;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kWave, lds) != hipSuccess) return GSR_E_LAUNCH;
  out.n_groups = n_groups;
  out.n_tiles = tiles;
  out.n_wave_tiles = wave_tiles;
  out.persistent = info->persistent ? 1 : 0;
  out.waves_per_simd = W;
  out.gjk_inline = G;
  out.phased = 0;
  out.grid = out.persistent ? std::max(1, std::min(tiles, n_cu * std::max(per_cu, 1))) : tiles;
  out.lds_bytes = (int32_t)lds;
  *info = out;
  return GSR_OK;
}

static bool fused_info_ok(const void* dev_table, const gsr_batch_info* info) {
  return dev_table && info && info->grid >= 1 && info->n_tiles >= 1 && info->n_types >= 1 &&
         info->n_types <= GSR_MAX_TYPES && (!info->phased || info->n_wave_tiles >= 1);
}

int gsr_fused_kernel_count(const gsr_batch_info* info) {
  if (!info) return GSR_E_ARG;
  if (info->phased) return 5;
  return info->gjk_inline ? 1 : 1 + info->n_types;
}

int gsr_launch_fused_kernel(const void* dev_table, const gsr_batch_info* info, int32_t k, hipStream_t stream) {
  if (!fused_info_ok(dev_table, info)) return GSR_E_ARG;
  const int n = gsr_fused_kernel_count(info);
  if (k < 0 || k >= n) return GSR_E_ARG;
  const char* t = static_cast<const char*>(dev_table);
  if (info->phased) {
    const int lds = k == 1 ? phase_lds(*info, 1) : k == 3 ? phase_lds(*info, 3) : 0;
    launch_phase(k, dim3(k == 2 || k == 3 ? info->n_wave_tiles : info->n_tiles), lds, stream, t);
  } else if (k == 0) {
    uint32_t* queue = reinterpret_cast<uint32_t*>(const_cast<char*>(t) + 2 * sizeof(int32_t));
    launch_fused(info->waves_per_simd == 4 ? 4 : 2, info->gjk_inline ? 1 : 0, dim3(info->grid), info->lds_bytes,
                 stream, t, queue, (int)info->persistent);
  } else {
    const gsr_model& m = info->models[k - 1];
    const gsr_state& st = info->states[k - 1];
    const size_t lds = m.solver_variant == 1 ? wave_lds_bytes(m) : 0;
    launch_gjk_resume(m, st, (unsigned)lds, stream);
  }
  return launch_status();
}

int gsr_step_fused(const void* dev_table, const gsr_batch_info* info, int32_t n_substeps, hipStream_t stream) {
  if (!fused_info_ok(dev_table, info) || n_substeps < 1) return GSR_E_ARG;
  const int n = gsr_fused_kernel_count(info);
  for (int i = 0; i < n_substeps; i++)
    for (int k = 0; k < n; k++) {
      const int rc = gsr_launch_fused_kernel(dev_table, info, k, stream);
      if (rc) return rc;
    }
  return GSR_OK;
}

int gsr_graph_capture(const void* dev_table, const gsr_batch_info* info, int32_t n_substeps, hipStream_t stream,
                      void** graph_exec) {
  if (!graph_exec || !stream) return GSR_E_ARG;
  *graph_exec = nullptr;
  if (hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return GSR_E_LAUNCH;
  int rc = gsr_step_fused(dev_table, info, n_substeps, stream);
  hipGraph_t graph = nullptr;
  hipError_t err = hipStreamEndCapture(stream, &graph);
  if (rc) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (err != hipSuccess || !graph) return GSR_E_LAUNCH;
  hipGraphExec_t exec = nullptr;
  err = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (err != hipSuccess) return GSR_E_LAUNCH;
  *graph_exec = exec;
  return GSR_OK;
}

int gsr_graph_launch(void* graph_exec, hipStream_t stream) {
  if (!graph_exec) return GSR_E_ARG;
  return hipGraphLaunch(static_cast<hipGraphExec_t>(graph_exec), stream) == hipSuccess ? GSR_OK : GSR_E_LAUNCH;
}

int gsr_graph_destroy(void* graph_exec) {
  if (!graph_exec) return GSR_E_ARG;
  return hipGraphExecDestroy(static_cast<hipGraphExec_t>(graph_exec)) == hipSuccess ? GSR_OK : GSR_E_LAUNCH;
}

// ---- serving form -------------------------------------------------------------------------------------------------
int gsr_serve_host_alloc(gsr_serve_host** out) {
  if (!out) return GSR_E_ARG;
  *out = nullptr;
  void* p = nullptr;
  if (hipHostMalloc(&p, sizeof(gsr_serve_host), hipHostMallocCoherent | hipHostMallocMapped) != hipSuccess || !p)
    return GSR_E_LAUNCH;
  std::memset(p, 0, sizeof(gsr_serve_host));
  *out = static_cast<gsr_serve_host*>(p);
  return GSR_OK;
}

int gsr_serve_host_free(gsr_serve_host* host) {
  if (!host) return GSR_E_ARG;
  return hipHostFree(host) == hipSuccess ? GSR_OK : GSR_E_LAUNCH;
}

int64_t gsr_serve_state_bytes(int32_t n_tiles) {
  return n_tiles < 1 ? GSR_E_ARG : (int64_t)serve_state_bytes(n_tiles);
}

int64_t gsr_serve_table_bytes(int32_t n_tiles) {
  return n_tiles < 1 ? GSR_E_ARG : (int64_t)serve_table_bytes(n_tiles);
}

int gsr_serve_clock_khz(void) {
  int dev = 0, khz = 0;
  if (hipGetDevice(&dev) != hipSuccess) return GSR_E_LAUNCH;
  if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess) return GSR_E_LAUNCH;
  return khz;
}

int gsr_serve_init(const gsr_serve_pool* pools, int32_t n_pools, double mem_budget, double tick_budget,
                   void* dev_state, int64_t state_bytes, void* dev_table, int64_t table_bytes, gsr_batch_info* info,
                   hipStream_t stream) {
  if (!pools || !dev_state || !dev_table || !info || n_pools < 1 || n_pools > GSR_MAX_TYPES) return GSR_E_ARG;
  gsr_batch_info out;
  std::memset(&out, 0, sizeof(out));
  int tiles = 0;
  size_t lds = 0;
  for (int p = 0; p < n_pools; p++) {
    const gsr_serve_pool& P = pools[p];
    const int rc = check(&P.model, &P.state);
    if (rc) return rc;
    if (P.envs_per_tile < 1 || P.envs_per_tile > kWave || P.n_tiles < 1 ||
        (int64_t)P.envs_per_tile * P.n_tiles > P.state.n_envs || !P.init_qpos)
      return GSR_E_ARG;
    if (P.ctrl_kind < 0 || P.ctrl_kind > 2) return GSR_E_ARG;
    if (P.ctrl_kind != 0 && (P.ctrl_dof_begin < 0 || P.ctrl_dof_end > P.model.n_dofs ||
                             P.ctrl_dof_begin >= P.ctrl_dof_end || (P.ctrl_kind == 1 && P.ctrl_period < 1) ||
                             (P.ctrl_kind == 1 && (!P.ctrl_lo || !P.ctrl_hi)) || (P.ctrl_kind == 2 && !P.ctrl_base)))
      return GSR_E_ARG;
    for (int q = 0; q < p; q++)
      if (pools[q].state.qpos == P.state.qpos) return GSR_E_ARG;  // one state per pool
    out.models[p] = P.model;
    out.states[p] = P.state;
    if (P.model.solver_variant == 1)
      lds = std::max(lds, wave_lds_bytes(P.model));
    if (P.model.collide_variant == 1) lds = std::max(lds, wave_collide_lds_bytes(P.model.n_geoms));
    tiles += P.n_tiles;
  }
  out.n_types = n_pools;
  if (state_bytes < (int64_t)serve_state_bytes(tiles) || table_bytes < (int64_t)serve_table_bytes(tiles))
    return GSR_E_ARG;
  const int rc = serve_init(pools, n_pools, mem_budget, tick_budget, dev_state, state_bytes, dev_table, table_bytes,
                            stream);
  if (rc) return rc;
  // every tile may be active: the launches cover all of them (idle ones return at once)
  return finish_launch_info(out, tiles, tiles, tiles, lds, info);
}

int gsr_serve_tick(void* dev_state, void* dev_table, const gsr_batch_info* info, gsr_serve_host* host,
                   hipStream_t stream) {
  if (!dev_state || !host || !fused_info_ok(dev_table, info)) return GSR_E_ARG;
  gsr_serve_host* hd = nullptr;
  if (hipHostGetDevicePointer((void**)&hd, host, 0) != hipSuccess || !hd) return GSR_E_ARG;
  const int n_tiles = info->n_tiles;
  launch_serve(0, dev_state, dev_table, hd, n_tiles, stream);
  launch_serve(1, dev_state, dev_table, hd, n_tiles, stream);
  launch_serve(2, dev_state, dev_table, hd, n_tiles, stream);
  const int n = gsr_fused_kernel_count(info);
  for (int k = 0; k < n; k++) {
    const int rc = gsr_launch_fused_kernel(dev_table, info, k, stream);
    if (rc) return rc;
  }
  launch_serve(3, dev_state, dev_table, hd, n_tiles, stream);
  return launch_status();
}

int gsr_serve_graph_capture(void* dev_state, void* dev_table, const gsr_batch_info* info, gsr_serve_host* host,
                            hipStream_t stream, void** graph_exec) {
  if (!graph_exec || !stream) return GSR_E_ARG;
  *graph_exec = nullptr;
  if (hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return GSR_E_LAUNCH;
  int rc = gsr_serve_tick(dev_state, dev_table, info, host, stream);
  hipGraph_t graph = nullptr;
  hipError_t err = hipStreamEndCapture(stream, &graph);
  if (rc) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (err != hipSuccess || !graph) return GSR_E_LAUNCH;
  hipGraphExec_t exec = nullptr;
  err = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (err != hipSuccess) return GSR_E_LAUNCH;
  *graph_exec = exec;
  return GSR_OK;
}

}  // extern "C"
