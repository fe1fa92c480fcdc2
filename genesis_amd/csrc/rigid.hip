// gfx950 kernels and C-ABI entry points of the batched rigid step (libgenesis_rigid.so).
//
// Plain drop-in form: one substep = five stream-ordered launches, one lane per environment over env-minor SoA state
// (consecutive lanes touch consecutive addresses for every per-env field):
//   k_step1      kinematics refresh (only when the pose cache is stale) + smooth dynamics (kernel_step_1)
//   k_collide    equality rows, AABBs, sweep-and-prune, narrowphase, inequality rows (_func_constraint_force)
//   k_solve      Newton solve + qacc/contact-force epilogue (ConstraintSolver.resolve)
//   k_step2      integration, NaN-guarded commit, end-of-step kinematics (kernel_step_2)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "genesis_rigid.h"
#include "rigid_kin.h"
#include "rigid_collide.h"
#include "rigid_solver.h"
#include "rigid_solver_wave.h"
#include "rigid_collide_wave.h"

using namespace gsr;

namespace {

constexpr int kBlock = 64;

__device__ __forceinline__ int env_id() { return blockIdx.x * blockDim.x + threadIdx.x; }

// Workgroup -> env for the one-env-per-workgroup kernels. The dispatcher deals workgroups to the 8 XCDs round-robin
// (workgroup i to XCD i % 8) and every XCD has its own L2, while the env-minor state arrays put 32 neighbouring envs
// in one 128-byte line. Within each chunk of 256 workgroups, XCD x therefore gets a run of 32 consecutive envs, so
// the waves it holds at once share their lines in its L2 instead of each XCD fetching every line. A bijection on
// [0, n): a partial last chunk is split the same way over its m workgroups.
__device__ __forceinline__ int xcd_env(int blk, int n) {
  constexpr int kRun = 32, kChunk = 8 * kRun;
  const int base = blk - blk % kChunk;
  const int m = min(kChunk, n - base);
  const int j = blk - base;
  const int q = m >> 3, r = m & 7, x = j & 7, k = j >> 3;
  return base + x * q + min(x, r) + k;
}

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

// step_1 / step_2 of one env on one wavefront: the env's entities are spread over the lanes (no entity reads another's
// state); only the geom poses, split by geom index, need the link poses of every lane, hence the barrier.
__device__ __forceinline__ void wave_step1(const Env& e) { dev_smooth_dynamics(e, threadIdx.x, kWave); }

__device__ __forceinline__ void wave_step2(const Env& e) {
  const int lane = threadIdx.x;
  bool valid = dev_integrate_next(e, lane, kWave);
  valid = __all(valid);
  dev_integrate_commit(e, lane, kWave, valid, lane == 0);
  dev_forward_kinematics(e, lane, kWave);
  dev_com_links(e, lane, kWave);
  __syncthreads();
  dev_update_geoms(e, lane, kWave);
  dev_forward_velocity(e, lane, kWave);
  __syncthreads();
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

// an env whose narrowphase deferred pairs to the GJK/EPA pass: its contact list and rows are completed there
__device__ __forceinline__ bool gjk_pending(const Env& e) { return e.I(e.s.gjk_defer_n, 0) > 0; }

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
__global__ void __launch_bounds__(kWave) k_gjk_fixup(gsr_model m, gsr_state s) {
  if ((int)blockIdx.x >= s.n_envs) return;
  const int b = xcd_env(blockIdx.x, s.n_envs);
  Env e(m, s, b);
  if (wave_gjk_fixup(e)) wave_rows(e);
}

__global__ void __launch_bounds__(kWave) k_collide_wave(gsr_model m, gsr_state s) {
  extern __shared__ char lds[];
  if ((int)blockIdx.x >= s.n_envs) return;  // uniform across the block
  const int b = xcd_env(blockIdx.x, s.n_envs);
  Env e(m, s, b);
  wave_collide(e, lds);
}

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


// ---- batching form ----------------------------------------------------------------------------------------------
// Device table of a heterogeneous batch: a header (work queue) then one record per group (env range). The model and
// state descriptors of the batch's scene types are kernel arguments, so every field read is a scalar load from the
// kernel-argument segment, as in the plain kernels.
struct BatchHeader {
  int32_t n_groups, n_tiles;
  uint32_t counter, done;  // work queue: next tile, finished workgroups (the last one resets both)
};
struct BatchRange {
  int32_t type, env_begin, n_envs, envs_per_tile, tile_begin, tile_end, pad0, pad1;
};
struct FusedTypes {
  gsr_model m[GSR_MAX_TYPES];
  gsr_state s[GSR_MAX_TYPES];
};
constexpr size_t kTypesOffset = 256;
constexpr size_t kRangesOffset = (kTypesOffset + sizeof(FusedTypes) + 255) & ~(size_t)255;
// The descriptors are read through the constant address space: every field read with a uniform type index is a
// scalar, invariant load, like a kernel argument, without copying the descriptors anywhere.
typedef const FusedTypes __attribute__((address_space(4))) ConstTypes;

// One substep of envs [first, first + n_here) of a group; the workgroup is one wavefront. Per-env stages run one
// lane per env; the wavefront collision and solver variants take the tile's envs one after the other.
// kGjk: GJK/EPA runs inside the collision stage (no env stops early, no resume launch) instead of being deferred
template <bool kGjk>
__device__ __forceinline__ void run_tile(const gsr_model& m, const gsr_state& s, int env0, int n_here, char* lds) {
  const int lane = threadIdx.x;
  const bool mine = lane < n_here;
  const bool cw = m.collide_variant == 1, sw = m.solver_variant == 1, dv = m.dyn_variant == 1;
  // an env with pairs deferred to GJK/EPA stops after its collision; k_gjk_resume finishes its substep
  if (!cw && !sw && !dv) {
    if (mine) {
      Env e(m, s, env0 + lane);
      dev_smooth_dynamics(e);
      dev_equality_rows(e);
      dev_collision<kGjk>(e);
      if (!gjk_pending(e)) {
        dev_inequality_rows(e);
        dev_solve(e);
        dev_integrate(e);
        dev_kinematics(e);
      }
    }
    return;
  }
  // collision reads only the pose state of the previous step, so it may run before or after the dynamics
  if (mine && (!dv || !cw)) {
    Env e(m, s, env0 + lane);
    if (!dv) dev_smooth_dynamics(e);
    if (!cw) {
      dev_equality_rows(e);
      dev_collision<kGjk>(e);
      if (!gjk_pending(e)) dev_inequality_rows(e);
    }
  }
  __syncthreads();
  for (int j = 0; j < n_here; j++) {
    Env e(m, s, env0 + j);
    if (dv) wave_step1(e);
    if (cw) wave_collide<kGjk>(e, lds);
    __syncthreads();
    if (sw && !gjk_pending(e)) wave_solve(e, lds);
    __syncthreads();
  }
  if (mine && (!sw || !dv)) {
    Env e(m, s, env0 + lane);
    if (!gjk_pending(e)) {
      if (!sw) dev_solve(e);
      if (!dv) {
        dev_integrate(e);
        dev_kinematics(e);
      }
    }
  }
  if (dv) {
    __syncthreads();
    for (int j = 0; j < n_here; j++) {
      Env e(m, s, env0 + j);
      if (!gjk_pending(e)) wave_step2(e);
    }
  }
}

// The rest of the substep for the envs of one scene type that the fused launch left at their collision: the GJK/EPA
// pass and inequality rows on lane 0, then the model's solver and integration variants. One wavefront per env; envs
// without deferred pairs return at once.
__global__ void __launch_bounds__(kWave) k_gjk_resume(gsr_model m, gsr_state s) {
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

template <bool kGjk>
__device__ __forceinline__ void fused_tile(ConstTypes* types, const BatchRange* __restrict__ R, int n_groups,
                                           int tile, char* lds) {
  // the range holding the tile: binary search over the cumulative tile ends
  int lo = 0, hi = n_groups - 1;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (tile < R[mid].tile_end) hi = mid;
    else lo = mid + 1;
  }
  const BatchRange g = R[lo];
  const int T = g.envs_per_tile;
  const int first = (tile - g.tile_begin) * T;
  const int n_here = min(T, g.n_envs - first);
  const int env0 = g.env_begin + first;
  // the type index is uniform, so the descriptor fields are scalar loads from the kernel-argument segment
  const int k = __builtin_amdgcn_readfirstlane(g.type);
  const FusedTypes* ty = (const FusedTypes*)types;
  run_tile<kGjk>(ty->m[k], ty->s[k], env0, n_here, lds);
}

// Phased form of the fused batch: one launch per stage of the plain substep, each over every tile of the batch, so
// each stage gets its own register allocation (the single fused kernel holds every stage's code and spills) while
// the scene types still share each launch. Stage P: 0 = step_1 dynamics, 1 = equality rows + collision + inequality
// rows (GJK/EPA pairs deferred), 2 = the GJK/EPA pass and rows of deferred envs, 3 = Newton solve, 4 = step_2. The
// device functions and their order per env are the plain step's, so results are bitwise those of gsr_step.
template <int P>
__device__ __forceinline__ void phase_tile(const gsr_model& m, const gsr_state& s, int env0, int n_here, char* lds) {
  const int lane = threadIdx.x;
  const bool mine = lane < n_here;
  const bool wave = P == 0 || P == 4 ? m.dyn_variant == 1 : P == 1 ? m.collide_variant == 1 : P == 3 ? m.solver_variant == 1 : true;
  if (!wave) {
    if (!mine) return;
    Env e(m, s, env0 + lane);
    if (P == 0) {
      dev_smooth_dynamics(e);
    } else if (P == 1) {
      dev_equality_rows(e);
      dev_collision(e);
      if (!gjk_pending(e)) dev_inequality_rows(e);
    } else if (P == 3) {
      dev_solve(e);
    } else if (P == 4) {
      dev_integrate(e);
      dev_kinematics(e);
    }
    return;
  }
  for (int j = 0; j < n_here; j++) {
    Env e(m, s, env0 + j);
    if (P == 0) wave_step1(e);
    if (P == 1) wave_collide(e, lds);
    if (P == 2 && wave_gjk_fixup(e)) wave_rows(e);
    if (P == 3) wave_solve(e, lds);
    if (P == 4) wave_step2(e);
    __syncthreads();
  }
}

template <int P>
__global__ void __launch_bounds__(kWave) k_phase(const char* __restrict__ table) {
  extern __shared__ char lds[];
  const BatchHeader* __restrict__ H = reinterpret_cast<const BatchHeader*>(table);
  ConstTypes* types = (ConstTypes*)(table + kTypesOffset);
  const BatchRange* __restrict__ R = reinterpret_cast<const BatchRange*>(table + kRangesOffset);
  const int n_tiles = H->n_tiles, n_groups = H->n_groups;
  if ((int)blockIdx.x >= n_tiles) return;
  const int tile = xcd_env(blockIdx.x, n_tiles);
  int lo = 0, hi = n_groups - 1;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (tile < R[mid].tile_end) hi = mid;
    else lo = mid + 1;
  }
  const BatchRange g = R[lo];
  const int T = g.envs_per_tile;
  const int first = (tile - g.tile_begin) * T;
  const int n_here = min(T, g.n_envs - first);
  const int k = __builtin_amdgcn_readfirstlane(g.type);
  const FusedTypes* ty = (const FusedTypes*)types;
  phase_tile<P>(ty->m[k], ty->s[k], g.env_begin + first, n_here, lds);
}

// W = resident wavefronts per SIMD the register allocation must allow (2: 256 VGPRs, 4: 128 VGPRs);
// G = GJK/EPA inside the launch (1) or deferred to k_gjk_resume (0)
template <int W, int G>
__global__ void __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(W, 8)))
k_fused(const char* __restrict__ table, uint32_t* queue, int persistent) {
  extern __shared__ char lds[];
  const BatchHeader* __restrict__ H = reinterpret_cast<const BatchHeader*>(table);
  ConstTypes* types = (ConstTypes*)(table + kTypesOffset);
  const BatchRange* __restrict__ R = reinterpret_cast<const BatchRange*>(table + kRangesOffset);
  const int n_tiles = H->n_tiles, n_groups = H->n_groups;
  if (!persistent) {
    if ((int)blockIdx.x < n_tiles) fused_tile<G != 0>(types, R, n_groups, xcd_env(blockIdx.x, n_tiles), lds);
    return;
  }
  for (;;) {
    uint32_t t = 0;
    if (threadIdx.x == 0) t = atomicAdd(&queue[0], 1u);
    const int tile = __builtin_amdgcn_readfirstlane((int)__shfl((int)t, 0, kWave));
    if (tile >= n_tiles) break;
    fused_tile<G != 0>(types, R, n_groups, tile, lds);
    __syncthreads();
  }
  // the last workgroup out resets the queue for the next launch (no host memset, so graph replays stay valid)
  if (threadIdx.x == 0) {
    __threadfence();
    uint32_t d = atomicAdd(&queue[1], 1u);
    if (d == gridDim.x - 1) {
      atomicExch(&queue[0], 0u);
      atomicExch(&queue[1], 0u);
    }
  }
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
__global__ void k_get_equalities(gsr_model m, gsr_state s, int32_t* iout, float* fout) {
  const int B = s.n_envs, E = m.n_equalities;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)B * E) return;
  const int i = (int)(t / B), b = (int)(t % B);
  const Env e(m, s, b);
  const int64_t row = (int64_t)b * E + i;
  iout[3 * row] = m.eq_type[i];
  iout[3 * row + 1] = m.eq_obj1[i];
  iout[3 * row + 2] = m.eq_obj2[i];
  // equalities of other types are rejected when the model is compiled; their rows would be 3 (CONNECT) or 6 (WELD)
  int c = 0;
  for (int j = 0; j < i; j++) c += m.eq_type[j] == EQ_JOINT;
  fout[6 * row] = m.eq_type[i] == EQ_JOINT ? e.EF(s.efc_force, c) : 0.f;
  for (int k = 1; k < 6; k++) fout[6 * row + k] = 0.f;
}

int check(const gsr_model* m, const gsr_state* s) {
  if (!m || !s) return GSR_E_ARG;
  if (s->n_envs <= 0) return GSR_E_ARG;
  if (m->row_width > kMaxRowWidth) return GSR_E_SHAPE;
  if (m->n_geoms >= (1 << 15)) return GSR_E_SHAPE;
  if (m->solver_variant == 1) {
    if (m->n_entities > kMaxPatternEntities) return GSR_E_SHAPE;
    if (m->wave_h_floats < 0 || wave_lds_bytes(m->n_dofs, m->n_entities, m->wave_h_floats) > 160 * 1024)
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
      if (m->collide_variant == 1)
        hipLaunchKernelGGL(k_collide_wave, dim3(s->n_envs), dim3(kWave), wave_collide_lds_bytes(m->n_geoms), st, *m,
                           *s);
      else
        hipLaunchKernelGGL(k_collide, g, dim3(kBlock), 0, st, *m, *s);
      hipLaunchKernelGGL(k_gjk_fixup, dim3(s->n_envs), dim3(kWave), 0, st, *m, *s);
      break;
    case 2:
      if (m->solver_variant == 1)
        hipLaunchKernelGGL(k_solve_wave, dim3(s->n_envs), dim3(kWave),
                           wave_lds_bytes(m->n_dofs, m->n_entities, m->wave_h_floats), st,
                           *m, *s);
      else
        hipLaunchKernelGGL(k_solve, g, dim3(kBlock), 0, st, *m, *s);
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
  const int64_t n = (int64_t)state->n_envs * model->n_equalities;
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
      lds = std::max(lds, (size_t)wave_lds_bytes(m.n_dofs, m.n_entities, m.wave_h_floats));
  }
  return (int)lds;
}

int gsr_batch_upload(const gsr_group* groups, int32_t n_groups, void* dev_table, int64_t table_bytes,
                     gsr_batch_info* info) {
  if (!groups || !dev_table || !info || n_groups < 1 || n_groups > GSR_MAX_GROUPS) return GSR_E_ARG;
  if (table_bytes < gsr_batch_table_bytes(n_groups)) return GSR_E_ARG;
  std::vector<BatchRange> ranges(n_groups);
  gsr_batch_info out;
  std::memset(&out, 0, sizeof(out));
  int tiles = 0;
  size_t lds = 0;
  for (int g = 0; g < n_groups; g++) {
    const gsr_group& G = groups[g];
    int rc = check(&G.model, &G.state);
    if (rc) return rc;
    if (G.n_envs < 1 || G.env_begin < 0 || G.env_begin + G.n_envs > G.state.n_envs) return GSR_E_ARG;
    if (G.envs_per_tile < 1 || G.envs_per_tile > kWave) return GSR_E_ARG;
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
      lds = std::max(lds, wave_lds_bytes(G.model.n_dofs, G.model.n_entities, G.model.wave_h_floats));
    if (G.model.collide_variant == 1) lds = std::max(lds, wave_collide_lds_bytes(G.model.n_geoms));
    int n_t = (G.n_envs + G.envs_per_tile - 1) / G.envs_per_tile;
    ranges[g] = BatchRange{type, G.env_begin, G.n_envs, G.envs_per_tile, tiles, tiles + n_t, 0, 0};
    tiles += n_t;
  }
  BatchHeader h{n_groups, tiles, 0u, 0u};
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
  if (hipMemcpy(dev_table, &h, 2 * sizeof(int32_t), hipMemcpyHostToDevice) != hipSuccess) return GSR_E_LAUNCH;
  int dev = 0, n_cu = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess) return GSR_E_LAUNCH;
  if (hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return GSR_E_LAUNCH;
  if (info->phased) {
    const int lc = phase_lds(out, 1), ls = phase_lds(out, 3);
    if (lc > 0 && hipFuncSetAttribute((const void*)k_phase<1>, hipFuncAttributeMaxDynamicSharedMemorySize, lc) !=
                      hipSuccess)
      return GSR_E_LAUNCH;
    if (ls > 0 && hipFuncSetAttribute((const void*)k_phase<3>, hipFuncAttributeMaxDynamicSharedMemorySize, ls) !=
                      hipSuccess)
      return GSR_E_LAUNCH;
    out.n_groups = n_groups;
    out.n_tiles = tiles;
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
  const void* fn = W == 4 ? (G ? (const void*)k_fused<4, 1> : (const void*)k_fused<4, 0>)
                          : (G ? (const void*)k_fused<2, 1> : (const void*)k_fused<2, 0>);
  if (lds > 0 && hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
    return GSR_E_LAUNCH;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kWave, lds) != hipSuccess) return GSR_E_LAUNCH;
  out.n_groups = n_groups;
  out.n_tiles = tiles;
  out.persistent = info->persistent ? 1 : 0;
  out.waves_per_simd = W;
  out.gjk_inline = G;
  out.phased = 0;
  out.grid = out.persistent ? std::max(1, std::min(tiles, n_cu * std::max(per_cu, 1))) : tiles;
  out.lds_bytes = (int32_t)lds;
  *info = out;
  return GSR_OK;
}

int gsr_step_fused(const void* dev_table, const gsr_batch_info* info, int32_t n_substeps, hipStream_t stream) {
  if (!dev_table || !info || n_substeps < 1 || info->grid < 1 || info->n_tiles < 1 || info->n_types < 1 ||
      info->n_types > GSR_MAX_TYPES)
    return GSR_E_ARG;
  const char* t = static_cast<const char*>(dev_table);
  uint32_t* queue = reinterpret_cast<uint32_t*>(const_cast<char*>(t) + 2 * sizeof(int32_t));
  if (info->phased) {
    const dim3 grid(info->n_tiles), block(kWave);
    const int lc = phase_lds(*info, 1), ls = phase_lds(*info, 3);
    for (int i = 0; i < n_substeps; i++) {
      hipLaunchKernelGGL(k_phase<0>, grid, block, 0, stream, t);
      hipLaunchKernelGGL(k_phase<1>, grid, block, lc, stream, t);
      hipLaunchKernelGGL(k_phase<2>, grid, block, 0, stream, t);
      hipLaunchKernelGGL(k_phase<3>, grid, block, ls, stream, t);
      hipLaunchKernelGGL(k_phase<4>, grid, block, 0, stream, t);
    }
    return launch_status();
  }
  for (int i = 0; i < n_substeps; i++) {
    const dim3 grid(info->grid), block(kWave);
    const int pers = (int)info->persistent;
    if (info->waves_per_simd == 4) {
      if (info->gjk_inline) hipLaunchKernelGGL((k_fused<4, 1>), grid, block, info->lds_bytes, stream, t, queue, pers);
      else hipLaunchKernelGGL((k_fused<4, 0>), grid, block, info->lds_bytes, stream, t, queue, pers);
    } else {
      if (info->gjk_inline) hipLaunchKernelGGL((k_fused<2, 1>), grid, block, info->lds_bytes, stream, t, queue, pers);
      else hipLaunchKernelGGL((k_fused<2, 0>), grid, block, info->lds_bytes, stream, t, queue, pers);
    }
    if (info->gjk_inline) continue;
    for (int k = 0; k < info->n_types; k++) {
      const gsr_model& m = info->models[k];
      const gsr_state& st = info->states[k];
      const size_t lds = m.solver_variant == 1 ? wave_lds_bytes(m.n_dofs, m.n_entities, m.wave_h_floats) : 0;
      hipLaunchKernelGGL(k_gjk_resume, dim3(st.n_envs), dim3(kWave), lds, stream, m, st);
    }
  }
  return launch_status();
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
    if (graph) hipGraphDestroy(graph);
    return rc;
  }
  if (err != hipSuccess || !graph) return GSR_E_LAUNCH;
  hipGraphExec_t exec = nullptr;
  err = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  hipGraphDestroy(graph);
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

}  // extern "C"
