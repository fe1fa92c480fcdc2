// Device code shared by the kernel translation units of libgenesis_rigid.so: env mapping, the per-wavefront
// step_1 / step_2 bodies, the batch table layout and the tile runners of the batching form, and the host-side
// launchers each kernel translation unit defines (the library is built from several units compiled in
// parallel; every kernel is launched only from the unit that defines it).
#pragma once
#include <hip/hip_runtime.h>

#include "genesis_rigid.h"
#include "rigid_kin.h"
#include "rigid_collide.h"
#include "rigid_solver.h"
#include "rigid_solver_wave.h"
#include "rigid_collide_wave.h"

namespace gsr {

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

// an env whose narrowphase deferred pairs to the GJK/EPA pass: its contact list and rows are completed there
__device__ __forceinline__ bool gjk_pending(const Env& e) { return e.I(e.s.gjk_defer_n, 0) > 0; }

// ---- batching form ----------------------------------------------------------------------------------------------
// Device table of a heterogeneous batch: a header (work queue) then one record per group (env range). The model and
// state descriptors of the batch's scene types are kernel arguments, so every field read is a scalar load from the
// kernel-argument segment, as in the plain kernels.
struct BatchHeader {
  int32_t n_groups, n_tiles;
  uint32_t counter, done;  // work queue: next tile, finished workgroups (the last one resets both)
  int32_t n_wave_tiles;    // tiles of the phased launch's wavefront-per-env stages (GJK/EPA pass, solve)
  int32_t pad0, pad1, pad2;
};
// A group's envs cut into tiles twice: envs_per_tile for the single fused kernel and the phased per-env stages, and
// wave_envs_per_tile for the phased GJK/EPA pass and solve, whose wavefront variants take a tile's envs one after
// the other (a lane-per-env collision wants 64 envs a tile, a wavefront solve few)
struct BatchRange {
  int32_t type, env_begin, n_envs, envs_per_tile, tile_begin, tile_end;
  int32_t wave_envs_per_tile, wave_tile_begin, wave_tile_end, pad0, pad1, pad2;
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

// two waves per SIMD at least: the collision phases would otherwise take 256 VGPRs plus AGPRs (one wave per SIMD)
template <int P>
__global__ void __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(2, 8))) k_phase(const char* __restrict__ table) {
  extern __shared__ char lds[];
  const BatchHeader* __restrict__ H = reinterpret_cast<const BatchHeader*>(table);
  ConstTypes* types = (ConstTypes*)(table + kTypesOffset);
  const BatchRange* __restrict__ R = reinterpret_cast<const BatchRange*>(table + kRangesOffset);
  constexpr bool kWaveTiles = P == 2 || P == 3;
  const int n_tiles = kWaveTiles ? H->n_wave_tiles : H->n_tiles, n_groups = H->n_groups;
  if ((int)blockIdx.x >= n_tiles) return;
  const int tile = xcd_env(blockIdx.x, n_tiles);
  int lo = 0, hi = n_groups - 1;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (tile < (kWaveTiles ? R[mid].wave_tile_end : R[mid].tile_end)) hi = mid;
    else lo = mid + 1;
  }
  const BatchRange g = R[lo];
  const int T = kWaveTiles ? g.wave_envs_per_tile : g.envs_per_tile;
  const int first = (tile - (kWaveTiles ? g.wave_tile_begin : g.tile_begin)) * T;
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

// ---- launchers (host code; each defined in the translation unit that holds its kernel) -----------------------
// plain step (rigid_k_collide.hip, rigid_k_solve.hip): the collision stage (collision kernel + GJK/EPA pass) and
// the Newton solve of the model's variants
void launch_collide(const gsr_model& m, const gsr_state& s, hipStream_t st);
void launch_solve(const gsr_model& m, const gsr_state& s, hipStream_t st);
// batching form: k_fused<W, G> (rigid_k_fused_wW_gG.hip), k_phase<P> (rigid_k_phase_P.hip), k_gjk_resume
// (rigid_k_resume.hip); fused_kernel / launch_fused / phase_kernel / launch_phase dispatch to them (rigid.hip)
#define GSR_FUSED_UNIT(W, G)                                                                                         \
  const void* fused_kernel_w##W##_g##G();                                                                            \
  void launch_fused_w##W##_g##G(dim3 grid, unsigned lds, hipStream_t st, const char* table, uint32_t* queue,        \
                                int persistent);
GSR_FUSED_UNIT(2, 0)
GSR_FUSED_UNIT(2, 1)
GSR_FUSED_UNIT(4, 0)
GSR_FUSED_UNIT(4, 1)
#undef GSR_FUSED_UNIT
#define GSR_PHASE_UNIT(P)                \
  const void* phase_kernel_##P();        \
  void launch_phase_##P(dim3 grid, unsigned lds, hipStream_t st, const char* table);
GSR_PHASE_UNIT(0)
GSR_PHASE_UNIT(1)
GSR_PHASE_UNIT(2)
GSR_PHASE_UNIT(3)
GSR_PHASE_UNIT(4)
#undef GSR_PHASE_UNIT
const void* fused_kernel(int W, int G);
void launch_fused(int W, int G, dim3 grid, unsigned lds, hipStream_t st, const char* table, uint32_t* queue,
                  int persistent);
const void* phase_kernel(int P);
void launch_phase(int P, dim3 grid, unsigned lds, hipStream_t st, const char* table);
void launch_gjk_resume(const gsr_model& m, const gsr_state& s, unsigned lds, hipStream_t st);
// serving form (rigid_serve.hip): scheduler state and tile table sizes, set-up, and the tick's own kernels
// (stage 0 schedule, 1 init, 2 control, 3 post)
size_t serve_state_bytes(int n_tiles);
size_t serve_table_bytes(int n_tiles);
int serve_init(const gsr_serve_pool* pools, int n_pools, double mem_budget, double tick_budget, void* dev_state,
               int64_t state_bytes, void* dev_table, int64_t table_bytes, hipStream_t stream);
void launch_serve(int stage, void* dev_state, void* dev_table, gsr_serve_host* host_dev, int n_tiles,
                  hipStream_t stream);

}  // namespace gsr
