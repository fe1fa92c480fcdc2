// Serving form: a device-side request scheduler over resident env pools (no counterpart in the reference).
//
// Each pool is the state of one scene type whose envs are cut into fixed tiles of `envs_per_tile` envs. A request
// (scene type, n_envs, horizon, seed) occupies ceil(n_envs / envs_per_tile) tiles of its pool, not necessarily
// adjacent: request-local env i lives in the tile whose `tile_first` <= i < tile_first + envs_per_tile. One tick is
// five stream-ordered launches, captured once in a hipGraph and replayed with no host synchronisation:
//   k_serve_schedule  one workgroup: admits arrived requests in arrival order while request slots, free tiles, the
//                     memory budget and the tick budget allow (a request is always admitted when nothing is active),
//                     stamps the tick start, and writes the fused launch's tile table: one range per active tile,
//                     pools in the given order (heaviest first)
//   k_serve_init      one wavefront per tile: envs of newly admitted tiles get a fresh state (the request-local
//                     env's initial pose from the pool, zero velocities and controls, cleared caches) and their
//                     kinematics
//   k_serve_control   one wavefront per tile: PD targets of active envs from a counter-based hash of (request
//                     seed, request-local env, request-local step, dof), so a request's controls do not depend on
//                     where its envs sit or what else runs
//   (fused substep)   the batching form's kernels over the tile table (gsr_launch_fused_kernel)
//   k_serve_post      one workgroup: fatal errno bits mark their request failed; every active request advances one
//                     step; a request at its horizon frees its tiles and posts a completion record; the tick end is
//                     stamped
// Host and device share a ring of arrivals and a ring of completions in coherent pinned host memory: the host
// appends arrivals and reads completions without waiting for the device; the device reads arrivals and appends
// completions with system-scope acquire / release accesses.
#include <hip/hip_runtime.h>

#include <cstring>

#include "rigid_tiles.h"

namespace gsr {
namespace {

constexpr int kSchedThreads = 256;
constexpr uint32_t kFatalErrno = GSR_ERR_BROAD_OVERFLOW | GSR_ERR_CONTACT_OVERFLOW | GSR_ERR_NAN_QACC |
                                 GSR_ERR_NAN_STATE;

struct ServeSlot {
  int32_t rid, pool, n_envs, horizon;
  int32_t steps, failed, active, pad;
  uint32_t seed, pad2;
  double mem, load;
};

// device-side scheduler state (caller-allocated device memory of gsr_serve_state_bytes)
struct ServeDev {
  int32_t n_pools, n_tiles, tick, n_active;
  int32_t arrivals_tail, done_count, pad0, pad1;
  double mem_used, load_used, mem_budget, load_budget;
  long long env_steps;
  long long pad2;
  gsr_serve_pool pools[GSR_MAX_TYPES];
  int32_t pool_tile0[GSR_MAX_TYPES];  // global index of each pool's first tile
  ServeSlot slots[GSR_SERVE_MAX_REQUESTS];
};

constexpr size_t kTileArraysOffset = (sizeof(ServeDev) + 255) & ~(size_t)255;

struct Tiles {
  int32_t *req, *first, *n, *init;
};

__device__ __forceinline__ Tiles tiles_of(ServeDev* S) {
  int32_t* base = reinterpret_cast<int32_t*>(reinterpret_cast<char*>(S) + kTileArraysOffset);
  const int T = S->n_tiles;
  return Tiles{base, base + T, base + 2 * T, base + 3 * T};
}

__device__ __forceinline__ int pool_of_tile(const ServeDev* S, int t) {
  int p = 0;
  while (p + 1 < S->n_pools && t >= S->pool_tile0[p + 1]) p++;
  return p;
}

template <class T>
__device__ __forceinline__ T load_acquire_sys(const T* p) {
  return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM);
}
template <class T>
__device__ __forceinline__ void store_release_sys(T* p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// block-wide exclusive prefix sum of one int per thread (kSchedThreads threads); returns the block total
__device__ __forceinline__ int block_scan(int v, int* sh, int& total) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int o = 1; o < kSchedThreads; o <<= 1) {
    int x = t >= o ? sh[t - o] : 0;
    __syncthreads();
    sh[t] += x;
    __syncthreads();
  }
  total = sh[kSchedThreads - 1];
  const int excl = sh[t] - v;
  __syncthreads();
  return excl;
}

__global__ void __launch_bounds__(kSchedThreads) k_serve_schedule(ServeDev* S, char* table, gsr_serve_host* host) {
  __shared__ int sh[kSchedThreads];
  __shared__ int s_go, s_slot, s_pool, s_need, s_free;
  __shared__ gsr_serve_request s_req;
  const int t = threadIdx.x;
  Tiles T = tiles_of(S);
  if (t == 0) host->tick_t0[S->tick % GSR_SERVE_RING] = (uint64_t)wall_clock64();
  // admission, in arrival order; stops at the first request that does not fit
  const int head = load_acquire_sys(&host->arrivals_head);
  for (;;) {
    if (t == 0) {
      s_go = 0;
      const int tail = S->arrivals_tail;
      if (tail < head) {
        s_req = host->arrivals[tail % GSR_SERVE_RING];
        const gsr_serve_request& r = s_req;
        int slot = -1;
        for (int i = 0; i < GSR_SERVE_MAX_REQUESTS && slot < 0; i++)
          if (!S->slots[i].active) slot = i;
        const gsr_serve_pool& P = S->pools[r.pool];
        const double mem = (double)r.n_envs * (double)P.bytes_per_env;
        const double load = (double)r.n_envs * (double)P.cost_per_env;
        const bool fits = slot >= 0 && S->mem_used + mem <= S->mem_budget &&
                          (S->n_active == 0 || S->load_used + load <= S->load_budget);
        if (fits) {
          s_go = 1;
          s_slot = slot;
          s_pool = r.pool;
          s_need = (r.n_envs + P.envs_per_tile - 1) / P.envs_per_tile;
        }
      }
    }
    __syncthreads();
    if (!s_go) break;
    const int p = s_pool, t0 = S->pool_tile0[p], nt = S->pools[p].n_tiles, ept = S->pools[p].envs_per_tile;
    // free tiles of the pool, counted and then ranked so the first `need` of them are taken in tile order
    int cnt = 0;
    for (int i = t; i < nt; i += kSchedThreads) cnt += T.req[t0 + i] < 0;
    int total;
    block_scan(cnt, sh, total);
    if (t == 0) s_free = total;
    __syncthreads();
    if (s_free < s_need) break;  // uniform: head-of-line request waits for tiles
    int base = 0;
    for (int c0 = 0; c0 < nt; c0 += kSchedThreads) {
      const int i = c0 + t;
      const int f = i < nt && T.req[t0 + i] < 0;
      int chunk;
      const int rank = base + block_scan(f, sh, chunk);
      if (f && rank < s_need) {
        const int g = t0 + i;
        T.req[g] = s_slot;
        T.first[g] = rank * ept;
        T.n[g] = min(ept, s_req.n_envs - rank * ept);
        T.init[g] = 1;
      }
      base += chunk;
      if (base >= s_need) break;  // uniform
    }
    if (t == 0) {
      ServeSlot& sl = S->slots[s_slot];
      const gsr_serve_pool& P = S->pools[p];
      sl.rid = s_req.rid;
      sl.pool = p;
      sl.n_envs = s_req.n_envs;
      sl.horizon = s_req.horizon;
      sl.seed = s_req.seed;
      sl.steps = 0;
      sl.failed = 0;
      sl.active = 1;
      sl.mem = (double)s_req.n_envs * (double)P.bytes_per_env;
      sl.load = (double)s_req.n_envs * (double)P.cost_per_env;
      S->mem_used += sl.mem;
      S->load_used += sl.load;
      S->n_active += 1;
      S->arrivals_tail += 1;
    }
    __syncthreads();
  }
  if (t == 0) store_release_sys(&host->arrivals_tail, S->arrivals_tail);
  // tile table of this tick: one range per active tile, pools in order
  BatchHeader* H = reinterpret_cast<BatchHeader*>(table);
  BatchRange* R = reinterpret_cast<BatchRange*>(table + kRangesOffset);
  int base = 0;
  for (int c0 = 0; c0 < S->n_tiles; c0 += kSchedThreads) {
    const int g = c0 + t;
    const int a = g < S->n_tiles && T.n[g] > 0;
    int chunk;
    const int idx = base + block_scan(a, sh, chunk);
    if (a) {
      const int p = pool_of_tile(S, g);
      const int ept = S->pools[p].envs_per_tile;
      R[idx] = BatchRange{p, (g - S->pool_tile0[p]) * ept, T.n[g], ept, idx, idx + 1, ept, idx, idx + 1, 0, 0, 0};
    }
    base += chunk;
  }
  if (t == 0) {
    H->n_groups = base;
    H->n_tiles = base;
    H->n_wave_tiles = base;
  }
}

// reset of one env to a fresh request-local env (the persistent fields init_state and set_qpos write), then its
// kinematics and AABBs
__device__ void serve_reset_env(const Env& e, const gsr_serve_pool& P, int local) {
  const gsr_model& m = e.m;
  const gsr_state& s = e.s;
  const int cap = P.n_tiles * P.envs_per_tile;
  for (int q = 0; q < m.n_qs; q++) e.F(s.qpos, q) = P.init_qpos[(size_t)q * cap + local];
  for (int d = 0; d < m.n_dofs; d++) {
    e.F(s.vel, d) = 0.f;
    e.F(s.acc, d) = 0.f;
    e.F(s.acc_smooth, d) = 0.f;
    e.F(s.qacc_ws, d) = 0.f;
    e.F(s.ctrl_pos, d) = 0.f;
    e.F(s.ctrl_vel, d) = 0.f;
    e.F(s.ctrl_force, d) = 0.f;
    e.I(s.ctrl_mode, d) = CTRL_FORCE;
  }
  for (int p = 0; p < m.n_pairs; p++) e.S3(s.cache_normal, p, zero3());
  for (int c = 0; c < m.max_rows; c++) e.EI(s.efc_active, c) = 0;
  for (int l = 0; l < m.n_links; l++) {
    e.F(s.mass_shift, l) = 0.f;
    e.S3(s.i_pos_shift, l, zero3());
  }
  for (int g = 0; g < m.n_geoms; g++) e.F(s.friction_ratio, g) = 1.f;
  e.I(s.is_warmstart, 0) = 0;
  e.I(s.first_time, 0) = 1;
  e.I((int32_t*)s.errno_env, 0) = 0;
  e.I(s.gjk_defer_n, 0) = 0;
  dev_kinematics(e);
  if (m.enable_collision) dev_update_aabbs(e);
}

__global__ void __launch_bounds__(kWave) k_serve_init(ServeDev* S) {
  const int g = blockIdx.x;
  if (g >= S->n_tiles) return;
  Tiles T = tiles_of(S);
  if (!T.init[g]) return;  // uniform across the block
  const int p = pool_of_tile(S, g);
  const gsr_serve_pool& P = S->pools[p];
  const int lane = threadIdx.x;
  if (lane < T.n[g]) {
    const Env e(P.model, P.state, (g - S->pool_tile0[p]) * P.envs_per_tile + lane);
    serve_reset_env(e, P, T.first[g] + lane);
  }
  __syncthreads();
  if (lane == 0) T.init[g] = 0;
}

// counter-based uniform in [0, 1): a 32-bit integer hash (xor-shift-multiply rounds of the MurmurHash3 finaliser)
// chained over the four key words, top 24 bits as the mantissa
__device__ __forceinline__ uint32_t mix32(uint32_t h) {
  h ^= h >> 16;
  h *= 0x85ebca6bu;
  h ^= h >> 13;
  h *= 0xc2b2ae35u;
  h ^= h >> 16;
  return h;
}
__device__ __forceinline__ float hash_uniform(uint32_t seed, uint32_t env, uint32_t step, uint32_t dof) {
  uint32_t h = mix32(seed ^ 0x9e3779b9u);
  h = mix32(h ^ (env * 0x27d4eb2fu));
  h = mix32(h ^ (step * 0x165667b1u));
  h = mix32(h ^ (dof * 0x61c88647u));
  return (float)(h >> 8) * (1.0f / 16777216.0f);
}

__global__ void __launch_bounds__(kWave) k_serve_control(ServeDev* S) {
  const int g = blockIdx.x;
  if (g >= S->n_tiles) return;
  Tiles T = tiles_of(S);
  const int n = T.n[g];
  if (n <= 0) return;
  const int p = pool_of_tile(S, g);
  const gsr_serve_pool& P = S->pools[p];
  if (P.ctrl_kind == 0) return;
  const ServeSlot& sl = S->slots[T.req[g]];
  const int step = sl.steps;
  if (P.ctrl_kind == 1 && step % P.ctrl_period != 0) return;
  const int lane = threadIdx.x;
  if (lane >= n) return;
  const Env e(P.model, P.state, (g - S->pool_tile0[p]) * P.envs_per_tile + lane);
  const uint32_t local = (uint32_t)(T.first[g] + lane);
  for (int d = P.ctrl_dof_begin; d < P.ctrl_dof_end; d++) {
    const int k = d - P.ctrl_dof_begin;
    const float u = hash_uniform(sl.seed, local, (uint32_t)step, (uint32_t)k);
    float target;
    if (P.ctrl_kind == 1) target = P.ctrl_lo[k] + (P.ctrl_hi[k] - P.ctrl_lo[k]) * u;
    else target = P.ctrl_base[k] + (u * (2.f * P.ctrl_amp) - P.ctrl_amp);
    e.F(e.s.ctrl_pos, d) = target;
    e.F(e.s.ctrl_vel, d) = 0.f;
    e.I(e.s.ctrl_mode, d) = CTRL_POSITION;
  }
}

__global__ void __launch_bounds__(kSchedThreads) k_serve_post(ServeDev* S, gsr_serve_host* host) {
  __shared__ int s_failed[GSR_SERVE_MAX_REQUESTS];
  __shared__ long long s_envs;
  const int t = threadIdx.x;
  Tiles T = tiles_of(S);
  for (int i = t; i < GSR_SERVE_MAX_REQUESTS; i += kSchedThreads) s_failed[i] = 0;
  if (t == 0) s_envs = 0;
  __syncthreads();
  // fatal errno bits of the stepped envs (one thread per tile)
  long long envs = 0;
  for (int g = t; g < S->n_tiles; g += kSchedThreads) {
    const int n = T.n[g];
    if (n <= 0) continue;
    envs += n;
    const int p = pool_of_tile(S, g);
    const gsr_serve_pool& P = S->pools[p];
    const gsr_state& st = P.state;
    const int b0 = (g - S->pool_tile0[p]) * P.envs_per_tile;
    uint32_t e = 0;
    for (int j = 0; j < n; j++) e |= st.errno_env[b0 + j];
    if (e & kFatalErrno) atomicOr(&s_failed[T.req[g]], 1);
  }
  atomicAdd((unsigned long long*)&s_envs, (unsigned long long)envs);
  __syncthreads();
  // retire requests at their horizon: their tiles become free; completion records in slot order
  __shared__ int s_done[GSR_SERVE_MAX_REQUESTS];
  for (int i = t; i < GSR_SERVE_MAX_REQUESTS; i += kSchedThreads) {
    ServeSlot& sl = S->slots[i];
    s_done[i] = 0;
    if (!sl.active) continue;
    sl.failed |= s_failed[i];
    sl.steps += 1;
    s_done[i] = sl.steps >= sl.horizon;
  }
  __syncthreads();
  for (int g = t; g < S->n_tiles; g += kSchedThreads) {
    const int r = T.req[g];
    if (r >= 0 && s_done[r]) {
      T.req[g] = -2 - r;  // free; remembers the retired slot until the tile is taken again
      T.n[g] = 0;
    }
  }
  __syncthreads();
  if (t == 0) {
    int dc = S->done_count;
    for (int i = 0; i < GSR_SERVE_MAX_REQUESTS; i++) {
      ServeSlot& sl = S->slots[i];
      if (!s_done[i]) continue;
      gsr_serve_done& d = host->done[dc % GSR_SERVE_RING];
      d.rid = sl.rid;
      d.tick = S->tick;
      d.failed = sl.failed;
      d.slot = i;
      dc++;
      sl.active = 0;
      S->mem_used -= sl.mem;
      S->load_used -= sl.load;
      S->n_active -= 1;
    }
    if (S->n_active == 0) {  // no drift of the budget counters across an idle point
      S->mem_used = 0.0;
      S->load_used = 0.0;
    }
    S->done_count = dc;
    S->env_steps += s_envs;
    host->tick_t1[S->tick % GSR_SERVE_RING] = (uint64_t)wall_clock64();
    host->tick_envs[S->tick % GSR_SERVE_RING] = (int32_t)s_envs;
    S->tick += 1;
    host->env_steps = S->env_steps;
    store_release_sys(&host->done_count, dc);
    store_release_sys(&host->ticks_done, S->tick);
  }
}

__global__ void k_serve_reset(ServeDev* S) {
  Tiles T = tiles_of(S);
  for (int g = blockIdx.x * blockDim.x + threadIdx.x; g < S->n_tiles; g += gridDim.x * blockDim.x) {
    T.req[g] = -1;
    T.first[g] = 0;
    T.n[g] = 0;
    T.init[g] = 0;
  }
}

}  // namespace

size_t serve_state_bytes(int n_tiles) { return kTileArraysOffset + sizeof(int32_t) * 4 * (size_t)n_tiles; }
size_t serve_table_bytes(int n_tiles) { return kRangesOffset + sizeof(BatchRange) * (size_t)n_tiles; }

int serve_init(const gsr_serve_pool* pools, int n_pools, double mem_budget, double tick_budget, void* dev_state,
               int64_t state_bytes, void* dev_table, int64_t table_bytes, hipStream_t stream) {
  if (n_pools < 1 || n_pools > GSR_MAX_TYPES) return GSR_E_ARG;
  ServeDev* h = new ServeDev;
  std::memset(h, 0, sizeof(ServeDev));
  h->n_pools = n_pools;
  int tiles = 0;
  for (int p = 0; p < n_pools; p++) {
    h->pools[p] = pools[p];
    h->pool_tile0[p] = tiles;
    tiles += pools[p].n_tiles;
  }
  h->n_tiles = tiles;
  h->mem_budget = mem_budget;
  h->load_budget = tick_budget;
  FusedTypes* ty = new FusedTypes;
  std::memset(ty, 0, sizeof(FusedTypes));
  for (int p = 0; p < n_pools; p++) {
    ty->m[p] = pools[p].model;
    ty->s[p] = pools[p].state;
  }
  int rc = GSR_OK;
  BatchHeader bh{0, 0, 0u, 0u, 0, 0, 0, 0};
  if ((size_t)state_bytes < serve_state_bytes(tiles) || (size_t)table_bytes < serve_table_bytes(tiles)) {
    rc = GSR_E_ARG;
  } else if (hipMemcpyAsync(dev_state, h, sizeof(ServeDev), hipMemcpyHostToDevice, stream) != hipSuccess ||
             hipMemcpyAsync(static_cast<char*>(dev_table) + kTypesOffset, ty, sizeof(FusedTypes),
                            hipMemcpyHostToDevice, stream) != hipSuccess ||
             hipMemcpyAsync(dev_table, &bh, sizeof(bh), hipMemcpyHostToDevice, stream) != hipSuccess) {
    rc = GSR_E_LAUNCH;
  } else {
    hipLaunchKernelGGL(k_serve_reset, dim3((tiles + 255) / 256), dim3(256), 0, stream,
                       static_cast<ServeDev*>(dev_state));
    // the copies read host memory freed below: finish them first
    if (hipStreamSynchronize(stream) != hipSuccess || hipGetLastError() != hipSuccess) rc = GSR_E_LAUNCH;
  }
  delete h;
  delete ty;
  return rc;
}

void launch_serve(int stage, void* dev_state, void* dev_table, gsr_serve_host* host_dev, int n_tiles,
                  hipStream_t stream) {
  ServeDev* S = static_cast<ServeDev*>(dev_state);
  char* table = static_cast<char*>(dev_table);
  switch (stage) {
    case 0: hipLaunchKernelGGL(k_serve_schedule, dim3(1), dim3(kSchedThreads), 0, stream, S, table, host_dev); break;
    case 1: hipLaunchKernelGGL(k_serve_init, dim3(n_tiles), dim3(kWave), 0, stream, S); break;
    case 2: hipLaunchKernelGGL(k_serve_control, dim3(n_tiles), dim3(kWave), 0, stream, S); break;
    default: hipLaunchKernelGGL(k_serve_post, dim3(1), dim3(kSchedThreads), 0, stream, S, host_dev);
  }
}

}  // namespace gsr
