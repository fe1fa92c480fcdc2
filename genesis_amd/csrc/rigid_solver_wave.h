// Newton solver with one wavefront (64 lanes) per environment and entity-block sparsity.
//
// Same algorithm as the lane-per-env solver in rigid_solver.h (func_init_solver / func_solve / func_linesearch /
// rank-1 Cholesky updates of constraint_solver_decomp.py), mapped onto a wavefront:
//   * per-row work (Jaref, activity, forces, line-search terms) is spread over the 64 lanes, sums over rows are
//     wave-shuffle tree reductions (every lane holds the identical result, so control flow stays uniform);
//   * per-dof work (M x, J^T f, Hessian rows) is spread over lanes by dof; every per-dof sum keeps the reference's
//     row order through per-entity row lists;
//   * the Hessian H = M + J^T D J lives only on its entity-block pattern: the symbolic Cholesky fill of the entity
//     graph of all constraint rows, so every entry outside it is an exact zero in the dense factorisation too and
//     skipping it changes no bit of the factors. When the pattern fits (gsr_model.wave_h_floats) H is kept in LDS,
//     row-compressed per entity; otherwise in the env's packed dense lower triangle in global memory.
// Vectors over dofs live in LDS; constraint rows live in global memory, env-major.
#pragma once
#include "rigid_env.h"

namespace gsr {

constexpr int kWave = 64;

// LDS pointers carry their address space so every access compiles to ds_* instructions (not flat), and the compiler
// knows they never alias global memory.
#define GSR_LDS __attribute__((address_space(3)))
typedef GSR_LDS float lds_f;
typedef GSR_LDS int lds_i;
typedef GSR_LDS unsigned long long lds_u64;
typedef GSR_LDS char lds_c;
constexpr int kMaxPatternEntities = 128;  // two 64-bit words per entity row of the pattern

__device__ __forceinline__ float wave_sum(float v) {
  for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}
// Sums over rows or dofs whose rounding steers discrete decisions (the line search's bracketing, the Newton stop
// test) are accumulated in double: the reference adds them one by one in FP32, an order a wavefront cannot follow,
// and a tree of FP32 partial sums has a different rounding error; the double sum is closer to both.
__device__ __forceinline__ double wave_sum_d(double v) {
  for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}
__device__ __forceinline__ int wave_or(int v) {
  for (int o = kWave / 2; o > 0; o >>= 1) v |= __shfl_xor(v, o, kWave);
  return v;
}

// A constraint row of up to kStage entries (every row of the arm, quadruped and box scenes) is loaded into
// registers with independent loads before it is used, so a lane does not wait for one global load per entry; wider
// rows take the generic loops.
constexpr int kStage = 12;
struct RowStage {
  float v[kStage];
  int d[kStage];
};
__device__ __forceinline__ void stage_row(const Env& e, int c, int nnz, RowStage& r) {
  const float* pv = &e.JV(c, 0);
  const int* pi = &e.JI(c, 0);
#pragma unroll
  for (int k = 0; k < kStage; k++) {
    r.v[k] = k < nnz ? pv[k] : 0.f;
    r.d[k] = k < nnz ? pi[k] : 0;
  }
}
// sum_k J[c, k] x[dof k] in entry order
__device__ __forceinline__ float row_dot_vec(const Env& e, int c, int nnz, float acc, const lds_f* x) {
  if (nnz <= kStage) {
    RowStage r;
    stage_row(e, c, nnz, r);
#pragma unroll
    for (int k = 0; k < kStage; k++)
      if (k < nnz) acc += r.v[k] * x[r.d[k]];
    return acc;
  }
  for (int k = 0; k < nnz; k++) acc += e.JV(c, k) * x[e.JI(c, k)];
  return acc;
}

// LDS layout of one env (floats / ints), sized on the host by wave_lds_bytes()
struct WaveLds {
  lds_f* qacc;
  lds_f* Ma;
  lds_f* search;
  lds_f* mv;
  lds_f* qfrc;
  lds_f* vec;  // factor-space work vector: rank-1 updates, then the gradient and the Newton direction
  // The factorisation works in elimination order: entity e is eliminated at position vof[e] (a minimum-degree order
  // of the env's entity graph), and "factor space" numbers entities by position and dofs position by position.
  // pat, es, ee, ent_of_dof, the levels and the LDS Hessian layout are in factor space; nent, nes, nee, moff are in
  // the model's own numbering.
  lds_u64* pat;  // [E][2] pattern bits (row v: positions v' with block (v, v') nonzero)
  lds_i* ent_of_dof;          // [D] factor space: position of the entity of each factor dof
  lds_i* nent;                // [D] entity of each model dof
  lds_i* pdof;                // [D] factor dof of each model dof
  lds_i* vof;                 // [E] position of each entity
  lds_i* ord;                 // [E] entity at each position
  lds_i* nes;                 // [E] first model dof of each entity
  lds_i* nee;                 // [E] end model dof of each entity
  lds_i* lvl_order;           // [E] entities sorted by forward level (elimination dependencies)
  lds_i* lvl_start;           // [E + 2]
  lds_i* rlvl_order;          // [E] entities sorted by backward level
  lds_i* rlvl_start;          // [E + 2]
  lds_i* cursor;              // [E] row-list fill cursors
  lds_i* n_lvl;               // [2] number of forward / backward levels
  // Hessian in LDS (when the pattern fits): row i of entity e starts at hoff[e] + r * below[e] + r * (r + 1) / 2
  // (r = i - es[e]) and holds the dofs of the pattern entities below e (entity k at pre[pstart[e] + rank of k in
  // the row]) followed by e's own dofs up to i.
  lds_f* hl;                // [gsr_model.wave_h_floats]
  lds_i* es;                  // [E] first factor dof of the entity at each position
  lds_i* ee;                  // [E] end factor dof of the entity at each position
  lds_i* hoff;                // [E]
  lds_i* below;               // [E]
  lds_i* pstart;              // [E]
  lds_i* pre;                 // [kWavePreSlots]
  lds_i* hmode;               // [1] 1 = H in LDS, 0 = H in global memory
  lds_i* moff;                // [E] offset of each entity's mass block (copy of ent_mass_offset, model numbering)
};

constexpr int kWavePreSlots = 256;   // off-diagonal pattern blocks addressable in LDS mode

// H = floats of LDS for the Hessian of one env (gsr_model.wave_h_floats): a pattern that needs more keeps H in global
// memory; a smaller budget lets more envs share a CU
__host__ __device__ inline size_t wave_lds_bytes(int D, int E, int H) {
  return sizeof(float) * (6 * (size_t)D + (size_t)H) + sizeof(unsigned long long) * 2 * (size_t)E +
         sizeof(int) * 3 * (size_t)D + sizeof(int) * (9 * (size_t)E + 6) + sizeof(int) * (6 * (size_t)E + kWavePreSlots + 1) +
         64;
}
__host__ __device__ inline size_t wave_lds_bytes(const gsr_model& m) {
  return wave_lds_bytes(m.n_dofs, m.n_entities, m.wave_h_floats);
}

__device__ inline WaveLds wave_lds_carve(lds_c* base, int D, int E, int H) {
  WaveLds w;
  lds_f* f = reinterpret_cast<lds_f*>(base);
  w.qacc = f; f += D;
  w.Ma = f; f += D;
  w.search = f; f += D;
  w.mv = f; f += D;
  w.qfrc = f; f += D;
  w.vec = f; f += D;
  w.hl = f; f += H;
  uint32_t p = (uint32_t)reinterpret_cast<uintptr_t>(f);
  lds_c* pc = reinterpret_cast<lds_c*>(f) + (((p + 7u) & ~7u) - p);
  w.pat = reinterpret_cast<lds_u64*>(pc);
  w.ent_of_dof = reinterpret_cast<lds_i*>(w.pat + 2 * E);
  w.nent = w.ent_of_dof + D;
  w.pdof = w.nent + D;
  w.vof = w.pdof + D;
  w.ord = w.vof + E;
  w.nes = w.ord + E;
  w.nee = w.nes + E;
  w.lvl_order = w.nee + E;
  w.lvl_start = w.lvl_order + E;
  w.rlvl_order = w.lvl_start + E + 2;
  w.rlvl_start = w.rlvl_order + E;
  w.cursor = w.rlvl_start + E + 2;
  w.n_lvl = w.cursor + E;
  w.es = w.n_lvl + 2;
  w.ee = w.es + E;
  w.hoff = w.ee + E;
  w.below = w.hoff + E;
  w.pstart = w.below + E;
  w.pre = w.pstart + E;
  w.hmode = w.pre + kWavePreSlots;
  w.moff = w.hmode + 1;
  return w;
}

__device__ __forceinline__ bool pat_has(const WaveLds& L, int e1, int e2) {
  return (L.pat[2 * e1 + (e2 >> 6)] >> (e2 & 63)) & 1ull;
}

struct WaveCtx {
  const Env& e;
  const WaveLds& L;
  int lane, D, E, n_con, ne, nef;
  // Models whose dofs all belong to one entity (an arm or a legged robot over a plane), at most 64 of them, with H
  // in LDS: H is one dense packed triangle (dense_h), and the rows staged per pass of the dense sums below (0: the
  // per-entity row walk) with their LDS staging area
  bool dense_h;
  int dense_rows;
  lds_f* dense_buf;
  // entity-block mass matrix entry (i, j) of entity ie (model numbering: entity and dofs), geometry read from LDS
  __device__ __forceinline__ float mm(int ie, int i, int j) const {
    int st = L.nes[ie], n = L.nee[ie] - st;
    return e.s.mass_blk[e.ix(L.moff[ie] + (i - st) * n + (j - st))];
  }
};

// entities touched by constraint row c (first and last stored dof; rows touch at most two entities), model numbering
__device__ __forceinline__ void row_entities(const WaveCtx& w, int c, int& ea, int& eb) {
  int nnz = w.e.EI(w.e.s.jac_nnz, c);
  ea = w.L.nent[w.e.JI(c, 0)];
  eb = w.L.nent[w.e.JI(c, nnz - 1)];
}

// Minimum-degree elimination order of the entity graph (entities as nodes, an edge per pair of entities sharing a
// constraint row), with its fill: eliminating a node joins its remaining neighbours into a clique. Ties go to the
// lower entity index; entities without dofs come last. Lanes own entities lane and lane + 64; adj0/adj1 hold their
// adjacency (two words each) on entry and the filled adjacency on exit. Writes L.ord and L.vof.
__device__ inline void wave_min_degree(const WaveCtx& w, unsigned long long (&adj0)[2], unsigned long long (&adj1)[2]) {
  const WaveLds& L = w.L;
  const int E = w.E, lane = w.lane;
  const int e0 = lane, e1 = lane + 64;
  const bool has0 = e0 < E && L.nee[e0] > L.nes[e0];
  const bool has1 = e1 < E && L.nee[e1] > L.nes[e1];
  unsigned long long al0 = __ballot(has0), al1 = __ballot(has1);
  int n_alive = __popcll(al0) + __popcll(al1);
  for (int step = 0; step < n_alive; step++) {
    // (degree, entity) minimum over the alive entities
    int k0 = has0 && ((al0 >> lane) & 1ull) ? ((__popcll(adj0[0] & al0) + __popcll(adj0[1] & al1)) << 8) | e0 : 0x7fffffff;
    int k1 = has1 && ((al1 >> lane) & 1ull) ? ((__popcll(adj1[0] & al0) + __popcll(adj1[1] & al1)) << 8) | e1 : 0x7fffffff;
    int k = min(k0, k1);
    for (int o = kWave / 2; o > 0; o >>= 1) k = min(k, __shfl_xor(k, o, kWave));
    const int v = k & 0xff;
    // v's live neighbours become a clique
    const int src = v & 63;
    unsigned long long n0 = __shfl(v < 64 ? adj0[0] : adj1[0], src, kWave);
    unsigned long long n1 = __shfl(v < 64 ? adj0[1] : adj1[1], src, kWave);
    if (v < 64) al0 &= ~(1ull << v); else al1 &= ~(1ull << (v - 64));
    n0 &= al0;
    n1 &= al1;
    if ((n0 >> lane) & 1ull) { adj0[0] |= n0 & ~(1ull << lane); adj0[1] |= n1; }
    if ((n1 >> lane) & 1ull) { adj1[0] |= n0; adj1[1] |= n1 & ~(1ull << lane); }
    if (lane == 0) {
      L.ord[step] = v;
      L.vof[v] = step;
    }
  }
  if (lane == 0) {
    int pos = n_alive;
    for (int ie = 0; ie < E; ie++)
      if (L.nee[ie] <= L.nes[ie]) {
        L.ord[pos] = ie;
        L.vof[ie] = pos++;
      }
  }
  __syncthreads();
}

// Build the factor-space structure: entity graph, minimum-degree order with its fill, pattern, dof maps, levels and
// LDS Hessian layout, plus the per-entity ordered row lists (model numbering).
// Row walks of multi-entity models (J^T f, J^T D J) give each lane the dofs it writes: one entity per lane, or one
// dof per lane when rows have at most kStage entries and the dofs fit two passes of the wavefront. A dof lane reads
// each of its entity's rows like an entity lane but does one entry's work per row instead of the row's whole pair
// set: measured on 20-box piles (120 dofs) the solve takes 15 % less time, on 64-box piles (384 dofs, six passes)
// 15 % more, hence the bound.
__device__ __forceinline__ bool wave_dof_lanes(const WaveCtx& w) {
  return w.dense_rows == 0 && w.e.m.row_width <= kStage && w.D <= 2 * kWave;
}

// The entries of entity ea (eb) in row c as one run: first dof | position << 16 | count << 24 when they are
// consecutive dofs at consecutive positions, else -1.
__device__ inline void row_runs(const WaveCtx& w, int c, int ea, int eb, int& run_a, int& run_b) {
  const Env& e = w.e;
  const int nnz = e.EI(e.s.jac_nnz, c);
  RowStage r;
  stage_row(e, c, e.m.row_width, r);
  int fa = -1, pa = 0, na = 0, fb = -1, pb = 0, nb = 0;
  bool oka = true, okb = true;
#pragma unroll
  for (int k = 0; k < kStage; k++) {
    if (k < nnz) {
      const int d = r.d[k], x = w.L.nent[d];
      if (x == ea) {
        if (na == 0) { fa = d; pa = k; } else oka = oka && d == fa + na && k == pa + na;
        na++;
      } else if (x == eb) {
        if (nb == 0) { fb = d; pb = k; } else okb = okb && d == fb + nb && k == pb + nb;
        nb++;
      }
    }
  }
  run_a = oka ? (fa | pa << 16 | na << 24) : -1;
  run_b = okb ? (fb | pb << 16 | nb << 24) : -1;
}

__device__ inline void wave_build_structure(const WaveCtx& w) {
  const Env& e = w.e;
  const gsr_model& m = e.m;
  const gsr_state& s = e.s;
  const WaveLds& L = w.L;
  for (int d = w.lane; d < w.D; d += kWave) L.nent[d] = m.dof_entity[d];
  for (int i = w.lane; i < w.E; i += kWave) {
    L.nes[i] = m.ent_dof_start[i];
    L.nee[i] = m.ent_dof_end[i];
    L.moff[i] = m.ent_mass_offset[i];
  }
  for (int i = w.lane; i < 2 * w.E; i += kWave) L.pat[i] = 0ull;
  __syncthreads();
  // entity graph in model numbering (pat as scratch)
  for (int c = w.lane; c < w.n_con; c += kWave) {
    int ea, eb;
    row_entities(w, c, ea, eb);
    if (ea != eb) {
      atomicOr((unsigned long long*)&L.pat[2 * ea + (eb >> 6)], 1ull << (eb & 63));
      atomicOr((unsigned long long*)&L.pat[2 * eb + (ea >> 6)], 1ull << (ea & 63));
    }
  }
  __syncthreads();
  unsigned long long adj0[2] = {0ull, 0ull}, adj1[2] = {0ull, 0ull};
  if (w.lane < w.E) { adj0[0] = L.pat[2 * w.lane]; adj0[1] = L.pat[2 * w.lane + 1]; }
  if (w.lane + 64 < w.E) { adj1[0] = L.pat[2 * (w.lane + 64)]; adj1[1] = L.pat[2 * (w.lane + 64) + 1]; }
  wave_min_degree(w, adj0, adj1);
  // factor-space pattern: row vof[e] gets the positions of e's filled neighbours and its own
  for (int i = w.lane; i < 2 * w.E; i += kWave) L.pat[i] = 0ull;
  __syncthreads();
  for (int h = 0; h < 2; h++) {
    const int ie = w.lane + 64 * h;
    if (ie >= w.E || L.nee[ie] <= L.nes[ie]) continue;
    const unsigned long long* a = h ? adj1 : adj0;
    const int vi = L.vof[ie];
    unsigned long long r0 = 0ull, r1 = 0ull;
    if (vi < 64) r0 |= 1ull << vi; else r1 |= 1ull << (vi - 64);
    for (int word = 0; word < 2; word++) {
      unsigned long long t = a[word];
      while (t) {
        int x = word * 64 + __ffsll((long long)t) - 1;
        t &= t - 1;
        int vx = L.vof[x];
        if (vx < 64) r0 |= 1ull << vx; else r1 |= 1ull << (vx - 64);
      }
    }
    L.pat[2 * vi] = r0;
    L.pat[2 * vi + 1] = r1;
  }
  // factor dof ranges by position, then the dof maps
  if (w.lane == 0) {
    int acc = 0;
    for (int v = 0; v < w.E; v++) {
      int ie = L.ord[v];
      L.es[v] = acc;
      acc += L.nee[ie] - L.nes[ie];
      L.ee[v] = acc;
    }
  }
  __syncthreads();
  for (int d = w.lane; d < w.D; d += kWave) {
    int ie = L.nent[d], v = L.vof[ie];
    int p = L.es[v] + (d - L.nes[ie]);
    L.pdof[d] = p;
    L.ent_of_dof[p] = v;
  }
  __syncthreads();
  // entity pair of every row, once (env-major scratch)
  int* rent = s.efc_ent + (size_t)e.b * m.max_rows;
  for (int c = w.lane; c < w.n_con; c += kWave) {
    int ea, eb;
    row_entities(w, c, ea, eb);
    rent[c] = ea | (eb << 16);
  }
  // ordered row lists per entity (CSR): counts, prefix, then a stable placement chunk by chunk
  int* ptr = s.ent_row_ptr + (size_t)e.b * (w.E + 1);
  int* lst = s.ent_row_list + (size_t)e.b * 2 * m.max_rows;
  for (int ie = w.lane; ie < w.E; ie += kWave) L.cursor[ie] = 0;
  __syncthreads();
  for (int c = w.lane; c < w.n_con; c += kWave) {
    int r = rent[c];
    int ea = r & 0xffff, eb = r >> 16;
    atomicAdd((int*)&L.cursor[ea], 1);
    if (eb != ea) atomicAdd((int*)&L.cursor[eb], 1);
  }
  __syncthreads();
  if (w.lane == 0) {
    int acc = 0;
    for (int ie = 0; ie < w.E; ie++) {
      int n = L.cursor[ie];
      ptr[ie] = acc;
      L.cursor[ie] = acc;
      acc += n;
    }
    ptr[w.E] = acc;
  }
  __syncthreads();
  int* run = s.ent_row_run + (size_t)e.b * 2 * m.max_rows;
  const bool runs = wave_dof_lanes(w);
  for (int c0 = 0; c0 < w.n_con; c0 += kWave) {
    int c = c0 + w.lane;
    bool valid = c < w.n_con;
    int r = valid ? rent[c] : -1;
    int ea = valid ? (r & 0xffff) : -1, eb = valid ? (r >> 16) : -2;
    // where each entity's entries sit in the row, for the dof-lane J^T f
    int run_a = -1, run_b = -1;
    if (runs && valid) row_runs(w, c, ea, eb, run_a, run_b);
    int rank_a = 0, rank_b = 0, cnt_a = 0, cnt_b = 0;
    for (int j = 0; j < kWave; j++) {
      int ra = __shfl(ea, j, kWave), rb = __shfl(eb, j, kWave);
      bool ta = valid && (ra == ea || rb == ea);
      bool tb = valid && (ra == eb || rb == eb);
      if (j < w.lane) { rank_a += ta; rank_b += tb; }
      cnt_a += ta;
      cnt_b += tb;
    }
    if (valid) {
      lst[L.cursor[ea] + rank_a] = c;
      if (eb != ea) lst[L.cursor[eb] + rank_b] = c;
      if (runs) {
        run[L.cursor[ea] + rank_a] = run_a;
        if (eb != ea) run[L.cursor[eb] + rank_b] = run_b;
      }
    }
    __syncthreads();
    // the last lane touching an entity in this chunk advances its cursor
    if (valid) {
      if (rank_a == cnt_a - 1) L.cursor[ea] += cnt_a;
      if (eb != ea && rank_b == cnt_b - 1) L.cursor[eb] += cnt_b;
    }
    __syncthreads();
  }
  // level schedules of the pattern: forward (predecessors first) and backward (successors first)
  if (w.lane == 0) {
    lds_i* lvl = L.cursor;  // reuse as level scratch
    int maxl = -1;
    for (int ie = 0; ie < w.E; ie++) {
      if (L.ee[ie] <= L.es[ie]) { lvl[ie] = -1; continue; }
      int lv = 0;
      for (int word = 0; word < 2; word++) {
        unsigned long long t = L.pat[2 * ie + word];
        while (t) {
          int k = word * 64 + __ffsll((long long)t) - 1;
          t &= t - 1;
          if (k >= ie) break;
          if (lvl[k] + 1 > lv) lv = lvl[k] + 1;
        }
      }
      lvl[ie] = lv;
      if (lv > maxl) maxl = lv;
    }
    int nl = maxl + 1;
    for (int l = 0; l <= nl; l++) L.lvl_start[l] = 0;
    for (int ie = 0; ie < w.E; ie++) if (lvl[ie] >= 0) L.lvl_start[lvl[ie] + 1]++;
    for (int l = 0; l < nl; l++) L.lvl_start[l + 1] += L.lvl_start[l];
    for (int l = 0; l < nl; l++) L.rlvl_start[l] = L.lvl_start[l];  // temp cursor
    for (int ie = 0; ie < w.E; ie++) if (lvl[ie] >= 0) L.lvl_order[L.rlvl_start[lvl[ie]]++] = ie;
    L.n_lvl[0] = nl;
    // backward levels
    maxl = -1;
    for (int ie = w.E - 1; ie >= 0; ie--) {
      if (L.ee[ie] <= L.es[ie]) { lvl[ie] = -1; continue; }
      int lv = 0;
      for (int word = 0; word < 2; word++) {
        unsigned long long t = L.pat[2 * ie + word];
        while (t) {
          int k = word * 64 + __ffsll((long long)t) - 1;
          t &= t - 1;
          if (k <= ie) continue;
          if (lvl[k] + 1 > lv) lv = lvl[k] + 1;
        }
      }
      lvl[ie] = lv;
      if (lv > maxl) maxl = lv;
    }
    nl = maxl + 1;
    for (int l = 0; l <= nl; l++) L.rlvl_start[l] = 0;
    for (int ie = 0; ie < w.E; ie++) if (lvl[ie] >= 0) L.rlvl_start[lvl[ie] + 1]++;
    for (int l = 0; l < nl; l++) L.rlvl_start[l + 1] += L.rlvl_start[l];
    // stable fill through a second pass (lvl_start space beyond nl is free scratch only when E is large; use
    // an O(E * nl) scan instead)
    int pos = 0;
    for (int l = 0; l < nl; l++)
      for (int ie = 0; ie < w.E; ie++)
        if (lvl[ie] == l) L.rlvl_order[pos++] = ie;
    L.n_lvl[1] = nl;
    // LDS layout of the Hessian on this pattern, if it fits
    int total = 0, slots = 0;
    bool ok = true;
    for (int ie = 0; ie < w.E; ie++) {
      int n = L.ee[ie] - L.es[ie];
      if (n <= 0) continue;
      L.hoff[ie] = total;
      L.pstart[ie] = slots;
      int acc = 0;
      for (int word = 0; word < 2; word++) {
        unsigned long long t = L.pat[2 * ie + word];
        while (t) {
          int k = word * 64 + __ffsll((long long)t) - 1;
          t &= t - 1;
          if (k >= ie) break;
          if (slots < kWavePreSlots) L.pre[slots] = acc;
          else ok = false;
          slots++;
          acc += L.ee[k] - L.es[k];
        }
      }
      L.below[ie] = acc;
      total += n * acc + n * (n + 1) / 2;
    }
    L.hmode[0] = (ok && total <= w.e.m.wave_h_floats) ? 1 : 0;
  }
  __syncthreads();
}

// ---- Hessian storage ----------------------------------------------------------------------------------------------
// Both stores hand out a pointer p with H(i, k) = p[k - es[ek]] for the dofs k of entity ek <= ent(i) in the pattern
// row of ent(i).
// Entry (i, j) of the factor, with j a dof of entity ek in the pattern row of ei = ent(i), sits at
// row(i, ei) + off(ei, ek) + j: the row part depends on the dof, the offset only on the entity pair;
// blk(i, ei, ek) = row(i, ei) + off(ei, ek) + es[ek] is the start of block (ei, ek) in row i.
struct HGlobal {
  // entries written by one lane and read by another must pass through a fence (global memory, unlike LDS, gives a
  // wavefront no ordering between its lanes' accesses)
  static constexpr bool kFence = true;
  typedef float* ptr;
  const Env& e;
  const lds_i* es;
  __device__ __forceinline__ float* blk(int i, int ei, int ek) const { return &e.H(i, es[ek]); }
  __device__ __forceinline__ float* row(int i, int ei) const { return &e.H(i, 0); }
  __device__ __forceinline__ int off(int ei, int ek) const { return 0; }
  // the wavefront solver's rows are env-major, so an env's packed triangle is contiguous from &H(0, 0)
  __device__ __forceinline__ float* base() const { return &e.H(0, 0); }
  __device__ __forceinline__ int row_off(int i, int ei) const { return i * (i + 1) / 2; }
};

struct HLds {
  static constexpr bool kFence = false;
  typedef lds_f* ptr;
  const WaveLds& L;
  __device__ __forceinline__ int off(int ei, int ek) const {
    if (ek == ei) return L.below[ei] - L.es[ei];
    unsigned long long w0 = L.pat[2 * ei], w1 = L.pat[2 * ei + 1];
    int pc = ek < 64 ? __popcll(w0 & ((1ull << ek) - 1ull)) : __popcll(w0) + __popcll(w1 & ((1ull << (ek - 64)) - 1ull));
    return L.pre[L.pstart[ei] + pc] - L.es[ek];
  }
  __device__ __forceinline__ int row_off(int i, int ei) const {
    int r = i - L.es[ei];
    return L.hoff[ei] + r * L.below[ei] + r * (r + 1) / 2;
  }
  __device__ __forceinline__ lds_f* base() const { return L.hl; }
  __device__ __forceinline__ lds_f* row(int i, int ei) const { return L.hl + row_off(i, ei); }
  __device__ __forceinline__ lds_f* blk(int i, int ei, int ek) const { return row(i, ei) + off(ei, ek) + L.es[ek]; }
};

// ---- single-entity models: rows staged densely --------------------------------------------------------------------
// When every dof belongs to one entity, every row touches only that entity and its row list is every row in
// ascending order. J^T f and the J^T D J part of the Hessian are then summed with one lane per dof (per entry) over
// rows staged densely in LDS, dense_rows at a time, instead of one lane walking every row through global memory.
// Each sum adds the general path's terms in the same row order from +0 (an absent dof contributes +-0, which changes
// no bit of a sum that starts at +0), so both paths give the same bits.

// lane-local: row c into staging slot `lane` (values over the model's dofs, zero elsewhere) with its scale, or a
// zero row when `use` is false
__device__ __forceinline__ void dense_stage_row(const WaveCtx& w, int c, bool use, float scale) {
  const Env& e = w.e;
  lds_f* row = w.dense_buf + w.lane * w.D;
  for (int d = 0; d < w.D; d++) row[d] = 0.f;
  if (use) {
    const int nnz = e.EI(e.s.jac_nnz, c);
    if (nnz <= kStage) {
      RowStage r;
      stage_row(e, c, nnz, r);
#pragma unroll
      for (int k = 0; k < kStage; k++)
        if (k < nnz) row[r.d[k]] = r.v[k];
    } else {
      for (int k = 0; k < nnz; k++) row[e.JI(c, k)] = e.JV(c, k);
    }
  }
  w.dense_buf[w.dense_rows * w.D + w.lane] = use ? scale : 0.f;
}

// J^T f into L.qfrc (lane d owns dof d)
__device__ inline void wave_jtf_dense(const WaveCtx& w) {
  const Env& e = w.e;
  const int D = w.D, R = w.dense_rows;
  const lds_f* sc = w.dense_buf + R * D;
  float acc = 0.f;
  for (int c0 = 0; c0 < w.n_con; c0 += R) {
    const int nr = min(R, w.n_con - c0);
    if (w.lane < nr) {
      const float f = e.EF(e.s.efc_force, c0 + w.lane);
      dense_stage_row(w, c0 + w.lane, f != 0.f, f);
    }
    __syncthreads();
    if (w.lane < D) {
      const lds_f* col = w.dense_buf + w.lane;
      for (int r = 0; r < nr; r++) acc = acc + col[r * D] * sc[r];
    }
    __syncthreads();
  }
  if (w.lane < D) w.L.qfrc[w.lane] = acc;
}

// H = M + J^T D J over the active rows, lanes over the entries of the lower triangle (entry t = i (i + 1) / 2 + j,
// j <= i, for t = lane, lane + 64, ...); each entry's partial sum waits in the entry itself between passes
__device__ __forceinline__ void dense_entry(int t, int& i, int& j) {
  i = (int)((sqrtf(8.f * (float)t + 1.f) - 1.f) * 0.5f);
  while (i * (i + 1) / 2 > t) i--;
  while ((i + 1) * (i + 2) / 2 <= t) i++;
  j = t - i * (i + 1) / 2;
}

template <class HS>
__device__ inline void wave_hessian_dense(const WaveCtx& w, const HS& h) {
  const Env& e = w.e;
  const gsr_state& s = e.s;
  const WaveLds& L = w.L;
  const int D = w.D, R = w.dense_rows, ntri = D * (D + 1) / 2;
  const float eps = e.m.eps;
  const lds_f* sc = w.dense_buf + R * D;
  const int v = L.ent_of_dof[0], ne = L.ord[v], n0 = L.nes[ne], f0 = L.es[v];
  for (int c0 = 0; c0 < w.n_con; c0 += R) {
    const int nr = min(R, w.n_con - c0);
    if (w.lane < nr) {
      const int c = c0 + w.lane;
      dense_stage_row(w, c, (e.EI(s.efc_active, c) & 1) != 0, e.EF(s.efc_D, c));
    }
    __syncthreads();
    for (int t = w.lane; t < ntri; t += kWave) {
      int i, j;
      dense_entry(t, i, j);
      // pair (row's later dof i, earlier dof j), skipped when |J[c, i]| <= eps, as the row walk skips it
      const lds_f* pi = w.dense_buf + i;
      const lds_f* pj = w.dense_buf + j;
      auto q = h.blk(f0 + i, v, v) + j;
      float a = c0 == 0 ? 0.f : *q;
      for (int r = 0; r < nr; r++) {
        const float vi = pi[r * D];
        const float vm = fabsf(vi) > eps ? vi : 0.f;
        a = a + pj[r * D] * vm * sc[r];
      }
      *q = a;
    }
    __syncthreads();
  }
  for (int t = w.lane; t < ntri; t += kWave) {
    int i, j;
    dense_entry(t, i, j);
    auto q = h.blk(f0 + i, v, v) + j;
    *q = (w.n_con > 0 ? *q : 0.f) + w.mm(ne, n0 + i, n0 + j);
  }
}

// M x (entity blocks), one dof per lane, serial inner sum in the reference order
__device__ inline void wave_mass_mul(const WaveCtx& w, const lds_f* x, lds_f* out) {
  const Env& e = w.e;
  for (int i = w.lane; i < w.D; i += kWave) {
    int ie = w.L.nent[i];
    float v = 0.f;
    for (int j = w.L.nes[ie]; j < w.L.nee[ie]; j++) v += w.mm(ie, i, j) * x[j];
    out[i] = v;
  }
}

// func_update_constraint; returns (cost, gauss)

__device__ inline void wave_update_constraint(const WaveCtx& w, double& cost, float& gauss) {
  const Env& e = w.e;
  const gsr_state& s = e.s;
  const WaveLds& L = w.L;
  double ci = 0.0;
  for (int c = w.lane; c < w.n_con; c += kWave) {
    int act_prev = e.EI(s.efc_active, c) & 1;
    int act = 1;
    float floss_force = 0.f;
    float ja = e.EF(s.efc_Jaref, c);
    if (w.ne <= c && c < w.nef) {
      float f = e.EF(s.efc_fl, c), r = e.EF(s.efc_diag, c), rf = r * f;
      bool ln = ja <= -rf, lp = ja >= rf;
      act = !(ln || lp);
      floss_force = (float)ln * f + (float)lp * -f;
      float fc = (float)ln * f * (-0.5f * rf - ja);
      fc += (float)lp * f * (-0.5f * rf + ja);
      ci += (double)fc;
    } else if (w.nef <= c) {
      act = ja < 0.f;
    }
    e.EI(s.efc_active, c) = act | (act_prev << 1);
    float Dc = e.EF(s.efc_D, c);
    e.EF(s.efc_force, c) = floss_force + (-ja * Dc * (float)act);
    ci += (double)(0.5f * (ja * ja * Dc * (float)act));
  }
  __syncthreads();
  // J^T f per dof, rows in ascending order through the entity row list; one lane per entity reads each of its
  // rows once and accumulates into the entity's own dofs (no other lane writes them)
  const int* ptr = s.ent_row_ptr + (size_t)e.b * (w.E + 1);
  const int* lst = s.ent_row_list + (size_t)e.b * 2 * e.m.max_rows;
  const bool split = wave_dof_lanes(w);
  if (w.dense_rows > 0) {
    wave_jtf_dense(w);
  } else if (split) {
    // one lane per dof walks the rows of its dof's entity, two rows per iteration (both rows' loads issued before
    // either is used, the next pair's indices loaded meanwhile), and sums entry(row, d) * force in row order from
    // +0 in a register: the per-entity walk's terms in its order, so the same bits
    const int* run = s.ent_row_run + (size_t)e.b * 2 * e.m.max_rows;
    for (int d = w.lane; d < w.D; d += kWave) {
      const int ie = L.nent[d];
      const int t_end = ptr[ie + 1];
      const int W = e.m.row_width;
      float acc = 0.f;
      // four rows per iteration: the rows' list entries and runs, then their forces and this dof's entries (one
      // load each where the entity's entries form a run), all issued before any is summed (in row order)
      for (int t = ptr[ie]; t < t_end; t += 4) {
        int c[4], rn[4];
        float f[4], v[4];
        bool h[4];
#pragma unroll
        for (int q = 0; q < 4; q++) {
          c[q] = t + q < t_end ? lst[t + q] : 0;
          rn[q] = t + q < t_end ? run[t + q] : 0;
        }
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const int f0 = rn[q] & 0xffff, p0 = (rn[q] >> 16) & 0xff, n0 = (rn[q] >> 24) & 0xff;
          h[q] = t + q < t_end && rn[q] >= 0 && d >= f0 && d < f0 + n0;
          f[q] = t + q < t_end ? e.EF(s.efc_force, c[q]) : 0.f;
          v[q] = h[q] ? e.JV(c[q], p0 + (d - f0)) : 0.f;
        }
#pragma unroll
        for (int q = 0; q < 4; q++) {
          if (t + q < t_end && rn[q] < 0) {
            // the entity's entries are not one run: find d among the row's entries
            const int nnz = e.EI(s.jac_nnz, c[q]);
            RowStage r;
            stage_row(e, c[q], W, r);
#pragma unroll
            for (int k = 0; k < kStage; k++)
              if (k < nnz && r.d[k] == d) { v[q] = r.v[k]; h[q] = true; }
          }
          // a zero force adds +-0 to a sum that starts at +0 and so never holds -0: skipping a row changes no bit
          if (h[q] && f[q] != 0.f) acc = acc + v[q] * f[q];
        }
      }
      L.qfrc[d] = acc;
    }
  } else {
    for (int d = w.lane; d < w.D; d += kWave) L.qfrc[d] = 0.f;
    __syncthreads();
  }
  for (int ie = w.lane; ie < w.E && w.dense_rows == 0 && !split; ie += kWave) {
    const int s0 = L.nes[ie], s1 = L.nee[ie];
    if (s1 <= s0) continue;
    // two rows per iteration: both rows' loads are issued before either is accumulated (in row order); the next
    // pair's row indices are loaded meanwhile, and every row slot is loaded at once with the force and entry count
    // (slots past the entry count are ignored), so a pair costs one round trip to memory
    const int t_end = ptr[ie + 1];
    const int W = e.m.row_width;
    int ca_n = ptr[ie] < t_end ? lst[ptr[ie]] : 0, cb_n = ptr[ie] + 1 < t_end ? lst[ptr[ie] + 1] : 0;
    for (int t = ptr[ie]; t < t_end; t += 2) {
      const bool two = t + 1 < t_end;
      const int ca = ca_n, cb = two ? cb_n : ca;
      if (t + 2 < t_end) ca_n = lst[t + 2];
      if (t + 3 < t_end) cb_n = lst[t + 3];
      // a zero force adds +-0 to sums that start at +0 and so never hold -0: skipping a row changes no bit
      const float fa = e.EF(s.efc_force, ca), fb = two ? e.EF(s.efc_force, cb) : 0.f;
      const int na = e.EI(s.jac_nnz, ca), nb = e.EI(s.jac_nnz, cb);
      if (W <= kStage) {
        RowStage ra, rb;
        stage_row(e, ca, W, ra);
        stage_row(e, cb, W, rb);
        // a row's entries are distinct dofs: all are read before any is written back
        float old_v[kStage];
#pragma unroll
        for (int k = 0; k < kStage; k++) {
          const int d = ra.d[k];
          old_v[k] = (k < na && fa != 0.f && d >= s0 && d < s1) ? L.qfrc[d] : 0.f;
        }
#pragma unroll
        for (int k = 0; k < kStage; k++) {
          const int d = ra.d[k];
          if (k < na && fa != 0.f && d >= s0 && d < s1) L.qfrc[d] = old_v[k] + ra.v[k] * fa;
        }
#pragma unroll
        for (int k = 0; k < kStage; k++) {
          const int d = rb.d[k];
          old_v[k] = (k < nb && fb != 0.f && d >= s0 && d < s1) ? L.qfrc[d] : 0.f;
        }
#pragma unroll
        for (int k = 0; k < kStage; k++) {
          const int d = rb.d[k];
          if (k < nb && fb != 0.f && d >= s0 && d < s1) L.qfrc[d] = old_v[k] + rb.v[k] * fb;
        }
      } else {
        for (int q = 0; q < (two ? 2 : 1); q++) {
          const int c = q ? cb : ca, nnz = q ? nb : na;
          const float f = q ? fb : fa;
          if (f == 0.f) continue;
          for (int k = 0; k < nnz; k++) {
            int d = e.JI(c, k);
            if (d >= s0 && d < s1) L.qfrc[d] = L.qfrc[d] + e.JV(c, k) * f;
          }
        }
      }
    }
  }
  double gi = 0.0;
  for (int d = w.lane; d < w.D; d += kWave) {
    float v = 0.5f * (L.Ma[d] - e.F(s.force, d)) * (L.qacc[d] - e.F(s.acc_smooth, d));
    gi += (double)v;
  }
  __syncthreads();
  const double gd = wave_sum_d(gi);
  gauss = (float)gd;
  cost = wave_sum_d(ci) + gd;
  __syncthreads();
}

// candidate rows j > i of column i inside the pattern, enumerated by index t: returns -1 past the end
__device__ __forceinline__ int wave_col_candidate(const WaveCtx& w, int i, int ei, int t) {
  const WaveLds& L = w.L;
  int rest = L.ee[ei] - i - 1;
  if (t < rest) return i + 1 + t;
  t -= rest;
  for (int word = (ei + 1) >> 6; word < 2; word++) {
    unsigned long long bits = L.pat[2 * ei + word];
    int lo = ei + 1 - word * 64;
    if (lo > 0) bits &= (lo >= 64) ? 0ull : (~0ull << lo);
    while (bits) {
      int ej = word * 64 + __ffsll((long long)bits) - 1;
      bits &= bits - 1;
      int n = L.ee[ej] - L.es[ej];
      if (t < n) return L.es[ej] + t;
      t -= n;
    }
  }
  return -1;
}

// The n entities of one level are independent; the wavefront is cut into groups of G lanes (the largest power of two
// with G * n <= 64, at least 1), group q taking entities q, q + 64 / G, ... of the level. Inside a group, the terms
// of every dot product are dealt round-robin to its lanes and summed by a butterfly, so all lanes of the group hold
// the same value; each entry is written by the group's lane 0. A wavefront's LDS accesses complete in order, so an
// entry written by one lane is seen by the reads any lane issues after it. Terms are dealt by a running term index
// t0 (lane r takes the terms t with t = r mod G).
struct LaneGroup {
  int G, q, r, ng;
};
__device__ __forceinline__ LaneGroup lane_group(int n, int lane) {
  int G = kWave;
  while (G > 1 && G * n > kWave) G >>= 1;
  return LaneGroup{G, lane / G, lane & (G - 1), kWave / G};
}
__device__ __forceinline__ float group_sum(float v, int G) {
  for (int o = G >> 1; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

// ---- single-entity models: the factor as one dense packed triangle ------------------------------------------------
// With every dof in one entity and H in LDS, H is the packed lower triangle at the start of the LDS store, H(i, j) at
// hl[i (i + 1) / 2 + j] (factor and model numbering agree). The factorisation, the substitutions and the rank-1
// updates then run with one lane per row and every sum serial over its index, the operation order of
// func_nt_chol_factor / func_nt_chol_solve / func_nt_hessian_incremental (constraint_solver_decomp.py), instead of
// the chunked per-entity forms below with their pattern lookups.
__device__ __forceinline__ lds_f* dense_row(const WaveLds& L, int i) { return L.hl + i * (i + 1) / 2; }

__device__ inline void wave_chol_factor_dense(const WaveCtx& w) {
  const float EPS = w.e.m.eps;
  const int D = w.D;
  for (int i = 0; i < D; i++) {
    lds_f* ri = dense_row(w.L, i);
    float tmp = ri[i];
#pragma unroll 8
    for (int j = 0; j < i; j++) {
      const float h = ri[j];
      tmp -= h * h;
    }
    const float hii = sqrtf(fmaxf(tmp, EPS));
    const float inv = 1.f / hii;
    // rows j > i, one per lane
    for (int j = i + 1 + w.lane; j < D; j += kWave) {
      lds_f* rj = dense_row(w.L, j);
      float dt = 0.f;
#pragma unroll 8
      for (int k = 0; k < i; k++) dt += rj[k] * ri[k];
      rj[i] = (rj[i] - dt) * inv;
    }
    if (w.lane == 0) ri[i] = hii;
    __syncthreads();
  }
}

// x = H^-1 x in place, lane d holding x[d] in a register: forward substitution (row i subtracts L(i, j) y(j) for j
// ascending as y(j) becomes final: the serial order of func_nt_chol_solve), then backward substitution column by
// column (row i subtracts L(j, i) x(j) for j descending as x(j) becomes final)
__device__ inline void wave_chol_solve_dense(const WaveCtx& w, lds_f* x) {
  const int D = w.D, lane = w.lane;
  __syncthreads();
  float a = lane < D ? x[lane] : 0.f;
  const lds_f* rl = dense_row(w.L, lane < D ? lane : 0);
  for (int j = 0; j < D; j++) {
    const float yj = __shfl(a, j, kWave) / dense_row(w.L, j)[j];
    if (lane == j) a = yj;
    else if (lane > j && lane < D) a = a - rl[j] * yj;
  }
  for (int i = D - 1; i >= 0; i--) {
    const lds_f* ri = dense_row(w.L, i);
    const float xi = __shfl(a, i, kWave) / ri[i];
    if (lane == i) a = xi;
    else if (lane < i) a = a - ri[lane] * xi;
  }
  if (lane < D) x[lane] = a;
  __syncthreads();
}

// rank-1 update / downdate of the factor with row c; returns false on rank loss (the update is then incomplete and
// the caller rebuilds the factor)
__device__ inline bool wave_rank1_dense(const WaveCtx& w, int c, float sgn) {
  const Env& e = w.e;
  const gsr_state& s = e.s;
  const float eps = e.m.eps;
  const int D = w.D, lane = w.lane;
  const float sq = sqrtf(e.EF(s.efc_D, c));
  lds_f* x = w.L.vec;
  for (int d = lane; d < D; d += kWave) x[d] = 0.f;
  __syncthreads();
  const int nnz = e.EI(s.jac_nnz, c);
  for (int k = lane; k < nnz; k += kWave) x[e.JI(c, k)] = e.JV(c, k) * sq;
  __syncthreads();
  float xi = lane < D ? x[lane] : 0.f;
  lds_f* rl = dense_row(w.L, lane < D ? lane : 0);
  bool ok = true;
  for (int k = 0; k < D; k++) {
    const float vk = __shfl(xi, k, kWave);
    if (!(fabsf(vk) > eps)) continue;
    lds_f* dk = dense_row(w.L, k) + k;
    const float Lkk = *dk;
    float tmp = Lkk * Lkk + vk * vk * sgn;
    if (tmp < eps) {
      tmp = eps;
      ok = false;
    }
    const float r = sqrtf(tmp), cc = r / Lkk, cinv = 1.f / cc, sv = vk / Lkk;
    if (lane > k && lane < D) {
      const float hv = (rl[k] + sv * xi * sgn) * cinv;
      rl[k] = hv;
      xi = xi * cc - sv * hv;
    }
    __syncthreads();
    if (lane == 0) *dk = r;
    __syncthreads();
  }
  return ok;
}

// Left-looking Cholesky over the pattern (func_nt_chol_factor's entries), level by level, lane groups per entity,
// kChunk dofs of an entity at a time:
//   1. the chunk's diagonal block less the products of its rows over the earlier columns (terms dealt round-robin
//      to the group, kChunk (kChunk + 1) / 2 sums reduced together);
//   2. the small Cholesky of that block, by every lane of the group alike (registers);
//   3. every row j below the chunk (later own dofs, then the pattern entities after it), one per lane of the group:
//      its kChunk entries less the products over the earlier columns the two rows share, then a forward solve with
//      the chunk's factor.
// Each row's earlier columns are read once for the whole chunk instead of once per column.
constexpr int kChunk = 6;  // dofs of an entity the factorisation and substitutions take at a time
constexpr int kChunkTri = kChunk * (kChunk + 1) / 2;
__device__ __forceinline__ constexpr int tri(int r, int q) { return r * (r + 1) / 2 + q; }

template <class HS>
__device__ __forceinline__ void wave_chol_factor(const WaveCtx& w, const HS& h) {
  const float EPS = w.e.m.eps;
  const WaveLds& L = w.L;
  int nl = L.n_lvl[0];
  for (int lv = 0; lv < nl; lv++) {
    const int i0 = L.lvl_start[lv], i1 = L.lvl_start[lv + 1];
    const LaneGroup g = lane_group(i1 - i0, w.lane);
    for (int idx = i0 + g.q; idx < i1; idx += g.ng) {
      const int ei = L.lvl_order[idx];
      const int es = L.es[ei], ee = L.ee[ei];
      const int od = h.off(ei, ei);
      for (int c0 = es; c0 < ee; c0 += kChunk) {
        const int nc = min(kChunk, ee - c0), c1 = c0 + nc;
        typename HS::ptr rp[kChunk];
#pragma unroll
        for (int r = 0; r < kChunk; r++) rp[r] = h.row(r < nc ? c0 + r : c0, ei);
        // 1. products of the chunk's rows over the columns before c0
        float dacc[kChunkTri];
#pragma unroll
        for (int t = 0; t < kChunkTri; t++) dacc[t] = 0.f;
        int t0 = 0;
        for (int word = 0; word < 2; word++) {
          unsigned long long bits = L.pat[2 * ei + word];
          while (bits) {
            const int ek = word * 64 + __ffsll((long long)bits) - 1;
            bits &= bits - 1;
            const int k0 = L.es[ek];
            if (k0 >= c0) break;
            const int k1 = min(L.ee[ek], c0);
            const int off = h.off(ei, ek);
            for (int k = k0 + ((g.r - t0) & (g.G - 1)); k < k1; k += g.G) {
              float v[kChunk];
#pragma unroll
              for (int r = 0; r < kChunk; r++) v[r] = r < nc ? rp[r][off + k] : 0.f;
#pragma unroll
              for (int r = 0; r < kChunk; r++)
#pragma unroll
                for (int q = 0; q <= r; q++) dacc[tri(r, q)] += v[r] * v[q];
            }
            t0 += k1 - k0;
          }
        }
#pragma unroll
        for (int t = 0; t < kChunkTri; t++) dacc[t] = group_sum(dacc[t], g.G);
        // 2. the chunk's own factor
        float lc[kChunkTri];
#pragma unroll
        for (int r = 0; r < kChunk; r++) {
#pragma unroll
          for (int q = 0; q <= r; q++) {
            float v = 0.f;
            if (r < nc) {
              v = rp[r][od + c0 + q] - dacc[tri(r, q)];
#pragma unroll
              for (int p2 = 0; p2 < q; p2++) v -= lc[tri(r, p2)] * lc[tri(q, p2)];
              v = (q == r) ? sqrtf(fmaxf(v, EPS)) : v / lc[tri(q, q)];
            }
            lc[tri(r, q)] = v;
          }
        }
        // 3. the rows below the chunk, one per lane of the group
        int u0 = 0;
        auto below_row = [&](int j, int ej) {
          const auto pj = h.row(j, ej);
          float acc[kChunk];
#pragma unroll
          for (int q = 0; q < kChunk; q++) acc[q] = 0.f;
          const unsigned long long m0 = L.pat[2 * ej] & L.pat[2 * ei], m1 = L.pat[2 * ej + 1] & L.pat[2 * ei + 1];
          for (int word = 0; word < 2; word++) {
            unsigned long long t = word ? m1 : m0;
            while (t) {
              const int ek = word * 64 + __ffsll((long long)t) - 1;
              t &= t - 1;
              const int k0 = L.es[ek];
              if (k0 >= c0) break;
              const int k1 = min(L.ee[ek], c0);
              const int offj = h.off(ej, ek), offi = h.off(ei, ek);
              for (int k = k0; k < k1; k++) {
                const float lj = pj[offj + k];
#pragma unroll
                for (int q = 0; q < kChunk; q++)
                  if (q < nc) acc[q] += lj * rp[q][offi + k];
              }
            }
          }
          const int oji = h.off(ej, ei) + c0;
          float y[kChunk];
#pragma unroll
          for (int q = 0; q < kChunk; q++) {
            if (q < nc) {
              float v = pj[oji + q] - acc[q];
#pragma unroll
              for (int p2 = 0; p2 < q; p2++) v -= y[p2] * lc[tri(q, p2)];
              y[q] = v / lc[tri(q, q)];
              pj[oji + q] = y[q];
            }
          }
        };
        for (int j = c1 + (g.r & (g.G - 1)); j < ee; j += g.G) below_row(j, ei);
        u0 = ee - c1;
        for (int word = (ei + 1) >> 6; word < 2; word++) {
          unsigned long long bits = L.pat[2 * ei + word];
          const int lo = ei + 1 - word * 64;
          if (lo > 0) bits &= (lo >= 64) ? 0ull : (~0ull << lo);
          while (bits) {
            const int ej = word * 64 + __ffsll((long long)bits) - 1;
            bits &= bits - 1;
            const int k0 = L.es[ej], k1 = L.ee[ej];
            for (int j = k0 + ((g.r - u0) & (g.G - 1)); j < k1; j += g.G) below_row(j, ej);
            u0 += k1 - k0;
          }
        }
        if (g.r == 0) {
#pragma unroll
          for (int r = 0; r < kChunk; r++)
#pragma unroll
            for (int q = 0; q <= r; q++)
              if (r < nc) rp[r][od + c0 + q] = lc[tri(r, q)];
        }
        if (HS::kFence) __threadfence_block();
      }
    }
    __syncthreads();
  }
}

// H = M + sum_c act D J^T J on the pattern (func_nt_hessian_direct), one lane per entity: the lane walks its
// entity's rows in ascending order once, so every entry still receives its terms in row order; then factor.
// J^T D J + M of a multi-entity model by dof lanes: the lane of factor dof i zeroes row i on the pattern, walks the
// rows of i's entity in ascending order and adds, for each active row holding i, the row's pairs (i, j <= i) in
// factor order, then adds the mass matrix. Each entry thus gets the per-entity walk's terms in the same order (the
// pair of row positions (hi, lo) contributes v[lo] v[hi] D when |v[hi]| > eps), so the bits do not change; the lanes
// of one entity load the same row addresses together.
template <class HS>
__device__ __forceinline__ void wave_hessian_by_dof(const WaveCtx& w, const HS& h, const int* ptr, const int* lst) {
  const Env& e = w.e;
  const gsr_model& m = e.m;
  const gsr_state& s = e.s;
  const WaveLds& L = w.L;
  const auto hb = h.base();
  for (int i = w.lane; i < w.D; i += kWave) {
    const int ei = L.ent_of_dof[i];
    const int s0 = L.es[ei];
    // zero row i on the pattern
    for (int word = 0; word < 2; word++) {
      unsigned long long t = L.pat[2 * ei + word];
      while (t) {
        int ek = word * 64 + __ffsll((long long)t) - 1;
        t &= t - 1;
        int k0 = L.es[ek], k1 = L.ee[ek];
        if (k0 > i) break;
        if (k1 > i + 1) k1 = i + 1;
        auto pk = h.blk(i, ei, ek);
        for (int k = 0; k < k1 - k0; k++) pk[k] = 0.f;
      }
    }
    const int ne = L.ord[ei];
    const int t_end = ptr[ne + 1];
    const int roi = h.row_off(i, ei), off_own = h.off(ei, ei);
    int c_next = ptr[ne] < t_end ? lst[ptr[ne]] : 0;
    for (int t = ptr[ne]; t < t_end; t++) {
      const int c = c_next;
      if (t + 1 < t_end) c_next = lst[t + 1];
      const int active = e.EI(s.efc_active, c) & 1;
      const int nnz = e.EI(s.jac_nnz, c);
      const float Dc = e.EF(s.efc_D, c), act = 1.f;
      RowStage r;
      stage_row(e, c, m.row_width, r);
      if (!active) continue;
      const int ent = s.efc_ent[(size_t)e.b * m.max_rows + c];
      const int va = L.vof[ent & 0xffff], vb = L.vof[ent >> 16];
      const int eo = va == ei ? vb : va;
      const int off_oth = eo < ei ? h.off(ei, eo) : 0;
      int pd[kStage], kp = -1;
#pragma unroll
      for (int k = 0; k < kStage; k++) {
        pd[k] = k < nnz ? L.pdof[r.d[k]] : 0x7fffffff;
        if (pd[k] == i) kp = k;
      }
      if (kp < 0) continue;
      float vp = 0.f;
#pragma unroll
      for (int k = 0; k < kStage; k++)
        if (k == kp) vp = r.v[k];
      // the entries (i, pd[k]) with pd[k] <= i are distinct: all are read before any is written back
      float old_v[kStage];
      int co[kStage];
      bool use[kStage];
#pragma unroll
      for (int k = 0; k < kStage; k++) {
        use[k] = pd[k] <= i && fabsf(k <= kp ? vp : r.v[k]) > m.eps;
        co[k] = use[k] ? roi + pd[k] + (pd[k] >= s0 ? off_own : off_oth) : 0;
        old_v[k] = use[k] ? hb[co[k]] : 0.f;
      }
#pragma unroll
      for (int k = 0; k < kStage; k++)
        if (use[k]) hb[co[k]] = old_v[k] + (k <= kp ? r.v[k] * vp : vp * r.v[k]) * Dc * act;
    }
    const int n0 = L.nes[ne];
    auto pi = h.blk(i, ei, ei);
    for (int j = s0; j <= i; j++) pi[j - s0] = pi[j - s0] + w.mm(ne, n0 + (i - s0), n0 + (j - s0));
  }
}

template <class HS>
__device__ __forceinline__ void wave_hessian_direct(const WaveCtx& w, const HS& h) {
  const Env& e = w.e;
  const gsr_model& m = e.m;
  const gsr_state& s = e.s;
  const WaveLds& L = w.L;
  const int* ptr = s.ent_row_ptr + (size_t)e.b * (w.E + 1);
  const int* lst = s.ent_row_list + (size_t)e.b * 2 * m.max_rows;
  if (w.dense_rows > 0) {
    wave_hessian_dense(w, h);
    __syncthreads();
    if (!HS::kFence && w.dense_h) wave_chol_factor_dense(w);
    else wave_chol_factor(w, h);
    return;
  }
  if (wave_dof_lanes(w)) {
    wave_hessian_by_dof(w, h, ptr, lst);
    __syncthreads();
    if (!HS::kFence && w.dense_h) wave_chol_factor_dense(w);
    else wave_chol_factor(w, h);
    return;
  }
  for (int ei = w.lane; ei < w.E; ei += kWave) {
    const int s0 = L.es[ei], s1 = L.ee[ei];
    if (s1 <= s0) continue;
    // zero the entity's rows on the pattern
    for (int i = s0; i < s1; i++) {
      for (int word = 0; word < 2; word++) {
        unsigned long long t = L.pat[2 * ei + word];
        while (t) {
          int ek = word * 64 + __ffsll((long long)t) - 1;
          t &= t - 1;
          int k0 = L.es[ek], k1 = L.ee[ek];
          if (k0 > i) break;
          if (k1 > i + 1) k1 = i + 1;
          auto pk = h.blk(i, ei, ek);
          for (int k = 0; k < k1 - k0; k++) pk[k] = 0.f;
        }
      }
    }
    // the entity's rows in ascending order; each unordered dof pair of a row once (k2 <= k1 in the row's model
    // order, as the reference walks it), stored at (max, min) of its factor dofs by the lane owning the max
    const int ne = L.ord[ei];
    const int t_end = ptr[ne + 1];
    int c_next = ptr[ne] < t_end ? lst[ptr[ne]] : 0;
    for (int t = ptr[ne]; t < t_end; t++) {
      // the next row's index is loaded while this row is summed; the row's slots are loaded at once with its
      // activity, entry count and weight (slots past the entry count are ignored)
      const int c = c_next;
      if (t + 1 < t_end) c_next = lst[t + 1];
      const int active = e.EI(s.efc_active, c) & 1;
      const int nnz = e.EI(s.jac_nnz, c);
      const float Dc = e.EF(s.efc_D, c), act = 1.f;
      if (m.row_width <= kStage) {
        RowStage r;
        stage_row(e, c, m.row_width, r);
        // an inactive row adds +-0 to entries that start at +0 and so never hold -0: skipping it changes no bit
        if (!active) continue;
        // the row's other entity (by position); its dofs are columns of this entity's rows only when it comes first.
        // The pair was stored by the structure pass (one load, beside the row's own, instead of a chain through nnz)
        const int ent = s.efc_ent[(size_t)e.b * m.max_rows + c];
        const int na = ent & 0xffff, nb = ent >> 16;
        const int va = L.vof[na], vb = L.vof[nb];
        const int eo = va == ei ? vb : va;
        const int off_own = h.off(ei, ei), off_oth = eo < ei ? h.off(ei, eo) : 0;
        // per entry: factor dof, whether it is this entity's (bit k of own), its column position, its row offset
        int pd[kStage], co[kStage], ro[kStage];
        unsigned own = 0u;
#pragma unroll
        for (int k = 0; k < kStage; k++) {
          pd[k] = k < nnz ? L.pdof[r.d[k]] : s0;
          const bool o = pd[k] >= s0 && pd[k] < s1;
          own |= (unsigned)o << k;
          co[k] = pd[k] + (o ? off_own : off_oth);
          ro[k] = h.row_off(o ? pd[k] : s0, ei);
        }
        const auto hb = h.base();
#pragma unroll
        for (int k1 = 0; k1 < kStage; k1++) {
          const float v1 = r.v[k1];
          if (k1 < nnz && fabsf(v1) > m.eps) {
            // the pairs (k1, k2 <= k1) are distinct entries: all of them are read before any is written back, so
            // the reads overlap instead of each waiting for the previous entry's write
            float old_v[kStage];
#pragma unroll
            for (int k2 = 0; k2 <= k1; k2++) {
              // entry (max, min) of the two factor dofs, kept by the lane whose entity holds the max
              const bool first = pd[k1] >= pd[k2];
              old_v[k2] = ((own >> (first ? k1 : k2)) & 1u) ? hb[first ? ro[k1] + co[k2] : ro[k2] + co[k1]] : 0.f;
            }
#pragma unroll
            for (int k2 = 0; k2 <= k1; k2++) {
              const bool first = pd[k1] >= pd[k2];
              if ((own >> (first ? k1 : k2)) & 1u)
                hb[first ? ro[k1] + co[k2] : ro[k2] + co[k1]] = old_v[k2] + r.v[k2] * v1 * Dc * act;
            }
          }
        }
        continue;
      }
      if (!active) continue;
      for (int k1 = 0; k1 < nnz; k1++) {
        float v1 = e.JV(c, k1);
        if (!(fabsf(v1) > m.eps)) continue;
        const int p1 = L.pdof[e.JI(c, k1)];
        for (int k2 = 0; k2 <= k1; k2++) {
          const int p2 = L.pdof[e.JI(c, k2)];
          const int pi = p1 >= p2 ? p1 : p2, pj = p1 >= p2 ? p2 : p1;
          if (pi < s0 || pi >= s1) continue;
          const int ej = L.ent_of_dof[pj];
          auto q = h.blk(pi, ei, ej) + (pj - L.es[ej]);
          *q = *q + e.JV(c, k2) * v1 * Dc * act;
        }
      }
    }
    const int n0 = L.nes[ne];
    for (int i = s0; i < s1; i++) {
      auto pi = h.blk(i, ei, ei);
      for (int j = s0; j <= i; j++) pi[j - s0] = pi[j - s0] + w.mm(ne, n0 + (i - s0), n0 + (j - s0));
    }
  }
  __syncthreads();
  if (!HS::kFence && w.dense_h) wave_chol_factor_dense(w);
  else wave_chol_factor(w, h);
}

// rank-1 update / downdate of the factor with row c (func_nt_hessian_incremental); returns false on rank loss
template <class HS>
__device__ __forceinline__ bool wave_rank1(const WaveCtx& w, const HS& h, int c, float sgn) {
  const Env& e = w.e;
  const gsr_model& m = e.m;
  const gsr_state& s = e.s;
  const WaveLds& L = w.L;
  float sq = sqrtf(e.EF(s.efc_D, c));
  int ea, eb;
  row_entities(w, c, ea, eb);
  ea = L.vof[ea];
  eb = L.vof[eb];
  // vec (factor space) is zero outside the entities reachable from ea / eb through the pattern
  for (int d = w.lane; d < w.D; d += kWave) L.vec[d] = 0.f;
  __syncthreads();
  int nnz = e.EI(s.jac_nnz, c);
  if (w.lane < nnz) L.vec[L.pdof[e.JI(c, w.lane)]] = e.JV(c, w.lane) * sq;
  __syncthreads();
  bool rank_ok = true;
  // Entities whose vec entries can be nonzero: the row's two, then every pattern entity below one that had a nonzero
  // entry. Popping the lowest one each time visits them in ascending order, like a sweep over all entities; the
  // entities skipped hold exact zeros, which that sweep skips too.
  unsigned long long act0 = 0ull, act1 = 0ull;
  if (ea < 64) act0 |= 1ull << ea; else act1 |= 1ull << (ea - 64);
  if (eb < 64) act0 |= 1ull << eb; else act1 |= 1ull << (eb - 64);
  for (;;) {
    int ek;
    if (act0) { ek = __ffsll((long long)act0) - 1; act0 &= act0 - 1; }
    else if (act1) { ek = 64 + __ffsll((long long)act1) - 1; act1 &= act1 - 1; }
    else break;
    const int k0 = L.es[ek], k1 = L.ee[ek];
    // this lane's column entry among the pattern entities below ek: the lane-th dof of their concatenation
    unsigned long long b0 = ek + 1 >= 64 ? 0ull : (L.pat[2 * ek] & (~0ull << (ek + 1)));
    unsigned long long b1 = L.pat[2 * ek + 1] & (ek + 1 <= 64 ? ~0ull : (ek + 1 >= 128 ? 0ull : (~0ull << (ek + 1 - 64))));
    int n_below = 0, j_below = -1;
    for (int word = 0; word < 2; word++) {
      unsigned long long t = word ? b1 : b0;
      while (t) {
        int ej = word * 64 + __ffsll((long long)t) - 1;
        t &= t - 1;
        int n = L.ee[ej] - L.es[ej];
        if (w.lane >= n_below && w.lane < n_below + n) j_below = L.es[ej] + (w.lane - n_below);
        n_below += n;
      }
    }
    const bool fast = (k1 - k0) + n_below <= kWave;
    bool any = false;
    for (int k = k0; k < k1; k++) {
      float vk = L.vec[k];
      if (!(fabsf(vk) > m.eps)) continue;
      any = true;
      auto dk = h.blk(k, ek, ek) + (k - k0);
      float Lkk = *dk;
      float tmp = Lkk * Lkk + vk * vk * sgn;
      if (tmp < m.eps) {
        tmp = m.eps;
        rank_ok = false;
      }
      float r = sqrtf(tmp), cc = r / Lkk, cinv = 1.f / cc, sv = vk / Lkk;
      __syncthreads();
      if (w.lane == 0) *dk = r;
      if (fast) {
        const int rest = k1 - k - 1;
        const int jb = __shfl(j_below, max(w.lane - rest, 0), kWave);
        const int i = w.lane < rest ? k + 1 + w.lane : (w.lane < rest + n_below ? jb : -1);
        if (i >= 0) {
          auto pi = h.blk(i, L.ent_of_dof[i], ek) + (k - k0);
          float hv = (*pi + sv * L.vec[i] * sgn) * cinv;
          *pi = hv;
          L.vec[i] = L.vec[i] * cc - sv * hv;
        }
      } else {
        for (int t = w.lane;; t += kWave) {
          int i = wave_col_candidate(w, k, ek, t);
          if (i < 0) break;
          auto pi = h.blk(i, L.ent_of_dof[i], ek) + (k - k0);
          float hv = (*pi + sv * L.vec[i] * sgn) * cinv;
          *pi = hv;
          L.vec[i] = L.vec[i] * cc - sv * hv;
        }
      }
      __syncthreads();
    }
    if (any) {
      act0 |= b0;
      act1 |= b1;
    }
  }
  return rank_ok;
}

// x = H^-1 x in place, x in factor space, through the factor (func_nt_chol_solve): forward over forward levels,
// backward over backward levels, lane groups per entity. An entity's dofs are taken kChunk at a time: the terms
// from outside the chunk (earlier entities and dofs going forward, later ones going back) are shared by every row of
// the chunk, so the group deals them out once and reduces kChunk sums together; the small triangle inside the chunk
// is then solved by every lane of the group alike. x[i] still holds its right-hand side when it is read, and every
// x[j] subtracted from it is final.
template <class HS>
__device__ __forceinline__ void wave_chol_solve(const WaveCtx& w, const HS& h, lds_f* x) {
  const WaveLds& L = w.L;
  __syncthreads();
  int nl = L.n_lvl[0];
  for (int lv = 0; lv < nl; lv++) {
    const int i0 = L.lvl_start[lv], i1 = L.lvl_start[lv + 1];
    const LaneGroup g = lane_group(i1 - i0, w.lane);
    for (int idx = i0 + g.q; idx < i1; idx += g.ng) {
      const int ei = L.lvl_order[idx];
      const int es = L.es[ei], ee = L.ee[ei];
      for (int c0 = es; c0 < ee; c0 += kChunk) {
        const int nc = min(kChunk, ee - c0);
        typename HS::ptr rp[kChunk];
#pragma unroll
        for (int r = 0; r < kChunk; r++) rp[r] = h.row(r < nc ? c0 + r : c0, ei);
        // sum_{k < c0 in the row pattern} L[c0 + r][k] x[k] for every row r of the chunk
        float acc[kChunk];
#pragma unroll
        for (int r = 0; r < kChunk; r++) acc[r] = 0.f;
        int t0 = 0;
        for (int word = 0; word < 2; word++) {
          unsigned long long t = L.pat[2 * ei + word];
          while (t) {
            const int ek = word * 64 + __ffsll((long long)t) - 1;
            t &= t - 1;
            const int k0 = L.es[ek];
            if (k0 >= c0) break;
            const int k1 = min(L.ee[ek], c0);
            const int off = h.off(ei, ek);
            for (int j = k0 + ((g.r - t0) & (g.G - 1)); j < k1; j += g.G) {
              const float xj = x[j];
#pragma unroll
              for (int r = 0; r < kChunk; r++) acc[r] += rp[r][off + j] * xj;
            }
            t0 += k1 - k0;
          }
        }
#pragma unroll
        for (int r = 0; r < kChunk; r++) acc[r] = group_sum(acc[r], g.G);
        // the chunk's own triangle
        const int od = h.off(ei, ei);
        float y[kChunk];
#pragma unroll
        for (int r = 0; r < kChunk; r++) {
          if (r < nc) {
            float v = x[c0 + r] - acc[r];
#pragma unroll
            for (int q = 0; q < r; q++) v -= rp[r][od + c0 + q] * y[q];
            y[r] = v / rp[r][od + c0 + r];
          }
        }
        if (g.r == 0) {
#pragma unroll
          for (int r = 0; r < kChunk; r++)
            if (r < nc) x[c0 + r] = y[r];
        }
      }
    }
    __syncthreads();
  }
  nl = L.n_lvl[1];
  for (int lv = 0; lv < nl; lv++) {
    const int i0 = L.rlvl_start[lv], i1 = L.rlvl_start[lv + 1];
    const LaneGroup g = lane_group(i1 - i0, w.lane);
    for (int idx = i0 + g.q; idx < i1; idx += g.ng) {
      const int ei = L.rlvl_order[idx];
      const int es = L.es[ei], ee = L.ee[ei];
      // chunks from the last one down; the last chunk holds the remainder
      const int n_ch = (ee - es + kChunk - 1) / kChunk;
      for (int ch = n_ch - 1; ch >= 0; ch--) {
        const int c0 = es + ch * kChunk, nc = min(kChunk, ee - c0), c1 = c0 + nc;
        // sum_{j >= c1 in column c0 + r's pattern} L[j][c0 + r] x[j] for every column r of the chunk: later own
        // dofs, then the dofs of the pattern entities after ei
        float acc[kChunk];
#pragma unroll
        for (int r = 0; r < kChunk; r++) acc[r] = 0.f;
        int t0 = 0;
        {
          const int off = h.off(ei, ei);
          for (int j = c1 + (g.r & (g.G - 1)); j < ee; j += g.G) {
            const float xj = x[j];
            const auto pj = h.row(j, ei) + off + c0;
#pragma unroll
            for (int r = 0; r < kChunk; r++)
              if (r < nc) acc[r] += pj[r] * xj;
          }
          t0 = ee - c1;
        }
        for (int word = (ei + 1) >> 6; word < 2; word++) {
          unsigned long long bits = L.pat[2 * ei + word];
          const int lo = ei + 1 - word * 64;
          if (lo > 0) bits &= (lo >= 64) ? 0ull : (~0ull << lo);
          while (bits) {
            const int ej = word * 64 + __ffsll((long long)bits) - 1;
            bits &= bits - 1;
            const int k0 = L.es[ej], k1 = L.ee[ej];
            const int off = h.off(ej, ei);
            for (int j = k0 + ((g.r - t0) & (g.G - 1)); j < k1; j += g.G) {
              const float xj = x[j];
              const auto pj = h.row(j, ej) + off + c0;
#pragma unroll
              for (int r = 0; r < kChunk; r++)
                if (r < nc) acc[r] += pj[r] * xj;
            }
            t0 += k1 - k0;
          }
        }
#pragma unroll
        for (int r = 0; r < kChunk; r++) acc[r] = group_sum(acc[r], g.G);
        const int od = h.off(ei, ei);
        float y[kChunk];
#pragma unroll
        for (int r = kChunk - 1; r >= 0; r--) {
          if (r < nc) {
            float v = x[c0 + r] - acc[r];
#pragma unroll
            for (int q = r + 1; q < kChunk; q++)
              if (q < nc) v -= h.row(c0 + q, ei)[od + c0 + r] * y[q];
            y[r] = v / h.row(c0 + r, ei)[od + c0 + r];
          }
        }
        if (g.r == 0) {
#pragma unroll
          for (int r = 0; r < kChunk; r++)
            if (r < nc) x[c0 + r] = y[r];
        }
      }
    }
    __syncthreads();
  }
}

// grad = M qacc - force - J^T f into vec (factor space), then vec = H^-1 grad in place; returns this lane's share
// of |grad|^2 (dofs lane, lane + 64, ... in that order). The Newton direction is -vec read back through pdof.
template <class HS>
__device__ __forceinline__ double wave_update_gradient(const WaveCtx& w, const HS& h) {
  const WaveLds& L = w.L;
  const Env& e = w.e;
  double g2 = 0.0;
  for (int d = w.lane; d < w.D; d += kWave) {
    float g = L.Ma[d] - e.F(e.s.force, d) - L.qfrc[d];
    g2 += (double)(g * g);
    L.vec[L.pdof[d]] = g;
  }
  if (!HS::kFence && w.dense_h) wave_chol_solve_dense(w, L.vec);
  else wave_chol_solve(w, h, L.vec);
  __syncthreads();
  return g2;
}

struct WLS {
  float alpha, cost, d0, d1;
};

// The line search's per-row terms of rows lane, lane + 64, ..., up to kLsRows of them per lane, stay in registers
// across its points; rows beyond are read back from the env's row arrays. Each point sums the same terms in the same
// order either way.
constexpr int kLsRows = 1;
struct LsCache {
  float ja[kLsRows], jv[kLsRows], q0[kLsRows], q1[kLsRows], q2[kLsRows];
};

__device__ __forceinline__ void ls_row_terms(const WaveCtx& w, int c, float ja, float jv, float q0, float q1, float q2,
                                             float alpha, double& t0, double& t1, double& t2) {
  const Env& e = w.e;
  const gsr_state& s = e.s;
  float x = ja + alpha * jv;
  bool act = true;
  if (w.ne <= c && c < w.nef) {
    float f = e.EF(s.efc_fl, c), r = e.EF(s.efc_diag, c), rf = r * f;
    bool ln = x <= -rf, lp = x >= rf;
    if (ln || lp) {
      q0 = (float)ln * f * (-0.5f * rf - ja) + (float)lp * f * (-0.5f * rf + ja);
      q1 = (float)ln * (-f * jv) + (float)lp * (f * jv);
      q2 = 0.f;
    }
  } else if (w.nef <= c) {
    act = x < 0.f;
  }
  t0 += (double)(q0 * (float)act);
  t1 += (double)(q1 * (float)act);
  t2 += (double)(q2 * (float)act);
}

__device__ __forceinline__ WLS wave_ls_point(const WaveCtx& w, const LsCache& rc, float qg0, float qg1, float qg2,
                                             float alpha, int& ls_it) {
  const Env& e = w.e;
  const gsr_state& s = e.s;
  double t0 = 0.0, t1 = 0.0, t2 = 0.0;
#pragma unroll
  for (int k = 0; k < kLsRows; k++) {
    const int c = w.lane + kWave * k;
    if (c < w.n_con) ls_row_terms(w, c, rc.ja[k], rc.jv[k], rc.q0[k], rc.q1[k], rc.q2[k], alpha, t0, t1, t2);
  }
  for (int c = w.lane + kWave * kLsRows; c < w.n_con; c += kWave)
    ls_row_terms(w, c, e.EF(s.efc_Jaref, c), e.EF(s.efc_jv, c), e.QD(c, 0), e.QD(c, 1), e.QD(c, 2), alpha, t0, t1, t2);
  const float T0 = (float)((double)qg0 + wave_sum_d(t0));
  const float T1 = (float)((double)qg1 + wave_sum_d(t1));
  const float T2 = (float)((double)qg2 + wave_sum_d(t2));
  WLS p;
  p.alpha = alpha;
  p.cost = alpha * alpha * T2 + alpha * T1 + T0;
  p.d0 = 2.f * alpha * T2 + T1;
  p.d1 = 2.f * T2;
  if (p.d1 <= 0.f) p.d1 = e.m.eps;
  ls_it++;
  return p;
}

__device__ inline float wave_linesearch(const WaveCtx& w, float gauss) {
  const Env& e = w.e;
  const gsr_model& m = e.m;
  const gsr_state& s = e.s;
  const WaveLds& L = w.L;
  double sn = 0.0;
  for (int d = w.lane; d < w.D; d += kWave) sn += (double)(L.search[d] * L.search[d]);
  float snorm = sqrtf((float)wave_sum_d(sn));
  float scale = e.meaninertia() * (float)(w.D > 1 ? w.D : 1);
  float gtol = m.tolerance * m.ls_tolerance * snorm * scale;
  if (snorm < m.eps) return 0.f;
  wave_mass_mul(w, L.search, L.mv);
  __syncthreads();
  LsCache rc;
#pragma unroll
  for (int k = 0; k < kLsRows; k++) {
    const int c = w.lane + kWave * k;
    float jv = 0.f, ja = 0.f, Dc = 0.f;
    if (c < w.n_con) {
      jv = row_dot_vec(e, c, e.EI(s.jac_nnz, c), jv, L.search);
      e.EF(s.efc_jv, c) = jv;
      Dc = e.EF(s.efc_D, c);
      ja = e.EF(s.efc_Jaref, c);
    }
    rc.ja[k] = ja;
    rc.jv[k] = jv;
    rc.q0[k] = Dc * (0.5f * ja * ja);
    rc.q1[k] = Dc * (jv * ja);
    rc.q2[k] = Dc * (0.5f * jv * jv);
  }
  for (int c = w.lane + kWave * kLsRows; c < w.n_con; c += kWave) {
    float jv = 0.f;
    int nnz = e.EI(s.jac_nnz, c);
    jv = row_dot_vec(e, c, nnz, jv, L.search);
    e.EF(s.efc_jv, c) = jv;
    float Dc = e.EF(s.efc_D, c), ja = e.EF(s.efc_Jaref, c);
    e.QD(c, 0) = Dc * (0.5f * ja * ja);
    e.QD(c, 1) = Dc * (jv * ja);
    e.QD(c, 2) = Dc * (0.5f * jv * jv);
  }
  double a1 = 0.0, a2 = 0.0;
  for (int d = w.lane; d < w.D; d += kWave) {
    float sd = L.search[d];
    a1 += (double)(sd * L.Ma[d] - sd * e.F(s.force, d));
    a2 += (double)(0.5f * sd * L.mv[d]);
  }
  float qg0 = gauss, qg1 = (float)wave_sum_d(a1), qg2 = (float)wave_sum_d(a2);
  __syncthreads();
  int ls_it = 0;
  WLS p0 = wave_ls_point(w, rc, qg0, qg1, qg2, 0.f, ls_it);
  WLS p1 = wave_ls_point(w, rc, qg0, qg1, qg2, p0.alpha - p0.d0 / p0.d1, ls_it);
  if (p0.cost < p1.cost) p1 = p0;
  if (fabsf(p1.d0) < gtol) return p1.alpha;
  float direction = (p1.d0 < 0.f) ? 1.f : -1.f;
  int p2update = 0;
  WLS p2 = p1;
  while (p1.d0 * direction <= -gtol && ls_it < m.ls_iterations) {
    p2 = p1;
    p2update = 1;
    p1 = wave_ls_point(w, rc, qg0, qg1, qg2, p1.alpha - p1.d0 / p1.d1, ls_it);
    if (fabsf(p1.d0) < gtol) return p1.alpha;
  }
  if (ls_it >= m.ls_iterations) return p1.alpha;
  if (!p2update) return p1.alpha;
  WLS p2n = p1;
  WLS p1n = wave_ls_point(w, rc, qg0, qg1, qg2, p1.alpha - p1.d0 / p1.d1, ls_it);
  while (ls_it < m.ls_iterations) {
    WLS pm = wave_ls_point(w, rc, qg0, qg1, qg2, (p1.alpha + p2.alpha) * 0.5f, ls_it);
    // candidates p1n, p2n, pm in this order (named values, not an indexed array, so they stay in registers)
    const WLS c0 = p1n, c1 = p2n, c2 = pm;
    int best = -1;
    float bcost = 0.f, balpha = 0.f;
    auto consider = [&](const WLS& c, int idx) {
      if (fabsf(c.d0) < gtol && (best < 0 || c.cost < bcost)) {
        bcost = c.cost;
        best = idx;
        balpha = c.alpha;
      }
    };
    consider(c0, 0);
    consider(c1, 1);
    consider(c2, 2);
    if (best >= 0) return balpha;
    // p1, then p2, moves to the candidates whose slope has its sign and is closer to zero
    auto narrow = [&](WLS& p) {
      int flag = 0;
      auto step = [&](const WLS& c) {
        if (p.d0 < 0.f && c.d0 < 0.f && p.d0 < c.d0) { p = c; flag = 1; }
        else if (p.d0 > 0.f && c.d0 > 0.f && p.d0 > c.d0) { p = c; flag = 2; }
      };
      step(c0);
      step(c1);
      step(c2);
      return flag;
    };
    const int f1 = narrow(p1);
    p1n = p1;
    if (f1 > 0) p1n = wave_ls_point(w, rc, qg0, qg1, qg2, p1.alpha - p1.d0 / p1.d1, ls_it);
    const int f2 = narrow(p2);
    p2n = p2;
    if (f2 > 0) p2n = wave_ls_point(w, rc, qg0, qg1, qg2, p2.alpha - p2.d0 / p2.d1, ls_it);
    if (f1 == 0 && f2 == 0) return pm.alpha;
  }
  if (p1.cost <= p2.cost && p1.cost < p0.cost) return p1.alpha;
  if (p2.cost <= p1.cost && p2.cost < p1.cost) return p2.alpha;
  return 0.f;
}

// Above this many activity changes in one Newton iteration the factor is rebuilt instead of updated row by row.
constexpr int kMaxRank1PerIter = 8;

// Newton iterations (func_solve / func_solve_body) from the initial factorisation on; returns the iteration count
template <class HS>
__device__ __forceinline__ int wave_newton(const WaveCtx& w, const HS& h, double cost, float gauss) {
  const Env& e = w.e;
  const gsr_model& m = e.m;
  const gsr_state& s = e.s;
  const WaveLds& L = w.L;
  const int lane = w.lane, D = w.D, n_con = w.n_con;
  wave_hessian_direct(w, h);
  wave_update_gradient(w, h);
  for (int d = lane; d < D; d += kWave) L.search[d] = -L.vec[L.pdof[d]];
  __syncthreads();
  int iters = 0;
  if (n_con <= 0) return 0;
  for (int it = 0; it < m.iterations; it++) {
    iters++;
    float alpha = wave_linesearch(w, gauss);
    bool improved;
    if (fabsf(alpha) < m.eps) {
      improved = false;
    } else {
      for (int d = lane; d < D; d += kWave) {
        L.qacc[d] = L.qacc[d] + L.search[d] * alpha;
        L.Ma[d] = L.Ma[d] + L.mv[d] * alpha;
      }
      for (int c = lane; c < n_con; c += kWave) e.EF(s.efc_Jaref, c) = e.EF(s.efc_Jaref, c) + e.EF(s.efc_jv, c) * alpha;
      __syncthreads();
      double prev_cost = cost;
      wave_update_constraint(w, cost, gauss);
      // activity changes in row order; rank-1 updates or a rebuild of the factor
      int n_changed = 0;
      for (int c = lane; c < n_con; c += kWave) {
        int a = e.EI(s.efc_active, c);
        n_changed += ((a & 1) != ((a >> 1) & 1));
      }
      for (int o = kWave / 2; o > 0; o >>= 1) n_changed += __shfl_xor(n_changed, o, kWave);
      if (n_changed > kMaxRank1PerIter) {
        wave_hessian_direct(w, h);
      } else if (n_changed > 0) {
        bool rebuilt = false;
        for (int c0 = 0; c0 < n_con && !rebuilt; c0 += kWave) {
          int c = c0 + lane;
          int a = c < n_con ? e.EI(s.efc_active, c) : 0;
          unsigned long long chg = __ballot((a & 1) != ((a >> 1) & 1));
          while (chg && !rebuilt) {
            int j = __ffsll((long long)chg) - 1;
            chg &= chg - 1;
            int aj = __shfl(a, j, kWave);
            const float sgn = (aj & 1) ? 1.f : -1.f;
            if (!((!HS::kFence && w.dense_h) ? wave_rank1_dense(w, c0 + j, sgn) : wave_rank1(w, h, c0 + j, sgn))) {
              wave_hessian_direct(w, h);
              rebuilt = true;
            }
          }
        }
      }
      double g2 = wave_update_gradient(w, h);
      float tol_scaled = (w.e.meaninertia() * (float)(D > 1 ? D : 1)) * m.tolerance;
      double improvement = prev_cost - cost;
      float gn = sqrtf((float)wave_sum_d(g2));
      if (gn < tol_scaled || improvement < tol_scaled) {
        improved = false;
      } else {
        improved = true;
        for (int d = lane; d < D; d += kWave) L.search[d] = -L.vec[L.pdof[d]];
      }
      __syncthreads();
    }
    if (!improved) break;
  }
  return iters;
}

__device__ __forceinline__ void wave_solve(const Env& e, char* lds) {
  const gsr_model& m = e.m;
  const gsr_state& s = e.s;
  const int lane = threadIdx.x;
  const int D = m.n_dofs, E = m.n_entities;
  WaveLds L = wave_lds_carve((lds_c*)lds, D, E, m.wave_h_floats);
  int n_con = e.I(s.n_con, 0), ne = e.I(s.n_con, 1), nef = ne + e.I(s.n_con, 2);
  // the dense row sums when one entity holds every dof (at most 64, with H in LDS and room after it for 8 or more
  // staged rows)
  bool dense_h = false;
  int dense_rows = 0;
  lds_f* dense_buf = nullptr;
  {
    int n_ent = 0;
    for (int i = 0; i < E; i++) n_ent += m.ent_dof_end[i] > m.ent_dof_start[i] ? 1 : 0;
    const int tri = D * (D + 1) / 2, tri4 = (tri + 3) & ~3;
    dense_h = n_ent == 1 && D <= kWave && tri <= m.wave_h_floats;
    if (dense_h && tri4 < m.wave_h_floats) {
      const int R = min(kWave, (m.wave_h_floats - tri4) / (D + 1));
      if (R >= 8) {
        dense_rows = R;
        dense_buf = L.hl + tri4;
      }
    }
  }
  WaveCtx w{e, L, lane, D, E, n_con, ne, nef, dense_h, dense_rows, dense_buf};
  int ws = e.I(s.is_warmstart, 0);
  // force (the smooth force until the epilogue) and acc_smooth are read from the state: nothing writes them here
  for (int d = lane; d < D; d += kWave) L.qacc[d] = (n_con > 0 && ws) ? e.F(s.qacc_ws, d) : e.F(s.acc_smooth, d);
  __syncthreads();
  wave_build_structure(w);
  wave_mass_mul(w, L.qacc, L.Ma);
  for (int c = lane; c < n_con; c += kWave) {
    float ja = -e.EF(s.efc_aref, c);
    int nnz = e.EI(s.jac_nnz, c);
    ja = row_dot_vec(e, c, nnz, ja, L.qacc);
    e.EF(s.efc_Jaref, c) = ja;
  }
  __syncthreads();
  double cost;
  float gauss;
  wave_update_constraint(w, cost, gauss);
  int iters = L.hmode[0] ? wave_newton(w, HLds{L}, cost, gauss) : wave_newton(w, HGlobal{e, L.es}, cost, gauss);
  // epilogue (func_update_qacc, func_update_contact_force)
  int bad = 0;
  for (int d = lane; d < D; d += kWave) {
    float q = L.qacc[d];
    e.F(s.acc, d) = q;
    e.F(s.force, d) = e.F(s.qf_smooth, d) + L.qfrc[d];
    e.F(s.qacc_ws, d) = q;
    bad |= isnan(q) ? 1 : 0;
  }
  bad = wave_or(bad);
  if (lane == 0) {
    e.I(s.solver_iters, 0) = iters;
    e.I(s.is_warmstart, 0) = 1;
    if (bad) e.I((int32_t*)s.errno_env, 0) |= (int)GSR_ERR_NAN_QACC;
  }
  int cs = ne + e.I(s.n_con, 2);
  int nc = m.enable_collision ? e.I(s.n_contacts, 0) : 0;
  for (int ic = lane; ic < nc; ic += kWave) {
    f3 cn = e.L3(s.contact_normal, ic);
    float fr = e.F(s.contact_friction, ic);
    f3 d1, d2;
    orthogonals(cn, d1, d2);
    f3 f = zero3();
    for (int i = 0; i < 4; i++) {
      f3 dd = (float)(2 * (i % 2) - 1) * (i < 2 ? d1 : d2);
      f3 n = dd * fr - cn;
      f = f + n * e.EF(s.efc_force, ic * 4 + i + cs);
    }
    e.S3(s.contact_force, ic, f);
  }
}

}  // namespace gsr
