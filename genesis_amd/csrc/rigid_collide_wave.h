// Collision detection and constraint rows of one environment on one wavefront (64 lanes).
//
// Same results as the one-lane-per-env path of rigid_collide.h / rigid_solver.h, spread over lanes:
//   * AABBs: one geom per lane (rigid_solver_decomp.py:5668-5689).
//   * Sweep-and-prune (collider_decomp.py:1331-1470): the reference warm-starts an insertion sort with strict `<`,
//     which is a stable sort of the previous order; here each lane ranks its entries (rank = #smaller + #equal
//     before it), which gives the same permutation. The sweep's active list at a min event holds the geoms whose
//     min event came earlier and whose max event comes later, in min-event order; each lane walks that set for its
//     own events and a wave prefix sum of the pair counts places every pair where the serial sweep puts it.
//   * Narrowphase: one broadphase pair per lane, contacts kept in the lane (LocalSink) and placed by a prefix sum,
//     convex pairs first and plane-box pairs second, both in broadphase order, as the reference appends them.
//   * Rows: equality, frictionloss and limit rows on lane 0 (few), the 4 rows of each contact on its own lane.
// Where the broadphase pair buffer overflows (errno bit 1) the serial sweep stops clearing the normal cache of
// later non-overlapping pairs; this path clears them all. Results differ only after that error.
#pragma once
#include "rigid_collide.h"
#include "rigid_solver.h"

namespace gsr {

constexpr int kWaveMaxGeoms = 1024;

__device__ __forceinline__ int wave_excl_scan(int v, int lane, int& total) {
  int x = v;
  for (int o = 1; o < 64; o <<= 1) {
    int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  total = __shfl(x, 63, 64);
  return x - v;
}

__host__ __device__ inline size_t wave_collide_lds_bytes(int G) {
  return sizeof(float) * 4 * (size_t)G + sizeof(int) * 6 * (size_t)G + 64;
}

__device__ inline void wave_update_aabbs(const Env& e, int lane) {
  const gsr_model& m = e.m;
  const gsr_state& s = e.s;
  for (int g = lane; g < m.n_geoms; g += 64) {
    f3 gp = e.L3(s.geoms_pos, g);
    q4 gq = e.L4(s.geoms_quat, g);
    f3 lo{INFINITY, INFINITY, INFINITY}, hi{-INFINITY, -INFINITY, -INFINITY};
    for (int c = 0; c < 8; c++) {
      f3 p = xform(T3(m.geom_init_aabb, 8 * g + c), gp, gq);
      lo = vmin(lo, p);
      hi = vmax(hi, p);
    }
    e.S3(s.aabb_min, g, lo);
    e.S3(s.aabb_max, g, hi);
  }
}

__device__ inline void wave_broad_phase(const Env& e, char* lds, int lane) {
  const gsr_model& m = e.m;
  const gsr_state& s = e.s;
  const int G = m.n_geoms, G2 = 2 * G;
  const int MAXB = 1 << 30;
  float* sv = reinterpret_cast<float*>(lds);
  float* tv = sv + G2;
  int* sig = reinterpret_cast<int*>(tv + G2);
  int* tig = sig + G2;
  int* minpos = tig + G2;
  int* maxpos = minpos + G;
  const bool first = e.I(s.first_time, 0) != 0;
  for (int i = lane; i < G2; i += 64) {
    int t;
    if (first) t = (i & 1) ? ((i >> 1) | MAXB) : (i >> 1);
    else t = e.I(s.sort_ig, i);
    int g = t & (MAXB - 1);
    sig[i] = t;
    sv[i] = (t & MAXB) ? e.F(s.aabb_max, 3 * g) : e.F(s.aabb_min, 3 * g);
  }
  __syncthreads();
  if (first && lane == 0) e.I(s.first_time, 0) = 0;
  // stable rank sort
  for (int i = lane; i < G2; i += 64) {
    float v = sv[i];
    int r = 0;
    for (int j = 0; j < G2; j++) {
      float u = sv[j];
      r += (u < v) || (u == v && j < i);
    }
    tv[r] = v;
    tig[r] = sig[i];
  }
  __syncthreads();
  for (int i = lane; i < G2; i += 64) {
    float v = tv[i];
    int t = tig[i];
    sv[i] = v;
    sig[i] = t;
    e.F(s.sort_value, i) = v;
    e.I(s.sort_ig, i) = t;
    if (t & MAXB) maxpos[t & (MAXB - 1)] = i;
    else minpos[t] = i;
  }
  __syncthreads();
  int base = 0;
  bool overflow = false;
  for (int p0 = 0; p0 < G2; p0 += 64) {
    const int p = p0 + lane;
    const bool is_min = p < G2 && !(sig[p] & MAXB);
    const int gi = is_min ? sig[p] : 0;
    int cnt = 0;
    if (is_min) {
      for (int j = 0; j < p; j++) {
        int t = sig[j];
        if ((t & MAXB) || maxpos[t] <= p) continue;
        int ga = t < gi ? t : gi, gb = t < gi ? gi : t;
        if (m.pair_idx[(size_t)ga * G + gb] == -1) continue;
        if (dev_aabb_overlap(e, ga, gb)) cnt++;
      }
    }
    int total;
    int k = base + wave_excl_scan(cnt, lane, total);
    if (is_min) {
      for (int j = 0; j < p; j++) {
        int t = sig[j];
        if ((t & MAXB) || maxpos[t] <= p) continue;
        int ga = t < gi ? t : gi, gb = t < gi ? gi : t;
        int pidx = m.pair_idx[(size_t)ga * G + gb];
        if (pidx == -1) continue;
        if (!dev_aabb_overlap(e, ga, gb)) {
          e.S3(s.cache_normal, pidx, zero3());
          continue;
        }
        if (k < m.max_broad) e.I(s.broad_pairs, k) = (ga << 16) | gb;
        else overflow = true;
        k++;
      }
    }
    base += total;
  }
  overflow = __any(overflow);
  if (lane == 0) {
    e.I(s.n_broad, 0) = min(base, m.max_broad);
    if (overflow) e.I((int32_t*)s.errno_env, 0) |= (int)GSR_ERR_BROAD_OVERFLOW;
  }
  __syncthreads();
}

// one pass over the broadphase pairs of one kind (convex pairs, then plane-box pairs), contacts appended at `base`
// kGjk: pairs that need GJK/EPA are redone in place, one lane at a time (they share the env's workspace), instead of
// being deferred to the GJK/EPA pass
template <bool kGjk>
__device__ inline int wave_narrow_pass(const Env& e, int lane, int nb, bool plane_box_pass, int base,
                                       uint32_t& err, int& n_def) {
  const gsr_model& m = e.m;
  const gsr_state& s = e.s;
  for (int p0 = 0; p0 < nb; p0 += 64) {
    const int ip = p0 + lane;
    PairContacts pc;
    pc.cnt = 0;
    pc.err = 0;
    pc.deferred = false;
    int ga = 0, gb = 0;
    bool pb = false;
    if (ip < nb) {
      int pr = e.I(s.broad_pairs, ip);
      ga = pr >> 16;
      gb = pr & 0xffff;
      if (m.geom_type[ga] > m.geom_type[gb]) { int t = ga; ga = gb; gb = t; }
      pb = m.geom_type[ga] == GEOM_PLANE && m.geom_type[gb] == GEOM_BOX;
      if (pb == plane_box_pass) {
        LocalSink<false> snk{pc};
        if (pb) dev_plane_box(e, ga, gb, snk);
        else dev_convex_convex(e, ga, gb, snk);
      }
    }
    if (kGjk) {
      unsigned long long redo = __ballot(pc.deferred);
      while (redo) {
        int j = __ffsll((long long)redo) - 1;
        redo &= redo - 1;
        if (lane == j) {
          pc.cnt = 0;
          pc.err = 0;
          pc.deferred = false;
          LocalSink<true> snk{pc};
          dev_convex_convex(e, ga, gb, snk);
        }
      }
    }
    // a pair that needs GJK/EPA contributes no contact here; it is recorded with the index its contacts go to
    if (pc.deferred) {
      pc.cnt = 0;
      pc.err = 0;
    }
    const unsigned long long def = __ballot(pc.deferred);
    int total;
    int k = base + wave_excl_scan(pc.cnt, lane, total);
    if (pc.deferred) {
      const int slot = n_def + __popcll(def & ((1ull << lane) - 1ull));
      e.I(s.gjk_defer, 2 * slot) = (ga << 16) | gb;
      e.I(s.gjk_defer, 2 * slot + 1) = min(k, m.max_contacts);
    }
    n_def += __popcll(def);
    float fr = fmaxf(fmaxf(m.geom_friction[ga] * e.F(s.friction_ratio, ga),
                           m.geom_friction[gb] * e.F(s.friction_ratio, gb)), 1e-2f);
    for (int c = 0; c < pc.cnt; c++, k++) {
      if (k < m.max_contacts) {
        e.I(s.contact_geoms, 2 * k) = ga;
        e.I(s.contact_geoms, 2 * k + 1) = gb;
        e.S3(s.contact_normal, k, pc.n[c]);
        e.S3(s.contact_pos, k, pc.p[c]);
        e.F(s.contact_pen, k) = pc.pen[c];
        e.F(s.contact_friction, k) = fr;
      } else {
        pc.err |= GSR_ERR_CONTACT_OVERFLOW;
      }
    }
    err |= pc.err;
    base += total;
  }
  return base;
}

template <bool kGjk>
__device__ inline void wave_narrow_phase(const Env& e, int lane) {
  const gsr_model& m = e.m;
  const gsr_state& s = e.s;
  const int nb = e.I(s.n_broad, 0);
  uint32_t err = 0;
  int n_def = 0;
  int n = wave_narrow_pass<kGjk>(e, lane, nb, false, 0, err, n_def);
  n = wave_narrow_pass<kGjk>(e, lane, nb, true, n, err, n_def);
  for (int o = 32; o > 0; o >>= 1) err |= (uint32_t)__shfl_xor((int)err, o, 64);
  if (lane == 0) {
    e.I(s.n_contacts, 0) = min(n, m.max_contacts);
    e.I(s.gjk_defer_n, 0) = n_def;
    if (err) e.I((int32_t*)s.errno_env, 0) |= (int)err;
  }
  __syncthreads();
}

// frictionloss rows on lane 0, the 4 rows of each contact on its own lane, limit rows on lane 0 after them
__device__ __forceinline__ void wave_rows(const Env& e) {
  const gsr_model& m = e.m;
  const gsr_state& s = e.s;
  const int lane = threadIdx.x;
  int n_con = 0, n_fl = 0;
  if (lane == 0) n_con = dev_frictionloss_rows(e, e.I(s.n_con, 0), n_fl);
  n_con = __shfl(n_con, 0, 64);
  if (m.enable_collision) {
    int nc = e.I(s.n_contacts, 0);
    for (int ic = lane; ic < nc; ic += 64) dev_contact_rows(e, ic, n_con + 4 * ic);
    n_con += 4 * nc;
  }
  if (lane == 0) {
    if (m.enable_joint_limit) n_con = dev_limit_rows(e, n_con);
    e.I(s.n_con, 0) = n_con;
    e.I(s.n_con, 2) = n_fl;
  }
  __syncthreads();
}

// equality rows, collision detection and inequality rows of one env (k_collide's work) on one wavefront
template <bool kGjk = false>
__device__ __forceinline__ void wave_collide(const Env& e, char* lds) {
  const gsr_model& m = e.m;
  const gsr_state& s = e.s;
  const int lane = threadIdx.x;
  if (lane == 0) {
    dev_equality_rows(e);
    e.I(s.n_contacts, 0) = 0;
    e.I(s.n_broad, 0) = 0;
    e.I(s.gjk_defer_n, 0) = 0;
  }
  if (m.enable_collision) {
    wave_update_aabbs(e, lane);
    __syncthreads();
    if (m.n_pairs > 0) {
      wave_broad_phase(e, lds, lane);
      wave_narrow_phase<kGjk>(e, lane);
    }
  }
  __syncthreads();
  // with deferred pairs the contact list is not final yet: the GJK/EPA pass builds the rows after inserting them
  if (e.I(s.gjk_defer_n, 0) > 0) return;
  wave_rows(e);
}

// The GJK/EPA pass of one env on one wavefront: lane 0 recomputes the deferred pairs with GJK/EPA (they share the
// env's workspace), last first; the contacts above each pair's insertion index move up across the lanes (64 at a
// time from the top: every lane reads before any writes), and lane 0 writes the pair's contacts in the gap.
// Contacts pushed past max_contacts are dropped with the overflow bit, as the one-pass narrowphase drops them.
// Returns whether the env had deferred pairs (uniform).
__device__ inline bool wave_gjk_fixup(const Env& e) {
  const gsr_model& m = e.m;
  const gsr_state& s = e.s;
  const int lane = threadIdx.x;
  const int nd = e.I(s.gjk_defer_n, 0);
  if (nd == 0) return false;
  __syncthreads();
  for (int j = nd - 1; j >= 0; j--) {
    const int pr = e.I(s.gjk_defer, 2 * j), at = e.I(s.gjk_defer, 2 * j + 1);
    const int ga = pr >> 16, gb = pr & 0xffff;
    PairContacts pc;
    pc.cnt = 0;
    pc.err = 0;
    pc.deferred = false;
    if (lane == 0) {
      LocalSink<true> snk{pc};
      dev_convex_convex(e, ga, gb, snk);
      if (pc.err) e.I((int32_t*)s.errno_env, 0) |= (int)pc.err;
    }
    const int cnt = __shfl(pc.cnt, 0, 64);
    if (cnt == 0) continue;
    const int n = e.I(s.n_contacts, 0);
    const int kept = min(n + cnt, m.max_contacts);
    for (int top = n; top > at; top -= 64) {
      const int i = top - 1 - lane;
      const bool act = i >= at;
      int g0 = 0, g1 = 0;
      f3 cp = zero3(), cn = zero3();
      float pe = 0.f, fr = 0.f;
      if (act) {
        g0 = e.I(s.contact_geoms, 2 * i);
        g1 = e.I(s.contact_geoms, 2 * i + 1);
        cp = e.L3(s.contact_pos, i);
        cn = e.L3(s.contact_normal, i);
        pe = e.F(s.contact_pen, i);
        fr = e.F(s.contact_friction, i);
      }
      __syncthreads();
      const int t = i + cnt;
      if (act && t < kept) {
        e.I(s.contact_geoms, 2 * t) = g0;
        e.I(s.contact_geoms, 2 * t + 1) = g1;
        e.S3(s.contact_pos, t, cp);
        e.S3(s.contact_normal, t, cn);
        e.F(s.contact_pen, t) = pe;
        e.F(s.contact_friction, t) = fr;
      }
      __syncthreads();
    }
    if (lane == 0) {
      if (n + cnt > m.max_contacts) e.I((int32_t*)s.errno_env, 0) |= (int)GSR_ERR_CONTACT_OVERFLOW;
      const float fr = fmaxf(fmaxf(m.geom_friction[ga] * e.F(s.friction_ratio, ga),
                                   m.geom_friction[gb] * e.F(s.friction_ratio, gb)), 1e-2f);
      for (int c = 0; c < cnt; c++) {
        const int t = at + c;
        if (t >= kept) break;
        e.I(s.contact_geoms, 2 * t) = ga;
        e.I(s.contact_geoms, 2 * t + 1) = gb;
        e.S3(s.contact_pos, t, pc.p[c]);
        e.S3(s.contact_normal, t, pc.n[c]);
        e.F(s.contact_pen, t) = pc.pen[c];
        e.F(s.contact_friction, t) = fr;
      }
      e.I(s.n_contacts, 0) = kept;
    }
    __syncthreads();
  }
  if (lane == 0) e.I(s.gjk_defer_n, 0) = 0;
  __syncthreads();
  return true;
}

}  // namespace gsr
