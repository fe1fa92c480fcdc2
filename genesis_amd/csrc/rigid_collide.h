// Collision detection of one environment: world AABBs (rigid_solver_decomp.py:5668-5689), sweep-and-prune broadphase
// with the warm-started insertion sort (collider_decomp.py:1331-1470), convex narrowphase through MPR with the
// cached-normal warm start and multi-contact perturbation (collider_decomp.py:2328-2725, mpr_decomp.py), and the
// closed-form plane-box contact (collider_decomp.py:2073-2142). Contacts are appended in the reference order: every
// convex pair in broadphase order first, plane-box pairs second.
#pragma once
#include "rigid_env.h"
#include "rigid_support.h"
#include "rigid_gjk.h"

namespace gsr {

constexpr float kCcdEps = 1e-9f;
constexpr float kCcdTol = 1e-6f;
constexpr int kCcdIters = 50;
// The reference's portal-refinement loops are unbounded; this cap only guards pathological inputs on the GPU.
constexpr int kMprLoopCap = 1000;

__device__ inline void dev_update_aabbs(const Env& e) {
  const gsr_model& m = e.m;
  const gsr_state& s = e.s;
  for (int g = 0; g < m.n_geoms; g++) {
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

__device__ inline bool dev_aabb_overlap(const Env& e, int a, int b) {
  const gsr_state& s = e.s;
  f3 amax = e.L3(s.aabb_max, a), amin = e.L3(s.aabb_min, a), bmax = e.L3(s.aabb_max, b), bmin = e.L3(s.aabb_min, b);
  bool sep = (amax.x <= bmin.x) || (amax.y <= bmin.y) || (amax.z <= bmin.z) || (amin.x >= bmax.x) ||
             (amin.y >= bmax.y) || (amin.z >= bmax.z);
  return !sep;
}

// sort entries: sort_ig holds geom index | (is_max << 30)
__device__ inline void dev_broad_phase(const Env& e) {
  const gsr_model& m = e.m;
  const gsr_state& s = e.s;
  const int G = m.n_geoms;
  const int MAXB = 1 << 30;
  if (e.I(s.first_time, 0)) {
    for (int i = 0; i < G; i++) {
      e.F(s.sort_value, 2 * i) = e.F(s.aabb_min, 3 * i);
      e.I(s.sort_ig, 2 * i) = i;
      e.F(s.sort_value, 2 * i + 1) = e.F(s.aabb_max, 3 * i);
      e.I(s.sort_ig, 2 * i + 1) = i | MAXB;
    }
    e.I(s.first_time, 0) = 0;
  } else {
    for (int i = 0; i < 2 * G; i++) {
      int t = e.I(s.sort_ig, i);
      int g = t & (MAXB - 1);
      e.F(s.sort_value, i) = (t & MAXB) ? e.F(s.aabb_max, 3 * g) : e.F(s.aabb_min, 3 * g);
    }
  }
  for (int i = 1; i < 2 * G; i++) {
    float kv = e.F(s.sort_value, i);
    int kt = e.I(s.sort_ig, i);
    int j = i - 1;
    while (j >= 0 && kv < e.F(s.sort_value, j)) {
      e.F(s.sort_value, j + 1) = e.F(s.sort_value, j);
      e.I(s.sort_ig, j + 1) = e.I(s.sort_ig, j);
      j--;
    }
    e.F(s.sort_value, j + 1) = kv;
    e.I(s.sort_ig, j + 1) = kt;
  }
  int n_broad = 0, n_active = 0;
  for (int i = 0; i < 2 * G; i++) {
    int t = e.I(s.sort_ig, i);
    int gi = t & (MAXB - 1);
    if (!(t & MAXB)) {
      for (int j = 0; j < n_active; j++) {
        int ga = e.I(s.active_buffer, j), gb = gi;
        if (ga > gb) { int tmp = ga; ga = gb; gb = tmp; }
        int pidx = m.pair_idx[(size_t)ga * G + gb];
        if (pidx == -1) continue;
        if (!dev_aabb_overlap(e, ga, gb)) {
          e.S3(s.cache_normal, pidx, zero3());
          continue;
        }
        if (n_broad == m.max_broad) {
          e.I((int32_t*)s.errno_env, 0) |= (int)GSR_ERR_BROAD_OVERFLOW;
          break;
        }
        e.I(s.broad_pairs, n_broad) = (ga << 16) | gb;
        n_broad++;
      }
      e.I(s.active_buffer, n_active) = gi;
      n_active++;
    } else {
      for (int j = 0; j < n_active; j++) {
        if (e.I(s.active_buffer, j) == gi) {
          for (int k = j; k < n_active - 1; k++) e.I(s.active_buffer, k) = e.I(s.active_buffer, k + 1);
          n_active--;
          break;
        }
      }
    }
  }
  e.I(s.n_broad, 0) = n_broad;
}

__device__ inline void dev_add_contact(const Env& e, int ga, int gb, f3 normal, f3 pos, float pen) {
  const gsr_model& m = e.m;
  const gsr_state& s = e.s;
  int i = e.I(s.n_contacts, 0);
  if (i < m.max_contacts) {
    e.I(s.contact_geoms, 2 * i) = ga;
    e.I(s.contact_geoms, 2 * i + 1) = gb;
    e.S3(s.contact_normal, i, normal);
    e.S3(s.contact_pos, i, pos);
    e.F(s.contact_pen, i) = pen;
    // per-env friction ratios (collider_decomp.py:2161-2162, 2200-2201)
    float fa = m.geom_friction[ga] * e.F(s.friction_ratio, ga);
    float fb = m.geom_friction[gb] * e.F(s.friction_ratio, gb);
    e.F(s.contact_friction, i) = fmaxf(fmaxf(fa, fb), 1e-2f);
    e.I(s.n_contacts, 0) = i + 1;
  } else {
    e.I((int32_t*)s.errno_env, 0) |= (int)GSR_ERR_CONTACT_OVERFLOW;
  }
}

// Where the narrowphase puts the contacts of one pair. GlobalSink appends to the env's contact list at once (one
// lane per env); LocalSink keeps them in the lane until the wavefront has counted every pair's contacts.
//
// The collision kernels contain no GJK/EPA code: its registers would lower the occupancy of every kernel holding the
// narrowphase (the fused batch kernel among them). A pair that needs it is deferred instead: it contributes no
// contact, its normal cache is left as it was, and (pair, index of the next contact) is recorded in the env's
// deferral list. The GJK/EPA pass (wave_gjk_fixup) then recomputes each deferred pair and inserts its contacts at the
// recorded index, which puts every contact where the one-pass narrowphase puts it.
// kGjk: run GJK/EPA in place (the lane owns its env's workspace), instead of deferring the pair
template <bool kGjk>
struct GlobalSink {
  const Env& e;
  int start;  // contacts stored before this pair
  __device__ explicit GlobalSink(const Env& e_) : e(e_), start(e_.I(e_.s.n_contacts, 0)) {}
  __device__ void add(int ga, int gb, f3 n, f3 p, float pen) { dev_add_contact(e, ga, gb, n, p, pen); }
  // position of the k-th most recent contact (k = 0: the last one)
  __device__ f3 recent_pos(int k) const { return e.L3(e.s.contact_pos, e.I(e.s.n_contacts, 0) - 1 - k); }
  __device__ void flag(uint32_t bits) { e.I((int32_t*)e.s.errno_env, 0) |= (int)bits; }
  __device__ static constexpr bool gjk_allowed() { return kGjk; }
  // the pair's contacts so far (its unperturbed detection) are withdrawn: the GJK/EPA pass recomputes all of them
  __device__ void defer(int ga, int gb) {
    e.I(e.s.n_contacts, 0) = start;
    int n = e.I(e.s.gjk_defer_n, 0);
    e.I(e.s.gjk_defer, 2 * n) = (ga << 16) | gb;
    e.I(e.s.gjk_defer, 2 * n + 1) = start;
    e.I(e.s.gjk_defer_n, 0) = n + 1;
  }
};

constexpr int kMaxPairContacts = 8;

struct PairContacts {
  f3 n[kMaxPairContacts], p[kMaxPairContacts];
  float pen[kMaxPairContacts];
  int cnt;
  uint32_t err;
  bool deferred;  // the pair needs GJK/EPA: redo it when the env's GJK workspace is free
};

// kGjk: this lane owns the env's GJK workspace; without it a pair that needs GJK/EPA is only marked deferred (and the
// GJK code is not compiled into that instance at all)
template <bool kGjk>
struct LocalSink {
  PairContacts& c;
  __device__ void add(int, int, f3 nn, f3 pp, float pe) {
    if (c.cnt < kMaxPairContacts) {
      c.n[c.cnt] = nn;
      c.p[c.cnt] = pp;
      c.pen[c.cnt] = pe;
      c.cnt++;
    }
  }
  __device__ f3 recent_pos(int k) const { return c.p[c.cnt - 1 - k]; }
  __device__ void flag(uint32_t bits) { c.err |= bits; }
  __device__ static constexpr bool gjk_allowed() { return kGjk; }
  __device__ void defer(int, int) { c.deferred = true; }
};

__device__ inline float dev_tolerance(const gsr_model& m, int ga, int gb, float tol) {
  float sb = norm(T3(m.geom_init_aabb, 8 * gb + 7) - T3(m.geom_init_aabb, 8 * gb));
  float sz = sb;
  if (m.geom_type[ga] != GEOM_PLANE) sz = fminf(norm(T3(m.geom_init_aabb, 8 * ga + 7) - T3(m.geom_init_aabb, 8 * ga)), sb);
  return 0.5f * tol * sz;
}

struct MPRSimplex {
  f3 v1[4], v2[4], v[4];
  int size;
};

__device__ inline void dev_mpr_support(const gsr_model& m, int ga, int gb, const GPose& A, const GPose& Bp, f3 dir, f3& v,
                                       f3& v1, f3& v2) {
  v1 = dev_support(m, ga, A, dir);
  v2 = dev_support(m, gb, Bp, -dir);
  v = v1 - v2;
}

__device__ inline f3 mpr_portal_dir(const MPRSimplex& s) { return normalized(cross(s.v[2] - s.v[1], s.v[3] - s.v[1])); }

__device__ inline bool mpr_reach_tol(const MPRSimplex& s, f3 v, f3 dir) {
  float dv1 = dot(s.v[1], dir), dv2 = dot(s.v[2], dir), dv3 = dot(s.v[3], dir), dv4 = dot(v, dir);
  float d1 = fminf(fminf(dv4 - dv1, dv4 - dv2), dv4 - dv3);
  return d1 < kCcdTol + kCcdEps * fmaxf(1.f, d1);
}

__device__ inline void mpr_expand(MPRSimplex& s, f3 v, f3 v1, f3 v2) {
  f3 v4v0 = cross(v, s.v[0]);
  float d = dot(s.v[1], v4v0);
  int i_s;
  if (d > 0.f) {
    d = dot(s.v[2], v4v0);
    i_s = d > 0.f ? 1 : 3;
  } else {
    d = dot(s.v[3], v4v0);
    i_s = d > 0.f ? 2 : 1;
  }
  // the register-resident simplex is indexed with a runtime slot
  if (i_s == 1) { s.v1[1] = v1; s.v2[1] = v2; s.v[1] = v; }
  else if (i_s == 2) { s.v1[2] = v1; s.v2[2] = v2; s.v[2] = v; }
  else { s.v1[3] = v1; s.v2[3] = v2; s.v[3] = v; }
}

// mpr_discover_portal (mpr_decomp.py:475-620)
__device__ inline int mpr_discover(const gsr_model& m, int ga, int gb, const GPose& A, const GPose& Bp, MPRSimplex& s,
                                   f3 ca, f3 cb) {
  const float E = kCcdEps;
  s.v1[0] = ca; s.v2[0] = cb; s.v[0] = ca - cb; s.size = 1;
  if (fabsf(s.v[0].x) < E && fabsf(s.v[0].y) < E && fabsf(s.v[0].z) < E) s.v[0].x += 10.f * E;
  f3 dir = -normalized(s.v[0]);
  f3 v, v1, v2;
  dev_mpr_support(m, ga, gb, A, Bp, dir, v, v1, v2);
  s.v1[1] = v1; s.v2[1] = v2; s.v[1] = v; s.size = 2;
  float d = dot(v, dir);
  if (d < E) return -1;
  dir = cross(s.v[0], s.v[1]);
  if (dot(dir, dir) < E) {
    if (fabsf(s.v[1].x) < E && fabsf(s.v[1].y) < E && fabsf(s.v[1].z) < E) return 1;
    return 2;
  }
  dir = normalized(dir);
  dev_mpr_support(m, ga, gb, A, Bp, dir, v, v1, v2);
  d = dot(v, dir);
  if (d < E) return -1;
  s.v1[2] = v1; s.v2[2] = v2; s.v[2] = v; s.size = 3;
  dir = normalized(cross(s.v[1] - s.v[0], s.v[2] - s.v[0]));
  d = dot(dir, s.v[0]);
  if (d > 0.f) {
    f3 t;
    t = s.v1[1]; s.v1[1] = s.v1[2]; s.v1[2] = t;
    t = s.v2[1]; s.v2[1] = s.v2[2]; s.v2[2] = t;
    t = s.v[1]; s.v[1] = s.v[2]; s.v[2] = t;
    dir = -dir;
  }
  int trials = 0;
  int ret = 0;
  while (s.size < 4) {
    dev_mpr_support(m, ga, gb, A, Bp, dir, v, v1, v2);
    d = dot(v, dir);
    if (d < E) { ret = -1; break; }
    bool cont = false;
    d = dot(cross(s.v[1], v), s.v[0]);
    if (d < -E) { s.v1[2] = v1; s.v2[2] = v2; s.v[2] = v; cont = true; }
    if (!cont) {
      d = dot(cross(v, s.v[2]), s.v[0]);
      if (d < -E) { s.v1[1] = v1; s.v2[1] = v2; s.v[1] = v; cont = true; }
    }
    if (cont) {
      dir = normalized(cross(s.v[1] - s.v[0], s.v[2] - s.v[0]));
      trials++;
      if (trials == 15) { ret = -1; break; }
    } else {
      s.v1[3] = v1; s.v2[3] = v2; s.v[3] = v; s.size = 4;
    }
  }
  return ret;
}

__device__ inline void dev_mpr_contact(const gsr_model& m, int ga, int gb, const GPose& A, const GPose& Bp, f3 nws,
                                       bool& is_col, f3& normal, float& pen, f3& pos) {
  // guess_geoms_center (mpr_decomp.py:623-705)
  f3 ca = xform(T3(m.geom_center, ga), A.p, A.q);
  f3 cb = xform(T3(m.geom_center, gb), Bp.p, Bp.q);
  const float E = kCcdEps;
  if (fabsf(nws.x) > E || fabsf(nws.y) > E || fabsf(nws.z) > E) {
    f3 a0 = T3(m.geom_init_aabb, 8 * ga), a7 = T3(m.geom_init_aabb, 8 * ga + 7);
    f3 b0 = T3(m.geom_init_aabb, 8 * gb), b7 = T3(m.geom_init_aabb, 8 * gb + 7);
    ca = xform(0.5f * (a7 + a0), A.p, A.q);
    cb = xform(0.5f * (b7 + b0), Bp.p, Bp.q);
    f3 delta = ca - cb;
    f3 nrm = normalized(delta);
    if (norm(cross(nws, nrm)) > 0.01f) {
      f3 off = dot(delta, nws) * nws - delta;
      float on = norm(off);
      if (on > m.eps) {
        f3 dof = off / on;
        f3 dla = qrot_inv(dof, A.q), dlb = qrot_inv(dof, Bp.q);
        float la = dot(a7 - a0, vabs(dla)), lb = dot(b7 - b0, vabs(dlb));
        float ratio = fminf(on / (la + lb), 0.5f);
        ca = ca + dof * la * ratio;
        cb = cb - dof * lb * ratio;
      }
    }
  }
  MPRSimplex s;
  int res = mpr_discover(m, ga, gb, A, Bp, s, ca, cb);
  is_col = false;
  pos = zero3();
  normal = zero3();
  pen = 0.f;
  if (res == 1) {
    is_col = true;
    normal = -normalized(s.v[0]);
    pos = (s.v1[1] + s.v2[1]) * 0.5f;
  } else if (res == 2) {
    is_col = true;
    pen = norm(s.v[1]);
    normal = -normalized(s.v[1]);
    pos = (s.v1[1] + s.v2[1]) * 0.5f;
  } else if (res == 0) {
    // mpr_refine_portal
    int rr = -1;
    for (int it = 0; it < kMprLoopCap; it++) {
      f3 dir = mpr_portal_dir(s);
      if (dot(s.v[1], dir) > -kCcdEps) { rr = 0; break; }
      f3 v, v1, v2;
      dev_mpr_support(m, ga, gb, A, Bp, dir, v, v1, v2);
      if (!(dot(v, dir) > -kCcdEps) || mpr_reach_tol(s, v, dir)) { rr = -1; break; }
      mpr_expand(s, v, v1, v2);
    }
    if (rr >= 0) {
      // mpr_find_penetration
      int iters = 0;
      for (int guard = 0; guard < kMprLoopCap; guard++) {
        f3 dir = mpr_portal_dir(s);
        f3 v, v1, v2;
        dev_mpr_support(m, ga, gb, A, Bp, dir, v, v1, v2);
        if (mpr_reach_tol(s, v, dir) || iters > kCcdIters) {
          pen = dot(dir, s.v[1]);
          normal = -dir;
          is_col = true;
          // mpr_find_pos, non-compatibility branch: weights from the portal direction
          float bw[4];
          bw[0] = 0.f;
          for (int i = 1; i < 4; i++) {
            int i1 = i % 3 + 1, i2 = (i + 1) % 3 + 1;
            bw[i] = dot(cross(s.v[i1], s.v[i2]), dir);
          }
          float sum = bw[0] + bw[1] + bw[2] + bw[3];
          f3 p1 = zero3(), p2 = zero3();
          for (int i = 0; i < 4; i++) {
            p1 = p1 + bw[i] * s.v1[i];
            p2 = p2 + bw[i] * s.v2[i];
          }
          pos = (0.5f / sum) * (p1 + p2);
          break;
        }
        mpr_expand(s, v, v1, v2);
        iters++;
      }
    }
  }
}

__device__ inline void dev_rotate_frame(GPose& P, f3 cpos, q4 rq) {
  P.q = qrot_q(P.q, rq);
  f3 rel = cpos - P.p;
  f3 vec = qrot(rel, rq) - rel;
  P.p = P.p - vec;
}

__device__ inline void dev_contact_orthogonals(const Env& e, int ga, int gb, f3 normal, f3& a0, f3& a1) {
  const gsr_model& m = e.m;
  int g = gb;
  if (m.geom_type[ga] != GEOM_PLANE) {
    f3 sa = T3(m.geom_init_aabb, 8 * ga + 7), sb = T3(m.geom_init_aabb, 8 * gb + 7);
    float va = sa.x * sa.y * sa.z, vb = sb.x * sb.y * sb.z;
    g = va < vb ? ga : gb;
  }
  int l = m.geom_link[g];
  m3 R = quat_to_R(e.L4(e.s.links_iquat, l), m.eps);
  int idx = 0;
  float amax = 0.f;
  for (int i = 0; i < 3; i++) {
    float a = fabsf(dot(mcol(R, i), normal));
    if (a > amax) { amax = a; idx = i; }
  }
  idx = (idx + 1) % 3;
  a0 = mcol(R, idx);
  a0 = normalized(a0 - dot(normal, a0) * normal);
  a1 = cross(normal, a0);
}

// func_convex_convex_contact (collider_decomp.py:2328-2725); MPR path with multi-contact
template <class Sink>
__device__ inline void dev_convex_convex(const Env& e, int ga, int gb, Sink& snk) {
  const gsr_model& m = e.m;
  const gsr_state& s = e.s;
  const float EPS = m.eps;
  int ta = m.geom_type[ga], tb = m.geom_type[gb];
  bool multi = m.enable_multi_contact && ta != GEOM_SPHERE && ta != GEOM_ELLIPSOID && tb != GEOM_SPHERE && tb != GEOM_ELLIPSOID;
  float tol = dev_tolerance(m, ga, gb, m.mc_tolerance);
  const GPose A0{e.L3(s.geoms_pos, ga), e.L4(s.geoms_quat, ga)};
  const GPose B0{e.L3(s.geoms_pos, gb), e.L4(s.geoms_quat, gb)};
  bool is_col0 = false, is_col = false;
  float pen0 = 0.f, pen = 0.f;
  f3 n0 = zero3(), cp0 = zero3(), normal = zero3(), cpos = zero3(), ax0 = zero3(), ax1 = zero3();
  q4 qr{0.f, 0.f, 0.f, 0.f};
  int n_con = 0;
  int gmin = ga < gb ? ga : gb, gmax = ga < gb ? gb : ga;
  int i_pair = m.pair_idx[(size_t)gmin * m.n_geoms + gmax];
  // the pair's normal cache, read once and written back at the end (det 0 sets it, later detections read it)
  f3 cache = e.L3(s.cache_normal, i_pair);
  for (int det = 0; det < 5; det++) {
    GPose A = A0, Bp = B0;
    if (multi && is_col0) {
      f3 axis = (float)(2 * (det % 2) - 1) * ax0 + (float)(1 - 2 * ((det / 2) % 2)) * ax1;
      qr = rotvec_to_quat(m.mc_perturbation * axis, EPS);
      dev_rotate_frame(A, cp0, qr);
      dev_rotate_frame(Bp, cp0, qinv(qr));
    }
    if ((multi && is_col0) || det == 0) {
      if (ta == GEOM_PLANE) {
        f3 pd = qrot(f3{m.geom_data[7 * ga], m.geom_data[7 * ga + 1], m.geom_data[7 * ga + 2]}, A.q);
        normal = -normalized(pd);
        f3 v1 = dev_support(m, gb, Bp, normal);
        pen = dot(normal, v1 - A.p);
        cpos = v1 - 0.5f * pen * normal;
        is_col = pen > 0.f;
      } else {
        bool updated = false;
        f3 nws = cache;
        bool guess = fabsf(nws.x) > EPS || fabsf(nws.y) > EPS || fabsf(nws.z) > EPS;
        for (int im = 0; im < 2; im++) {
          if (im == 1 && det == 0 && !is_col && guess) {
            nws = zero3();
            guess = false;
            updated = false;
          }
          if (!updated) {
            dev_mpr_contact(m, ga, gb, A, Bp, nws, is_col, normal, pen, cpos);
            updated = true;
          }
        }
        if (pen > tol) {
          bool prefer_gjk = !guess || (m.mc_tolerance * pen >= m.mpr_to_gjk * tol);
          if (prefer_gjk) {
            if (!snk.gjk_allowed()) {
              snk.defer(ga, gb);
              return;
            }
            snk.flag(GSR_ERR_DEEP_PENETRATION);
            float gpen;
            f3 gn, gp;
            is_col = dev_gjk_contact(m, gjk_ws(e), ga, gb, A, Bp, gpen, gn, gp);
            pen = gpen;
            if (is_col) {
              cpos = gp;
              normal = gn;
            }
          }
        }
      }
    }
    if (det == 0) {
      is_col0 = is_col; n0 = normal; pen0 = pen; cp0 = cpos;
      if (is_col0) {
        snk.add(ga, gb, n0, cp0, pen0);
        if (multi) {
          dev_contact_orthogonals(e, ga, gb, normal, ax0, ax1);
          n_con = 1;
        }
        cache = normal;
      } else {
        cache = zero3();
      }
    } else if (multi && is_col0 && is_col) {
      f3 cpa = qrot((cpos - 0.5f * pen * normal) - cp0, qinv(qr)) + cp0;
      f3 cpb = qrot((cpos + 0.5f * pen * normal) - cp0, qr) + cp0;
      cpos = 0.5f * (cpa + cpb);
      f3 tw = cross(normal, n0);
      float mp = m.mc_perturbation;
      tw = f3{fminf(fmaxf(tw.x, -mp), mp), fminf(fmaxf(tw.y, -mp), mp), fminf(fmaxf(tw.z, -mp), mp)};
      normal = normal + cross(tw, normal);
      pen = dot(normal, cpb - cpa);
      bool repeated = false;
      for (int ic = 0; ic < n_con; ic++) {
        if (!repeated) {
          if (norm(cpos - snk.recent_pos(ic)) < tol) repeated = true;
        }
      }
      if (!repeated && pen > -tol) {
        pen = fmaxf(pen, 0.f);
        snk.add(ga, gb, normal, cpos, pen);
        n_con++;
      }
    }
  }
  e.S3(s.cache_normal, i_pair, cache);
}

template <class Sink>
__device__ inline void dev_plane_box(const Env& e, int ga, int gb, Sink& snk) {
  const gsr_model& m = e.m;
  const gsr_state& s = e.s;
  f3 gap = e.L3(s.geoms_pos, ga);
  q4 gaq = e.L4(s.geoms_quat, ga);
  GPose Bp{e.L3(s.geoms_pos, gb), e.L4(s.geoms_quat, gb)};
  f3 pd = qrot(f3{m.geom_data[7 * ga], m.geom_data[7 * ga + 1], m.geom_data[7 * ga + 2]}, gaq);
  f3 normal = -normalized(pd);
  f3 v1 = dev_support(m, gb, Bp, normal);
  float pen = dot(normal, v1 - gap);
  if (pen > 0.f) {
    f3 cpos = v1 - 0.5f * pen * normal;
    snk.add(ga, gb, normal, cpos, pen);
    if (m.enable_multi_contact) {
      int n_con = 1;
      f3 cp0 = cpos;
      float tol = dev_tolerance(m, ga, gb, m.mc_tolerance);
      for (int iv = m.geom_vert_start[gb]; iv < m.geom_vert_end[gb]; iv++) {
        if (n_con < m.n_contacts_per_pair) {
          f3 pc = xform(T3(m.verts, iv), Bp.p, Bp.q);
          pen = dot(normal, pc - gap);
          if (pen > 0.f) {
            cpos = pc - 0.5f * pen * normal;
            if (norm(cpos - cp0) > tol) {
              snk.add(ga, gb, normal, cpos, pen);
              n_con++;
            }
          }
        }
      }
    }
  }
}

template <bool kGjk>
__device__ inline void dev_narrow_phase(const Env& e) {
  const gsr_model& m = e.m;
  const gsr_state& s = e.s;
  int nb = e.I(s.n_broad, 0);
  for (int ip = 0; ip < nb; ip++) {
    int p = e.I(s.broad_pairs, ip);
    int ga = p >> 16, gb = p & 0xffff;
    if (m.geom_type[ga] > m.geom_type[gb]) { int t = ga; ga = gb; gb = t; }
    GlobalSink<kGjk> snk(e);
    if (!(m.geom_type[ga] == GEOM_PLANE && m.geom_type[gb] == GEOM_BOX)) dev_convex_convex(e, ga, gb, snk);
  }
  for (int ip = 0; ip < nb; ip++) {
    int p = e.I(s.broad_pairs, ip);
    int ga = p >> 16, gb = p & 0xffff;
    if (m.geom_type[ga] > m.geom_type[gb]) { int t = ga; ga = gb; gb = t; }
    GlobalSink<kGjk> snk(e);
    if (m.geom_type[ga] == GEOM_PLANE && m.geom_type[gb] == GEOM_BOX) dev_plane_box(e, ga, gb, snk);
  }
}

template <bool kGjk = false>
__device__ inline void dev_collision(const Env& e) {
  e.I(e.s.n_contacts, 0) = 0;
  e.I(e.s.n_broad, 0) = 0;
  e.I(e.s.gjk_defer_n, 0) = 0;
  if (!e.m.enable_collision) return;
  dev_update_aabbs(e);
  if (e.m.n_pairs > 0) {
    dev_broad_phase(e);
    dev_narrow_phase<kGjk>(e);
  }
}

}  // namespace gsr
