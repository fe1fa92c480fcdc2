// GJK + EPA for deep penetrations, the "safe" (non-MuJoCo-compatible) path of gjk_decomp.py, run by the convex
// narrowphase when MPR reports a penetration beyond the multi-contact tolerance and either no cached normal was
// available or the overlap is large (collider_decomp.py:2498-2605).
//
//   func_safe_gjk :3656-3856 (initial tetrahedron from +-z / +-y supports, then face walking towards the origin),
//   func_safe_gjk_support :4167-4280 (support with EPS-perturbed directions to avoid degenerate simplices),
//   func_search_valid_simplex_vertex :3982-4076, func_safe_epa_init :4604-4653, func_safe_epa :4338-4540,
//   func_safe_attach_face_to_polytope :4657-4739, func_plane_normal :4742-4780, func_epa_horizon :1317-1383,
//   func_safe_epa_witness :4543-4600, func_triangle_affine_coords :2037-2094, func_gjk_contact :384-417.
//
// The simplex (4 vertices), the polytope (55 vertices, 300 faces) and the horizon buffers live in a per-env workspace
// of global memory (gsr_state.gjk_ws, env-major), used by one pair at a time; keeping them out of registers keeps
// this rarely taken path from raising the register count of the collision kernels that contain it.
#pragma once
#include "rigid_env.h"
#include "rigid_support.h"

namespace gsr {

constexpr int kGjkIters = 50, kEpaIters = 50, kPolyFaces = 6 * kEpaIters, kPolyVerts = 5 + kEpaIters;
constexpr int kHorizon = 6 + kEpaIters, kHStack = 3 * kPolyFaces;
// workspace words per env: polytope vertices (9 floats + 2 ints), faces (4 floats + 7 ints), face map, horizon,
// stack, simplex vertices (9 floats + 2 ints)
constexpr int kGjkWsWords = kPolyVerts * 11 + kPolyFaces * 12 + kHorizon * 2 + kHStack * 2 + 4 * 11;

constexpr float kGjkFloatMin = 1e-15f, kGjkFloatMax = 1e15f, kGjkFloatMinSq = 1e-30f, kGjkFloatMaxSq = 1e30f;
constexpr float kGjkTolerance = 1e-6f, kGjkDegenSq = 1e-10f, kGjkMaxReproj = 1e-5f;

struct GVert {
  f3 o1, o2, mk;
  int id1, id2;
};

// one env's slice of the workspace
struct GjkWs {
  float* pv;   // [kPolyVerts][9]: o1, o2, mk
  int* pvid;   // [kPolyVerts][2]
  float* fn;   // [kPolyFaces][4]: normal, dist2
  int* fi;     // [kPolyFaces][7]: verts[3], adj[3], map_idx
  int* fmap;   // [kPolyFaces]
  int* hd;     // [kHorizon][2]: face, edge
  int* st;     // [kHStack][2]
  float* sv;   // [4][9]: simplex o1, o2, mk
  int* sid;    // [4][2]
  __device__ f3 mk(int v) const { return f3{pv[9 * v + 6], pv[9 * v + 7], pv[9 * v + 8]}; }
  __device__ f3 o1(int v) const { return f3{pv[9 * v], pv[9 * v + 1], pv[9 * v + 2]}; }
  __device__ f3 o2(int v) const { return f3{pv[9 * v + 3], pv[9 * v + 4], pv[9 * v + 5]}; }
  __device__ void set_v(int v, const GVert& g) const {
    pv[9 * v] = g.o1.x; pv[9 * v + 1] = g.o1.y; pv[9 * v + 2] = g.o1.z;
    pv[9 * v + 3] = g.o2.x; pv[9 * v + 4] = g.o2.y; pv[9 * v + 5] = g.o2.z;
    pv[9 * v + 6] = g.mk.x; pv[9 * v + 7] = g.mk.y; pv[9 * v + 8] = g.mk.z;
    pvid[2 * v] = g.id1;
    pvid[2 * v + 1] = g.id2;
  }
  __device__ f3 fnorm(int f) const { return f3{fn[4 * f], fn[4 * f + 1], fn[4 * f + 2]}; }
  __device__ float& dist2(int f) const { return fn[4 * f + 3]; }
  __device__ int& fv(int f, int k) const { return fi[7 * f + k]; }
  __device__ int& fadj(int f, int k) const { return fi[7 * f + 3 + k]; }
  __device__ int& fmapidx(int f) const { return fi[7 * f + 6]; }
};

__device__ inline GjkWs gjk_ws(const Env& e) {
  int* base = reinterpret_cast<int*>(e.s.gjk_ws) + (size_t)e.b * kGjkWsWords;
  GjkWs w;
  w.pv = reinterpret_cast<float*>(base);
  w.pvid = base + kPolyVerts * 9;
  w.fn = reinterpret_cast<float*>(w.pvid + kPolyVerts * 2);
  w.fi = reinterpret_cast<int*>(w.fn + kPolyFaces * 4);
  w.fmap = w.fi + kPolyFaces * 7;
  w.hd = w.fmap + kPolyFaces;
  w.st = w.hd + kHorizon * 2;
  w.sv = reinterpret_cast<float*>(w.st + kHStack * 2);
  w.sid = reinterpret_cast<int*>(w.sv + 4 * 9);
  return w;
}

struct GSimplex {
  GjkWs w;  // vertex storage: w.sv, w.sid
  int ns;
  int last_searched;
  __device__ f3 mk(int i) const { return f3{w.sv[9 * i + 6], w.sv[9 * i + 7], w.sv[9 * i + 8]}; }
  __device__ int id1(int i) const { return w.sid[2 * i]; }
  __device__ int id2(int i) const { return w.sid[2 * i + 1]; }
  __device__ GVert get(int i) const {
    GVert g;
    g.o1 = f3{w.sv[9 * i], w.sv[9 * i + 1], w.sv[9 * i + 2]};
    g.o2 = f3{w.sv[9 * i + 3], w.sv[9 * i + 4], w.sv[9 * i + 5]};
    g.mk = mk(i);
    g.id1 = id1(i);
    g.id2 = id2(i);
    return g;
  }
  __device__ void set(int i, const GVert& g) const {
    float* v = w.sv + 9 * i;
    v[0] = g.o1.x; v[1] = g.o1.y; v[2] = g.o1.z;
    v[3] = g.o2.x; v[4] = g.o2.y; v[5] = g.o2.z;
    v[6] = g.mk.x; v[7] = g.mk.y; v[8] = g.mk.z;
    w.sid[2 * i] = g.id1;
    w.sid[2 * i + 1] = g.id2;
  }
};

__device__ __forceinline__ bool gjk_discrete(int t) { return t == GEOM_BOX || t == GEOM_MESH; }

// support_driver (gjk_decomp.py:3605-3652); boxes also return the vertex id (support_field_decomp.py:289-311)
__device__ inline void gjk_support_one(const gsr_model& m, int g, const GPose& P, f3 d, f3& v, int& vid) {
  vid = -1;
  if (m.geom_type[g] == GEOM_BOX) {
    f3 db = qrot_inv(d, P.q);
    const float* data = m.geom_data + 7 * g;
    f3 vl{(db.x < 0.f ? -1.f : 1.f) * data[0] * 0.5f, (db.y < 0.f ? -1.f : 1.f) * data[1] * 0.5f,
          (db.z < 0.f ? -1.f : 1.f) * data[2] * 0.5f};
    vid = (vl.x > 0.f ? 1 : 0) + (vl.y > 0.f ? 2 : 0) + (vl.z > 0.f ? 4 : 0) + m.geom_vert_start[g];
    v = xform(vl, P.p, P.q);
  } else if (m.geom_type[g] == GEOM_MESH) {
    // the support field, not an exhaustive vertex search (that one is the MuJoCo-compatibility path)
    v = dev_support_mesh(m, g, P, d, vid);
  } else {
    v = dev_support(m, g, P, d);
  }
}

__device__ inline GVert gjk_support(const gsr_model& m, int ga, int gb, const GPose& A, const GPose& B, f3 d) {
  GVert r;
  gjk_support_one(m, ga, A, d, r.o1, r.id1);
  gjk_support_one(m, gb, B, -d, r.o2, r.id2);
  r.mk = r.o1 - r.o2;
  return r;
}

__device__ inline int gjk_count_supports(const gsr_model& m, int g, const GPose& P, f3 d) {
  if (m.geom_type[g] != GEOM_BOX) return 1;
  f3 db = qrot_inv(d, P.q);
  int z = (db.x == 0.f) + (db.y == 0.f) + (db.z == 0.f);
  return 1 << z;
}

__device__ inline bool gjk_duplicate(const GSimplex& S, int id1, int id2) {
  for (int i = 0; i < S.ns; i++) {
    if (id1 == -1 || S.id1(i) != id1) continue;
    if (id2 == -1 || S.id2(i) != id2) continue;
    return true;
  }
  return false;
}

__device__ inline bool gjk_degenerate(const GSimplex& S, f3 mk) {
  for (int i = 0; i < S.ns; i++) {
    f3 d = S.mk(i) - mk;
    if (dot(d, d) < kGjkDegenSq) return true;
  }
  if (S.ns == 2) {
    f3 e1 = S.mk(1) - S.mk(0), e2 = mk - S.mk(0);
    f3 n = cross(e1, e2);
    return dot(n, n) < kGjkDegenSq * dot(e1, e1) * dot(e2, e2);
  }
  if (S.ns == 3) {
    f3 e1 = normalized(S.mk(1) - S.mk(0)), e2 = normalized(S.mk(2) - S.mk(0));
    f3 n = cross(e1, e2);
    f3 diff = mk - S.mk(0);
    float nd = dot(n, diff);
    return nd * nd < kGjkDegenSq * dot(n, n) * dot(diff, diff);
  }
  return false;
}

__device__ __forceinline__ bool gjk_valid(const GSimplex& S, const GVert& v) {
  return !gjk_duplicate(S, v.id1, v.id2) && !gjk_degenerate(S, v.mk);
}

__device__ inline GVert gjk_safe_support(const gsr_model& m, const GSimplex& S, int ga, int gb, const GPose& A,
                                         const GPose& B, f3 dir) {
  const float EPS = m.eps;
  GVert r{};
  r.id1 = r.id2 = -1;
  #pragma unroll 1
  for (int i = 0; i < 9; i++) {
    f3 nd = dir;
    if (i > 0) {
      int j = i - 1;
      // the reference scales by 1 - 2 * (j & bit) with the bit's value, not a 0/1 flag
      nd.x += -(float)(1 - 2 * (j & 1)) * EPS;
      nd.y += -(float)(1 - 2 * (j & 2)) * EPS;
      nd.z += -(float)(1 - 2 * (j & 4)) * EPS;
    }
    nd = nd * (2.f - dot(nd, dir));
    int num = gjk_count_supports(m, ga, A, nd) * gjk_count_supports(m, gb, B, -nd);
    if (i > 0 && num > 1) continue;
    r = gjk_support(m, ga, gb, A, B, nd);
    if (i == 0) {
      if (num > 1) continue;
      break;
    }
    if (i == 8) break;
    if (gjk_valid(S, r)) break;
  }
  return r;
}

__device__ inline f3 gjk_discrete_vertex(const gsr_model& m, int g, const GPose& P, int iv) {
  f3 vl;
  if (m.geom_type[g] == GEOM_BOX) {
    const float* data = m.geom_data + 7 * g;
    vl = f3{((iv & 1) ? 1.f : -1.f) * data[0] * 0.5f, ((iv & 2) ? 1.f : -1.f) * data[1] * 0.5f,
            ((iv & 4) ? 1.f : -1.f) * data[2] * 0.5f};
  } else {
    vl = T3(m.verts, m.geom_vert_start[g] + iv);
  }
  return xform(vl, P.p, P.q);
}

__device__ inline bool gjk_search_valid(const gsr_model& m, GSimplex& S, int ga, int gb, const GPose& A,
                                        const GPose& B, GVert& out) {
  if (gjk_discrete(m.geom_type[ga]) && gjk_discrete(m.geom_type[gb])) {
    int n0 = m.geom_vert_end[ga] - m.geom_vert_start[ga], n1 = m.geom_vert_end[gb] - m.geom_vert_start[gb];
    int cases = n0 * n1;
    #pragma unroll 1
    for (int k = 0; k < cases; k++) {
      int mm = (k + S.last_searched) % cases;
      int i = mm / n1, j = mm % n1;
      out.id1 = m.geom_vert_start[ga] + i;
      out.id2 = m.geom_vert_start[gb] + j;
      out.o1 = gjk_discrete_vertex(m, ga, A, i);
      out.o2 = gjk_discrete_vertex(m, gb, B, j);
      out.mk = out.o1 - out.o2;
      if (gjk_valid(S, out)) {
        S.last_searched = (mm + 1) % cases;
        return true;
      }
    }
    return false;
  }
  if (S.ns == 3) {
    f3 dir = normalized(cross(S.mk(2) - S.mk(0), S.mk(1) - S.mk(0)));
    #pragma unroll 1
    for (int i = 0; i < 2; i++) {
      out = gjk_safe_support(m, S, ga, gb, A, B, i == 0 ? dir : -dir);
      if (gjk_valid(S, out)) return true;
    }
  }
  return false;
}

// true when the origin is inside the final simplex (the shapes intersect)
__device__ inline bool gjk_safe(const gsr_model& m, GSimplex& S, int ga, int gb, const GPose& A, const GPose& B) {
  bool init_ok = true;
  S.ns = 0;
  #pragma unroll 1
  for (int i = 0; i < 4; i++) {
    f3 dir = zero3();
    float val = (float)(1 - 2 * (i % 2));
    if (i < 2) dir.z = val;
    else dir.y = val;
    GVert v = gjk_safe_support(m, S, ga, gb, A, B, dir);
    if (!gjk_valid(S, v)) {
      if (!gjk_search_valid(m, S, ga, gb, A, B, v)) {
        init_ok = false;
        break;
      }
    }
    S.set(i, v);
    S.ns++;
  }
  if (!init_ok) return false;
  #pragma unroll 1
  for (int it = 0; it < kGjkIters; it++) {
    f3 best_n = zero3();
    float best_sd = 0.f;
    int mi = 0;
    #pragma unroll 1
    for (int j = 0; j < 4; j++) {
      int s0 = 2, s1 = 1, s2 = 3, ap = 0;
      if (j == 1) { s0 = 0; s1 = 2; s2 = 3; ap = 1; }
      else if (j == 2) { s0 = 1; s1 = 0; s2 = 3; ap = 2; }
      else if (j == 3) { s0 = 0; s1 = 1; s2 = 2; ap = 3; }
      f3 v1 = S.mk(s0), v2 = S.mk(s1), v3 = S.mk(s2), va = S.mk(ap);
      f3 n = normalized(cross(v3 - v1, v2 - v1));
      if (dot(n, va - v1) > 0.f) n = -n;
      float sd = dot(n, v1);
      if (j == 0 || sd < best_sd) {
        best_sd = sd;
        best_n = n;
        mi = j;
      }
    }
    if (best_sd >= 0.f) return true;
    S.ns = 3;
    if (mi != 3) S.set(mi, S.get(3));
    GVert v = gjk_safe_support(m, S, ga, gb, A, B, best_n);
    if (gjk_duplicate(S, v.id1, v.id2)) return false;
    if (gjk_degenerate(S, v.mk)) return false;
    if (dot(v.mk, best_n) < 0.f) return false;
    S.set(3, v);
    S.ns = 4;
  }
  return false;
}

__device__ inline bool gjk_plane_normal(f3 v1, f3 v2, f3 v3, f3& normal) {
  f3 d21 = v2 - v1, d31 = v3 - v1, d32 = v3 - v2;
  #pragma unroll 1
  for (int i = 0; i < 3; i++) {
    f3 n = i == 0 ? cross(d32, d21) : (i == 1 ? cross(d21, d31) : cross(d31, d32));
    float nn = norm(n);
    if (nn == 0.f) return false;
    if (nn > kGjkFloatMin) {
      normal = normalized(n);
      return true;
    }
  }
  return false;
}

__device__ inline bool epa_attach(const GjkWs& w, int& nf, int npv, int v1, int v2, int v3, int a1, int a2, int a3) {
  int n = nf++;
  w.fv(n, 0) = v1; w.fv(n, 1) = v2; w.fv(n, 2) = v3;
  w.fadj(n, 0) = a1; w.fadj(n, 1) = a2; w.fadj(n, 2) = a3;
  f3 normal;
  bool ok = gjk_plane_normal(w.mk(v3), w.mk(v2), w.mk(v1), normal);
  if (ok) {
    f3 center = (w.mk(v1) + w.mk(v2) + w.mk(v3)) / 3.f;
    float max_orient = -dot(normal, center), max_abs = fabsf(max_orient);
    #pragma unroll 1
    for (int iv = 0; iv < npv; iv++) {
      if (iv == v1 || iv == v2 || iv == v3) continue;
      float orient = dot(normal, w.mk(iv) - center);
      if (fabsf(orient) > max_abs) {
        max_abs = fabsf(orient);
        max_orient = orient;
      }
    }
    if (max_orient > 0.f) normal = -normal;
    w.fn[4 * n] = normal.x; w.fn[4 * n + 1] = normal.y; w.fn[4 * n + 2] = normal.z;
    float md = kGjkFloatMax;
    for (int i = 0; i < 3; i++) {
      float d = dot(normal, w.mk(w.fv(n, i)));
      float d2 = d * d;
      if (d2 < md) md = d2;
    }
    w.dist2(n) = md;
    w.fmapidx(n) = -1;
  } else {
    // degenerate face: never the nearest (the reference leaves the slot's previous contents)
    w.fn[4 * n] = 0.f; w.fn[4 * n + 1] = 0.f; w.fn[4 * n + 2] = 0.f;
    w.dist2(n) = kGjkFloatMaxSq;
  }
  return ok;
}

__device__ inline void epa_delete_face(const GjkWs& w, int& nfmap, int i_f) {
  int idx = w.fmapidx(i_f);
  if (idx >= 0) {
    int last = w.fmap[nfmap - 1];
    w.fmap[idx] = last;
    w.fmapidx(last) = idx;
    nfmap--;
  }
  w.fmapidx(i_f) = -2;
}

// returns true on error (horizon not a closed loop, or a buffer full)
__device__ inline bool epa_horizon(const GjkWs& w, int& nfmap, int& nh, int nearest, f3 wv) {
  w.st[0] = nearest;
  w.st[1] = 0;
  int top = 1;
  bool first = true;
  while (top > 0) {
    int i_f = w.st[2 * (top - 1)], i_e = w.st[2 * (top - 1) + 1];
    f3 v = w.mk(w.fv(i_f, 0));
    top--;
    bool deleted = w.fmapidx(i_f) == -2;
    if (!first && deleted) continue;
    bool visible = dot(w.fnorm(i_f), wv - v) > kGjkFloatMin;
    if (visible || first) {
      epa_delete_face(w, nfmap, i_f);
      for (int k = first ? 0 : 1; k < 3; k++) {
        int e2 = (i_e + k) % 3;
        int adj = w.fadj(i_f, e2);
        if (w.fmapidx(adj) != -2) {
          int sv = w.fv(i_f, (e2 + 1) % 3);
          int ae = w.fv(adj, 0) == sv ? 0 : (w.fv(adj, 1) == sv ? 1 : 2);
          if (top >= kHStack) return true;
          w.st[2 * top] = adj;
          w.st[2 * top + 1] = ae;
          top++;
        }
      }
    } else {
      if (nh >= kHorizon) return true;
      w.hd[2 * nh] = i_f;
      w.hd[2 * nh + 1] = i_e;
      nh++;
    }
    first = false;
  }
  return false;
}

__device__ inline f3 gjk_project_origin_to_plane(f3 v1, f3 v2, f3 v3) {
  f3 point = zero3();
  f3 d21 = v2 - v1, d31 = v3 - v1, d32 = v3 - v2;
  #pragma unroll 1
  for (int i = 0; i < 3; i++) {
    f3 n = i == 0 ? cross(d32, d21) : (i == 1 ? cross(d21, d31) : cross(d31, d32));
    f3 v = i == 0 ? v2 : (i == 1 ? v1 : v3);
    float nv = dot(n, v), nn = dot(n, n);
    if (nn == 0.f) return point;
    if (nn > kGjkFloatMin) return n * (nv / nn);
    if (i == 2 && !(nn < kGjkFloatMin)) point = n * (nv / nn);
  }
  return point;
}

__device__ inline f3 gjk_triangle_affine_coords(f3 p, f3 t1, f3 t2, f3 t3) {
  float ms[3];
  for (int i = 0; i < 3; i++) {
    int i1 = (i + 1) % 3, i2 = (i + 2) % 3;
    if (i == 1) { int t = i1; i1 = i2; i2 = t; }
    ms[i] = comp(t2, i1) * comp(t3, i2) - comp(t2, i2) * comp(t3, i1) - comp(t1, i1) * comp(t3, i2) +
            comp(t1, i2) * comp(t3, i1) + comp(t1, i1) * comp(t2, i2) - comp(t1, i2) * comp(t2, i1);
  }
  float m_max = 0.f;
  int ix = 0, iy = 0;
  for (int i = 0; i < 3; i++) {
    float a = fabsf(ms[i]);
    if (a >= fabsf(ms[(i + 1) % 3]) && a >= fabsf(ms[(i + 2) % 3])) {
      m_max = ms[i];
      ix = (i + 1) % 3;
      iy = (i + 2) % 3;
      if (i == 1) { int t = ix; ix = iy; iy = t; }
      break;
    }
  }
  float cs[3];
  for (int i = 0; i < 3; i++) {
    f3 tv1 = i == 0 ? t2 : (i == 1 ? t3 : t1), tv2 = i == 0 ? t3 : (i == 1 ? t1 : t2);
    cs[i] = comp(p, ix) * comp(tv1, iy) + comp(p, iy) * comp(tv2, ix) + comp(tv1, ix) * comp(tv2, iy) -
            comp(p, ix) * comp(tv2, iy) - comp(p, iy) * comp(tv1, ix) - comp(tv2, ix) * comp(tv1, iy);
  }
  return f3{cs[0] / m_max, cs[1] / m_max, cs[2] / m_max};
}

__device__ inline bool epa_witness(const GjkWs& w, int i_f, f3& w1, f3& w2) {
  int i1 = w.fv(i_f, 0), i2 = w.fv(i_f, 1), i3 = w.fv(i_f, 2);
  f3 v1 = w.mk(i1), v2 = w.mk(i2), v3 = w.mk(i3);
  f3 proj = gjk_project_origin_to_plane(v1, v2, v3);
  f3 l = gjk_triangle_affine_coords(proj, v1, v2, v3);
  f3 pl = v1 * l.x + v2 * l.y + v3 * l.z;
  float err = norm(proj - pl);
  float me = fmaxf(fmaxf(dot(v1 - v2, v1 - v2), dot(v2 - v3, v2 - v3)), fmaxf(dot(v3 - v1, v3 - v1), kGjkFloatMinSq));
  float inv = 1.f / sqrtf(me);
  if (err * inv > kGjkMaxReproj) return false;
  w1 = w.o1(i1) * l.x + w.o1(i2) * l.y + w.o1(i3) * l.z;
  w2 = w.o2(i1) * l.x + w.o2(i2) * l.y + w.o2(i3) * l.z;
  return true;
}

// func_safe_epa_init + func_safe_epa; returns the signed distance (negative: penetration) and the witness points
__device__ inline float epa_safe(const gsr_model& m, const GjkWs& w, const GSimplex& S, int ga, int gb,
                                 const GPose& A, const GPose& B, int& n_witness, f3& w1, f3& w2) {
  const float EPS = m.eps;
  int npv = 0, nf = 0, nfmap = 0, nh = 0;
  for (int i = 0; i < 4; i++) w.set_v(npv++, S.get(i));
  const int fvx[4][3] = {{0, 1, 2}, {0, 3, 1}, {0, 2, 3}, {3, 2, 1}};
  const int fax[4][3] = {{1, 3, 2}, {2, 3, 0}, {0, 3, 1}, {2, 0, 1}};
  #pragma unroll 1
  for (int i = 0; i < 4; i++) epa_attach(w, nf, npv, fvx[i][0], fvx[i][1], fvx[i][2], fax[i][0], fax[i][1], fax[i][2]);
  for (int i = 0; i < 4; i++) {
    w.fmap[i] = i;
    w.fmapidx(i) = i;
  }
  nfmap = 4;
  float upper = kGjkFloatMax, upper2 = kGjkFloatMaxSq, lower = 0.f, tol = kGjkTolerance;
  int nearest = -1, prev = -1;
  bool discrete = gjk_discrete(m.geom_type[ga]) && gjk_discrete(m.geom_type[gb]);
  if (discrete) tol = EPS;
  #pragma unroll 1
  for (int k = 0; k < kEpaIters; k++) {
    prev = nearest;
    float lower2 = kGjkFloatMaxSq;
    for (int i = 0; i < nfmap; i++) {
      int i_f = w.fmap[i];
      if (w.dist2(i_f) < lower2) {
        lower2 = w.dist2(i_f);
        nearest = i_f;
      }
    }
    if (lower2 > upper2 || nearest == -1) {
      nearest = prev;
      break;
    }
    lower = sqrtf(lower2);
    f3 dir = w.fnorm(nearest);
    int wi = npv;
    w.set_v(npv++, gjk_support(m, ga, gb, A, B, dir / 1.f));
    f3 wv = w.mk(wi);
    float upper_k = dot(wv, dir);
    if (upper_k < upper) {
      upper = upper_k;
      upper2 = upper * upper;
    }
    if (upper - lower < tol) break;
    if (discrete) {
      bool repeated = false;
      int id1 = w.pvid[2 * wi], id2 = w.pvid[2 * wi + 1];
      for (int i = 0; i < npv; i++) {
        if (i == wi) continue;
        if (w.pvid[2 * i] == id1 && w.pvid[2 * i + 1] == id2) {
          repeated = true;
          break;
        }
      }
      if (repeated) break;
    }
    if (epa_horizon(w, nfmap, nh, nearest, wv)) {
      nearest = -1;
      break;
    }
    if (nh < 3) {
      nearest = -1;
      break;
    }
    int nfaces = nf, nedges = nh;
    if (nfaces + nedges >= kPolyFaces) break;
    bool ok = true;
    #pragma unroll 1
    for (int i = 0; i < nedges; i++) {
      int f0 = nfaces + i, f1 = nfaces + (i + 1) % nedges;
      int hf = w.hd[2 * i], he = w.hd[2 * i + 1];
      int hv1 = w.fv(hf, he), hv2 = w.fv(hf, (he + 1) % 3);
      w.fadj(hf, he) = f0;
      int adj0 = i > 0 ? f0 - 1 : nfaces + nedges - 1;
      ok = epa_attach(w, nf, npv, wi, hv2, hv1, f1, hf, adj0);
      if (!ok) break;
      float d2 = w.dist2(nf - 1);
      if (d2 >= lower2 - EPS && d2 <= upper2 + EPS) {
        w.fmap[nfmap] = f0;
        w.fmapidx(f0) = nfmap;
        nfmap++;
      }
    }
    if (!ok) break;
    nh = 0;
    if (nfmap == 0 || nearest == -1) break;
  }
  n_witness = 0;
  if (nearest != -1) {
    float d2 = w.dist2(nearest);
    if (epa_witness(w, nearest, w1, w2)) {
      n_witness = 1;
      return -sqrtf(d2);
    }
  }
  return 0.f;
}

// func_gjk_contact, safe path: is_col with penetration, normal and contact position
__device__ inline bool dev_gjk_contact(const gsr_model& m, const GjkWs& w, int ga, int gb, const GPose& A,
                                       const GPose& B, float& pen, f3& normal, f3& pos) {
  GSimplex S;
  S.w = w;
  S.ns = 0;
  S.last_searched = 0;
  float distance = kGjkFloatMax;
  int n_witness = 0;
  f3 w1 = zero3(), w2 = zero3();
  if (gjk_safe(m, S, ga, gb, A, B)) distance = epa_safe(m, w, S, ga, gb, A, B, n_witness, w1, w2);
  bool is_col = distance < 0.f;
  pen = is_col ? -distance : 0.f;
  int n = 0;
  if (is_col) {
    for (int i = 0; i < n_witness; i++) {
      f3 p = 0.5f * (w1 + w2);
      f3 nn = w2 - w1;
      float len = norm(nn);
      if (len < kGjkFloatMin) continue;
      normal = nn / len;
      pos = p;
      n++;
    }
  }
  if (n == 0) {
    is_col = false;
    pen = 0.f;
  }
  return is_col;
}

}  // namespace gsr
