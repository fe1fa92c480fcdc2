// Constraint rows and the Newton solver of one environment.
//
// Rows are laid out [equality | frictionloss | 4 pyramid rows per contact | joint limits] exactly as
// constraint_solver_decomp.py builds them (add_equality_constraints :721-778, add_frictionloss_constraints
// :1075-1124, add_collision_constraints :428-525, add_joint_limit_constraints :1017-1072). Each row is stored sparse:
// up to row_width (dof, value) pairs sorted by dof. Every dense sum of the reference over all dofs is evaluated over
// the stored nonzeros in ascending dof order, which skips only exact +0 terms, so the sums round identically.
// The solver restates func_init_solver (:2047-2429, direct Hessian path), func_solve (:1409-1437),
// func_solve_body (:1795-1894), the exact line search (:1440-1792), the rank-1 Cholesky updates (:1127-1236) and
// func_update_qacc / func_update_contact_force (:1343-1406). The Newton Hessian is dense, packed lower triangle.
#pragma once
#include "rigid_env.h"

namespace gsr {

struct RowBuf {
  int idx[kMaxRowWidth];
  float val[kMaxRowWidth];
  int n;
};

__device__ inline void row_store(const Env& e, int c, const RowBuf& r) {
  const gsr_state& s = e.s;
  const int W = e.m.row_width;
  for (int k = 0; k < r.n; k++) {
    e.JI(c, k) = r.idx[k];
    e.JV(c, k) = r.val[k];
  }
  e.EI(s.jac_nnz, c) = r.n;
}

__device__ inline void row_add(RowBuf& r, int d, float v) {
  for (int k = 0; k < r.n; k++)
    if (r.idx[k] == d) {
      r.val[k] = r.val[k] + v;
      return;
    }
  r.idx[r.n] = d;
  r.val[r.n] = 0.f + v;
  r.n++;
}

__device__ inline void row_sort(RowBuf& r) {
  for (int i = 1; i < r.n; i++) {
    int ki = r.idx[i];
    float kv = r.val[i];
    int j = i - 1;
    while (j >= 0 && r.idx[j] > ki) {
      r.idx[j + 1] = r.idx[j];
      r.val[j + 1] = r.val[j];
      j--;
    }
    r.idx[j + 1] = ki;
    r.val[j + 1] = kv;
  }
}

// Row of component i3 of the velocity difference of two points, p1 fixed to link1 (sign +1) and p2 fixed to link2
// (sign -1): the dofs of both root chains, link by link from the link up, each link's dofs last first, as
// func_equality_connect walks them (constraint_solver_decomp.py:586-616); returns J qvel summed in that order.
__device__ inline float point_diff_row(const Env& e, int l1, f3 p1, int l2, f3 p2, int i3, RowBuf& r) {
  const gsr_model& m = e.m;
  const gsr_state& s = e.s;
  r.n = 0;
  float jac_qvel = 0.f;
  for (int iab = 0; iab < 2; iab++) {
    const float sign = iab == 0 ? 1.f : -1.f;
    int link = iab == 0 ? l1 : l2;
    const f3 p = iab == 0 ? p1 : p2;
    while (link > -1) {
      const int de = m.link_dof_end[link], nd = de - m.link_dof_start[link];
      const f3 tpos = p - e.L3(s.root_com, link);
      for (int k = 0; k < nd; k++) {
        const int d = de - 1 - k;
        const f3 vel = qrot(e.L3(s.cdof_vel, d) - cross(tpos, e.L3(s.cdof_ang, d)), qid());
        const float jac = comp(sign * vel, i3);
        jac_qvel = jac_qvel + jac * e.F(s.vel, d);
        row_add(r, d, jac);
      }
      link = m.link_parent[link];
    }
  }
  row_sort(r);
  return jac_qvel;
}

// stores row c with its impedance terms: imp_aref(sol_params, -pen, J qvel, pos), diag = invweight (1 - imp) / imp
__device__ inline void eq_row_finish(const Env& e, int c, const RowBuf& r, const float* sp, size_t sps, float pen,
                                     float jac_qvel, float pos, float invw) {
  const gsr_state& s = e.s;
  row_store(e, c, r);
  float imp, aref;
  imp_aref(sp, sps, -pen, jac_qvel, pos, imp, aref);
  const float diag = fmaxf(invw * (1.f - imp) / imp, e.m.eps);
  e.EF(s.efc_diag, c) = diag;
  e.EF(s.efc_aref, c) = aref;
  e.EF(s.efc_D, c) = 1.f / diag;
}

// CONNECT (func_equality_connect, constraint_solver_decomp.py:528-641): anchor data[0:3] on link1 and data[3:6] on
// link2 coincide; 3 rows from row c, returns the next row (solver parameters sp[k * sps])
__device__ inline int dev_connect_rows(const Env& e, int c, int l1, int l2, const float* data, const float* sp,
                                       size_t sps) {
  const gsr_model& m = e.m;
  const gsr_state& s = e.s;
  const f3 ga1 = xform(mk3(data[0], data[1], data[2]), e.L3(s.links_pos, l1), e.L4(s.links_quat, l1));
  const f3 ga2 = xform(mk3(data[3], data[4], data[5]), e.L3(s.links_pos, l2), e.L4(s.links_quat, l2));
  const float invw = e.LP(m.link_invweight, 2 * l1) + e.LP(m.link_invweight, 2 * l2);
  for (int i3 = 0; i3 < 3; i3++) {
    RowBuf r;
    const float jq = point_diff_row(e, l1, ga1, l2, ga2, i3, r);
    const f3 pd = ga1 - ga2;
    eq_row_finish(e, c + i3, r, sp, sps, norm(pd), jq, comp(pd, i3), invw);
  }
  return c + 3;
}

// q * (0, v): ti_quat_mul_axis (genesis/utils/geom.py:195-204)
__device__ __forceinline__ q4 quat_mul_axis(q4 q, f3 v) {
  return q4{-q.x * v.x - q.y * v.y - q.z * v.z, +q.w * v.x + q.y * v.z - q.z * v.y,
            +q.w * v.y + q.z * v.x - q.x * v.z, +q.w * v.z + q.x * v.y - q.y * v.x};
}

// angular Jacobian difference of the two chains at dof d (a dof of either chain): link1's chain adds cdof_ang, link2's
// subtracts it, summed from zero in that order (a dof on both chains cancels)
__device__ inline f3 weld_ang(const Env& e, int l1, int l2, int d) {
  const gsr_model& m = e.m;
  f3 a = zero3();
  for (int iab = 0; iab < 2; iab++)
    for (int link = iab == 0 ? l1 : l2; link > -1; link = m.link_parent[link])
      if (d >= m.link_dof_start[link] && d < m.link_dof_end[link]) {
        a = a + (iab == 0 ? 1.f : -1.f) * e.L3(e.s.cdof_ang, d);
        break;
      }
  return a;
}

// WELD (func_equality_weld, constraint_solver_decomp.py:824-1015): data[0:3] anchor on link2, data[3:6] anchor on
// link1, data[6:10] pose of link2 relative to link1, data[10] torque scale; 3 position rows then 3 orientation rows
// from row c, returns the next row (solver parameters sp[k * sps])
__device__ inline int dev_weld_rows(const Env& e, int c, int l1, int l2, const float* data, const float* sp,
                                    size_t sps) {
  const gsr_model& m = e.m;
  const gsr_state& s = e.s;
  const q4 q1 = e.L4(s.links_quat, l1), q2 = e.L4(s.links_quat, l2);
  const q4 rel = q4{data[6], data[7], data[8], data[9]};
  const float ts = data[10];
  const f3 ga1 = xform(mk3(data[3], data[4], data[5]), e.L3(s.links_pos, l1), q1);
  const f3 ga2 = xform(mk3(data[0], data[1], data[2]), e.L3(s.links_pos, l2), q2);
  const f3 pe = ga1 - ga2;
  const q4 q = qmul(q1, rel);
  const q4 iq2 = qinv(q2);
  const q4 eq = qmul(iq2, q);
  const f3 re = mk3(eq.x, eq.y, eq.z) * ts;
  // norm of the 6-vector (position error, orientation error)
  const float pos_imp = sqrtf(pe.x * pe.x + pe.y * pe.y + pe.z * pe.z + re.x * re.x + re.y * re.y + re.z * re.z);
  const float invw0 = e.LP(m.link_invweight, 2 * l1) + e.LP(m.link_invweight, 2 * l2);
  const float invw1 = e.LP(m.link_invweight, 2 * l1 + 1) + e.LP(m.link_invweight, 2 * l2 + 1);
  RowBuf r;
  for (int i = 0; i < 3; i++) {
    const float jq = point_diff_row(e, l1, ga1, l2, ga2, i, r);
    eq_row_finish(e, c + i, r, sp, sps, pos_imp, jq, comp(pe, i), invw0);
  }
  // orientation rows over the same sorted dofs (the dofs of both chains, which r holds now): the angular Jacobian
  // difference of the two chains rotated by inv(q2) ... q, one row at a time (the per-dof quaternion products are
  // recomputed per row: the same values)
  for (int i = 0; i < 3; i++) {
    float jq = 0.f;
    for (int k = 0; k < r.n; k++) {
      const int d = r.idx[k];
      const q4 q3 = qmul(quat_mul_axis(iq2, weld_ang(e, l1, l2, d)), q);
      const float jac = 0.5f * (i == 0 ? q3.x : i == 1 ? q3.y : q3.z) * ts;
      r.val[k] = jac;
      jq = jq + jac * e.F(s.vel, d);
    }
    eq_row_finish(e, c + 3 + i, r, sp, sps, pos_imp, jq, comp(re, i), invw1);
  }
  return c + 6;
}

// equality rows (add_equality_constraints, constraint_solver_decomp.py:778-822): the model's equalities in order,
// then the env's own welds (their solver parameters as kernel_add_weld_constraint sets them, :2475-2477)
__device__ inline void dev_equality_rows(const Env& e) {
  const gsr_model& m = e.m;
  const gsr_state& s = e.s;
  int n_con = 0;
  for (int ie = 0; ie < m.n_equalities; ie++) {
    size_t sps;
    const float* sp = e.eq_sp(ie, sps);
    const float* data = m.eq_data + 11 * ie;
    if (m.eq_type[ie] == EQ_CONNECT) {
      n_con = dev_connect_rows(e, n_con, m.eq_obj1[ie], m.eq_obj2[ie], data, sp, sps);
      continue;
    }
    if (m.eq_type[ie] == EQ_WELD) {
      n_con = dev_weld_rows(e, n_con, m.eq_obj1[ie], m.eq_obj2[ie], data, sp, sps);
      continue;
    }
    int j1 = m.eq_obj1[ie], j2 = m.eq_obj2[ie];
    int q1 = m.joint_q_start[j1], q2 = m.joint_q_start[j2], d1 = m.joint_dof_start[j1], d2 = m.joint_dof_start[j2];
    int c = n_con++;
    float pos1 = e.F(s.qpos, q1), pos2 = e.F(s.qpos, q2);
    float diff = pos2 - m.qpos0[q2], pos = pos1 - m.qpos0[q1], deriv = 0.f;
    for (int i5 = 0; i5 < 5; i5++) {
      float dp = crpow(diff, (float)i5);
      pos = pos - dp * data[i5];
      if (i5 < 4) deriv = deriv + data[i5 + 1] * dp * (float)(i5 + 1);
    }
    RowBuf r;
    r.n = 0;
    if (d1 < d2) {
      r.idx[0] = d1; r.val[0] = 1.f; r.idx[1] = d2; r.val[1] = -deriv;
    } else {
      r.idx[0] = d2; r.val[0] = -deriv; r.idx[1] = d1; r.val[1] = 1.f;
    }
    r.n = 2;
    row_store(e, c, r);
    float jac_qvel = 1.f * e.F(s.vel, d1) + -deriv * e.F(s.vel, d2);
    float invw = e.DP(m.dof_invweight, d1) + e.DP(m.dof_invweight, d2);
    float imp, aref;
    imp_aref(sp, sps, -fabsf(pos), jac_qvel, pos, imp, aref);
    float diag = fmaxf(invw * (1.f - imp) / imp, m.eps);
    e.EF(s.efc_diag, c) = diag;
    e.EF(s.efc_aref, c) = aref;
    e.EF(s.efc_D, c) = 1.f / diag;
  }
  const int n_dyn = m.max_dyn_welds > 0 ? e.I(s.eq_dyn_n, 0) : 0;
  for (int k = 0; k < n_dyn; k++) {
    float data[11];
    for (int j = 0; j < 11; j++) data[j] = e.F(s.eq_dyn_data, 11 * k + j);
    const float sp[7] = {2.f * m.dt, 1.f, 0.9f, 0.95f, 0.001f, 0.5f, 2.f};
    n_con = dev_weld_rows(e, n_con, e.I(s.eq_dyn_obj, 2 * k), e.I(s.eq_dyn_obj, 2 * k + 1), data, sp, 1);
  }
  e.I(s.n_con, 0) = n_con;
  e.I(s.n_con, 1) = n_con;
}

// frictionloss rows (add_frictionloss_constraints), appended from row n_con; returns the new row count
__device__ inline int dev_frictionloss_rows(const Env& e, int n_con, int& n_fl) {
  const gsr_model& m = e.m;
  const gsr_state& s = e.s;
  n_fl = 0;
  for (int l = 0; l < m.n_links; l++)
    for (int j = m.link_joint_start[l]; j < m.link_joint_end[l]; j++)
      for (int d = m.joint_dof_start[j]; d < m.joint_dof_end[j]; d++) {
        if (e.DP(m.dof_frictionloss, d) > m.eps) {
          float imp, aref;
          size_t sps;
          imp_aref(e.joint_sp(j, sps), sps, 0.f, 1.f * e.F(s.vel, d), 0.f, imp, aref);
          float diag = fmaxf(e.DP(m.dof_invweight, d) * (1.f - imp) / imp, m.eps);
          int c = n_con++;
          n_fl++;
          e.EF(s.efc_diag, c) = diag;
          e.EF(s.efc_aref, c) = aref;
          e.EF(s.efc_D, c) = 1.f / diag;
          e.EF(s.efc_fl, c) = e.DP(m.dof_frictionloss, d);
          RowBuf r;
          r.n = 1; r.idx[0] = d; r.val[0] = 1.f;
          row_store(e, c, r);
        }
      }
  return n_con;
}

// the 4 pyramid rows of contact ic, stored at rows c0 .. c0 + 3 (add_collision_constraints)
// Dofs of the chain from link l to its root: count, smallest and largest index, and whether the walk (the link's
// dofs last first, then its parent's, ...) visits them in strictly descending order.
__device__ __forceinline__ void chain_span(const gsr_model& m, int l, int& n, int& lo, int& hi, bool& desc) {
  n = 0;
  lo = 0x7fffffff;
  hi = -1;
  desc = true;
  int prev = 0x7fffffff;
  for (int link = l; link > -1; link = m.link_parent[link]) {
    const int ds = m.link_dof_start[link], de = m.link_dof_end[link];
    if (de <= ds) continue;
    desc = desc && de <= prev;
    prev = ds;
    n += de - ds;
    lo = min(lo, ds);
    hi = max(hi, de - 1);
  }
}

// The 4 pyramid rows of a contact whose two link chains have no dof in common and whose walks descend (every chain of
// a tree whose dofs grow from the root outwards): each row is then the two chains' dofs in ascending order, the lower
// chain first, so each walked dof has a known slot and the rows are written in place, all four from one walk of the
// chains, with the values the general build gives (0 + jac per entry, J qvel summed in walk order).
__device__ inline void contact_rows_disjoint(const Env& e, int c0, int la, int lb, f3 cp, const f3 (&n)[4], int na,
                                             int nb, int off_a, int off_b, float (&jq)[4]) {
  const gsr_model& m = e.m;
  const gsr_state& s = e.s;
  for (int i = 0; i < 4; i++) jq[i] = 0.f;
  for (int iab = 0; iab < 2; iab++) {
    const float sign = iab == 0 ? -1.f : 1.f;
    int slot = iab == 0 ? off_a + na - 1 : off_b + nb - 1;  // the walk fills the chain's slots from the top
    for (int link = iab == 0 ? la : lb; link > -1; link = m.link_parent[link]) {
      const int de = m.link_dof_end[link], nd = de - m.link_dof_start[link];
      const f3 tpos = cp - e.L3(s.root_com, link);
      for (int k = 0; k < nd; k++, slot--) {
        const int d = de - 1 - k;
        const f3 vel = qrot(e.L3(s.cdof_vel, d) - cross(tpos, e.L3(s.cdof_ang, d)), qid());
        const float qv = e.F(s.vel, d);
        for (int i = 0; i < 4; i++) {
          const float jac = dot(sign * vel, n[i]);
          jq[i] = jq[i] + jac * qv;
          e.JI(c0 + i, slot) = d;
          e.JV(c0 + i, slot) = 0.f + jac;
        }
      }
    }
  }
}

__device__ inline void dev_contact_rows(const Env& e, int ic, int c0) {
  const gsr_model& m = e.m;
  const gsr_state& s = e.s;
  int ga = e.I(s.contact_geoms, 2 * ic), gb = e.I(s.contact_geoms, 2 * ic + 1);
  int la = m.geom_link[ga], lb = m.geom_link[gb];
  f3 cn = e.L3(s.contact_normal, ic), cp = e.L3(s.contact_pos, ic);
  float fr = e.F(s.contact_friction, ic), pen = e.F(s.contact_pen, ic);
  float sp[7];
  for (int k = 0; k < 7; k++) sp[k] = 0.5f * (m.geom_sol_params[7 * ga + k] + m.geom_sol_params[7 * gb + k]);
  f3 d1, d2;
  orthogonals(cn, d1, d2);
  float invw = e.LP(m.link_invweight, 2 * la);
  if (lb > -1) invw = invw + e.LP(m.link_invweight, 2 * lb);
  {
    int na, nb, lo_a, hi_a, lo_b, hi_b;
    bool desc_a, desc_b;
    chain_span(m, la, na, lo_a, hi_a, desc_a);
    chain_span(m, lb, nb, lo_b, hi_b, desc_b);
    if (desc_a && desc_b && (hi_a < lo_b || hi_b < lo_a)) {
      f3 n[4];
      for (int i = 0; i < 4; i++) n[i] = (float)(2 * (i % 2) - 1) * (i < 2 ? d1 : d2) * fr - cn;
      float jq[4];
      const bool a_first = na == 0 || nb == 0 || hi_a < lo_b;
      contact_rows_disjoint(e, c0, la, lb, cp, n, na, nb, a_first ? 0 : nb, a_first ? na : 0, jq);
      for (int i = 0; i < 4; i++) {
        const int c = c0 + i;
        e.EI(s.jac_nnz, c) = na + nb;
        float imp, aref;
        imp_aref(sp, -pen, jq[i], -pen, imp, aref);
        float diag = invw + fr * fr * invw;
        diag *= 2.f * fr * fr * (1.f - imp) / imp;
        diag = fmaxf(diag, m.eps);
        e.EF(s.efc_diag, c) = diag;
        e.EF(s.efc_aref, c) = aref;
        e.EF(s.efc_D, c) = 1.f / diag;
      }
      return;
    }
  }
  for (int i = 0; i < 4; i++) {
    f3 dd = (float)(2 * (i % 2) - 1) * (i < 2 ? d1 : d2);
    f3 n = dd * fr - cn;
    int c = c0 + i;
    RowBuf r;
    r.n = 0;
    float jac_qvel = 0.f;
    for (int iab = 0; iab < 2; iab++) {
      float sign = iab == 0 ? -1.f : 1.f;
      int link = iab == 0 ? la : lb;
      while (link > -1) {
        int de = m.link_dof_end[link], nd = de - m.link_dof_start[link];
        f3 tpos = cp - e.L3(s.root_com, link);
        for (int k = 0; k < nd; k++) {
          int d = de - 1 - k;
          f3 vel = qrot(e.L3(s.cdof_vel, d) - cross(tpos, e.L3(s.cdof_ang, d)), qid());
          float jac = dot(sign * vel, n);
          jac_qvel = jac_qvel + jac * e.F(s.vel, d);
          row_add(r, d, jac);
        }
        link = m.link_parent[link];
      }
    }
    row_sort(r);
    row_store(e, c, r);
    float imp, aref;
    imp_aref(sp, -pen, jac_qvel, -pen, imp, aref);
    float diag = invw + fr * fr * invw;
    diag *= 2.f * fr * fr * (1.f - imp) / imp;
    diag = fmaxf(diag, m.eps);
    e.EF(s.efc_diag, c) = diag;
    e.EF(s.efc_aref, c) = aref;
    e.EF(s.efc_D, c) = 1.f / diag;
  }
}

// joint-limit rows (add_joint_limit_constraints), appended from row n_con; returns the new row count
__device__ inline int dev_limit_rows(const Env& e, int n_con) {
  const gsr_model& m = e.m;
  const gsr_state& s = e.s;
  for (int l = 0; l < m.n_links; l++)
    for (int j = m.link_joint_start[l]; j < m.link_joint_end[l]; j++) {
      int jt = m.joint_type[j];
      if (jt != JOINT_REVOLUTE && jt != JOINT_PRISMATIC) continue;
      int q = m.joint_q_start[j], d = m.joint_dof_start[j];
      float qv = e.F(s.qpos, q);
      float dmin = qv - e.DP(m.dof_limit, 2 * d), dmax = e.DP(m.dof_limit, 2 * d + 1) - qv;
      float pd = fminf(dmin, dmax);
      if (pd < 0.f) {
        float jac = (dmin < dmax) ? 1.f : -1.f;
        float imp, aref;
        size_t sps;
        imp_aref(e.joint_sp(j, sps), sps, pd, jac * e.F(s.vel, d), pd, imp, aref);
        float diag = fmaxf(e.DP(m.dof_invweight, d) * (1.f - imp) / imp, m.eps);
        int c = n_con++;
        e.EF(s.efc_diag, c) = diag;
        e.EF(s.efc_aref, c) = aref;
        e.EF(s.efc_D, c) = 1.f / diag;
        RowBuf r;
        r.n = 1; r.idx[0] = d; r.val[0] = jac;
        row_store(e, c, r);
      }
    }
  return n_con;
}

__device__ inline void dev_inequality_rows(const Env& e) {
  const gsr_model& m = e.m;
  const gsr_state& s = e.s;
  int n_fl = 0;
  int n_con = dev_frictionloss_rows(e, e.I(s.n_con, 0), n_fl);
  if (m.enable_collision) {
    int nc = e.I(s.n_contacts, 0);
    for (int ic = 0; ic < nc; ic++) dev_contact_rows(e, ic, n_con + 4 * ic);
    n_con += 4 * nc;
  }
  if (m.enable_joint_limit) n_con = dev_limit_rows(e, n_con);
  e.I(s.n_con, 0) = n_con;
  e.I(s.n_con, 2) = n_fl;
}

// ---------------------------------------------------------------------------------------------------------------
// Newton solver
// ---------------------------------------------------------------------------------------------------------------
struct SolverScalars {
  float gauss, cost, prev_cost;
  float qg0, qg1, qg2;
  int ls_it;
};

__device__ inline void dev_update_constraint(const Env& e, SolverScalars& sc) {
  const gsr_model& m = e.m;
  const gsr_state& s = e.s;
  const int D = m.n_dofs, W = m.row_width;
  int n_con = e.I(s.n_con, 0), ne = e.I(s.n_con, 1), nef = ne + e.I(s.n_con, 2);
  sc.prev_cost = sc.cost;
  float cost_i = 0.f, gauss_i = 0.f;
  for (int c = 0; c < n_con; c++) {
    int act_prev = e.EI(s.efc_active, c) & 1;
    int act = 1;
    float floss_force = 0.f;
    float ja = e.EF(s.efc_Jaref, c);
    if (ne <= c && c < nef) {
      float f = e.EF(s.efc_fl, c), r = e.EF(s.efc_diag, c), rf = r * f;
      bool ln = ja <= -rf, lp = ja >= rf;
      act = !(ln || lp);
      floss_force = (float)ln * f + (float)lp * -f;
      float fc = (float)ln * f * (-0.5f * rf - ja);
      fc += (float)lp * f * (-0.5f * rf + ja);
      cost_i += fc;
    } else if (nef <= c) {
      act = ja < 0.f;
    }
    e.EI(s.efc_active, c) = act | (act_prev << 1);
    e.EF(s.efc_force, c) = floss_force + (-ja * e.EF(s.efc_D, c) * (float)act);
  }
  for (int d = 0; d < D; d++) e.F(s.qfrc_constraint, d) = 0.f;
  for (int c = 0; c < n_con; c++) {
    float f = e.EF(s.efc_force, c);
    int nnz = e.EI(s.jac_nnz, c);
    for (int k = 0; k < nnz; k++) {
      int d = e.JI(c, k);
      e.F(s.qfrc_constraint, d) = e.F(s.qfrc_constraint, d) + e.JV(c, k) * f;
    }
  }
  for (int d = 0; d < D; d++) {
    float v = 0.5f * (e.F(s.Ma, d) - e.F(s.force, d)) * (e.F(s.qacc, d) - e.F(s.acc_smooth, d));
    gauss_i += v;
    cost_i += v;
  }
  for (int c = 0; c < n_con; c++) {
    float ja = e.EF(s.efc_Jaref, c);
    cost_i += 0.5f * (ja * ja * e.EF(s.efc_D, c) * (float)(e.EI(s.efc_active, c) & 1));
  }
  sc.gauss = gauss_i;
  sc.cost = cost_i;
}

__device__ inline void dev_chol_factor(const Env& e) {
  const int D = e.m.n_dofs;
  const float EPS = e.m.eps;
  for (int i = 0; i < D; i++) {
    float tmp = e.H(i, i);
    for (int j = 0; j < i; j++) {
      float h = e.H(i, j);
      tmp -= h * h;
    }
    float hii = sqrtf(fmaxf(tmp, EPS));
    e.H(i, i) = hii;
    float inv = 1.f / hii;
    for (int j = i + 1; j < D; j++) {
      float dt = 0.f;
      for (int k = 0; k < i; k++) dt += e.H(j, k) * e.H(i, k);
      e.H(j, i) = (e.H(j, i) - dt) * inv;
    }
  }
}

__device__ inline void dev_hessian_direct(const Env& e) {
  const gsr_model& m = e.m;
  const gsr_state& s = e.s;
  const int D = m.n_dofs, W = m.row_width;
  int n_con = e.I(s.n_con, 0);
  for (int i = 0; i < D; i++)
    for (int j = 0; j <= i; j++) e.H(i, j) = 0.f;
  for (int c = 0; c < n_con; c++) {
    int nnz = e.EI(s.jac_nnz, c);
    float Dc = e.EF(s.efc_D, c), act = (float)(e.EI(s.efc_active, c) & 1);
    for (int k1 = 0; k1 < nnz; k1++) {
      float v1 = e.JV(c, k1);
      if (!(fabsf(v1) > m.eps)) continue;
      int i1 = e.JI(c, k1);
      for (int k2 = 0; k2 <= k1; k2++) {
        int i2 = e.JI(c, k2);
        e.H(i1, i2) = e.H(i1, i2) + e.JV(c, k2) * v1 * Dc * act;
      }
    }
  }
  for (int ie = 0; ie < m.n_entities; ie++)
    for (int i = m.ent_dof_start[ie]; i < m.ent_dof_end[ie]; i++)
      for (int j = m.ent_dof_start[ie]; j <= i; j++) e.H(i, j) = e.H(i, j) + e.Mm(ie, i, j);
  dev_chol_factor(e);
}

__device__ inline void dev_hessian_incremental(const Env& e) {
  const gsr_model& m = e.m;
  const gsr_state& s = e.s;
  const int D = m.n_dofs, W = m.row_width;
  int n_con = e.I(s.n_con, 0);
  for (int c = 0; c < n_con; c++) {
    int a = e.EI(s.efc_active, c);
    int act = a & 1, prev = (a >> 1) & 1;
    int flag = -1;
    if (prev == 0 && act == 1) flag = 1;
    if (prev == 1 && act == 0) flag = 0;
    if (flag == -1) continue;
    float sq = sqrtf(e.EF(s.efc_D, c));
    for (int d = 0; d < D; d++) e.F(s.nt_vec, d) = 0.f * sq;
    int nnz = e.EI(s.jac_nnz, c);
    for (int k = 0; k < nnz; k++) e.F(s.nt_vec, e.JI(c, k)) = e.JV(c, k) * sq;
    int rank = D;
    float sgn = (float)(flag * 2 - 1);
    for (int k = 0; k < D; k++) {
      float vk = e.F(s.nt_vec, k);
      if (fabsf(vk) > m.eps) {
        float Lkk = e.H(k, k);
        float tmp = Lkk * Lkk + vk * vk * sgn;
        if (tmp < m.eps) {
          tmp = m.eps;
          rank--;
        }
        float r = sqrtf(tmp), cc = r / Lkk, cinv = 1.f / cc, sv = vk / Lkk;
        e.H(k, k) = r;
        for (int i = k + 1; i < D; i++) e.H(i, k) = (e.H(i, k) + sv * e.F(s.nt_vec, i) * sgn) * cinv;
        for (int i = k + 1; i < D; i++) e.F(s.nt_vec, i) = e.F(s.nt_vec, i) * cc - sv * e.H(i, k);
      }
    }
    if (rank < D) {
      dev_hessian_direct(e);
      return;
    }
  }
}

__device__ inline void dev_update_gradient(const Env& e) {
  const gsr_state& s = e.s;
  const int D = e.m.n_dofs;
  for (int d = 0; d < D; d++) e.F(s.grad, d) = e.F(s.Ma, d) - e.F(s.force, d) - e.F(s.qfrc_constraint, d);
  for (int i = 0; i < D; i++) {
    float g = e.F(s.grad, i);
    for (int j = 0; j < i; j++) g = g - e.H(i, j) * e.F(s.Mgrad, j);
    e.F(s.Mgrad, i) = g / e.H(i, i);
  }
  for (int ii = 0; ii < D; ii++) {
    int i = D - 1 - ii;
    float g = e.F(s.Mgrad, i);
    for (int j = i + 1; j < D; j++) g = g - e.H(j, i) * e.F(s.Mgrad, j);
    e.F(s.Mgrad, i) = g / e.H(i, i);
  }
}

struct LSPoint {
  float alpha, cost, d0, d1;
};

__device__ inline LSPoint dev_ls_point(const Env& e, SolverScalars& sc, float alpha) {
  const gsr_state& s = e.s;
  int n_con = e.I(s.n_con, 0), ne = e.I(s.n_con, 1), nef = ne + e.I(s.n_con, 2);
  float t0 = sc.qg0, t1 = sc.qg1, t2 = sc.qg2;
  for (int c = 0; c < n_con; c++) {
    float ja = e.EF(s.efc_Jaref, c), jv = e.EF(s.efc_jv, c);
    float x = ja + alpha * jv;
    float q0 = e.QD(c, 0), q1 = e.QD(c, 1), q2 = e.QD(c, 2);
    bool act = true;
    if (ne <= c && c < nef) {
      float f = e.EF(s.efc_fl, c), r = e.EF(s.efc_diag, c), rf = r * f;
      bool ln = x <= -rf, lp = x >= rf;
      if (ln || lp) {
        q0 = (float)ln * f * (-0.5f * rf - ja) + (float)lp * f * (-0.5f * rf + ja);
        q1 = (float)ln * (-f * jv) + (float)lp * (f * jv);
        q2 = 0.f;
      }
    } else if (nef <= c) {
      act = x < 0.f;
    }
    t0 += q0 * (float)act;
    t1 += q1 * (float)act;
    t2 += q2 * (float)act;
  }
  LSPoint p;
  p.alpha = alpha;
  p.cost = alpha * alpha * t2 + alpha * t1 + t0;
  p.d0 = 2.f * alpha * t2 + t1;
  p.d1 = 2.f * t2;
  if (p.d1 <= 0.f) p.d1 = e.m.eps;
  sc.ls_it++;
  return p;
}

// moves p to the candidates (c0, c1, c2 in this order) whose slope has its sign and is closer to zero; pn is the
// next point from p when it moved. Named candidates, not an array, so they stay in registers.
__device__ inline void dev_update_bracket(const Env& e, SolverScalars& sc, const LSPoint& c0, const LSPoint& c1,
                                          const LSPoint& c2, LSPoint& p, int& flag, LSPoint& pn) {
  flag = 0;
  auto step = [&](const LSPoint& c) {
    if (p.d0 < 0.f && c.d0 < 0.f && p.d0 < c.d0) {
      p = c;
      flag = 1;
    } else if (p.d0 > 0.f && c.d0 > 0.f && p.d0 > c.d0) {
      p = c;
      flag = 2;
    }
  };
  step(c0);
  step(c1);
  step(c2);
  pn = p;
  if (flag > 0) pn = dev_ls_point(e, sc, p.alpha - p.d0 / p.d1);
}

__device__ inline float dev_linesearch(const Env& e, SolverScalars& sc) {
  const gsr_model& m = e.m;
  const gsr_state& s = e.s;
  const int D = m.n_dofs, W = m.row_width;
  float snorm = 0.f;
  for (int d = 0; d < D; d++) {
    float v = e.F(s.search, d);
    snorm += v * v;
  }
  snorm = sqrtf(snorm);
  float scale = e.meaninertia() * (float)(D > 1 ? D : 1);
  float gtol = m.tolerance * m.ls_tolerance * snorm * scale;
  sc.ls_it = 0;
  if (snorm < m.eps) return 0.f;
  // func_ls_init
  for (int ie = 0; ie < m.n_entities; ie++)
    for (int i = m.ent_dof_start[ie]; i < m.ent_dof_end[ie]; i++) {
      float mv = 0.f;
      for (int j = m.ent_dof_start[ie]; j < m.ent_dof_end[ie]; j++) mv += e.Mm(ie, i, j) * e.F(s.search, j);
      e.F(s.mv, i) = mv;
    }
  int n_con = e.I(s.n_con, 0);
  for (int c = 0; c < n_con; c++) {
    float jv = 0.f;
    int nnz = e.EI(s.jac_nnz, c);
    for (int k = 0; k < nnz; k++) jv += e.JV(c, k) * e.F(s.search, e.JI(c, k));
    e.EF(s.efc_jv, c) = jv;
  }
  float qg1 = 0.f, qg2 = 0.f;
  for (int d = 0; d < D; d++) {
    float sd = e.F(s.search, d);
    qg1 += sd * e.F(s.Ma, d) - sd * e.F(s.force, d);
    qg2 += 0.5f * sd * e.F(s.mv, d);
  }
  sc.qg0 = sc.gauss;
  sc.qg1 = qg1;
  sc.qg2 = qg2;
  for (int c = 0; c < n_con; c++) {
    float Dc = e.EF(s.efc_D, c), ja = e.EF(s.efc_Jaref, c), jv = e.EF(s.efc_jv, c);
    e.QD(c, 0) = Dc * (0.5f * ja * ja);
    e.QD(c, 1) = Dc * (jv * ja);
    e.QD(c, 2) = Dc * (0.5f * jv * jv);
  }
  LSPoint p0 = dev_ls_point(e, sc, 0.f);
  LSPoint p1 = dev_ls_point(e, sc, p0.alpha - p0.d0 / p0.d1);
  if (p0.cost < p1.cost) p1 = p0;
  if (fabsf(p1.d0) < gtol) return p1.alpha;
  float direction = (p1.d0 < 0.f) ? 1.f : -1.f;
  int p2update = 0;
  LSPoint p2 = p1;
  while (p1.d0 * direction <= -gtol && sc.ls_it < m.ls_iterations) {
    p2 = p1;
    p2update = 1;
    p1 = dev_ls_point(e, sc, p1.alpha - p1.d0 / p1.d1);
    if (fabsf(p1.d0) < gtol) return p1.alpha;
  }
  if (sc.ls_it >= m.ls_iterations) return p1.alpha;
  if (!p2update) return p1.alpha;
  LSPoint p2n = p1;
  LSPoint p1n = dev_ls_point(e, sc, p1.alpha - p1.d0 / p1.d1);
  while (sc.ls_it < m.ls_iterations) {
    LSPoint pm = dev_ls_point(e, sc, (p1.alpha + p2.alpha) * 0.5f);
    const LSPoint c0 = p1n, c1 = p2n, c2 = pm;
    int best = -1;
    float bcost = 0.f, balpha = 0.f;
    auto consider = [&](const LSPoint& c, int idx) {
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
    int b1, b2;
    dev_update_bracket(e, sc, c0, c1, c2, p1, b1, p1n);
    dev_update_bracket(e, sc, c0, c1, c2, p2, b2, p2n);
    if (b1 == 0 && b2 == 0) return pm.alpha;
  }
  if (p1.cost <= p2.cost && p1.cost < p0.cost) return p1.alpha;
  if (p2.cost <= p1.cost && p2.cost < p1.cost) return p2.alpha;
  return 0.f;
}

__device__ inline void dev_solve(const Env& e) {
  const gsr_model& m = e.m;
  const gsr_state& s = e.s;
  const int D = m.n_dofs, W = m.row_width;
  int n_con = e.I(s.n_con, 0);
  SolverScalars sc;
  sc.cost = 0.f;
  sc.prev_cost = 0.f;
  sc.gauss = 0.f;
  sc.ls_it = 0;
  // func_init_solver, non-compatibility warm start
  int ws = e.I(s.is_warmstart, 0);
  for (int d = 0; d < D; d++) e.F(s.qacc, d) = (n_con > 0 && ws) ? e.F(s.qacc_ws, d) : e.F(s.acc_smooth, d);
  for (int i = 0; i < D; i++) {
    int ie = m.dof_entity[i];
    float v = 0.f;
    for (int j = m.ent_dof_start[ie]; j < m.ent_dof_end[ie]; j++) v += e.Mm(ie, i, j) * e.F(s.qacc, j);
    e.F(s.Ma, i) = v;
  }
  for (int c = 0; c < n_con; c++) {
    float ja = -e.EF(s.efc_aref, c);
    int nnz = e.EI(s.jac_nnz, c);
    for (int k = 0; k < nnz; k++) ja += e.JV(c, k) * e.F(s.qacc, e.JI(c, k));
    e.EF(s.efc_Jaref, c) = ja;
  }
  dev_update_constraint(e, sc);
  dev_hessian_direct(e);
  dev_update_gradient(e);
  for (int d = 0; d < D; d++) e.F(s.search, d) = -e.F(s.Mgrad, d);
  int iters = 0;
  if (n_con > 0) {
    for (int it = 0; it < m.iterations; it++) {
      iters++;
      float alpha = dev_linesearch(e, sc);
      bool improved;
      if (fabsf(alpha) < m.eps) {
        improved = false;
      } else {
        for (int d = 0; d < D; d++) {
          e.F(s.qacc, d) = e.F(s.qacc, d) + e.F(s.search, d) * alpha;
          e.F(s.Ma, d) = e.F(s.Ma, d) + e.F(s.mv, d) * alpha;
        }
        for (int c = 0; c < n_con; c++) e.EF(s.efc_Jaref, c) = e.EF(s.efc_Jaref, c) + e.EF(s.efc_jv, c) * alpha;
        dev_update_constraint(e, sc);
        dev_hessian_incremental(e);
        dev_update_gradient(e);
        float tol_scaled = (e.meaninertia() * (float)(D > 1 ? D : 1)) * m.tolerance;
        float improvement = sc.prev_cost - sc.cost;
        float gn = 0.f;
        for (int d = 0; d < D; d++) {
          float g = e.F(s.grad, d);
          gn += g * g;
        }
        gn = sqrtf(gn);
        if (gn < tol_scaled || improvement < tol_scaled) {
          improved = false;
        } else {
          improved = true;
          for (int d = 0; d < D; d++) e.F(s.search, d) = -e.F(s.Mgrad, d);
        }
      }
      if (!improved) break;
    }
  }
  e.I(s.solver_iters, 0) = iters;
  // func_update_qacc
  bool nan_qacc = false;
  for (int d = 0; d < D; d++) {
    float q = e.F(s.qacc, d);
    e.F(s.acc, d) = q;
    e.F(s.force, d) = e.F(s.qf_smooth, d) + e.F(s.qfrc_constraint, d);
    e.F(s.qacc_ws, d) = q;
    nan_qacc |= isnan(q);
  }
  if (nan_qacc) e.I((int32_t*)s.errno_env, 0) |= (int)GSR_ERR_NAN_QACC;
  e.I(s.is_warmstart, 0) = 1;
  // func_update_contact_force (per-contact force only; per-link sums are derived on read)
  int cs = e.I(s.n_con, 1) + e.I(s.n_con, 2);
  int nc = m.enable_collision ? e.I(s.n_contacts, 0) : 0;
  for (int ic = 0; ic < nc; ic++) {
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

// func_integrate (rsd:6566-6708) and the NaN-guarded commit func_copy_next_to_curr (rsd:6711-6740)
// semi-implicit Euler into vel_next / qpos_next for the entities (first, stride); returns false on a NaN
__device__ inline bool dev_integrate_next(const Env& e, int e0, int es) {
  const gsr_model& m = e.m;
  const gsr_state& s = e.s;
  const float dt = m.dt;
  bool valid = true;
  for (int ie = e0; ie < m.n_entities; ie += es) {
    for (int d = m.ent_dof_start[ie]; d < m.ent_dof_end[ie]; d++) {
      float v = e.F(s.vel, d) + e.F(s.acc, d) * dt;
      e.F(s.vel_next, d) = v;
      valid &= !isnan(v);
    }
    for (int l = m.ent_link_start[ie]; l < m.ent_link_end[ie]; l++) {
      int ds = m.link_dof_start[l], qs = m.link_q_start[l], qe = m.link_q_end[l];
      if (m.link_dof_end[l] - ds <= 0) continue;
      int jt = m.joint_type[m.link_joint_start[l]];
      if (jt == JOINT_FREE) {
        for (int j = 0; j < 3; j++) e.F(s.qpos_next, qs + j) = e.F(s.qpos, qs + j) + e.F(s.vel_next, ds + j) * dt;
      }
      if (jt == JOINT_SPHERICAL || jt == JOINT_FREE) {
        int ro = jt == JOINT_FREE ? 3 : 0;
        q4 rot0{e.F(s.qpos, qs + ro), e.F(s.qpos, qs + ro + 1), e.F(s.qpos, qs + ro + 2), e.F(s.qpos, qs + ro + 3)};
        f3 ang = f3{e.F(s.vel_next, ds + ro), e.F(s.vel_next, ds + ro + 1), e.F(s.vel_next, ds + ro + 2)} * dt;
        q4 rot = qrot_q(rotvec_to_quat(ang, m.eps), rot0);  // normalize(rot0 * qrot)
        e.F(s.qpos_next, qs + ro) = rot.w;
        e.F(s.qpos_next, qs + ro + 1) = rot.x;
        e.F(s.qpos_next, qs + ro + 2) = rot.y;
        e.F(s.qpos_next, qs + ro + 3) = rot.z;
      } else {
        for (int j = 0; j < qe - qs; j++) e.F(s.qpos_next, qs + j) = e.F(s.qpos, qs + j) + e.F(s.vel_next, ds + j) * dt;
      }
      for (int q = qs; q < qe; q++) valid &= !isnan(e.F(s.qpos_next, q));
    }
  }
  return valid;
}

// NaN-guarded commit (func_copy_next_to_curr): the env's whole state moves only if no entity produced a NaN
__device__ inline void dev_integrate_commit(const Env& e, int e0, int es, bool valid, bool flag_lane) {
  const gsr_model& m = e.m;
  const gsr_state& s = e.s;
  if (valid) {
    for (int ie = e0; ie < m.n_entities; ie += es) {
      for (int d = m.ent_dof_start[ie]; d < m.ent_dof_end[ie]; d++) e.F(s.vel, d) = e.F(s.vel_next, d);
      for (int l = m.ent_link_start[ie]; l < m.ent_link_end[ie]; l++)
        for (int q = m.link_q_start[l]; q < m.link_q_end[l]; q++) e.F(s.qpos, q) = e.F(s.qpos_next, q);
    }
  } else if (flag_lane) {
    e.I((int32_t*)s.errno_env, 0) |= (int)GSR_ERR_NAN_STATE;
  }
}

// func_integrate (rsd:6566-6708) + func_copy_next_to_curr (rsd:6711-6740)
__device__ inline void dev_integrate(const Env& e) {
  bool valid = dev_integrate_next(e, 0, 1);
  dev_integrate_commit(e, 0, 1, valid, true);
}

}  // namespace gsr
