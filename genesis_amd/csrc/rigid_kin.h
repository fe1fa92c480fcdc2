// Kinematics and smooth dynamics of one environment (one lane per env, env-minor global state).
//
// Stage order and arithmetic follow rigid_solver_decomp.py: FK (func_forward_kinematics_entity :5219-5376),
// COM and motion subspaces (func_COM_links :4746-5078), geom poses (func_update_geoms :5533-5595), velocity FK
// (func_forward_velocity_entity :5379-5503), CRBA (func_compute_mass_matrix :3422-3678), LDL^T (func_factor_mass
// :3681-3807), actuation (func_torque_and_passive_force :5960-6177), bias accelerations and forces (func_update_acc
// :6180-6281, func_update_force :6284-6398, func_bias_force :6425-6485) and M^-1 f (func_solve_mass_batched :3932-4042).
#pragma once
#include "rigid_env.h"

namespace gsr {

// Every function below takes the entities it works on as (first, stride): (0, 1) is the whole env on one lane,
// (lane, 64) splits the env's entities over a wavefront. Entities share no state, so the split changes no result.
#define GSR_FOR_ENT_LINKS(ie, l)                                      \
  for (int ie = e0; ie < m.n_entities; ie += es)                      \
    for (int l = m.ent_link_start[ie]; l < m.ent_link_end[ie]; l++)
#define GSR_FOR_ENT_DOFS(ie, d)                                       \
  for (int ie = e0; ie < m.n_entities; ie += es)                      \
    for (int d = m.ent_dof_start[ie]; d < m.ent_dof_end[ie]; d++)

__device__ inline void dev_forward_kinematics(const Env& e, int e0 = 0, int es = 1) {
  const gsr_model& m = e.m;
  const gsr_state& s = e.s;
  const float EPS = m.eps;
  GSR_FOR_ENT_LINKS(ie, l) {
    f3 pos = T3(m.link_pos, l);
    q4 quat = T4(m.link_quat, l);
    int par = m.link_parent[l];
    if (par != -1) {
      f3 ppos = e.L3(s.links_pos, par);
      q4 pq = e.L4(s.links_quat, par);
      pos = ppos + qrot(T3(m.link_pos, l), pq);
      quat = qrot_q(T4(m.link_quat, l), pq);
    }
    for (int j = m.link_joint_start[l]; j < m.link_joint_end[l]; j++) {
      int jt = m.joint_type[j], qs = m.joint_q_start[j], ds = m.joint_dof_start[j];
      if (jt == JOINT_FREE) {
        e.S3(s.xanchor, j, f3{e.F(s.qpos, qs), e.F(s.qpos, qs + 1), e.F(s.qpos, qs + 2)});
        e.S3(s.xaxis, j, f3{0.f, 0.f, 1.f});
      } else if (jt != JOINT_FIXED) {
        f3 axis{0.f, 0.f, 1.f};
        if (jt == JOINT_REVOLUTE) axis = T3(m.dof_motion_ang, ds);
        else if (jt == JOINT_PRISMATIC) axis = T3(m.dof_motion_vel, ds);
        e.S3(s.xanchor, j, qrot(T3(m.joint_pos, j), quat) + pos);
        e.S3(s.xaxis, j, qrot(axis, quat));
      }
      if (jt == JOINT_FREE) {
        pos = f3{e.F(s.qpos, qs), e.F(s.qpos, qs + 1), e.F(s.qpos, qs + 2)};
        quat = q4{e.F(s.qpos, qs + 3), e.F(s.qpos, qs + 4), e.F(s.qpos, qs + 5), e.F(s.qpos, qs + 6)};
        f3 xyz = quat_to_xyz(quat, EPS);
        e.F(s.dofs_pos, ds) = pos.x; e.F(s.dofs_pos, ds + 1) = pos.y; e.F(s.dofs_pos, ds + 2) = pos.z;
        e.F(s.dofs_pos, ds + 3) = xyz.x; e.F(s.dofs_pos, ds + 4) = xyz.y; e.F(s.dofs_pos, ds + 5) = xyz.z;
      } else if (jt == JOINT_SPHERICAL) {
        q4 qloc{e.F(s.qpos, qs), e.F(s.qpos, qs + 1), e.F(s.qpos, qs + 2), e.F(s.qpos, qs + 3)};
        f3 xyz = quat_to_xyz(qloc, EPS);
        e.F(s.dofs_pos, ds) = xyz.x; e.F(s.dofs_pos, ds + 1) = xyz.y; e.F(s.dofs_pos, ds + 2) = xyz.z;
        quat = qrot_q(qloc, quat);
        pos = e.L3(s.xanchor, j) - qrot(T3(m.joint_pos, j), quat);
      } else if (jt == JOINT_REVOLUTE) {
        f3 axis = T3(m.dof_motion_ang, ds);
        float dp = e.F(s.qpos, qs) - m.qpos0[qs];
        e.F(s.dofs_pos, ds) = dp;
        q4 qloc = rotvec_to_quat(axis * dp, EPS);
        quat = qrot_q(qloc, quat);
        pos = e.L3(s.xanchor, j) - qrot(T3(m.joint_pos, j), quat);
      } else if (jt == JOINT_PRISMATIC) {
        float dp = e.F(s.qpos, qs) - m.qpos0[qs];
        e.F(s.dofs_pos, ds) = dp;
        pos = pos + e.L3(s.xaxis, j) * dp;
      }
    }
    // fixed root links keep their build-time pose
    if (par == -1 && m.link_is_fixed[l]) {
      pos = T3(m.link_pos, l);
      quat = T4(m.link_quat, l);
    }
    e.S3(s.links_pos, l, pos);
    e.S4(s.links_quat, l, quat);
  }
}

__device__ inline void dev_com_links(const Env& e, int e0 = 0, int es = 1) {
  const gsr_model& m = e.m;
  const gsr_state& s = e.s;
  const float EPS = m.eps;
  // subtree-root mass and mass-weighted COM, accumulated in link order (crb_* used as scratch)
  GSR_FOR_ENT_LINKS(ie, l) {
    e.S3(s.root_com, l, zero3());
    e.F(s.crb_mass, l) = 0.f;
  }
  GSR_FOR_ENT_LINKS(ie, l) {
    // per-env mass and COM shifts (rsd:4812-4817)
    float mass = e.LP(m.link_inertial_mass, l) + e.F(s.mass_shift, l);
    f3 lp = e.L3(s.links_pos, l);
    q4 lq = e.L4(s.links_quat, l);
    f3 ipos_bw = lp + qrot(T3(m.link_inertial_pos, l) + e.L3(s.i_pos_shift, l), lq);
    e.S4(s.links_iquat, l, qrot_q(T4(m.link_inertial_quat, l), lq));
    e.S3(s.crb_pos, l, ipos_bw);
    int r = m.link_root[l];
    e.F(s.crb_mass, r) = e.F(s.crb_mass, r) + mass;
    e.S3(s.root_com, r, e.L3(s.root_com, r) + mass * ipos_bw);
  }
  GSR_FOR_ENT_LINKS(ie, l)
    if (m.link_root[l] == l) e.S3(s.root_com, l, e.L3(s.root_com, l) / e.F(s.crb_mass, l));
  GSR_FOR_ENT_LINKS(ie, l) {
    int r = m.link_root[l];
    if (r != l) e.S3(s.root_com, l, e.L3(s.root_com, r));
  }
  GSR_FOR_ENT_LINKS(ie, l) {
    f3 t = e.L3(s.crb_pos, l) - e.L3(s.root_com, l);  // i_pos
    float mass = e.LP(m.link_inertial_mass, l) + e.F(s.mass_shift, l);  // rsd:4910
    float x = t.x, y = t.y, z = t.z;
    float xx = x * x, xy = x * y, xz = x * z, yy = y * y, yz = y * z, zz = z * z;
    float hhT[9] = {yy + zz, -xy, -xz, -xy, xx + zz, -yz, -xz, -yz, xx + yy};
    m3 R = quat_to_R(e.L4(s.links_iquat, l), EPS);
    const float* I0 = m.link_inertial_i + 9 * l;
    float RI[9];
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) RI[3 * i + j] = R.a[3 * i] * I0[j] + R.a[3 * i + 1] * I0[3 + j] + R.a[3 * i + 2] * I0[6 + j];
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) {
        float v = RI[3 * i] * R.a[3 * j] + RI[3 * i + 1] * R.a[3 * j + 1] + RI[3 * i + 2] * R.a[3 * j + 2];
        e.F(s.cinr_inertial, 9 * l + 3 * i + j) = v + hhT[3 * i + j] * mass;
      }
    e.S3(s.cinr_pos, l, t * mass);
    e.F(s.cinr_mass, l) = mass;
  }
  GSR_FOR_ENT_LINKS(ie, l) {
    if (m.link_dof_end[l] - m.link_dof_start[l] <= 0) continue;
    for (int j = m.link_joint_start[l]; j < m.link_joint_end[l]; j++) {
      int jt = m.joint_type[j], ds = m.joint_dof_start[j];
      if (jt == JOINT_FIXED) continue;
      f3 xaxis = e.L3(s.xaxis, j);
      f3 off = e.L3(s.root_com, l) - e.L3(s.xanchor, j);
      if (jt == JOINT_REVOLUTE) {
        e.S3(s.cdof_ang, ds, xaxis);
        e.S3(s.cdof_vel, ds, cross(xaxis, off));
      } else if (jt == JOINT_PRISMATIC) {
        e.S3(s.cdof_ang, ds, zero3());
        e.S3(s.cdof_vel, ds, xaxis);
      } else if (jt == JOINT_SPHERICAL) {
        m3 R = quat_to_R(e.L4(s.links_quat, l), EPS);
        for (int i = 0; i < 3; i++) {
          f3 r = mcol(R, i);  // row i of R^T
          e.S3(s.cdof_ang, ds + i, r);
          e.S3(s.cdof_vel, ds + i, cross(r, off));
        }
      } else {  // FREE
        for (int i = 0; i < 3; i++) {
          e.S3(s.cdof_ang, ds + i, zero3());
          e.S3(s.cdof_vel, ds + i, f3{i == 0 ? 1.f : 0.f, i == 1 ? 1.f : 0.f, i == 2 ? 1.f : 0.f});
        }
        m3 R = quat_to_R(e.L4(s.links_quat, l), EPS);
        for (int i = 0; i < 3; i++) {
          f3 r = mcol(R, i);
          e.S3(s.cdof_ang, ds + 3 + i, r);
          e.S3(s.cdof_vel, ds + 3 + i, cross(r, off));
        }
      }
    }
  }
}

__device__ inline void dev_update_geoms(const Env& e, int e0 = 0, int es = 1) {
  const gsr_model& m = e.m;
  const gsr_state& s = e.s;
  for (int g = e0; g < m.n_geoms; g += es) {
    int l = m.geom_link[g];
    f3 lp = e.L3(s.links_pos, l);
    q4 lq = e.L4(s.links_quat, l);
    e.S3(s.geoms_pos, g, lp + qrot(T3(m.geom_pos, g), lq));
    e.S4(s.geoms_quat, g, qrot_q(T4(m.geom_quat, g), lq));
  }
}

__device__ inline void dev_forward_velocity(const Env& e, int e0 = 0, int es = 1) {
  const gsr_model& m = e.m;
  const gsr_state& s = e.s;
  GSR_FOR_ENT_LINKS(ie, l) {
    f3 cv = zero3(), ca = zero3();
    int par = m.link_parent[l];
    if (par != -1) {
      cv = e.L3(s.cd_vel, par);
      ca = e.L3(s.cd_ang, par);
    }
    for (int j = m.link_joint_start[l]; j < m.link_joint_end[l]; j++) {
      int jt = m.joint_type[j], ds = m.joint_dof_start[j], de = m.joint_dof_end[j];
      if (jt == JOINT_FREE) {
        for (int i = 0; i < 3; i++) {
          float v = e.F(s.vel, ds + i);
          cv = cv + e.L3(s.cdof_vel, ds + i) * v;
          ca = ca + e.L3(s.cdof_ang, ds + i) * v;
        }
        for (int i = 0; i < 3; i++) {
          e.S3(s.cdofd_ang, ds + i, zero3());
          e.S3(s.cdofd_vel, ds + i, zero3());
          f3 oa, ov;
          motion_cross_motion(ca, cv, e.L3(s.cdof_ang, ds + 3 + i), e.L3(s.cdof_vel, ds + 3 + i), oa, ov);
          e.S3(s.cdofd_ang, ds + 3 + i, oa);
          e.S3(s.cdofd_vel, ds + 3 + i, ov);
        }
        for (int i = 0; i < 3; i++) {
          float v = e.F(s.vel, ds + 3 + i);
          cv = cv + e.L3(s.cdof_vel, ds + 3 + i) * v;
          ca = ca + e.L3(s.cdof_ang, ds + 3 + i) * v;
        }
      } else {
        for (int d = ds; d < de; d++) {
          f3 oa, ov;
          motion_cross_motion(ca, cv, e.L3(s.cdof_ang, d), e.L3(s.cdof_vel, d), oa, ov);
          e.S3(s.cdofd_ang, d, oa);
          e.S3(s.cdofd_vel, d, ov);
        }
        for (int d = ds; d < de; d++) {
          float v = e.F(s.vel, d);
          cv = cv + e.L3(s.cdof_vel, d) * v;
          ca = ca + e.L3(s.cdof_ang, d) * v;
        }
      }
    }
    e.S3(s.cd_vel, l, cv);
    e.S3(s.cd_ang, l, ca);
  }
}

// geoms are split by index, so a split run needs a barrier between link poses and geom poses
__device__ inline void dev_kinematics(const Env& e) {
  dev_forward_kinematics(e);
  dev_com_links(e);
  dev_update_geoms(e);
  dev_forward_velocity(e);
}

// M^-1 vec per entity block through the LDL^T factors (rsd:3932-4042)
__device__ inline void dev_solve_mass(const Env& e, const float* vec, float* out, int e0 = 0, int es = 1) {
  const gsr_model& m = e.m;
  const gsr_state& s = e.s;
  for (int ie = e0; ie < m.n_entities; ie += es) {
    int st = m.ent_dof_start[ie], en = m.ent_dof_end[ie], n = en - st;
    for (int ii = 0; ii < n; ii++) {
      int i = en - ii - 1;
      float o = e.F(const_cast<float*>(vec), i);
      for (int j = i + 1; j < en; j++) o -= e.ML(ie, j, i) * e.F(out, j);
      e.F(out, i) = o;
    }
    for (int i = st; i < en; i++) e.F(out, i) = e.F(out, i) * e.F(s.mass_D_inv, i);
    for (int i = st; i < en; i++) {
      float o = e.F(out, i);
      for (int j = st; j < i; j++) o -= e.ML(ie, i, j) * e.F(out, j);
      e.F(out, i) = o;
    }
  }
}

__device__ inline void dev_smooth_dynamics(const Env& e, int e0 = 0, int es = 1) {
  const gsr_model& m = e.m;
  const gsr_state& s = e.s;
  const int D = m.n_dofs, E = m.n_entities;
  const float EPS = m.eps, dt = m.dt;
  // --- CRBA (rsd:3422-3678) ---
  GSR_FOR_ENT_LINKS(ie0, l) {
    for (int k = 0; k < 9; k++) e.F(s.crb_inertial, 9 * l + k) = e.F(s.cinr_inertial, 9 * l + k);
    e.S3(s.crb_pos, l, e.L3(s.cinr_pos, l));
    e.F(s.crb_mass, l) = e.F(s.cinr_mass, l);
  }
  for (int ie = e0; ie < E; ie += es) {
    int ls = m.ent_link_start[ie], le = m.ent_link_end[ie];
    for (int l = le - 1; l >= ls; l--) {
      int p = m.link_parent[l];
      if (p != -1) {
        for (int k = 0; k < 9; k++) e.F(s.crb_inertial, 9 * p + k) = e.F(s.crb_inertial, 9 * p + k) + e.F(s.crb_inertial, 9 * l + k);
        e.F(s.crb_mass, p) = e.F(s.crb_mass, p) + e.F(s.crb_mass, l);
        e.S3(s.crb_pos, p, e.L3(s.crb_pos, p) + e.L3(s.crb_pos, l));
      }
    }
  }
  GSR_FOR_ENT_LINKS(ie1, l) {
    int d0 = m.link_dof_start[l], d1 = m.link_dof_end[l];
    if (d1 <= d0) continue;
    float I[9];
    for (int k = 0; k < 9; k++) I[k] = e.F(s.crb_inertial, 9 * l + k);
    f3 cp = e.L3(s.crb_pos, l);
    float cm = e.F(s.crb_mass, l);
    for (int d = d0; d < d1; d++) {
      f3 fa, fv;
      inertial_mul(cp, I, cm, e.L3(s.cdof_vel, d), e.L3(s.cdof_ang, d), fa, fv);
      e.S3(s.f_ang, d, fa);
      e.S3(s.f_vel, d, fv);
    }
  }
  for (int ie = e0; ie < E; ie += es) {
    int st = m.ent_dof_start[ie], en = m.ent_dof_end[ie];
    for (int i = st; i < en; i++) {
      f3 fa = e.L3(s.f_ang, i), fv = e.L3(s.f_vel, i);
      for (int j = st; j <= i; j++) {
        float v = (dot(fa, e.L3(s.cdof_ang, j)) + dot(fv, e.L3(s.cdof_vel, j))) * m.mass_parent_mask[(size_t)i * D + j];
        e.Mm(ie, i, j) = v;
      }
    }
    for (int i = st; i < en; i++)
      for (int j = i + 1; j < en; j++) e.Mm(ie, i, j) = e.Mm(ie, j, i);
    // armature, then the implicit damping terms of approximate_implicitfast
    for (int d = st; d < en; d++) e.Mm(ie, d, d) = e.Mm(ie, d, d) + e.DP(m.dof_armature, d);
    for (int d = st; d < en; d++) {
      float v = e.Mm(ie, d, d) + e.DP(m.dof_damping, d) * dt;
      int cm = e.I(s.ctrl_mode, d);
      if (cm == CTRL_POSITION || cm == CTRL_VELOCITY) v = v + e.DP(m.dof_kv, d) * dt;
      e.Mm(ie, d, d) = v;
    }
    // --- LDL^T (rsd:3697-3760) ---
    int n = en - st;
    for (int i = st; i < en; i++)
      for (int j = st; j <= i; j++) e.ML(ie, i, j) = e.Mm(ie, i, j);
    for (int ii = 0; ii < n; ii++) {
      int i = en - ii - 1;
      float dinv = 1.f / e.ML(ie, i, i);
      e.F(s.mass_D_inv, i) = dinv;
      for (int jj = 0; jj < i - st; jj++) {
        int j = i - jj - 1;
        float a = e.ML(ie, i, j) * dinv;
        for (int k = st; k <= j; k++) e.ML(ie, j, k) = e.ML(ie, j, k) - a * e.ML(ie, i, k);
        e.ML(ie, i, j) = a;
      }
      e.ML(ie, i, i) = 1.f;
    }
  }
  // --- actuation and passive forces (rsd:5960-6177) ---
  GSR_FOR_ENT_LINKS(ie2, l) {
    int dstart = m.link_dof_start[l], dend = m.link_dof_end[l];
    if (dend <= dstart) continue;
    int jt = m.joint_type[m.link_joint_start[l]];
    for (int d = dstart; d < dend; d++) {
      float f = 0.f;
      int cm = e.I(s.ctrl_mode, d);
      if (cm == CTRL_FORCE) f = e.F(s.ctrl_force, d);
      else if (cm == CTRL_VELOCITY) f = e.DP(m.dof_kv, d) * (e.F(s.ctrl_vel, d) - e.F(s.vel, d));
      else if (cm == CTRL_POSITION && !(jt == JOINT_FREE && d >= dstart + 3))
        f = e.DP(m.dof_kp, d) * (e.F(s.ctrl_pos, d) - e.F(s.dofs_pos, d)) +
            e.DP(m.dof_kv, d) * (e.F(s.ctrl_vel, d) - e.F(s.vel, d));
      e.F(s.qf_applied, d) =
          fminf(fmaxf(f, e.DP(m.dof_force_range, 2 * d)), e.DP(m.dof_force_range, 2 * d + 1));
    }
    if (jt == JOINT_FREE && (e.I(s.ctrl_mode, dstart + 3) == CTRL_POSITION || e.I(s.ctrl_mode, dstart + 4) == CTRL_POSITION ||
                             e.I(s.ctrl_mode, dstart + 5) == CTRL_POSITION)) {
      f3 xyz{e.F(s.dofs_pos, dstart + 3), e.F(s.dofs_pos, dstart + 4), e.F(s.dofs_pos, dstart + 5)};
      f3 cxyz{e.F(s.ctrl_pos, dstart + 3), e.F(s.ctrl_pos, dstart + 4), e.F(s.ctrl_pos, dstart + 5)};
      q4 qd = qrot_q(xyz_to_quat(cxyz), qinv(xyz_to_quat(xyz)));
      f3 rv = quat_to_rotvec(qd, EPS);
      for (int j = 0; j < 3; j++) {
        int d = dstart + 3 + j;
        float f = e.DP(m.dof_kp, d) * comp(rv, j) - e.DP(m.dof_kv, d) * e.F(s.vel, d);
        e.F(s.qf_applied, d) =
          fminf(fmaxf(f, e.DP(m.dof_force_range, 2 * d)), e.DP(m.dof_force_range, 2 * d + 1));
      }
    }
  }
  GSR_FOR_ENT_DOFS(ie3, d) e.F(s.qf_passive, d) = -e.DP(m.dof_damping, d) * e.F(s.vel, d);
  GSR_FOR_ENT_LINKS(ie4, l) {
    if (m.link_dof_end[l] - m.link_dof_start[l] <= 0) continue;
    int jt = m.joint_type[m.link_joint_start[l]];
    if (jt != JOINT_FREE && jt != JOINT_FIXED)
      for (int d = m.link_dof_start[l]; d < m.link_dof_end[l]; d++)
        e.F(s.qf_passive, d) = e.F(s.qf_passive, d) + -e.F(s.dofs_pos, d) * e.DP(m.dof_stiffness, d);
  }
  // --- bias accelerations (rsd:6180-6281) ---
  for (int ie = e0; ie < E; ie += es) {
    for (int l = m.ent_link_start[ie]; l < m.ent_link_end[ie]; l++) {
      int p = m.link_parent[l];
      f3 cv, ca;
      if (p == -1) {
        cv = -f3{m.gravity[0], m.gravity[1], m.gravity[2]} * (1.f - m.ent_gravity_compensation[ie]);
        ca = zero3();
      } else {
        cv = e.L3(s.cdd_vel, p);
        ca = e.L3(s.cdd_ang, p);
      }
      for (int d = m.link_dof_start[l]; d < m.link_dof_end[l]; d++) {
        float v = e.F(s.vel, d);
        cv = cv + e.L3(s.cdofd_vel, d) * v;
        ca = ca + e.L3(s.cdofd_ang, d) * v;
      }
      e.S3(s.cdd_vel, l, cv);
      e.S3(s.cdd_ang, l, ca);
    }
  }
  // --- link forces and backward accumulation (rsd:6284-6398) ---
  GSR_FOR_ENT_LINKS(ie5, l) {
    float I[9];
    for (int k = 0; k < 9; k++) I[k] = e.F(s.cinr_inertial, 9 * l + k);
    f3 cp = e.L3(s.cinr_pos, l);
    float cm = e.F(s.cinr_mass, l);
    f3 f1a, f1v, f2a, f2v, f3a, f3v;
    f3 cdv = e.L3(s.cd_vel, l), cda = e.L3(s.cd_ang, l);
    inertial_mul(cp, I, cm, e.L3(s.cdd_vel, l), e.L3(s.cdd_ang, l), f1a, f1v);
    inertial_mul(cp, I, cm, cdv, cda, f2a, f2v);
    motion_cross_force(cda, cdv, f2a, f2v, f3a, f3v);
    // applied and coupling forces are zero on this path; their additions are kept for the same rounding
    e.S3(s.cfrc_vel, l, f1v + f3v + zero3() + zero3());
    e.S3(s.cfrc_ang, l, f1a + f3a + zero3() + zero3());
  }
  for (int ie = e0; ie < E; ie += es) {
    int ls = m.ent_link_start[ie], le = m.ent_link_end[ie];
    for (int l = le - 1; l >= ls; l--) {
      int p = m.link_parent[l];
      if (p != -1) {
        e.S3(s.cfrc_vel, p, e.L3(s.cfrc_vel, p) + e.L3(s.cfrc_vel, l));
        e.S3(s.cfrc_ang, p, e.L3(s.cfrc_ang, p) + e.L3(s.cfrc_ang, l));
      }
    }
  }
  // --- generalized bias force, total force, smooth acceleration (rsd:6425-6563) ---
  GSR_FOR_ENT_LINKS(ie6, l) {
    f3 fa = e.L3(s.cfrc_ang, l), fv = e.L3(s.cfrc_vel, l);
    for (int d = m.link_dof_start[l]; d < m.link_dof_end[l]; d++) {
      float qb = dot(e.L3(s.cdof_ang, d), fa) + dot(e.L3(s.cdof_vel, d), fv);
      e.F(s.qf_bias, d) = qb;
      float f = e.F(s.qf_passive, d) - qb + e.F(s.qf_applied, d);
      e.F(s.force, d) = f;
      e.F(s.qf_smooth, d) = f;
    }
  }
  dev_solve_mass(e, s.force, s.acc_smooth, e0, es);
  GSR_FOR_ENT_DOFS(ie7, d) e.F(s.acc, d) = e.F(s.acc_smooth, d);
}

}  // namespace gsr
