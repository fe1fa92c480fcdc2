// CPU oracle for the batched rigid step -- TEST INFRASTRUCTURE ONLY.
//
// A scalar, per-environment restatement of the reference's rigid step
// (genesis/engine/solvers/rigid/rigid_solver_decomp.py, collider_decomp.py, mpr_decomp.py,
//  support_field_decomp.py, constraint_solver_decomp.py), in the reference's loop order, templated on the float
// type so the same code gives the FP32 parity target and the FP64 drift reference. Only tests/, the smoke check and
// bench.py's cpu_baseline leg may load it; the product never does.
//
// Functions without a Taichi counterpart are marked. Parity status per function is listed in oracle/README.md.
#include <cstdint>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>
#include <thread>
#include <limits>

#include "oracle_math.h"

using namespace orc;

namespace {

enum { GEOM_PLANE = 0, GEOM_SPHERE = 1, GEOM_ELLIPSOID = 2, GEOM_CYLINDER = 3, GEOM_CAPSULE = 4, GEOM_BOX = 5, GEOM_MESH = 6 };
enum { JOINT_FIXED = 0, JOINT_REVOLUTE = 1, JOINT_PRISMATIC = 2, JOINT_SPHERICAL = 3, JOINT_FREE = 4 };
enum { EQ_CONNECT = 0, EQ_WELD = 1, EQ_JOINT = 2 };
enum { CTRL_FORCE = 0, CTRL_VELOCITY = 1, CTRL_POSITION = 2 };

// ------------------------------------------------------------------------------------------------------------
// Model tables (host copies, converted to T)
// ------------------------------------------------------------------------------------------------------------
struct RawModel {
  std::unordered_map<std::string, std::vector<int32_t>> ints;
  std::unordered_map<std::string, std::vector<double>> flts;
  const std::vector<int32_t>& I(const char* k) const {
    auto it = ints.find(k);
    if (it == ints.end()) { static std::vector<int32_t> e; return e; }
    return it->second;
  }
  const std::vector<double>& F(const char* k) const {
    auto it = flts.find(k);
    if (it == flts.end()) { static std::vector<double> e; return e; }
    return it->second;
  }
  int scalar_i(const char* k) const { return I(k).empty() ? 0 : I(k)[0]; }
  double scalar_f(const char* k) const { return F(k).empty() ? 0.0 : F(k)[0]; }
};

template <class T>
struct Model {
  int L, J, D, Q, G, E, NEQ, P, max_contacts, max_broad, max_rows, max_dyn_welds;
  int iterations, ls_iterations, n_contacts_per_pair;
  int enable_collision, enable_joint_limit, enable_multi_contact;
  T dt, tolerance, ls_tolerance, EPS, meaninertia, mc_perturbation, mc_tolerance, mpr_to_gjk;
  V3<T> gravity;
  std::vector<int32_t> link_parent, link_root, link_q_start, link_q_end, link_dof_start, link_dof_end,
      link_joint_start, link_joint_end, link_is_fixed, link_entity;
  std::vector<V3<T>> link_pos, link_inertial_pos;
  std::vector<Q4<T>> link_quat, link_inertial_quat;
  std::vector<M3<T>> link_inertial_i;
  std::vector<T> link_inertial_mass, link_invweight;  // invweight (L,2)
  std::vector<int32_t> joint_type, joint_q_start, joint_dof_start, joint_dof_end;
  std::vector<V3<T>> joint_pos;
  std::vector<T> joint_sol_params;  // (J,7)
  std::vector<V3<T>> dof_motion_ang, dof_motion_vel;
  std::vector<T> dof_limit, dof_force_range, dof_kp, dof_kv, dof_damping, dof_armature, dof_stiffness,
      dof_frictionloss, dof_invweight;
  std::vector<int32_t> dof_entity;
  std::vector<T> qpos0;
  std::vector<int32_t> ent_link_start, ent_link_end, ent_dof_start, ent_dof_end;
  std::vector<T> ent_gravity_compensation;
  std::vector<int32_t> geom_type, geom_link, geom_is_fixed, geom_vert_start, geom_vert_end;
  std::vector<V3<T>> geom_pos, geom_center;
  std::vector<Q4<T>> geom_quat;
  std::vector<T> geom_data, geom_friction, geom_sol_params;  // data (G,7), sol (G,7)
  std::vector<V3<T>> geom_init_aabb;                           // (G,8)
  std::vector<V3<T>> verts;
  std::vector<int32_t> eq_type, eq_obj1, eq_obj2;
  std::vector<T> eq_data, eq_sol_params;  // (NEQ,11), (NEQ,7)
  std::vector<T> mass_parent_mask;        // (D,D)
  std::vector<int32_t> pair_idx;          // (G,G)
  // mesh support field (support_field_decomp.py:24-94): 180 x 180 direction cells per geom from its first cell
  std::vector<int32_t> support_cell_start, support_vid;
  std::vector<V3<T>> support_v;

  void load(const RawModel& r) {
    L = r.scalar_i("n_links"); J = r.scalar_i("n_joints"); D = r.scalar_i("n_dofs"); Q = r.scalar_i("n_qs");
    G = r.scalar_i("n_geoms"); E = r.scalar_i("n_entities"); NEQ = r.scalar_i("n_equalities");
    P = r.scalar_i("n_pairs"); max_contacts = r.scalar_i("max_contacts"); max_broad = r.scalar_i("max_broad");
    max_rows = r.scalar_i("max_rows");
    max_dyn_welds = r.scalar_i("max_dyn_welds");
    iterations = r.scalar_i("iterations"); ls_iterations = r.scalar_i("ls_iterations");
    n_contacts_per_pair = r.scalar_i("n_contacts_per_pair");
    enable_collision = r.scalar_i("enable_collision"); enable_joint_limit = r.scalar_i("enable_joint_limit");
    enable_multi_contact = r.scalar_i("enable_multi_contact");
    dt = T(r.scalar_f("dt")); tolerance = T(r.scalar_f("tolerance")); ls_tolerance = T(r.scalar_f("ls_tolerance"));
    EPS = T(r.scalar_f("EPS")); meaninertia = T(r.scalar_f("meaninertia"));
    mc_perturbation = T(r.scalar_f("mc_perturbation")); mc_tolerance = T(r.scalar_f("mc_tolerance"));
    mpr_to_gjk = T(r.scalar_f("mpr_to_gjk_overlap_ratio"));
    auto g = r.F("gravity"); gravity = V3<T>(T(g[0]), T(g[1]), T(g[2]));
    auto cp = [&](const char* k, std::vector<int32_t>& dst) { dst = r.I(k); };
    auto cf = [&](const char* k, std::vector<T>& dst) { auto& s = r.F(k); dst.assign(s.begin(), s.end()); };
    auto cv3 = [&](const char* k, std::vector<V3<T>>& dst) {
      auto& s = r.F(k); dst.resize(s.size() / 3);
      for (size_t i = 0; i < dst.size(); i++) dst[i] = V3<T>(T(s[3 * i]), T(s[3 * i + 1]), T(s[3 * i + 2]));
    };
    auto cq4 = [&](const char* k, std::vector<Q4<T>>& dst) {
      auto& s = r.F(k); dst.resize(s.size() / 4);
      for (size_t i = 0; i < dst.size(); i++) dst[i] = Q4<T>(T(s[4 * i]), T(s[4 * i + 1]), T(s[4 * i + 2]), T(s[4 * i + 3]));
    };
    cp("link_parent", link_parent); cp("link_root", link_root); cp("link_q_start", link_q_start);
    cp("link_q_end", link_q_end); cp("link_dof_start", link_dof_start); cp("link_dof_end", link_dof_end);
    cp("link_joint_start", link_joint_start); cp("link_joint_end", link_joint_end);
    cp("link_is_fixed", link_is_fixed); cp("link_entity", link_entity);
    cv3("link_pos", link_pos); cv3("link_inertial_pos", link_inertial_pos);
    cq4("link_quat", link_quat); cq4("link_inertial_quat", link_inertial_quat);
    {
      auto& s = r.F("link_inertial_i"); link_inertial_i.resize(L);
      for (int l = 0; l < L; l++) for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) link_inertial_i[l].m[i][j] = T(s[9 * l + 3 * i + j]);
    }
    cf("link_inertial_mass", link_inertial_mass); cf("link_invweight", link_invweight);
    cp("joint_type", joint_type); cp("joint_q_start", joint_q_start); cp("joint_dof_start", joint_dof_start);
    cp("joint_dof_end", joint_dof_end); cv3("joint_pos", joint_pos); cf("joint_sol_params", joint_sol_params);
    cv3("dof_motion_ang", dof_motion_ang); cv3("dof_motion_vel", dof_motion_vel);
    cf("dof_limit", dof_limit); cf("dof_force_range", dof_force_range); cf("dof_kp", dof_kp); cf("dof_kv", dof_kv);
    cf("dof_damping", dof_damping); cf("dof_armature", dof_armature); cf("dof_stiffness", dof_stiffness);
    cf("dof_frictionloss", dof_frictionloss); cf("dof_invweight", dof_invweight); cp("dof_entity", dof_entity);
    cf("qpos0", qpos0);
    cp("ent_link_start", ent_link_start); cp("ent_link_end", ent_link_end); cp("ent_dof_start", ent_dof_start);
    cp("ent_dof_end", ent_dof_end); cf("ent_gravity_compensation", ent_gravity_compensation);
    cp("geom_type", geom_type); cp("geom_link", geom_link); cp("geom_is_fixed", geom_is_fixed);
    cp("geom_vert_start", geom_vert_start); cp("geom_vert_end", geom_vert_end);
    cv3("geom_pos", geom_pos); cv3("geom_center", geom_center); cq4("geom_quat", geom_quat);
    cf("geom_data", geom_data); cf("geom_friction", geom_friction); cf("geom_sol_params", geom_sol_params);
    cv3("geom_init_aabb", geom_init_aabb); cv3("verts", verts);
    cp("eq_type", eq_type); cp("eq_obj1", eq_obj1); cp("eq_obj2", eq_obj2);
    cf("eq_data", eq_data); cf("eq_sol_params", eq_sol_params);
    cf("mass_parent_mask", mass_parent_mask); cp("pair_idx", pair_idx);
    cp("support_cell_start", support_cell_start); cp("support_vid", support_vid); cv3("support_v", support_v);
  }
};

// ------------------------------------------------------------------------------------------------------------
// Per-environment state
// ------------------------------------------------------------------------------------------------------------
template <class T>
struct Contact {
  int geom_a, geom_b, link_a, link_b;
  V3<T> pos, normal, force;
  T penetration, friction;
  T sol_params[7];
};

template <class T>
struct Env {
  // persistent
  std::vector<T> qpos, vel, qacc_ws;
  int is_warmstart = 0;
  std::vector<int32_t> ctrl_mode;
  std::vector<T> ctrl_pos, ctrl_vel, ctrl_force;
  std::vector<V3<T>> cache_normal;  // (P)
  std::vector<T> mass_shift, friction_ratio;  // links_state.mass_shift (L), geoms_state.friction_ratio (G)
  std::vector<V3<T>> i_pos_shift;             // links_state.i_pos_shift (L)
  std::vector<T> sort_value;
  std::vector<int32_t> sort_ig, sort_ismax;
  // welds registered after the build (constraint_state.ti_n_equalities beyond the model's): count, link pairs, data
  int eq_dyn_n = 0;
  std::vector<int32_t> eq_dyn_obj;
  std::vector<T> eq_dyn_data;
  int first_time = 1;
  uint32_t errno_ = 0;
  // kinematics
  std::vector<V3<T>> l_pos, l_ipos_bw, l_ipos, root_com, root_com_bw, j_pos, cd_vel, cd_ang, cdd_vel, cdd_ang,
      cinr_pos, cfrc_vel, cfrc_ang, crb_pos;
  std::vector<Q4<T>> l_quat, l_iquat, j_quat;
  std::vector<T> mass_sum, cinr_mass, crb_mass;
  std::vector<M3<T>> cinr_I, crb_I;
  std::vector<V3<T>> xanchor, xaxis;
  std::vector<V3<T>> cdof_ang, cdof_vel, cdofd_ang, cdofd_vel, f_ang, f_vel;
  std::vector<T> dofs_pos;
  std::vector<V3<T>> g_pos, aabb_min, aabb_max;
  std::vector<Q4<T>> g_quat;
  // dynamics
  std::vector<T> mass_mat, mass_L, mass_D_inv;
  std::vector<T> qf_applied, qf_passive, qf_bias, force, qf_smooth, acc_smooth, acc;
  // collision
  std::vector<int32_t> active_buffer;
  std::vector<int32_t> broad_a, broad_b;
  int n_broad = 0;
  std::vector<Contact<T>> contacts;
  int n_contacts = 0;
  // constraints
  int n_con = 0, n_con_eq = 0, n_con_fl = 0;
  std::vector<T> jac, diag, aref, efc_D, efc_fl, efc_force, Jaref, jv, quad;
  std::vector<int32_t> active, prev_active;
  std::vector<T> qacc, Ma, grad, Mgrad, search, mv, qfrc_constraint, nt_H, nt_vec;
  T gauss = 0, cost = 0, prev_cost = 0;
  T quad_gauss[3] = {0, 0, 0};
  int improved = 0, ls_it = 0, n_newton_iters = 0;
  T candidates[12];
  std::vector<T> vel_next, qpos_next;

  void alloc(const Model<T>& m) {
    int L = m.L, D = m.D, G = m.G, R = m.max_rows;
    qpos.assign(m.Q, 0); vel.assign(D, 0); qacc_ws.assign(D, 0);
    ctrl_mode.assign(D, CTRL_FORCE); ctrl_pos.assign(D, 0); ctrl_vel.assign(D, 0); ctrl_force.assign(D, 0);
    cache_normal.assign(std::max(m.P, 1), V3<T>());
    mass_shift.assign(m.L, 0); i_pos_shift.assign(m.L, V3<T>()); friction_ratio.assign(m.G, 1);
    sort_value.assign(2 * G, 0); sort_ig.assign(2 * G, 0); sort_ismax.assign(2 * G, 0);
    eq_dyn_obj.assign(2 * (size_t)m.max_dyn_welds, 0); eq_dyn_data.assign(11 * (size_t)m.max_dyn_welds, 0);
    for (auto* v : {&l_pos, &l_ipos_bw, &l_ipos, &root_com, &root_com_bw, &j_pos, &cd_vel, &cd_ang, &cdd_vel, &cdd_ang,
                    &cinr_pos, &cfrc_vel, &cfrc_ang, &crb_pos})
      v->assign(L, V3<T>());
    l_quat.assign(L, Q4<T>()); l_iquat.assign(L, Q4<T>()); j_quat.assign(L, Q4<T>());
    mass_sum.assign(L, 0); cinr_mass.assign(L, 0); crb_mass.assign(L, 0);
    cinr_I.assign(L, M3<T>()); crb_I.assign(L, M3<T>());
    xanchor.assign(m.J, V3<T>()); xaxis.assign(m.J, V3<T>());
    for (auto* v : {&cdof_ang, &cdof_vel, &cdofd_ang, &cdofd_vel, &f_ang, &f_vel}) v->assign(D, V3<T>());
    dofs_pos.assign(D, 0);
    g_pos.assign(G, V3<T>()); aabb_min.assign(G, V3<T>()); aabb_max.assign(G, V3<T>()); g_quat.assign(G, Q4<T>());
    mass_mat.assign((size_t)D * D, 0); mass_L.assign((size_t)D * D, 0); mass_D_inv.assign(D, 0);
    for (auto* v : {&qf_applied, &qf_passive, &qf_bias, &force, &qf_smooth, &acc_smooth, &acc}) v->assign(D, 0);
    active_buffer.assign(G, 0);
    broad_a.assign(std::max(m.max_broad, 1), 0); broad_b.assign(std::max(m.max_broad, 1), 0);
    contacts.assign(std::max(m.max_contacts, 1), Contact<T>());
    jac.assign((size_t)R * D, 0);
    for (auto* v : {&diag, &aref, &efc_D, &efc_fl, &efc_force, &Jaref, &jv}) v->assign(R, 0);
    quad.assign((size_t)R * 3, 0);
    active.assign(R, 0); prev_active.assign(R, 0);
    for (auto* v : {&qacc, &Ma, &grad, &Mgrad, &search, &mv, &qfrc_constraint, &nt_vec, &vel_next}) v->assign(D, 0);
    nt_H.assign((size_t)D * D, 0);
    qpos_next.assign(m.Q, 0);
  }
};

// ------------------------------------------------------------------------------------------------------------
// Kinematics: func_forward_kinematics_entity (rsd:5219-5376)
// ------------------------------------------------------------------------------------------------------------
template <class T>
void forward_kinematics(const Model<T>& m, Env<T>& e) {
  const T EPS = m.EPS;
  for (int i_e = 0; i_e < m.E; i_e++) {
    for (int i_l = m.ent_link_start[i_e]; i_l < m.ent_link_end[i_e]; i_l++) {
      V3<T> pos = m.link_pos[i_l];
      Q4<T> quat = m.link_quat[i_l];
      int par = m.link_parent[i_l];
      if (par != -1) {
        V3<T> ppos = e.l_pos[par];
        Q4<T> pquat = e.l_quat[par];
        pos = ppos + transform_by_quat(m.link_pos[i_l], pquat);
        quat = transform_quat_by_quat(m.link_quat[i_l], pquat);
      }
      for (int i_j = m.link_joint_start[i_l]; i_j < m.link_joint_end[i_l]; i_j++) {
        int jt = m.joint_type[i_j], qs = m.joint_q_start[i_j], ds = m.joint_dof_start[i_j];
        if (jt == JOINT_FREE) {
          e.xanchor[i_j] = V3<T>(e.qpos[qs], e.qpos[qs + 1], e.qpos[qs + 2]);
          e.xaxis[i_j] = V3<T>(0, 0, 1);
        } else if (jt != JOINT_FIXED) {
          V3<T> axis(0, 0, 1);
          if (jt == JOINT_REVOLUTE) axis = m.dof_motion_ang[ds];
          else if (jt == JOINT_PRISMATIC) axis = m.dof_motion_vel[ds];
          e.xanchor[i_j] = transform_by_quat(m.joint_pos[i_j], quat) + pos;
          e.xaxis[i_j] = transform_by_quat(axis, quat);
        }
        if (jt == JOINT_FREE) {
          pos = V3<T>(e.qpos[qs], e.qpos[qs + 1], e.qpos[qs + 2]);
          quat = Q4<T>(e.qpos[qs + 3], e.qpos[qs + 4], e.qpos[qs + 5], e.qpos[qs + 6]);
          V3<T> xyz = quat_to_xyz(quat, EPS);
          for (int j = 0; j < 3; j++) { e.dofs_pos[ds + j] = pos[j]; e.dofs_pos[ds + 3 + j] = xyz[j]; }
        } else if (jt == JOINT_FIXED) {
        } else if (jt == JOINT_SPHERICAL) {
          Q4<T> qloc(e.qpos[qs], e.qpos[qs + 1], e.qpos[qs + 2], e.qpos[qs + 3]);
          V3<T> xyz = quat_to_xyz(qloc, EPS);
          for (int j = 0; j < 3; j++) e.dofs_pos[ds + j] = xyz[j];
          quat = transform_quat_by_quat(qloc, quat);
          pos = e.xanchor[i_j] - transform_by_quat(m.joint_pos[i_j], quat);
        } else if (jt == JOINT_REVOLUTE) {
          V3<T> axis = m.dof_motion_ang[ds];
          e.dofs_pos[ds] = e.qpos[qs] - m.qpos0[qs];
          Q4<T> qloc = rotvec_to_quat(axis * e.dofs_pos[ds], EPS);
          quat = transform_quat_by_quat(qloc, quat);
          pos = e.xanchor[i_j] - transform_by_quat(m.joint_pos[i_j], quat);
        } else {  // PRISMATIC
          e.dofs_pos[ds] = e.qpos[qs] - m.qpos0[qs];
          pos = pos + e.xaxis[i_j] * e.dofs_pos[ds];
        }
      }
      if (!(par == -1 && m.link_is_fixed[i_l])) {
        e.l_pos[i_l] = pos;
        e.l_quat[i_l] = quat;
      } else {
        // fixed root links keep their build-time pose (set once by the reference's init kernels)
        e.l_pos[i_l] = m.link_pos[i_l];
        e.l_quat[i_l] = m.link_quat[i_l];
      }
    }
  }
}

// func_COM_links (rsd:4746-5078)
template <class T>
void com_links(const Model<T>& m, Env<T>& e) {
  const T EPS = m.EPS;
  for (int l = 0; l < m.L; l++) { e.root_com_bw[l] = V3<T>(); e.mass_sum[l] = 0; }
  for (int l = 0; l < m.L; l++) {
    // per-env mass and COM shifts (rsd:4812-4817)
    T mass = m.link_inertial_mass[l] + e.mass_shift[l];
    e.l_ipos_bw[l] = e.l_pos[l] + transform_by_quat(m.link_inertial_pos[l] + e.i_pos_shift[l], e.l_quat[l]);
    e.l_iquat[l] = transform_quat_by_quat(m.link_inertial_quat[l], e.l_quat[l]);
    int r = m.link_root[l];
    e.mass_sum[r] += mass;
    e.root_com_bw[r] = e.root_com_bw[r] + mass * e.l_ipos_bw[l];
  }
  for (int l = 0; l < m.L; l++)
    if (m.link_root[l] == l) e.root_com[l] = e.root_com_bw[l] / e.mass_sum[l];
  for (int l = 0; l < m.L; l++) e.root_com[l] = e.root_com[m.link_root[l]];
  for (int l = 0; l < m.L; l++) {
    e.l_ipos[l] = e.l_ipos_bw[l] - e.root_com[l];
    // ti_transform_inertia_by_trans_quat (geom.py:293-309)
    V3<T> t = e.l_ipos[l];
    T mass = m.link_inertial_mass[l] + e.mass_shift[l];  // rsd:4910
    T x = t.x, y = t.y, z = t.z;
    T xx = x * x, xy = x * y, xz = x * z, yy = y * y, yz = y * z, zz = z * z;
    M3<T> hhT;
    hhT.m[0][0] = yy + zz; hhT.m[0][1] = -xy; hhT.m[0][2] = -xz;
    hhT.m[1][0] = -xy; hhT.m[1][1] = xx + zz; hhT.m[1][2] = -yz;
    hhT.m[2][0] = -xz; hhT.m[2][1] = -yz; hhT.m[2][2] = xx + yy;
    M3<T> R = quat_to_R(e.l_iquat[l], EPS);
    M3<T> I = mul(mul(R, m.link_inertial_i[l]), transpose(R));
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) I.m[i][j] = I.m[i][j] + hhT.m[i][j] * mass;
    e.cinr_I[l] = I;
    e.cinr_pos[l] = t * mass;
    e.cinr_mass[l] = mass;
  }
  // joint frames: only j_pos/j_quat, unused by the step path; skipped.
  for (int l = 0; l < m.L; l++) {
    if (m.link_dof_end[l] - m.link_dof_start[l] <= 0) continue;
    for (int i_j = m.link_joint_start[l]; i_j < m.link_joint_end[l]; i_j++) {
      V3<T> off = e.root_com[l] - e.xanchor[i_j];
      int jt = m.joint_type[i_j], ds = m.joint_dof_start[i_j];
      if (jt == JOINT_REVOLUTE) {
        e.cdof_ang[ds] = e.xaxis[i_j];
        e.cdof_vel[ds] = cross(e.xaxis[i_j], off);
      } else if (jt == JOINT_PRISMATIC) {
        e.cdof_ang[ds] = V3<T>();
        e.cdof_vel[ds] = e.xaxis[i_j];
      } else if (jt == JOINT_SPHERICAL) {
        M3<T> xT = transpose(quat_to_R(e.l_quat[l], EPS));
        for (int i = 0; i < 3; i++) { e.cdof_ang[ds + i] = xT.row(i); e.cdof_vel[ds + i] = cross(xT.row(i), off); }
      } else if (jt == JOINT_FREE) {
        for (int i = 0; i < 3; i++) { e.cdof_ang[ds + i] = V3<T>(); V3<T> v; v[i] = 1; e.cdof_vel[ds + i] = v; }
        M3<T> xT = transpose(quat_to_R(e.l_quat[l], EPS));
        for (int i = 0; i < 3; i++) { e.cdof_ang[ds + i + 3] = xT.row(i); e.cdof_vel[ds + i + 3] = cross(xT.row(i), off); }
      }
    }
  }
}

// func_update_geoms (rsd:5533-5595) -- fixed geoms are computed once at build (force_update_fixed_geoms)
template <class T>
void update_geoms(const Model<T>& m, Env<T>& e) {
  for (int g = 0; g < m.G; g++) {
    int l = m.geom_link[g];
    e.g_pos[g] = e.l_pos[l] + transform_by_quat(m.geom_pos[g], e.l_quat[l]);
    e.g_quat[g] = transform_quat_by_quat(m.geom_quat[g], e.l_quat[l]);
  }
}

// func_forward_velocity_entity (rsd:5379-5503)
template <class T>
void forward_velocity(const Model<T>& m, Env<T>& e) {
  for (int i_e = 0; i_e < m.E; i_e++) {
    for (int l = m.ent_link_start[i_e]; l < m.ent_link_end[i_e]; l++) {
      V3<T> cvel_vel, cvel_ang;
      int par = m.link_parent[l];
      if (par != -1) { cvel_vel = e.cd_vel[par]; cvel_ang = e.cd_ang[par]; }
      for (int i_j = m.link_joint_start[l]; i_j < m.link_joint_end[l]; i_j++) {
        int jt = m.joint_type[i_j], ds = m.joint_dof_start[i_j], de = m.joint_dof_end[i_j];
        if (jt == JOINT_FREE) {
          for (int i = 0; i < 3; i++) {
            cvel_vel = cvel_vel + e.cdof_vel[ds + i] * e.vel[ds + i];
            cvel_ang = cvel_ang + e.cdof_ang[ds + i] * e.vel[ds + i];
          }
          for (int i = 0; i < 3; i++) {
            e.cdofd_ang[ds + i] = V3<T>(); e.cdofd_vel[ds + i] = V3<T>();
            motion_cross_motion(cvel_ang, cvel_vel, e.cdof_ang[ds + i + 3], e.cdof_vel[ds + i + 3], e.cdofd_ang[ds + i + 3], e.cdofd_vel[ds + i + 3]);
          }
          for (int i = 0; i < 3; i++) {
            cvel_vel = cvel_vel + e.cdof_vel[ds + i + 3] * e.vel[ds + i + 3];
            cvel_ang = cvel_ang + e.cdof_ang[ds + i + 3] * e.vel[ds + i + 3];
          }
        } else {
          for (int d = ds; d < de; d++) motion_cross_motion(cvel_ang, cvel_vel, e.cdof_ang[d], e.cdof_vel[d], e.cdofd_ang[d], e.cdofd_vel[d]);
          for (int d = ds; d < de; d++) {
            cvel_vel = cvel_vel + e.cdof_vel[d] * e.vel[d];
            cvel_ang = cvel_ang + e.cdof_ang[d] * e.vel[d];
          }
        }
      }
      e.cd_vel[l] = cvel_vel;
      e.cd_ang[l] = cvel_ang;
    }
  }
}

// ------------------------------------------------------------------------------------------------------------
// Dynamics
// ------------------------------------------------------------------------------------------------------------
// func_compute_mass_matrix (rsd:3422-3678), implicit_damping for approximate_implicitfast
template <class T>
void compute_mass_matrix(const Model<T>& m, Env<T>& e) {
  const int D = m.D;
  for (int l = 0; l < m.L; l++) {
    e.crb_I[l] = e.cinr_I[l]; e.crb_pos[l] = e.cinr_pos[l]; e.crb_mass[l] = e.cinr_mass[l];
  }
  for (int i_e = 0; i_e < m.E; i_e++) {
    int nl = m.ent_link_end[i_e] - m.ent_link_start[i_e];
    for (int i = 0; i < nl; i++) {
      int l = m.ent_link_end[i_e] - 1 - i;
      int p = m.link_parent[l];
      if (p != -1) {
        for (int a = 0; a < 3; a++) for (int b = 0; b < 3; b++) e.crb_I[p].m[a][b] = e.crb_I[p].m[a][b] + e.crb_I[l].m[a][b];
        e.crb_mass[p] = e.crb_mass[p] + e.crb_mass[l];
        e.crb_pos[p] = e.crb_pos[p] + e.crb_pos[l];
      }
    }
  }
  for (int l = 0; l < m.L; l++)
    for (int d = m.link_dof_start[l]; d < m.link_dof_end[l]; d++)
      inertial_mul(e.crb_pos[l], e.crb_I[l], e.crb_mass[l], e.cdof_vel[d], e.cdof_ang[d], e.f_ang[d], e.f_vel[d]);
  for (int i_e = 0; i_e < m.E; i_e++) {
    int s = m.ent_dof_start[i_e], t = m.ent_dof_end[i_e];
    for (int i = s; i < t; i++)
      for (int j = s; j < t; j++)
        e.mass_mat[(size_t)i * D + j] = (dot(e.f_ang[i], e.cdof_ang[j]) + dot(e.f_vel[i], e.cdof_vel[j])) * m.mass_parent_mask[(size_t)i * D + j];
    for (int i = s; i < t; i++)
      for (int j = i + 1; j < t; j++) e.mass_mat[(size_t)i * D + j] = e.mass_mat[(size_t)j * D + i];
  }
  for (int d = 0; d < D; d++) e.mass_mat[(size_t)d * D + d] += m.dof_armature[d];
  for (int d = 0; d < D; d++) {
    e.mass_mat[(size_t)d * D + d] += m.dof_damping[d] * m.dt;
    if (e.ctrl_mode[d] == CTRL_POSITION || e.ctrl_mode[d] == CTRL_VELOCITY) e.mass_mat[(size_t)d * D + d] += m.dof_kv[d] * m.dt;
  }
}

// func_factor_mass forward path (rsd:3697-3760), implicit_damping=False
template <class T>
void factor_mass(const Model<T>& m, Env<T>& e) {
  const int D = m.D;
  for (int i_e = 0; i_e < m.E; i_e++) {
    int s = m.ent_dof_start[i_e], t = m.ent_dof_end[i_e], n = t - s;
    for (int i = s; i < t; i++)
      for (int j = s; j <= i; j++) e.mass_L[(size_t)i * D + j] = e.mass_mat[(size_t)i * D + j];
    for (int ii = 0; ii < n; ii++) {
      int i = t - ii - 1;
      e.mass_D_inv[i] = T(1) / e.mass_L[(size_t)i * D + i];
      for (int jj = 0; jj < i - s; jj++) {
        int j = i - jj - 1;
        T a = e.mass_L[(size_t)i * D + j] * e.mass_D_inv[i];
        for (int k = s; k <= j; k++) e.mass_L[(size_t)j * D + k] -= a * e.mass_L[(size_t)i * D + k];
        e.mass_L[(size_t)i * D + j] = a;
      }
      e.mass_L[(size_t)i * D + i] = 1;
    }
  }
}

// func_solve_mass_batched (rsd:3932-4042)
template <class T>
void solve_mass(const Model<T>& m, const Env<T>& e, const std::vector<T>& vec, std::vector<T>& out) {
  const int D = m.D;
  for (int i_e = 0; i_e < m.E; i_e++) {
    int s = m.ent_dof_start[i_e], t = m.ent_dof_end[i_e], n = t - s;
    for (int ii = 0; ii < n; ii++) {
      int i = t - ii - 1;
      out[i] = vec[i];
      for (int j = i + 1; j < t; j++) out[i] -= e.mass_L[(size_t)j * D + i] * out[j];
    }
    for (int i = s; i < t; i++) out[i] *= e.mass_D_inv[i];
    for (int i = s; i < t; i++)
      for (int j = s; j < i; j++) out[i] -= e.mass_L[(size_t)i * D + j] * out[j];
  }
}

// func_torque_and_passive_force (rsd:5960-6177)
template <class T>
void torque_and_passive(const Model<T>& m, Env<T>& e) {
  const T EPS = m.EPS;
  for (int i_e = 0; i_e < m.E; i_e++) {
    for (int l = m.ent_link_start[i_e]; l < m.ent_link_end[i_e]; l++) {
      if (m.link_dof_end[l] - m.link_dof_start[l] <= 0) continue;
      int jt = m.joint_type[m.link_joint_start[l]];
      int dstart = m.link_dof_start[l];
      for (int d = dstart; d < m.link_dof_end[l]; d++) {
        T f = 0;
        if (e.ctrl_mode[d] == CTRL_FORCE) f = e.ctrl_force[d];
        else if (e.ctrl_mode[d] == CTRL_VELOCITY) f = m.dof_kv[d] * (e.ctrl_vel[d] - e.vel[d]);
        else if (e.ctrl_mode[d] == CTRL_POSITION && !(jt == JOINT_FREE && d >= dstart + 3))
          f = m.dof_kp[d] * (e.ctrl_pos[d] - e.dofs_pos[d]) + m.dof_kv[d] * (e.ctrl_vel[d] - e.vel[d]);
        e.qf_applied[d] = std::min(std::max(f, m.dof_force_range[2 * d]), m.dof_force_range[2 * d + 1]);
      }
      if (jt == JOINT_FREE && (e.ctrl_mode[dstart + 3] == CTRL_POSITION || e.ctrl_mode[dstart + 4] == CTRL_POSITION ||
                               e.ctrl_mode[dstart + 5] == CTRL_POSITION)) {
        V3<T> xyz(e.dofs_pos[dstart + 3], e.dofs_pos[dstart + 4], e.dofs_pos[dstart + 5]);
        V3<T> cxyz(e.ctrl_pos[dstart + 3], e.ctrl_pos[dstart + 4], e.ctrl_pos[dstart + 5]);
        Q4<T> q = xyz_to_quat(xyz), cq = xyz_to_quat(cxyz);
        Q4<T> qd = transform_quat_by_quat(cq, inv_quat(q));
        V3<T> rv = quat_to_rotvec(qd, EPS);
        for (int j = 0; j < 3; j++) {
          int d = dstart + 3 + j;
          T f = m.dof_kp[d] * rv[j] - m.dof_kv[d] * e.vel[d];
          e.qf_applied[d] = std::min(std::max(f, m.dof_force_range[2 * d]), m.dof_force_range[2 * d + 1]);
        }
      }
    }
  }
  for (int d = 0; d < m.D; d++) e.qf_passive[d] = -m.dof_damping[d] * e.vel[d];
  for (int l = 0; l < m.L; l++) {
    if (m.link_dof_end[l] - m.link_dof_start[l] <= 0) continue;
    int jt = m.joint_type[m.link_joint_start[l]];
    if (jt != JOINT_FREE && jt != JOINT_FIXED)
      for (int d = m.link_dof_start[l]; d < m.link_dof_end[l]; d++) e.qf_passive[d] += -e.dofs_pos[d] * m.dof_stiffness[d];
  }
}

// func_update_acc (rsd:6180-6281), update_cacc=False
template <class T>
void update_acc(const Model<T>& m, Env<T>& e) {
  for (int i_e = 0; i_e < m.E; i_e++) {
    for (int l = m.ent_link_start[i_e]; l < m.ent_link_end[i_e]; l++) {
      int p = m.link_parent[l];
      if (p == -1) {
        e.cdd_vel[l] = -m.gravity * (T(1) - m.ent_gravity_compensation[i_e]);
        e.cdd_ang[l] = V3<T>();
      } else {
        e.cdd_vel[l] = e.cdd_vel[p];
        e.cdd_ang[l] = e.cdd_ang[p];
      }
      for (int d = m.link_dof_start[l]; d < m.link_dof_end[l]; d++) {
        e.cdd_vel[l] = e.cdd_vel[l] + e.cdofd_vel[d] * e.vel[d];
        e.cdd_ang[l] = e.cdd_ang[l] + e.cdofd_ang[d] * e.vel[d];
      }
    }
  }
}

// func_update_force (rsd:6284-6398); applied and coupling forces are zero on this path
template <class T>
void update_force(const Model<T>& m, Env<T>& e) {
  for (int l = 0; l < m.L; l++) {
    V3<T> f1a, f1v, f2a, f2v, f3a, f3v;
    inertial_mul(e.cinr_pos[l], e.cinr_I[l], e.cinr_mass[l], e.cdd_vel[l], e.cdd_ang[l], f1a, f1v);
    inertial_mul(e.cinr_pos[l], e.cinr_I[l], e.cinr_mass[l], e.cd_vel[l], e.cd_ang[l], f2a, f2v);
    motion_cross_force(e.cd_ang[l], e.cd_vel[l], f2a, f2v, f3a, f3v);
    V3<T> zero;
    e.cfrc_vel[l] = f1v + f3v + zero + zero;
    e.cfrc_ang[l] = f1a + f3a + zero + zero;
  }
  for (int i_e = 0; i_e < m.E; i_e++) {
    int nl = m.ent_link_end[i_e] - m.ent_link_start[i_e];
    for (int i = 0; i < nl; i++) {
      int l = m.ent_link_end[i_e] - 1 - i;
      int p = m.link_parent[l];
      if (p != -1) { e.cfrc_vel[p] = e.cfrc_vel[p] + e.cfrc_vel[l]; e.cfrc_ang[p] = e.cfrc_ang[p] + e.cfrc_ang[l]; }
    }
  }
}

// func_bias_force (rsd:6425-6485) + func_compute_qacc (rsd:6505-6563)
template <class T>
void bias_and_qacc(const Model<T>& m, Env<T>& e) {
  for (int l = 0; l < m.L; l++)
    for (int d = m.link_dof_start[l]; d < m.link_dof_end[l]; d++) {
      e.qf_bias[d] = dot(e.cdof_ang[d], e.cfrc_ang[l]) + dot(e.cdof_vel[d], e.cfrc_vel[l]);
      e.force[d] = e.qf_passive[d] - e.qf_bias[d] + e.qf_applied[d];
      e.qf_smooth[d] = e.force[d];
    }
  solve_mass(m, e, e.force, e.acc_smooth);
  for (int d = 0; d < m.D; d++) e.acc[d] = e.acc_smooth[d];
}

// ------------------------------------------------------------------------------------------------------------
// Collision
// ------------------------------------------------------------------------------------------------------------
// kernel_update_geom_aabbs (rsd:5668-5689)
template <class T>
void update_aabbs(const Model<T>& m, Env<T>& e) {
  T inf = std::numeric_limits<T>::infinity();
  for (int g = 0; g < m.G; g++) {
    V3<T> lo(inf, inf, inf), hi(-inf, -inf, -inf);
    for (int c = 0; c < 8; c++) {
      V3<T> p = transform_by_trans_quat(m.geom_init_aabb[8 * g + c], e.g_pos[g], e.g_quat[g]);
      lo = vmin(lo, p);
      hi = vmax(hi, p);
    }
    e.aabb_min[g] = lo;
    e.aabb_max[g] = hi;
  }
}

template <class T>
bool aabb_overlap(const Env<T>& e, int a, int b) {
  // collider_decomp.py:738-749 (strict)
  const V3<T>&amax = e.aabb_max[a], &amin = e.aabb_min[a], &bmax = e.aabb_max[b], &bmin = e.aabb_min[b];
  bool sep = (amax.x <= bmin.x) || (amax.y <= bmin.y) || (amax.z <= bmin.z) || (amin.x >= bmax.x) ||
             (amin.y >= bmax.y) || (amin.z >= bmax.z);
  return !sep;
}

// func_broad_phase (collider_decomp.py:1331-1470), no hibernation, no dynamic welds
template <class T>
void broad_phase(const Model<T>& m, Env<T>& e) {
  const int G = m.G;
  e.n_contacts = 0;
  if (e.first_time) {
    for (int i = 0; i < G; i++) {
      e.sort_value[2 * i] = e.aabb_min[i].x; e.sort_ig[2 * i] = i; e.sort_ismax[2 * i] = 0;
      e.sort_value[2 * i + 1] = e.aabb_max[i].x; e.sort_ig[2 * i + 1] = i; e.sort_ismax[2 * i + 1] = 1;
    }
    e.first_time = 0;
  } else {
    for (int i = 0; i < 2 * G; i++)
      e.sort_value[i] = e.sort_ismax[i] ? e.aabb_max[e.sort_ig[i]].x : e.aabb_min[e.sort_ig[i]].x;
  }
  for (int i = 1; i < 2 * G; i++) {
    T kv = e.sort_value[i];
    int kmax = e.sort_ismax[i], kig = e.sort_ig[i];
    int j = i - 1;
    while (j >= 0 && kv < e.sort_value[j]) {
      e.sort_value[j + 1] = e.sort_value[j]; e.sort_ismax[j + 1] = e.sort_ismax[j]; e.sort_ig[j + 1] = e.sort_ig[j];
      j--;
    }
    e.sort_value[j + 1] = kv; e.sort_ismax[j + 1] = kmax; e.sort_ig[j + 1] = kig;
  }
  int n_broad = 0, n_active = 0;
  for (int i = 0; i < 2 * G; i++) {
    if (!e.sort_ismax[i]) {
      for (int j = 0; j < n_active; j++) {
        int ga = e.active_buffer[j], gb = e.sort_ig[i];
        if (ga > gb) std::swap(ga, gb);
        if (m.pair_idx[(size_t)ga * G + gb] == -1) continue;
        if (!aabb_overlap(e, ga, gb)) {
          e.cache_normal[m.pair_idx[(size_t)ga * G + gb]] = V3<T>();
          continue;
        }
        if (n_broad == m.max_broad) { e.errno_ |= 1u; break; }
        e.broad_a[n_broad] = ga; e.broad_b[n_broad] = gb; n_broad++;
      }
      e.active_buffer[n_active++] = e.sort_ig[i];
    } else {
      int rm = e.sort_ig[i];
      for (int j = 0; j < n_active; j++) {
        if (e.active_buffer[j] == rm) {
          for (int k = j; k < n_active - 1; k++) e.active_buffer[k] = e.active_buffer[k + 1];
          n_active--;
          break;
        }
      }
    }
  }
  e.n_broad = n_broad;
}

// func_add_contact (collider_decomp.py:2145-2179)
template <class T>
void add_contact(const Model<T>& m, Env<T>& e, int ga, int gb, V3<T> normal, V3<T> pos, T pen) {
  int i = e.n_contacts;
  if (i < m.max_contacts) {
    Contact<T>& c = e.contacts[i];
    // per-env friction ratios (collider_decomp.py:2161-2162)
    T fa = m.geom_friction[ga] * e.friction_ratio[ga], fb = m.geom_friction[gb] * e.friction_ratio[gb];
    c.geom_a = ga; c.geom_b = gb; c.normal = normal; c.pos = pos; c.penetration = pen;
    c.friction = std::max(std::max(fa, fb), T(1e-2));
    for (int k = 0; k < 7; k++) c.sol_params[k] = T(0.5) * (m.geom_sol_params[7 * ga + k] + m.geom_sol_params[7 * gb + k]);
    c.link_a = m.geom_link[ga]; c.link_b = m.geom_link[gb];
    c.force = V3<T>();
    e.n_contacts = i + 1;
  } else {
    e.errno_ |= 2u;
  }
}

// func_compute_tolerance (collider_decomp.py:2244-2263)
template <class T>
T compute_tolerance(const Model<T>& m, int ga, int gb, T tol) {
  T sb = norm(m.geom_init_aabb[8 * gb + 7] - m.geom_init_aabb[8 * gb + 0]);
  T s = sb;
  if (m.geom_type[ga] != GEOM_PLANE) s = std::min(norm(m.geom_init_aabb[8 * ga + 7] - m.geom_init_aabb[8 * ga + 0]), sb);
  return T(0.5) * tol * s;
}

// support functions (support_field_decomp.py:201-311)
template <class T>
V3<T> support_box(const Model<T>& m, const V3<T>& gpos, const Q4<T>& gquat, V3<T> d, int g) {
  V3<T> db = inv_transform_by_quat(d, gquat);
  const T* data = &m.geom_data[7 * g];
  V3<T> v((db.x < 0 ? T(-1) : T(1)) * data[0] * T(0.5), (db.y < 0 ? T(-1) : T(1)) * data[1] * T(0.5),
          (db.z < 0 ? T(-1) : T(1)) * data[2] * T(0.5));
  return transform_by_trans_quat(v, gpos, gquat);
}
template <class T>
V3<T> support_sphere(const Model<T>& m, const V3<T>& gpos, V3<T> d, int g) {
  return gpos + d * m.geom_data[7 * g];
}
template <class T>
V3<T> support_capsule(const Model<T>& m, const V3<T>& gpos, const Q4<T>& gquat, V3<T> d, int g) {
  T r = m.geom_data[7 * g], hl = T(0.5) * m.geom_data[7 * g + 1];
  V3<T> axis = transform_by_quat(V3<T>(0, 0, 1), gquat);
  T side = dot(d, axis) < 0 ? T(-1) : T(1);
  V3<T> ep = gpos + (hl * side) * axis;
  return ep + d * r;
}
// Mesh support from the precomputed direction grid (support_field_decomp.py:145-186, _func_support_mesh): the
// mesh-frame direction's spherical angles select a cell; of the four cells at the floor / ceil corners (theta
// wrapped, phi kept off the poles) the candidate vertex with the largest projection wins, the first on ties.
const int SUPPORT_RES = 180;
template <class T>
V3<T> support_mesh_local(const Model<T>& m, V3<T> dm, int g, int& vid) {
  const T pi = T(M_PI);
  const T theta = tatan2(dm.y, dm.x);
  // a unit direction's z can round past +-1; clamped so the pole cells are looked up instead of a NaN angle
  const T phi = tacos(std::min(std::max(dm.z, T(-1)), T(1)));
  const T ii = (theta + pi) / pi / T(2) * T(SUPPORT_RES);
  const T jj = phi / pi * T(SUPPORT_RES);
  T dot_max = T(-1e20);
  V3<T> v(0, 0, 0);
  vid = 0;
  for (int i4 = 0; i4 < 4; i4++) {
    int i = (i4 % 2) ? int(std::ceil(ii)) % SUPPORT_RES : int(std::floor(ii)) % SUPPORT_RES;
    int j;
    if (i4 / 2 > 0) {
      j = std::min(std::max(int(std::ceil(jj)), 0), SUPPORT_RES - 1);
      if (j == SUPPORT_RES - 1) j = SUPPORT_RES - 2;
    } else {
      j = std::min(std::max(int(std::floor(jj)), 0), SUPPORT_RES - 1);
      if (j == 0) j = 1;
    }
    const int idx = m.support_cell_start[g] + i * SUPPORT_RES + j;
    const V3<T> pos = m.support_v[idx];
    const T dp = dot(pos, dm);
    if (dp > dot_max) {
      v = pos;
      dot_max = dp;
      vid = m.support_vid[idx];
    }
  }
  return v;
}
// world-frame mesh support (support_field_decomp.py:125-141, _func_support_world)
template <class T>
V3<T> support_mesh(const Model<T>& m, const V3<T>& gpos, const Q4<T>& gquat, V3<T> d, int g, int& vid) {
  const V3<T> dm = transform_by_quat(d, inv_quat(gquat));
  return transform_by_trans_quat(support_mesh_local(m, dm, g, vid), gpos, gquat);
}

// mpr support_driver (mpr_decomp.py:171-197) over the (possibly perturbed) pose held in the env
template <class T>
V3<T> support_driver(const Model<T>& m, const Env<T>& e, V3<T> d, int g) {
  int t = m.geom_type[g];
  if (t == GEOM_SPHERE) return support_sphere(m, e.g_pos[g], d, g);
  if (t == GEOM_CAPSULE) return support_capsule(m, e.g_pos[g], e.g_quat[g], d, g);
  if (t == GEOM_BOX) return support_box(m, e.g_pos[g], e.g_quat[g], d, g);
  if (t == GEOM_MESH) {
    int vid;
    return support_mesh(m, e.g_pos[g], e.g_quat[g], d, g, vid);
  }
  return e.g_pos[g];
}

template <class T>
struct MPRS {
  V3<T> v1[4], v2[4], v[4];
  int size;
};

template <class T>
void compute_support(const Model<T>& m, const Env<T>& e, V3<T> dir, int ga, int gb, V3<T>& v, V3<T>& v1, V3<T>& v2) {
  v1 = support_driver(m, e, dir, ga);
  v2 = support_driver(m, e, -dir, gb);
  v = v1 - v2;
}

const double CCD_EPS_F = 1e-9, CCD_TOL = 1e-6;
const int CCD_ITERS = 50;
// Portal-refinement loops of the reference carry no bound; this cap only guards pathological inputs.
const int MPR_LOOP_CAP = 1000;

template <class T>
V3<T> mpr_portal_dir(const MPRS<T>& s) {
  V3<T> a = s.v[2] - s.v[1], b = s.v[3] - s.v[1];
  return normalized(cross(a, b));
}
template <class T>
bool mpr_reach_tol(const MPRS<T>& s, V3<T> v, V3<T> dir) {
  T dv1 = dot(s.v[1], dir), dv2 = dot(s.v[2], dir), dv3 = dot(s.v[3], dir), dv4 = dot(v, dir);
  T dot1 = std::min(std::min(dv4 - dv1, dv4 - dv2), dv4 - dv3);
  return dot1 < T(CCD_TOL) + T(CCD_EPS_F) * std::max(T(1), dot1);
}
template <class T>
void mpr_expand(MPRS<T>& s, V3<T> v, V3<T> v1, V3<T> v2) {
  V3<T> v4v0 = cross(v, s.v[0]);
  T d = dot(s.v[1], v4v0);
  int i_s;
  if (d > 0) { d = dot(s.v[2], v4v0); i_s = d > 0 ? 1 : 3; }
  else { d = dot(s.v[3], v4v0); i_s = d > 0 ? 2 : 1; }
  s.v1[i_s] = v1; s.v2[i_s] = v2; s.v[i_s] = v;
}

// mpr_discover_portal (mpr_decomp.py:475-620)
template <class T>
int mpr_discover(const Model<T>& m, const Env<T>& e, MPRS<T>& s, int ga, int gb, V3<T> ca, V3<T> cb) {
  const T E = T(CCD_EPS_F);
  s.v1[0] = ca; s.v2[0] = cb; s.v[0] = ca - cb; s.size = 1;
  if (std::fabs(s.v[0].x) < E && std::fabs(s.v[0].y) < E && std::fabs(s.v[0].z) < E) s.v[0].x += T(10) * E;
  V3<T> dir = -normalized(s.v[0]);
  V3<T> v, v1, v2;
  compute_support(m, e, dir, ga, gb, v, v1, v2);
  s.v1[1] = v1; s.v2[1] = v2; s.v[1] = v; s.size = 2;
  T d = dot(v, dir);
  int ret = 0;
  if (d < E) return -1;
  dir = cross(s.v[0], s.v[1]);
  if (dot(dir, dir) < E) {
    if (std::fabs(s.v[1].x) < E && std::fabs(s.v[1].y) < E && std::fabs(s.v[1].z) < E) return 1;
    return 2;
  }
  dir = normalized(dir);
  compute_support(m, e, dir, ga, gb, v, v1, v2);
  d = dot(v, dir);
  if (d < E) return -1;
  s.v1[2] = v1; s.v2[2] = v2; s.v[2] = v; s.size = 3;
  V3<T> va = s.v[1] - s.v[0], vb = s.v[2] - s.v[0];
  dir = normalized(cross(va, vb));
  d = dot(dir, s.v[0]);
  if (d > 0) {
    std::swap(s.v1[1], s.v1[2]); std::swap(s.v2[1], s.v2[2]); std::swap(s.v[1], s.v[2]);
    dir = -dir;
  }
  int trials = 0;
  while (s.size < 4) {
    compute_support(m, e, dir, ga, gb, v, v1, v2);
    d = dot(v, dir);
    if (d < E) { ret = -1; break; }
    bool cont = false;
    va = cross(s.v[1], v);
    d = dot(va, s.v[0]);
    if (d < -E) { s.v1[2] = v1; s.v2[2] = v2; s.v[2] = v; cont = true; }
    if (!cont) {
      va = cross(v, s.v[2]);
      d = dot(va, s.v[0]);
      if (d < -E) { s.v1[1] = v1; s.v2[1] = v2; s.v[1] = v; cont = true; }
    }
    if (cont) {
      va = s.v[1] - s.v[0]; vb = s.v[2] - s.v[0];
      dir = normalized(cross(va, vb));
      trials++;
      if (trials == 15) { ret = -1; break; }
    } else {
      s.v1[3] = v1; s.v2[3] = v2; s.v[3] = v; s.size = 4;
    }
  }
  return ret;
}

template <class T>
int mpr_refine(const Model<T>& m, const Env<T>& e, MPRS<T>& s, int ga, int gb) {
  for (int it = 0; it < MPR_LOOP_CAP; it++) {
    V3<T> dir = mpr_portal_dir(s);
    if (dot(s.v[1], dir) > -T(CCD_EPS_F)) return 0;
    V3<T> v, v1, v2;
    compute_support(m, e, dir, ga, gb, v, v1, v2);
    if (!(dot(v, dir) > -T(CCD_EPS_F)) || mpr_reach_tol(s, v, dir)) return -1;
    mpr_expand(s, v, v1, v2);
  }
  return -1;
}

template <class T>
V3<T> mpr_find_pos(const MPRS<T>& s) {
  // enable_mujoco_compatibility=False: barycentric weights from the portal direction (mpr_decomp.py:318-342)
  T b[4] = {0, 0, 0, 0};
  T sum = 0;
  V3<T> dir = mpr_portal_dir(s);
  for (int i = 1; i < 4; i++) {
    int i1 = i % 3 + 1, i2 = (i + 1) % 3 + 1;
    b[i] = dot(cross(s.v[i1], s.v[i2]), dir);
  }
  sum = b[0] + b[1] + b[2] + b[3];
  V3<T> p1, p2;
  for (int i = 0; i < 4; i++) { p1 = p1 + b[i] * s.v1[i]; p2 = p2 + b[i] * s.v2[i]; }
  return (T(0.5) / sum) * (p1 + p2);
}

template <class T>
void mpr_find_penetration(const Model<T>& m, const Env<T>& e, MPRS<T>& s, int ga, int gb, bool& is_col, V3<T>& normal, T& pen, V3<T>& pos) {
  int iters = 0;
  for (int guard = 0; guard < MPR_LOOP_CAP; guard++) {
    V3<T> dir = mpr_portal_dir(s);
    V3<T> v, v1, v2;
    compute_support(m, e, dir, ga, gb, v, v1, v2);
    if (mpr_reach_tol(s, v, dir) || iters > CCD_ITERS) {
      pen = dot(dir, s.v[1]);
      normal = -dir;
      is_col = true;
      pos = mpr_find_pos(s);
      return;
    }
    mpr_expand(s, v, v1, v2);
    iters++;
  }
  is_col = false;
}

// guess_geoms_center (mpr_decomp.py:623-705)
template <class T>
void guess_centers(const Model<T>& m, const Env<T>& e, int ga, int gb, V3<T> nws, V3<T>& ca, V3<T>& cb) {
  V3<T> pa = e.g_pos[ga], pb = e.g_pos[gb];
  Q4<T> qa = e.g_quat[ga], qb = e.g_quat[gb];
  ca = transform_by_trans_quat(m.geom_center[ga], pa, qa);
  cb = transform_by_trans_quat(m.geom_center[gb], pb, qb);
  const T E = T(CCD_EPS_F);
  if (std::fabs(nws.x) > E || std::fabs(nws.y) > E || std::fabs(nws.z) > E) {
    V3<T> cal = T(0.5) * (m.geom_init_aabb[8 * ga + 7] + m.geom_init_aabb[8 * ga + 0]);
    ca = transform_by_trans_quat(cal, pa, qa);
    V3<T> cbl = T(0.5) * (m.geom_init_aabb[8 * gb + 7] + m.geom_init_aabb[8 * gb + 0]);
    cb = transform_by_trans_quat(cbl, pb, qb);
    V3<T> delta = ca - cb;
    V3<T> nrm = normalized(delta);
    if (norm(cross(nws, nrm)) > T(0.01)) {
      V3<T> off = dot(delta, nws) * nws - delta;
      T on = norm(off);
      if (on > m.EPS) {
        V3<T> dof = off / on;
        V3<T> dla = inv_transform_by_quat(dof, qa), dlb = inv_transform_by_quat(dof, qb);
        V3<T> bsa = m.geom_init_aabb[8 * ga + 7] - m.geom_init_aabb[8 * ga + 0];
        V3<T> bsb = m.geom_init_aabb[8 * gb + 7] - m.geom_init_aabb[8 * gb + 0];
        T la = dot(bsa, vabs(dla)), lb = dot(bsb, vabs(dlb));
        T ratio = std::min(on / (la + lb), T(0.5));
        ca = ca + dof * la * ratio;
        cb = cb - dof * lb * ratio;
      }
    }
  }
}

// func_mpr_contact (mpr_decomp.py:708-826)
template <class T>
void mpr_contact(const Model<T>& m, const Env<T>& e, int ga, int gb, V3<T> nws, bool& is_col, V3<T>& normal, T& pen, V3<T>& pos) {
  V3<T> ca, cb;
  guess_centers(m, e, ga, gb, nws, ca, cb);
  MPRS<T> s;
  int res = mpr_discover(m, e, s, ga, gb, ca, cb);
  is_col = false; pos = V3<T>(); normal = V3<T>(); pen = 0;
  if (res == 1) {
    is_col = true; pen = 0; normal = -normalized(s.v[0]); pos = (s.v1[1] + s.v2[1]) * T(0.5);
  } else if (res == 2) {
    is_col = true; pen = norm(s.v[1]); normal = -normalized(s.v[1]); pos = (s.v1[1] + s.v2[1]) * T(0.5);
  } else if (res == 0) {
    res = mpr_refine(m, e, s, ga, gb);
    if (res >= 0) mpr_find_penetration(m, e, s, ga, gb, is_col, normal, pen, pos);
  }
}

// func_contact_orthogonals (collider_decomp.py:2266-2325), non-compat branch
template <class T>
void contact_orthogonals(const Model<T>& m, const Env<T>& e, int ga, int gb, V3<T> normal, V3<T>& a0, V3<T>& a1) {
  int g = gb;
  if (m.geom_type[ga] != GEOM_PLANE) {
    V3<T> sa = m.geom_init_aabb[8 * ga + 7], sb = m.geom_init_aabb[8 * gb + 7];
    T va = sa.x * sa.y * sa.z, vb = sb.x * sb.y * sb.z;
    g = va < vb ? ga : gb;
  }
  int l = m.geom_link[g];
  M3<T> R = quat_to_R(e.l_iquat[l], m.EPS);
  int idx = 0;
  T amax = 0;
  for (int i = 0; i < 3; i++) {
    T a = std::fabs(dot(R.col(i), normal));
    if (a > amax) { amax = a; idx = i; }
  }
  idx = (idx + 1) % 3;
  a0 = R.col(idx);
  a0 = normalized(a0 - dot(normal, a0) * normal);
  a1 = cross(normal, a0);
}

template <class T>
void rotate_frame(Env<T>& e, int g, V3<T> cpos, Q4<T> qrot) {
  e.g_quat[g] = transform_quat_by_quat(e.g_quat[g], qrot);
  V3<T> rel = cpos - e.g_pos[g];
  V3<T> vec = transform_by_quat(rel, qrot) - rel;
  e.g_pos[g] = e.g_pos[g] - vec;
}

// func_convex_convex_contact (collider_decomp.py:2328-2725) -- MPR path; the GJK/EPA fallback is not restated
// ------------------------------------------------------------------------------------------------------------
// GJK + EPA, the "safe" (non-MuJoCo-compatible) path taken when MPR reports a deep penetration
// (gjk_decomp.py: func_gjk_contact :165-417, func_safe_gjk :3656-3856, func_safe_gjk_support :4167-4280,
//  func_search_valid_simplex_vertex :3982-4076, func_safe_epa :4338-4540, func_safe_epa_witness :4543-4600,
//  func_safe_epa_init :4604-4653, func_safe_attach_face_to_polytope :4657-4739, func_plane_normal :4742-4780,
//  func_epa_horizon :1317-1383, func_triangle_affine_coords :2037-2094, func_project_origin_to_plane :3414-3464)
// ------------------------------------------------------------------------------------------------------------
constexpr int kGjkIters = 50, kEpaIters = 50, kPolyFaces = 6 * kEpaIters, kPolyVerts = 5 + kEpaIters;
constexpr int kHorizon = 6 + kEpaIters, kHStack = 3 * kPolyFaces;

template <class T>
struct GjkConst {  // gjk_decomp.py:87-113
  T FLOAT_MIN = T(1e-15), FLOAT_MAX = T(1e15), FLOAT_MIN_SQ = T(1e-30), FLOAT_MAX_SQ = T(1e30);
  T tolerance = T(1e-6), degen_sq = T(1e-10), max_reproj = T(1e-5);
};

template <class T>
struct GVert {
  V3<T> o1, o2, mk;
  int id1 = -1, id2 = -1;
};

template <class T>
struct GFace {
  int v[3] = {0, 0, 0}, adj[3] = {0, 0, 0};
  V3<T> n;
  T dist2 = 0;
  int map_idx = -1;
};

template <class T>
struct GjkWork {
  GVert<T> s[4];
  int ns = 0, last_searched = 0;
  GVert<T> pv[kPolyVerts];
  int npv = 0;
  GFace<T> f[kPolyFaces];
  int nf = 0;
  int fmap[kPolyFaces];
  int nfmap = 0;
  int hf[kHorizon], he[kHorizon];
  int nh = 0;
  int sf[kHStack], se[kHStack];
  V3<T> w1, w2;
  int n_witness = 0;
  T distance = 0;
};

inline bool gjk_discrete(int t) { return t == GEOM_BOX || t == GEOM_MESH; }

// support_driver (gjk_decomp.py:3605-3652) with the box vertex id of support_field_decomp.py:289-311
template <class T>
void gjk_support_one(const Model<T>& m, const Env<T>& e, V3<T> d, int g, V3<T>& v, int& vid) {
  int t = m.geom_type[g];
  vid = -1;
  if (t == GEOM_BOX) {
    V3<T> db = inv_transform_by_quat(d, e.g_quat[g]);
    const T* data = &m.geom_data[7 * g];
    V3<T> vl((db.x < 0 ? T(-1) : T(1)) * data[0] * T(0.5), (db.y < 0 ? T(-1) : T(1)) * data[1] * T(0.5),
             (db.z < 0 ? T(-1) : T(1)) * data[2] * T(0.5));
    vid = (vl.x > 0 ? 1 : 0) + (vl.y > 0 ? 2 : 0) + (vl.z > 0 ? 4 : 0) + m.geom_vert_start[g];
    v = transform_by_trans_quat(vl, e.g_pos[g], e.g_quat[g]);
  } else if (t == GEOM_MESH) {
    // the support field, not an exhaustive vertex search (that one is the MuJoCo-compatibility path)
    v = support_mesh(m, e.g_pos[g], e.g_quat[g], d, g, vid);
  } else {
    v = support_driver(m, e, d, g);
  }
}

template <class T>
GVert<T> gjk_support(const Model<T>& m, const Env<T>& e, int ga, int gb, V3<T> d) {  // func_support
  GVert<T> r;
  gjk_support_one(m, e, d, ga, r.o1, r.id1);
  gjk_support_one(m, e, -d, gb, r.o2, r.id2);
  r.mk = r.o1 - r.o2;
  return r;
}

template <class T>
int gjk_count_supports(const Model<T>& m, const Env<T>& e, V3<T> d, int g) {  // count_support_driver
  if (m.geom_type[g] != GEOM_BOX) return 1;
  V3<T> db = inv_transform_by_quat(d, e.g_quat[g]);
  int z = (db.x == 0) + (db.y == 0) + (db.z == 0);
  return 1 << z;
}

template <class T>
bool gjk_duplicate(const GjkWork<T>& w, int id1, int id2) {
  for (int i = 0; i < w.ns; i++) {
    if (id1 == -1 || w.s[i].id1 != id1) continue;
    if (id2 == -1 || w.s[i].id2 != id2) continue;
    return true;
  }
  return false;
}

template <class T>
bool gjk_degenerate(const GjkWork<T>& w, const GjkConst<T>& c, V3<T> mk) {
  for (int i = 0; i < w.ns; i++) {
    V3<T> d = w.s[i].mk - mk;
    if (dot(d, d) < c.degen_sq) return true;
  }
  if (w.ns == 2) {  // func_is_colinear
    V3<T> e1 = w.s[1].mk - w.s[0].mk, e2 = mk - w.s[0].mk;
    V3<T> n = cross(e1, e2);
    return dot(n, n) < c.degen_sq * dot(e1, e1) * dot(e2, e2);
  }
  if (w.ns == 3) {  // func_is_coplanar
    V3<T> e1 = normalized(w.s[1].mk - w.s[0].mk), e2 = normalized(w.s[2].mk - w.s[0].mk);
    V3<T> n = cross(e1, e2);
    V3<T> diff = mk - w.s[0].mk;
    T nd = dot(n, diff);
    return nd * nd < c.degen_sq * dot(n, n) * dot(diff, diff);
  }
  return false;
}

template <class T>
bool gjk_valid(const GjkWork<T>& w, const GjkConst<T>& c, const GVert<T>& v) {
  return !gjk_duplicate(w, v.id1, v.id2) && !gjk_degenerate(w, c, v.mk);
}

template <class T>
GVert<T> gjk_safe_support(const Model<T>& m, const Env<T>& e, const GjkWork<T>& w, const GjkConst<T>& c, int ga,
                          int gb, V3<T> dir) {
  const T EPS = m.EPS;
  GVert<T> r;
  for (int i = 0; i < 9; i++) {
    V3<T> nd = dir;
    if (i > 0) {
      int j = i - 1;
      // the reference scales by 1 - 2 * (j & bit) with the bit's value, not a 0/1 flag
      nd.x += -T(1 - 2 * (j & 1)) * EPS;
      nd.y += -T(1 - 2 * (j & 2)) * EPS;
      nd.z += -T(1 - 2 * (j & 4)) * EPS;
    }
    nd = nd * (T(2) - dot(nd, dir));
    int num = gjk_count_supports(m, e, nd, ga) * gjk_count_supports(m, e, -nd, gb);
    if (i > 0 && num > 1) continue;
    r = gjk_support(m, e, ga, gb, nd);
    if (i == 0) {
      if (num > 1) continue;
      break;
    }
    if (i == 8) break;
    if (gjk_valid(w, c, r)) break;
  }
  return r;
}

template <class T>
V3<T> gjk_discrete_vertex(const Model<T>& m, const Env<T>& e, int g, int iv) {
  V3<T> vl;
  if (m.geom_type[g] == GEOM_BOX) {
    const T* data = &m.geom_data[7 * g];
    vl = V3<T>(((iv & 1) ? T(1) : T(-1)) * data[0] * T(0.5), ((iv & 2) ? T(1) : T(-1)) * data[1] * T(0.5),
               ((iv & 4) ? T(1) : T(-1)) * data[2] * T(0.5));
  } else {
    vl = m.verts[m.geom_vert_start[g] + iv];
  }
  return transform_by_trans_quat(vl, e.g_pos[g], e.g_quat[g]);
}

template <class T>
bool gjk_search_valid(const Model<T>& m, const Env<T>& e, GjkWork<T>& w, const GjkConst<T>& c, int ga, int gb,
                      GVert<T>& out) {
  if (gjk_discrete(m.geom_type[ga]) && gjk_discrete(m.geom_type[gb])) {
    int n0 = m.geom_vert_end[ga] - m.geom_vert_start[ga], n1 = m.geom_vert_end[gb] - m.geom_vert_start[gb];
    int cases = n0 * n1;
    for (int k = 0; k < cases; k++) {
      int mm = (k + w.last_searched) % cases;
      int i = mm / n1, j = mm % n1;
      out.id1 = m.geom_vert_start[ga] + i;
      out.id2 = m.geom_vert_start[gb] + j;
      out.o1 = gjk_discrete_vertex(m, e, ga, i);
      out.o2 = gjk_discrete_vertex(m, e, gb, j);
      out.mk = out.o1 - out.o2;
      if (gjk_valid(w, c, out)) {
        w.last_searched = (mm + 1) % cases;
        return true;
      }
    }
    return false;
  }
  if (w.ns == 3) {
    V3<T> dir = normalized(cross(w.s[2].mk - w.s[0].mk, w.s[1].mk - w.s[0].mk));
    for (int i = 0; i < 2; i++) {
      out = gjk_safe_support(m, e, w, c, ga, gb, i == 0 ? dir : -dir);
      if (gjk_valid(w, c, out)) return true;
    }
  }
  return false;
}

// returns 1 when the origin is inside the simplex (intersection), 0 otherwise
template <class T>
int gjk_safe(const Model<T>& m, const Env<T>& e, GjkWork<T>& w, const GjkConst<T>& c, int ga, int gb) {
  bool init_ok = true;
  w.ns = 0;
  for (int i = 0; i < 4; i++) {
    V3<T> dir;
    T val = T(1 - 2 * (i % 2));
    if (2 - i / 2 == 2) dir.z = val;
    else dir.y = val;
    GVert<T> v = gjk_safe_support(m, e, w, c, ga, gb, dir);
    if (!gjk_valid(w, c, v)) {
      if (!gjk_search_valid(m, e, w, c, ga, gb, v)) {
        init_ok = false;
        break;
      }
    }
    w.s[i] = v;
    w.ns++;
  }
  int flag = 0;
  if (init_ok) {
    for (int it = 0; it < kGjkIters; it++) {
      V3<T> nrm[4];
      T sd[4];
      for (int j = 0; j < 4; j++) {
        int s0 = 2, s1 = 1, s2 = 3, ap = 0;
        if (j == 1) { s0 = 0; s1 = 2; s2 = 3; ap = 1; }
        else if (j == 2) { s0 = 1; s1 = 0; s2 = 3; ap = 2; }
        else if (j == 3) { s0 = 0; s1 = 1; s2 = 2; ap = 3; }
        V3<T> v1 = w.s[s0].mk, v2 = w.s[s1].mk, v3 = w.s[s2].mk, va = w.s[ap].mk;
        V3<T> n = normalized(cross(v3 - v1, v2 - v1));
        if (dot(n, va - v1) > 0) n = -n;
        nrm[j] = n;
        sd[j] = dot(n, v1);
      }
      int mi = 0;
      for (int j = 1; j < 4; j++)
        if (sd[j] < sd[mi]) mi = j;
      if (sd[mi] >= 0) { flag = 1; break; }
      w.ns = 3;
      if (mi != 3) w.s[mi] = w.s[3];
      GVert<T> v = gjk_safe_support(m, e, w, c, ga, gb, nrm[mi]);
      if (gjk_duplicate(w, v.id1, v.id2)) break;
      if (gjk_degenerate(w, c, v.mk)) break;
      if (dot(v.mk, nrm[mi]) < 0) break;
      w.s[3] = v;
      w.ns = 4;
    }
  }
  w.distance = flag ? T(0) : c.FLOAT_MAX;
  return flag;
}

template <class T>
bool plane_normal(const GjkConst<T>& c, V3<T> v1, V3<T> v2, V3<T> v3, V3<T>& normal) {
  V3<T> d21 = v2 - v1, d31 = v3 - v1, d32 = v3 - v2;
  for (int i = 0; i < 3; i++) {
    V3<T> n = i == 0 ? cross(d32, d21) : (i == 1 ? cross(d21, d31) : cross(d31, d32));
    T nn = norm(n);
    if (nn == 0) return false;
    if (nn > c.FLOAT_MIN) {
      normal = normalized(n);
      return true;
    }
  }
  return false;
}

template <class T>
bool epa_attach(GjkWork<T>& w, const GjkConst<T>& c, int v1, int v2, int v3, int a1, int a2, int a3) {
  int n = w.nf;
  GFace<T>& F = w.f[n];
  F.v[0] = v1; F.v[1] = v2; F.v[2] = v3;
  F.adj[0] = a1; F.adj[1] = a2; F.adj[2] = a3;
  w.nf++;
  V3<T> normal;
  bool ok = plane_normal(c, w.pv[v3].mk, w.pv[v2].mk, w.pv[v1].mk, normal);
  if (ok) {
    V3<T> center = (w.pv[v1].mk + w.pv[v2].mk + w.pv[v3].mk) / T(3);
    T max_orient = -dot(normal, center), max_abs = std::fabs(max_orient);
    for (int iv = 0; iv < w.npv; iv++) {
      if (iv == v1 || iv == v2 || iv == v3) continue;
      T orient = dot(normal, w.pv[iv].mk - center);
      if (std::fabs(orient) > max_abs) {
        max_abs = std::fabs(orient);
        max_orient = orient;
      }
    }
    if (max_orient > 0) normal = -normal;
    F.n = normal;
    T md = c.FLOAT_MAX;
    for (int i = 0; i < 3; i++) {
      T d = dot(normal, w.pv[F.v[i]].mk);
      T d2 = d * d;
      if (d2 < md) md = d2;
    }
    F.dist2 = md;
    F.map_idx = -1;
  } else {
    // degenerate face: never the nearest (the reference leaves the slot's previous contents)
    F.n = V3<T>();
    F.dist2 = c.FLOAT_MAX_SQ;
  }
  return ok;
}

template <class T>
void epa_delete_face(GjkWork<T>& w, int i_f) {
  int idx = w.f[i_f].map_idx;
  if (idx >= 0) {
    int last = w.fmap[w.nfmap - 1];
    w.fmap[idx] = last;
    w.f[last].map_idx = idx;
    w.nfmap--;
  }
  w.f[i_f].map_idx = -2;
}

// returns true on error (horizon not a closed loop, or a buffer full)
template <class T>
bool epa_horizon(GjkWork<T>& w, const GjkConst<T>& c, int nearest, V3<T> wv) {
  w.sf[0] = nearest;
  w.se[0] = 0;
  int top = 1;
  bool first = true;
  while (top > 0) {
    int i_f = w.sf[top - 1], i_e = w.se[top - 1];
    V3<T> v = w.pv[w.f[i_f].v[0]].mk;
    top--;
    bool deleted = w.f[i_f].map_idx == -2;
    if (!first && deleted) continue;
    bool visible = dot(w.f[i_f].n, wv - v) > c.FLOAT_MIN;
    if (visible || first) {
      epa_delete_face(w, i_f);
      for (int k = first ? 0 : 1; k < 3; k++) {
        int e2 = (i_e + k) % 3;
        int adj = w.f[i_f].adj[e2];
        if (w.f[adj].map_idx != -2) {
          int sv = w.f[i_f].v[(e2 + 1) % 3];
          const int* vv = w.f[adj].v;
          int ae = vv[0] == sv ? 0 : (vv[1] == sv ? 1 : 2);
          if (top >= kHStack) return true;
          w.sf[top] = adj;
          w.se[top] = ae;
          top++;
        }
      }
    } else {
      if (w.nh >= kHorizon) return true;
      w.hf[w.nh] = i_f;
      w.he[w.nh] = i_e;
      w.nh++;
    }
    first = false;
  }
  return false;
}

template <class T>
void project_origin_to_plane(const GjkConst<T>& c, V3<T> v1, V3<T> v2, V3<T> v3, V3<T>& point) {
  V3<T> d21 = v2 - v1, d31 = v3 - v1, d32 = v3 - v2;
  for (int i = 0; i < 3; i++) {
    V3<T> n = i == 0 ? cross(d32, d21) : (i == 1 ? cross(d21, d31) : cross(d31, d32));
    V3<T> v = i == 0 ? v2 : (i == 1 ? v1 : v3);
    T nv = dot(n, v), nn = dot(n, n);
    if (nn == 0) return;
    if (nn > c.FLOAT_MIN) {
      point = n * (nv / nn);
      return;
    }
    if (i == 2) {
      if (!(nn < c.FLOAT_MIN)) point = n * (nv / nn);
    }
  }
}

template <class T>
V3<T> triangle_affine_coords(V3<T> p, V3<T> t1, V3<T> t2, V3<T> t3) {
  T ms[3];
  for (int i = 0; i < 3; i++) {
    int i1 = (i + 1) % 3, i2 = (i + 2) % 3;
    if (i == 1) std::swap(i1, i2);
    ms[i] = t2[i1] * t3[i2] - t2[i2] * t3[i1] - t1[i1] * t3[i2] + t1[i2] * t3[i1] + t1[i1] * t2[i2] - t1[i2] * t2[i1];
  }
  T m_max = 0;
  int ix = 0, iy = 0;
  for (int i = 0; i < 3; i++) {
    T a = std::fabs(ms[i]);
    if (a >= std::fabs(ms[(i + 1) % 3]) && a >= std::fabs(ms[(i + 2) % 3])) {
      m_max = ms[i];
      ix = (i + 1) % 3;
      iy = (i + 2) % 3;
      if (i == 1) std::swap(ix, iy);
      break;
    }
  }
  T cs[3];
  for (int i = 0; i < 3; i++) {
    V3<T> tv1 = i == 0 ? t2 : (i == 1 ? t3 : t1), tv2 = i == 0 ? t3 : (i == 1 ? t1 : t2);
    cs[i] = p[ix] * tv1[iy] + p[iy] * tv2[ix] + tv1[ix] * tv2[iy] - p[ix] * tv2[iy] - p[iy] * tv1[ix] - tv2[ix] * tv1[iy];
  }
  return V3<T>(cs[0] / m_max, cs[1] / m_max, cs[2] / m_max);
}

template <class T>
bool epa_witness(GjkWork<T>& w, const GjkConst<T>& c, int i_f) {
  const int* iv = w.f[i_f].v;
  V3<T> v1 = w.pv[iv[0]].mk, v2 = w.pv[iv[1]].mk, v3 = w.pv[iv[2]].mk;
  V3<T> proj;
  project_origin_to_plane(c, v1, v2, v3, proj);
  V3<T> l = triangle_affine_coords(proj, v1, v2, v3);
  V3<T> pl = v1 * l.x + v2 * l.y + v3 * l.z;
  T err = norm(proj - pl);
  T me = std::max(std::max(dot(v1 - v2, v1 - v2), dot(v2 - v3, v2 - v3)), std::max(dot(v3 - v1, v3 - v1), c.FLOAT_MIN_SQ));
  T inv = T(1) / std::sqrt(me);
  if (err * inv > c.max_reproj) return false;
  w.w1 = w.pv[iv[0]].o1 * l.x + w.pv[iv[1]].o1 * l.y + w.pv[iv[2]].o1 * l.z;
  w.w2 = w.pv[iv[0]].o2 * l.x + w.pv[iv[1]].o2 * l.y + w.pv[iv[2]].o2 * l.z;
  return true;
}

template <class T>
void epa_safe(const Model<T>& m, const Env<T>& e, GjkWork<T>& w, const GjkConst<T>& c, int ga, int gb) {
  const T EPS = m.EPS;
  // func_safe_epa_init: the tetrahedron's 4 faces
  w.npv = 0; w.nf = 0; w.nfmap = 0; w.nh = 0;
  int vi[4];
  for (int i = 0; i < 4; i++) { w.pv[w.npv] = w.s[i]; vi[i] = w.npv++; }
  const int fv[4][3] = {{vi[0], vi[1], vi[2]}, {vi[0], vi[3], vi[1]}, {vi[0], vi[2], vi[3]}, {vi[3], vi[2], vi[1]}};
  const int fa[4][3] = {{1, 3, 2}, {2, 3, 0}, {0, 3, 1}, {2, 0, 1}};
  for (int i = 0; i < 4; i++) epa_attach(w, c, fv[i][0], fv[i][1], fv[i][2], fa[i][0], fa[i][1], fa[i][2]);
  for (int i = 0; i < 4; i++) { w.fmap[i] = i; w.f[i].map_idx = i; }
  w.nfmap = 4;
  // func_safe_epa
  T upper = c.FLOAT_MAX, upper2 = c.FLOAT_MAX_SQ, lower = 0, tol = c.tolerance;
  int nearest = -1, prev = -1;
  bool discrete = gjk_discrete(m.geom_type[ga]) && gjk_discrete(m.geom_type[gb]);
  if (discrete) tol = EPS;
  for (int k = 0; k < kEpaIters; k++) {
    prev = nearest;
    T lower2 = c.FLOAT_MAX_SQ;
    for (int i = 0; i < w.nfmap; i++) {
      int i_f = w.fmap[i];
      if (w.f[i_f].dist2 < lower2) { lower2 = w.f[i_f].dist2; nearest = i_f; }
    }
    if (lower2 > upper2 || nearest == -1) { nearest = prev; break; }
    lower = std::sqrt(lower2);
    V3<T> dir = w.f[nearest].n;
    // func_epa_support with dir_norm = 1 (dir / 1)
    int wi = w.npv;
    w.pv[w.npv++] = gjk_support(m, e, ga, gb, dir / T(1));
    V3<T> wv = w.pv[wi].mk;
    T upper_k = dot(wv, dir);
    if (upper_k < upper) { upper = upper_k; upper2 = upper * upper; }
    if (upper - lower < tol) break;
    if (discrete) {
      bool repeated = false;
      for (int i = 0; i < w.npv; i++) {
        if (i == wi) continue;
        if (w.pv[i].id1 == w.pv[wi].id1 && w.pv[i].id2 == w.pv[wi].id2) { repeated = true; break; }
      }
      if (repeated) break;
    }
    if (epa_horizon(w, c, nearest, wv)) { nearest = -1; break; }
    if (w.nh < 3) { nearest = -1; break; }
    int nfaces = w.nf, nedges = w.nh;
    if (nfaces + nedges >= kPolyFaces) break;
    bool ok = true;
    for (int i = 0; i < nedges; i++) {
      int f0 = nfaces + i, f1 = nfaces + (i + 1) % nedges;
      int hf = w.hf[i], he = w.he[i];
      int hv1 = w.f[hf].v[he], hv2 = w.f[hf].v[(he + 1) % 3];
      w.f[hf].adj[he] = f0;
      int adj0 = i > 0 ? f0 - 1 : nfaces + nedges - 1;
      ok = epa_attach(w, c, wi, hv2, hv1, f1, hf, adj0);
      if (!ok) break;
      T d2 = w.f[w.nf - 1].dist2;
      if (d2 >= lower2 - EPS && d2 <= upper2 + EPS) {
        w.fmap[w.nfmap] = f0;
        w.f[f0].map_idx = w.nfmap;
        w.nfmap++;
      }
    }
    if (!ok) break;
    w.nh = 0;
    if (w.nfmap == 0 || nearest == -1) break;
  }
  if (nearest != -1) {
    T d2 = w.f[nearest].dist2;
    if (epa_witness(w, c, nearest)) {
      w.n_witness = 1;
      w.distance = -std::sqrt(d2);
    } else {
      w.n_witness = 0;
      w.distance = 0;
    }
  } else {
    w.n_witness = 0;
    w.distance = 0;
  }
}

// func_gjk_contact, safe path: returns is_col with penetration, normal and contact position
template <class T>
bool gjk_contact(const Model<T>& m, const Env<T>& e, int ga, int gb, T& pen, V3<T>& normal, V3<T>& pos) {
  static thread_local GjkWork<T> w;
  GjkConst<T> c;
  w.last_searched = 0;
  w.n_witness = 0;
  if (gjk_safe(m, e, w, c, ga, gb)) epa_safe(m, e, w, c, ga, gb);
  bool is_col = w.distance < 0;
  pen = is_col ? -w.distance : T(0);
  int n = 0;
  if (is_col) {
    for (int i = 0; i < w.n_witness; i++) {
      V3<T> p = T(0.5) * (w.w1 + w.w2);
      V3<T> nn = w.w2 - w.w1;
      T len = norm(nn);
      if (len < c.FLOAT_MIN) continue;
      normal = nn / len;
      pos = p;
      n++;
    }
  }
  if (n == 0) {
    is_col = false;
    pen = 0;
  }
  return is_col;
}

template <class T>
void convex_convex_contact(const Model<T>& m, Env<T>& e, int ga, int gb) {
  const T EPS = m.EPS;
  int ta = m.geom_type[ga], tb = m.geom_type[gb];
  bool multi = m.enable_multi_contact && ta != GEOM_SPHERE && ta != GEOM_ELLIPSOID && tb != GEOM_SPHERE && tb != GEOM_ELLIPSOID;
  T tol = compute_tolerance(m, ga, gb, m.mc_tolerance);
  V3<T> ga_pos = e.g_pos[ga], gb_pos = e.g_pos[gb];
  Q4<T> ga_quat = e.g_quat[ga], gb_quat = e.g_quat[gb];
  bool is_col0 = false, is_col = false;
  T pen0 = 0, pen = 0;
  V3<T> n0, cp0, normal, cpos, ax0, ax1;
  Q4<T> qrot(0, 0, 0, 0);
  int n_con = 0;
  int i_pair = m.pair_idx[(size_t)std::min(ga, gb) * m.G + std::max(ga, gb)];
  for (int det = 0; det < 5; det++) {
    if (multi && is_col0) {
      V3<T> axis = T(2 * (det % 2) - 1) * ax0 + T(1 - 2 * ((det / 2) % 2)) * ax1;
      qrot = rotvec_to_quat(m.mc_perturbation * axis, EPS);
      rotate_frame(e, ga, cp0, qrot);
      rotate_frame(e, gb, cp0, inv_quat(qrot));
    }
    if ((multi && is_col0) || det == 0) {
      if (ta == GEOM_PLANE) {
        V3<T> pd(m.geom_data[7 * ga], m.geom_data[7 * ga + 1], m.geom_data[7 * ga + 2]);
        pd = transform_by_quat(pd, e.g_quat[ga]);
        normal = -normalized(pd);
        V3<T> v1 = support_driver(m, e, normal, gb);
        pen = dot(normal, v1 - e.g_pos[ga]);
        cpos = v1 - T(0.5) * pen * normal;
        is_col = pen > 0;
      } else {
        bool updated = false;
        V3<T> nws = e.cache_normal[i_pair];
        bool guess = std::fabs(nws.x) > EPS || std::fabs(nws.y) > EPS || std::fabs(nws.z) > EPS;
        for (int im = 0; im < 2; im++) {
          if (im == 1) {
            if (det == 0 && !is_col && guess) { nws = V3<T>(); guess = false; updated = false; }
          }
          if (!updated) { mpr_contact(m, e, ga, gb, nws, is_col, normal, pen, cpos); updated = true; }
        }
        // GJK/EPA fallback (collider_decomp.py:2498-2503, 2513-2605) when MPR reports a deep penetration and
        // either no cached normal was available or the overlap is large
        if (pen > tol) {
          bool prefer_gjk = !guess || (m.mc_tolerance * pen >= m.mpr_to_gjk * tol);
          if (prefer_gjk) {
            e.errno_ |= 0x100u;  // informational: the GJK/EPA fallback ran
            T gpen;
            V3<T> gn, gp;
            is_col = gjk_contact(m, e, ga, gb, gpen, gn, gp);
            pen = gpen;
            if (is_col) { cpos = gp; normal = gn; }
          }
        }
      }
    }
    if (det == 0) {
      is_col0 = is_col; n0 = normal; pen0 = pen; cp0 = cpos;
      if (is_col0) {
        add_contact(m, e, ga, gb, n0, cp0, pen0);
        if (multi) { contact_orthogonals(m, e, ga, gb, normal, ax0, ax1); n_con = 1; }
        e.cache_normal[i_pair] = normal;
      } else {
        e.cache_normal[i_pair] = V3<T>();
      }
    } else if (multi && is_col0 && is_col) {
      V3<T> cpa = transform_by_quat((cpos - T(0.5) * pen * normal) - cp0, inv_quat(qrot)) + cp0;
      V3<T> cpb = transform_by_quat((cpos + T(0.5) * pen * normal) - cp0, qrot) + cp0;
      cpos = T(0.5) * (cpa + cpb);
      V3<T> tw = cross(normal, n0);
      tw = V3<T>(std::min(std::max(tw.x, -m.mc_perturbation), m.mc_perturbation),
                 std::min(std::max(tw.y, -m.mc_perturbation), m.mc_perturbation),
                 std::min(std::max(tw.z, -m.mc_perturbation), m.mc_perturbation));
      normal = normal + cross(tw, normal);
      pen = dot(normal, cpb - cpa);
      bool repeated = false;
      for (int ic = 0; ic < n_con; ic++) {
        if (!repeated) {
          int idx = e.n_contacts - 1 - ic;
          if (norm(cpos - e.contacts[idx].pos) < tol) repeated = true;
        }
      }
      if (!repeated && pen > -tol) {
        pen = std::max(pen, T(0));
        add_contact(m, e, ga, gb, normal, cpos, pen);
        n_con++;
      }
    }
    if (multi && is_col0) {
      e.g_pos[ga] = ga_pos; e.g_quat[ga] = ga_quat; e.g_pos[gb] = gb_pos; e.g_quat[gb] = gb_quat;
    }
  }
}

// func_plane_box_contact (collider_decomp.py:2073-2142)
template <class T>
void plane_box_contact(const Model<T>& m, Env<T>& e, int ga, int gb) {
  V3<T> ga_pos = e.g_pos[ga];
  Q4<T> ga_quat = e.g_quat[ga];
  V3<T> gb_pos = e.g_pos[gb];
  Q4<T> gb_quat = e.g_quat[gb];
  V3<T> pd(m.geom_data[7 * ga], m.geom_data[7 * ga + 1], m.geom_data[7 * ga + 2]);
  pd = transform_by_quat(pd, ga_quat);
  V3<T> normal = -normalized(pd);
  V3<T> v1 = support_box(m, gb_pos, gb_quat, normal, gb);
  T pen = dot(normal, v1 - ga_pos);
  if (pen > 0) {
    V3<T> cpos = v1 - T(0.5) * pen * normal;
    add_contact(m, e, ga, gb, normal, cpos, pen);
    if (m.enable_multi_contact) {
      int n_con = 1;
      V3<T> cp0 = cpos;
      T tol = compute_tolerance(m, ga, gb, m.mc_tolerance);
      for (int iv = m.geom_vert_start[gb]; iv < m.geom_vert_end[gb]; iv++) {
        if (n_con < m.n_contacts_per_pair) {
          V3<T> pc = transform_by_trans_quat(m.verts[iv], gb_pos, gb_quat);
          pen = dot(normal, pc - ga_pos);
          if (pen > 0) {
            cpos = pc - T(0.5) * pen * normal;
            if (norm(cpos - cp0) > tol) { add_contact(m, e, ga, gb, normal, cpos, pen); n_con++; }
          }
        }
      }
    }
  }
}

// Collider.detection narrowphase order on Linux: convex pass (skipping plane-box), then specialisations
template <class T>
void narrow_phase(const Model<T>& m, Env<T>& e) {
  for (int ip = 0; ip < e.n_broad; ip++) {
    int ga = e.broad_a[ip], gb = e.broad_b[ip];
    if (m.geom_type[ga] > m.geom_type[gb]) std::swap(ga, gb);
    if (!(m.geom_type[ga] == GEOM_PLANE && m.geom_type[gb] == GEOM_BOX)) convex_convex_contact(m, e, ga, gb);
  }
  for (int ip = 0; ip < e.n_broad; ip++) {
    int ga = e.broad_a[ip], gb = e.broad_b[ip];
    if (m.geom_type[ga] > m.geom_type[gb]) std::swap(ga, gb);
    if (m.geom_type[ga] == GEOM_PLANE && m.geom_type[gb] == GEOM_BOX) plane_box_contact(m, e, ga, gb);
  }
}

// ------------------------------------------------------------------------------------------------------------
// Constraints (constraint_solver_decomp.py)
// ------------------------------------------------------------------------------------------------------------
// Row n: component i3 of the velocity difference of the point p1 fixed to link l1 (sign +1) and p2 fixed to l2
// (sign -1), both root chains walked from the link up, each link's dofs last first (constraint_solver_decomp.py
// 586-616); returns J qvel in that order.
template <class T>
T point_diff_row(const Model<T>& m, Env<T>& e, int n, int l1, V3<T> p1, int l2, V3<T> p2, int i3) {
  const int D = m.D;
  for (int d = 0; d < D; d++) e.jac[(size_t)n * D + d] = 0;
  T jac_qvel = 0;
  for (int iab = 0; iab < 2; iab++) {
    const T sign = iab == 0 ? T(1) : T(-1);
    int link = iab == 0 ? l1 : l2;
    const V3<T> pos = iab == 0 ? p1 : p2;
    while (link > -1) {
      const int nd = m.link_dof_end[link] - m.link_dof_start[link];
      for (int k = 0; k < nd; k++) {
        const int d = m.link_dof_end[link] - 1 - k;
        const V3<T> tpos = pos - e.root_com[link];
        const V3<T> vel = transform_by_quat(e.cdof_vel[d] - cross(tpos, e.cdof_ang[d]), Q4<T>(1, 0, 0, 0));
        const V3<T> diff = sign * vel;
        const T jac = diff[i3];
        jac_qvel = jac_qvel + jac * e.vel[d];
        e.jac[(size_t)n * D + d] = e.jac[(size_t)n * D + d] + jac;
      }
      link = m.link_parent[link];
    }
  }
  return jac_qvel;
}

template <class T>
void eq_row_finish(const Model<T>& m, Env<T>& e, int n, const T* sp, T pen, T jac_qvel, T pos, T invw) {
  T imp, aref;
  imp_aref(sp, -pen, jac_qvel, pos, imp, aref);
  const T diag = std::max(invw * (T(1) - imp) / imp, m.EPS);
  e.diag[n] = diag; e.aref[n] = aref; e.efc_D[n] = T(1) / diag;
}

// func_equality_connect (constraint_solver_decomp.py:528-641): data[0:3] on link l1 and data[3:6] on l2 coincide
template <class T>
void add_connect(const Model<T>& m, Env<T>& e, int l1, int l2, const T* data, const T* sp) {
  const V3<T> ga1 = transform_by_trans_quat(V3<T>(data[0], data[1], data[2]), e.l_pos[l1], e.l_quat[l1]);
  const V3<T> ga2 = transform_by_trans_quat(V3<T>(data[3], data[4], data[5]), e.l_pos[l2], e.l_quat[l2]);
  const T invw = m.link_invweight[2 * l1] + m.link_invweight[2 * l2];
  for (int i3 = 0; i3 < 3; i3++) {
    const int n = e.n_con++;
    e.n_con_eq++;
    const T jq = point_diff_row(m, e, n, l1, ga1, l2, ga2, i3);
    const V3<T> pd = ga1 - ga2;
    eq_row_finish(m, e, n, sp, norm(pd), jq, pd[i3], invw);
  }
}

// ti_quat_mul_axis (genesis/utils/geom.py:195-204): q * (0, v)
template <class T>
Q4<T> quat_mul_axis(Q4<T> q, V3<T> v) {
  return Q4<T>(-q.x * v[0] - q.y * v[1] - q.z * v[2], +q.w * v[0] + q.y * v[2] - q.z * v[1],
               +q.w * v[1] + q.z * v[0] - q.x * v[2], +q.w * v[2] + q.x * v[1] - q.y * v[0]);
}

// func_equality_weld (constraint_solver_decomp.py:824-1015): data[0:3] anchor on l2, data[3:6] anchor on l1,
// data[6:10] pose of l2 relative to l1, data[10] torque scale; 3 position rows, then 3 orientation rows
template <class T>
void add_weld(const Model<T>& m, Env<T>& e, int l1, int l2, const T* data, const T* sp) {
  const int D = m.D;
  const Q4<T> q1 = e.l_quat[l1], q2 = e.l_quat[l2];
  const Q4<T> rel(data[6], data[7], data[8], data[9]);
  const T ts = data[10];
  const V3<T> ga1 = transform_by_trans_quat(V3<T>(data[3], data[4], data[5]), e.l_pos[l1], q1);
  const V3<T> ga2 = transform_by_trans_quat(V3<T>(data[0], data[1], data[2]), e.l_pos[l2], q2);
  const V3<T> pe = ga1 - ga2;
  const Q4<T> q = quat_mul(q1, rel);
  const Q4<T> iq2 = inv_quat(q2);
  const Q4<T> eq = quat_mul(iq2, q);
  const V3<T> re = V3<T>(eq.x, eq.y, eq.z) * ts;
  const T pos_imp = std::sqrt(pe[0] * pe[0] + pe[1] * pe[1] + pe[2] * pe[2] + re[0] * re[0] + re[1] * re[1] + re[2] * re[2]);
  const T invw0 = m.link_invweight[2 * l1] + m.link_invweight[2 * l2];
  const T invw1 = m.link_invweight[2 * l1 + 1] + m.link_invweight[2 * l2 + 1];
  for (int i = 0; i < 3; i++) {
    const int n = e.n_con++;
    e.n_con_eq++;
    const T jq = point_diff_row(m, e, n, l1, ga1, l2, ga2, i);
    eq_row_finish(m, e, n, sp, pos_imp, jq, pe[i], invw0);
  }
  const int n0 = e.n_con;
  e.n_con += 3;
  e.n_con_eq += 3;
  for (int i = 0; i < 3; i++)
    for (int d = 0; d < D; d++) e.jac[(size_t)(n0 + i) * D + d] = 0;
  for (int iab = 0; iab < 2; iab++) {
    const T sign = iab == 0 ? T(1) : T(-1);
    int link = iab == 0 ? l1 : l2;
    while (link > -1) {
      const int nd = m.link_dof_end[link] - m.link_dof_start[link];
      for (int k = 0; k < nd; k++) {
        const int d = m.link_dof_end[link] - 1 - k;
        const V3<T> jac = sign * e.cdof_ang[d];
        for (int i = 0; i < 3; i++) e.jac[(size_t)(n0 + i) * D + d] = e.jac[(size_t)(n0 + i) * D + d] + jac[i];
      }
      link = m.link_parent[link];
    }
  }
  T jq[3] = {0, 0, 0};
  for (int d = 0; d < D; d++) {
    const V3<T> jr(e.jac[(size_t)n0 * D + d], e.jac[(size_t)(n0 + 1) * D + d], e.jac[(size_t)(n0 + 2) * D + d]);
    const Q4<T> q3 = quat_mul(quat_mul_axis(iq2, jr), q);
    const T v3[3] = {q3.x, q3.y, q3.z};
    for (int i = 0; i < 3; i++) {
      e.jac[(size_t)(n0 + i) * D + d] = T(0.5) * v3[i] * ts;
      jq[i] = jq[i] + e.jac[(size_t)(n0 + i) * D + d] * e.vel[d];
    }
  }
  for (int i = 0; i < 3; i++) eq_row_finish(m, e, n0 + i, sp, pos_imp, jq[i], re[i], invw1);
}

// add_equality_constraints (constraint_solver_decomp.py:778-822): the model's equalities, then the env's own welds
template <class T>
void add_equality(const Model<T>& m, Env<T>& e) {
  const int D = m.D;
  e.n_con = 0; e.n_con_eq = 0;
  for (int ie = 0; ie < m.NEQ; ie++) {
    const T* sp = &m.eq_sol_params[7 * ie];
    const T* data = &m.eq_data[11 * ie];
    if (m.eq_type[ie] == EQ_CONNECT) { add_connect(m, e, m.eq_obj1[ie], m.eq_obj2[ie], data, sp); continue; }
    if (m.eq_type[ie] == EQ_WELD) { add_weld(m, e, m.eq_obj1[ie], m.eq_obj2[ie], data, sp); continue; }
    int j1 = m.eq_obj1[ie], j2 = m.eq_obj2[ie];
    int q1 = m.joint_q_start[j1], q2 = m.joint_q_start[j2], d1 = m.joint_dof_start[j1], d2 = m.joint_dof_start[j2];
    int n = e.n_con++;
    e.n_con_eq++;
    for (int d = 0; d < D; d++) e.jac[(size_t)n * D + d] = 0;
    T pos1 = e.qpos[q1], pos2 = e.qpos[q2], ref1 = m.qpos0[q1], ref2 = m.qpos0[q2];
    T diff = pos2 - ref2, pos = pos1 - ref1, deriv = 0;
    for (int i5 = 0; i5 < 5; i5++) {
      T dp = tpow(diff, T(i5));
      pos = pos - dp * data[i5];
      if (i5 < 4) deriv = deriv + data[i5 + 1] * dp * T(i5 + 1);
    }
    e.jac[(size_t)n * D + d1] = 1;
    e.jac[(size_t)n * D + d2] = -deriv;
    T jac_qvel = e.jac[(size_t)n * D + d1] * e.vel[d1] + e.jac[(size_t)n * D + d2] * e.vel[d2];
    T invw = m.dof_invweight[d1] + m.dof_invweight[d2];
    T imp, aref;
    imp_aref(sp, -std::fabs(pos), jac_qvel, pos, imp, aref);
    T diag = std::max(invw * (T(1) - imp) / imp, m.EPS);
    e.diag[n] = diag; e.aref[n] = aref; e.efc_D[n] = T(1) / diag;
  }
  // welds registered per env (kernel_add_weld_constraint's solver parameters, constraint_solver_decomp.py:2475-2477)
  const T sp_dyn[7] = {T(2) * m.dt, T(1), T(0.9f), T(0.95f), T(0.001f), T(0.5f), T(2)};
  for (int k = 0; k < e.eq_dyn_n; k++)
    add_weld(m, e, e.eq_dyn_obj[2 * k], e.eq_dyn_obj[2 * k + 1], &e.eq_dyn_data[11 * (size_t)k], sp_dyn);
}

template <class T>
void add_frictionloss(const Model<T>& m, Env<T>& e) {
  const int D = m.D;
  e.n_con_fl = 0;
  for (int l = 0; l < m.L; l++)
    for (int j = m.link_joint_start[l]; j < m.link_joint_end[l]; j++)
      for (int d = m.joint_dof_start[j]; d < m.joint_dof_end[j]; d++) {
        if (m.dof_frictionloss[d] > m.EPS) {
          T jac_qvel = e.vel[d];
          T imp, aref;
          imp_aref(&m.joint_sol_params[7 * j], T(0), jac_qvel, T(0), imp, aref);
          T diag = std::max(m.dof_invweight[d] * (T(1) - imp) / imp, m.EPS);
          int n = e.n_con++;
          e.n_con_fl++;
          e.diag[n] = diag; e.aref[n] = aref; e.efc_D[n] = T(1) / diag; e.efc_fl[n] = m.dof_frictionloss[d];
          for (int dd = 0; dd < D; dd++) e.jac[(size_t)n * D + dd] = 0;
          e.jac[(size_t)n * D + d] = 1;
        }
      }
}

template <class T>
void add_collision_constraints(const Model<T>& m, Env<T>& e) {
  const int D = m.D;
  for (int ic = 0; ic < e.n_contacts; ic++) {
    const Contact<T>& c = e.contacts[ic];
    V3<T> d1, d2;
    orthogonals(c.normal, d1, d2);
    T invw = m.link_invweight[2 * c.link_a];
    if (c.link_b > -1) invw = invw + m.link_invweight[2 * c.link_b];
    for (int i = 0; i < 4; i++) {
      V3<T> dd = T(2 * (i % 2) - 1) * (i < 2 ? d1 : d2);
      V3<T> n = dd * c.friction - c.normal;
      int nc = e.n_con++;
      for (int d = 0; d < D; d++) e.jac[(size_t)nc * D + d] = 0;
      T jac_qvel = 0;
      for (int iab = 0; iab < 2; iab++) {
        T sign = iab == 0 ? T(-1) : T(1);
        int link = iab == 0 ? c.link_a : c.link_b;
        while (link > -1) {
          int nd = m.link_dof_end[link] - m.link_dof_start[link];
          for (int k = 0; k < nd; k++) {
            int d = m.link_dof_end[link] - 1 - k;
            V3<T> tpos = c.pos - e.root_com[link];
            // ti_transform_motion_by_trans_quat with identity quat: vel = cdof_vel - tpos x cdof_ang
            V3<T> vel = transform_by_quat(e.cdof_vel[d] - cross(tpos, e.cdof_ang[d]), Q4<T>(1, 0, 0, 0));
            V3<T> diff = sign * vel;
            T jac = dot(diff, n);
            jac_qvel = jac_qvel + jac * e.vel[d];
            e.jac[(size_t)nc * D + d] = e.jac[(size_t)nc * D + d] + jac;
          }
          link = m.link_parent[link];
        }
      }
      T imp, aref;
      imp_aref(c.sol_params, -c.penetration, jac_qvel, -c.penetration, imp, aref);
      T diag = invw + c.friction * c.friction * invw;
      diag *= T(2) * c.friction * c.friction * (T(1) - imp) / imp;
      diag = std::max(diag, m.EPS);
      e.diag[nc] = diag; e.aref[nc] = aref; e.efc_D[nc] = T(1) / diag;
    }
  }
}

template <class T>
void add_joint_limits(const Model<T>& m, Env<T>& e) {
  const int D = m.D;
  for (int l = 0; l < m.L; l++)
    for (int j = m.link_joint_start[l]; j < m.link_joint_end[l]; j++) {
      int jt = m.joint_type[j];
      if (jt != JOINT_REVOLUTE && jt != JOINT_PRISMATIC) continue;
      int q = m.joint_q_start[j], d = m.joint_dof_start[j];
      T dmin = e.qpos[q] - m.dof_limit[2 * d], dmax = m.dof_limit[2 * d + 1] - e.qpos[q];
      T pd = std::min(dmin, dmax);
      if (pd < 0) {
        T jac = (dmin < dmax) ? T(1) : T(-1);
        T jac_qvel = jac * e.vel[d];
        T imp, aref;
        imp_aref(&m.joint_sol_params[7 * j], pd, jac_qvel, pd, imp, aref);
        T diag = std::max(m.dof_invweight[d] * (T(1) - imp) / imp, m.EPS);
        int n = e.n_con++;
        e.diag[n] = diag; e.aref[n] = aref; e.efc_D[n] = T(1) / diag;
        for (int dd = 0; dd < D; dd++) e.jac[(size_t)n * D + dd] = 0;
        e.jac[(size_t)n * D + d] = jac;
      }
    }
}

// --- Newton solver ---
template <class T>
void update_constraint(const Model<T>& m, Env<T>& e, const std::vector<T>& qacc, const std::vector<T>& Ma, T& cost) {
  const int D = m.D;
  int ne = e.n_con_eq, nef = ne + e.n_con_fl;
  e.prev_cost = cost;
  T cost_i = 0, gauss_i = 0;
  for (int c = 0; c < e.n_con; c++) {
    e.prev_active[c] = e.active[c];
    e.active[c] = 1;
    T floss_force = 0;
    if (ne <= c && c < nef) {
      T f = e.efc_fl[c], r = e.diag[c], rf = r * f;
      bool ln = e.Jaref[c] <= -rf, lp = e.Jaref[c] >= rf;
      e.active[c] = !(ln || lp);
      floss_force = T(ln) * f + T(lp) * -f;
      T fc = T(ln) * f * (T(-0.5) * rf - e.Jaref[c]);
      fc += T(lp) * f * (T(-0.5) * rf + e.Jaref[c]);
      cost_i += fc;
    } else if (nef <= c) {
      e.active[c] = e.Jaref[c] < 0;
    }
    e.efc_force[c] = floss_force + (-e.Jaref[c] * e.efc_D[c] * T(e.active[c]));
  }
  for (int d = 0; d < D; d++) {
    T q = 0;
    for (int c = 0; c < e.n_con; c++) q += e.jac[(size_t)c * D + d] * e.efc_force[c];
    e.qfrc_constraint[d] = q;
  }
  for (int d = 0; d < D; d++) {
    T v = T(0.5) * (Ma[d] - e.force[d]) * (qacc[d] - e.acc_smooth[d]);
    gauss_i += v;
    cost_i += v;
  }
  for (int c = 0; c < e.n_con; c++) cost_i += T(0.5) * (e.Jaref[c] * e.Jaref[c] * e.efc_D[c] * T(e.active[c]));
  e.gauss = gauss_i;
  cost = cost_i;
}

template <class T>
void chol_factor(const Model<T>& m, Env<T>& e) {
  const int D = m.D;
  std::vector<T>& H = e.nt_H;
  for (int i = 0; i < D; i++) {
    T tmp = H[(size_t)i * D + i];
    for (int j = 0; j < i; j++) tmp -= H[(size_t)i * D + j] * H[(size_t)i * D + j];
    H[(size_t)i * D + i] = std::sqrt(std::max(tmp, m.EPS));
    tmp = T(1) / H[(size_t)i * D + i];
    for (int j = i + 1; j < D; j++) {
      T dt = 0;
      for (int k = 0; k < i; k++) dt += H[(size_t)j * D + k] * H[(size_t)i * D + k];
      H[(size_t)j * D + i] = (H[(size_t)j * D + i] - dt) * tmp;
    }
  }
}

template <class T>
void hessian_direct(const Model<T>& m, Env<T>& e) {
  const int D = m.D;
  std::vector<T>& H = e.nt_H;
  for (int i = 0; i < D; i++) for (int j = 0; j <= i; j++) H[(size_t)i * D + j] = 0;
  for (int i = 0; i < D; i++)
    for (int c = 0; c < e.n_con; c++) {
      if (std::fabs(e.jac[(size_t)c * D + i]) > m.EPS)
        for (int j = 0; j <= i; j++)
          H[(size_t)i * D + j] = H[(size_t)i * D + j] + e.jac[(size_t)c * D + j] * e.jac[(size_t)c * D + i] * e.efc_D[c] * T(e.active[c]);
    }
  for (int ie = 0; ie < m.E; ie++)
    for (int i = m.ent_dof_start[ie]; i < m.ent_dof_end[ie]; i++)
      for (int j = m.ent_dof_start[ie]; j <= i; j++) H[(size_t)i * D + j] = H[(size_t)i * D + j] + e.mass_mat[(size_t)i * D + j];
  chol_factor(m, e);
}

template <class T>
void hessian_incremental(const Model<T>& m, Env<T>& e) {
  const int D = m.D;
  std::vector<T>& H = e.nt_H;
  bool updated = false;
  for (int c = 0; c < e.n_con; c++) {
    if (updated) continue;
    int flag = -1;
    if (e.prev_active[c] == 0 && e.active[c] == 1) flag = 1;
    if (e.prev_active[c] == 1 && e.active[c] == 0) flag = 0;
    if (flag == -1) continue;
    for (int d = 0; d < D; d++) e.nt_vec[d] = e.jac[(size_t)c * D + d] * std::sqrt(e.efc_D[c]);
    int rank = D;
    T sgn = T(flag * 2 - 1);
    for (int k = 0; k < D; k++) {
      if (std::fabs(e.nt_vec[k]) > m.EPS) {
        T Lkk = H[(size_t)k * D + k];
        T tmp = Lkk * Lkk + e.nt_vec[k] * e.nt_vec[k] * sgn;
        if (tmp < m.EPS) { tmp = m.EPS; rank--; }
        T r = std::sqrt(tmp), cc = r / Lkk, cinv = T(1) / cc, s = e.nt_vec[k] / Lkk;
        H[(size_t)k * D + k] = r;
        for (int i = k + 1; i < D; i++) H[(size_t)i * D + k] = (H[(size_t)i * D + k] + s * e.nt_vec[i] * sgn) * cinv;
        for (int i = k + 1; i < D; i++) e.nt_vec[i] = e.nt_vec[i] * cc - s * H[(size_t)i * D + k];
      }
    }
    if (rank < D) { hessian_direct(m, e); updated = true; }
  }
}

template <class T>
void chol_solve(const Model<T>& m, Env<T>& e) {
  const int D = m.D;
  const std::vector<T>& H = e.nt_H;
  for (int i = 0; i < D; i++) e.Mgrad[i] = e.grad[i];
  for (int i = 0; i < D; i++) {
    for (int j = 0; j < i; j++) e.Mgrad[i] = e.Mgrad[i] - H[(size_t)i * D + j] * e.Mgrad[j];
    e.Mgrad[i] = e.Mgrad[i] / H[(size_t)i * D + i];
  }
  for (int ii = 0; ii < D; ii++) {
    int i = D - 1 - ii;
    for (int j = i + 1; j < D; j++) e.Mgrad[i] = e.Mgrad[i] - H[(size_t)j * D + i] * e.Mgrad[j];
    e.Mgrad[i] = e.Mgrad[i] / H[(size_t)i * D + i];
  }
}

template <class T>
void update_gradient(const Model<T>& m, Env<T>& e) {
  for (int d = 0; d < m.D; d++) e.grad[d] = e.Ma[d] - e.force[d] - e.qfrc_constraint[d];
  chol_solve(m, e);
}

template <class T>
void ls_init(const Model<T>& m, Env<T>& e) {
  const int D = m.D;
  for (int ie = 0; ie < m.E; ie++)
    for (int i = m.ent_dof_start[ie]; i < m.ent_dof_end[ie]; i++) {
      T mv = 0;
      for (int j = m.ent_dof_start[ie]; j < m.ent_dof_end[ie]; j++) mv += e.mass_mat[(size_t)i * D + j] * e.search[j];
      e.mv[i] = mv;
    }
  for (int c = 0; c < e.n_con; c++) {
    T jv = 0;
    for (int d = 0; d < D; d++) jv += e.jac[(size_t)c * D + d] * e.search[d];
    e.jv[c] = jv;
  }
  T qg1 = 0, qg2 = 0;
  for (int d = 0; d < D; d++) {
    qg1 += e.search[d] * e.Ma[d] - e.search[d] * e.force[d];
    qg2 += T(0.5) * e.search[d] * e.mv[d];
  }
  e.quad_gauss[0] = e.gauss; e.quad_gauss[1] = qg1; e.quad_gauss[2] = qg2;
  for (int c = 0; c < e.n_con; c++) {
    e.quad[3 * c + 0] = e.efc_D[c] * (T(0.5) * e.Jaref[c] * e.Jaref[c]);
    e.quad[3 * c + 1] = e.efc_D[c] * (e.jv[c] * e.Jaref[c]);
    e.quad[3 * c + 2] = e.efc_D[c] * (T(0.5) * e.jv[c] * e.jv[c]);
  }
}

template <class T>
struct LSP { T alpha, cost, d0, d1; };

template <class T>
LSP<T> ls_point(const Model<T>& m, Env<T>& e, T alpha) {
  int ne = e.n_con_eq, nef = ne + e.n_con_fl;
  T t0 = e.quad_gauss[0], t1 = e.quad_gauss[1], t2 = e.quad_gauss[2];
  for (int c = 0; c < e.n_con; c++) {
    T x = e.Jaref[c] + alpha * e.jv[c];
    T q0 = e.quad[3 * c], q1 = e.quad[3 * c + 1], q2 = e.quad[3 * c + 2];
    bool act = true;
    if (ne <= c && c < nef) {
      T f = e.efc_fl[c], r = e.diag[c], rf = r * f;
      bool ln = x <= -rf, lp = x >= rf;
      if (ln || lp) {
        q0 = T(ln) * f * (T(-0.5) * rf - e.Jaref[c]) + T(lp) * f * (T(-0.5) * rf + e.Jaref[c]);
        q1 = T(ln) * (-f * e.jv[c]) + T(lp) * (f * e.jv[c]);
        q2 = 0;
      }
    } else if (nef <= c) {
      act = x < 0;
    }
    t0 += q0 * T(act); t1 += q1 * T(act); t2 += q2 * T(act);
  }
  T cost = alpha * alpha * t2 + alpha * t1 + t0;
  T d0 = T(2) * alpha * t2 + t1;
  T d1 = T(2) * t2;
  if (d1 <= 0) d1 = m.EPS;
  e.ls_it++;
  return {alpha, cost, d0, d1};
}

template <class T>
void update_bracket(const Model<T>& m, Env<T>& e, LSP<T>& p, int& flag, LSP<T>& pn) {
  flag = 0;
  for (int i = 0; i < 3; i++) {
    T* c = &e.candidates[4 * i];
    if (p.d0 < 0 && c[2] < 0 && p.d0 < c[2]) { p = {c[0], c[1], c[2], c[3]}; flag = 1; }
    else if (p.d0 > 0 && c[2] > 0 && p.d0 > c[2]) { p = {c[0], c[1], c[2], c[3]}; flag = 2; }
  }
  pn = p;
  if (flag > 0) pn = ls_point(m, e, p.alpha - p.d0 / p.d1);
}

// func_linesearch (cs:1565-1743)
template <class T>
T linesearch(const Model<T>& m, Env<T>& e) {
  const int D = m.D;
  T snorm = 0;
  for (int d = 0; d < D; d++) snorm += e.search[d] * e.search[d];
  snorm = std::sqrt(snorm);
  T scale = m.meaninertia * T(std::max(1, D));
  T gtol = m.tolerance * m.ls_tolerance * snorm * scale;
  e.ls_it = 0;
  T res = 0;
  if (snorm < m.EPS) return T(0);
  ls_init(m, e);
  LSP<T> p0 = ls_point(m, e, T(0));
  LSP<T> p1 = ls_point(m, e, p0.alpha - p0.d0 / p0.d1);
  if (p0.cost < p1.cost) p1 = p0;
  if (std::fabs(p1.d0) < gtol) return p1.alpha;
  T direction = (p1.d0 < 0) ? T(1) : T(-1);
  int p2update = 0;
  LSP<T> p2 = p1;
  bool done = false;
  while (p1.d0 * direction <= -gtol && e.ls_it < m.ls_iterations) {
    p2 = p1;
    p2update = 1;
    p1 = ls_point(m, e, p1.alpha - p1.d0 / p1.d1);
    if (std::fabs(p1.d0) < gtol) { res = p1.alpha; done = true; break; }
  }
  if (!done) {
    if (e.ls_it >= m.ls_iterations) { res = p1.alpha; done = true; }
    if (!p2update && !done) { res = p1.alpha; done = true; }
    if (!done) {
      LSP<T> p2n = p1;
      LSP<T> p1n = ls_point(m, e, p1.alpha - p1.d0 / p1.d1);
      while (e.ls_it < m.ls_iterations) {
        LSP<T> pm = ls_point(m, e, (p1.alpha + p2.alpha) * T(0.5));
        T* c = e.candidates;
        c[0] = p1n.alpha; c[1] = p1n.cost; c[2] = p1n.d0; c[3] = p1n.d1;
        c[4] = p2n.alpha; c[5] = p2n.cost; c[6] = p2n.d0; c[7] = p2n.d1;
        c[8] = pm.alpha; c[9] = pm.cost; c[10] = pm.d0; c[11] = pm.d1;
        int best = -1;
        T bcost = 0;
        for (int ii = 0; ii < 3; ii++)
          if (std::fabs(c[4 * ii + 2]) < gtol && (best < 0 || c[4 * ii + 1] < bcost)) { bcost = c[4 * ii + 1]; best = ii; }
        if (best >= 0) { res = c[4 * best]; done = true; break; }
        int b1, b2;
        update_bracket(m, e, p1, b1, p1n);
        update_bracket(m, e, p2, b2, p2n);
        if (b1 == 0 && b2 == 0) { res = pm.alpha; done = true; break; }
      }
      if (!done) {
        if (p1.cost <= p2.cost && p1.cost < p0.cost) res = p1.alpha;
        else if (p2.cost <= p1.cost && p2.cost < p1.cost) res = p2.alpha;
        else res = 0;
      }
    }
  }
  return res;
}

template <class T>
void init_solver(const Model<T>& m, Env<T>& e) {
  const int D = m.D;
  for (int d = 0; d < D; d++) e.qacc[d] = (e.n_con > 0 && e.is_warmstart) ? e.qacc_ws[d] : e.acc_smooth[d];
  for (int i = 0; i < D; i++) {
    int ie = m.dof_entity[i];
    T s = 0;
    for (int j = m.ent_dof_start[ie]; j < m.ent_dof_end[ie]; j++) s += e.mass_mat[(size_t)i * D + j] * e.qacc[j];
    e.Ma[i] = s;
  }
  for (int c = 0; c < e.n_con; c++) {
    T ja = -e.aref[c];
    for (int d = 0; d < D; d++) ja += e.jac[(size_t)c * D + d] * e.qacc[d];
    e.Jaref[c] = ja;
  }
  update_constraint(m, e, e.qacc, e.Ma, e.cost);
  hessian_direct(m, e);
  update_gradient(m, e);
  for (int d = 0; d < D; d++) e.search[d] = -e.Mgrad[d];
}

template <class T>
void solve(const Model<T>& m, Env<T>& e) {
  const int D = m.D;
  e.n_newton_iters = 0;
  if (e.n_con <= 0) { e.improved = 0; return; }
  for (int it = 0; it < m.iterations; it++) {
    e.n_newton_iters++;
    T alpha = linesearch(m, e);
    if (std::fabs(alpha) < m.EPS) {
      e.improved = 0;
    } else {
      for (int d = 0; d < D; d++) { e.qacc[d] = e.qacc[d] + e.search[d] * alpha; e.Ma[d] = e.Ma[d] + e.mv[d] * alpha; }
      for (int c = 0; c < e.n_con; c++) e.Jaref[c] = e.Jaref[c] + e.jv[c] * alpha;
      update_constraint(m, e, e.qacc, e.Ma, e.cost);
      hessian_incremental(m, e);
      update_gradient(m, e);
      T tol_scaled = (m.meaninertia * T(std::max(1, D))) * m.tolerance;
      T improvement = e.prev_cost - e.cost;
      T gn = 0;
      for (int d = 0; d < D; d++) gn += e.grad[d] * e.grad[d];
      gn = std::sqrt(gn);
      if (gn < tol_scaled || improvement < tol_scaled) e.improved = 0;
      else {
        e.improved = 1;
        for (int d = 0; d < D; d++) e.search[d] = -e.Mgrad[d];
      }
    }
    if (!e.improved) break;
  }
}

// func_update_qacc (cs:1385-1406) and func_update_contact_force (cs:1343-1382)
template <class T>
void update_qacc_and_forces(const Model<T>& m, Env<T>& e) {
  for (int d = 0; d < m.D; d++) {
    e.acc[d] = e.qacc[d];
    e.force[d] = e.qf_smooth[d] + e.qfrc_constraint[d];
    e.qacc_ws[d] = e.qacc[d];
    if (std::isnan(e.qacc[d])) e.errno_ |= 4u;
  }
  e.is_warmstart = 1;
  int cs = e.n_con_eq + e.n_con_fl;
  for (int ic = 0; ic < e.n_contacts; ic++) {
    Contact<T>& c = e.contacts[ic];
    V3<T> f, d1, d2;
    orthogonals(c.normal, d1, d2);
    for (int i = 0; i < 4; i++) {
      V3<T> dd = T(2 * (i % 2) - 1) * (i < 2 ? d1 : d2);
      V3<T> n = dd * c.friction - c.normal;
      f = f + n * e.efc_force[ic * 4 + i + cs];
    }
    c.force = f;
  }
}

// func_integrate (rsd:6566-6708) + func_copy_next_to_curr (rsd:6711-6740)
template <class T>
void integrate(const Model<T>& m, Env<T>& e) {
  const T dt = m.dt;
  for (int d = 0; d < m.D; d++) e.vel_next[d] = e.vel[d] + e.acc[d] * dt;
  for (int l = 0; l < m.L; l++) {
    if (m.link_dof_end[l] - m.link_dof_start[l] <= 0) continue;
    int ds = m.link_dof_start[l], qs = m.link_q_start[l], qe = m.link_q_end[l];
    int jt = m.joint_type[m.link_joint_start[l]];
    if (jt == JOINT_FREE) {
      V3<T> pos(e.qpos[qs], e.qpos[qs + 1], e.qpos[qs + 2]);
      V3<T> v(e.vel_next[ds], e.vel_next[ds + 1], e.vel_next[ds + 2]);
      pos = pos + v * dt;
      for (int j = 0; j < 3; j++) e.qpos_next[qs + j] = pos[j];
    }
    if (jt == JOINT_SPHERICAL || jt == JOINT_FREE) {
      int ro = jt == JOINT_FREE ? 3 : 0;
      Q4<T> rot0(e.qpos[qs + ro], e.qpos[qs + ro + 1], e.qpos[qs + ro + 2], e.qpos[qs + ro + 3]);
      V3<T> ang = V3<T>(e.vel_next[ds + ro], e.vel_next[ds + ro + 1], e.vel_next[ds + ro + 2]) * dt;
      Q4<T> qrot = rotvec_to_quat(ang, m.EPS);
      Q4<T> rot = transform_quat_by_quat(qrot, rot0);
      e.qpos_next[qs + ro] = rot.w; e.qpos_next[qs + ro + 1] = rot.x; e.qpos_next[qs + ro + 2] = rot.y; e.qpos_next[qs + ro + 3] = rot.z;
    } else {
      for (int j = 0; j < qe - qs; j++) e.qpos_next[qs + j] = e.qpos[qs + j] + e.vel_next[ds + j] * dt;
    }
  }
  // links without dofs keep qpos (they have no qs)
  bool valid = true;
  for (int d = 0; d < m.D; d++) valid &= !std::isnan(e.vel_next[d]);
  for (int q = 0; q < m.Q; q++) valid &= !std::isnan(e.qpos_next[q]);
  if (valid) {
    for (int d = 0; d < m.D; d++) e.vel[d] = e.vel_next[d];
    for (int q = 0; q < m.Q; q++) e.qpos[q] = e.qpos_next[q];
  } else {
    e.errno_ |= 8u;
  }
}

template <class T>
void kinematics(const Model<T>& m, Env<T>& e) {
  forward_kinematics(m, e);
  com_links(m, e);
  update_geoms(m, e);
  forward_velocity(m, e);
}

template <class T>
void step(const Model<T>& m, Env<T>& e) {
  // kernel_step_1 (FK state is a pure function of qpos/vel, recomputed here)
  kinematics(m, e);
  compute_mass_matrix(m, e);
  factor_mass(m, e);
  torque_and_passive(m, e);
  update_acc(m, e);
  update_force(m, e);
  bias_and_qacc(m, e);
  // _func_constraint_force (rsd:973-989)
  add_equality(m, e);
  e.n_broad = 0;
  e.n_contacts = 0;
  if (m.enable_collision) {
    update_aabbs(m, e);
    if (m.P > 0) {
      broad_phase(m, e);
      narrow_phase(m, e);
    }
  }
  add_frictionloss(m, e);
  if (m.enable_collision) add_collision_constraints(m, e);
  if (m.enable_joint_limit) add_joint_limits(m, e);
  init_solver(m, e);
  solve(m, e);
  update_qacc_and_forces(m, e);
  // kernel_step_2
  integrate(m, e);
  kinematics(m, e);
}

}  // namespace

// ------------------------------------------------------------------------------------------------------------
// C interface (ctypes)
// ------------------------------------------------------------------------------------------------------------
struct OrcHandle {
  RawModel raw;
  int precision = 32;
  Model<float> mf;
  Model<double> md;
  std::vector<Env<float>> ef;
  std::vector<Env<double>> ed;
  int B = 0;
};

extern "C" void* orc_new() { return new OrcHandle(); }
extern "C" void orc_free(void* h) { delete static_cast<OrcHandle*>(h); }
extern "C" void orc_set_i(void* h, const char* name, const int32_t* p, int64_t n) {
  static_cast<OrcHandle*>(h)->raw.ints[name].assign(p, p + n);
}
extern "C" void orc_set_f(void* h, const char* name, const double* p, int64_t n) {
  static_cast<OrcHandle*>(h)->raw.flts[name].assign(p, p + n);
}

// Finalize the model and allocate B env states.
extern "C" void orc_build(void* hp, int B, int precision) {
  OrcHandle* h = static_cast<OrcHandle*>(hp);
  h->precision = precision;
  h->B = B;
  if (precision == 64) {
    h->md.load(h->raw);
    h->ed.assign(B, Env<double>());
    for (auto& e : h->ed) e.alloc(h->md);
  } else {
    h->mf.load(h->raw);
    h->ef.assign(B, Env<float>());
    for (auto& e : h->ef) e.alloc(h->mf);
  }
}

// Generic state access: field names below; arrays are env-major (B, n) float64 / int32.
#define ORC_FIELDS_F(X) X(qpos) X(vel) X(qacc_ws) X(ctrl_pos) X(ctrl_vel) X(ctrl_force) X(acc) X(acc_smooth) X(dofs_pos) X(sort_value) \
  X(mass_shift) X(friction_ratio) X(eq_dyn_data)
#define ORC_FIELDS_I(X) X(ctrl_mode) X(sort_ig) X(sort_ismax) X(eq_dyn_obj)

template <class T>
static void get_set_env(Env<T>& e, const Model<T>& m, const char* name, double* f, int32_t* iv, int set) {
  std::string n(name);
#define X(fld) if (n == #fld) { auto& v = e.fld; if (set) for (size_t i = 0; i < v.size(); i++) v[i] = T(f[i]); else for (size_t i = 0; i < v.size(); i++) f[i] = double(v[i]); return; }
  ORC_FIELDS_F(X)
#undef X
#define X(fld) if (n == #fld) { auto& v = e.fld; if (set) for (size_t i = 0; i < v.size(); i++) v[i] = iv[i]; else for (size_t i = 0; i < v.size(); i++) iv[i] = v[i]; return; }
  ORC_FIELDS_I(X)
#undef X
  if (n == "is_warmstart") { if (set) e.is_warmstart = iv[0]; else iv[0] = e.is_warmstart; return; }
  if (n == "first_time") { if (set) e.first_time = iv[0]; else iv[0] = e.first_time; return; }
  if (n == "eq_dyn_n") { if (set) e.eq_dyn_n = iv[0]; else iv[0] = e.eq_dyn_n; return; }
  if (n == "errno") { if (set) e.errno_ = uint32_t(iv[0]); else iv[0] = int32_t(e.errno_); return; }
  if (n == "n_contacts") { if (set) e.n_contacts = iv[0]; else iv[0] = e.n_contacts; return; }
  if (n == "n_broad") { if (set) e.n_broad = iv[0]; else iv[0] = e.n_broad; return; }
  if (n == "n_con") { iv[0] = e.n_con; return; }
  if (n == "n_con_eq") { iv[0] = e.n_con_eq; return; }
  if (n == "n_con_fl") { iv[0] = e.n_con_fl; return; }
  // constraint rows (reads only): dense Jacobian rows [n_con][D] and per-row scalars [n_con]
  if (n == "efc_jac") { for (size_t i = 0; i < (size_t)e.n_con * m.D; i++) f[i] = double(e.jac[i]); return; }
  if (n == "efc_D") { for (int i = 0; i < e.n_con; i++) f[i] = double(e.efc_D[i]); return; }
  if (n == "efc_aref") { for (int i = 0; i < e.n_con; i++) f[i] = double(e.aref[i]); return; }
  if (n == "efc_diag") { for (int i = 0; i < e.n_con; i++) f[i] = double(e.diag[i]); return; }
  if (n == "efc_fl") { for (int i = 0; i < e.n_con; i++) f[i] = double(e.efc_fl[i]); return; }
  if (n == "efc_force") { for (int i = 0; i < e.n_con; i++) f[i] = double(e.efc_force[i]); return; }
  if (n == "qf_smooth") { for (int i = 0; i < m.D; i++) f[i] = double(e.qf_smooth[i]); return; }
  if (n == "qacc") { for (int i = 0; i < m.D; i++) f[i] = double(e.qacc[i]); return; }
  if (n == "n_newton_iters") { iv[0] = e.n_newton_iters; return; }
  if (n == "cache_normal") {
    for (size_t i = 0; i < e.cache_normal.size(); i++)
      for (int k = 0; k < 3; k++) { if (set) e.cache_normal[i][k] = T(f[3 * i + k]); else f[3 * i + k] = double(e.cache_normal[i][k]); }
    return;
  }
  auto v3 = [&](std::vector<V3<T>>& v) {
    for (size_t i = 0; i < v.size(); i++) for (int k = 0; k < 3; k++) { if (set) v[i][k] = T(f[3 * i + k]); else f[3 * i + k] = double(v[i][k]); }
  };
  auto q4 = [&](std::vector<Q4<T>>& v) {
    for (size_t i = 0; i < v.size(); i++) {
      if (set) v[i] = Q4<T>(T(f[4 * i]), T(f[4 * i + 1]), T(f[4 * i + 2]), T(f[4 * i + 3]));
      else { f[4 * i] = v[i].w; f[4 * i + 1] = v[i].x; f[4 * i + 2] = v[i].y; f[4 * i + 3] = v[i].z; }
    }
  };
  if (n == "links_pos") return v3(e.l_pos);
  if (n == "i_pos_shift") return v3(e.i_pos_shift);
  if (n == "links_quat") return q4(e.l_quat);
  if (n == "geoms_pos") return v3(e.g_pos);
  if (n == "geoms_quat") return q4(e.g_quat);
  if (n == "aabb_min") return v3(e.aabb_min);
  if (n == "aabb_max") return v3(e.aabb_max);
  if (n == "cdof_ang") return v3(e.cdof_ang);
  if (n == "cdof_vel") return v3(e.cdof_vel);
  if (n == "cd_vel") return v3(e.cd_vel);
  if (n == "cd_ang") return v3(e.cd_ang);
  if (n == "root_com") return v3(e.root_com);
  if (n == "mass_mat") { for (size_t i = 0; i < e.mass_mat.size(); i++) f[i] = double(e.mass_mat[i]); return; }
  if (n == "broad_pairs") { for (int i = 0; i < e.n_broad; i++) { iv[2 * i] = e.broad_a[i]; iv[2 * i + 1] = e.broad_b[i]; } return; }
  if (n == "contacts") {
    // 16 doubles per contact: geom_a, geom_b, link_a, link_b, pos(3), normal(3), penetration, friction, force(3), pad
    for (int i = 0; i < e.n_contacts; i++) {
      const Contact<T>& c = e.contacts[i];
      double* o = f + 16 * i;
      o[0] = c.geom_a; o[1] = c.geom_b; o[2] = c.link_a; o[3] = c.link_b;
      for (int k = 0; k < 3; k++) { o[4 + k] = c.pos[k]; o[7 + k] = c.normal[k]; o[12 + k] = c.force[k]; }
      o[10] = c.penetration; o[11] = c.friction; o[15] = 0;
    }
    return;
  }
}

// Read (set=0) or write (set=1) field `name` of env b.
extern "C" int orc_env_io(void* hp, int b, const char* name, double* f, int32_t* iv, int set) {
  OrcHandle* h = static_cast<OrcHandle*>(hp);
  if (b < 0 || b >= h->B) return -1;
  if (h->precision == 64) get_set_env(h->ed[b], h->md, name, f, iv, set);
  else get_set_env(h->ef[b], h->mf, name, f, iv, set);
  return 0;
}

// Stages for tests: 0 = full step, 1 = kinematics only, 2 = step_1 (kinematics + smooth dynamics),
// 3 = step_1 + collision detection (AABB/broad/narrow), 4 = everything before integration (rows + Newton solve).
template <class T>
static void run_stage(const Model<T>& m, Env<T>& e, int stage) {
  if (stage == 0) { step(m, e); return; }
  kinematics(m, e);
  if (stage == 1) return;
  compute_mass_matrix(m, e); factor_mass(m, e); torque_and_passive(m, e); update_acc(m, e); update_force(m, e);
  bias_and_qacc(m, e);
  if (stage == 2) return;
  add_equality(m, e);
  e.n_broad = 0; e.n_contacts = 0;
  if (m.enable_collision) {
    update_aabbs(m, e);
    if (m.P > 0) { broad_phase(m, e); narrow_phase(m, e); }
  }
  if (stage == 3) return;
  add_frictionloss(m, e);
  if (m.enable_collision) add_collision_constraints(m, e);
  if (m.enable_joint_limit) add_joint_limits(m, e);
  init_solver(m, e);
  solve(m, e);
  update_qacc_and_forces(m, e);
}

// Advance envs [b0, b1) by n_steps with n_threads host threads (env-chunked).
extern "C" void orc_run(void* hp, int b0, int b1, int n_steps, int stage, int n_threads) {
  OrcHandle* h = static_cast<OrcHandle*>(hp);
  if (n_threads < 1) n_threads = 1;
  auto work = [&](int lo, int hi) {
    for (int b = lo; b < hi; b++)
      for (int s = 0; s < n_steps; s++) {
        if (h->precision == 64) run_stage(h->md, h->ed[b], stage);
        else run_stage(h->mf, h->ef[b], stage);
      }
  };
  if (n_threads == 1) { work(b0, b1); return; }
  std::vector<std::thread> th;
  int n = b1 - b0, chunk = (n + n_threads - 1) / n_threads;
  for (int t = 0; t < n_threads; t++) {
    int lo = b0 + t * chunk, hi = std::min(b1, lo + chunk);
    if (lo < hi) th.emplace_back(work, lo, hi);
  }
  for (auto& t : th) t.join();
}

// Broadphase alone on caller-supplied AABBs (for bit-exact pair-set checks from identical inputs).
extern "C" int orc_broadphase_from_aabbs(void* hp, int b, const double* aabb_min, const double* aabb_max, int32_t* pairs_out) {
  OrcHandle* h = static_cast<OrcHandle*>(hp);
  auto run = [&](auto& m, auto& e) {
    using TT = typename std::decay<decltype(e.qpos[0])>::type;
    for (int g = 0; g < m.G; g++)
      for (int k = 0; k < 3; k++) { e.aabb_min[g][k] = TT(aabb_min[3 * g + k]); e.aabb_max[g][k] = TT(aabb_max[3 * g + k]); }
    broad_phase(m, e);
    for (int i = 0; i < e.n_broad; i++) { pairs_out[2 * i] = e.broad_a[i]; pairs_out[2 * i + 1] = e.broad_b[i]; }
    return e.n_broad;
  };
  if (h->precision == 64) return run(h->md, h->ed[b]);
  return run(h->mf, h->ef[b]);
}


// Mesh-frame support-field lookup of geom g for direction d (test hook for the grid lookup):
// writes the support vertex and returns its index within the geom.
extern "C" int orc_support_mesh(void* hp, int g, const double* d, double* v_out) {
  OrcHandle* h = static_cast<OrcHandle*>(hp);
  auto run = [&](auto& m) {
    using TT = typename std::decay<decltype(m.dt)>::type;
    int vid = 0;
    V3<TT> v = support_mesh_local(m, V3<TT>(TT(d[0]), TT(d[1]), TT(d[2])), g, vid);
    v_out[0] = double(v.x); v_out[1] = double(v.y); v_out[2] = double(v.z);
    return vid;
  };
  if (h->precision == 64) return run(h->md);
  return run(h->mf);
}
