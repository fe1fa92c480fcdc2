// Per-environment view of the env-minor SoA state used by all gfx950 rigid-step kernels.
#pragma once
#include "rigid_math.h"
#include "genesis_rigid.h"

namespace gsr {

enum { GEOM_PLANE = 0, GEOM_SPHERE = 1, GEOM_ELLIPSOID = 2, GEOM_CYLINDER = 3, GEOM_CAPSULE = 4, GEOM_BOX = 5, GEOM_MESH = 6 };
enum { JOINT_FIXED = 0, JOINT_REVOLUTE = 1, JOINT_PRISMATIC = 2, JOINT_SPHERICAL = 3, JOINT_FREE = 4 };
enum { EQ_CONNECT = 0, EQ_WELD = 1, EQ_JOINT = 2 };
enum { CTRL_FORCE = 0, CTRL_VELOCITY = 1, CTRL_POSITION = 2 };

// compile-time bounds of per-thread scratch (checked on the host before every launch)
constexpr int kMaxRowWidth = 32;

struct Env {
  const gsr_model& m;
  const gsr_state& s;
  int b;
  int B;
  __device__ Env(const gsr_model& m_, const gsr_state& s_, int b_) : m(m_), s(s_), b(b_), B(s_.n_envs) {}

  __device__ __forceinline__ size_t ix(int i) const { return (size_t)i * (size_t)B + (size_t)b; }
  __device__ __forceinline__ float& F(float* a, int i) const { return a[ix(i)]; }
  __device__ __forceinline__ int32_t& I(int32_t* a, int i) const { return a[ix(i)]; }
  // per-dof parameter table entry i of this env: shared by every env, or env-minor when the model batches dofs
  __device__ __forceinline__ float DP(const float* a, int i) const { return m.batch_dofs ? a[ix(i)] : a[i]; }
  // per-link parameter table entry i (link_inertial_mass, link_invweight), per env when the model batches links
  __device__ __forceinline__ float LP(const float* a, int i) const { return m.batch_links ? a[ix(i)] : a[i]; }
  // the 7 solver parameters of joint j / equality i: first entry, stride between entries returned in st
  __device__ __forceinline__ const float* joint_sp(int j, size_t& st) const {
    st = m.batch_joints ? (size_t)B : 1;
    return m.joint_sol_params + (m.batch_joints ? ix(7 * j) : (size_t)(7 * j));
  }
  __device__ __forceinline__ const float* eq_sp(int i, size_t& st) const {
    st = m.batch_eqs ? (size_t)B : 1;
    return m.eq_sol_params + (m.batch_eqs ? ix(7 * i) : (size_t)(7 * i));
  }
  // the solver's tolerance scale of this env (rigid_global_info.meaninertia[i_b])
  __device__ __forceinline__ float meaninertia() const { return m.env_meaninertia ? m.env_meaninertia[b] : m.meaninertia; }
  __device__ __forceinline__ f3 L3(const float* a, int i) const {
    return f3{a[ix(3 * i)], a[ix(3 * i + 1)], a[ix(3 * i + 2)]};
  }
  __device__ __forceinline__ void S3(float* a, int i, f3 v) const {
    a[ix(3 * i)] = v.x;
    a[ix(3 * i + 1)] = v.y;
    a[ix(3 * i + 2)] = v.z;
  }
  __device__ __forceinline__ q4 L4(const float* a, int i) const {
    return q4{a[ix(4 * i)], a[ix(4 * i + 1)], a[ix(4 * i + 2)], a[ix(4 * i + 3)]};
  }
  __device__ __forceinline__ void S4(float* a, int i, q4 v) const {
    a[ix(4 * i)] = v.w;
    a[ix(4 * i + 1)] = v.x;
    a[ix(4 * i + 2)] = v.y;
    a[ix(4 * i + 3)] = v.z;
  }
  // Constraint rows and the Newton Hessian use one of two layouts, chosen per model by the solver variant:
  // env-minor like all other state (one lane per env: lanes read consecutive envs) or env-major (one wave per
  // env: the lanes of an env read consecutive rows).
  __device__ __forceinline__ size_t rx(int i, size_t per_env) const {
    return m.solver_variant ? (size_t)b * per_env + (size_t)i : (size_t)i * (size_t)B + (size_t)b;
  }
  __device__ __forceinline__ float& EF(float* a, int c) const { return a[rx(c, (size_t)m.max_rows)]; }
  __device__ __forceinline__ int32_t& EI(int32_t* a, int c) const { return a[rx(c, (size_t)m.max_rows)]; }
  __device__ __forceinline__ float& QD(int c, int k) const { return s.efc_quad[rx(3 * c + k, 3 * (size_t)m.max_rows)]; }
  __device__ __forceinline__ float& JV(int c, int k) const {
    return s.jac_val[rx(c * m.row_width + k, (size_t)m.max_rows * m.row_width)];
  }
  __device__ __forceinline__ int32_t& JI(int c, int k) const {
    return s.jac_idx[rx(c * m.row_width + k, (size_t)m.max_rows * m.row_width)];
  }
  // packed lower triangle of the Newton Hessian, j <= i
  __device__ __forceinline__ float& H(int i, int j) const {
    return s.nt_H[rx(i * (i + 1) / 2 + j, (size_t)m.n_dofs * (m.n_dofs + 1) / 2)];
  }
  // entity-block packed mass matrix
  __device__ __forceinline__ float& Mm(int e, int i, int j) const {
    int st = m.ent_dof_start[e], n = m.ent_dof_end[e] - st;
    return s.mass_blk[ix(m.ent_mass_offset[e] + (i - st) * n + (j - st))];
  }
  __device__ __forceinline__ float& ML(int e, int i, int j) const {
    int st = m.ent_dof_start[e], n = m.ent_dof_end[e] - st;
    return s.mass_L_blk[ix(m.ent_mass_offset[e] + (i - st) * n + (j - st))];
  }
};

// static tables (shared by all envs)
__device__ __forceinline__ f3 T3(const float* a, int i) { return f3{a[3 * i], a[3 * i + 1], a[3 * i + 2]}; }
__device__ __forceinline__ q4 T4(const float* a, int i) { return q4{a[4 * i], a[4 * i + 1], a[4 * i + 2], a[4 * i + 3]}; }

}  // namespace gsr
