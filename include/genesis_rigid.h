/*
 * genesis_rigid.h -- C-ABI of the MI355X batched rigid-body step (libgenesis_rigid.so).
 *
 * The reference has no FFI for this path: its boundary is the Python solver object. Each entry point below replaces
 * one call the reference's Simulator / RigidSolver makes on that object:
 *
 *   gsr_step               <- Simulator.step rigid-only branch (genesis/engine/simulator.py:268-292) calling
 *                             RigidSolver.substep (rigid_solver_decomp.py:865-933) = kernel_step_1 ->
 *                             _func_constraint_force (collider detection + constraint solver) -> kernel_step_2.
 *   gsr_forward_kinematics <- RigidSolver.set_qpos's FK refresh (rigid_solver_decomp.py:1801-1857,
 *                             kernel_forward_kinematics_links_geoms rsd:4599).
 *   gsr_reset_envs         <- RigidSolver.set_state cache resets: collider.clear(envs_idx) (collider_decomp.py:328-366)
 *                             and constraint_solver.clear/reset(envs_idx) (constraint_solver_decomp.py:110-145).
 *   gsr_get_errno          <- RigidSolver.check_errno (rigid_solver_decomp.py:935-956).
 *   gsr_step_fused         <- no counterpart (batching form: one launch per substep over a heterogeneous batch).
 *
 * Conventions: every pointer in gsr_model / gsr_state is device memory owned by the caller (PyTorch tensors). The
 * library never allocates or frees on the step path. Per-env arrays are env-minor: element i of env b lives at
 * [i * n_envs + b]; a vector field of k components stores component c of element i at [(i * k + c) * n_envs + b].
 * Calls are stream-ordered and never synchronise the host. Return 0 on success, a negative GSR_E* code otherwise.
 */
#ifndef GENESIS_RIGID_H
#define GENESIS_RIGID_H

#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GSR_ABI_VERSION 1

#define GSR_OK 0
#define GSR_E_ARG -1      /* null pointer, bad count */
#define GSR_E_SHAPE -2    /* model sizes exceed compiled-in limits */
#define GSR_E_LAUNCH -3   /* HIP launch error */

/* errno bits, same meaning as the reference's errno word (rigid_solver_decomp.py:941-956) */
#define GSR_ERR_BROAD_OVERFLOW 1u
#define GSR_ERR_CONTACT_OVERFLOW 2u
#define GSR_ERR_NAN_QACC 4u
#define GSR_ERR_NAN_STATE 8u
#define GSR_ERR_DEEP_PENETRATION 0x100u /* informational: the GJK/EPA fallback ran for a deep penetration */

/* reset flags for gsr_reset_envs */
#define GSR_RESET_COLLIDER 1u  /* contact-normal cache + first_time (sort warm start) */
#define GSR_RESET_WARMSTART 2u /* qacc_ws = 0, is_warmstart = 0 */

/* Static tables of one scene type (shared by all its envs). Index arrays int32, real arrays float32. */
typedef struct gsr_model {
  int32_t n_links, n_joints, n_dofs, n_qs, n_geoms, n_entities, n_equalities, n_verts, n_pairs;
  int32_t max_contacts, max_broad, max_rows, row_width, mass_blk_size, max_entity_dofs;
  int32_t iterations, ls_iterations, n_contacts_per_pair;
  int32_t enable_collision, enable_joint_limit, enable_multi_contact;
  int32_t solver_variant;  /* 0: one lane per env (rows env-minor), 1: one wavefront per env (rows env-major) */
  int32_t collide_variant; /* 0: one lane per env, 1: one wavefront per env (pairs across lanes) */
  int32_t dyn_variant;     /* step_1 / step_2: 0 one lane per env, 1 one wavefront per env (entities across lanes) */
  int32_t batch_dofs;      /* 1: dof_kp, dof_kv, dof_force_range, dof_stiffness, dof_damping, dof_frictionloss,
                              dof_limit, dof_armature and dof_invweight hold one value per env, env-minor like the state
                              (component c of dof d of env b at [(d * k + c) * n_envs + b]); 0: one value per dof
                              shared by every env (RigidOptions.batch_dofs_info) */
  int32_t batch_links;     /* 1: link_inertial_mass and link_invweight hold one value per env, env-minor
                              (RigidOptions.batch_links_info) */
  int32_t batch_joints;    /* 1: joint_sol_params holds one value per env, env-minor (RigidOptions.batch_joints_info) */
  int32_t batch_eqs;       /* 1: eq_sol_params holds one value per env, env-minor (the reference's equality solver
                              parameters are always per env: RigidSolver.set_sol_params(eqs_idx, envs_idx)) */
  int32_t wave_h_floats;   /* wavefront solver: LDS floats reserved for one env's Hessian (a larger pattern keeps it
                              in global memory); sets the solver's LDS per env, so how many envs share a CU */
  int32_t max_dyn_welds;   /* capacity of the per-env weld list (gsr_state.eq_dyn_*), RigidOptions.max_dynamic_constraints */
  float dt, tolerance, ls_tolerance, eps, meaninertia, mc_perturbation, mc_tolerance, mpr_to_gjk;
  float gravity[4];
  /* links */
  const int32_t *link_parent, *link_root, *link_q_start, *link_q_end, *link_dof_start, *link_dof_end;
  const int32_t *link_joint_start, *link_joint_end, *link_is_fixed;
  const float *link_pos, *link_quat, *link_inertial_pos, *link_inertial_quat, *link_inertial_i, *link_inertial_mass;
  const float *link_invweight;
  /* joints */
  const int32_t *joint_type, *joint_q_start, *joint_dof_start, *joint_dof_end;
  const float *joint_pos, *joint_sol_params;
  /* dofs */
  const float *dof_motion_ang, *dof_motion_vel, *dof_limit, *dof_force_range, *dof_kp, *dof_kv, *dof_damping;
  const float *dof_armature, *dof_stiffness, *dof_frictionloss, *dof_invweight;
  const int32_t *dof_entity;
  const float *qpos0;
  /* entities */
  const int32_t *ent_link_start, *ent_link_end, *ent_dof_start, *ent_dof_end, *ent_mass_offset;
  const float *ent_gravity_compensation;
  /* geoms */
  const int32_t *geom_type, *geom_link, *geom_vert_start, *geom_vert_end;
  const float *geom_pos, *geom_quat, *geom_data, *geom_friction, *geom_sol_params, *geom_center, *geom_init_aabb;
  const float *verts;
  /* mesh support field: per mesh geom, 180 x 180 direction cells (theta-major) from support_cell_start[g], each the
     support vertex's position (3 floats) and its index within the geom */
  const int32_t *support_cell_start, *support_vid;
  const float *support_v;
  /* equalities */
  const int32_t *eq_type, *eq_obj1, *eq_obj2;
  const float *eq_data, *eq_sol_params;
  /* collision pair table (G*G, -1 = never collide) and CRBA ancestor mask (D*D) */
  const int32_t *pair_idx;
  const float *mass_parent_mask;
  /* per-env mean inertia (n_envs; the solver's tolerance scale, rigid_global_info.meaninertia[i_b]) when per-env
     armature or link masses make it differ between envs, else null and meaninertia above holds for every env */
  const float *env_meaninertia;
} gsr_model;

/* Per-env state and scratch of n_envs envs of one scene type (all env-minor). */
typedef struct gsr_state {
  int32_t n_envs, pad0;
  /* persistent state and controls */
  float *qpos, *vel, *acc, *acc_smooth, *qacc_ws;
  int32_t *is_warmstart, *ctrl_mode;
  float *ctrl_pos, *ctrl_vel, *ctrl_force;
  float *dofs_pos, *links_pos, *links_quat, *geoms_pos, *geoms_quat, *aabb_min, *aabb_max;
  float *cache_normal, *sort_value;
  int32_t *sort_ig, *first_time;
  uint32_t *errno_env;
  /* per-env model shifts: links_state.mass_shift (L), links_state.i_pos_shift (L x 3), geoms_state.friction_ratio (G) */
  float *mass_shift, *i_pos_shift, *friction_ratio;
  /* welds registered per env after the build (ConstraintSolver.add_weld_constraint), after the model's own equalities:
     count (1), (link1, link2) pairs (2 x max_dyn_welds), data in the model's eq_data layout (11 x max_dyn_welds) */
  int32_t *eq_dyn_n, *eq_dyn_obj;
  float *eq_dyn_data;
  /* kinematics / dynamics intermediates */
  float *cdof_ang, *cdof_vel, *cdofd_ang, *cdofd_vel, *root_com, *cd_vel, *cd_ang, *cinr_inertial, *cinr_pos;
  float *cinr_mass, *links_iquat, *mass_blk, *mass_L_blk, *mass_D_inv, *qf_smooth, *force;
  float *xanchor, *xaxis, *crb_inertial, *crb_pos, *crb_mass, *f_ang, *f_vel, *cdd_vel, *cdd_ang, *cfrc_vel;
  float *cfrc_ang, *qf_applied, *qf_passive, *qf_bias;
  /* collision */
  int32_t *n_broad, *broad_pairs, *n_contacts, *contact_geoms;
  float *contact_pos, *contact_normal, *contact_pen, *contact_friction, *contact_force;
  int32_t *active_buffer;
  /* constraint rows (sparse, row_width nonzeros per row) */
  int32_t *n_con; /* 3 per env: total, equality, frictionloss */
  float *jac_val;
  int32_t *jac_idx, *jac_nnz;
  float *efc_diag, *efc_aref, *efc_D, *efc_fl, *efc_force, *efc_Jaref, *efc_jv, *efc_quad;
  int32_t *efc_active;
  /* Newton solver */
  float *qacc, *Ma, *grad, *Mgrad, *search, *mv, *qfrc_constraint, *nt_vec, *nt_H;
  int32_t *solver_iters;
  int32_t *ent_row_ptr, *ent_row_list; /* per-entity ordered constraint-row lists (wavefront solver) */
  int32_t *efc_ent;                    /* entity pair of each row (wavefront solver) */
  int32_t *ent_row_run;                /* beside each ent_row_list entry: where the entity's entries sit in that row
                                          (first dof | position << 16 | count << 24), -1 when not one run of
                                          consecutive dofs (wavefront solver, dof-lane row walks) */
  float *gjk_ws;                       /* GJK/EPA polytope workspace, gsr_gjk_workspace_words() per env, env-major */
  /* pairs the narrowphase handed to the GJK/EPA pass of this substep: count, then (ga << 16 | gb, contact index the
     pair's contacts are inserted at) per pair, max_broad pairs per env */
  int32_t *gjk_defer_n, *gjk_defer;
  /* integration scratch */
  float *qpos_next, *vel_next;
} gsr_state;

int gsr_abi_version(void);
/* sizeof(gsr_model), sizeof(gsr_state), sizeof(gsr_group), sizeof(gsr_batch_info), for binding checks */
int gsr_model_size(void);
int gsr_state_size(void);
int gsr_group_size(void);
int gsr_batch_info_size(void);
/* 32-bit words of GJK/EPA workspace one env needs (size of gsr_state.gjk_ws per env) */
int gsr_gjk_workspace_words(void);

int gsr_step(const gsr_model* model, const gsr_state* state, int32_t n_substeps, hipStream_t stream);
/* ---- Allocation-free step -------------------------------------------------------------------------------------
 * No counterpart in the reference. For callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc): gsr_step with the per-substep scratch -- constraint rows
 * (jac_val, jac_idx, jac_nnz, efc_diag, efc_aref, efc_D, efc_fl, efc_Jaref, efc_jv, efc_quad, efc_active), the
 * Newton solver's vectors and Hessian (qacc, Ma, grad, Mgrad, search, mv, qfrc_constraint, nt_vec, nt_H,
 * solver_iters), the per-entity row lists (ent_row_ptr, ent_row_list, efc_ent, ent_row_run), the GJK/EPA workspace and deferral
 * list (gjk_ws, gjk_defer_n, gjk_defer) and the integration scratch (qpos_next, vel_next) -- taken from a static
 * device array of GSR_NOALLOC_SCRATCH_BYTES per GPU instead of the state (those state pointers are ignored and may
 * be NULL). The step allocates and frees nothing; results are those of gsr_step. Every call on one GPU shares that
 * GPU's array: calls that overlap in time on one GPU (two streams, two host threads) race. Returns GSR_E_SHAPE when
 * the scratch of n_envs envs exceeds GSR_NOALLOC_SCRATCH_BYTES (gsr_noalloc_scratch_bytes gives the need). */
#define GSR_NOALLOC_SCRATCH_BYTES (256u << 20)
int64_t gsr_noalloc_scratch_bytes(const gsr_model* model, int32_t n_envs);
int gsr_step_noalloc(const gsr_model* model, const gsr_state* state, int32_t n_substeps, hipStream_t stream);

int gsr_forward_kinematics(const gsr_model* model, const gsr_state* state, const int32_t* envs_idx, int32_t n_idx,
                           hipStream_t stream);
int gsr_reset_envs(const gsr_model* model, const gsr_state* state, const int32_t* envs_idx, int32_t n_idx,
                   uint32_t what, hipStream_t stream);
int gsr_get_errno(const gsr_state* state, uint32_t* dev_out, hipStream_t stream);

/* ---- Readback of the last step's contacts and constraint forces ---------------------------------------------- */
/* Replaces collider_kernel_get_contacts (collider_decomp.py:687-724) behind Collider.get_contacts (:463-565).
 * Output row r holds iout[4r..] = {link_a, link_b, geom_a, geom_b} and fout[10r..] = {penetration, position[3],
 * normal[3], force[3]}. Padded (dev_starts == NULL): contact i of env b at row b * n_max + i, rows past the env's
 * count set to -1 / 0; out_rows must be at least n_envs * n_max. Ragged: row dev_starts[b] + i, dev_starts being the
 * exclusive prefix sum of the per-env counts. Rows at or past out_rows are never written. */
int gsr_get_contacts(const gsr_model* model, const gsr_state* state, const int32_t* dev_starts, int32_t n_max,
                     int64_t out_rows, int32_t* dev_iout, float* dev_fout, hipStream_t stream);
/* Replaces links_state.contact_force of func_update_contact_force (constraint_solver_decomp.py:1350-1379), read by
 * RigidEntity.get_links_net_contact_force (rigid_entity.py:2838-2850): dev_out[k][l][3] = net contact force on link l
 * of env dev_envs_idx[k] (of env k when dev_envs_idx is NULL, n_idx then ignored). */
int gsr_links_contact_force(const gsr_model* model, const gsr_state* state, const int32_t* dev_envs_idx,
                            int32_t n_idx, float* dev_out, hipStream_t stream);
/* Replaces kernel_get_equality_constraints (constraint_solver_decomp.py:2510-2558) behind
 * ConstraintSolver.get_equality_constraints (:235-287): row b * (n_equalities + max_dyn_welds) + i holds
 * iout[3r..] = {type, obj1, obj2} and fout[6r..] = the equality's row forces (1 JOINT, 3 CONNECT, 6 WELD, the rest
 * zero): the model's equalities, then the env's own welds; rows past the env's count hold type -1 and zeros. */
int gsr_get_equalities(const gsr_model* model, const gsr_state* state, int32_t* dev_iout, float* dev_fout,
                       hipStream_t stream);

/* One kernel of the plain substep (0 = step_1 dynamics, 1 = collision + constraint rows, 2 = Newton solve,
 * 3 = integration + end-of-step kinematics); gsr_step launches exactly these four in order. No counterpart in the
 * reference: used to time each kernel on the caller's stream. */
int gsr_launch_kernel(const gsr_model* model, const gsr_state* state, int32_t kernel, hipStream_t stream);

/* Stage runner for parity tests: 1 = kinematics, 2 = + smooth dynamics, 3 = + collision detection, 0 = full step. */
int gsr_run_stage(const gsr_model* model, const gsr_state* state, int32_t stage, hipStream_t stream);

/* ---- Batching form (no counterpart in the reference, whose batches are homogeneous) ----------------------------
 * A heterogeneous batch is a list of groups: each group is one scene type (model) and a contiguous range of envs of
 * one state. One persistent launch per substep steps every group: workgroups (one wavefront each) pull tiles of
 * `envs_per_tile` envs from a device work queue, heaviest group first, and run the whole substep (step_1, collision,
 * constraint rows, Newton solve, step_2) for their tile. The per-env arithmetic is the plain path's, so a result
 * does not depend on what else is in the batch. */
#define GSR_MAX_GROUPS 4096 /* env ranges per batch */
#define GSR_MAX_TYPES 4     /* distinct (model, state) pairs per batch: their descriptors travel as kernel arguments */

typedef struct gsr_group {
  gsr_model model;         /* by value; model.solver_variant / collide_variant select the per-env code paths */
  gsr_state state;         /* by value */
  int32_t env_begin;       /* first env of the group inside `state` */
  int32_t n_envs;          /* envs of the group: [env_begin, env_begin + n_envs) */
  int32_t envs_per_tile;   /* 1..64; the per-lane stages use one lane per env of the tile */
  int32_t tile_begin;      /* written by gsr_batch_upload: first tile index of the group */
  int32_t wave_envs_per_tile; /* 0..64, phased launch only: envs per tile of the GJK/EPA pass and the solve, whose
                                 wavefront variants take a tile's envs one after the other (0 = envs_per_tile) */
  int32_t pad;
} gsr_group;

typedef struct gsr_batch_info {
  int32_t n_groups, n_tiles; /* groups and tiles in the batch */
  int32_t grid;              /* workgroups per launch */
  int32_t lds_bytes;         /* dynamic LDS per workgroup (largest wavefront-variant group) */
  int32_t n_types;           /* distinct (model, state) pairs, in first-appearance order */
  int32_t persistent;        /* input: 1 = CUs x resident workgroups pull tiles from a queue, 0 = one per tile */
  int32_t waves_per_simd;    /* input: register budget of the fused kernel, 2 (256 VGPRs) or 4 (128 VGPRs) */
  int32_t gjk_inline;        /* input: 1 = GJK/EPA inside the fused kernel, 0 = deferred to a per-type resume launch */
  int32_t phased;            /* input: 1 = one launch per substep stage over every tile (persistent, waves_per_simd
                                and gjk_inline are then ignored), 0 = one fused launch per substep */
  int32_t n_wave_tiles;      /* phased: tiles of the GJK/EPA pass and the solve (wave_envs_per_tile) */
  gsr_model models[GSR_MAX_TYPES];
  gsr_state states[GSR_MAX_TYPES];
} gsr_batch_info;

/* Device bytes the batch table of n_groups needs (the caller allocates it, 256-byte aligned, zero-filled once). */
int64_t gsr_batch_table_bytes(int32_t n_groups);
/* Validate groups, assign tiles and copy the table to `dev_table` (synchronously); fills `info` (whose `persistent`
 * and `waves_per_simd` are read as inputs), which the launch functions need. Groups are stepped in the given order,
 * so list the most expensive first. Groups over one state (same `state.qpos`) must use the same solver, collide
 * and dyn variants and disjoint env ranges (GSR_E_ARG otherwise). No launch that uses `dev_table` may be in
 * flight. */
int gsr_batch_upload(const gsr_group* groups, int32_t n_groups, void* dev_table, int64_t table_bytes,
                     gsr_batch_info* info);
/* n_substeps fused launches of the uploaded batch on `stream`. */
int gsr_step_fused(const void* dev_table, const gsr_batch_info* info, int32_t n_substeps, hipStream_t stream);
/* Kernels one fused substep launches, in order: the phased form's five stage kernels, or the single fused kernel
 * followed (GJK/EPA deferred) by one resume kernel per scene type. No counterpart in the reference: used to time
 * each kernel of the fused substep on the caller's stream. */
int gsr_fused_kernel_count(const gsr_batch_info* info);
/* Launch kernel k (0 <= k < gsr_fused_kernel_count) of one fused substep; gsr_step_fused launches exactly these. */
int gsr_launch_fused_kernel(const void* dev_table, const gsr_batch_info* info, int32_t k, hipStream_t stream);
/* Capture n_substeps fused launches into a hipGraph (stream must not be the legacy default stream) and return the
 * instantiated executable; launch it with gsr_graph_launch, free it with gsr_graph_destroy. */
int gsr_graph_capture(const void* dev_table, const gsr_batch_info* info, int32_t n_substeps, hipStream_t stream,
                      void** graph_exec);
int gsr_graph_launch(void* graph_exec, hipStream_t stream);
int gsr_graph_destroy(void* graph_exec);

/* ---- Serving form: device-side request scheduler (no counterpart in the reference) ------------------------------
 * Resident pools of env slots (one per scene type, cut into tiles of envs_per_tile envs) serve rollout requests
 * (scene type, n_envs, horizon, seed). A tick (gsr_serve_tick, or its hipGraph from gsr_serve_graph_capture) admits
 * arrived requests in arrival order while request slots, free tiles and the memory and tick budgets allow, resets
 * the admitted envs, applies each active request's controls, runs one fused substep over the active tiles, advances
 * every active request by one step and retires those at their horizon. Host and device exchange arrivals,
 * completions and per-tick GPU timestamps through a gsr_serve_host block in coherent pinned host memory, with no
 * host synchronisation: the host appends arrivals (entry first, then arrivals_head) and reads completions and
 * ticks_done; the device reads arrivals and writes the rest. A request's results do not depend on its slots, the
 * tile size or the other requests (same per-env device code; controls from a hash of request-local indices). */
#define GSR_SERVE_MAX_REQUESTS 1024 /* requests admitted at once */
#define GSR_SERVE_RING 4096         /* entries of the arrival, completion and tick-timestamp rings */

typedef struct gsr_serve_pool {
  gsr_model model;          /* by value, with the execution variants the pool runs */
  gsr_state state;          /* by value: the pool's envs are envs [0, envs_per_tile * n_tiles) of this state */
  int32_t envs_per_tile, n_tiles;
  int32_t ctrl_kind;        /* 0 none; 1 position targets U[ctrl_lo, ctrl_hi] per controlled dof, redrawn when the
                               request's step is a multiple of ctrl_period; 2 position targets ctrl_base +
                               U[-ctrl_amp, ctrl_amp] every step */
  int32_t ctrl_period, ctrl_dof_begin, ctrl_dof_end;
  float ctrl_amp, pad;
  const float *ctrl_lo, *ctrl_hi, *ctrl_base; /* device, one per controlled dof */
  const float *init_qpos;   /* device, env-minor [n_qs][envs_per_tile * n_tiles]: initial qpos of request-local env i */
  double bytes_per_env;     /* memory-budget bytes one admitted env holds */
  double cost_per_env;      /* tick-budget units (e.g. predicted ms) one admitted env adds to a tick */
} gsr_serve_pool;

typedef struct gsr_serve_request {
  int32_t rid, pool, n_envs, horizon;
  uint32_t seed, pad0, pad1, pad2;
} gsr_serve_request;

typedef struct gsr_serve_done {
  int32_t rid, tick, failed, slot; /* failed: an env of the request hit an errno bit the reference raises on */
} gsr_serve_done;

typedef struct gsr_serve_host {
  int32_t arrivals_head; /* host: requests appended so far */
  int32_t arrivals_tail; /* device: requests admitted so far */
  int32_t done_count;    /* device: completion records written so far */
  int32_t ticks_done;    /* device: ticks finished */
  int64_t env_steps;     /* device: env-steps stepped so far */
  gsr_serve_request arrivals[GSR_SERVE_RING];
  gsr_serve_done done[GSR_SERVE_RING];
  uint64_t tick_t0[GSR_SERVE_RING], tick_t1[GSR_SERVE_RING]; /* GPU wall clock at tick start / end */
  int32_t tick_envs[GSR_SERVE_RING];                                   /* envs stepped by each tick */
} gsr_serve_host;

/* Coherent pinned host memory for one gsr_serve_host (zero-filled); free with gsr_serve_host_free. */
int gsr_serve_host_alloc(gsr_serve_host** out);
int gsr_serve_host_free(gsr_serve_host* host);
/* Device bytes of the scheduler state and of the tile table for n_tiles tiles over all pools. */
int64_t gsr_serve_state_bytes(int32_t n_tiles);
int64_t gsr_serve_table_bytes(int32_t n_tiles);
/* GPU wall-clock rate of the tick timestamps, in kHz. */
int gsr_serve_clock_khz(void);
/* Set up the scheduler over n_pools pools (<= GSR_MAX_TYPES, heaviest first): memory and tick budgets, all tiles
 * free. `info` is read for the launch form (persistent, waves_per_simd, gjk_inline, phased) and filled like
 * gsr_batch_upload fills it. Synchronises the stream. */
int gsr_serve_init(const gsr_serve_pool* pools, int32_t n_pools, double mem_budget, double tick_budget,
                   void* dev_state, int64_t state_bytes, void* dev_table, int64_t table_bytes, gsr_batch_info* info,
                   hipStream_t stream);
/* One tick on `stream` (no host synchronisation). */
int gsr_serve_tick(void* dev_state, void* dev_table, const gsr_batch_info* info, gsr_serve_host* host,
                   hipStream_t stream);
/* Capture one tick into a hipGraph (launch with gsr_graph_launch, free with gsr_graph_destroy). */
int gsr_serve_graph_capture(void* dev_state, void* dev_table, const gsr_batch_info* info, gsr_serve_host* host,
                            hipStream_t stream, void** graph_exec);

#ifdef __cplusplus
}
#endif
#endif /* GENESIS_RIGID_H */
