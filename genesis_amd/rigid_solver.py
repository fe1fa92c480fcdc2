"""Host side of the MI355X rigid step: model tables and per-env state in PyTorch tensors, steps via the C-ABI.

Mirrors the parts of the reference's RigidSolver (genesis/engine/solvers/rigid/rigid_solver_decomp.py) that sit on
the batched step path: build (:190-352), substep (:865-933), check_errno (:935-956), control_dofs_* (:2123-2201),
get_dofs_* (:2366-2376), set_qpos (:1801-1857), set_dofs_velocity, set_state/get_state (:1442-1512). State tensors
are stored env-minor, shape (n, B), and exposed as zero-copy (B, n) views like ti_to_torch(transpose=True)
(genesis/utils/misc.py:615-659).
"""
from __future__ import annotations

import ctypes
from typing import Optional

import numpy as np
import torch

from . import _lib
from .model import CTRL_FORCE, CTRL_POSITION, CTRL_VELOCITY, RigidModel

RATE_CHECK_ERRNO = 10  # genesis/engine/simulator.py:47

GSR_RESET_COLLIDER = 1
GSR_RESET_WARMSTART = 2

ERRNO_MESSAGES = [
    (1, "Exceeding max number of broad phase candidate contact pairs. Increase 'multiplier_collision_broad_phase'."),
    (2, "Exceeding max number of contact pairs. Increase 'max_collision_pairs'."),
    (4, "Invalid constraint forces causing 'nan'. Please decrease Rigid simulation timestep."),
    (8, "Invalid accelerations causing 'nan'. Please decrease Rigid simulation timestep."),
]


class GenesisException(Exception):
    pass


MODEL_TABLES_I = [
    "link_parent", "link_root", "link_q_start", "link_q_end", "link_dof_start", "link_dof_end", "link_joint_start",
    "link_joint_end", "link_is_fixed", "joint_type", "joint_q_start", "joint_dof_start", "joint_dof_end", "dof_entity",
    "ent_link_start", "ent_link_end", "ent_dof_start", "ent_dof_end", "ent_mass_offset", "geom_type", "geom_link",
    "geom_vert_start", "geom_vert_end", "eq_type", "eq_obj1", "eq_obj2", "pair_idx", "support_cell_start",
    "support_vid",
]
MODEL_TABLES_F = [
    "link_pos", "link_quat", "link_inertial_pos", "link_inertial_quat", "link_inertial_i", "link_inertial_mass",
    "link_invweight", "joint_pos", "joint_sol_params", "dof_motion_ang", "dof_motion_vel", "dof_limit",
    "dof_force_range", "dof_kp", "dof_kv", "dof_damping", "dof_armature", "dof_stiffness", "dof_frictionloss",
    "dof_invweight", "qpos0", "ent_gravity_compensation", "geom_pos", "geom_quat", "geom_data", "geom_friction",
    "geom_sol_params", "geom_center", "geom_init_aabb", "verts", "eq_data", "eq_sol_params", "mass_parent_mask",
    "support_v",
]

INT_STATE = {"is_warmstart", "ctrl_mode", "sort_ig", "first_time", "errno_env", "n_broad", "broad_pairs",
             "n_contacts", "contact_geoms", "active_buffer", "n_con", "jac_idx", "jac_nnz", "efc_active",
             "solver_iters", "ent_row_ptr", "ent_row_list", "efc_ent", "ent_row_run", "gjk_defer_n", "gjk_defer",
             "eq_dyn_n", "eq_dyn_obj"}

# Entity counts above this cannot use the wavefront solver's pattern bitsets.
WAVE_MAX_ENTITIES = 128


# Up to this many envs, one wavefront per env still fills the chip (8 waves per CU at 2,048 envs) and cuts the
# per-env latency that a few envs' step is made of; beyond it, the dynamics and collision stages of a small
# articulation run one lane per env.
WAVE_MAX_SMALL_ENVS = 2048


def default_solver_variant(M: RigidModel, n_envs: int = None) -> int:
    """1 = one wavefront per env (every model the pattern bitsets cover: the Newton solve of a many-contact env is
    latency-bound, and the wavefront solver is the faster one at every env count measured), 0 = one lane per env."""
    return 0 if M.n_entities > WAVE_MAX_ENTITIES else 1


def default_stage_variant(M: RigidModel, n_envs: int = None) -> int:
    """Variant of the dynamics and collision stages: 1 = one wavefront per env (large systems, or few envs), 0 = one
    lane per env (many envs of a small articulation)."""
    return 1 if (M.n_dofs >= 48 or (n_envs is not None and n_envs <= WAVE_MAX_SMALL_ENVS)) else 0


def default_collide_variant(M: RigidModel, n_envs: int = None) -> int:
    """As default_stage_variant, except that mesh geoms take the wavefront collision (support-field lookups diverge
    across a lane-per-env wavefront: the Panda's meshes collide 10x slower one lane per env at 5,492 envs)."""
    from .model import GEOM_MESH
    return 1 if (np.asarray(M.geom_type) == GEOM_MESH).any() else default_stage_variant(M, n_envs)


def state_sizes(M: RigidModel) -> dict:
    """Per-env element count of every gsr_state array."""
    L, J, D, Q, G = M.n_links, M.n_joints, M.n_dofs, M.n_qs, M.n_geoms
    C, R, W = max(M.max_contacts, 1), max(M.max_rows, 1), M.row_width
    P = max(M.n_pairs, 1)
    sz = dict(
        qpos=Q, vel=D, acc=D, acc_smooth=D, qacc_ws=D, is_warmstart=1, ctrl_mode=D, ctrl_pos=D, ctrl_vel=D,
        ctrl_force=D, dofs_pos=D, links_pos=3 * L, links_quat=4 * L, geoms_pos=3 * G, geoms_quat=4 * G,
        aabb_min=3 * G, aabb_max=3 * G, cache_normal=3 * P, sort_value=2 * G, sort_ig=2 * G, first_time=1,
        errno_env=1, mass_shift=L, i_pos_shift=3 * L, friction_ratio=G,
        eq_dyn_n=1, eq_dyn_obj=2 * M.max_dyn_welds, eq_dyn_data=11 * M.max_dyn_welds,
        cdof_ang=3 * D, cdof_vel=3 * D, cdofd_ang=3 * D, cdofd_vel=3 * D, root_com=3 * L, cd_vel=3 * L,
        cd_ang=3 * L, cinr_inertial=9 * L, cinr_pos=3 * L, cinr_mass=L, links_iquat=4 * L,
        mass_blk=M.mass_blk_size, mass_L_blk=M.mass_blk_size, mass_D_inv=D, qf_smooth=D, force=D,
        xanchor=3 * J, xaxis=3 * J, crb_inertial=9 * L, crb_pos=3 * L, crb_mass=L, f_ang=3 * D, f_vel=3 * D,
        cdd_vel=3 * L, cdd_ang=3 * L, cfrc_vel=3 * L, cfrc_ang=3 * L, qf_applied=D, qf_passive=D, qf_bias=D,
        n_broad=1, broad_pairs=max(M.max_broad, 1), n_contacts=1, contact_geoms=2 * C, contact_pos=3 * C,
        contact_normal=3 * C, contact_pen=C, contact_friction=C, contact_force=3 * C, active_buffer=G,
        n_con=3, jac_val=R * W, jac_idx=R * W, jac_nnz=R, efc_diag=R, efc_aref=R, efc_D=R, efc_fl=R,
        efc_force=R, efc_Jaref=R, efc_jv=R, efc_quad=3 * R, efc_active=R,
        qacc=D, Ma=D, grad=D, Mgrad=D, search=D, mv=D, qfrc_constraint=D, nt_vec=D, nt_H=D * (D + 1) // 2,
        solver_iters=1, qpos_next=Q, vel_next=D, ent_row_ptr=M.n_entities + 1, ent_row_list=2 * R, efc_ent=R,
        ent_row_run=2 * R, gjk_ws=_lib.load().gsr_gjk_workspace_words(), gjk_defer_n=1, gjk_defer=2 * max(M.max_broad, 1),
    )
    missing = set(_lib.STATE_POINTERS) - set(sz)
    assert not missing, f"state arrays without a size rule: {missing}"
    return {k: max(int(v), 1) for k, v in sz.items()}


def finalize_model(M: RigidModel):
    """Derived tables the library needs beyond the compiler output."""
    ns = (M.ent_dof_end - M.ent_dof_start).astype(np.int64)
    off = np.concatenate([[0], np.cumsum(ns * ns)[:-1]]) if len(ns) else np.zeros(0)
    M.ent_mass_offset = off.astype(np.int32)
    M.mass_blk_size = int((ns * ns).sum())


# LDS floats the wavefront solver may reserve for one env's Hessian: enough for every 64-box pile Hessian measured
WAVE_H_FLOATS = 5120


def wave_h_floats_default(M: RigidModel) -> int:
    """Hessian LDS budget of one env: 4x the entity-diagonal blocks (room for the contact-coupled off-diagonal
    blocks), in [1024, WAVE_H_FLOATS]. A Hessian that does not fit is kept in global memory instead (same bits,
    slower); a smaller budget lets more envs share a CU (a 20-box pile: 19 KB of LDS per env instead of 27 KB)."""
    n = (np.asarray(M.ent_dof_end) - np.asarray(M.ent_dof_start)).astype(np.int64)
    diag = int((n * (n + 1) // 2).sum())
    return int(min(WAVE_H_FLOATS, max(1024, -(-4 * diag // 256) * 256)))

# per-dof / per-link / per-joint tables that batch_dofs_info / batch_links_info / batch_joints_info make per-env
# (env-minor, like the state; options/solvers.py:373-375); equality solver parameters are always per env
BATCHED_DOF_TABLES = ("dof_kp", "dof_kv", "dof_force_range", "dof_stiffness", "dof_damping", "dof_frictionloss",
                      "dof_limit", "dof_armature", "dof_invweight")
BATCHED_LINK_TABLES = ("link_inertial_mass", "link_invweight")
BATCHED_JOINT_TABLES = ("joint_sol_params",)
BATCHED_EQ_TABLES = ("eq_sol_params",)


class DeviceModel:
    """Static tables of one scene type uploaded once (torch-owned device memory) + the gsr_model descriptor."""

    def __init__(self, M: RigidModel, device, solver_variant=None, collide_variant=None, dyn_variant=None,
                 n_envs=None, wave_h_floats=None):
        finalize_model(M)
        self.M = M
        self.device = torch.device(device)
        self.tensors = {}
        for k in MODEL_TABLES_I:
            a = np.ascontiguousarray(np.asarray(getattr(M, k)), dtype=np.int32).ravel()
            self.tensors[k] = torch.from_numpy(a if a.size else np.zeros(1, np.int32)).to(self.device)
        o = M.options
        self.batch_dofs = bool(getattr(o, "batch_dofs_info", False))
        self.batch_links = bool(getattr(o, "batch_links_info", False))
        self.batch_joints = bool(getattr(o, "batch_joints_info", False))
        self.B = max(1, int(n_envs or 0))
        for k in MODEL_TABLES_F:
            a = np.ascontiguousarray(np.asarray(getattr(M, k)), dtype=np.float32).ravel()
            if self.batched(k):
                a = np.ascontiguousarray(np.repeat(a[:, None], self.B, axis=1)).ravel()  # (n, B) env-minor
            self.tensors[k] = torch.from_numpy(a if a.size else np.zeros(1, np.float32)).to(self.device)
        # per-env armature or link masses give each env its own mean inertia (rigid_global_info.meaninertia[i_b])
        self.env_meaninertia = None
        if self.batch_dofs or self.batch_links:
            self.env_meaninertia = torch.full((self.B,), float(M.meaninertia), dtype=torch.float32, device=self.device)
        d = _lib.GsrModel()
        d.n_links, d.n_joints, d.n_dofs, d.n_qs = M.n_links, M.n_joints, M.n_dofs, M.n_qs
        d.n_geoms, d.n_entities, d.n_equalities, d.n_verts, d.n_pairs = (M.n_geoms, M.n_entities, M.n_equalities,
                                                                         M.n_verts, M.n_pairs)
        d.max_contacts, d.max_broad, d.max_rows, d.row_width = M.max_contacts, M.max_broad, M.max_rows, M.row_width
        d.max_dyn_welds = M.max_dyn_welds
        d.mass_blk_size, d.max_entity_dofs = M.mass_blk_size, M.max_entity_dofs
        d.iterations, d.ls_iterations, d.n_contacts_per_pair = o.iterations, o.ls_iterations, o.n_contacts_per_pair
        d.enable_collision, d.enable_joint_limit = int(o.enable_collision), int(o.enable_joint_limit)
        d.enable_multi_contact = int(o.enable_multi_contact)
        # an explicit solver variant is also the default of the other stages; otherwise each takes its own default
        explicit = solver_variant is not None
        d.solver_variant = int(solver_variant) if explicit else default_solver_variant(M, n_envs)
        if d.solver_variant == 1 and M.n_entities > WAVE_MAX_ENTITIES:
            raise ValueError(f"wavefront solver supports at most {WAVE_MAX_ENTITIES} entities")
        if collide_variant is None:
            collide_variant = d.solver_variant if explicit else default_collide_variant(M, n_envs)
        if dyn_variant is None:
            dyn_variant = d.solver_variant if explicit else default_stage_variant(M, n_envs)
        d.collide_variant, d.dyn_variant = int(collide_variant), int(dyn_variant)
        d.batch_dofs, d.batch_links, d.batch_joints = int(self.batch_dofs), int(self.batch_links), int(self.batch_joints)
        d.batch_eqs = 1
        d.wave_h_floats = wave_h_floats_default(M) if wave_h_floats is None else int(wave_h_floats)
        d.dt, d.tolerance, d.ls_tolerance = o.dt, o.tolerance, o.ls_tolerance
        d.eps = float(np.finfo(np.float32).eps)
        d.meaninertia = float(M.meaninertia)
        d.mc_perturbation, d.mc_tolerance, d.mpr_to_gjk = 1e-2, 1e-2, 0.25
        for i in range(3):
            d.gravity[i] = float(M.gravity[i])
        for k in _lib.MODEL_POINTERS:
            if k == "env_meaninertia":
                d.env_meaninertia = None if self.env_meaninertia is None else self.env_meaninertia.data_ptr()
            else:
                setattr(d, k, self.tensors[k].data_ptr())
        self.desc = d

    def batched(self, name) -> bool:
        """Whether the device table `name` holds one value per env (env-minor)."""
        return (name in BATCHED_EQ_TABLES or (self.batch_dofs and name in BATCHED_DOF_TABLES)
                or (self.batch_links and name in BATCHED_LINK_TABLES)
                or (self.batch_joints and name in BATCHED_JOINT_TABLES))


class DeviceState:
    """Per-env state of n_envs envs of one model (torch tensors, env-minor) + the gsr_state descriptor."""

    def __init__(self, dm: DeviceModel, n_envs: int):
        self.dm = dm
        self.B = int(n_envs)
        M = dm.M
        self.sizes = state_sizes(M)
        self.t = {}
        for k, n in self.sizes.items():
            dt = torch.int32 if k in INT_STATE else torch.float32
            self.t[k] = torch.zeros((n, self.B), dtype=dt, device=dm.device)
        d = _lib.GsrState()
        d.n_envs = self.B
        for k in _lib.STATE_POINTERS:
            setattr(d, k, self.t[k].data_ptr())
        self.desc = d
        self.init_state()

    def init_state(self, envs=None):
        M = self.dm.M
        sl = slice(None) if envs is None else envs
        self.t["qpos"][:, sl] = torch.as_tensor(M.qpos0, device=self.dm.device)[:, None]
        for k in ("vel", "acc", "acc_smooth", "qacc_ws", "ctrl_pos", "ctrl_vel", "ctrl_force", "cache_normal",
                  "efc_active"):
            self.t[k][:, sl] = 0
        self.t["ctrl_mode"][:, sl] = CTRL_FORCE
        self.t["is_warmstart"][:, sl] = 0
        self.t["first_time"][:, sl] = 1
        self.t["errno_env"][:, sl] = 0
        self.t["mass_shift"][:, sl] = 0
        self.t["i_pos_shift"][:, sl] = 0
        self.t["friction_ratio"][:, sl] = 1
        self.t["eq_dyn_n"][:, sl] = 0

    def memory_bytes(self):
        return sum(t.numel() * t.element_size() for t in self.t.values())


# State arrays a substep writes before it reads them (constraint rows, solver work vectors, Hessian, collision
# scratch): a checkpoint leaves them out. Everything else (poses, velocities, controls, warm starts, the contact-normal
# and sort caches, kinematics caches, the last step's contacts and errno) is saved.
CHECKPOINT_SKIP = frozenset({
    "n_broad", "broad_pairs", "active_buffer", "jac_val", "jac_idx", "jac_nnz", "efc_diag", "efc_aref", "efc_D",
    "efc_fl", "efc_force", "efc_Jaref", "efc_jv", "efc_quad", "efc_active", "n_con", "Ma", "grad", "Mgrad", "search",
    "mv", "qfrc_constraint", "nt_vec", "nt_H", "solver_iters", "ent_row_ptr", "ent_row_list", "efc_ent", "ent_row_run",
    "gjk_ws",
    "gjk_defer_n", "gjk_defer", "qpos_next", "vel_next",
})


class RigidSolverState:
    """Dynamic state of a RigidSolver, env-major, mirroring the reference's RigidSolverState
    (genesis/engine/states/solvers.py:46-85). The COM/mass shifts and friction ratios are the neutral values: per-env
    model parameters are not built."""

    def __init__(self, B: int, model: RigidModel, device, s_global: int = 0):
        f = dict(dtype=torch.float32, device=device)
        L = model.n_links
        self.s_global = s_global
        self.qpos = torch.zeros((B, model.n_qs), **f)
        self.dofs_vel = torch.zeros((B, model.n_dofs), **f)
        self.dofs_acc = torch.zeros((B, model.n_dofs), **f)
        self.links_pos = torch.zeros((B, L, 3), **f)
        self.links_quat = torch.zeros((B, L, 4), **f)
        self.i_pos_shift = torch.zeros((B, L, 3), **f)
        self.mass_shift = torch.zeros((B, L), **f)
        self.friction_ratio = torch.ones((B, model.n_geoms), **f)

    def serializable(self):
        for k in ("qpos", "dofs_vel", "dofs_acc", "links_pos", "links_quat", "i_pos_shift", "mass_shift",
                  "friction_ratio"):
            setattr(self, k, getattr(self, k).detach())
        return self


class RigidSolver:
    """Batched rigid solver for one scene type on one device."""

    def __init__(self, model: RigidModel, n_envs: int, device="cuda", solver_variant=None, collide_variant=None,
                 dyn_variant=None, wave_h_floats=None):
        self.model = model
        self.n_envs = n_envs
        self._B = max(1, n_envs)
        self.device = torch.device(device)
        if self.device.type != "cuda":
            raise RuntimeError("the rigid step runs on a HIP device; there is no CPU execution path")
        self._lib = _lib.load()
        self.dm = DeviceModel(model, self.device, solver_variant, collide_variant, dyn_variant, n_envs=n_envs,
                              wave_h_floats=wave_h_floats)
        self.st = DeviceState(self.dm, self._B)
        self._errno_dev = torch.zeros(1, dtype=torch.int32, device=self.device)
        self._substep = 0
        self._mass_valid = False  # mass_blk holds the last substep's mass matrix (none before the first substep)
        self.forward_kinematics()

    # -- raw launches -------------------------------------------------------------------------------------------
    def _stream(self):
        return ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream) if self.device.type == "cuda" else None

    def forward_kinematics(self, envs_idx: Optional[torch.Tensor] = None):
        if envs_idx is None:
            rc = self._lib.gsr_forward_kinematics(self.dm.desc, self.st.desc, None, 0, self._stream())
        else:
            idx = envs_idx.to(device=self.device, dtype=torch.int32).contiguous()
            rc = self._lib.gsr_forward_kinematics(self.dm.desc, self.st.desc, ctypes.c_void_p(idx.data_ptr()),
                                                  int(idx.numel()), self._stream())
        _lib.check(rc, "gsr_forward_kinematics")

    def reset_caches(self, envs_idx: Optional[torch.Tensor] = None, what=GSR_RESET_COLLIDER | GSR_RESET_WARMSTART):
        if envs_idx is None:
            rc = self._lib.gsr_reset_envs(self.dm.desc, self.st.desc, None, 0, what, self._stream())
        else:
            idx = envs_idx.to(device=self.device, dtype=torch.int32).contiguous()
            rc = self._lib.gsr_reset_envs(self.dm.desc, self.st.desc, ctypes.c_void_p(idx.data_ptr()),
                                          int(idx.numel()), what, self._stream())
        _lib.check(rc, "gsr_reset_envs")

    def substep(self):
        _lib.check(self._lib.gsr_step(self.dm.desc, self.st.desc, 1, self._stream()), "gsr_step")
        self._mass_valid = True
        self._substep += 1
        if self._substep % RATE_CHECK_ERRNO == 0:
            self.check_errno()

    def set_variants(self, solver_variant: Optional[int] = None, collide_variant: Optional[int] = None,
                     dyn_variant: Optional[int] = None):
        """Switch the execution variants of this solver's kernels (between substeps: every variant reads and writes
        the same persistent state; only the per-substep constraint rows and Hessian take the solver variant's
        layout, and every later launch and readback then uses it)."""
        d = self.dm.desc
        if solver_variant is not None:
            if int(solver_variant) == 1 and self.model.n_entities > WAVE_MAX_ENTITIES:
                raise ValueError(f"wavefront solver supports at most {WAVE_MAX_ENTITIES} entities")
            d.solver_variant = int(solver_variant)
        if collide_variant is not None:
            d.collide_variant = int(collide_variant)
        if dyn_variant is not None:
            d.dyn_variant = int(dyn_variant)

    def launch_kernel(self, k: int):
        _lib.check(self._lib.gsr_launch_kernel(self.dm.desc, self.st.desc, k, self._stream()), "gsr_launch_kernel")

    def kernel_name(self, k: int) -> str:
        """Device kernel that launch_kernel(k) runs for this model's variant switches."""
        d = self.dm.desc
        wave = [d.dyn_variant, d.collide_variant, d.solver_variant, d.dyn_variant][k]
        return ["k_step1", "k_collide", "k_solve", "k_step2"][k] + ("_wave" if wave else "")

    def run_stage(self, stage: int):
        _lib.check(self._lib.gsr_run_stage(self.dm.desc, self.st.desc, stage, self._stream()), "gsr_run_stage")

    def check_errno(self):
        e = self._reduce_errno()
        for bit, msg in ERRNO_MESSAGES:
            if e & bit:
                raise GenesisException(msg)

    def _reduce_errno(self):
        rc = self._lib.gsr_get_errno(self.st.desc, ctypes.c_void_p(self._errno_dev.data_ptr()), self._stream())
        _lib.check(rc, "gsr_get_errno")
        return int(self._errno_dev.item())

    # -- zero-copy views ----------------------------------------------------------------------------------------
    def view(self, name: str) -> torch.Tensor:
        """(B, n) view of a per-env array (env-minor storage)."""
        return self.st.t[name].T

    def _idx(self, idx, n, device=None):
        if idx is None:
            return torch.arange(n, device=device or self.device)
        t = torch.as_tensor(idx, device=device or self.device)
        if t.dtype == torch.bool:
            t = t.nonzero().flatten()
        return t.long().reshape(-1)

    # -- control and readback (rsd:2123-2201, 2366-2376) --------------------------------------------------------
    def _write(self, name, values, dofs_idx, envs_idx, n):
        d = self._idx(dofs_idx, n)
        e = self._idx(envs_idx, self._B)
        v = torch.as_tensor(values, dtype=torch.float32, device=self.device)
        if self.n_envs == 0:
            v = v.reshape(1, -1)
        v = v.reshape(len(e), len(d))
        self.st.t[name][d[:, None], e[None, :]] = v.T

    def control_dofs_position(self, position, dofs_idx=None, envs_idx=None):
        d = self._idx(dofs_idx, self.model.n_dofs)
        e = self._idx(envs_idx, self._B)
        self.st.t["ctrl_mode"][d[:, None], e[None, :]] = CTRL_POSITION
        self._write("ctrl_pos", position, d, e, self.model.n_dofs)
        self.st.t["ctrl_vel"][d[:, None], e[None, :]] = 0.0

    def control_dofs_velocity(self, velocity, dofs_idx=None, envs_idx=None):
        d = self._idx(dofs_idx, self.model.n_dofs)
        e = self._idx(envs_idx, self._B)
        self.st.t["ctrl_mode"][d[:, None], e[None, :]] = CTRL_VELOCITY
        self._write("ctrl_vel", velocity, d, e, self.model.n_dofs)

    def control_dofs_force(self, force, dofs_idx=None, envs_idx=None):
        d = self._idx(dofs_idx, self.model.n_dofs)
        e = self._idx(envs_idx, self._B)
        self.st.t["ctrl_mode"][d[:, None], e[None, :]] = CTRL_FORCE
        self._write("ctrl_force", force, d, e, self.model.n_dofs)

    def _read(self, name, idx, envs_idx, n):
        d = self._idx(idx, n)
        v = self.st.t[name][d]
        if envs_idx is not None:
            v = v[:, self._idx(envs_idx, self._B)]
        v = v.T
        return v[0] if self.n_envs == 0 else v

    def get_dofs_position(self, dofs_idx=None, envs_idx=None):
        return self._read("dofs_pos", dofs_idx, envs_idx, self.model.n_dofs)

    def get_dofs_velocity(self, dofs_idx=None, envs_idx=None):
        return self._read("vel", dofs_idx, envs_idx, self.model.n_dofs)

    def get_qpos(self, qs_idx=None, envs_idx=None):
        return self._read("qpos", qs_idx, envs_idx, self.model.n_qs)

    def set_qpos(self, qpos, qs_idx=None, envs_idx=None, skip_forward=False):
        q = self._idx(qs_idx, self.model.n_qs)
        e = self._idx(envs_idx, self._B)
        v = torch.as_tensor(qpos, dtype=torch.float32, device=self.device).reshape(len(e), len(q))
        self.st.t["qpos"][q[:, None], e[None, :]] = v.T
        self.reset_caches(e.to(torch.int32))
        if not skip_forward:
            self.forward_kinematics(e.to(torch.int32))

    def set_dofs_velocity(self, velocity, dofs_idx=None, envs_idx=None, skip_forward=False):
        d = self._idx(dofs_idx, self.model.n_dofs)
        e = self._idx(envs_idx, self._B)
        self._write("vel", velocity, d, e, self.model.n_dofs)
        self.reset_caches(e.to(torch.int32), GSR_RESET_WARMSTART)
        if not skip_forward:
            self.forward_kinematics(e.to(torch.int32))

    def get_links_pos(self, envs_idx=None):
        v = self.view("links_pos").reshape(self._B, self.model.n_links, 3)
        return v if envs_idx is None else v[self._idx(envs_idx, self._B)]

    def get_links_quat(self, envs_idx=None):
        v = self.view("links_quat").reshape(self._B, self.model.n_links, 4)
        return v if envs_idx is None else v[self._idx(envs_idx, self._B)]

    # -- state snapshot (rsd:1442-1512) ------------------------------------------------------------------------
    def get_state(self, s_global: int = 0) -> "RigidSolverState":
        """Snapshot of the dynamic state (kernel_get_state, rsd:7165-7209), env-major copies."""
        B, L = self._B, self.model.n_links
        st = RigidSolverState(B, self.model, self.device, s_global)
        st.qpos.copy_(self.st.t["qpos"].T)
        st.dofs_vel.copy_(self.st.t["vel"].T)
        st.dofs_acc.copy_(self.st.t["acc"].T)
        st.links_pos.copy_(self.st.t["links_pos"].T.reshape(B, L, 3))
        st.links_quat.copy_(self.st.t["links_quat"].T.reshape(B, L, 4))
        st.i_pos_shift.copy_(self.st.t["i_pos_shift"].T.reshape(B, L, 3))
        st.mass_shift.copy_(self.st.t["mass_shift"].T)
        st.friction_ratio.copy_(self.st.t["friction_ratio"].T)
        return st

    def set_state(self, state: "RigidSolverState", envs_idx=None):
        """kernel_set_state (rsd:7212-7258) for the listed envs: qpos, dofs velocity and acceleration, COM / mass
        shifts and friction ratios, controls back to zero force; then forward kinematics of the new poses and cleared
        collider / warm-start caches and errno (rsd:1478-1512). Link poses follow from qpos through the kinematics, as
        in the reference."""
        e = self._idx(envs_idx, self._B)
        t = self.st.t
        B, L = self._B, self.model.n_links
        dev = self.device
        ed = e.to(dev)
        t["qpos"][:, ed] = state.qpos.to(dev, torch.float32)[ed].T
        t["vel"][:, ed] = state.dofs_vel.to(dev, torch.float32)[ed].T
        t["acc"][:, ed] = state.dofs_acc.to(dev, torch.float32)[ed].T
        t["i_pos_shift"][:, ed] = state.i_pos_shift.to(dev, torch.float32).reshape(B, 3 * L)[ed].T
        t["mass_shift"][:, ed] = state.mass_shift.to(dev, torch.float32)[ed].T
        t["friction_ratio"][:, ed] = state.friction_ratio.to(dev, torch.float32)[ed].T
        t["ctrl_force"][:, ed] = 0.0
        t["ctrl_mode"][:, ed] = CTRL_FORCE
        ei = ed.to(torch.int32)
        self.forward_kinematics(ei)
        self.reset_caches(ei)

    # -- checkpoint (scene.py:1283-1344) -------------------------------------------------------------------------
    def checkpoint_arrays(self):
        """Host copies of every state array that outlives a substep (the rest is rewritten by each substep before
        it is read), keyed "RigidSolver.<field>", and of every model parameter table ("RigidSolver.model.<table>",
        including per-env tables and the parameters setters changed, plus meaninertia), like the reference's
        dump_ckpt_to_numpy, which saves every solver field (base_solver.py:63-88)."""
        out = {f"RigidSolver.{k}": v.detach().cpu().numpy() for k, v in self.st.t.items()
               if k not in CHECKPOINT_SKIP}
        for k in MODEL_TABLES_F:
            out[f"RigidSolver.model.{k}"] = self.dm.tensors[k].detach().cpu().numpy()
        out["RigidSolver.model.meaninertia"] = np.float32(self.model.meaninertia)
        if self.dm.env_meaninertia is not None:
            out["RigidSolver.model.env_meaninertia"] = self.dm.env_meaninertia.detach().cpu().numpy()
        return out

    def load_checkpoint_arrays(self, arrays):
        def fits(a, v):
            return tuple(a.shape) == tuple(v.shape) and np.dtype(a.dtype) == np.dtype(str(v.dtype).replace("torch.", ""))

        pending = []
        for k, v in self.st.t.items():
            if k in CHECKPOINT_SKIP:
                continue
            a = arrays.get(f"RigidSolver.{k}")
            if a is None:
                raise KeyError(f"checkpoint has no array for state field {k}")
            if not fits(a, v):
                raise ValueError(f"checkpoint array {k}: {a.dtype}{tuple(a.shape)} does not fit {v.dtype}{tuple(v.shape)}")
            pending.append((v, a))
        for k in MODEL_TABLES_F:
            a = arrays.get(f"RigidSolver.model.{k}")
            if a is None:
                raise KeyError(f"checkpoint has no model table {k}")
            if not fits(a, self.dm.tensors[k]):
                raise ValueError(f"checkpoint model table {k} does not fit this scene")
        for v, a in pending:
            v.copy_(torch.from_numpy(np.ascontiguousarray(a)).to(v.device))
        # model tables: the device tables in place (every descriptor keeps its pointers) and the host model arrays
        # (shared tables; a per-env table has no host copy)
        for k in MODEL_TABLES_F:
            a = np.ascontiguousarray(arrays[f"RigidSolver.model.{k}"])
            self.dm.tensors[k].copy_(torch.from_numpy(a).to(self.device))
            host = getattr(self.model, k)
            # a per-env table has no host copy; an empty table is uploaded as one padding element
            if not self.dm.batched(k) and host.size:
                host[...] = a.reshape(host.shape).astype(host.dtype)
        mi = arrays.get("RigidSolver.model.meaninertia")
        if mi is not None:
            self.model.meaninertia = np.float32(mi)
            self.dm.desc.meaninertia = float(mi)
        emi = arrays.get("RigidSolver.model.env_meaninertia")
        if emi is not None and self.dm.env_meaninertia is not None and emi.shape == tuple(self.dm.env_meaninertia.shape):
            self.dm.env_meaninertia.copy_(torch.from_numpy(np.ascontiguousarray(emi, np.float32)).to(self.device))

    # -- model parameters (rsd:1777-2010 setters, 1880-1925 / 2200-2225 sol params, 2334-2452 getters) -----------
    # A table holds one value per dof / link / joint shared by every env (the reference's default batch_dofs_info =
    # batch_links_info = batch_joints_info = False, options/solvers.py:373-375) or, when its kind is batched, one
    # value per env (env-minor device table, no host copy); equality solver parameters are always per env. Shared
    # tables keep the host model array and the device table equal; both are updated in place, so every descriptor
    # holding the table pointers sees the new values at the next launch.
    _KIND_N = {"dofs": "n_dofs", "links": "n_links", "joints": "n_joints", "geoms": "n_geoms", "eqs": "n_equalities"}

    def _model_write(self, name, values, idx, n, width=0):
        arr = getattr(self.model, name)
        i = self._idx(idx, n, device="cpu").numpy()
        if ((i < 0) | (i >= n)).any():
            raise IndexError(f"{name}: index out of range [0, {n})")
        v = np.asarray(values.detach().cpu() if torch.is_tensor(values) else values, np.float64)
        v = v.reshape((len(i), width) if width else (len(i),))
        arr[i] = v.astype(arr.dtype)
        self.dm.tensors[name].copy_(torch.from_numpy(np.ascontiguousarray(arr, np.float32).ravel()))

    def _model_read(self, name, idx, n):
        t = self.dm.tensors[name]
        w = t.numel() // max(n, 1)
        t = t.reshape(n, w) if w > 1 else t
        i = self._idx(idx, n)
        if bool(((i < 0) | (i >= n)).any()):
            raise IndexError(f"{name}: index out of range [0, {n})")
        return t[i].clone()

    @staticmethod
    def _no_envs(envs_idx, what):
        if envs_idx is not None:
            raise GenesisException(f"`envs_idx` cannot be specified for non-batched {what} info.")

    def _set_table(self, name, values, idx, envs_idx, kind, width=0):
        """Write `values` for entries idx (all if None) of table `name`: (n_envs, n[, width]) or (n[, width]) values
        broadcast over envs_idx for a per-env table; (n[, width]) for a shared one, where envs_idx must be None."""
        n = int(getattr(self.model, self._KIND_N[kind]))
        if self.dm.batched(name):
            k, w = len(self._idx(idx, n)), max(width, 1)
            e = self._idx(envs_idx, self._B)
            v = torch.as_tensor(values, dtype=torch.float32, device=self.device).reshape(-1, k, w)
            self._write_env_table(self.dm.tensors[name].view(n * w, self._B), v.expand(len(e), k, w), idx, n, e, w)
            return
        self._no_envs(envs_idx, kind)
        self._model_write(name, values, idx, n, width)

    def _write_env_table(self, table, v, idx, n, e, width):
        i = self._idx(idx, n)
        if bool(((i < 0) | (i >= n)).any()):
            raise IndexError(f"index out of range [0, {n})")
        if bool(((e < 0) | (e >= self._B)).any()):
            raise IndexError(f"envs_idx out of range [0, {self._B})")
        rows = (i[:, None] * width + torch.arange(width, device=self.device)[None]).reshape(-1)
        table[rows[:, None], e[None, :]] = v.reshape(len(e), -1).T

    def _get_table(self, name, idx, envs_idx, kind, width=1):
        """Values of entries idx: (n_envs, n[, width]) for a per-env table ((n[, width]) without batch envs), else
        (n[, width]), where envs_idx must be None."""
        n = int(getattr(self.model, self._KIND_N[kind]))
        if self.dm.batched(name):
            i = self._idx(idx, n)
            if bool(((i < 0) | (i >= n)).any()):
                raise IndexError(f"{name}: index out of range [0, {n})")
            rows = (i[:, None] * width + torch.arange(width, device=self.device)[None]).reshape(-1)
            v = self.dm.tensors[name].view(n * width, self._B)[rows][:, self._idx(envs_idx, self._B)].T
            v = v.reshape(-1, len(i), width)
            v = v[..., 0] if width == 1 else v
            return v[0] if self.n_envs == 0 else v
        self._no_envs(envs_idx, kind)
        return self._model_read(name, idx, n)

    # kept under their former names for the dof setters below
    def _set_dofs(self, name, values, dofs_idx, envs_idx, width=0):
        self._set_table(name, values, dofs_idx, envs_idx, "dofs", width)

    def _get_dofs(self, name, dofs_idx, envs_idx, width=1):
        return self._get_table(name, dofs_idx, envs_idx, "dofs", width)

    def set_dofs_kp(self, kp, dofs_idx=None, envs_idx=None):
        self._set_dofs("dof_kp", kp, dofs_idx, envs_idx)

    def set_dofs_kv(self, kv, dofs_idx=None, envs_idx=None):
        self._set_dofs("dof_kv", kv, dofs_idx, envs_idx)

    def set_dofs_stiffness(self, stiffness, dofs_idx=None, envs_idx=None):
        self._set_dofs("dof_stiffness", stiffness, dofs_idx, envs_idx)

    def set_dofs_damping(self, damping, dofs_idx=None, envs_idx=None):
        self._set_dofs("dof_damping", damping, dofs_idx, envs_idx)

    def set_dofs_frictionloss(self, frictionloss, dofs_idx=None, envs_idx=None):
        self._set_dofs("dof_frictionloss", frictionloss, dofs_idx, envs_idx)

    def set_dofs_force_range(self, lower, upper, dofs_idx=None, envs_idx=None):
        lo, hi = (torch.as_tensor(x, dtype=torch.float64).cpu() for x in (lower, upper))
        self._set_dofs("dof_force_range", torch.stack([lo, hi], -1), dofs_idx, envs_idx, 2)

    def set_dofs_limit(self, lower, upper, dofs_idx=None, envs_idx=None):
        lo, hi = (torch.as_tensor(x, dtype=torch.float64).cpu() for x in (lower, upper))
        self._set_dofs("dof_limit", torch.stack([lo, hi], -1), dofs_idx, envs_idx, 2)

    def set_dofs_armature(self, armature, dofs_idx=None, envs_idx=None):
        """Armature enters the mass matrix, so the invweights and meaninertia are recomputed at the neutral pose
        (rsd:1964-1969 -> _init_invweight_and_meaninertia :354-507); the envs keep their qpos and lose their
        contact and warm-start caches, as through set_qpos. Choosing envs needs batched dofs and links (rsd:360-364)."""
        if envs_idx is not None and not (self.dm.batch_dofs and self.dm.batch_links):
            raise GenesisException("Links and dofs must be batched to selectively update invweight and meaninertia "
                                   "for some environment.")
        self._set_dofs("dof_armature", armature, dofs_idx, envs_idx)
        self._update_invweight_and_meaninertia(envs_idx)

    def _update_invweight_and_meaninertia(self, envs_idx=None):
        """Link / dof inverse weights and mean inertia of envs_idx (all if None) from the float64 host restatement
        of _init_invweight_and_meaninertia (rsd:354-515) on each env's armature and link masses at qpos0: per-env
        tables get each env's values, shared tables the first env's (the reference's loop stops after it when
        dofs and links are not both batched); per-env mean inertia goes to every env. Envs with equal parameters
        share one evaluation."""
        import copy

        from .model import _init_invweight_and_meaninertia
        dm, M = self.dm, self.model
        e = self._idx(envs_idx, self._B).cpu().numpy()
        if dm.env_meaninertia is None:
            _init_invweight_and_meaninertia(M)
            for k in ("dof_invweight", "link_invweight"):
                dm.tensors[k].copy_(torch.from_numpy(np.ascontiguousarray(getattr(M, k), np.float32).ravel()))
            dm.desc.meaninertia = float(M.meaninertia)
        else:
            D, L, B = M.n_dofs, M.n_links, self._B
            arm = dm.tensors["dof_armature"].view(D, B).cpu().numpy() if dm.batched("dof_armature") else None
            mass = (dm.tensors["link_inertial_mass"].view(L, B).cpu().numpy()
                    if dm.batched("link_inertial_mass") else None)
            memo, out = {}, []
            for b in e:
                key = (b"" if arm is None else arm[:, b].tobytes(), b"" if mass is None else mass[:, b].tobytes())
                if key not in memo:
                    Mb = copy.copy(M)
                    if arm is not None:
                        Mb.dof_armature = arm[:, b].copy()
                    if mass is not None:
                        Mb.link_inertial_mass = mass[:, b].copy()
                    _init_invweight_and_meaninertia(Mb)
                    memo[key] = (Mb.dof_invweight, Mb.link_invweight, np.float32(Mb.meaninertia))
                out.append(memo[key])
            ed = torch.as_tensor(e, device=self.device)
            for k, j, n in (("dof_invweight", 0, D), ("link_invweight", 1, 2 * L)):
                if dm.batched(k):
                    vals = torch.as_tensor(np.stack([o[j].reshape(-1) for o in out], 1), device=self.device)
                    dm.tensors[k].view(n, B)[:, ed] = vals
                else:
                    setattr(M, k, out[0][j].copy())
                    dm.tensors[k].copy_(torch.from_numpy(np.ascontiguousarray(out[0][j], np.float32).ravel()))
            dm.env_meaninertia[ed] = torch.as_tensor(np.array([o[2] for o in out], np.float32), device=self.device)
        ei = None if envs_idx is None else torch.as_tensor(e, dtype=torch.int32, device=self.device)
        self.reset_caches(ei)
        self.forward_kinematics(ei)

    def set_links_inertial_mass(self, mass, links_idx=None, envs_idx=None):
        """Link masses enter the COM and composite inertias that the end-of-step kinematics keep for the next step,
        so those are recomputed here (for envs_idx with batched links, else for every env); like the reference,
        the inverse weights keep their build values."""
        self._set_table("link_inertial_mass", mass, links_idx, envs_idx, "links")
        self.forward_kinematics(None if envs_idx is None else self._idx(envs_idx, self._B).to(torch.int32))

    def set_geoms_friction(self, friction, geoms_idx=None):
        self._model_write("geom_friction", friction, geoms_idx, self.model.n_geoms)

    def set_sol_params(self, sol_params, geoms_idx=None, envs_idx=None, *, joints_idx=None, eqs_idx=None):
        """Constraint solver parameters (timeconst, dampratio, dmin, dmax, width, mid, power) of geoms (shared),
        joints (per env with batch_joints_info) or equalities (always per env), sanitised with the minimum time
        constant 2 dt as their default (rsd:1880-1925, 49-77)."""
        from .model import TIME_CONSTANT_SAFETY_FACTOR, sanitize_sol_params
        if sum(x is not None for x in (geoms_idx, joints_idx, eqs_idx)) > 1:
            raise GenesisException("Cannot set more than one constraint type at once.")
        kind, idx, name = (("eqs", eqs_idx, "eq_sol_params") if eqs_idx is not None else
                           ("joints", joints_idx, "joint_sol_params") if joints_idx is not None else
                           ("geoms", geoms_idx, "geom_sol_params"))
        min_tc = TIME_CONSTANT_SAFETY_FACTOR * float(self.model.options.dt)
        sp = np.asarray(sol_params.detach().cpu() if torch.is_tensor(sol_params) else sol_params, np.float64)
        sp = sanitize_sol_params(sp, min_tc, min_tc).reshape(sp.shape)
        self._set_table(name, sp, idx, envs_idx, kind, 7)

    def get_sol_params(self, geoms_idx=None, envs_idx=None, *, joints_idx=None, eqs_idx=None):
        kind, idx, name = (("eqs", eqs_idx, "eq_sol_params") if eqs_idx is not None else
                           ("joints", joints_idx, "joint_sol_params") if joints_idx is not None else
                           ("geoms", geoms_idx, "geom_sol_params"))
        return self._get_table(name, idx, envs_idx, kind, 7)

    def get_dofs_kp(self, dofs_idx=None, envs_idx=None):
        return self._get_dofs("dof_kp", dofs_idx, envs_idx)

    def get_dofs_kv(self, dofs_idx=None, envs_idx=None):
        return self._get_dofs("dof_kv", dofs_idx, envs_idx)

    def get_dofs_stiffness(self, dofs_idx=None, envs_idx=None):
        return self._get_dofs("dof_stiffness", dofs_idx, envs_idx)

    def get_dofs_damping(self, dofs_idx=None, envs_idx=None):
        return self._get_dofs("dof_damping", dofs_idx, envs_idx)

    def get_dofs_frictionloss(self, dofs_idx=None, envs_idx=None):
        return self._get_dofs("dof_frictionloss", dofs_idx, envs_idx)

    def get_dofs_armature(self, dofs_idx=None, envs_idx=None):
        return self._get_dofs("dof_armature", dofs_idx, envs_idx)

    def get_dofs_invweight(self, dofs_idx=None, envs_idx=None):
        return self._get_dofs("dof_invweight", dofs_idx, envs_idx)

    def get_dofs_force_range(self, dofs_idx=None, envs_idx=None):
        t = self._get_dofs("dof_force_range", dofs_idx, envs_idx, 2)
        return t[..., 0], t[..., 1]

    def get_dofs_limit(self, dofs_idx=None, envs_idx=None):
        t = self._get_dofs("dof_limit", dofs_idx, envs_idx, 2)
        return t[..., 0], t[..., 1]

    def get_links_inertial_mass(self, links_idx=None, envs_idx=None):
        return self._get_table("link_inertial_mass", links_idx, envs_idx, "links")

    def get_links_invweight(self, links_idx=None, envs_idx=None):
        return self._get_table("link_invweight", links_idx, envs_idx, "links", 2)

    def get_meaninertia(self, envs_idx=None):
        """The solver's per-env mean inertia (rigid_global_info.meaninertia)."""
        e = self._idx(envs_idx, self._B)
        if self.dm.env_meaninertia is not None:
            return self.dm.env_meaninertia[e].clone()
        return torch.full((len(e),), float(self.dm.desc.meaninertia), dtype=torch.float32, device=self.device)

    def get_geoms_friction(self, geoms_idx=None):
        return self._model_read("geom_friction", geoms_idx, self.model.n_geoms)

    def get_mass_mat(self, dofs_idx=None, envs_idx=None, decompose=False):
        """Joint-space mass matrix (B, n, n) the solver holds (rsd:2436-2449): the one the last substep's dynamics
        built, at the pose before that substep's integration, with armature and the implicit damping terms (dof
        damping, and kv for position / velocity controlled dofs, times dt). Before the first substep the dynamics
        kernel builds it at the current pose; its other outputs are recomputed from the same inputs by the next
        step, so the call changes no result. decompose: (L, D_inv) with unit-diagonal L and
        M = L^T diag(1 / D_inv) L."""
        if not self._mass_valid:
            self.launch_kernel(0)
            self._mass_valid = True
        Mo, B = self.model, self._B
        src = self.st.t["mass_L_blk" if decompose else "mass_blk"]
        out = torch.zeros((B, Mo.n_dofs, Mo.n_dofs), dtype=torch.float32, device=self.device)
        for ie in range(Mo.n_entities):
            st, en = int(Mo.ent_dof_start[ie]), int(Mo.ent_dof_end[ie])
            n = en - st
            if n == 0:
                continue
            off = int(Mo.ent_mass_offset[ie])
            blk = src[off:off + n * n].T.reshape(B, n, n)
            out[:, st:en, st:en] = torch.tril(blk) if decompose else blk
        e = None if envs_idx is None else self._idx(envs_idx, B)
        if e is not None:
            out = out[e]
        if dofs_idx is not None:
            d = self._idx(dofs_idx, Mo.n_dofs)
            out = out[:, d][:, :, d]
        if decompose:
            dinv = self.st.t["mass_D_inv"].T
            dinv = dinv if e is None else dinv[e]
            dinv = dinv if dofs_idx is None else dinv[:, self._idx(dofs_idx, Mo.n_dofs)]
            return (out[0], dinv[0]) if self.n_envs == 0 else (out, dinv)
        return out[0] if self.n_envs == 0 else out

    # -- per-env shifts (rsd:1755-1799 setters, 2326-2350 getters): links_state / geoms_state, (n, B) like all state
    def _write_env(self, name, values, idx, n, envs_idx, width=1):
        i = self._idx(idx, n)
        if bool(((i < 0) | (i >= n)).any()):
            raise IndexError(f"{name}: index out of range [0, {n})")
        e = self._idx(envs_idx, self._B)
        v = torch.as_tensor(values, dtype=torch.float32, device=self.device)
        if v.numel() == len(i) * width:  # one value set for every listed env
            v = v.reshape(1, len(i), width).expand(len(e), len(i), width)
        v = v.reshape(len(e), len(i), width)
        rows = (i[:, None] * width + torch.arange(width, device=self.device)[None]).reshape(-1)
        self.st.t[name][rows[:, None], e[None, :]] = v.reshape(len(e), -1).T
        return e

    def _read_env(self, name, idx, n, envs_idx, width=1):
        i = self._idx(idx, n)
        rows = (i[:, None] * width + torch.arange(width, device=self.device)[None]).reshape(-1)
        v = self.st.t[name][rows][:, self._idx(envs_idx, self._B)].T.reshape(-1, len(i), width)
        v = v[..., 0] if width == 1 else v
        return v[0] if self.n_envs == 0 else v

    def set_links_mass_shift(self, mass, links_idx=None, envs_idx=None):
        """Added to each listed link's mass in the listed envs; the COM caches are recomputed for those envs."""
        e = self._write_env("mass_shift", mass, links_idx, self.model.n_links, envs_idx)
        self.forward_kinematics(e.to(torch.int32))

    def set_links_COM_shift(self, com, links_idx=None, envs_idx=None):
        """Added to each listed link's inertial position (link frame) in the listed envs."""
        e = self._write_env("i_pos_shift", com, links_idx, self.model.n_links, envs_idx, 3)
        self.forward_kinematics(e.to(torch.int32))

    def set_geoms_friction_ratio(self, friction_ratio, geoms_idx=None, envs_idx=None):
        """Scales each listed geom's friction in the listed envs (read when contacts are created)."""
        self._write_env("friction_ratio", friction_ratio, geoms_idx, self.model.n_geoms, envs_idx)

    def get_links_mass_shift(self, links_idx=None, envs_idx=None):
        return self._read_env("mass_shift", links_idx, self.model.n_links, envs_idx)

    def get_links_COM_shift(self, links_idx=None, envs_idx=None):
        return self._read_env("i_pos_shift", links_idx, self.model.n_links, envs_idx, 3)

    def get_geoms_friction_ratio(self, geoms_idx=None, envs_idx=None):
        return self._read_env("friction_ratio", geoms_idx, self.model.n_geoms, envs_idx)

    # -- contact and constraint readback (col:463-565, cs:235-322, rigid_entity.py:2838-2850) ---------------------
    @property
    def collider(self) -> "Collider":
        return Collider(self)

    @property
    def constraint_solver(self) -> "ConstraintSolver":
        return ConstraintSolver(self)

    def get_links_net_contact_force(self, envs_idx=None, links_idx=None):
        """Net force of the last step's contacts on each link, (n_envs, n_links, 3), or (n_links, 3) when
        n_envs == 0: the per-link sums of func_update_contact_force (constraint_solver_decomp.py:1350-1379)."""
        L = self.model.n_links
        if envs_idx is None:
            n, idx = self._B, None
        else:
            idx = self._idx(envs_idx, self._B).to(torch.int32).contiguous()
            if ((idx < 0) | (idx >= self._B)).any():
                raise IndexError("envs_idx out of range")
            n = int(idx.numel())
        out = torch.empty((n, L, 3), dtype=torch.float32, device=self.device)
        rc = self._lib.gsr_links_contact_force(self.dm.desc, self.st.desc,
                                               None if idx is None else ctypes.c_void_p(idx.data_ptr()), n,
                                               ctypes.c_void_p(out.data_ptr()), self._stream())
        _lib.check(rc, "gsr_links_contact_force")
        if links_idx is not None:
            out = out[:, self._idx(links_idx, L)]
        return out[0] if self.n_envs == 0 else out


def _to_numpy(values):
    return tuple(v.cpu().numpy() for v in values)


class Collider:
    """Contact readback of the reference's Collider (collider_decomp.py:463-565): the contacts found by the last
    step's collision detection, with the per-contact force the last Newton solve gave them."""

    KEYS = ("link_a", "link_b", "geom_a", "geom_b", "penetration", "position", "normal", "force")

    def __init__(self, solver: RigidSolver):
        self._solver = solver

    def get_contacts(self, as_tensor: bool = True, to_torch: bool = True, keep_batch_dim: bool = False):
        """Dict of the contact fields. as_tensor: (n_envs, n_contacts_max, ...) padded with -1 / 0 (no env axis
        when n_envs == 0 unless keep_batch_dim); otherwise one entry per env in a tuple (flat when n_envs == 0)."""
        S = self._solver
        n_envs, B = S.n_envs, S._B
        counts_dev = S.st.t["n_contacts"][0].clamp(max=S.model.max_contacts)
        counts = counts_dev.cpu()
        n_max = int(counts.max()) if B else 0
        if as_tensor or n_envs == 0:
            starts, out_size = None, n_max * B
        else:
            starts = (torch.cumsum(counts_dev, 0) - counts_dev).to(torch.int32).contiguous()
            out_size = int(counts.sum())
        iout = torch.empty((out_size, 4), dtype=torch.int32, device=S.device)
        fout = torch.empty((out_size, 10), dtype=torch.float32, device=S.device)
        if out_size > 0:
            rc = S._lib.gsr_get_contacts(S.dm.desc, S.st.desc, None if starts is None else ctypes.c_void_p(starts.data_ptr()),
                                         n_max, out_size, ctypes.c_void_p(iout.data_ptr()),
                                         ctypes.c_void_p(fout.data_ptr()), S._stream())
            _lib.check(rc, "gsr_get_contacts")

        def fields(i, f):
            return (i[..., 0], i[..., 1], i[..., 2], i[..., 3], f[..., 0], f[..., 1:4], f[..., 4:7], f[..., 7:])

        if as_tensor or n_envs == 0:
            if n_envs > 0 or keep_batch_dim:
                iout, fout = iout.reshape(B, n_max, 4), fout.reshape(B, n_max, 10)
            values = fields(iout, fout)
            if not to_torch:
                values = _to_numpy(values)
        else:
            n = counts.tolist()
            per_env = [fields(i, f) for i, f in zip(torch.split(iout, n), torch.split(fout, n))]
            if keep_batch_dim:
                per_env = [tuple(v[None] for v in env) for env in per_env]
            if not to_torch:
                per_env = [_to_numpy(env) for env in per_env]
            values = tuple(zip(*per_env)) if per_env else tuple(() for _ in self.KEYS)
        return dict(zip(self.KEYS, values))


EQUALITY_WELD = 1  # gs.EQUALITY_TYPE: CONNECT 0, WELD 1, JOINT 2


# quaternion helpers over (4, n) / vectors over (3, n) float32 tensors, in the operation order of
# genesis/utils/geom.py (ti_inv_quat, ti_quat_mul, ti_transform_by_quat)
def _inv_quat(q):
    return torch.stack([q[0], -q[1], -q[2], -q[3]])


def _quat_mul(u, v):
    w = u[0] * v[0] - u[1] * v[1] - u[2] * v[2] - u[3] * v[3]
    x = u[0] * v[1] + u[1] * v[0] + u[2] * v[3] - u[3] * v[2]
    y = u[0] * v[2] - u[1] * v[3] + u[2] * v[0] + u[3] * v[1]
    z = u[0] * v[3] + u[1] * v[2] - u[2] * v[1] + u[3] * v[0]
    return torch.stack([w, x, y, z])


def _transform_by_quat(v, q):
    qw, qx, qy, qz = q
    xx, xy, xz, wx = qx * qx, qx * qy, qx * qz, qx * qw
    yy, yz, wy = qy * qy, qy * qz, qy * qw
    zz, wz = qz * qz, qz * qw
    ww = qw * qw
    out = torch.stack([
        v[0] * (xx + ww - yy - zz) + v[1] * (2.0 * xy - 2.0 * wz) + v[2] * (2.0 * xz + 2.0 * wy),
        v[0] * (2.0 * wz + 2.0 * xy) + v[1] * (ww - xx + yy - zz) + v[2] * (-2.0 * wx + 2.0 * yz),
        v[0] * (-2.0 * wy + 2.0 * xz) + v[1] * (2.0 * wx + 2.0 * yz) + v[2] * (ww - xx - yy + zz),
    ])
    return out / (ww + xx + yy + zz)


class ConstraintSolver:
    """Equality-constraint readback of the reference's ConstraintSolver (constraint_solver_decomp.py:235-322)."""

    def __init__(self, solver: RigidSolver):
        self._solver = solver

    def get_equality_constraints(self, as_tensor: bool = True, to_torch: bool = True):
        """{type, obj_a, obj_b, force}: (n_envs, max equalities of an env[, 6]) with as_tensor (type -1 past an env's
        count), else a tuple per env. An env holds the model's equalities, in model order, then its own welds."""
        S = self._solver
        B, E, K = S._B, S.model.n_equalities, S.model.max_dyn_welds
        N = E + K
        iout = torch.full((B * N, 3), -1, dtype=torch.int32, device=S.device)
        fout = torch.zeros((B * N, 6), dtype=torch.float32, device=S.device)
        if B * N > 0:
            rc = S._lib.gsr_get_equalities(S.dm.desc, S.st.desc, ctypes.c_void_p(iout.data_ptr()),
                                           ctypes.c_void_p(fout.data_ptr()), S._stream())
            _lib.check(rc, "gsr_get_equalities")
        counts = E + (S.st.t["eq_dyn_n"][0].to(torch.int64) if K > 0 else torch.zeros(B, dtype=torch.int64,
                                                                                          device=S.device))
        n_max = int(counts.max().item()) if B > 0 else 0
        iout, fout = iout.reshape(B, N, 3)[:, :n_max], fout.reshape(B, N, 6)[:, :n_max]
        if as_tensor:
            values = (iout[..., 0], iout[..., 1], iout[..., 2], fout)
            if not to_torch:
                values = _to_numpy(values)
        else:
            cnt = counts.tolist()
            per_env = [(iout[b, :cnt[b], 0], iout[b, :cnt[b], 1], iout[b, :cnt[b], 2], fout[b, :cnt[b]])
                       for b in range(B)]
            if not to_torch:
                per_env = [_to_numpy(env) for env in per_env]
            values = tuple(zip(*per_env))
        return dict(zip(("type", "obj_a", "obj_b", "force"), values))

    def add_weld_constraint(self, link1_idx, link2_idx, envs_idx=None):
        """Registers a weld of link2 to link1 in their current relative pose (anchors at link1's origin, relative
        quaternion inv(q1) q2, torque scale 1; kernel_add_weld_constraint, constraint_solver_decomp.py:2433-2480)
        in the listed envs. The pair must not be welded already; an env whose list is full
        (RigidOptions.max_dynamic_constraints) keeps its welds and is reported in a warning."""
        S = self._solver
        l1, l2 = int(link1_idx), int(link2_idx)
        if not (0 <= l1 < S.model.n_links and 0 <= l2 < S.model.n_links):
            raise IndexError("link index out of range")
        info = self.get_weld_constraints(as_tensor=True, to_torch=True)
        la, lb = info["link_a"], info["link_b"]
        if bool((((la == l1) | (lb == l1)) & ((la == l2) | (lb == l2))).any()):
            raise GenesisException(f"links {l1} and {l2} are welded already")
        K = S.model.max_dyn_welds
        envs = S._idx(envs_idx, S._B)
        t = S.st.t
        n = t["eq_dyn_n"][0, envs].to(torch.int64)
        full = n >= K
        if bool(full.any()):
            import warnings
            warnings.warn(f"{int(full.sum())} env(s) hold max_dynamic_constraints = {K} welds already; the weld is "
                          "not added there")
        envs, n = envs[~full], n[~full]
        if envs.numel() == 0:
            return
        pos = t["links_pos"].view(S.model.n_links, 3, -1)
        quat = t["links_quat"].view(S.model.n_links, 4, -1)
        p1, p2 = pos[l1][:, envs], pos[l2][:, envs]
        q1, q2 = quat[l1][:, envs], quat[l2][:, envs]
        shared = p1
        data = torch.empty((11, envs.numel()), dtype=torch.float32, device=S.device)
        data[3:6] = _transform_by_quat(shared - p1, _inv_quat(q1))
        data[0:3] = _transform_by_quat(shared - p2, _inv_quat(q2))
        data[6:10] = _quat_mul(_inv_quat(q1), q2)
        data[10] = 1.0
        rows = 11 * n[None, :] + torch.arange(11, device=S.device)[:, None]
        t["eq_dyn_data"][rows, envs[None, :]] = data
        t["eq_dyn_obj"][2 * n, envs] = l1
        t["eq_dyn_obj"][2 * n + 1, envs] = l2
        t["eq_dyn_n"][0, envs] = (n + 1).to(torch.int32)

    def delete_weld_constraint(self, link1_idx, link2_idx, envs_idx=None):
        """Removes the weld of (link1, link2) registered by add_weld_constraint from the listed envs; the env's last
        weld takes its place (kernel_delete_weld_constraint, constraint_solver_decomp.py:2484-2506, moves only the
        type of the last entry: here the whole entry moves, so the remaining welds are the ones still registered)."""
        S = self._solver
        l1, l2 = int(link1_idx), int(link2_idx)
        K = S.model.max_dyn_welds
        if K == 0:
            return
        t = S.st.t
        envs = S._idx(envs_idx, S._B)
        obj = t["eq_dyn_obj"][:, envs].view(K, 2, -1)
        n = t["eq_dyn_n"][0, envs].to(torch.int64)
        slots = torch.arange(K, device=S.device)[:, None]
        match = (obj[:, 0] == l1) & (obj[:, 1] == l2) & (slots < n[None, :])
        for b_local in torch.nonzero(match.any(0)).flatten().tolist():
            b = int(envs[b_local])
            k = int(torch.nonzero(match[:, b_local])[0])
            last = int(n[b_local]) - 1
            if k < last:
                t["eq_dyn_obj"][2 * k:2 * k + 2, b] = t["eq_dyn_obj"][2 * last:2 * last + 2, b]
                t["eq_dyn_data"][11 * k:11 * k + 11, b] = t["eq_dyn_data"][11 * last:11 * last + 11, b]
            t["eq_dyn_n"][0, b] = last

    def get_weld_constraints(self, as_tensor: bool = True, to_torch: bool = True):
        """The WELD subset of get_equality_constraints, with obj_a / obj_b renamed link_a / link_b
        (constraint_solver_decomp.py:289-322)."""
        info = self.get_equality_constraints(as_tensor, to_torch)
        eq_type = info.pop("type")
        out = {}
        if as_tensor:
            mask = eq_type == EQUALITY_WELD
            n_welds = mask.sum(-1) if to_torch else np.sum(mask, axis=-1)
            n_max = int(n_welds.max()) if len(n_welds) else 0
            for key, value in info.items():
                shape = (len(mask), n_max, *value.shape[2:])
                if to_torch:
                    w = (torch.zeros(shape, dtype=value.dtype, device=value.device) if torch.is_floating_point(value)
                         else torch.full(shape, -1, dtype=value.dtype, device=value.device))
                else:
                    w = np.zeros(shape, value.dtype) if np.issubdtype(value.dtype, np.floating) else np.full(shape, -1, value.dtype)
                for b in range(len(mask)):
                    k = int(n_welds[b])
                    w[b, :k] = value[b][mask[b]]
                out[key] = w
        else:
            masks = tuple(t == EQUALITY_WELD for t in eq_type)
            for key, value in info.items():
                out[key] = tuple(v[m] for m, v in zip(masks, value))
        out["link_a"] = out.pop("obj_a")
        out["link_b"] = out.pop("obj_b")
        return out
