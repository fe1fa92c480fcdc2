"""gs.Scene / RigidEntity surface of the batched rigid step.

Mirrors the reference's user API on this path: Scene (genesis/engine/scene.py: add_entity :326, build :740-804,
step :916-936, reset :857-895, get_state :904) and RigidEntity (rigid_entity.py: control_dofs_position :2450,
get_dofs_position :2571, get_dofs_velocity :2551, set_qpos :2259, local->global index mapping :1780-1825). Option
classes keep the reference's names and defaults for the fields this path reads (genesis/options/solvers.py:15-51,
281-407). The step itself runs in the HIP library through RigidSolver.
"""
from __future__ import annotations

import os
import time
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np
import torch

from . import loaders
from .model import (CompileOptions, EntitySpec, GeomSpec, JOINT_FIXED, JOINT_FREE, JointSpec, LinkSpec, GEOM_BOX,
                    GEOM_PLANE, GEOM_SPHERE, compile_model)
from .rigid_solver import GenesisException, RigidSolver, RATE_CHECK_ERRNO

ASSETS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "assets")
# Reference asset paths whose compiled descriptions ship with the package.
BUILTIN_ASSETS = {
    "xml/franka_emika_panda/panda.xml": "panda",
    "urdf/go2/urdf/go2.urdf": "go2",
}


# ------------------------------------------------------------------------------------------------------------
# options
# ------------------------------------------------------------------------------------------------------------
@dataclass
class SimOptions:
    dt: float = 1e-2
    substeps: int = 1
    gravity: tuple = (0.0, 0.0, -9.81)
    requires_grad: bool = False


@dataclass
class RigidOptions:
    dt: Optional[float] = None
    gravity: Optional[tuple] = None
    enable_collision: bool = True
    enable_joint_limit: bool = True
    enable_self_collision: bool = True
    enable_adjacent_collision: bool = False
    disable_constraint: bool = False
    max_collision_pairs: int = 150
    multiplier_collision_broad_phase: int = 8
    integrator: str = "approximate_implicitfast"
    constraint_solver: str = "Newton"
    iterations: int = 50
    tolerance: float = 1e-6
    ls_iterations: int = 50
    ls_tolerance: float = 1e-2
    constraint_timeconst: float = 0.01
    enable_multi_contact: bool = True
    enable_mujoco_compatibility: bool = False
    box_box_detection: bool = False
    use_gjk_collision: Optional[bool] = None
    max_dynamic_constraints: int = 8
    batch_dofs_info: bool = False
    batch_links_info: bool = False
    batch_joints_info: bool = False


class options:  # noqa: N801 - namespace mirroring gs.options
    SimOptions = SimOptions
    RigidOptions = RigidOptions


# ------------------------------------------------------------------------------------------------------------
# morphs and materials
# ------------------------------------------------------------------------------------------------------------
@dataclass
class Morph:
    pos: tuple = (0.0, 0.0, 0.0)
    quat: tuple = (1.0, 0.0, 0.0, 0.0)
    fixed: bool = False
    collision: bool = True
    contype: int = 1
    conaffinity: int = 1


@dataclass
class Plane(Morph):
    normal: tuple = (0.0, 0.0, 1.0)
    fixed: bool = True


@dataclass
class Box(Morph):
    size: tuple = (1.0, 1.0, 1.0)


@dataclass
class Sphere(Morph):
    radius: float = 0.5


@dataclass
class URDF(Morph):
    file: str = ""
    merge_fixed_links: bool = True
    links_to_keep: tuple = ()


@dataclass
class MJCF(Morph):
    file: str = ""


class morphs:  # noqa: N801 - namespace mirroring gs.morphs
    Plane = Plane
    Box = Box
    Sphere = Sphere
    URDF = URDF
    MJCF = MJCF


@dataclass
class RigidMaterial:
    rho: float = 200.0
    friction: Optional[float] = None


class materials:  # noqa: N801
    Rigid = RigidMaterial


def _spec_from_morph(morph: Morph, material: RigidMaterial) -> EntitySpec:
    pos = np.asarray(morph.pos, np.float64)
    quat = np.asarray(morph.quat, np.float64)
    quat = quat / np.linalg.norm(quat)
    if isinstance(morph, Plane):
        link = LinkSpec("plane_baselink", -1, pos, quat, [JointSpec("plane_baselink_joint", JOINT_FIXED)],
                        [GeomSpec(GEOM_PLANE, data=tuple(morph.normal), contype=morph.contype,
                                  conaffinity=morph.conaffinity)] if morph.collision else [])
        return EntitySpec([link], rho=material.rho, friction=material.friction, name="plane")
    if isinstance(morph, (Box, Sphere)):
        g = GeomSpec(GEOM_BOX, data=tuple(morph.size)) if isinstance(morph, Box) else GeomSpec(GEOM_SPHERE, data=(morph.radius,))
        g.contype, g.conaffinity = morph.contype, morph.conaffinity
        name = "box" if isinstance(morph, Box) else "sphere"
        j = JointSpec(f"{name}_baselink_joint", JOINT_FIXED) if morph.fixed else JointSpec(
            f"{name}_baselink_joint", JOINT_FREE, kp=0.0, kv=0.0, init_qpos=np.concatenate([pos, quat]))
        link = LinkSpec(f"{name}_baselink", -1, pos, quat, [j], [g] if morph.collision else [])
        return EntitySpec([link], rho=material.rho, friction=material.friction, name=name)
    if isinstance(morph, (URDF, MJCF)):
        key = morph.file.replace("\\", "/")
        if key in BUILTIN_ASSETS and not os.path.exists(key):
            ent = loaders.entity_from_json(__import__("json").load(open(os.path.join(ASSETS, BUILTIN_ASSETS[key] + ".json"))))
        elif isinstance(morph, URDF):
            ent = loaders.parse_urdf(morph.file, fixed=morph.fixed, merge_fixed_links=morph.merge_fixed_links,
                                     links_to_keep=morph.links_to_keep)
        else:
            ent = loaders.parse_mjcf(morph.file)
        # place the root link (and free-joint initial pose) at the morph pose
        root = ent.links[0]
        for j in root.joints:
            if j.type == JOINT_FREE:
                j.init_qpos = np.concatenate([pos + (j.init_qpos[:3] if j.init_qpos is not None else 0),
                                              quat if j.init_qpos is None else j.init_qpos[3:]])
        if not any(j.type == JOINT_FREE for j in root.joints):
            root.pos = pos + np.asarray(root.pos)
            root.quat = quat
        if isinstance(morph, URDF) and morph.fixed:
            for j in root.joints:
                if j.type == JOINT_FREE:
                    j.type = JOINT_FIXED
        ent.rho = material.rho if material.rho != 200.0 else ent.rho
        ent.friction = material.friction
        return ent
    raise TypeError(f"unsupported morph {type(morph).__name__}")


# ------------------------------------------------------------------------------------------------------------
# entity
# ------------------------------------------------------------------------------------------------------------
class RigidEntity:
    def __init__(self, scene: "Scene", idx: int, morph, material, spec: EntitySpec):
        self._scene = scene
        self.idx = idx
        self.morph = morph
        self.material = material
        self.spec = spec

    # filled at build
    def _bind(self, M):
        self.link_start, self.link_end = int(M.ent_link_start[self.idx]), int(M.ent_link_end[self.idx])
        self.dof_start, self.dof_end = int(M.ent_dof_start[self.idx]), int(M.ent_dof_end[self.idx])
        self.geom_start, self.geom_end = int(M.ent_geom_start[self.idx]), int(M.ent_geom_end[self.idx])
        qs = [M.link_q_start[l] for l in range(self.link_start, self.link_end)]
        qe = [M.link_q_end[l] for l in range(self.link_start, self.link_end)]
        self.q_start, self.q_end = (int(min(qs)), int(max(qe))) if qs else (0, 0)
        self.n_dofs = self.dof_end - self.dof_start
        self.n_qs = self.q_end - self.q_start
        self.n_links = self.link_end - self.link_start

    @property
    def _solver(self) -> RigidSolver:
        return self._scene.rigid_solver

    def _global(self, idx_local, n, start):
        if idx_local is None:
            return torch.arange(start, start + n)
        t = torch.as_tensor(idx_local).long().reshape(-1)
        if ((t < 0) | (t >= n)).any():
            raise IndexError("local index out of range")
        return t + start

    def control_dofs_position(self, position, dofs_idx_local=None, envs_idx=None):
        self._solver.control_dofs_position(position, self._global(dofs_idx_local, self.n_dofs, self.dof_start), envs_idx)

    def control_dofs_velocity(self, velocity, dofs_idx_local=None, envs_idx=None):
        self._solver.control_dofs_velocity(velocity, self._global(dofs_idx_local, self.n_dofs, self.dof_start), envs_idx)

    def control_dofs_force(self, force, dofs_idx_local=None, envs_idx=None):
        self._solver.control_dofs_force(force, self._global(dofs_idx_local, self.n_dofs, self.dof_start), envs_idx)

    def get_dofs_position(self, dofs_idx_local=None, envs_idx=None):
        return self._solver.get_dofs_position(self._global(dofs_idx_local, self.n_dofs, self.dof_start), envs_idx)

    def get_dofs_velocity(self, dofs_idx_local=None, envs_idx=None):
        return self._solver.get_dofs_velocity(self._global(dofs_idx_local, self.n_dofs, self.dof_start), envs_idx)

    def get_qpos(self, qs_idx_local=None, envs_idx=None):
        return self._solver.get_qpos(self._global(qs_idx_local, self.n_qs, self.q_start), envs_idx)

    def set_qpos(self, qpos, qs_idx_local=None, envs_idx=None, zero_velocity=True):
        qidx = self._global(qs_idx_local, self.n_qs, self.q_start)
        if zero_velocity:
            self._solver.set_dofs_velocity(torch.zeros((self._solver._idx(envs_idx, self._solver._B).numel(), self.n_dofs)),
                                           self._global(None, self.n_dofs, self.dof_start), envs_idx, skip_forward=True)
        self._solver.set_qpos(qpos, qidx, envs_idx)

    def set_dofs_velocity(self, velocity, dofs_idx_local=None, envs_idx=None):
        self._solver.set_dofs_velocity(velocity, self._global(dofs_idx_local, self.n_dofs, self.dof_start), envs_idx)

    def get_links_pos(self, envs_idx=None):
        return self._solver.get_links_pos(envs_idx)[..., self.link_start:self.link_end, :]

    def get_links_quat(self, envs_idx=None):
        return self._solver.get_links_quat(envs_idx)[..., self.link_start:self.link_end, :]

    def get_pos(self, envs_idx=None):
        return self.get_links_pos(envs_idx)[..., 0, :]

    def get_quat(self, envs_idx=None):
        return self.get_links_quat(envs_idx)[..., 0, :]

    # model parameters of the entity's dofs / links (rigid_entity.py set_dofs_kp ... get_links_inertial_mass)
    def _dofs(self, dofs_idx_local):
        return self._global(dofs_idx_local, self.n_dofs, self.dof_start)

    def set_dofs_kp(self, kp, dofs_idx_local=None, envs_idx=None):
        self._solver.set_dofs_kp(kp, self._dofs(dofs_idx_local), envs_idx)

    def set_dofs_kv(self, kv, dofs_idx_local=None, envs_idx=None):
        self._solver.set_dofs_kv(kv, self._dofs(dofs_idx_local), envs_idx)

    def set_dofs_force_range(self, lower, upper, dofs_idx_local=None, envs_idx=None):
        self._solver.set_dofs_force_range(lower, upper, self._dofs(dofs_idx_local), envs_idx)

    def set_dofs_limit(self, lower, upper, dofs_idx_local=None, envs_idx=None):
        self._solver.set_dofs_limit(lower, upper, self._dofs(dofs_idx_local), envs_idx)

    def set_dofs_stiffness(self, stiffness, dofs_idx_local=None, envs_idx=None):
        self._solver.set_dofs_stiffness(stiffness, self._dofs(dofs_idx_local), envs_idx)

    def set_dofs_damping(self, damping, dofs_idx_local=None, envs_idx=None):
        self._solver.set_dofs_damping(damping, self._dofs(dofs_idx_local), envs_idx)

    def set_dofs_frictionloss(self, frictionloss, dofs_idx_local=None, envs_idx=None):
        self._solver.set_dofs_frictionloss(frictionloss, self._dofs(dofs_idx_local), envs_idx)

    def set_dofs_armature(self, armature, dofs_idx_local=None, envs_idx=None):
        self._solver.set_dofs_armature(armature, self._dofs(dofs_idx_local), envs_idx)

    def get_dofs_kp(self, dofs_idx_local=None, envs_idx=None):
        return self._solver.get_dofs_kp(self._dofs(dofs_idx_local), envs_idx)

    def get_dofs_kv(self, dofs_idx_local=None, envs_idx=None):
        return self._solver.get_dofs_kv(self._dofs(dofs_idx_local), envs_idx)

    def get_dofs_force_range(self, dofs_idx_local=None, envs_idx=None):
        return self._solver.get_dofs_force_range(self._dofs(dofs_idx_local), envs_idx)

    def get_dofs_limit(self, dofs_idx_local=None, envs_idx=None):
        return self._solver.get_dofs_limit(self._dofs(dofs_idx_local), envs_idx)

    def get_dofs_stiffness(self, dofs_idx_local=None, envs_idx=None):
        return self._solver.get_dofs_stiffness(self._dofs(dofs_idx_local), envs_idx)

    def get_dofs_damping(self, dofs_idx_local=None, envs_idx=None):
        return self._solver.get_dofs_damping(self._dofs(dofs_idx_local), envs_idx)

    def get_dofs_frictionloss(self, dofs_idx_local=None, envs_idx=None):
        return self._solver.get_dofs_frictionloss(self._dofs(dofs_idx_local), envs_idx)

    def get_dofs_armature(self, dofs_idx_local=None, envs_idx=None):
        return self._solver.get_dofs_armature(self._dofs(dofs_idx_local), envs_idx)

    def get_links_inertial_mass(self, links_idx_local=None, envs_idx=None):
        return self._solver.get_links_inertial_mass(self._global(links_idx_local, self.n_links, self.link_start),
                                                    envs_idx)

    def set_links_inertial_mass(self, mass, links_idx_local=None, envs_idx=None):
        self._solver.set_links_inertial_mass(mass, self._global(links_idx_local, self.n_links, self.link_start),
                                             envs_idx)

    def get_links_invweight(self, links_idx_local=None, envs_idx=None):
        return self._solver.get_links_invweight(self._global(links_idx_local, self.n_links, self.link_start),
                                                envs_idx)

    def get_dofs_invweight(self, dofs_idx_local=None, envs_idx=None):
        return self._solver.get_dofs_invweight(self._dofs(dofs_idx_local), envs_idx)

    # per-env shifts (rigid_entity.py:2852-2935)
    def set_friction_ratio(self, friction_ratio, links_idx_local=None, envs_idx=None):
        """Friction ratio per listed link, applied to every geom of the link (shape (n_envs, n_links))."""
        links = self._global(links_idx_local, self.n_links, self.link_start).tolist()
        gl = self._scene.model.geom_link
        geoms = [[g for g in range(len(gl)) if gl[g] == l] for l in links]
        counts = torch.tensor([len(g) for g in geoms])
        ratio = torch.as_tensor(friction_ratio, dtype=torch.float32).cpu()
        ratio = ratio.reshape(-1, len(links)) if ratio.dim() > 1 or self._solver.n_envs else ratio.reshape(1, -1)
        self._solver.set_geoms_friction_ratio(torch.repeat_interleave(ratio, counts, dim=-1),
                                              [g for gs_ in geoms for g in gs_], envs_idx)

    def set_friction(self, friction):
        """Friction coefficient of every geom of the entity (shared by all envs)."""
        gl = self._scene.model.geom_link
        geoms = [g for g in range(len(gl)) if self.link_start <= gl[g] < self.link_end]
        self._solver.set_geoms_friction(np.full(len(geoms), float(friction)), geoms)

    def set_mass_shift(self, mass_shift, links_idx_local=None, envs_idx=None):
        self._solver.set_links_mass_shift(mass_shift, self._global(links_idx_local, self.n_links, self.link_start),
                                          envs_idx)

    def set_COM_shift(self, com_shift, links_idx_local=None, envs_idx=None):
        self._solver.set_links_COM_shift(com_shift, self._global(links_idx_local, self.n_links, self.link_start),
                                         envs_idx)

    def get_mass_mat(self, envs_idx=None, decompose=False):
        """The entity's block of the mass matrix (rigid_entity.py:2715-2717)."""
        return self._solver.get_mass_mat(self._dofs(None), envs_idx, decompose)

    def get_mass(self):
        """Total mass of the entity's links (rigid_entity.py:2965)."""
        return float(self._scene.model.link_inertial_mass[self.link_start:self.link_end].astype(np.float64).sum())

    def get_contacts(self, with_entity=None, exclude_self_contact=False):
        """Contacts of the last step involving this entity (rigid_entity.py:2756-2836): the collider's padded
        fields, plus `valid_mask` (n_envs > 0) or filtered rows (n_envs == 0), with `force` split into
        force_a = -force and force_b = +force."""
        info = self._solver.collider.get_contacts(as_tensor=True, to_torch=True)
        op = torch.logical_xor if exclude_self_contact else torch.logical_or
        if with_entity is not None and self.idx == with_entity.idx:
            if exclude_self_contact:
                raise GenesisException("`with_entity` is self but `exclude_self_contact` is True.")
            op = torch.logical_and

        def inside(g, ent):
            return (g >= ent.geom_start) & (g < ent.geom_end)

        valid = op(inside(info["geom_a"], self), inside(info["geom_b"], self))
        if with_entity is not None and self.idx != with_entity.idx:
            valid = valid & (inside(info["geom_a"], with_entity) | inside(info["geom_b"], with_entity))
        if self._solver.n_envs == 0:
            info = {k: v[valid] for k, v in info.items()}
        else:
            info["valid_mask"] = valid
        info["force_a"] = -info["force"]
        info["force_b"] = +info["force"]
        del info["force"]
        return info

    def get_links_net_contact_force(self, envs_idx=None):
        """Net contact force on each link of the entity, (n_envs, n_links, 3) or (n_links, 3) when n_envs == 0
        (rigid_entity.py:2838-2850)."""
        return self._solver.get_links_net_contact_force(envs_idx)[..., self.link_start:self.link_end, :]


# ------------------------------------------------------------------------------------------------------------
# scene
# ------------------------------------------------------------------------------------------------------------
class Scene:
    def __init__(self, sim_options: Optional[SimOptions] = None, rigid_options: Optional[RigidOptions] = None,
                 show_viewer: bool = False, **_ignored):
        self.sim_options = sim_options or SimOptions()
        self.rigid_options = rigid_options or RigidOptions()
        self.entities: List[RigidEntity] = []
        self.is_built = False
        self.rigid_solver: Optional[RigidSolver] = None
        self.t = 0

    def add_entity(self, morph, material: Optional[RigidMaterial] = None, **_ignored) -> RigidEntity:
        if self.is_built:
            raise RuntimeError("Scene is already built")
        material = material or RigidMaterial()
        ent = RigidEntity(self, len(self.entities), morph, material, _spec_from_morph(morph, material))
        self.entities.append(ent)
        return ent

    def _compile_options(self) -> CompileOptions:
        so, ro = self.sim_options, self.rigid_options
        dt = (ro.dt if ro.dt is not None else so.dt) / max(1, so.substeps)
        return CompileOptions(
            dt=dt, gravity=tuple(ro.gravity if ro.gravity is not None else so.gravity), iterations=ro.iterations,
            tolerance=ro.tolerance, ls_iterations=ro.ls_iterations, ls_tolerance=ro.ls_tolerance,
            constraint_timeconst=ro.constraint_timeconst, max_collision_pairs=ro.max_collision_pairs,
            multiplier_collision_broad_phase=ro.multiplier_collision_broad_phase,
            enable_collision=ro.enable_collision and not ro.disable_constraint, enable_joint_limit=ro.enable_joint_limit,
            enable_self_collision=ro.enable_self_collision, enable_adjacent_collision=ro.enable_adjacent_collision,
            enable_multi_contact=ro.enable_multi_contact, max_dynamic_constraints=ro.max_dynamic_constraints,
            batch_dofs_info=ro.batch_dofs_info, batch_links_info=ro.batch_links_info,
            batch_joints_info=ro.batch_joints_info)

    def build(self, n_envs: int = 0, device="cuda", solver_variant=None):
        ro = self.rigid_options
        if ro.integrator != "approximate_implicitfast" or ro.constraint_solver != "Newton":
            raise NotImplementedError("only the default approximate_implicitfast integrator and Newton solver are built")
        if ro.enable_mujoco_compatibility:
            raise NotImplementedError("enable_mujoco_compatibility=True is not built")
        self.model = compile_model([e.spec for e in self.entities], self._compile_options())
        self.n_envs = n_envs
        self.rigid_solver = RigidSolver(self.model, n_envs, device, solver_variant=solver_variant)
        for e in self.entities:
            e._bind(self.model)
        self._init_qpos = self.rigid_solver.st.t["qpos"].clone()
        self._init_state = None
        self.is_built = True

    @property
    def envs_idx(self):
        return torch.arange(max(1, self.n_envs))

    def step(self):
        for _ in range(max(1, self.sim_options.substeps)):
            self.rigid_solver.substep()
        self.t += 1

    def get_state(self):
        return self.rigid_solver.get_state(self.t)

    def reset(self, state=None, envs_idx=None):
        """Back to the registered initial state (scene.py:857-885): the build-time state, or the last `state` given,
        which becomes the new initial state; contact caches, the solver warm start and errno are cleared."""
        S = self.rigid_solver
        if state is not None:
            self._init_state = state
        if self._init_state is None:
            e = S._idx(envs_idx, S._B)
            S.st.init_state(e)
            S.st.t["qpos"][:, e] = self._init_qpos[:, e]
            ei = e.to(torch.int32)
            S.reset_caches(ei)
            S.forward_kinematics(ei)
        else:
            S.set_state(self._init_state, envs_idx)
        self.t = 0

    # -- checkpoint (scene.py:1283-1344) ---------------------------------------------------------------------------
    def dump_ckpt_to_numpy(self):
        """Flat {"RigidSolver.<field>": array} of every state array that outlives a substep."""
        return self.rigid_solver.checkpoint_arrays()

    def save_checkpoint(self, path):
        """Write the physics state to one .npz file: the arrays, `step_index` and `timestamp`. The reference pickles
        the same dictionary; an npz holds no code, so loading it runs nothing from the file."""
        arrays = self.dump_ckpt_to_numpy()
        with open(path, "wb") as f:
            np.savez(f, __step_index__=np.int64(self.t), __timestamp__=np.float64(time.time()), **arrays)

    def load_checkpoint(self, path):
        """Restore a file written by save_checkpoint into this built scene (same entities and n_envs)."""
        with np.load(path, allow_pickle=False) as z:
            arrays = {k: z[k] for k in z.files}
        self.rigid_solver.load_checkpoint_arrays(arrays)
        self.t = int(arrays.get("__step_index__", self.t))
