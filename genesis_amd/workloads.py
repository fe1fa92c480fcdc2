"""Seeded synthetic scenes of the benchmark workloads (entity specs + per-env initial states + control schedules).

* arm:       9-DoF Panda arm (7 revolute + 2 prismatic fingers, 11 links, capsule/box collision) on a plane,
             self-collision off, PD targets uniform within joint limits, resampled every 100 steps.
* pile:      free boxes (edges U[0.05, 0.2] m, density 1000 kg/m^3, friction 1.0) dropped in a 1 m column onto a plane.
* quadruped: go2 floating base + 4 legs x 3 revolute joints (18 DoF, 13 bodies) on a plane,
             PD targets = standing pose + U(-0.3, 0.3) per dof per step.
"""
from __future__ import annotations

import copy
import json
import os
from dataclasses import dataclass
from typing import Callable, List, Optional

import numpy as np

from .loaders import entity_from_json
from .model import GEOM_BOX, GEOM_PLANE, JOINT_FIXED, JOINT_FREE, CompileOptions, EntitySpec, GeomSpec, JointSpec, LinkSpec

ASSETS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "assets")

GO2_STANDING = np.array([0.0, 0.0, 0.0, 0.0, 0.8, 0.8, 1.0, 1.0, -1.5, -1.5, -1.5, -1.5])
GO2_INIT_BASE = np.array([0.0, 0.0, 0.42, 1.0, 0.0, 0.0, 0.0])


def plane_entity(normal=(0.0, 0.0, 1.0)) -> EntitySpec:
    """Plane primitive: fixed link, one PLANE geom (rigid_entity.py:155-242)."""
    link = LinkSpec("plane_baselink", -1, np.zeros(3), np.array([1.0, 0, 0, 0]),
                    [JointSpec("plane_baselink_joint", JOINT_FIXED)], [GeomSpec(GEOM_PLANE, data=tuple(normal))])
    return EntitySpec([link], name="plane")


def box_entity(size, pos, quat=(1.0, 0, 0, 0), fixed=False, rho=200.0, friction=None) -> EntitySpec:
    """Box primitive: one link, FREE joint with init qpos = (pos, quat) unless fixed (rigid_entity.py:155-242)."""
    pos = np.asarray(pos, np.float64)
    quat = np.asarray(quat, np.float64)
    j = JointSpec("box_baselink_joint", JOINT_FIXED) if fixed else JointSpec(
        "box_baselink_joint", JOINT_FREE, kp=0.0, kv=0.0, init_qpos=np.concatenate([pos, quat]))
    link = LinkSpec("box_baselink", -1, pos, quat, [j], [GeomSpec(GEOM_BOX, data=tuple(size))])
    return EntitySpec([link], rho=rho, friction=friction, name="box")


def load_asset(name: str) -> EntitySpec:
    with open(os.path.join(ASSETS, f"{name}.json")) as f:
        return entity_from_json(json.load(f))


def panda_entity() -> EntitySpec:
    return load_asset("panda")


def go2_entity() -> EntitySpec:
    ent = load_asset("go2")
    ent = copy.deepcopy(ent)
    # initial base pose (0, 0, 0.42) and the standing joint configuration
    for l in ent.links:
        for j in l.joints:
            if j.type == JOINT_FREE:
                j.init_qpos = GO2_INIT_BASE.copy()
    return ent


def _uniform_quats(rng: np.random.Generator, n):
    u = rng.random((n, 3))
    q = np.stack([np.sqrt(1 - u[:, 0]) * np.sin(2 * np.pi * u[:, 1]), np.sqrt(1 - u[:, 0]) * np.cos(2 * np.pi * u[:, 1]),
                  np.sqrt(u[:, 0]) * np.sin(2 * np.pi * u[:, 2]), np.sqrt(u[:, 0]) * np.cos(2 * np.pi * u[:, 2])], 1)
    return q[:, [3, 0, 1, 2]]


@dataclass
class Workload:
    name: str
    entities: List[EntitySpec]
    options: CompileOptions
    n_envs: int
    n_steps: int
    init_qpos: Optional[Callable] = None  # (n_envs) -> (n_envs, n_qs) float64, or None for qpos0
    control: Optional[Callable] = None  # (step, state) -> None, applies PD targets


def arm_workload(n_envs=64, n_steps=1000, seed=0) -> Workload:
    opts = CompileOptions(dt=0.01, enable_self_collision=False)
    return Workload("arm", [plane_entity(), panda_entity()], opts, n_envs, n_steps)


def quadruped_workload(n_envs=32768, n_steps=1000, seed=2, dt=0.005) -> Workload:
    opts = CompileOptions(dt=dt)
    return Workload("quadruped", [plane_entity(), go2_entity()], opts, n_envs, n_steps)


def pile_entities(n_boxes: int, rng: np.random.Generator, rho=1000.0, friction=1.0):
    ents = [plane_entity()]
    sizes = rng.uniform(0.05, 0.2, (n_boxes, 3))
    for i in range(n_boxes):
        ents.append(box_entity(sizes[i], (0.0, 0.0, 0.5 + i), rho=rho, friction=friction))
    return ents


def _quat_to_mat(q):
    w, x, y, z = q[..., 0], q[..., 1], q[..., 2], q[..., 3]
    return np.stack([np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)], -1),
                     np.stack([2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)], -1),
                     np.stack([2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], -1)], -2)


def boxes_overlap(c1, R1, h1, c2, R2, h2):
    """Separating-axis test for oriented boxes (broadcast over leading dims): centres c (..., 3), rotations R
    (..., 3, 3) whose columns are the box axes, half extents h (..., 3). True where no axis among the 6 face normals
    and the 9 edge cross products separates them."""
    A = np.swapaxes(R1, -1, -2)  # (..., 3 axes, 3)
    B = np.swapaxes(R2, -1, -2)
    A, B = np.broadcast_arrays(A, B)
    cr = np.cross(A[..., :, None, :], B[..., None, :, :]).reshape(*A.shape[:-2], 9, 3)
    n = np.linalg.norm(cr, axis=-1, keepdims=True)
    cr = np.where(n > 1e-9, cr / np.maximum(n, 1e-30), 0.0)  # parallel edges: a zero axis never separates
    L = np.concatenate([A, B, cr], -2)  # (..., 15, 3)
    T = (c2 - c1)[..., None, :]
    r1 = (np.abs(L @ np.swapaxes(A, -1, -2)) * h1[..., None, :]).sum(-1)
    r2 = (np.abs(L @ np.swapaxes(B, -1, -2)) * h2[..., None, :]).sum(-1)
    sep = np.abs((L * T).sum(-1)) > r1 + r2
    return ~sep.any(-1)


PILE_BLOCK = 256  # envs per independently seeded block of initial poses


def pile_poses(n_envs, sizes, seed, lo=(-0.5, -0.5, 0.1), hi=(0.5, 0.5, 1.1), gap=0.005):
    """Seeded random poses of boxes `sizes` (n_boxes, 3) inside the box [lo, hi] (1 m^3 by default) for n_envs envs,
    drawn by rejection: uniform position and uniform orientation per box, kept when the box (grown by `gap`) overlaps
    none of the boxes placed before it and stays above the plane z = 0. Env i depends only on (seed, i): poses come
    in blocks of PILE_BLOCK envs, block k drawn from default_rng((seed, k)). Returns pos (n, b, 3), quat (n, b, 4)."""
    sizes = np.asarray(sizes, np.float64)
    nb = len(sizes)
    h = sizes / 2
    rad = np.linalg.norm(h, axis=-1)
    lo, hi = np.asarray(lo, np.float64), np.asarray(hi, np.float64)
    n_blocks = -(-n_envs // PILE_BLOCK)
    N = n_blocks * PILE_BLOCK
    rngs = [np.random.default_rng((seed, k)) for k in range(n_blocks)]
    pos = np.zeros((N, nb, 3))
    quat = np.zeros((N, nb, 4))
    R = np.zeros((N, nb, 3, 3))
    for i in range(nb):
        todo = np.arange(N)
        for _ in range(10000):
            # each block draws its own candidates, in env order, from its own generator
            blk = todo // PILE_BLOCK
            cnt = np.bincount(blk, minlength=n_blocks)
            p = np.concatenate([rngs[k].uniform(lo, hi, (cnt[k], 3)) for k in range(n_blocks) if cnt[k]])
            q = np.concatenate([_uniform_quats(rngs[k], cnt[k]) for k in range(n_blocks) if cnt[k]])
            Rm = _quat_to_mat(q)
            ok = p[:, 2] - (np.abs(Rm[:, 2, :]) * h[i]).sum(-1) > gap
            if i > 0:
                # separating-axis test only where the bounding spheres meet
                d = np.linalg.norm(pos[todo, :i] - p[:, None], axis=-1)
                ii, jj = np.nonzero(d < rad[i] + rad[:i] + 2 * gap)
                if ii.size:
                    ov = boxes_overlap(p[ii], Rm[ii], h[i] + gap, pos[todo[ii], jj], R[todo[ii], jj], h[jj] + gap)
                    bad = np.zeros(todo.size, bool)
                    bad[ii[ov]] = True
                    ok &= ~bad
            acc = todo[ok]
            pos[acc, i], quat[acc, i], R[acc, i] = p[ok], q[ok], Rm[ok]
            todo = todo[~ok]
            if todo.size == 0:
                break
        if todo.size:
            raise RuntimeError(f"no room for box {i} in {todo.size} envs")
    return pos[:n_envs], quat[:n_envs]


def pile_workload(n_boxes=64, n_envs=4096, n_steps=500, seed=1, max_collision_pairs=1000) -> Workload:
    """Box sizes are shared by every env of one scene type; poses differ per env (seeded, non-overlapping, in a
    1 m^3 column above the plane)."""
    opts = CompileOptions(dt=0.01, max_collision_pairs=max_collision_pairs)
    ents = pile_entities(n_boxes, np.random.default_rng(seed))
    sizes = np.array([e.links[0].geoms[0].data[:3] for e in ents[1:]], np.float64)
    cache = {}

    def init_qpos(n, n_qs):
        if cache.get("n", 0) < n:
            cache["pos"], cache["quat"] = pile_poses(n, sizes, seed + 1)
            cache["n"] = n
        q = np.concatenate([cache["pos"][:n], cache["quat"][:n]], -1).reshape(n, n_boxes * 7)
        assert q.shape[1] == n_qs
        return q

    return Workload("pile", ents, opts, n_envs, n_steps, init_qpos=init_qpos)


MIXED_TYPES = ("arm", "quadruped", "pile20")


def mixed_workloads(n_envs=16384, seed=3, dt=0.01):
    """Heterogeneous batch: a seeded multinomial split of n_envs among the arm, the quadruped and a 20-box pile,
    all at the same dt. Returns [(Workload, count)] in MIXED_TYPES order."""
    import torch
    g = torch.Generator().manual_seed(seed)
    pick = torch.multinomial(torch.ones(3), n_envs, replacement=True, generator=g)
    counts = torch.bincount(pick, minlength=3).tolist()
    ws = [arm_workload(n_envs=counts[0]), quadruped_workload(n_envs=counts[1], dt=dt),
          pile_workload(n_boxes=20, n_envs=counts[2])]
    ws[2].name = "pile20"
    return list(zip(ws, counts))


def workload_by_name(name: str, n_envs: int, dt: float = None) -> Workload:
    if name == "arm":
        return arm_workload(n_envs=n_envs)
    if name == "quadruped":
        return quadruped_workload(n_envs=n_envs, **({} if dt is None else {"dt": dt}))
    if name == "pile":
        return pile_workload(n_envs=n_envs)
    if name == "pile20":
        w = pile_workload(n_boxes=20, n_envs=n_envs)
        w.name = "pile20"
        return w
    raise ValueError(f"unknown workload {name}")


class Control:
    """Per-step PD targets of a workload for envs [env_begin, env_begin + n_envs) of a solver, drawn on the device
    from a seeded generator: arm targets U(joint limits) resampled every 100 steps; quadruped targets = standing
    pose + U(-0.3, 0.3) every step; piles are uncontrolled."""

    def __init__(self, w: Workload, M, S, seed, env_begin=0, n_envs=None):
        import torch
        self.w, self.M, self.S = w, M, S
        self.b0 = env_begin
        self.n = S._B - env_begin if n_envs is None else n_envs
        self.envs = torch.arange(self.b0, self.b0 + self.n, device=S.device)
        self.gen = torch.Generator(device=S.device).manual_seed(seed)
        self.n_ctrl = 0
        if w.name == "arm":
            self.lo = torch.as_tensor(M.dof_limit[:, 0], device=S.device)
            self.hi = torch.as_tensor(M.dof_limit[:, 1], device=S.device)
            self.n_ctrl = M.n_dofs
        elif w.name == "quadruped":
            self.stand = torch.as_tensor(GO2_STANDING, dtype=torch.float32, device=S.device)
            self.dofs = torch.arange(6, 18, device=S.device)
            self.n_ctrl = 12

    def __call__(self, step):
        import torch
        if self.w.name == "arm" and step % 100 == 0:
            u = torch.rand((self.n, self.M.n_dofs), generator=self.gen, device=self.S.device)
            self.S.control_dofs_position(self.lo + (self.hi - self.lo) * u, envs_idx=self.envs)
        elif self.w.name == "quadruped":
            u = torch.rand((self.n, 12), generator=self.gen, device=self.S.device)
            self.S.control_dofs_position(self.stand + (u * 0.6 - 0.3), dofs_idx=self.dofs, envs_idx=self.envs)


def init_state(w: Workload, M, S, env_begin=0, n_envs=None):
    """Fresh state for envs [env_begin, env_begin + n_envs): zero velocities, force control, neutral shifts, cleared
    caches and errno, and the workload's initial poses (qpos0 unless the workload says). Nothing of a previous
    occupant of those envs survives (a reused serving slot starts like a new env)."""
    import torch
    n = S._B - env_begin if n_envs is None else n_envs
    envs = torch.arange(env_begin, env_begin + n, device=S.device)
    S.st.init_state(envs)
    if w.init_qpos is not None:
        S.set_qpos(torch.as_tensor(w.init_qpos(n, M.n_qs), dtype=torch.float32), envs_idx=envs)
    elif w.name == "quadruped":
        q = np.tile(np.concatenate([GO2_INIT_BASE, GO2_STANDING]), (n, 1))
        S.set_qpos(torch.as_tensor(q, dtype=torch.float32), envs_idx=envs)
    else:
        S.set_qpos(torch.as_tensor(np.tile(M.qpos0, (n, 1)), dtype=torch.float32), envs_idx=envs)


# errno bits the reference raises on (rigid_solver_decomp.py:941-956): broad-phase or contact overflow, NaN
FATAL_ERRNO = 0xF


def reset_envs(w: Workload, M, S, envs):
    """Env i of the listed envs (solver indices) gets the fresh state init_state gives env i at the start of the
    workload: its own initial pose, zero velocities, force control, cleared caches and errno."""
    import torch
    envs = torch.as_tensor(envs, device=S.device, dtype=torch.long).reshape(-1)
    if envs.numel() == 0:
        return
    S.st.init_state(envs)
    idx = envs.cpu().numpy()
    if w.init_qpos is not None:
        q = np.asarray(w.init_qpos(int(idx.max()) + 1, M.n_qs))[idx]
    elif w.name == "quadruped":
        q = np.tile(np.concatenate([GO2_INIT_BASE, GO2_STANDING]), (len(idx), 1))
    else:
        q = np.tile(M.qpos0, (len(idx), 1))
    S.set_qpos(torch.as_tensor(q, dtype=torch.float32), envs_idx=envs)


def reset_failed_envs(w: Workload, M, S) -> int:
    """The episode-termination rule of the benchmark loops: envs whose last steps set an errno bit the reference
    raises on (a diverged env: NaN accelerations or state, or a capacity overflow) restart from their initial state,
    as a reinforcement-learning loop resets a terminated env, instead of stopping every other env. Returns how many
    envs were reset (one host synchronisation)."""
    import torch
    bad = torch.nonzero((S.st.t["errno_env"][0] & FATAL_ERRNO) != 0).flatten()
    n = int(bad.numel())
    if n:
        reset_envs(w, M, S, bad)
    return n
