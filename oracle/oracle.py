"""ctypes wrapper of the CPU oracle (oracle/oracle.cpp). TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this module, and only as the checker
or the reported CPU baseline; the product path (genesis_amd/) never loads it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")
_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.orc_new.restype = ctypes.c_void_p
        L.orc_free.argtypes = [ctypes.c_void_p]
        L.orc_set_i.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_void_p, ctypes.c_int64]
        L.orc_set_f.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_void_p, ctypes.c_int64]
        L.orc_build.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        L.orc_env_io.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
        L.orc_env_io.restype = ctypes.c_int
        L.orc_run.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.orc_broadphase_from_aabbs.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_broadphase_from_aabbs.restype = ctypes.c_int
        L.orc_support_mesh.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
        L.orc_support_mesh.restype = ctypes.c_int
        _lib = L
    return _lib


INT_TABLES = [
    "link_parent", "link_root", "link_q_start", "link_q_end", "link_dof_start", "link_dof_end", "link_joint_start",
    "link_joint_end", "link_is_fixed", "link_entity", "joint_type", "joint_q_start", "joint_dof_start", "joint_dof_end",
    "dof_entity", "ent_link_start", "ent_link_end", "ent_dof_start", "ent_dof_end", "geom_type", "geom_link",
    "geom_is_fixed", "geom_vert_start", "geom_vert_end", "eq_type", "eq_obj1", "eq_obj2", "pair_idx",
    "support_cell_start", "support_vid",
]
FLOAT_TABLES = [
    "link_pos", "link_quat", "link_inertial_pos", "link_inertial_quat", "link_inertial_i", "link_inertial_mass",
    "link_invweight", "joint_pos", "joint_sol_params", "dof_motion_ang", "dof_motion_vel", "dof_limit",
    "dof_force_range", "dof_kp", "dof_kv", "dof_damping", "dof_armature", "dof_stiffness", "dof_frictionloss",
    "dof_invweight", "qpos0", "ent_gravity_compensation", "geom_pos", "geom_quat", "geom_center", "geom_data",
    "geom_friction", "geom_sol_params", "geom_init_aabb", "verts", "eq_data", "eq_sol_params", "mass_parent_mask",
    "gravity", "support_v",
]

# state field -> (kind, per-env shape builder)
FIELDS_F = {"qpos": "Q", "vel": "D", "qacc_ws": "D", "ctrl_pos": "D", "ctrl_vel": "D", "ctrl_force": "D", "acc": "D",
            "acc_smooth": "D", "dofs_pos": "D", "sort_value": "2G", "mass_shift": "L", "friction_ratio": "G",
            "eq_dyn_data": "11K"}
FIELDS_I = {"ctrl_mode": "D", "sort_ig": "2G", "sort_ismax": "2G", "eq_dyn_obj": "2K"}
SCALAR_I = ["is_warmstart", "first_time", "errno", "n_contacts", "n_broad", "n_con", "n_newton_iters", "n_con_eq",
            "n_con_fl", "eq_dyn_n"]
ROW_F = ["efc_D", "efc_aref", "efc_diag", "efc_fl", "efc_force"]


class Oracle:
    """The CPU oracle loaded with one compiled model and B env states (float32 or float64 arithmetic)."""

    def __init__(self, model, n_envs: int, precision: int = 32):
        self.m = model
        self.B = n_envs
        self.precision = precision
        L = lib()
        self.h = ctypes.c_void_p(L.orc_new())
        self._keep = []
        for k in INT_TABLES:
            a = np.ascontiguousarray(np.asarray(getattr(model, k)), dtype=np.int32).ravel()
            self._keep.append(a)
            L.orc_set_i(self.h, k.encode(), a.ctypes.data, a.size)
        for k in FLOAT_TABLES:
            a = np.ascontiguousarray(np.asarray(getattr(model, k), dtype=np.float32).astype(np.float64)).ravel()
            self._keep.append(a)
            L.orc_set_f(self.h, k.encode(), a.ctypes.data, a.size)
        o = model.options
        ints = dict(n_links=model.n_links, n_joints=model.n_joints, n_dofs=model.n_dofs, n_qs=model.n_qs,
                    n_geoms=model.n_geoms, n_entities=model.n_entities, n_equalities=model.n_equalities,
                    n_pairs=model.n_pairs, max_contacts=model.max_contacts, max_broad=model.max_broad,
                    max_rows=model.max_rows, max_dyn_welds=model.max_dyn_welds, iterations=o.iterations,
                    ls_iterations=o.ls_iterations,
                    n_contacts_per_pair=o.n_contacts_per_pair, enable_collision=int(o.enable_collision),
                    enable_joint_limit=int(o.enable_joint_limit), enable_multi_contact=int(o.enable_multi_contact))
        f32 = lambda x: float(np.float32(x))  # noqa: E731
        flts = dict(dt=f32(o.dt), tolerance=f32(o.tolerance), ls_tolerance=f32(o.ls_tolerance),
                    EPS=float(np.finfo(np.float32 if precision == 32 else np.float64).eps),
                    meaninertia=float(model.meaninertia), mc_perturbation=f32(1e-2), mc_tolerance=f32(1e-2),
                    mpr_to_gjk_overlap_ratio=f32(0.25))
        for k, v in ints.items():
            a = np.array([v], np.int32)
            L.orc_set_i(self.h, k.encode(), a.ctypes.data, 1)
        for k, v in flts.items():
            a = np.array([v], np.float64)
            L.orc_set_f(self.h, k.encode(), a.ctypes.data, 1)
        L.orc_build(self.h, n_envs, precision)
        self.sizes = dict(Q=model.n_qs, D=model.n_dofs, G=model.n_geoms, L=model.n_links)
        # initial state: qpos0, zero velocity, FORCE control
        self.set("qpos", np.tile(model.qpos0.astype(np.float64), (n_envs, 1)))

    def __del__(self):
        try:
            lib().orc_free(self.h)
        except Exception:
            pass

    def _n(self, key):
        if key.endswith("K"):
            return int(key[:-1]) * self.m.max_dyn_welds
        return 2 * self.sizes["G"] if key == "2G" else self.sizes[key]

    def set(self, name, arr, envs=None):
        envs = range(self.B) if envs is None else envs
        L = lib()
        for i, b in enumerate(envs):
            row = np.asarray(arr)[i]
            if name in FIELDS_I or name in SCALAR_I:
                a = np.ascontiguousarray(np.atleast_1d(row), dtype=np.int32)
                L.orc_env_io(self.h, b, name.encode(), None, a.ctypes.data, 1)
            else:
                a = np.ascontiguousarray(np.atleast_1d(row), dtype=np.float64)
                L.orc_env_io(self.h, b, name.encode(), a.ctypes.data, None, 1)

    def get(self, name, envs=None):
        envs = list(range(self.B)) if envs is None else list(envs)
        L = lib()
        m = self.m
        shapes = {"links_pos": (m.n_links, 3), "i_pos_shift": (m.n_links, 3), "links_quat": (m.n_links, 4), "geoms_pos": (m.n_geoms, 3),
                  "geoms_quat": (m.n_geoms, 4), "aabb_min": (m.n_geoms, 3), "aabb_max": (m.n_geoms, 3),
                  "cdof_ang": (m.n_dofs, 3), "cdof_vel": (m.n_dofs, 3), "cd_vel": (m.n_links, 3),
                  "cd_ang": (m.n_links, 3), "root_com": (m.n_links, 3), "mass_mat": (m.n_dofs, m.n_dofs),
                  "cache_normal": (max(m.n_pairs, 1), 3)}
        out = []
        for b in envs:
            if name in FIELDS_F:
                a = np.zeros(self._n(FIELDS_F[name]), np.float64)
                L.orc_env_io(self.h, b, name.encode(), a.ctypes.data, None, 0)
            elif name in FIELDS_I:
                a = np.zeros(self._n(FIELDS_I[name]), np.int32)
                L.orc_env_io(self.h, b, name.encode(), None, a.ctypes.data, 0)
            elif name in SCALAR_I:
                a = np.zeros(1, np.int32)
                L.orc_env_io(self.h, b, name.encode(), None, a.ctypes.data, 0)
                a = a[0]
            elif name == "broad_pairs":
                n = np.zeros(1, np.int32)
                L.orc_env_io(self.h, b, b"n_broad", None, n.ctypes.data, 0)
                a = np.zeros((max(int(n[0]), 1), 2), np.int32)
                L.orc_env_io(self.h, b, b"broad_pairs", None, a.ctypes.data, 0)
                a = a[: int(n[0])]
            elif name in ROW_F or name == "efc_jac":
                n = np.zeros(1, np.int32)
                L.orc_env_io(self.h, b, b"n_con", None, n.ctypes.data, 0)
                shape = (int(n[0]), m.n_dofs) if name == "efc_jac" else (int(n[0]),)
                a = np.zeros(max(int(np.prod(shape)), 1), np.float64)
                L.orc_env_io(self.h, b, name.encode(), a.ctypes.data, None, 0)
                a = a[: int(np.prod(shape))].reshape(shape)
            elif name in ("qf_smooth", "qacc"):
                a = np.zeros(m.n_dofs, np.float64)
                L.orc_env_io(self.h, b, name.encode(), a.ctypes.data, None, 0)
            elif name == "contacts":
                n = np.zeros(1, np.int32)
                L.orc_env_io(self.h, b, b"n_contacts", None, n.ctypes.data, 0)
                a = np.zeros((max(int(n[0]), 1), 16), np.float64)
                L.orc_env_io(self.h, b, b"contacts", a.ctypes.data, None, 0)
                a = a[: int(n[0])]
            else:
                a = np.zeros(shapes[name], np.float64)
                L.orc_env_io(self.h, b, name.encode(), a.ctypes.data, None, 0)
            out.append(a)
        if name in ("broad_pairs", "contacts", "efc_jac") or name in ROW_F:
            return out
        return np.stack(out)

    def run(self, n_steps=1, stage=0, envs=None, n_threads=1):
        b0, b1 = (0, self.B) if envs is None else envs
        lib().orc_run(self.h, b0, b1, n_steps, stage, n_threads)

    def support_mesh(self, g, d):
        """Support-field lookup of mesh geom g for a mesh-frame direction d: (vertex, its index in the geom)."""
        dv = np.ascontiguousarray(d, np.float64)
        out = np.zeros(3, np.float64)
        vid = lib().orc_support_mesh(self.h, g, dv.ctypes.data, out.ctypes.data)
        return out, vid

    def broadphase_from_aabbs(self, b, aabb_min, aabb_max):
        amin = np.ascontiguousarray(aabb_min, np.float64).ravel()
        amax = np.ascontiguousarray(aabb_max, np.float64).ravel()
        out = np.zeros((max(self.m.max_broad, 1), 2), np.int32)
        n = lib().orc_broadphase_from_aabbs(self.h, b, amin.ctypes.data, amax.ctypes.data, out.ctypes.data)
        return out[:n]


def links_net_contact_force(link_a, link_b, force, n_links):
    """Net contact force per link of one env (constraint_solver_decomp.py:1350-1379): start from zero and, contact
    by contact in list order, subtract the contact's force from link_a and add it to link_b. Accumulates in
    force's dtype, so float32 inputs give the reference's float32 sums."""
    force = np.asarray(force)
    out = np.zeros((n_links, 3), force.dtype)
    for la, lb, f in zip(np.asarray(link_a).tolist(), np.asarray(link_b).tolist(), force):
        out[la] = out[la] - f
        out[lb] = out[lb] + f
    return out
