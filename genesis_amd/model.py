"""Model compiler: morphs -> flat, env-shared tables for the batched rigid step.

The tables mirror the reference's static "info" structures (links/joints/dofs/geoms/entities/equalities,
genesis/utils/array_class.py:1141-1527) and the build rules of genesis/engine/entities/rigid_entity
(_load_primitive rigid_entity.py:155-242, _add_by_info rigid_entity.py:659-848, link inertia
rigid_link.py:125-188, geom init AABB rigid_solver_decomp.py:3206-3221) and of the collider
(_compute_collision_pair_idx collider_decomp.py:184-264, capacities collider_decomp.py:296-304) and the
constraint solver (row capacity constraint_solver_decomp.py:42-48).

Everything here runs once per model on the host in float64 and is then frozen to float32/int32 arrays that
the HIP library and the CPU oracle both consume unchanged.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

# ---------------------------------------------------------------------------------------------------------
# Enumerations (values follow genesis/constants.py:19-64)
# ---------------------------------------------------------------------------------------------------------
GEOM_PLANE, GEOM_SPHERE, GEOM_ELLIPSOID, GEOM_CYLINDER, GEOM_CAPSULE, GEOM_BOX, GEOM_MESH, GEOM_TERRAIN = range(8)
JOINT_FIXED, JOINT_REVOLUTE, JOINT_PRISMATIC, JOINT_SPHERICAL, JOINT_FREE = range(5)
EQ_CONNECT, EQ_WELD, EQ_JOINT = range(3)
CTRL_FORCE, CTRL_VELOCITY, CTRL_POSITION = range(3)

EPS32 = float(np.finfo(np.float32).eps)  # genesis/__init__.py:206 (max(1e-15, finfo.eps))
IMP_MIN, IMP_MAX = 1e-4, 0.9999  # rigid_solver_decomp.py:40-43
TIME_CONSTANT_SAFETY_FACTOR = 2.0  # rigid_solver_decomp.py:46


def default_solver_params():
    """(timeconst, dampratio, dmin, dmax, width, mid, power), genesis/utils/geom.py:1708-1714."""
    return np.array([0.0, 1.0, 0.9, 0.95, 1e-3, 0.5, 2.0])


def sanitize_sol_params(sp, min_timeconst, default_timeconst):
    """Clamp solver parameters like rigid_solver_decomp.py:49-77."""
    sp = np.array(sp, dtype=np.float64).reshape(-1, 7).copy()
    tc = sp[:, 0]
    tc[tc < EPS32] = default_timeconst
    sp[:, 0] = np.maximum(tc, min_timeconst)
    sp[:, 1] = np.maximum(sp[:, 1], 0.0)
    sp[:, 2] = np.clip(sp[:, 2], IMP_MIN, IMP_MAX)
    sp[:, 3] = np.clip(sp[:, 3], IMP_MIN, IMP_MAX)
    sp[:, 5] = np.clip(sp[:, 5], IMP_MIN, IMP_MAX)
    sp[:, 4] = np.maximum(sp[:, 4], 0.0)
    sp[:, 6] = np.maximum(sp[:, 6], 1.0)
    return sp


# ---------------------------------------------------------------------------------------------------------
# Small float64 rotation helpers (host side only)
# ---------------------------------------------------------------------------------------------------------
def quat_mul(u, v):
    w1, x1, y1, z1 = u
    w2, x2, y2, z2 = v
    return np.array([
        w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2,
        w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
        w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2,
        w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2,
    ])


def quat_to_R(q):
    q = np.asarray(q, dtype=np.float64)
    d = q @ q
    if d < 1e-30:
        return np.eye(3)
    s = 2.0 / d
    w, x, y, z = q
    return np.array([
        [1 - s * (y * y + z * z), s * (x * y - w * z), s * (x * z + w * y)],
        [s * (x * y + w * z), 1 - s * (x * x + z * z), s * (y * z - w * x)],
        [s * (x * z - w * y), s * (y * z + w * x), 1 - s * (x * x + y * y)],
    ])


def R_to_quat(R):
    R = np.asarray(R, dtype=np.float64)
    tr = np.trace(R)
    if tr > 0:
        s = math.sqrt(tr + 1.0) * 2
        q = [0.25 * s, (R[2, 1] - R[1, 2]) / s, (R[0, 2] - R[2, 0]) / s, (R[1, 0] - R[0, 1]) / s]
    elif R[0, 0] > R[1, 1] and R[0, 0] > R[2, 2]:
        s = math.sqrt(1.0 + R[0, 0] - R[1, 1] - R[2, 2]) * 2
        q = [(R[2, 1] - R[1, 2]) / s, 0.25 * s, (R[0, 1] + R[1, 0]) / s, (R[0, 2] + R[2, 0]) / s]
    elif R[1, 1] > R[2, 2]:
        s = math.sqrt(1.0 + R[1, 1] - R[0, 0] - R[2, 2]) * 2
        q = [(R[0, 2] - R[2, 0]) / s, (R[0, 1] + R[1, 0]) / s, 0.25 * s, (R[1, 2] + R[2, 1]) / s]
    else:
        s = math.sqrt(1.0 + R[2, 2] - R[0, 0] - R[1, 1]) * 2
        q = [(R[1, 0] - R[0, 1]) / s, (R[0, 2] + R[2, 0]) / s, (R[1, 2] + R[2, 1]) / s, 0.25 * s]
    q = np.array(q)
    if q[0] < 0:
        q = -q
    return q / np.linalg.norm(q)


def rpy_to_quat(rpy):
    r, p, y = rpy
    qx = np.array([math.cos(r / 2), math.sin(r / 2), 0, 0])
    qy = np.array([math.cos(p / 2), 0, math.sin(p / 2), 0])
    qz = np.array([math.cos(y / 2), 0, 0, math.sin(y / 2)])
    return quat_mul(qz, quat_mul(qy, qx))


def transform_by_quat(v, q):
    return quat_to_R(q) @ np.asarray(v, dtype=np.float64)


# ---------------------------------------------------------------------------------------------------------
# Specs (what a morph expands into before flattening)
# ---------------------------------------------------------------------------------------------------------
@dataclass
class GeomSpec:
    type: int
    pos: np.ndarray = field(default_factory=lambda: np.zeros(3))
    quat: np.ndarray = field(default_factory=lambda: np.array([1.0, 0, 0, 0]))
    data: Sequence[float] = ()  # box: full extents; sphere: radius; capsule: (radius, length); plane: normal
    friction: Optional[float] = None
    sol_params: Optional[np.ndarray] = None
    contype: int = 1
    conaffinity: int = 1
    verts: Optional[np.ndarray] = None  # mesh: vertices of the convex collision hull, geom frame


@dataclass
class JointSpec:
    name: str
    type: int
    pos: np.ndarray = field(default_factory=lambda: np.zeros(3))
    axis: np.ndarray = field(default_factory=lambda: np.array([0.0, 0.0, 1.0]))
    limit: Optional[np.ndarray] = None  # (n_dofs, 2)
    damping: float = 0.0
    armature: float = 0.0
    stiffness: float = 0.0
    frictionloss: float = 0.0
    kp: Optional[float] = None
    kv: Optional[float] = None
    force_range: Optional[np.ndarray] = None  # (2,)
    sol_params: Optional[np.ndarray] = None
    init_qpos: Optional[np.ndarray] = None


@dataclass
class LinkSpec:
    name: str
    parent: int  # index inside the entity, -1 for root
    pos: np.ndarray
    quat: np.ndarray
    joints: List[JointSpec]
    geoms: List[GeomSpec]
    inertial_pos: Optional[np.ndarray] = None
    inertial_quat: Optional[np.ndarray] = None
    inertial_i: Optional[np.ndarray] = None
    inertial_mass: Optional[float] = None


@dataclass
class EqualitySpec:
    """JOINT: obj1 / obj2 joint names, data = the coupling polynomial's 5 coefficients. CONNECT: link names, data[0:3]
    = the anchor in obj1's frame (its obj2-frame position is taken from the build pose). WELD: link names; the
    anchors and relative pose come from the build pose unless explicit=True, when data holds all 11 values
    (anchor on obj2, anchor on obj1, relative quaternion, torque scale). A build-pose weld is welded at obj1's origin
    (add_weld_constraint's choice), or at anchor2 on obj2 when given (MJCF's weld anchor), with torque scale
    torquescale."""
    type: int
    obj1: str
    obj2: str
    data: np.ndarray = field(default_factory=lambda: np.array([0.0, 1.0, 0, 0, 0, 0, 0, 0, 0, 0, 0]))
    sol_params: Optional[np.ndarray] = None
    explicit: bool = False
    anchor2: Optional[np.ndarray] = None
    torquescale: float = 1.0


@dataclass
class EntitySpec:
    links: List[LinkSpec]
    equalities: List[EqualitySpec] = field(default_factory=list)
    rho: float = 200.0  # genesis/engine/materials/rigid.py:44
    friction: Optional[float] = None
    name: str = "entity"
    local_collision_mask: bool = False  # True for MJCF entities (rigid_entity.py:152)


# ---------------------------------------------------------------------------------------------------------
# Mesh helpers for primitive geometry (vertex order matters: multi-contact walks vertices in order)
# ---------------------------------------------------------------------------------------------------------
def box_vertices(extents):
    """8 box corners; corner k sits at ((k>>2)&1, (k>>1)&1, k&1) - 0.5, times the extents.

    This is the corner order of the unit box mesh the reference builds for Box morphs
    (genesis/utils/mesh.py:814-820), which func_plane_box_contact walks in order (collider_decomp.py:2104-2110)."""
    k = np.arange(8)
    v = np.stack([(k >> 2) & 1, (k >> 1) & 1, k & 1], axis=1).astype(np.float64) - 0.5
    return v * np.asarray(extents, dtype=np.float64)


def sphere_vertices(radius, n_lat=8, n_lon=16):
    """Sample points on the sphere surface (only used for the init AABB and the vertex centre)."""
    pts = [[0, 0, radius], [0, 0, -radius]]
    for i in range(1, n_lat):
        th = math.pi * i / n_lat
        for j in range(n_lon):
            ph = 2 * math.pi * j / n_lon
            pts.append([radius * math.sin(th) * math.cos(ph), radius * math.sin(th) * math.sin(ph), radius * math.cos(th)])
    pts = np.array(pts)
    # make the extreme points exact so that the AABB equals the analytic one
    pts = np.concatenate([pts, np.array([[radius, 0, 0], [-radius, 0, 0], [0, radius, 0], [0, -radius, 0]])])
    return pts


def capsule_vertices(radius, length, n=16):
    h = 0.5 * length
    pts = []
    for zc, sgn in ((h, 1.0), (-h, -1.0)):
        pts.append([0, 0, zc + sgn * radius])
        for j in range(n):
            ph = 2 * math.pi * j / n
            pts.append([radius * math.cos(ph), radius * math.sin(ph), zc])
    pts.extend([[radius, 0, 0], [-radius, 0, 0], [0, radius, 0], [0, -radius, 0]])
    return np.array(pts, dtype=np.float64)


def plane_vertices(plane_size=(1e3, 1e3), thickness=1e-2):
    """Collision box of the plane: (size_x, size_y, thickness) shifted down by thickness/2 (mesh.py:894-898)."""
    v = box_vertices([plane_size[0], plane_size[1], thickness])
    v[:, 2] -= thickness / 2
    return v


def geom_mass_properties(g: GeomSpec, rho: float):
    """(mass, com_local, inertia_local) of a solid primitive of density rho."""
    t = g.type
    if t == GEOM_BOX:
        a, b, c = g.data[:3]
        m = rho * a * b * c
        I = m / 12.0 * np.diag([b * b + c * c, a * a + c * c, a * a + b * b])
        return m, np.zeros(3), I
    if t == GEOM_SPHERE:
        r = g.data[0]
        m = rho * 4.0 / 3.0 * math.pi * r ** 3
        return m, np.zeros(3), np.eye(3) * 0.4 * m * r * r
    if t == GEOM_CAPSULE:
        r, L = g.data[0], g.data[1]
        mc = rho * math.pi * r * r * L
        ms = rho * 4.0 / 3.0 * math.pi * r ** 3
        Izz = 0.5 * mc * r * r + 0.4 * ms * r * r
        Ixx = mc * (3 * r * r + L * L) / 12.0 + ms * (0.4 * r * r + L * L / 4.0 + 3.0 * L * r / 8.0)
        return mc + ms, np.zeros(3), np.diag([Ixx, Ixx, Izz])
    if t == GEOM_MESH:
        from .meshes import hull_mass_properties
        return hull_mass_properties(g.verts, rho)
    raise ValueError(f"no mass model for geom type {t}")


def _compose(m1, c1, I1, m2, c2, I2):
    """Parallel-axis composition (genesis/utils/urdf.py:422-446)."""
    m = m1 + m2
    c = (m1 * c1 + m2 * c2) / m

    def shift(I, mm, d):
        return I + mm * (np.dot(d, d) * np.eye(3) - np.outer(d, d))

    return m, c, shift(I1, m1, c - c1) + shift(I2, m2, c - c2)


# ---------------------------------------------------------------------------------------------------------
# Compiled model
# ---------------------------------------------------------------------------------------------------------
class RigidModel:
    """Flat tables of one scene type. Float arrays are float32, index arrays int32."""

    def __init__(self):
        pass

    def summary(self):
        return (f"links={self.n_links} joints={self.n_joints} dofs={self.n_dofs} qs={self.n_qs} geoms={self.n_geoms} "
                f"entities={self.n_entities} pairs={self.n_pairs} max_contacts={self.max_contacts} rows={self.max_rows}")


@dataclass
class CompileOptions:
    dt: float = 0.01
    gravity: tuple = (0.0, 0.0, -9.81)
    iterations: int = 50
    tolerance: float = 1e-6
    ls_iterations: int = 50
    ls_tolerance: float = 1e-2
    constraint_timeconst: float = 0.01
    max_collision_pairs: int = 150
    multiplier_collision_broad_phase: int = 8
    enable_collision: bool = True
    enable_joint_limit: bool = True
    enable_self_collision: bool = True
    enable_adjacent_collision: bool = False
    enable_multi_contact: bool = True
    max_dynamic_constraints: int = 8
    n_contacts_per_pair: int = 5
    batch_dofs_info: bool = False  # per-env kp, kv, force range, stiffness, damping, frictionloss, limits, armature
    batch_links_info: bool = False  # per-env link masses and link inverse weights
    batch_joints_info: bool = False  # per-env joint solver parameters


def compile_model(entities: List[EntitySpec], opt: CompileOptions) -> RigidModel:
    M = RigidModel()
    M.options = opt
    substep_dt = opt.dt
    min_tc = TIME_CONSTANT_SAFETY_FACTOR * substep_dt
    default_tc = max(opt.constraint_timeconst, min_tc)

    links, joints, geoms, eqs = [], [], [], []
    ent_rows = []
    q0 = []
    n_q = n_d = 0
    for i_e, ent in enumerate(entities):
        link_start, dof_start_e, geom_start_e = len(links), n_d, len(geoms)
        name_to_joint = {}
        for i_l_local, ls in enumerate(ent.links):
            i_l = len(links)
            parent = ls.parent + link_start if ls.parent >= 0 else -1
            # top-most ancestor (rigid_link.py:69-78)
            r = i_l_local
            while ent.links[r].parent >= 0:
                r = ent.links[r].parent
            root = r + link_start
            joint_start = len(joints)
            lq_start, ld_start = n_q, n_d
            is_fixed = True
            rr = i_l_local
            while True:
                is_fixed &= all(j.type == JOINT_FIXED for j in ent.links[rr].joints)
                if ent.links[rr].parent < 0:
                    break
                rr = ent.links[rr].parent
            for js in ls.joints:
                nd = {JOINT_FIXED: 0, JOINT_REVOLUTE: 1, JOINT_PRISMATIC: 1, JOINT_SPHERICAL: 3, JOINT_FREE: 6}[js.type]
                nq = {JOINT_FIXED: 0, JOINT_REVOLUTE: 1, JOINT_PRISMATIC: 1, JOINT_SPHERICAL: 4, JOINT_FREE: 7}[js.type]
                if js.type == JOINT_FREE:
                    m_ang = np.eye(6, 3, -3)
                    m_vel = np.eye(6, 3)
                elif js.type == JOINT_REVOLUTE:
                    m_ang = np.asarray(js.axis, dtype=np.float64).reshape(1, 3)
                    m_vel = np.zeros((1, 3))
                elif js.type == JOINT_PRISMATIC:
                    m_ang = np.zeros((1, 3))
                    m_vel = np.asarray(js.axis, dtype=np.float64).reshape(1, 3)
                elif js.type == JOINT_SPHERICAL:
                    m_ang = np.eye(3)
                    m_vel = np.zeros((3, 3))
                else:
                    m_ang = np.zeros((0, 3))
                    m_vel = np.zeros((0, 3))
                lim = np.tile([[-np.inf, np.inf]], (nd, 1)) if js.limit is None else np.asarray(js.limit, np.float64).reshape(nd, 2)
                fr = np.tile([[-np.inf, np.inf]], (nd, 1)) if js.force_range is None else np.tile(np.asarray(js.force_range, np.float64).reshape(1, 2), (nd, 1))
                kp = np.full(nd, 0.0 if js.kp is None else js.kp)
                kv = np.full(nd, 0.0 if js.kv is None else js.kv)
                if js.init_qpos is not None:
                    iq = np.asarray(js.init_qpos, dtype=np.float64)
                elif js.type == JOINT_FREE:
                    iq = np.concatenate([ls.pos, ls.quat])
                elif js.type == JOINT_SPHERICAL:
                    iq = np.array([1.0, 0, 0, 0])
                else:
                    iq = np.zeros(nq)
                sp = js.sol_params if js.sol_params is not None else default_solver_params()
                joints.append(dict(type=js.type, link=i_l, q_start=n_q, dof_start=n_d, n_dofs=nd, n_qs=nq,
                                   pos=np.asarray(js.pos, np.float64), sol_params=sanitize_sol_params(sp, min_tc, default_tc)[0],
                                   m_ang=m_ang, m_vel=m_vel, limit=lim, force_range=fr, kp=kp, kv=kv,
                                   damping=np.full(nd, js.damping), armature=np.full(nd, js.armature),
                                   stiffness=np.full(nd, js.stiffness), frictionloss=np.full(nd, js.frictionloss),
                                   entity=i_e, name=js.name))
                name_to_joint[js.name] = len(joints) - 1
                q0.append(iq)
                n_q += nq
                n_d += nd
            # geoms
            g_start = len(geoms)
            for gs_ in ls.geoms:
                if gs_.type == GEOM_BOX:
                    verts = box_vertices(gs_.data[:3])
                    data = list(gs_.data[:3])
                elif gs_.type == GEOM_PLANE:
                    verts = plane_vertices()
                    data = list(gs_.data[:3]) if len(gs_.data) else [0.0, 0.0, 1.0]
                elif gs_.type == GEOM_SPHERE:
                    verts = sphere_vertices(gs_.data[0])
                    data = [gs_.data[0]]
                elif gs_.type == GEOM_CAPSULE:
                    verts = capsule_vertices(gs_.data[0], gs_.data[1])
                    data = [gs_.data[0], gs_.data[1]]
                elif gs_.type == GEOM_MESH:
                    if gs_.verts is None or len(gs_.verts) < 4:
                        raise ValueError("a mesh geom needs the vertices of its convex hull")
                    verts = np.asarray(gs_.verts, np.float64)
                    data = []
                else:
                    raise ValueError(f"unsupported geom type {gs_.type}")
                fric = ent.friction if ent.friction is not None else (gs_.friction if gs_.friction is not None else 1.0)
                sp = gs_.sol_params if gs_.sol_params is not None else default_solver_params()
                geoms.append(dict(type=gs_.type, link=i_l, pos=np.asarray(gs_.pos, np.float64),
                                  quat=np.asarray(gs_.quat, np.float64), data=np.pad(np.asarray(data, np.float64), (0, 7 - len(data))),
                                  friction=fric, sol_params=sanitize_sol_params(sp, min_tc, default_tc)[0],
                                  contype=gs_.contype, conaffinity=gs_.conaffinity, verts=verts,
                                  center=verts.mean(0), is_fixed=is_fixed, entity=i_e))
            # inertia
            if ls.inertial_mass is None or ls.inertial_i is None:
                tm, tc, ti = EPS32, np.zeros(3), np.eye(3) * EPS32
                for gs_ in ls.geoms:
                    if gs_.type == GEOM_PLANE:
                        # plane collision mesh is a 1000x1000x0.01 slab (mesh.py:894-898)
                        sl = GeomSpec(GEOM_BOX, data=(1e3, 1e3, 1e-2))
                        m, c, I = geom_mass_properties(sl, ent.rho)
                        c = c + np.array([0, 0, -5e-3])
                    else:
                        m, c, I = geom_mass_properties(gs_, ent.rho)
                    Rg = quat_to_R(gs_.quat)
                    tm, tc, ti = _compose(tm, tc, ti, m, Rg @ c + gs_.pos, Rg @ I @ Rg.T)
                imass, ipos, iI, iquat = tm, tc, ti, np.array([1.0, 0, 0, 0])
            else:
                imass = float(ls.inertial_mass)
                ipos = np.asarray(ls.inertial_pos if ls.inertial_pos is not None else np.zeros(3), np.float64)
                iI = np.asarray(ls.inertial_i, np.float64).reshape(3, 3)
                iquat = np.asarray(ls.inertial_quat if ls.inertial_quat is not None else [1.0, 0, 0, 0], np.float64)
            links.append(dict(parent=parent, root=root, pos=np.asarray(ls.pos, np.float64), quat=np.asarray(ls.quat, np.float64),
                              joint_start=joint_start, joint_end=len(joints), q_start=lq_start, q_end=n_q,
                              dof_start=ld_start, dof_end=n_d, geom_start=g_start, geom_end=len(geoms),
                              inertial_pos=ipos, inertial_quat=iquat, inertial_i=iI, inertial_mass=imass,
                              is_fixed=is_fixed, entity=i_e, name=ls.name))
        for eq in ent.equalities:
            sp = eq.sol_params if eq.sol_params is not None else default_solver_params()
            sp = sanitize_sol_params(sp, min_tc, default_tc)[0]
            if eq.type == EQ_JOINT:
                eqs.append(dict(type=eq.type, obj1=name_to_joint[eq.obj1], obj2=name_to_joint[eq.obj2],
                                data=np.pad(np.asarray(eq.data, np.float64), (0, 11 - len(eq.data))), sol_params=sp))
            elif eq.type in (EQ_CONNECT, EQ_WELD):
                # link names resolve once every entity is built: this entity's links first, then any entity's
                eqs.append(dict(type=eq.type, obj1=eq.obj1, obj2=eq.obj2, link_start=link_start, spec=eq,
                                data=None, sol_params=sp))
            else:
                raise ValueError(f"unknown equality type {eq.type}")
        ent_rows.append(dict(link_start=link_start, link_end=len(links), dof_start=dof_start_e, dof_end=n_d,
                             geom_start=geom_start_e, geom_end=len(geoms)))

    L, J, D, Q, G, E = len(links), len(joints), n_d, n_q, len(geoms), len(entities)
    for e in eqs:
        if e["type"] == EQ_JOINT:
            continue
        local = {links[i]["name"]: i for i in range(e["link_start"], L)}
        glob = {}
        for i in range(L):
            glob.setdefault(links[i]["name"], i)
        for k in ("obj1", "obj2"):
            name = e[k]
            if name not in local and name not in glob:
                raise KeyError(f"equality link {name!r} not found")
            e[k] = local.get(name, glob.get(name))
    M.n_links, M.n_joints, M.n_dofs, M.n_qs, M.n_geoms, M.n_entities = L, J, D, Q, G, E
    M.n_equalities = len(eqs)
    f32, i32 = np.float32, np.int32

    # links
    M.link_parent = np.array([l["parent"] for l in links], i32)
    M.link_root = np.array([l["root"] for l in links], i32)
    M.link_q_start = np.array([l["q_start"] for l in links], i32)
    M.link_q_end = np.array([l["q_end"] for l in links], i32)
    M.link_dof_start = np.array([l["dof_start"] for l in links], i32)
    M.link_dof_end = np.array([l["dof_end"] for l in links], i32)
    M.link_joint_start = np.array([l["joint_start"] for l in links], i32)
    M.link_joint_end = np.array([l["joint_end"] for l in links], i32)
    M.link_is_fixed = np.array([l["is_fixed"] for l in links], i32)
    M.link_entity = np.array([l["entity"] for l in links], i32)
    M.link_pos = np.array([l["pos"] for l in links], f32).reshape(L, 3)
    M.link_quat = np.array([l["quat"] for l in links], f32).reshape(L, 4)
    M.link_inertial_pos = np.array([l["inertial_pos"] for l in links], f32).reshape(L, 3)
    M.link_inertial_quat = np.array([l["inertial_quat"] for l in links], f32).reshape(L, 4)
    M.link_inertial_i = np.array([l["inertial_i"] for l in links], f32).reshape(L, 9)
    M.link_inertial_mass = np.array([l["inertial_mass"] for l in links], f32)
    M.link_names = [l["name"] for l in links]
    # joints
    M.joint_type = np.array([j["type"] for j in joints], i32)
    M.joint_q_start = np.array([j["q_start"] for j in joints], i32)
    M.joint_dof_start = np.array([j["dof_start"] for j in joints], i32)
    M.joint_dof_end = np.array([j["dof_start"] + j["n_dofs"] for j in joints], i32)
    M.joint_pos = np.array([j["pos"] for j in joints], f32).reshape(J, 3)
    M.joint_sol_params = np.array([j["sol_params"] for j in joints], f32).reshape(J, 7)
    M.joint_link = np.array([j["link"] for j in joints], i32)
    M.joint_names = [j["name"] for j in joints]
    # dofs

    def cat(key, shape):
        arrs = [j[key] for j in joints if j["n_dofs"] > 0]
        return np.concatenate(arrs).astype(f32).reshape(shape) if arrs else np.zeros(shape, f32)

    M.dof_motion_ang = cat("m_ang", (D, 3))
    M.dof_motion_vel = cat("m_vel", (D, 3))
    M.dof_limit = cat("limit", (D, 2))
    M.dof_force_range = cat("force_range", (D, 2))
    M.dof_kp = cat("kp", (D,))
    M.dof_kv = cat("kv", (D,))
    M.dof_damping = cat("damping", (D,))
    M.dof_armature = cat("armature", (D,))
    M.dof_stiffness = cat("stiffness", (D,))
    M.dof_frictionloss = cat("frictionloss", (D,))
    M.dof_entity = np.array(sum([[j["entity"]] * j["n_dofs"] for j in joints], []), i32)
    M.dof_link = np.array(sum([[j["link"]] * j["n_dofs"] for j in joints], []), i32)
    M.dof_joint_type = np.array(sum([[j["type"]] * j["n_dofs"] for j in joints], []), i32)
    M.qpos0 = (np.concatenate(q0) if q0 else np.zeros(0)).astype(f32)
    # entities
    M.ent_link_start = np.array([e["link_start"] for e in ent_rows], i32)
    M.ent_link_end = np.array([e["link_end"] for e in ent_rows], i32)
    M.ent_dof_start = np.array([e["dof_start"] for e in ent_rows], i32)
    M.ent_dof_end = np.array([e["dof_end"] for e in ent_rows], i32)
    M.ent_geom_start = np.array([e["geom_start"] for e in ent_rows], i32)
    M.ent_geom_end = np.array([e["geom_end"] for e in ent_rows], i32)
    M.ent_gravity_compensation = np.zeros(E, f32)
    # geoms + verts
    M.geom_type = np.array([g["type"] for g in geoms], i32)
    M.geom_link = np.array([g["link"] for g in geoms], i32)
    M.geom_pos = np.array([g["pos"] for g in geoms], f32).reshape(G, 3)
    M.geom_quat = np.array([g["quat"] for g in geoms], f32).reshape(G, 4)
    M.geom_data = np.array([g["data"] for g in geoms], f32).reshape(G, 7)
    M.geom_friction = np.array([g["friction"] for g in geoms], f32)
    M.geom_sol_params = np.array([g["sol_params"] for g in geoms], f32).reshape(G, 7)
    M.geom_is_fixed = np.array([g["is_fixed"] for g in geoms], i32)
    M.geom_center = np.array([g["center"] for g in geoms], f32).reshape(G, 3)
    vs, ve, allv = [], [], []
    nv = 0
    for g in geoms:
        vs.append(nv)
        allv.append(g["verts"])
        nv += len(g["verts"])
        ve.append(nv)
    M.geom_vert_start = np.array(vs, i32)
    M.geom_vert_end = np.array(ve, i32)
    M.verts = (np.concatenate(allv) if allv else np.zeros((0, 3))).astype(f32).reshape(-1, 3)
    M.n_verts = nv
    # init AABB corners in the fixed order of rigid_solver_decomp.py:3210-3221 (computed from float32 verts)
    M.geom_init_aabb = np.zeros((G, 8, 3), f32)
    for i_g in range(G):
        v = M.verts[vs[i_g]:ve[i_g]]
        lo, hi = v.min(0), v.max(0)
        for c in range(8):
            M.geom_init_aabb[i_g, c] = [hi[0] if c & 4 else lo[0], hi[1] if c & 2 else lo[1], hi[2] if c & 1 else lo[2]]
    # mesh support fields (support_field_decomp.py:24-94), from the float32 vertices: only mesh geoms look them up
    from .meshes import SUPPORT_RES, support_field
    cell_start, sv, sid, cells = np.zeros(G, i32), [], [], 0
    for i_g in range(G):
        if geoms[i_g]["type"] != GEOM_MESH:
            continue
        pos, vid = support_field(M.verts[vs[i_g]:ve[i_g]])
        cell_start[i_g] = cells
        sv.append(pos)
        sid.append(vid)
        cells += SUPPORT_RES * SUPPORT_RES
    M.support_cell_start = cell_start
    M.support_v = (np.concatenate(sv) if sv else np.zeros((1, 3))).astype(f32)
    M.support_vid = (np.concatenate(sid) if sid else np.zeros(1)).astype(i32)
    # equalities
    M.eq_type = np.array([e["type"] for e in eqs], i32)
    M.eq_obj1 = np.array([e["obj1"] for e in eqs], i32)
    M.eq_obj2 = np.array([e["obj2"] for e in eqs], i32)
    M.eq_sol_params = np.array([e["sol_params"] for e in eqs], f32).reshape(len(eqs), 7)
    M.max_dyn_welds = int(opt.max_dynamic_constraints)

    # mass parent mask (rigid_solver_decomp.py:512-523)
    mask = np.zeros((D, D), f32)
    for i_l in range(L):
        j_l = i_l
        while j_l != -1:
            mask[M.link_dof_start[i_l]:M.link_dof_end[i_l], M.link_dof_start[j_l]:M.link_dof_end[j_l]] = 1.0
            j_l = M.link_parent[j_l]
    M.mass_parent_mask = mask

    # collision pair table (collider_decomp.py:184-264)
    pair = np.full((G, G), -1, i32)
    n_pairs = 0
    for a in range(G):
        for b in range(a + 1, G):
            la, lb = M.geom_link[a], M.geom_link[b]
            if la == lb:
                continue
            if M.link_root[la] == M.link_root[lb]:
                if not opt.enable_self_collision:
                    continue
                if not opt.enable_adjacent_collision and (M.link_parent[la] == lb or M.link_parent[lb] == la):
                    continue
            ea, eb = M.link_entity[la], M.link_entity[lb]
            ga, gb = geoms[a], geoms[b]
            local_a, local_b = entities[ea].local_collision_mask, entities[eb].local_collision_mask
            if (ea == eb or not (local_a or local_b)) and not (
                (ga["contype"] & gb["conaffinity"]) or (gb["contype"] & ga["conaffinity"])
            ):
                continue
            if M.link_is_fixed[la] and M.link_is_fixed[lb]:
                continue
            pair[a, b] = n_pairs
            n_pairs += 1
    M.pair_idx = pair
    M.n_pairs = n_pairs
    max_collision_pairs = min(opt.max_collision_pairs, n_pairs)
    M.max_contacts = max_collision_pairs * opt.n_contacts_per_pair
    M.max_broad = max_collision_pairs * opt.multiplier_collision_broad_phase
    n_1dof = int(sum(1 for j in joints if j["type"] in (JOINT_REVOLUTE, JOINT_PRISMATIC)))
    n_cand_eq = max(1, len(eqs) + opt.max_dynamic_constraints)
    M.max_rows = 4 * M.max_contacts + n_1dof + D + 6 * n_cand_eq
    # widest constraint row: dofs on the union of two root chains
    chain_dofs = []
    for i_l in range(L):
        s = set()
        j = i_l
        while j != -1:
            s.update(range(M.link_dof_start[j], M.link_dof_end[j]))
            j = M.link_parent[j]
        chain_dofs.append(s)
    width = 2
    for a in range(G):
        for b in range(a + 1, G):
            if pair[a, b] >= 0:
                width = max(width, len(chain_dofs[M.geom_link[a]] | chain_dofs[M.geom_link[b]]))
    # equality rows span the union of their two links' chains; welds registered after the build may join any two
    for e in eqs:
        if e["type"] != EQ_JOINT:
            width = max(width, len(chain_dofs[e["obj1"]] | chain_dofs[e["obj2"]]))
    if M.max_dyn_welds > 0:
        for a in range(L):
            for b in range(a + 1, L):
                width = max(width, len(chain_dofs[a] | chain_dofs[b]))
    M.row_width = width
    M.max_entity_dofs = int(max([e["dof_end"] - e["dof_start"] for e in ent_rows] + [1]))

    M.gravity = np.array(opt.gravity, f32)
    # CONNECT / WELD data from the build pose when the spec gives none (MJCF's compiled anchors, and the weld's
    # relative pose as kernel_add_weld_constraint takes it, constraint_solver_decomp.py:2452-2473)
    lpos = lquat = None
    for e in eqs:
        if e["type"] == EQ_JOINT:
            continue
        if e["spec"].data is not None and len(e["spec"].data) == 11 and e["type"] == EQ_WELD and e["spec"].explicit:
            e["data"] = np.asarray(e["spec"].data, np.float64)
            continue
        if lpos is None:
            lpos, lquat = _host_fk_crba(M, M.qpos0.astype(np.float64))[:2]
        l1, l2 = e["obj1"], e["obj2"]
        inv = lambda p, t, q: transform_by_quat(p - t, np.array([q[0], -q[1], -q[2], -q[3]]))  # noqa: E731
        d = np.zeros(11)
        if e["type"] == EQ_CONNECT:
            a1 = np.asarray(e["spec"].data[:3], np.float64)
            d[0:3] = a1
            d[3:6] = inv(lpos[l1] + transform_by_quat(a1, lquat[l1]), lpos[l2], lquat[l2])
        else:
            a2 = e["spec"].anchor2
            shared = lpos[l1] if a2 is None else lpos[l2] + transform_by_quat(np.asarray(a2, np.float64), lquat[l2])
            d[3:6] = inv(shared, lpos[l1], lquat[l1])
            d[0:3] = inv(shared, lpos[l2], lquat[l2]) if a2 is None else a2
            d[6:10] = quat_mul(np.array([lquat[l1][0], -lquat[l1][1], -lquat[l1][2], -lquat[l1][3]]), lquat[l2])
            d[10] = e["spec"].torquescale
        e["data"] = d
    M.eq_data = np.array([e["data"] for e in eqs], f32).reshape(len(eqs), 11)
    _init_invweight_and_meaninertia(M)
    return M


# ---------------------------------------------------------------------------------------------------------
# Build-time inverse weights and mean inertia (rigid_solver_decomp.py:354-507, 2854-2875), float64 on host.
# ---------------------------------------------------------------------------------------------------------
def _host_fk_crba(M: RigidModel, qpos):
    """Float64 FK + composite-rigid-body mass matrix at one configuration.

    Returns (link_pos, link_quat, i_pos (COM relative to the subtree root COM), cdof_ang, cdof_vel, mass_mat)."""
    L, D = M.n_links, M.n_dofs
    lpos = np.zeros((L, 3))
    lquat = np.zeros((L, 4))
    xanchor = np.zeros((M.n_joints, 3))
    xaxis = np.zeros((M.n_joints, 3))
    for i_l in range(L):
        pos = M.link_pos[i_l].astype(np.float64)
        quat = M.link_quat[i_l].astype(np.float64)
        p = M.link_parent[i_l]
        if p != -1:
            pos = lpos[p] + transform_by_quat(pos, lquat[p])
            quat = quat_mul(lquat[p], quat)
            quat /= np.linalg.norm(quat)
        for i_j in range(M.link_joint_start[i_l], M.link_joint_end[i_l]):
            jt = M.joint_type[i_j]
            qs = M.joint_q_start[i_j]
            ds = M.joint_dof_start[i_j]
            if jt == JOINT_FREE:
                xanchor[i_j] = qpos[qs:qs + 3]
                xaxis[i_j] = [0, 0, 1]
                pos = np.array(qpos[qs:qs + 3], np.float64)
                quat = np.array(qpos[qs + 3:qs + 7], np.float64)
            elif jt == JOINT_FIXED:
                pass
            else:
                axis = np.array([0, 0, 1.0])
                if jt == JOINT_REVOLUTE:
                    axis = M.dof_motion_ang[ds].astype(np.float64)
                elif jt == JOINT_PRISMATIC:
                    axis = M.dof_motion_vel[ds].astype(np.float64)
                xanchor[i_j] = transform_by_quat(M.joint_pos[i_j], quat) + pos
                xaxis[i_j] = transform_by_quat(axis, quat)
                if jt == JOINT_REVOLUTE:
                    ang = qpos[qs] - M.qpos0[qs]
                    rv = axis * ang
                    th = np.linalg.norm(rv)
                    qloc = np.array([1.0, 0, 0, 0]) if th < 1e-12 else np.concatenate([[math.cos(th / 2)], math.sin(th / 2) * rv / th])
                    quat = quat_mul(quat, qloc)
                    quat /= np.linalg.norm(quat)
                    pos = xanchor[i_j] - transform_by_quat(M.joint_pos[i_j], quat)
                elif jt == JOINT_PRISMATIC:
                    pos = pos + xaxis[i_j] * (qpos[qs] - M.qpos0[qs])
                elif jt == JOINT_SPHERICAL:
                    qloc = np.array(qpos[qs:qs + 4], np.float64)
                    quat = quat_mul(quat, qloc)
                    quat /= np.linalg.norm(quat)
                    pos = xanchor[i_j] - transform_by_quat(M.joint_pos[i_j], quat)
        lpos[i_l], lquat[i_l] = pos, quat
    # COM
    mass = M.link_inertial_mass.astype(np.float64)
    ipos_w = np.array([lpos[l] + transform_by_quat(M.link_inertial_pos[l], lquat[l]) for l in range(L)])
    iquat_w = np.array([quat_mul(lquat[l], M.link_inertial_quat[l]) for l in range(L)])
    root_com = np.zeros((L, 3))
    for r in set(M.link_root.tolist()):
        sel = M.link_root == r
        root_com[sel] = (mass[sel, None] * ipos_w[sel]).sum(0) / mass[sel].sum()
    i_pos = ipos_w - root_com
    # cinr (COM-frame spatial inertia about root COM)
    cinr_I = np.zeros((L, 3, 3))
    for l in range(L):
        R = quat_to_R(iquat_w[l])
        x = i_pos[l]
        hhT = np.array([[x[1] ** 2 + x[2] ** 2, -x[0] * x[1], -x[0] * x[2]],
                        [-x[0] * x[1], x[0] ** 2 + x[2] ** 2, -x[1] * x[2]],
                        [-x[0] * x[2], -x[1] * x[2], x[0] ** 2 + x[1] ** 2]])
        cinr_I[l] = R @ M.link_inertial_i[l].reshape(3, 3) @ R.T + hhT * mass[l]
    cinr_pos = i_pos * mass[:, None]
    # cdof
    cdof_ang = np.zeros((D, 3))
    cdof_vel = np.zeros((D, 3))
    for l in range(L):
        for i_j in range(M.link_joint_start[l], M.link_joint_end[l]):
            jt = M.joint_type[i_j]
            ds = M.joint_dof_start[i_j]
            off = root_com[l] - xanchor[i_j]
            if jt == JOINT_REVOLUTE:
                cdof_ang[ds] = xaxis[i_j]
                cdof_vel[ds] = np.cross(xaxis[i_j], off)
            elif jt == JOINT_PRISMATIC:
                cdof_vel[ds] = xaxis[i_j]
            elif jt == JOINT_SPHERICAL:
                RT = quat_to_R(lquat[l]).T
                for i in range(3):
                    cdof_ang[ds + i] = RT[i]
                    cdof_vel[ds + i] = np.cross(RT[i], off)
            elif jt == JOINT_FREE:
                for i in range(3):
                    cdof_vel[ds + i, i] = 1.0
                RT = quat_to_R(lquat[l]).T
                for i in range(3):
                    cdof_ang[ds + 3 + i] = RT[i]
                    cdof_vel[ds + 3 + i] = np.cross(RT[i], off)
    # CRBA
    crb_I, crb_m, crb_p = cinr_I.copy(), mass.copy(), cinr_pos.copy()
    for e in range(M.n_entities):
        for l in range(M.ent_link_end[e] - 1, M.ent_link_start[e] - 1, -1):
            p = M.link_parent[l]
            if p != -1:
                crb_I[p] += crb_I[l]
                crb_m[p] += crb_m[l]
                crb_p[p] += crb_p[l]
    f_ang = np.zeros((D, 3))
    f_vel = np.zeros((D, 3))
    for l in range(L):
        for d in range(M.link_dof_start[l], M.link_dof_end[l]):
            f_ang[d] = crb_I[l] @ cdof_ang[d] + np.cross(crb_p[l], cdof_vel[d])
            f_vel[d] = crb_m[l] * cdof_vel[d] - np.cross(crb_p[l], cdof_ang[d])
    Mm = np.zeros((D, D))
    for e in range(M.n_entities):
        s, t = M.ent_dof_start[e], M.ent_dof_end[e]
        for i in range(s, t):
            for j in range(s, i + 1):
                Mm[i, j] = (f_ang[i] @ cdof_ang[j] + f_vel[i] @ cdof_vel[j]) * M.mass_parent_mask[i, j]
                Mm[j, i] = Mm[i, j]
    Mm += np.diag(M.dof_armature.astype(np.float64))
    return lpos, lquat, i_pos, cdof_ang, cdof_vel, Mm


def _init_invweight_and_meaninertia(M: RigidModel):
    L, D = M.n_links, M.n_dofs
    M.link_invweight = np.zeros((L, 2), np.float32)
    M.dof_invweight = np.zeros(D, np.float32)
    if D == 0:
        M.meaninertia = np.float32(1.0)
        return
    _, _, i_pos, cdof_ang, cdof_vel, Mm = _host_fk_crba(M, M.qpos0.astype(np.float64))
    Minv = np.zeros((D, D))
    for e in range(M.n_entities):
        s, t = M.ent_dof_start[e], M.ent_dof_end[e]
        if t > s:
            Minv[s:t, s:t] = np.linalg.inv(Mm[s:t, s:t])
    for l in range(L):
        jacp = np.zeros((3, D))
        jacr = np.zeros((3, D))
        off = i_pos[l]
        j = l
        while j != -1:
            for d in range(M.link_dof_start[j], M.link_dof_end[j]):
                jacp[:, d] = cdof_vel[d] + np.cross(cdof_ang[d], off)
                jacr[:, d] = cdof_ang[d]
            j = M.link_parent[j]
        jac = np.concatenate([jacp, jacr], 0)
        A = np.diag(jac @ Minv @ jac.T)
        M.link_invweight[l] = [A[:3].mean(), A[3:].mean()]
    for i_j in range(M.n_joints):
        jt = M.joint_type[i_j]
        if jt == JOINT_FIXED:
            continue
        ds, de = M.joint_dof_start[i_j], M.joint_dof_end[i_j]
        A = np.diag(Minv)[ds:de]
        if jt == JOINT_FREE:
            M.dof_invweight[ds:ds + 3] = A[:3].mean()
            M.dof_invweight[ds + 3:ds + 6] = A[3:].mean()
        elif jt == JOINT_SPHERICAL:
            M.dof_invweight[ds:ds + 3] = A[:3].mean()
        else:
            M.dof_invweight[ds] = A[0]
    # meaninertia: the division by n_dofs sits inside the entity loop (rigid_solver_decomp.py:2866-2872)
    mi = np.float32(0.0)
    diag = np.diag(Mm).astype(np.float32)
    for e in range(M.n_entities):
        for d in range(M.ent_dof_start[e], M.ent_dof_end[e]):
            mi = np.float32(mi + diag[d])
        mi = np.float32(mi / np.float32(D))
    M.meaninertia = mi
