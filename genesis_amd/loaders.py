"""Asset loaders for the rigid step: a URDF parser and an MJCF subset parser producing EntitySpec trees.

Semantics follow the reference's loaders where they are pure bookkeeping:
* URDF: genesis/utils/urdf.py:48-300 (link order by tree depth `_order_links` :11-45, fixed-link merging
  `merge_fixed_links` :367-410, joint types/limits/effort, kp/kv defaults 100/10 from genesis/utils/geom.py:1721-1726,
  damping/friction from <dynamics>), plus the free root joint that rigid_entity.py:421-461 adds for a floating base.
* MJCF: joint defaults/classes, hinge/slide/free/ball joints, inertials, `general` actuators with affine bias mapped to
  kp = -gear*biasprm[1], kv = -gear*biasprm[2] and forcerange (genesis/utils/mjcf.py:295-370), JOINT equalities.

The reference routes both formats through MuJoCo's compiler (rigid_entity.py:366-418), which is not available here,
so these loaders restate the documented file semantics directly. An MJCF collision mesh becomes a MESH geom holding
the vertices of its convex hull in the mesh file's frame (the reference convexifies collision meshes with a hull,
genesis/utils/mesh.py:299-475; its optional decimation above 500 faces is not applied); a URDF cylinder becomes a
capsule of the same radius and length (the quadruped workload specifies capsule/sphere collision geometry).
"""
from __future__ import annotations

import math
import os
import struct
import xml.etree.ElementTree as ET
from typing import Dict, List, Optional

import numpy as np

from .model import (
    EQ_CONNECT, EQ_JOINT, EQ_WELD, GEOM_BOX, GEOM_CAPSULE, GEOM_MESH, GEOM_SPHERE, JOINT_FIXED, JOINT_FREE, JOINT_PRISMATIC, JOINT_REVOLUTE,
    JOINT_SPHERICAL, EntitySpec, EqualitySpec, GeomSpec, JointSpec, LinkSpec, R_to_quat, quat_mul, quat_to_R,
    rpy_to_quat,
)


def _floats(s, n=None, default=None):
    if s is None:
        return None if default is None else np.array(default, dtype=np.float64)
    v = np.array([float(x) for x in s.split()], dtype=np.float64)
    if n is not None and len(v) != n:
        raise ValueError(f"expected {n} numbers, got {s!r}")
    return v


# ---------------------------------------------------------------------------------------------------------
# mesh -> capsule fit
# ---------------------------------------------------------------------------------------------------------
def read_mesh_vertices(path: str) -> np.ndarray:
    ext = os.path.splitext(path)[1].lower()
    if ext == ".obj":
        vs = []
        with open(path) as f:
            for line in f:
                if line.startswith("v "):
                    vs.append([float(x) for x in line.split()[1:4]])
        return np.array(vs, dtype=np.float64)
    if ext == ".stl":
        with open(path, "rb") as f:
            data = f.read()
        if data[:5].lower() == b"solid" and b"facet" in data[:512]:
            vs = []
            for line in data.decode(errors="ignore").splitlines():
                t = line.strip().split()
                if t and t[0] == "vertex":
                    vs.append([float(x) for x in t[1:4]])
            return np.array(vs, dtype=np.float64)
        n = struct.unpack("<I", data[80:84])[0]
        tri = np.frombuffer(data[84:84 + 50 * n], dtype=np.dtype([("n", "<f4", 3), ("v", "<f4", (3, 3)), ("a", "<u2")]))
        return np.unique(tri["v"].reshape(-1, 3).astype(np.float64), axis=0)
    raise ValueError(f"unsupported mesh format {path}")


def fit_capsule(verts: np.ndarray):
    """Capsule (center, quat, radius, length) around a vertex cloud: axis = principal direction of the points."""
    c = 0.5 * (verts.min(0) + verts.max(0))
    X = verts - c
    _, _, vt = np.linalg.svd(X, full_matrices=False)
    axis = vt[0]
    t = X @ axis
    radial = X - np.outer(t, axis)
    r = float(np.percentile(np.linalg.norm(radial, axis=1), 90))
    lo, hi = float(t.min()) + r, float(t.max()) - r
    if hi < lo:
        lo = hi = 0.5 * (lo + hi)
    center = c + axis * 0.5 * (lo + hi)
    z = np.array([0.0, 0.0, 1.0])
    v = np.cross(z, axis)
    s, cth = np.linalg.norm(v), float(z @ axis)
    if s < 1e-12:
        q = np.array([1.0, 0, 0, 0]) if cth > 0 else np.array([0.0, 1.0, 0, 0])
    else:
        ang = math.atan2(s, cth)
        q = np.concatenate([[math.cos(ang / 2)], math.sin(ang / 2) * v / s])
    return center, q, r, hi - lo


# ---------------------------------------------------------------------------------------------------------
# URDF
# ---------------------------------------------------------------------------------------------------------
def _origin(el):
    if el is None:
        return np.zeros(3), np.array([1.0, 0, 0, 0])
    xyz = _floats(el.get("xyz"), 3, [0, 0, 0])
    rpy = _floats(el.get("rpy"), 3, [0, 0, 0])
    return xyz, rpy_to_quat(rpy)


def parse_urdf(path: str, fixed: bool = False, merge_fixed_links: bool = True, links_to_keep=()) -> EntitySpec:
    root = ET.parse(path).getroot()
    links: Dict[str, dict] = {}
    order = []
    for le in root.findall("link"):
        name = le.get("name")
        inertial = le.find("inertial")
        d = dict(name=name, geoms=[], mass=None, ipos=np.zeros(3), iquat=np.array([1.0, 0, 0, 0]), I=None)
        if inertial is not None:
            d["ipos"], d["iquat"] = _origin(inertial.find("origin"))
            d["mass"] = float(inertial.find("mass").get("value"))
            ie = inertial.find("inertia")
            ixx, ixy, ixz, iyy, iyz, izz = (float(ie.get(k, 0.0)) for k in ("ixx", "ixy", "ixz", "iyy", "iyz", "izz"))
            d["I"] = np.array([[ixx, ixy, ixz], [ixy, iyy, iyz], [ixz, iyz, izz]])
        for ce in le.findall("collision"):
            pos, quat = _origin(ce.find("origin"))
            g = ce.find("geometry")
            if g.find("box") is not None:
                d["geoms"].append(GeomSpec(GEOM_BOX, pos, quat, tuple(_floats(g.find("box").get("size"), 3))))
            elif g.find("sphere") is not None:
                d["geoms"].append(GeomSpec(GEOM_SPHERE, pos, quat, (float(g.find("sphere").get("radius")),)))
            elif g.find("cylinder") is not None:
                cy = g.find("cylinder")
                d["geoms"].append(GeomSpec(GEOM_CAPSULE, pos, quat, (float(cy.get("radius")), float(cy.get("length")))))
            elif g.find("capsule") is not None:
                cy = g.find("capsule")
                d["geoms"].append(GeomSpec(GEOM_CAPSULE, pos, quat, (float(cy.get("radius")), float(cy.get("length")))))
        links[name] = d
        order.append(name)
    joints = []
    for je in root.findall("joint"):
        pos, quat = _origin(je.find("origin"))
        lim = je.find("limit")
        dyn = je.find("dynamics")
        joints.append(dict(
            name=je.get("name"), type=je.get("type"), parent=je.find("parent").get("link"),
            child=je.find("child").get("link"), pos=pos, quat=quat,
            axis=_floats(je.find("axis").get("xyz"), 3) if je.find("axis") is not None else np.array([1.0, 0, 0]),
            lower=float(lim.get("lower")) if lim is not None and lim.get("lower") is not None else None,
            upper=float(lim.get("upper")) if lim is not None and lim.get("upper") is not None else None,
            effort=float(lim.get("effort")) if lim is not None and lim.get("effort") is not None else None,
            damping=float(dyn.get("damping", 0.0)) if dyn is not None else 0.0,
            friction=float(dyn.get("friction", 0.0)) if dyn is not None else 0.0,
            mimic=je.find("mimic"),
        ))
    # merge links attached through fixed joints into their parent (genesis/utils/urdf.py:367-410)
    if merge_fixed_links:
        while True:
            jf = next((j for j in joints if j["type"] == "fixed" and j["child"] not in links_to_keep), None)
            if jf is None:
                break
            p, c = links[jf["parent"]], links[jf["child"]]
            R = quat_to_R(jf["quat"])
            # geoms of the child move into the parent frame
            for g in c["geoms"]:
                g.pos = jf["pos"] + R @ g.pos
                g.quat = quat_mul(jf["quat"], g.quat)
                p["geoms"].append(g)
            # inertia merge (urdf.py:449-475)
            if c["mass"] is not None:
                cI_p = R @ quat_to_R(c["iquat"]) @ c["I"] @ quat_to_R(c["iquat"]).T @ R.T
                ccom_p = jf["pos"] + R @ c["ipos"]
                if p["mass"] is None:
                    p["mass"], p["ipos"], p["I"], p["iquat"] = c["mass"], ccom_p, cI_p, np.array([1.0, 0, 0, 0])
                else:
                    pI = quat_to_R(p["iquat"]) @ p["I"] @ quat_to_R(p["iquat"]).T
                    m = p["mass"] + c["mass"]
                    com = (p["mass"] * p["ipos"] + c["mass"] * ccom_p) / m

                    def sh(I, mm, d):
                        return I + mm * (d @ d * np.eye(3) - np.outer(d, d))

                    p["I"] = sh(pI, p["mass"], com - p["ipos"]) + sh(cI_p, c["mass"], com - ccom_p)
                    p["mass"], p["ipos"], p["iquat"] = m, com, np.array([1.0, 0, 0, 0])
            # re-parent grandchildren
            for j in joints:
                if j["parent"] == jf["child"]:
                    j["pos"] = jf["pos"] + R @ j["pos"]
                    j["quat"] = quat_mul(jf["quat"], j["quat"])
                    j["parent"] = jf["parent"]
            joints.remove(jf)
            del links[jf["child"]]
            order.remove(jf["child"])
    child_joint = {j["child"]: j for j in joints}
    # order by depth (urdf.py:11-45)
    children = {n: [] for n in order}
    for j in joints:
        children[j["parent"]].append(j["child"])
    level = [n for n in order if n not in child_joint]
    ordered = []
    while level:
        ordered.extend(level)
        nxt = []
        for n in level:
            nxt.extend(sorted(children[n], key=order.index))
        level = nxt
    idx = {n: i for i, n in enumerate(ordered)}
    out: List[LinkSpec] = []
    for n in ordered:
        d = links[n]
        j = child_joint.get(n)
        jspecs = []
        if j is None:
            pos, quat, parent = np.zeros(3), np.array([1.0, 0, 0, 0]), -1
            if not fixed:
                jspecs.append(JointSpec("root_joint", JOINT_FREE, kp=0.0, kv=0.0))
            else:
                jspecs.append(JointSpec(f"{n}_joint", JOINT_FIXED))
        else:
            pos, quat, parent = j["pos"], j["quat"], idx[j["parent"]]
            t = j["type"]
            common = dict(damping=j["damping"], frictionloss=j["friction"], kp=100.0, kv=10.0,
                          force_range=None if j["effort"] is None else np.array([-j["effort"], j["effort"]]))
            if t in ("revolute", "continuous"):
                lim = None if t == "continuous" else np.array([[j["lower"] if j["lower"] is not None else -np.inf,
                                                                j["upper"] if j["upper"] is not None else np.inf]])
                jspecs.append(JointSpec(j["name"], JOINT_REVOLUTE, axis=j["axis"], limit=lim, **common))
            elif t == "prismatic":
                lim = np.array([[j["lower"] if j["lower"] is not None else -np.inf,
                                 j["upper"] if j["upper"] is not None else np.inf]])
                jspecs.append(JointSpec(j["name"], JOINT_PRISMATIC, axis=j["axis"], limit=lim, **common))
            elif t == "fixed":
                jspecs.append(JointSpec(j["name"], JOINT_FIXED))
            elif t == "floating":
                jspecs.append(JointSpec(j["name"], JOINT_FREE, **common))
            else:
                raise ValueError(f"unsupported URDF joint type {t}")
        out.append(LinkSpec(n, parent, pos, quat, jspecs, d["geoms"], d["ipos"], d["iquat"], d["I"], d["mass"]))
    eqs = []
    for j in joints:
        if j["mimic"] is not None:
            mult = float(j["mimic"].get("multiplier", 1.0))
            off = float(j["mimic"].get("offset", 0.0))
            eqs.append(EqualitySpec(EQ_JOINT, j["name"], j["mimic"].get("joint"),
                                    data=np.array([off, mult, 0, 0, 0, 0, 0, 0, 0, 0, 0])))
    return EntitySpec(out, eqs, name=os.path.splitext(os.path.basename(path))[0])


# ---------------------------------------------------------------------------------------------------------
# MJCF subset
# ---------------------------------------------------------------------------------------------------------
def _mj_quat(attrs):
    """Orientation of an MJCF element from its attribute mapping (quat or euler, radians)."""
    if attrs.get("quat") is not None:
        q = _floats(attrs.get("quat"), 4)
        return q / np.linalg.norm(q)
    if attrs.get("euler") is not None:
        return rpy_to_quat(_floats(attrs.get("euler"), 3))
    return np.array([1.0, 0, 0, 0])


def parse_mjcf(path: str, mesh_to_capsule: bool = False) -> EntitySpec:
    """mesh_to_capsule: replace each collision mesh by a capsule fitted to its vertices instead of its convex hull."""
    root = ET.parse(path).getroot()
    base_dir = os.path.dirname(path)
    comp = root.find("compiler")
    meshdir = os.path.join(base_dir, comp.get("meshdir", "")) if comp is not None else base_dir
    meshes = {}
    asset = root.find("asset")
    if asset is not None:
        for m in asset.findall("mesh"):
            name = m.get("name") or os.path.splitext(os.path.basename(m.get("file")))[0]
            meshes[name] = os.path.join(meshdir, m.get("file"))
    # default classes: class -> {tag: attrs}; nested classes inherit
    classes: Dict[str, Dict[str, dict]] = {}

    def walk_default(el, parent_attrs):
        cname = el.get("class", "main")
        attrs = {k: dict(v) for k, v in parent_attrs.items()}
        for ch in el:
            if ch.tag != "default":
                attrs.setdefault(ch.tag, {}).update(ch.attrib)
        classes[cname] = attrs
        for ch in el:
            if ch.tag == "default":
                walk_default(ch, attrs)

    d0 = root.find("default")
    if d0 is not None:
        walk_default(d0, {})

    def attrs_of(el, tag, inherited_class):
        cls = el.get("class", inherited_class)
        a = dict(classes.get("main", {}).get(tag, {}))
        a.update(classes.get(cls, {}).get(tag, {}))
        a.update(el.attrib)
        return a

    links: List[LinkSpec] = []
    name_to_joint = {}

    def add_body(el, parent_idx, childclass):
        cc = el.get("childclass", childclass)
        pos = _floats(el.get("pos"), 3, [0, 0, 0])
        quat = _mj_quat(el)
        inert = el.find("inertial")
        mass = ipos = iquat = I = None
        if inert is not None:
            mass = float(inert.get("mass"))
            ipos = _floats(inert.get("pos"), 3, [0, 0, 0])
            iquat = _mj_quat(inert)
            if inert.get("fullinertia") is not None:
                ixx, iyy, izz, ixy, ixz, iyz = _floats(inert.get("fullinertia"), 6)
                I = np.array([[ixx, ixy, ixz], [ixy, iyy, iyz], [ixz, iyz, izz]])
            else:
                I = np.diag(_floats(inert.get("diaginertia"), 3))
        jspecs = []
        for je in el.findall("joint"):
            a = attrs_of(je, "joint", cc)
            t = a.get("type", "hinge")
            axis = _floats(a.get("axis"), 3, [0, 0, 1])
            rng = _floats(a.get("range"), 2)
            limited = a.get("limited", "auto")
            lim = None
            if rng is not None and limited != "false":
                lim = rng.reshape(1, 2)
            common = dict(pos=_floats(a.get("pos"), 3, [0, 0, 0]), axis=axis, damping=float(a.get("damping", 0.0)),
                          armature=float(a.get("armature", 0.0)), stiffness=float(a.get("stiffness", 0.0)),
                          frictionloss=float(a.get("frictionloss", 0.0)), kp=0.0, kv=0.0, limit=lim)
            jtype = {"hinge": JOINT_REVOLUTE, "slide": JOINT_PRISMATIC, "ball": JOINT_SPHERICAL, "free": JOINT_FREE}[t]
            if jtype in (JOINT_FREE, JOINT_SPHERICAL):
                common["limit"] = None
            jspecs.append(JointSpec(je.get("name"), jtype, **common))
        if el.find("freejoint") is not None:
            jspecs.append(JointSpec(el.find("freejoint").get("name", "root_joint"), JOINT_FREE, kp=0.0, kv=0.0))
        if not jspecs:
            jspecs.append(JointSpec(f"{el.get('name')}_joint", JOINT_FIXED))
        geoms = []
        for ge in el.findall("geom"):
            a = attrs_of(ge, "geom", cc)
            if a.get("contype", "1") == "0" and a.get("conaffinity", "1") == "0":
                continue
            gtype = a.get("type", "sphere")
            gpos = _floats(a.get("pos"), 3, [0, 0, 0])
            gq = _mj_quat(a)
            size = _floats(a.get("size"))
            fric = float(a["friction"].split()[0]) if "friction" in a else None
            if gtype == "box":
                geoms.append(GeomSpec(GEOM_BOX, gpos, gq, tuple(2 * size[:3]), friction=fric))
            elif gtype == "sphere":
                geoms.append(GeomSpec(GEOM_SPHERE, gpos, gq, (size[0],), friction=fric))
            elif gtype == "capsule":
                geoms.append(GeomSpec(GEOM_CAPSULE, gpos, gq, (size[0], 2 * size[1]), friction=fric))
            elif gtype == "mesh" and mesh_to_capsule:
                v = read_mesh_vertices(meshes[a["mesh"]])
                c, q, r, length = fit_capsule(v)
                Rg = quat_to_R(gq)
                geoms.append(GeomSpec(GEOM_CAPSULE, gpos + Rg @ c, quat_mul(gq, q), (r, length), friction=fric))
            elif gtype == "mesh":
                from .meshes import convex_hull_vertices
                hull = convex_hull_vertices(read_mesh_vertices(meshes[a["mesh"]]))
                geoms.append(GeomSpec(GEOM_MESH, gpos, gq, (), friction=fric, verts=hull))
        my_idx = len(links)
        links.append(LinkSpec(el.get("name"), parent_idx, pos, quat, jspecs, geoms, ipos, iquat, I, mass))
        for j in jspecs:
            name_to_joint[j.name] = j
        for ch in el.findall("body"):
            add_body(ch, my_idx, cc)

    wb = root.find("worldbody")
    for b in wb.findall("body"):
        add_body(b, -1, "main")
    eq = root.find("equality")
    if eq is not None and all(l.name != "world" for l in links) and any(
            e.tag in ("connect", "weld") and (e.get("body2") or "world") == "world" for e in eq):
        # a CONNECT / WELD to the world: a fixed, massless link at the origin stands for the world frame
        links.append(LinkSpec("world", -1, np.zeros(3), np.array([1.0, 0.0, 0.0, 0.0]),
                              [JointSpec("world_joint", JOINT_FIXED)], [], inertial_mass=0.0))
    # re-order by depth like the URDF path
    depth = []
    for i, l in enumerate(links):
        dd, p = 0, l.parent
        while p >= 0:
            dd += 1
            p = links[p].parent
        depth.append(dd)
    order = sorted(range(len(links)), key=lambda i: (depth[i], i))
    remap = {o: n for n, o in enumerate(order)}
    new_links = []
    for o in order:
        l = links[o]
        l.parent = remap[l.parent] if l.parent >= 0 else -1
        new_links.append(l)
    # actuators (mjcf.py:295-370)
    act = root.find("actuator")
    tendons = {}
    ten = root.find("tendon")
    if ten is not None:
        for t in ten.findall("fixed"):
            tendons[t.get("name")] = [j.get("joint") for j in t.findall("joint")]
    if act is not None:
        for ae in act:
            a = attrs_of(ae, "general", ae.get("class", "main"))
            targets = [a["joint"]] if "joint" in a else tendons.get(a.get("tendon"), [])
            gain = _floats(a.get("gainprm"), None, [1.0])
            bias = _floats(a.get("biasprm"), None, [0.0, 0.0, 0.0])
            gain = np.pad(gain, (0, max(0, 3 - len(gain))))
            bias = np.pad(bias, (0, max(0, 3 - len(bias))))
            gear = 1.0
            if a.get("biastype", "none") == "none":
                kp = kv = 0.0
            elif gain[1:].any() or bias[0]:
                kp, kv = 100.0, 10.0
            elif gain[0] != -bias[1]:
                kp, kv = -min(-gain[0], bias[1]), -bias[2]
            else:
                kp, kv = -bias[1], -bias[2]
            fr = _floats(a.get("forcerange"), 2)
            for jn in targets:
                j = name_to_joint[jn]
                j.kp, j.kv = gear * kp, gear * kv
                if fr is not None and a.get("forcelimited", "auto") != "false":
                    j.force_range = fr
    eqs = []
    eq = root.find("equality")
    if eq is not None:
        def sol_params(e):
            solref = _floats(e.get("solref"), 2, [0.02, 1.0])
            solimp = _floats(e.get("solimp"), None, [0.9, 0.95, 0.001, 0.5, 2.0])
            solimp = np.concatenate([solimp, [0.9, 0.95, 0.001, 0.5, 2.0][len(solimp):]])
            return np.concatenate([solref, solimp])

        # in document order, like MuJoCo's equality list
        for e in eq:
            if e.tag == "joint":
                poly = _floats(e.get("polycoef"), None, [0.0, 1.0, 0.0, 0.0, 0.0])
                data = np.zeros(11)
                data[:len(poly)] = poly
                eqs.append(EqualitySpec(EQ_JOINT, e.get("joint1"), e.get("joint2"), data=data,
                                        sol_params=sol_params(e)))
            elif e.tag == "connect":
                # anchor in body1's frame; its body2-frame position comes from the build pose
                anchor = _floats(e.get("anchor"), 3, [0.0, 0.0, 0.0])
                eqs.append(EqualitySpec(EQ_CONNECT, e.get("body1"), e.get("body2") or "world", data=anchor,
                                        sol_params=sol_params(e)))
            elif e.tag == "weld":
                # anchor: the weld point on body2. relpose: body2's pose in body1's frame; its all-zero quaternion
                # (the default) means the build pose's, otherwise body1's anchor is the point relpose puts there
                rel = _floats(e.get("relpose"), 7, [0.0, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0])
                anchor = _floats(e.get("anchor"), 3, [0.0, 0.0, 0.0])
                ts = float(e.get("torquescale", 1.0))
                b1, b2 = e.get("body1"), e.get("body2") or "world"
                if np.any(rel[3:] != 0):
                    q = rel[3:] / np.linalg.norm(rel[3:])
                    data = np.concatenate([anchor, rel[:3] + quat_to_R(q) @ anchor, q, [ts]])
                    eqs.append(EqualitySpec(EQ_WELD, b1, b2, data=data, sol_params=sol_params(e), explicit=True))
                else:
                    eqs.append(EqualitySpec(EQ_WELD, b1, b2, sol_params=sol_params(e), anchor2=anchor, torquescale=ts))
    return EntitySpec(new_links, eqs, name=os.path.splitext(os.path.basename(path))[0], local_collision_mask=True)


# ---------------------------------------------------------------------------------------------------------
# JSON round trip of EntitySpec (the committed workload assets are stored this way)
# ---------------------------------------------------------------------------------------------------------
def entity_to_json(ent: EntitySpec) -> dict:
    def arr(x):
        return None if x is None else np.asarray(x, dtype=np.float64).tolist()

    return dict(
        name=ent.name, rho=ent.rho, friction=ent.friction, local_collision_mask=ent.local_collision_mask,
        links=[dict(name=l.name, parent=l.parent, pos=arr(l.pos), quat=arr(l.quat), inertial_pos=arr(l.inertial_pos),
                    inertial_quat=arr(l.inertial_quat), inertial_i=arr(l.inertial_i), inertial_mass=l.inertial_mass,
                    joints=[dict(name=j.name, type=j.type, pos=arr(j.pos), axis=arr(j.axis), limit=arr(j.limit),
                                 damping=j.damping, armature=j.armature, stiffness=j.stiffness,
                                 frictionloss=j.frictionloss, kp=j.kp, kv=j.kv, force_range=arr(j.force_range),
                                 sol_params=arr(j.sol_params), init_qpos=arr(j.init_qpos)) for j in l.joints],
                    geoms=[dict(type=g.type, pos=arr(g.pos), quat=arr(g.quat), data=list(map(float, g.data)),
                                friction=g.friction, sol_params=arr(g.sol_params), contype=g.contype,
                                conaffinity=g.conaffinity, verts=arr(g.verts)) for g in l.geoms]) for l in ent.links],
        equalities=[dict(type=e.type, obj1=e.obj1, obj2=e.obj2, data=arr(e.data), sol_params=arr(e.sol_params),
                         explicit=e.explicit, anchor2=arr(e.anchor2), torquescale=e.torquescale)
                    for e in ent.equalities],
    )


def entity_from_json(d: dict) -> EntitySpec:
    def a(x):
        return None if x is None else np.array(x, dtype=np.float64)

    links = []
    for l in d["links"]:
        js = [JointSpec(j["name"], j["type"], a(j["pos"]), a(j["axis"]), a(j["limit"]), j["damping"], j["armature"],
                        j["stiffness"], j["frictionloss"], j["kp"], j["kv"], a(j["force_range"]), a(j["sol_params"]),
                        a(j["init_qpos"])) for j in l["joints"]]
        gs = [GeomSpec(g["type"], a(g["pos"]), a(g["quat"]), tuple(g["data"]), g["friction"], a(g["sol_params"]),
                       g["contype"], g["conaffinity"], a(g.get("verts"))) for g in l["geoms"]]
        links.append(LinkSpec(l["name"], l["parent"], a(l["pos"]), a(l["quat"]), js, gs, a(l["inertial_pos"]),
                              a(l["inertial_quat"]), a(l["inertial_i"]), l["inertial_mass"]))
    eqs = [EqualitySpec(e["type"], e["obj1"], e["obj2"], a(e["data"]), a(e["sol_params"]), e.get("explicit", False),
                        a(e.get("anchor2")), e.get("torquescale", 1.0)) for e in d["equalities"]]
    return EntitySpec(links, eqs, d["rho"], d["friction"], d["name"], d["local_collision_mask"])
