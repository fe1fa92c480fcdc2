"""Known answers for the oracle's constraint rows, build-time inverse weights and box-box multi-contact (no GPU).

Each test builds a scene whose row inputs have closed forms (Jacobian, J v, inverse weight, penetration), reads the
rows the FP64 oracle builds (efc_aref, efc_D, efc_diag, efc_jac) and compares them with numbers worked out by hand
from the reference's definitions:

* impedance and reference acceleration, `imp_aref` (genesis/utils/geom.py:377-398) with the default solver
  parameters (timeconst 0 -> 2 dt, dampratio 1, dmin 0.9, dmax 0.95, width 1e-3, mid 0.5, power 2;
  geom.py:1708-1714 sanitised by rigid_solver_decomp.py:49-77);
* contact rows (constraint_solver_decomp.py:428-525): J v, diag = (iw + mu^2 iw) 2 mu^2 (1 - imp) / imp, D = 1/diag;
* joint-limit rows (constraint_solver_decomp.py:1017-1072) and the JOINT equality row (:641-718), whose quartic
  coupling and residual are evaluated by hand;
* link and dof inverse weights (rigid_solver_decomp.py:354-507): J M^-1 J^T from finite-difference Jacobians and a
  brute-force mass matrix, and the free-body closed form.

Impedance by hand (x = |pos| / width; y = x^2 / mid^(p-1) for x < mid, else 1 - (1-x)^2 / (1-mid)^(p-1);
imp = dmin + y (dmax - dmin), imp = dmax for x > 1):
    x = 0.2 -> y = 0.08 -> imp = 0.904       x = 0.5 -> y = 0.5 -> imp = 0.925
    x = 0.7 -> y = 0.82 -> imp = 0.941       x = 2.5 -> imp = dmax = 0.95
With dt = 0.01 the time constant is 0.02: b = 2 / (0.95 * 0.02), k = 1 / (0.95^2 * 0.02^2), aref = -b J v - k imp pos.
"""
import numpy as np
import pytest

from genesis_amd import workloads as W
from genesis_amd.model import (EQ_JOINT, CompileOptions, EntitySpec, EqualitySpec, GeomSpec, JointSpec, LinkSpec,
                               JOINT_FIXED, JOINT_PRISMATIC, JOINT_REVOLUTE, compile_model, quat_to_R)

IMP = {0.2: 0.904, 0.5: 0.925, 0.7: 0.941, 2.5: 0.95}
WIDTH = 1e-3
B_DAMP = 2.0 / (0.95 * 0.02)
K_SPRING = 1.0 / (0.95 ** 2 * 0.02 ** 2)


@pytest.fixture(scope="module")
def Oracle(oracle_lib):
    return oracle_lib.Oracle


def f32(x):
    """Model tables are stored in float32: the oracle sees the float32 value of every constant."""
    return float(np.float32(x))


@pytest.mark.parametrize("x", [0.2, 0.5, 0.7, 2.5])
def test_contact_rows_known_answer(Oracle, x):
    """A 0.2 m box sinking into the plane at vz = -0.05 m/s with penetration x * width. Every pyramid row of every
    corner contact has J v = vz (the directions d mu - n have unit z and horizontal d), inverse weight 1/m (the plane's
    link has none) and the contact friction mu = max(1.5, 1.0)."""
    pen = x * WIDTH
    vz = -0.05
    ents = [W.plane_entity(), W.box_entity((0.2, 0.2, 0.2), (0, 0, 0.1 - pen), friction=1.5)]
    M = compile_model(ents, CompileOptions(dt=0.01))
    o = Oracle(M, 1, 64)
    o.set("vel", np.array([[0, 0, vz, 0, 0, 0]]))
    o.run(1, stage=4)
    assert o.get("n_contacts")[0] == 4
    c = o.get("contacts")[0]
    np.testing.assert_allclose(c[:, 10], pen, atol=1e-8)  # float32 box extents
    mu = c[0, 11]
    assert mu == f32(1.5)
    m = 200.0 * 0.2 ** 3  # default density 200 kg/m^3
    iw = 1.0 / m
    assert o.get("n_con")[0] == 16 and o.get("n_con_eq")[0] == 0 and o.get("n_con_fl")[0] == 0
    J = o.get("efc_jac")[0]
    v = np.array([0, 0, vz, 0, 0, 0])
    np.testing.assert_allclose(J @ v, vz, rtol=1e-6)
    imp = IMP[x]
    pen_c = c[:, 10].repeat(4)  # the rows of contact i use that contact's own penetration
    aref_expected = -B_DAMP * vz - K_SPRING * imp * (-pen_c)
    diag_expected = (iw + mu * mu * iw) * 2 * mu * mu * (1 - imp) / imp
    np.testing.assert_allclose(o.get("efc_aref")[0], aref_expected, rtol=2e-5)
    np.testing.assert_allclose(o.get("efc_diag")[0], diag_expected, rtol=2e-5)
    np.testing.assert_allclose(o.get("efc_D")[0], 1.0 / diag_expected, rtol=2e-5)
    # the four rows of a contact differ only in the friction direction: same aref, same diag
    a = o.get("efc_aref")[0].reshape(4, 4)
    np.testing.assert_array_equal(a, a[:, :1].repeat(4, 1))


def _hinge(limit=(-0.5, 0.5), length=0.5, mass=1.0):
    base = LinkSpec("base", -1, np.zeros(3), np.array([1.0, 0, 0, 0]), [JointSpec("fix", JOINT_FIXED)], [])
    arm = LinkSpec("arm", 0, np.array([0, 0, 1.0]), np.array([1.0, 0, 0, 0]),
                   [JointSpec("hinge", JOINT_REVOLUTE, axis=np.array([0, 1.0, 0]), limit=np.array([limit]))], [],
                   inertial_pos=np.array([length, 0, 0]), inertial_quat=np.array([1.0, 0, 0, 0]),
                   inertial_i=np.eye(3) * 1e-6, inertial_mass=mass)
    return EntitySpec([base, arm], name="hinge")


@pytest.mark.parametrize("x", [0.2, 0.7, 2.5])
@pytest.mark.parametrize("side", [1, -1])
def test_joint_limit_row_known_answer(Oracle, x, side):
    """A hinge past its upper (side 1) or lower (side -1) limit of +-0.5 rad by x * width, turning at w = 0.3 rad/s.
    pos_delta = -x width; the row's Jacobian is -1 past the upper limit and +1 past the lower one; the dof inverse
    weight is 1 / (m l^2 + I_yy) for the point-mass hinge."""
    M = compile_model([_hinge()], CompileOptions(dt=0.01, enable_collision=False))
    o = Oracle(M, 1, 64)
    delta = x * WIDTH
    w = 0.3
    o.set("qpos", np.array([[side * (0.5 + delta)]]))
    o.set("vel", np.array([[w]]))
    o.run(1, stage=4)
    assert o.get("n_con")[0] == 1
    jac = -1.0 if side == 1 else 1.0
    np.testing.assert_allclose(o.get("efc_jac")[0], [[jac]])
    # pos_delta in float arithmetic on the float32 limit
    pos_delta = f32(0.5) - (0.5 + delta) if side == 1 else -(0.5 + delta) - f32(-0.5)
    imp = IMP[x]
    iw = 1.0 / (f32(1.0) * f32(0.5) ** 2 + f32(1e-6))
    assert M.dof_invweight[0] == pytest.approx(iw, rel=1e-6)
    aref_expected = -B_DAMP * (jac * w) - K_SPRING * imp * pos_delta
    diag_expected = iw * (1 - imp) / imp
    np.testing.assert_allclose(o.get("efc_aref")[0], [aref_expected], rtol=2e-5)
    np.testing.assert_allclose(o.get("efc_diag")[0], [diag_expected], rtol=2e-5)
    np.testing.assert_allclose(o.get("efc_D")[0], [1.0 / diag_expected], rtol=2e-5)


POLY = np.array([0.01, 0.5, -0.3, 0.2, 0.1])  # y - y0 = a0 + a1 d + a2 d^2 + a3 d^3 + a4 d^4


def _slides(m1=1.5, m2=0.5):
    """Slide x carrying slide y: M = [[m1 + m2, 0], [0, m2]], so the dof inverse weights are 1/(m1+m2) and 1/m2."""
    I3 = np.eye(3) * 1e-3
    base = LinkSpec("base", -1, np.zeros(3), np.array([1.0, 0, 0, 0]), [JointSpec("fix", JOINT_FIXED)], [])
    l1 = LinkSpec("l1", 0, np.zeros(3), np.array([1.0, 0, 0, 0]),
                  [JointSpec("j1", JOINT_PRISMATIC, axis=np.array([1.0, 0, 0]))], [],
                  inertial_pos=np.zeros(3), inertial_quat=np.array([1.0, 0, 0, 0]), inertial_i=I3, inertial_mass=m1)
    l2 = LinkSpec("l2", 1, np.zeros(3), np.array([1.0, 0, 0, 0]),
                  [JointSpec("j2", JOINT_PRISMATIC, axis=np.array([0, 1.0, 0]))], [],
                  inertial_pos=np.zeros(3), inertial_quat=np.array([1.0, 0, 0, 0]), inertial_i=I3, inertial_mass=m2)
    eq = EqualitySpec(EQ_JOINT, "j1", "j2", data=np.concatenate([POLY, np.zeros(6)]))
    return EntitySpec([base, l1, l2], equalities=[eq], name="slides")


@pytest.mark.parametrize("resid", [2e-4, -7e-4, 2.5e-3])
def test_joint_equality_row_known_answer(Oracle, resid):
    """JOINT equality q1 - q1_0 = poly(q2 - q2_0) on two orthogonal slides. At q2 = 0.2 (hand-evaluated):
    poly(0.2) = 0.01 + 0.1 - 0.012 + 0.0016 + 0.00016 = 0.09976, poly'(0.2) = 0.5 - 0.12 + 0.024 + 0.0032 = 0.4072.
    q1 = 0.09976 + resid gives the residual `resid`; the row is [1, -poly'], J v = v1 - poly' v2, aref =
    -b J v - k imp resid with imp from |resid| / width, diag = (iw1 + iw2) (1 - imp) / imp."""
    M = compile_model([_slides()], CompileOptions(dt=0.01, enable_collision=False, gravity=(0.0, 0.0, 0.0)))
    o = Oracle(M, 1, 64)
    q2 = 0.2
    a = [f32(c) for c in POLY]
    poly = sum(a[i] * q2 ** i for i in range(5))
    deriv = sum(a[i + 1] * (i + 1) * q2 ** i for i in range(4))
    assert poly == pytest.approx(0.09976, abs=1e-8) and deriv == pytest.approx(0.4072, abs=1e-8)
    v = np.array([0.07, -0.2])
    o.set("qpos", np.array([[poly + resid, q2]]))
    o.set("vel", v[None])
    o.run(1, stage=4)
    assert o.get("n_con")[0] == 1 and o.get("n_con_eq")[0] == 1
    J = o.get("efc_jac")[0]
    np.testing.assert_allclose(J, [[1.0, -deriv]], rtol=1e-7)
    iw1, iw2 = 1.0 / (1.5 + 0.5), 1.0 / 0.5
    np.testing.assert_allclose(M.dof_invweight, [iw1, iw2], rtol=1e-6)
    imp = IMP[round(abs(resid) / WIDTH, 6)]
    jv = v[0] - deriv * v[1]
    aref_expected = -B_DAMP * jv - K_SPRING * imp * resid
    diag_expected = (iw1 + iw2) * (1 - imp) / imp
    np.testing.assert_allclose(o.get("efc_aref")[0], [aref_expected], rtol=1e-4)
    np.testing.assert_allclose(o.get("efc_diag")[0], [diag_expected], rtol=2e-5)
    np.testing.assert_allclose(o.get("efc_D")[0], [1.0 / diag_expected], rtol=2e-5)


def _link_jacobians(M, o, b, eps=1e-6):
    """Central-difference Jacobians (6 x D: COM linear velocity, angular velocity) of every link at the env's qpos
    (hinge / slide models, so qpos and dofs coincide)."""
    q0 = o.get("qpos", [b])[0].copy()
    D = M.n_dofs
    assert M.n_qs == D

    def pose(q):
        o.set("qpos", q[None], [b])
        o.run(stage=1, envs=(b, b + 1))
        p, qt = o.get("links_pos", [b])[0], o.get("links_quat", [b])[0]
        R = [quat_to_R(qt[l]) for l in range(M.n_links)]
        c = [p[l] + R[l] @ M.link_inertial_pos[l] for l in range(M.n_links)]
        return c, R

    c0, R0 = pose(q0)
    J = np.zeros((M.n_links, 6, D))
    for d in range(D):
        qp, qm = q0.copy(), q0.copy()
        qp[d] += eps
        qm[d] -= eps
        cp, Rp = pose(qp)
        cm, Rm = pose(qm)
        for l in range(M.n_links):
            J[l, :3, d] = (cp[l] - cm[l]) / (2 * eps)
            Wm = (Rp[l] - Rm[l]) / (2 * eps) @ R0[l].T
            J[l, 3:, d] = [Wm[2, 1], Wm[0, 2], Wm[1, 0]]
    o.set("qpos", q0[None], [b])
    return J, R0


def test_invweight_matches_brute_force_on_arm(Oracle):
    """rigid_solver_decomp.py:354-507 at qpos0: link invweight = (mean of the linear, mean of the angular diagonal
    of J M^-1 J^T) for the link's COM Jacobian; 1-dof joints' dof invweight = diag(M^-1). M is assembled by brute
    force from the finite-difference Jacobians, sum m Jp^T Jp + Jr^T I_w Jr, plus armature (no implicit damping)."""
    w = W.arm_workload()
    M = compile_model(w.entities, w.options)
    o = Oracle(M, 1, 64)
    o.set("qpos", M.qpos0[None].astype(np.float64))
    J, R0 = _link_jacobians(M, o, 0)
    D = M.n_dofs
    Mb = np.zeros((D, D))
    for l in range(M.n_links):
        R = R0[l] @ quat_to_R(M.link_inertial_quat[l])
        Iw = R @ M.link_inertial_i[l].reshape(3, 3).astype(np.float64) @ R.T
        Mb += float(M.link_inertial_mass[l]) * J[l, :3].T @ J[l, :3] + J[l, 3:].T @ Iw @ J[l, 3:]
    Mb += np.diag(M.dof_armature.astype(np.float64))
    Minv = np.linalg.inv(Mb)
    for l in range(M.n_links):
        A = np.diag(J[l] @ Minv @ J[l].T)
        np.testing.assert_allclose(M.link_invweight[l], [A[:3].mean(), A[3:].mean()], rtol=2e-5, atol=1e-9,
                                   err_msg=f"link {l}")
    np.testing.assert_allclose(M.dof_invweight, np.diag(Minv), rtol=2e-5)
    # the arm's links carry real weight: a plausible mistake (the link origin instead of its COM, angular terms
    # dropped) changes these by far more than the tolerance
    assert (M.link_invweight[2:, 0] > 0).all() and (M.link_invweight[2:, 1] > 0).all()


def test_invweight_free_box_closed_form():
    """A free box (joint at its COM): link invweight = (1/m, mean(1/I_xx, 1/I_yy, 1/I_zz)), dof invweight 1/m on the
    three translations and the same angular mean on the three rotations; the fixed plane's link has none. Mass and
    inertia start from EPS before the geoms are added (rigid_link.py:160-162)."""
    M = compile_model([W.plane_entity(), W.box_entity((0.1, 0.2, 0.3), (0, 0, 1.0))], CompileOptions(dt=0.01))
    eps = float(np.finfo(np.float32).eps)
    mb = 200.0 * 0.1 * 0.2 * 0.3
    m = mb + eps
    I = mb / 12.0 * np.array([0.2 ** 2 + 0.3 ** 2, 0.1 ** 2 + 0.3 ** 2, 0.1 ** 2 + 0.2 ** 2]) + eps
    ang = np.mean(1.0 / I)
    np.testing.assert_allclose(M.link_invweight[1], [1 / m, ang], rtol=1e-5)
    np.testing.assert_allclose(M.link_invweight[0], [0.0, 0.0], atol=0)
    np.testing.assert_allclose(M.dof_invweight, [1 / m] * 3 + [ang] * 3, rtol=1e-5)


@pytest.mark.parametrize("pen", [1e-3, 4e-3])
def test_box_on_box_multi_contact_known_answer(Oracle, pen):
    """A 0.2 x 0.3 x 0.1 box resting centred and axis-aligned on a fixed 0.6 x 0.6 x 0.2 box, sunk by `pen`
    (collider_decomp.py:2328-2725: MPR, then 4 detections with geom a rotated by +-mc_perturbation about two
    axes orthogonal to the normal, rotated back, de-duplicated within 0.5 * mc_tolerance * min AABB diagonal):
    the first contact sits on the axis of the boxes with depth pen; the four perturbed detections land on the small
    box's four bottom corners, each at depth pen, at z = 0.2 - pen / 2 (midway between the two surfaces), with
    normal -z to first order (|n| - 1 = O(mc_perturbation^2))."""
    big = W.box_entity((0.6, 0.6, 0.2), (0, 0, 0.1), fixed=True)
    small = W.box_entity((0.2, 0.3, 0.1), (0, 0, 0.25 - pen))
    M = compile_model([big, small], CompileOptions(dt=0.01, gravity=(0.0, 0.0, 0.0)))
    o = Oracle(M, 1, 64)
    o.run(1, stage=3)
    c = o.get("contacts")[0]
    assert len(c) == 5
    np.testing.assert_allclose(c[0, 10], pen, rtol=1e-5)
    # the perturbed depths are measured along the first-order corrected normal: pen (1 + O(mc_perturbation^2))
    np.testing.assert_allclose(c[1:, 10], pen, rtol=2e-4)
    np.testing.assert_allclose(c[:, 6], 0.2 - pen / 2, atol=1e-6)
    np.testing.assert_allclose(c[0, 7:10], [0, 0, -1], atol=1e-6)
    np.testing.assert_allclose(c[1:, 7:10], np.tile([0, 0, -1.0], (4, 1)), atol=2e-4)
    assert np.hypot(c[0, 4], c[0, 5]) < 0.16  # inside the overlap
    corners = sorted(map(tuple, np.sign(c[1:, 4:6]).astype(int).tolist()))
    assert corners == [(-1, -1), (-1, 1), (1, -1), (1, 1)]
    np.testing.assert_allclose(np.abs(c[1:, 4:6]), np.tile([0.1, 0.15], (4, 1)), atol=1e-4)


def test_multi_contact_dedup_keeps_one_point_contact(Oracle):
    """De-duplication radius: a sphere resting on a box touches at one point; the four perturbed detections find
    that point again (within far less than the radius 0.5 * 1e-2 * min AABB diagonal), so exactly one contact is
    kept. A missing or zero radius would keep five."""
    from genesis_amd.model import GEOM_SPHERE, JOINT_FREE
    big = W.box_entity((0.6, 0.6, 0.2), (0, 0, 0.1), fixed=True)
    pos = np.array([0.05, -0.02, 0.2 + 0.1 - 5e-4])
    ball = EntitySpec([LinkSpec("b", -1, pos, np.array([1.0, 0, 0, 0]),
                                [JointSpec("f", JOINT_FREE, init_qpos=np.concatenate([pos, [1, 0, 0, 0]]))],
                                [GeomSpec(GEOM_SPHERE, data=(0.1,))])])
    M = compile_model([big, ball], CompileOptions(dt=0.01, gravity=(0.0, 0.0, 0.0)))
    o = Oracle(M, 1, 64)
    o.run(1, stage=3)
    c = o.get("contacts")[0]
    assert len(c) == 1
    np.testing.assert_allclose(c[0, 10], 5e-4, rtol=1e-3)
    np.testing.assert_allclose(c[0, 4:6], pos[:2], atol=1e-6)


def _mesh_model(verts, pos=(0.0, 0.0, 1.0), gravity=(0.0, 0.0, -9.81)):
    from genesis_amd.model import GEOM_MESH, JOINT_FREE
    from genesis_amd.meshes import convex_hull_vertices
    hull = convex_hull_vertices(verts)
    pos = np.asarray(pos, float)
    body = EntitySpec([LinkSpec("m", -1, pos, np.array([1.0, 0, 0, 0]),
                                [JointSpec("f", JOINT_FREE, init_qpos=np.concatenate([pos, [1, 0, 0, 0]]))],
                                [GeomSpec(GEOM_MESH, verts=hull)])], rho=500.0)
    return compile_model([W.plane_entity(), body], CompileOptions(dt=0.01, gravity=gravity)), hull


def test_mesh_support_field_known_answers(Oracle):
    """support_field_decomp.py:24-94, 145-186 on a random convex hull. Along a grid direction (theta_i, phi_j) the
    lookup's four candidate cells include that direction's own cell, whose precomputed vertex is the exact
    maximiser, so the lookup attains the brute-force maximum projection; along any direction it returns a hull
    vertex, its index matches, and it falls short of the maximum by at most the grid's angular resolution."""
    from genesis_amd.meshes import SUPPORT_RES
    rng = np.random.default_rng(3)
    M, hull = _mesh_model(rng.normal(size=(300, 3)) * [0.1, 0.2, 0.05])
    g = 1
    v32 = M.verts[M.geom_vert_start[g]:M.geom_vert_end[g]].astype(np.float64)
    o = Oracle(M, 1, 64)
    R = np.linalg.norm(v32, axis=1).max()
    for i in range(0, SUPPORT_RES, 7):
        for j in range(3, SUPPORT_RES - 3, 11):
            th, ph = 2 * np.pi * i / SUPPORT_RES - np.pi, np.pi * j / SUPPORT_RES
            d = np.array([np.sin(ph) * np.cos(th), np.sin(ph) * np.sin(th), np.cos(ph)])
            v, vid = o.support_mesh(g, d)
            np.testing.assert_array_equal(v, v32[vid])
            assert v @ d >= (v32 @ d).max() - 1e-12, (i, j)
    step = 2 * np.pi / SUPPORT_RES
    for d in rng.normal(size=(400, 3)):
        d /= np.linalg.norm(d)
        v, vid = o.support_mesh(g, d)
        np.testing.assert_array_equal(v, v32[vid])
        assert v @ d >= (v32 @ d).max() - 2 * R * step


def test_mesh_cube_on_plane_known_answer(Oracle):
    """A cube given as a mesh (its 8 corners, through the support field instead of the box closed form) sunk 2 mm
    into the plane: the deepest-vertex contact and the perturbed detections give the 4 bottom corners, each at
    depth 2 mm, normal -z, contact points midway (z = -1 mm); a missing de-duplication would keep more."""
    import itertools
    s, pen = 0.1, 2e-3
    corners = np.array(list(itertools.product([-s, s], repeat=3)))
    M, _ = _mesh_model(corners, pos=(0.0, 0.0, s - pen), gravity=(0.0, 0.0, 0.0))
    o = Oracle(M, 1, 64)
    o.run(1, stage=3)
    c = o.get("contacts")[0]
    assert len(c) == 4
    np.testing.assert_allclose(c[0, 10], pen, atol=1e-8)  # float32 vertices
    # perturbed detections measure along the first-order corrected normal: pen (1 + O(mc_perturbation^2))
    np.testing.assert_allclose(c[1:, 10], pen, rtol=2e-4)
    np.testing.assert_allclose(c[:, 6], -pen / 2, atol=1e-6)
    np.testing.assert_allclose(c[:, 7:10], np.tile([0, 0, -1.0], (4, 1)), atol=2e-4)
    xy = sorted(map(tuple, np.round(c[:, 4:6] / s).astype(int).tolist()))
    assert xy == [(-1, -1), (-1, 1), (1, -1), (1, 1)]


def _two_arms(eqs):
    """Two 2-dof hinge arms on fixed bases (hinges about y then x), so every dof is an angle and a row's Jacobian is
    the derivative of its residual with respect to qpos."""
    def arm(name, base):
        fix = LinkSpec(f"{name}0", -1, np.asarray(base, float), np.array([1.0, 0, 0, 0]), [JointSpec("f", JOINT_FIXED)], [])
        l1 = LinkSpec(f"{name}1", 0, np.array([0, 0, 0.3]), np.array([1.0, 0, 0, 0]),
                      [JointSpec(f"{name}h1", JOINT_REVOLUTE, axis=np.array([0, 1.0, 0]))], [],
                      inertial_pos=np.array([0.2, 0, 0]), inertial_quat=np.array([1.0, 0, 0, 0]),
                      inertial_i=np.eye(3) * 1e-3, inertial_mass=1.0)
        l2 = LinkSpec(f"{name}2", 1, np.array([0.4, 0, 0]), np.array([1.0, 0, 0, 0]),
                      [JointSpec(f"{name}h2", JOINT_REVOLUTE, axis=np.array([1.0, 0, 0]))], [],
                      inertial_pos=np.array([0.1, 0.05, 0]), inertial_quat=np.array([1.0, 0, 0, 0]),
                      inertial_i=np.eye(3) * 1e-3, inertial_mass=0.5)
        return [fix, l1, l2]
    a = EntitySpec(arm("a", [0, 0, 0]), equalities=eqs, name="a")
    b = EntitySpec(arm("b", [0.9, 0.2, 0]), name="b")
    return compile_model([a, b], CompileOptions(dt=0.01, enable_collision=False, enable_joint_limit=False))


def _anchor_gap(M, q, l1, a1, l2, a2):
    from genesis_amd.model import _host_fk_crba, transform_by_quat
    lp, lq = _host_fk_crba(M, q)[:2]
    return lp[l1] + transform_by_quat(a1, lq[l1]) - (lp[l2] + transform_by_quat(a2, lq[l2])), lq


def test_connect_rows_known_answer(Oracle):
    """CONNECT (constraint_solver_decomp.py:528-641): anchor data[0:3] on link a2 and data[3:6] on link b2 (taken from
    the build pose, where the two points coincide). At another pose each row i is the derivative of the anchor gap's
    component i with respect to the joint angles (central differences of the float64 kinematics), and aref is
    imp_aref(-|gap|, J qvel, gap_i) evaluated by hand."""
    from genesis_amd.model import EQ_CONNECT
    M = _two_arms([EqualitySpec(EQ_CONNECT, "a2", "b2", data=np.array([0.15, 0.0, 0.0]))])
    l1, l2 = M.eq_obj1[0], M.eq_obj2[0]
    d = M.eq_data[0].astype(np.float64)
    gap0, _ = _anchor_gap(M, M.qpos0.astype(np.float64), l1, d[0:3], l2, d[3:6])
    np.testing.assert_allclose(gap0, 0, atol=1e-6)
    q = np.array([0.3, -0.2, 0.1, 0.25])
    qd = np.array([0.4, -0.7, 0.2, 0.5])
    o = Oracle(M, 1, 64)
    o.set("qpos", q[None])
    o.set("vel", qd[None])
    o.run(1, stage=4)
    assert o.get("n_con_eq")[0] == 3
    J = o.get("efc_jac")[0][:3]
    h = 1e-6
    Jfd = np.stack([(_anchor_gap(M, q + h * np.eye(4)[k], l1, d[0:3], l2, d[3:6])[0]
                     - _anchor_gap(M, q - h * np.eye(4)[k], l1, d[0:3], l2, d[3:6])[0]) / (2 * h) for k in range(4)], 1)
    np.testing.assert_allclose(J, Jfd, atol=1e-7)
    gap, _ = _anchor_gap(M, q, l1, d[0:3], l2, d[3:6])
    sp = M.eq_sol_params[0].astype(np.float64)
    x = min(np.linalg.norm(gap) / sp[4], 1.0)
    y = x ** 2 / sp[5] if x < sp[5] else 1 - (1 - x) ** 2 / (1 - sp[5])
    imp = sp[2] + y * (sp[3] - sp[2]) if x < 1 else sp[3]
    tc = max(sp[0], 2 * f32(0.01))
    b, k = 2 / (sp[3] * tc), 1 / (sp[3] ** 2 * tc ** 2 * sp[1] ** 2)
    np.testing.assert_allclose(o.get("efc_aref")[0][:3], -b * (J @ qd) - k * imp * gap, rtol=1e-6)
    iw = M.link_invweight[l1, 0] + M.link_invweight[l2, 0]
    np.testing.assert_allclose(o.get("efc_diag")[0][:3], iw * (1 - imp) / imp, rtol=1e-6)


def test_weld_rows_known_answer(Oracle):
    """WELD (constraint_solver_decomp.py:824-1015) of b2 to a1 from the build pose (anchor at a1's origin, relative
    quaternion inv(q_a1) q_b2, torque scale 1): the position rows are the derivatives of the anchor gap, the
    orientation rows the derivatives of the vector part of the error quaternion inv(q_b2) q_a1 rel (its time
    derivative is 0.5 inv(q_b2) (0, w_a1 - w_b2) q_a1 rel, which the rows restate), both by central differences;
    at the build pose every residual is zero."""
    from genesis_amd.model import EQ_WELD, quat_mul
    M = _two_arms([EqualitySpec(EQ_WELD, "a1", "b2")])
    l1, l2 = M.eq_obj1[0], M.eq_obj2[0]
    d = M.eq_data[0].astype(np.float64)
    np.testing.assert_allclose(d[3:6], 0, atol=1e-7)
    assert d[10] == 1.0

    def residual(q):
        gap, lq = _anchor_gap(M, q, l1, d[3:6], l2, d[0:3])
        e = quat_mul(np.array([lq[l2][0], -lq[l2][1], -lq[l2][2], -lq[l2][3]]), quat_mul(lq[l1], d[6:10]))
        return np.concatenate([gap, e[1:] * d[10]])

    np.testing.assert_allclose(residual(M.qpos0.astype(np.float64)), 0, atol=1e-6)
    q = np.array([0.3, -0.2, 0.1, 0.25])
    o = Oracle(M, 1, 64)
    o.set("qpos", q[None])
    o.run(1, stage=4)
    assert o.get("n_con_eq")[0] == 6
    J = o.get("efc_jac")[0][:6]
    h = 1e-6
    Jfd = np.stack([(residual(q + h * np.eye(4)[k]) - residual(q - h * np.eye(4)[k])) / (2 * h) for k in range(4)], 1)
    np.testing.assert_allclose(J, Jfd, atol=1e-7)
    # orientation rows weigh by the rotational inverse weights, position rows by the translational ones
    sp = M.eq_sol_params[0].astype(np.float64)
    r = residual(q)
    x = min(np.linalg.norm(r) / sp[4], 1.0)
    imp = sp[3] if x >= 1 else None
    assert imp is not None  # the residual is far beyond the impedance width here
    iw0 = M.link_invweight[l1, 0] + M.link_invweight[l2, 0]
    iw1 = M.link_invweight[l1, 1] + M.link_invweight[l2, 1]
    np.testing.assert_allclose(o.get("efc_diag")[0][:6], np.repeat([iw0, iw1], 3) * (1 - imp) / imp, rtol=1e-6)


def test_welded_box_falls_with_its_carrier(Oracle):
    """A free box welded (from the build pose) to a second free box: under gravity the pair falls together like one
    free body (the semi-implicit Euler drop g dt^2 n (n + 1) / 2 after n steps) and keeps its relative pose."""
    from genesis_amd.model import EQ_WELD
    a = W.box_entity((0.1, 0.1, 0.1), (0, 0, 2.0))
    b = W.box_entity((0.1, 0.1, 0.1), (0.15, 0.0, 2.05))
    a.links[0].name, b.links[0].name = "boxa", "boxb"
    a.equalities = [EqualitySpec(EQ_WELD, "boxa", "boxb")]
    M = compile_model([a, b], CompileOptions(dt=0.01, enable_collision=False))
    o = Oracle(M, 1, 64)
    for _ in range(40):
        o.run(1)
    q = o.get("qpos")[0]
    drop = 9.81 * 0.01 ** 2 * 40 * 41 / 2
    np.testing.assert_allclose([2.0 - q[2], 2.05 - q[9]], drop, rtol=1e-5)
    np.testing.assert_allclose(q[7:10] - q[0:3], [0.15, 0.0, 0.05], atol=1e-6)
    np.testing.assert_allclose(q[10:14], q[3:7], atol=1e-6)


def test_box_welded_to_the_world_holds_against_gravity(Oracle):
    """A free box welded to the plane's fixed link 1 m up, and one connected to it at a corner: after 2 s both are at
    rest within a few millimetres of the build pose (the soft constraint settles where its spring balances gravity),
    and the connected box hangs from its corner (its COM below the anchor)."""
    from genesis_amd.model import EQ_CONNECT, EQ_WELD
    plane = W.plane_entity()
    a = W.box_entity((0.1, 0.1, 0.1), (0.0, 0.0, 1.0))
    b = W.box_entity((0.1, 0.1, 0.1), (0.5, 0.0, 1.0))
    a.links[0].name, b.links[0].name = "boxa", "boxb"
    plane.equalities = [EqualitySpec(EQ_WELD, "plane_baselink", "boxa"),
                        EqualitySpec(EQ_CONNECT, "plane_baselink", "boxb", data=np.array([0.55, 0.05, 1.05]))]
    M = compile_model([plane, a, b], CompileOptions(dt=0.01, enable_collision=False))
    o = Oracle(M, 1, 64)
    for _ in range(200):
        o.run(1)
    q, v = o.get("qpos")[0], o.get("vel")[0]
    np.testing.assert_allclose(q[0:3], [0, 0, 1.0], atol=5e-3)
    np.testing.assert_allclose(v[0:6], 0, atol=1e-3)
    # the connected box's corner stays at the anchor; the box swings below it
    from genesis_amd.model import transform_by_quat
    corner = q[7:10] + transform_by_quat(np.array([0.05, 0.05, 0.05]), q[10:14])
    np.testing.assert_allclose(corner, [0.55, 0.05, 1.05], atol=5e-3)
    assert q[9] < 1.05 - 0.05


MJCF_EQ = """<mujoco>
  <worldbody>
    <body name="post" pos="0 0 1">
      <inertial pos="0 0 0" mass="1" diaginertia="0.01 0.01 0.01"/>
      <body name="arm" pos="0 0 0">
        <joint name="hinge" type="hinge" axis="0 1 0"/>
        <inertial pos="0.2 0 0" mass="1" diaginertia="0.01 0.01 0.01"/>
      </body>
    </body>
    <body name="bob" pos="0.4 0 1">
      <freejoint/>
      <inertial pos="0 0 0" mass="0.5" diaginertia="0.002 0.002 0.002"/>
    </body>
    <body name="lid" pos="0.4 0.3 1.2" quat="0.9238795 0 0 0.3826834">
      <freejoint/>
      <inertial pos="0 0 0" mass="0.2" diaginertia="0.001 0.001 0.001"/>
    </body>
    <body name="cap" pos="-0.5 0 0.5">
      <freejoint/>
      <inertial pos="0 0 0" mass="0.3" diaginertia="0.001 0.001 0.001"/>
    </body>
  </worldbody>
  <equality>
    <connect body1="arm" body2="bob" anchor="0.4 0 0"/>
    <weld body1="bob" body2="lid" anchor="0.05 0 0" torquescale="0.5"/>
    <weld body1="lid" body2="cap" anchor="0 0.1 0" relpose="0 0 -0.2 1 0 0 1"/>
    <connect body1="cap" anchor="0 0 0.1" solref="0.05 1"/>
  </equality>
</mujoco>
"""


def test_mjcf_connect_and_weld_compile_to_build_pose_data(tmp_path, Oracle):
    """MJCF <connect> / <weld> (mjcf.py:700-725 takes them from MuJoCo's compiled eq_data): body1's CONNECT anchor in
    data[0:3] and its body2-frame image at the build pose in data[3:6]; a WELD without relpose welds body2's anchor
    where it is at the build pose (zero residual there, torquescale kept); a WELD with relpose puts body1's anchor at
    relpose's position plus the rotated anchor, with relpose's normalised quaternion; an omitted body2 is the world
    frame (a fixed link at the origin). The build data survive the JSON round trip, and the oracle's CONNECT to the
    world holds the free body it carries up against gravity."""
    from genesis_amd.loaders import entity_from_json, entity_to_json, parse_mjcf
    from genesis_amd.model import EQ_CONNECT, EQ_WELD, quat_mul
    p = tmp_path / "eq.xml"
    p.write_text(MJCF_EQ)
    ent = parse_mjcf(str(p))
    assert [e.type for e in ent.equalities] == [EQ_CONNECT, EQ_WELD, EQ_WELD, EQ_CONNECT]
    M = compile_model([ent], CompileOptions(dt=0.01, enable_collision=False))
    d = M.eq_data.astype(np.float64)
    q0 = M.qpos0.astype(np.float64)
    for i in (0, 3):  # CONNECT: body1's anchor and body2's image coincide at the build pose
        np.testing.assert_allclose(_anchor_gap(M, q0, M.eq_obj1[i], d[i, 0:3], M.eq_obj2[i], d[i, 3:6])[0], 0,
                                   atol=1e-6)
    np.testing.assert_allclose(d[0, 0:3], [0.4, 0, 0])
    np.testing.assert_allclose(d[3, 0:3], [0, 0, 0.1])
    np.testing.assert_allclose(d[3, 3:6], [-0.5, 0, 0.6], atol=1e-6)  # the world frame is the identity
    # WELD without relpose: anchor on body2, zero position and orientation residual at the build pose
    np.testing.assert_allclose(d[1, 0:3], [0.05, 0, 0])
    gap, lq = _anchor_gap(M, q0, M.eq_obj1[1], d[1, 3:6], M.eq_obj2[1], d[1, 0:3])
    np.testing.assert_allclose(gap, 0, atol=1e-6)
    l1, l2 = M.eq_obj1[1], M.eq_obj2[1]
    err = quat_mul(np.array([lq[l2][0], -lq[l2][1], -lq[l2][2], -lq[l2][3]]), quat_mul(lq[l1], d[1, 6:10]))
    np.testing.assert_allclose(np.abs(err), [1, 0, 0, 0], atol=1e-6)
    assert d[1, 10] == 0.5
    # WELD with relpose (a 90 degree turn about z, quaternion given unnormalised): body1's anchor = relpose position
    # + the turned anchor (0, 0.1, 0) -> (-0.1, 0, 0)
    s = np.sqrt(0.5)
    np.testing.assert_allclose(d[2, 6:10], [s, 0, 0, s], atol=1e-7)
    np.testing.assert_allclose(d[2, 3:6], [-0.1, 0, -0.2], atol=1e-7)
    np.testing.assert_allclose(d[2, 0:3], [0, 0.1, 0])
    assert M.eq_sol_params[3][0] == np.float32(0.05)
    M2 = compile_model([entity_from_json(entity_to_json(ent))], CompileOptions(dt=0.01, enable_collision=False))
    np.testing.assert_array_equal(M2.eq_data, M.eq_data)
    # under gravity the chain (hinged arm - bob - lid - cap - world) holds together: after 2 s every anchor pair is
    # within a few millimetres
    o = Oracle(M, 1, 64)
    for _ in range(200):
        o.run(1)
    q = o.get("qpos")[0].astype(np.float64)
    for i, (a1, a2) in enumerate(((0, 3), (3, 0), (3, 0), (0, 3))):
        gap = _anchor_gap(M, q, M.eq_obj1[i], d[i, a1:a1 + 3], M.eq_obj2[i], d[i, a2:a2 + 3])[0]
        assert np.linalg.norm(gap) < 5e-3, (i, gap)
