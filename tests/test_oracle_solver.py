"""Pins the oracle's constraint rows, Newton solver, bias forces and integrator against independent mathematics.

* Newton (constraint_solver_decomp.py:1409-1894, 2047-2429): the solver minimises MuJoCo's convex primal cost
  ½(a-a0)ᵀM(a-a0) + Σ s_i(J_i a - aref_i) (equality rows: ½Dr²; contact/limit rows: ½Dr² when r < 0; friction-loss
  rows: Huber with threshold R·f). The cost is written here from that published definition, minimised with scipy's
  trust-region Newton, and the oracle's qacc must be its minimiser (the cost is C¹ and strictly convex, so a zero
  gradient is the global optimum).
* Contact rows (constraint_solver_decomp.py:428-525): each row times any joint velocity equals a fixed direction
  dotted with the relative velocity of the contact point (finite differences of forward kinematics), and the four
  directions of a contact form the friction pyramid -n ± μ t1, -n ± μ t2.
* Bias forces (rigid_solver_decomp.py:6180-6485): qf_smooth = -(C(q, q̇) q̇ + ∂V/∂q) from the Lagrangian, with the
  Christoffel symbols taken from finite differences of the (separately pinned) mass matrix.
* Integrator (rigid_solver_decomp.py:6566-6708): a torque-free cube with angular velocity ω turns by exp([ω]× dt).
  Free-joint angular velocity is expressed in the body frame: the new orientation is quatmul(rot0, exp(ω dt))
  (rigid_solver_decomp.py:6685-6686, geometry helper geom.py:218-224), i.e. R' = R exp([ω]× dt).
"""
import numpy as np
import pytest
import scipy.linalg as sl
import scipy.optimize as so

from genesis_amd import workloads as W
from genesis_amd.model import (CompileOptions, EntitySpec, GeomSpec, JointSpec, LinkSpec, GEOM_BOX, JOINT_FIXED,
                               JOINT_PRISMATIC, JOINT_REVOLUTE, compile_model, quat_to_R)


@pytest.fixture(scope="module")
def Oracle(oracle_lib):
    return oracle_lib.Oracle


def _skew(w):
    return np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]])


def _R_to_quat(R):
    w = np.sqrt(max(0.0, 1 + R[0, 0] + R[1, 1] + R[2, 2])) / 2
    x = np.copysign(np.sqrt(max(0.0, 1 + R[0, 0] - R[1, 1] - R[2, 2])) / 2, R[2, 1] - R[1, 2])
    y = np.copysign(np.sqrt(max(0.0, 1 - R[0, 0] + R[1, 1] - R[2, 2])) / 2, R[0, 2] - R[2, 0])
    z = np.copysign(np.sqrt(max(0.0, 1 - R[0, 0] - R[1, 1] + R[2, 2])) / 2, R[1, 0] - R[0, 1])
    return np.array([w, x, y, z])


def _primal(o, b):
    """Convex primal cost of env b from the oracle's rows: returns (f, grad, hess, a0)."""
    J = o.get("efc_jac", [b])[0]
    Dd = o.get("efc_D", [b])[0]
    aref = o.get("efc_aref", [b])[0]
    R = o.get("efc_diag", [b])[0]
    fl = o.get("efc_fl", [b])[0]
    ne, nf = int(o.get("n_con_eq", [b])[0]), int(o.get("n_con_fl", [b])[0])
    Mm = o.get("mass_mat", [b])[0]
    a0 = o.get("acc_smooth", [b])[0]
    kind = np.full(len(Dd), 2)
    kind[:ne] = 0
    kind[ne:ne + nf] = 1

    def parts(a):
        r = J @ a - aref
        # per-row cost, slope and curvature
        c = np.zeros_like(r)
        s = np.zeros_like(r)
        k = np.zeros_like(r)
        q = kind == 0
        c[q], s[q], k[q] = 0.5 * Dd[q] * r[q] ** 2, Dd[q] * r[q], Dd[q]
        q = (kind == 2) & (r < 0)
        c[q], s[q], k[q] = 0.5 * Dd[q] * r[q] ** 2, Dd[q] * r[q], Dd[q]
        for i in np.nonzero(kind == 1)[0]:
            th = R[i] * fl[i]
            if abs(r[i]) < th:
                c[i], s[i], k[i] = 0.5 * r[i] ** 2 / R[i], r[i] / R[i], 1 / R[i]
            else:
                c[i], s[i], k[i] = fl[i] * abs(r[i]) - 0.5 * R[i] * fl[i] ** 2, fl[i] * np.sign(r[i]), 0.0
        return c, s, k

    def f(a):
        return 0.5 * (a - a0) @ Mm @ (a - a0) + parts(a)[0].sum()

    def g(a):
        return Mm @ (a - a0) + J.T @ parts(a)[1]

    def h(a):
        return Mm + J.T @ (parts(a)[2][:, None] * J)

    return f, g, h, a0, (kind, Dd, R)


def _check_newton(o, b, rtol=1e-9):
    f, g, h, a0, (kind, Dd, R) = _primal(o, b)
    fr = kind == 1
    np.testing.assert_allclose(Dd[fr] * R[fr], 1.0, rtol=1e-12)  # friction-loss rows: D = 1/R
    qa = o.get("qacc", [b])[0]
    scale = max(1.0, np.abs(qa).max())
    res = so.minimize(f, a0, jac=g, hess=h, method="trust-exact", options=dict(gtol=1e-11, maxiter=500))
    # the oracle's point must be stationary for the independent cost, and agree with scipy's minimiser
    Mscale = np.abs(h(qa)).max()
    assert np.linalg.norm(g(qa)) <= 1e-9 * Mscale * scale, np.linalg.norm(g(qa))
    np.testing.assert_allclose(qa, res.x, rtol=0, atol=rtol * scale * 1e3)
    assert f(qa) <= res.fun + 1e-12 * abs(res.fun) + 1e-12


def _box_trio():
    ents = [W.plane_entity(),
            W.box_entity((0.2, 0.2, 0.2), (0, 0, 0.09), rho=1000),
            W.box_entity((0.2, 0.3, 0.1), (0.02, 0.01, 0.23), rho=1000),
            W.box_entity((0.15, 0.15, 0.15), (0.5, 0, 0.07), quat=(0.98, 0.1, 0.15, 0.0), rho=1000)]
    return compile_model(ents, CompileOptions(dt=0.01))


def _box_trio_state(M, o, seed):
    q = np.tile(M.qpos0, (o.B, 1))
    rng = np.random.default_rng(seed)
    for b in range(o.B):
        for i in range(3):
            qq = q[b, 7 * i + 3:7 * i + 7] + rng.normal(0, 0.05, 4)
            q[b, 7 * i + 3:7 * i + 7] = qq / np.linalg.norm(qq)
    o.set("qpos", q)
    o.set("vel", rng.normal(0, 0.3, (o.B, M.n_dofs)))


def test_newton_minimises_convex_primal_with_contacts(Oracle):
    M = _box_trio()
    o = Oracle(M, 4, 64)
    _box_trio_state(M, o, 0)
    o.run(1, stage=4)
    for b in range(4):
        assert o.get("n_contacts", [b])[0] >= 4
        assert o.get("n_newton_iters", [b])[0] >= 2
        _check_newton(o, b)


def test_newton_minimises_convex_primal_with_limits_and_equality(Oracle):
    wl = W.arm_workload(n_envs=4)
    M = compile_model(wl.entities, wl.options)
    o = Oracle(M, 4, 64)
    rng = np.random.default_rng(1)
    lo, hi = M.dof_limit[:, 0], M.dof_limit[:, 1]
    q = np.tile(M.qpos0, (4, 1))
    q[:, :9] = rng.uniform(lo - 0.2, hi + 0.2, (4, 9))
    o.set("qpos", q)
    o.set("vel", rng.normal(0, 1, (4, 9)))
    o.run(1, stage=4)
    for b in range(4):
        ne = o.get("n_con_eq", [b])[0]
        assert ne == 1 and o.get("n_con", [b])[0] > ne  # the finger JOINT equality plus limit rows
        _check_newton(o, b)


def _chain(frictionloss=(0.3, 0.5, 1.0)):
    """Hinge (y) -> hinge (x) -> slide (z) chain with friction loss and limits, boxes on each link."""
    I3 = np.eye(3)
    base = LinkSpec("base", -1, np.array([0, 0, 1.0]), np.array([1.0, 0, 0, 0]), [JointSpec("fix", JOINT_FIXED)], [])
    l1 = LinkSpec("l1", 0, np.zeros(3), np.array([1.0, 0, 0, 0]),
                  [JointSpec("h1", JOINT_REVOLUTE, axis=np.array([0, 1.0, 0]), frictionloss=frictionloss[0],
                             limit=np.array([[-0.5, 0.5]]))],
                  [GeomSpec(GEOM_BOX, data=(0.3, 0.05, 0.05), pos=np.array([0.15, 0, 0]))],
                  inertial_pos=np.array([0.15, 0, 0.01]), inertial_quat=np.array([1.0, 0, 0, 0]),
                  inertial_i=np.diag([0.01, 0.02, 0.03]), inertial_mass=1.2)
    l2 = LinkSpec("l2", 1, np.array([0.3, 0, 0]), np.array([0.9950042, 0, 0.0998334, 0]),
                  [JointSpec("h2", JOINT_REVOLUTE, axis=np.array([1.0, 0, 0]), frictionloss=frictionloss[1])],
                  [GeomSpec(GEOM_BOX, data=(0.05, 0.3, 0.05), pos=np.array([0, 0.15, 0]))],
                  inertial_pos=np.array([0.02, 0.15, 0]), inertial_quat=np.array([0.9238795, 0, 0, 0.3826834]),
                  inertial_i=np.diag([0.03, 0.005, 0.02]), inertial_mass=0.8)
    l3 = LinkSpec("l3", 2, np.array([0, 0.3, 0]), np.array([1.0, 0, 0, 0]),
                  [JointSpec("s3", JOINT_PRISMATIC, axis=np.array([0, 0, 1.0]), frictionloss=frictionloss[2],
                             limit=np.array([[-0.1, 0.1]]))],
                  [GeomSpec(GEOM_BOX, data=(0.05, 0.05, 0.2))],
                  inertial_pos=np.array([0.0, 0.01, -0.05]), inertial_quat=np.array([1.0, 0, 0, 0]),
                  inertial_i=I3 * 0.004, inertial_mass=0.5)
    return EntitySpec([base, l1, l2, l3], name="chain")


def test_newton_minimises_convex_primal_with_frictionloss(Oracle):
    M = compile_model([_chain()], CompileOptions(dt=0.01, enable_collision=False))
    o = Oracle(M, 6, 64)
    rng = np.random.default_rng(3)
    o.set("qpos", rng.uniform([-0.7, -1.0, -0.2], [0.7, 1.0, 0.2], (6, 3)))
    o.set("vel", rng.normal(0, 2.0, (6, 3)) * np.array([[1], [1], [1], [1e-3], [1e-3], [1e-3]]))
    o.run(1, stage=4)
    saw_linear = saw_quad = False
    for b in range(6):
        assert o.get("n_con_fl", [b])[0] == 3
        f, g, h, a0, (kind, Dd, R) = _primal(o, b)
        qa = o.get("qacc", [b])[0]
        J = o.get("efc_jac", [b])[0]
        aref = o.get("efc_aref", [b])[0]
        fl = o.get("efc_fl", [b])[0]
        r = J @ qa - aref
        for i in np.nonzero(kind == 1)[0]:
            if abs(r[i]) >= R[i] * fl[i]:
                saw_linear = True
            else:
                saw_quad = True
        _check_newton(o, b)
    assert saw_linear and saw_quad  # both Huber zones are exercised


def test_contact_rows_are_friction_pyramid_of_point_velocity(Oracle):
    M = _box_trio()
    o = Oracle(M, 1, 64)
    _box_trio_state(M, o, 5)
    o.run(1, stage=4)
    q0 = o.get("qpos")[0].copy()
    nc = int(o.get("n_contacts")[0])
    cons = o.get("contacts")[0]
    J = o.get("efc_jac")[0]
    off = int(o.get("n_con_eq")[0] + o.get("n_con_fl")[0])
    o.run(1, stage=1)
    lp0, lq0 = o.get("links_pos")[0], o.get("links_quat")[0]
    rng = np.random.default_rng(7)
    h = 1e-6

    def shifted(v, s):
        q = q0.copy()
        for i in range(3):  # free joints: position + v h, orientation R exp([w] s h) (body-frame w)
            q[7 * i:7 * i + 3] += s * h * v[6 * i:6 * i + 3]
            R0 = quat_to_R(q0[7 * i + 3:7 * i + 7])
            q[7 * i + 3:7 * i + 7] = _R_to_quat(R0 @ sl.expm(_skew(v[6 * i + 3:6 * i + 6]) * s * h))
        o.set("qpos", q[None])
        o.run(1, stage=1)
        return o.get("links_pos")[0], o.get("links_quat")[0]

    samples = []
    for _ in range(8):
        v = rng.normal(0, 1, M.n_dofs)
        (pp, qp), (pm, qm) = shifted(v, 1), shifted(v, -1)
        vel = []
        for ic in range(nc):
            p = cons[ic, 4:7]
            vl = []
            for link in (int(cons[ic, 2]), int(cons[ic, 3])):
                R0 = quat_to_R(lq0[link])
                body = R0.T @ (p - lp0[link])
                xp = pp[link] + quat_to_R(qp[link]) @ body
                xm = pm[link] + quat_to_R(qm[link]) @ body
                vl.append((xp - xm) / (2 * h))
            vel.append(vl[1] - vl[0])  # velocity of b's point relative to a's
        samples.append((v, vel))
    o.set("qpos", q0[None])
    for ic in range(nc):
        n = cons[ic, 7:10]
        mu = cons[ic, 11]
        dirs = []
        for k in range(4):
            row = J[off + 4 * ic + k]
            A = np.array([s[1][ic] for s in samples])
            y = np.array([row @ s[0] for s in samples])
            d, res, *_ = np.linalg.lstsq(A, y, rcond=None)
            np.testing.assert_allclose(A @ d, y, atol=1e-6 * max(1, np.abs(y).max()))
            dirs.append(d)
        dirs = np.array(dirs)
        c = dirs.mean(0)
        np.testing.assert_allclose(np.linalg.norm(c), 1.0, atol=1e-6)
        np.testing.assert_allclose(abs(c @ n), 1.0, atol=1e-6)
        t = dirs - c
        np.testing.assert_allclose(np.linalg.norm(t, axis=1), mu, atol=1e-6)
        np.testing.assert_allclose(t[0], -t[1], atol=1e-6)
        np.testing.assert_allclose(t[2], -t[3], atol=1e-6)
        assert abs(t[0] @ t[2]) < 1e-6 and np.abs(t @ n).max() < 1e-6


def test_contact_rows_signal_approach_as_negative(Oracle):
    """A box falling flat onto the plane: every contact row sees J v < 0 (approach activates the row)."""
    M = compile_model([W.plane_entity(), W.box_entity((0.2, 0.2, 0.2), (0, 0, 0.099))], CompileOptions(dt=0.01))
    o = Oracle(M, 1, 64)
    o.set("vel", np.array([[0.0, 0, -1.0, 0, 0, 0]]))
    o.run(1, stage=4)
    nc = int(o.get("n_contacts")[0])
    assert nc == 4
    J = o.get("efc_jac")[0]
    jv = J @ np.array([0.0, 0, -1.0, 0, 0, 0])
    assert (jv < 0).all()


def test_bias_force_matches_lagrangian(Oracle):
    M = compile_model([_chain(frictionloss=(0, 0, 0))], CompileOptions(dt=0.01, enable_collision=False,
                                                                      enable_joint_limit=False))
    o = Oracle(M, 1, 64)
    rng = np.random.default_rng(11)
    g = np.array([0, 0, -9.81])
    eps = 1e-5

    def mass(q):
        o.set("qpos", q[None])
        o.set("vel", np.zeros((1, 3)))
        o.run(1, stage=2)
        return o.get("mass_mat")[0]

    def potential(q):
        o.set("qpos", q[None])
        o.run(1, stage=1)
        lp, lq = o.get("links_pos")[0], o.get("links_quat")[0]
        V = 0.0
        for link in range(M.n_links):
            com = lp[link] + quat_to_R(lq[link]) @ M.link_inertial_pos[link]
            V -= M.link_inertial_mass[link] * g @ com
        return V

    for _ in range(3):
        q = rng.uniform(-0.4, 0.4, 3)
        qd = rng.normal(0, 2, 3)
        dM = np.zeros((3, 3, 3))  # dM[:, :, k] = dM/dq_k
        dV = np.zeros(3)
        for k in range(3):
            e = np.zeros(3)
            e[k] = eps
            dM[:, :, k] = (mass(q + e) - mass(q - e)) / (2 * eps)
            dV[k] = (potential(q + e) - potential(q - e)) / (2 * eps)
        Gam = dM - 0.5 * np.transpose(dM, (2, 0, 1))  # Γ_ijk = ∂M_ij/∂q_k - ½ ∂M_jk/∂q_i
        bias = np.einsum("ijk,j,k->i", Gam, qd, qd) + dV
        o.set("qpos", q[None])
        o.set("vel", qd[None])
        o.run(1, stage=2)
        np.testing.assert_allclose(o.get("qf_smooth")[0], -bias, rtol=1e-6, atol=1e-6)


def test_torque_free_cube_turns_by_exponential_map(Oracle):
    ent = W.box_entity((0.3, 0.3, 0.3), (0.2, -0.1, 2.0), quat=(0.9, 0.3, -0.2, 0.1))
    M = compile_model([ent], CompileOptions(dt=0.01, gravity=(0.0, 0.0, 0.0), enable_collision=False))
    o = Oracle(M, 1, 64)
    w = np.array([1.3, -0.7, 2.1])
    v = np.array([0.4, 0.1, -0.3])
    o.set("vel", np.concatenate([v, w])[None])
    q0 = o.get("qpos")[0].copy()
    o.run(1)
    q1 = o.get("qpos")[0]
    R_expected = quat_to_R(q0[3:] / np.linalg.norm(q0[3:])) @ sl.expm(_skew(w) * 0.01)
    np.testing.assert_allclose(quat_to_R(q1[3:]), R_expected, atol=1e-12)
    np.testing.assert_allclose(q1[:3], q0[:3] + v * 0.01, atol=1e-14)
    np.testing.assert_allclose(o.get("vel")[0], np.concatenate([v, w]), atol=1e-12)


def test_mass_solve_inverts_mass_matrix(Oracle):
    """LDLᵀ factor/solve (rigid_solver_decomp.py:3681-4042): M acc_smooth = qf_smooth on the arm and a free-body pile."""
    wl = W.arm_workload(n_envs=3)
    for M, seed in ((compile_model(wl.entities, wl.options), 1), (_box_trio(), 2)):
        o = Oracle(M, 3, 64)
        rng = np.random.default_rng(seed)
        if M.n_qs == M.n_dofs:
            o.set("qpos", np.tile(M.qpos0, (3, 1)) + rng.normal(0, 0.3, (3, M.n_qs)))
        else:
            _box_trio_state(M, o, seed)
        o.set("vel", rng.normal(0, 1, (3, M.n_dofs)))
        o.run(1, stage=2)
        Mm, a, f = o.get("mass_mat"), o.get("acc_smooth"), o.get("qf_smooth")
        for b in range(3):
            np.testing.assert_allclose(Mm[b] @ a[b], f[b], rtol=1e-10, atol=1e-10 * np.abs(f[b]).max())
