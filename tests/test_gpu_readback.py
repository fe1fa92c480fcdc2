"""Contact and constraint readback on the GPU: Collider.get_contacts (collider_decomp.py:463-565), the per-link net
contact force (constraint_solver_decomp.py:1350-1379, rigid_entity.py:2838-2850) and the equality-constraint readback
(constraint_solver_decomp.py:235-322, 2510-2558).

The gathers are checked bit-exact against the solver's own state; the contact list and its forces against the
oracle over one resynced step; the per-link sums bit-exact against the oracle's float32 restatement applied to the
GPU's contact list; and the reference's resting-cube test (test_rigid_physics.py:1626-1675) through the Scene API."""
import numpy as np
import pytest
import torch

import genesis_amd as gs
from genesis_amd import workloads as W
from genesis_amd.model import compile_model
from tests.sync import gpu_to_oracle

pytestmark = pytest.mark.gpu

TOL = 1e-4
KEYS = ("link_a", "link_b", "geom_a", "geom_b", "penetration", "position", "normal", "force")


def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


def _landed_pile(B, variant, steps=80, n_boxes=8):
    from genesis_amd.rigid_solver import RigidSolver
    w = W.pile_workload(n_boxes=n_boxes, n_envs=B, max_collision_pairs=200)
    M = compile_model(w.entities, w.options)
    S = RigidSolver(M, B, "cuda", solver_variant=variant)
    S.set_qpos(torch.as_tensor(w.init_qpos(B, M.n_qs), dtype=torch.float32))
    for _ in range(steps):
        S.substep()
    torch.cuda.synchronize()
    return M, S


def _state_contacts(S):
    t, B = S.st.t, S._B
    return dict(n=t["n_contacts"][0].cpu().numpy(),
                geoms=t["contact_geoms"].T.cpu().numpy().reshape(B, -1, 2),
                pos=t["contact_pos"].T.cpu().numpy().reshape(B, -1, 3),
                normal=t["contact_normal"].T.cpu().numpy().reshape(B, -1, 3),
                pen=t["contact_pen"].T.cpu().numpy(),
                force=t["contact_force"].T.cpu().numpy().reshape(B, -1, 3))


@pytest.mark.parametrize("variant", [0, 1])
def test_get_contacts_padded_and_ragged_gather_the_state(variant):
    _need_gpu()
    B = 16
    M, S = _landed_pile(B, variant)
    st = _state_contacts(S)
    n = st["n"]
    assert n.sum() > 0 and (n > 0).any()
    info = S.collider.get_contacts()
    assert set(info) == set(KEYS)
    n_max = int(n.max())
    assert tuple(info["geom_a"].shape) == (B, n_max) and info["geom_a"].dtype == torch.int32
    assert tuple(info["position"].shape) == (B, n_max, 3) and info["force"].dtype == torch.float32
    c = {k: v.cpu().numpy() for k, v in info.items()}
    for b in range(B):
        k = int(n[b])
        np.testing.assert_array_equal(c["geom_a"][b, :k], st["geoms"][b, :k, 0])
        np.testing.assert_array_equal(c["geom_b"][b, :k], st["geoms"][b, :k, 1])
        np.testing.assert_array_equal(c["link_a"][b, :k], M.geom_link[st["geoms"][b, :k, 0]])
        np.testing.assert_array_equal(c["link_b"][b, :k], M.geom_link[st["geoms"][b, :k, 1]])
        np.testing.assert_array_equal(c["penetration"][b, :k], st["pen"][b, :k])
        for key, f in (("position", "pos"), ("normal", "normal"), ("force", "force")):
            np.testing.assert_array_equal(c[key][b, :k], st[f][b, :k])
        for key in ("link_a", "link_b", "geom_a", "geom_b"):
            assert (c[key][b, k:] == -1).all()
        for key in ("penetration", "position", "normal", "force"):
            assert (c[key][b, k:] == 0).all()
    ragged = S.collider.get_contacts(as_tensor=False)
    for key in KEYS:
        assert len(ragged[key]) == B
        for b in range(B):
            assert torch.equal(ragged[key][b], info[key][b, :int(n[b])])
    as_np = S.collider.get_contacts(to_torch=False)
    for key in KEYS:
        assert isinstance(as_np[key], np.ndarray)
        np.testing.assert_array_equal(as_np[key], c[key])


@pytest.mark.parametrize("variant", [0, 1])
def test_links_net_contact_force_sums_in_reference_order(variant):
    _need_gpu()
    from oracle.oracle import links_net_contact_force
    B = 16
    M, S = _landed_pile(B, variant)
    info = {k: v.cpu().numpy() for k, v in S.collider.get_contacts().items()}
    n = S.st.t["n_contacts"][0].cpu().numpy()
    net = S.get_links_net_contact_force()
    assert tuple(net.shape) == (B, M.n_links, 3)
    net = net.cpu().numpy()
    for b in range(B):
        k = int(n[b])
        want = links_net_contact_force(info["link_a"][b, :k], info["link_b"][b, :k], info["force"][b, :k], M.n_links)
        np.testing.assert_array_equal(net[b], want)
    sub = S.get_links_net_contact_force(envs_idx=[5, 2, 11]).cpu().numpy()
    np.testing.assert_array_equal(sub, net[[5, 2, 11]])


@pytest.mark.parametrize("variant", [0, 1])
def test_contacts_and_forces_match_oracle_over_one_resynced_step(variant):
    """GPU and oracle take one step from the same state; their contact lists (pairs exact, geometry within TOL) and
    contact forces agree. Forces are held to 1e-3 of the env's largest force, widened in an env whose FP32 solve is
    ill-conditioned on that step to 4x the FP32-vs-FP64 oracle spread of that env's forces."""
    _need_gpu()
    from oracle.oracle import Oracle
    B = 16
    M, S = _landed_pile(B, variant)
    O, O64 = Oracle(M, B, 32), Oracle(M, B, 64)
    gpu_to_oracle(S, O)
    gpu_to_oracle(S, O64)
    S.substep()
    O.run(1)
    O64.run(1)
    torch.cuda.synchronize()
    info = {k: v.cpu().numpy() for k, v in S.collider.get_contacts().items()}
    oc, oc64 = O.get("contacts"), O64.get("contacts")
    n = S.st.t["n_contacts"][0].cpu().numpy()
    worst = 0.0
    for b in range(B):
        k = int(n[b])
        o = oc[b]
        assert k == len(o), f"env {b}: {k} contacts vs oracle {len(o)}"
        np.testing.assert_array_equal(info["geom_a"][b, :k], o[:, 0].astype(np.int32))
        np.testing.assert_array_equal(info["geom_b"][b, :k], o[:, 1].astype(np.int32))
        np.testing.assert_array_equal(info["link_a"][b, :k], o[:, 2].astype(np.int32))
        np.testing.assert_array_equal(info["link_b"][b, :k], o[:, 3].astype(np.int32))
        for key, cols in (("position", slice(4, 7)), ("normal", slice(7, 10)), ("penetration", 10)):
            np.testing.assert_allclose(info[key][b, :k], o[:, cols], atol=TOL, rtol=TOL, err_msg=f"env {b} {key}")
        if k == 0:
            continue
        f, fo, f64 = info["force"][b, :k], o[:, 12:15], oc64[b][:, 12:15]
        bound = max(1e-3 * np.abs(fo).max(), 4 * np.abs(fo - f64).max()) + TOL
        err = np.abs(f - fo).max()
        worst = max(worst, err / bound)
        assert err <= bound, f"env {b}: force error {err:.3g} above {bound:.3g}"
    print(f"v{variant}: {int(n.sum())} contacts, worst force error / bound {worst:.3g}")


def test_equality_constraint_readback():
    """The arm's finger equality (a JOINT equality): type, joints and its row force, padded per env."""
    _need_gpu()
    from genesis_amd.rigid_solver import RigidSolver
    w = W.arm_workload(n_envs=4)
    M = compile_model(w.entities, w.options)
    S = RigidSolver(M, 4, "cuda")
    S.control_dofs_position(torch.full((4, 2), 0.01), dofs_idx=[7, 8])
    for _ in range(20):
        S.substep()
    torch.cuda.synchronize()
    eq = S.constraint_solver.get_equality_constraints()
    assert set(eq) == {"type", "obj_a", "obj_b", "force"}
    assert tuple(eq["type"].shape) == (4, 1) and tuple(eq["force"].shape) == (4, 1, 6)
    assert (eq["type"] == 2).all()
    assert (eq["obj_a"] == int(M.eq_obj1[0])).all() and (eq["obj_b"] == int(M.eq_obj2[0])).all()
    t = S.st.t["efc_force"]
    row0 = (t.reshape(4, -1)[:, 0] if int(S.dm.desc.solver_variant) else t.reshape(-1, 4)[0]).cpu()
    assert torch.equal(eq["force"][:, 0, 0].cpu(), row0)
    assert (eq["force"][:, 0, 1:] == 0).all()
    ragged = S.constraint_solver.get_equality_constraints(as_tensor=False)
    assert len(ragged["force"]) == 4 and tuple(ragged["force"][0].shape) == (1, 6)
    welds = S.constraint_solver.get_weld_constraints()
    assert set(welds) == {"link_a", "link_b", "force"}
    assert tuple(welds["link_a"].shape) == (4, 0) and tuple(welds["force"].shape) == (4, 0, 6)


@pytest.mark.parametrize("n_envs", [0, 3])
def test_resting_cube_net_contact_force_equals_weight(n_envs):
    """test_rigid_physics.py:1626-1675, hold phase: a 4 cm cube resting on the plane for 50 steps carries its weight
    upward within 1e-5 N; entity.get_contacts reports its 4 contacts with force_b pushing it up."""
    _need_gpu()
    gs.init(backend=gs.gpu, seed=0)
    try:
        scene = gs.Scene(sim_options=gs.options.SimOptions(dt=0.01))
        scene.add_entity(gs.morphs.Plane())
        cube = scene.add_entity(gs.morphs.Box(size=(0.04, 0.04, 0.04), pos=(0.65, 0.0, 0.02)))
        scene.build(n_envs=n_envs)
        for _ in range(50):
            scene.step()
        weight = 9.81 * cube.get_mass()
        f = cube.get_links_net_contact_force().cpu().numpy()
        if n_envs == 0:
            assert f.shape == (1, 3)
            np.testing.assert_allclose(f[0], [0.0, 0.0, weight], atol=1e-5)
        else:
            assert f.shape == (n_envs, 1, 3)
            np.testing.assert_allclose(f[:, 0], np.tile([0.0, 0.0, weight], (n_envs, 1)), atol=1e-5)
        c = cube.get_contacts()
        if n_envs == 0:
            assert c["geom_a"].shape[0] == 4
            assert (c["force_b"][:, 2] > 0).all()
        else:
            assert bool(c["valid_mask"].all()) and c["valid_mask"].shape[-1] == 4
            assert (c["force_b"][..., 2] > 0).all()
        np.testing.assert_allclose((c["force_b"][..., 2].sum(-1)).cpu().numpy(), weight, atol=1e-5)
    finally:
        gs.destroy()
