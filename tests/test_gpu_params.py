"""Model parameters shared by every env (the reference's default batch_dofs_info = batch_links_info = False):
RigidSolver.set_dofs_* / set_links_inertial_mass / set_geoms_friction and their getters (rigid_solver_decomp.py
1777-2010, 2334-2452). After the changes the GPU step is held to the oracle built from the changed model, over
resynced steps, with the tolerance of the step parity tests."""
import numpy as np
import pytest
import torch

from genesis_amd import workloads as W
from genesis_amd.model import compile_model
from genesis_amd.rigid_solver import GenesisException, RigidSolver
from tests.sync import gpu_to_oracle

pytestmark = pytest.mark.gpu

TOL = 1e-4


def _need_gpu():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


def _arm(B, variant):
    w = W.arm_workload(n_envs=B)
    M = compile_model(w.entities, w.options)
    return M, RigidSolver(M, B, "cuda", solver_variant=variant)


def _change_parameters(S):
    S.set_dofs_kp(np.linspace(2000.0, 6000.0, 7), dofs_idx=np.arange(7))
    S.set_dofs_kv([150.0, 120.0], dofs_idx=[1, 3])
    S.set_dofs_force_range([-40.0] * 3, [40.0] * 3, dofs_idx=[0, 1, 2])
    S.set_dofs_damping([2.0, 2.0], dofs_idx=[4, 5])
    S.set_dofs_armature([0.3, 0.2, 0.1], dofs_idx=[0, 1, 2])
    S.set_links_inertial_mass([3.0], links_idx=[4])


@pytest.mark.parametrize("variant", [0, 1])
def test_shared_parameters_change_the_step_like_the_oracle(variant):
    _need_gpu()
    from oracle.oracle import Oracle
    B = 16
    M, S = _arm(B, variant)
    M0, S0 = _arm(B, variant)
    inv_before = S.get_dofs_invweight().cpu().numpy()
    _change_parameters(S)
    assert (S.get_dofs_invweight().cpu().numpy()[:3] < inv_before[:3]).all()  # more armature, less inverse weight
    rng = np.random.default_rng(5)
    lo = np.where(np.isfinite(M.dof_limit[:, 0]), M.dof_limit[:, 0], -1)
    hi = np.where(np.isfinite(M.dof_limit[:, 1]), M.dof_limit[:, 1], 1)
    tgt = torch.as_tensor(rng.uniform(lo, hi, (B, M.n_dofs)), dtype=torch.float32)
    S.control_dofs_position(tgt)
    S0.control_dofs_position(tgt)
    O = Oracle(S.model, B, 32)
    O64 = Oracle(S.model, B, 64) if variant == 1 else None
    for step in range(60):
        torch.cuda.synchronize()
        gpu_to_oracle(S, O)
        if O64 is not None:
            gpu_to_oracle(S, O64)
        S.substep()
        S0.substep()
        O.run(1)
        if O64 is not None:
            O64.run(1)
        torch.cuda.synchronize()
        for name, g_, o_ in (("qpos", S.st.t["qpos"].T.cpu().numpy(), O.get("qpos")),
                             ("vel", S.st.t["vel"].T.cpu().numpy(), O.get("vel"))):
            err = np.abs(g_ - o_)
            bound = TOL + TOL * np.abs(o_)
            if O64 is not None:
                o64 = O64.get(name)
                bound = np.maximum(bound, 4 * np.abs(o_ - o64).max(axis=1, keepdims=True) + TOL)
            assert (err <= bound).all(), f"step {step} {name}: max excess {(err - bound).max():.3g}"
    # the changed parameters took effect: the unchanged arm went elsewhere
    assert float((S.st.t["qpos"] - S0.st.t["qpos"]).abs().max()) > 1e-2


def test_parameter_getters_and_errors():
    _need_gpu()
    M, S = _arm(4, None)
    _change_parameters(S)
    np.testing.assert_allclose(S.get_dofs_kp(np.arange(7)).cpu().numpy(), np.linspace(2000.0, 6000.0, 7), rtol=1e-6)
    lo, hi = S.get_dofs_force_range([0, 1, 2])
    assert torch.equal(lo.cpu(), torch.full((3,), -40.0)) and torch.equal(hi.cpu(), torch.full((3,), 40.0))
    assert float(S.get_links_inertial_mass([4])[0]) == 3.0
    assert float(S.get_dofs_armature([1])[0]) == pytest.approx(0.2)
    assert float(S.dm.desc.meaninertia) == pytest.approx(float(S.model.meaninertia))
    with pytest.raises(GenesisException):
        S.set_dofs_kp([1.0], dofs_idx=[0], envs_idx=[1])
    with pytest.raises(IndexError):
        S.set_dofs_kv([1.0], dofs_idx=[M.n_dofs])
    with pytest.raises(IndexError):
        S.get_dofs_kp([-1])


def _landed_pile(B, variant, steps=60):
    w = W.pile_workload(n_boxes=8, n_envs=B, max_collision_pairs=200)
    M = compile_model(w.entities, w.options)
    S = RigidSolver(M, B, "cuda", solver_variant=variant)
    S.set_qpos(torch.as_tensor(w.init_qpos(B, M.n_qs), dtype=torch.float32))
    for _ in range(steps):
        S.substep()
    return M, S


@pytest.mark.parametrize("variant", [0, 1])
def test_per_env_shifts_match_oracle(variant):
    """Random per-env mass shifts, COM shifts and friction ratios on a landed 8-box pile: the GPU step is held to the
    oracle (which receives the shifts with the rest of the state) over resynced steps."""
    _need_gpu()
    from oracle.oracle import Oracle
    B = 16
    M, S = _landed_pile(B, variant)
    rng = np.random.default_rng(9)
    boxes = np.arange(1, M.n_links)
    S.set_links_mass_shift(torch.as_tensor(rng.uniform(0.0, 1.0, (B, len(boxes)))), links_idx=boxes)
    S.set_links_COM_shift(torch.as_tensor(rng.uniform(-0.01, 0.01, (B, len(boxes), 3))), links_idx=boxes)
    S.set_geoms_friction_ratio(torch.as_tensor(rng.uniform(0.2, 1.0, (B, M.n_geoms))))
    np.testing.assert_allclose(S.get_links_mass_shift(boxes).cpu().numpy().shape, (B, len(boxes)))
    assert tuple(S.get_links_COM_shift(boxes).shape) == (B, len(boxes), 3)
    O = Oracle(M, B, 32)
    O64 = Oracle(M, B, 64) if variant == 1 else None
    for step in range(40):
        torch.cuda.synchronize()
        gpu_to_oracle(S, O)
        if O64 is not None:
            gpu_to_oracle(S, O64)
        S.substep()
        O.run(1)
        if O64 is not None:
            O64.run(1)
        torch.cuda.synchronize()
        ncg, nco = S.st.t["n_contacts"][0].cpu().numpy(), O.get("n_contacts")
        assert (ncg == nco).all(), f"step {step}: contact counts differ"
        fr = S.st.t["contact_friction"].T.cpu().numpy()
        for b in range(B):
            oc = O.get("contacts", [b])[0]
            np.testing.assert_allclose(fr[b, :len(oc)], oc[:, 11], rtol=1e-6)
        for name, g_, o_ in (("qpos", S.st.t["qpos"].T.cpu().numpy(), O.get("qpos")),
                             ("vel", S.st.t["vel"].T.cpu().numpy(), O.get("vel"))):
            err = np.abs(g_ - o_)
            bound = TOL + TOL * np.abs(o_)
            if O64 is not None:
                o64 = O64.get(name)
                bound = np.maximum(bound, 4 * np.abs(o_ - o64).max(axis=1, keepdims=True) + TOL)
            assert (err <= bound).all(), f"step {step} {name}: max excess {(err - bound).max():.3g}"


def test_entity_mass_shift_changes_resting_contact_force():
    """A 4 cm cube at rest with per-env mass shifts 0 / 10 g / 50 g carries (m + dm) g (rigid_entity.py:2905-2919)."""
    _need_gpu()
    import genesis_amd as gs
    gs.init(backend=gs.gpu, seed=0)
    try:
        scene = gs.Scene(sim_options=gs.options.SimOptions(dt=0.01))
        scene.add_entity(gs.morphs.Plane())
        cube = scene.add_entity(gs.morphs.Box(size=(0.04, 0.04, 0.04), pos=(0.65, 0.0, 0.02)))
        scene.build(n_envs=3)
        dm = torch.tensor([[0.0], [0.01], [0.05]])
        cube.set_mass_shift(dm)
        cube.set_friction_ratio(torch.tensor([[1.0], [0.5], [0.25]]))
        for _ in range(60):
            scene.step()
        f = cube.get_links_net_contact_force().cpu().numpy()[:, 0]
        want = 9.81 * (cube.get_mass() + dm[:, 0].numpy())
        np.testing.assert_allclose(f[:, 2], want, atol=1e-5)
        np.testing.assert_allclose(f[:, :2], 0.0, atol=1e-5)
        fr = scene.rigid_solver.st.t["contact_friction"][:4].T.cpu().numpy()
        mu = scene.model.geom_friction
        np.testing.assert_allclose(fr[1], max(mu[0], 0.5 * mu[1]), rtol=1e-6)
    finally:
        gs.destroy()


@pytest.mark.parametrize("variant", [0, 1])
def test_batched_dofs_info_matches_per_env_oracles(variant):
    """batch_dofs_info=True (options/solvers.py:375): per-env kp, kv, force range, damping, stiffness and
    frictionloss. Env b is held over resynced steps to an oracle of the shared-parameter model carrying env b's
    values."""
    _need_gpu()
    import copy
    import dataclasses
    from oracle.oracle import Oracle
    B = 6
    w = W.arm_workload(n_envs=B)
    M = compile_model(w.entities, dataclasses.replace(w.options, batch_dofs_info=True))
    S = RigidSolver(M, B, "cuda", solver_variant=variant)
    rng = np.random.default_rng(11)
    arm7 = np.arange(7)
    S.set_dofs_kp(rng.uniform(1000.0, 6000.0, (B, 7)), dofs_idx=arm7)
    S.set_dofs_kv(rng.uniform(50.0, 300.0, (B, 7)), dofs_idx=arm7)
    f = rng.uniform(20.0, 80.0, (B, 3))
    S.set_dofs_force_range(-f, f, dofs_idx=[0, 1, 2])
    S.set_dofs_damping(rng.uniform(0.0, 3.0, (B, M.n_dofs)))
    S.set_dofs_stiffness(rng.uniform(0.0, 5.0, (B, M.n_dofs)))
    S.set_dofs_frictionloss(rng.uniform(0.0, 0.5, (B, M.n_dofs)))
    S.set_dofs_kp(np.full(7, 100.0), dofs_idx=arm7, envs_idx=[2])  # one env, values broadcast
    kp = S.get_dofs_kp()
    assert tuple(kp.shape) == (B, M.n_dofs)
    assert bool((kp[2, :7] == 100.0).all()) and not bool((kp[3, :7] == 100.0).all())
    lo, hi = S.get_dofs_force_range(dofs_idx=[0, 1, 2])
    np.testing.assert_allclose(hi.cpu().numpy(), f.astype(np.float32))
    oracles = []
    for b in range(B):
        Mb = copy.copy(M)
        for name, get in (("dof_kp", S.get_dofs_kp), ("dof_kv", S.get_dofs_kv), ("dof_damping", S.get_dofs_damping),
                          ("dof_stiffness", S.get_dofs_stiffness), ("dof_frictionloss", S.get_dofs_frictionloss)):
            setattr(Mb, name, get()[b].cpu().numpy().astype(np.float32))
        for name, get in (("dof_force_range", S.get_dofs_force_range), ("dof_limit", S.get_dofs_limit)):
            a, c = get()
            setattr(Mb, name, np.stack([a[b].cpu().numpy(), c[b].cpu().numpy()], -1).astype(np.float32))
        oracles.append((Oracle(Mb, 1, 32), Oracle(Mb, 1, 64)))
    lo_, hi_ = (np.where(np.isfinite(M.dof_limit[:, k]), M.dof_limit[:, k], (-1, 1)[k]) for k in (0, 1))
    S.control_dofs_position(torch.as_tensor(rng.uniform(lo_, hi_, (B, M.n_dofs)), dtype=torch.float32))
    for step in range(40):
        torch.cuda.synchronize()
        for b, (O, O64) in enumerate(oracles):
            gpu_to_oracle(S, O, envs=slice(b, b + 1))
            gpu_to_oracle(S, O64, envs=slice(b, b + 1))
        S.substep()
        for O, O64 in oracles:
            O.run(1)
            O64.run(1)
        torch.cuda.synchronize()
        qg, vg = S.st.t["qpos"].T.cpu().numpy(), S.st.t["vel"].T.cpu().numpy()
        for b, (O, O64) in enumerate(oracles):
            for name, g_ in (("qpos", qg[b]), ("vel", vg[b])):
                o_, o64 = O.get(name)[0], O64.get(name)[0]
                err = np.abs(g_ - o_)
                bound = np.maximum(TOL + TOL * np.abs(o_), 4 * np.abs(o_ - o64).max() + TOL)
                assert (err <= bound).all(), f"step {step} env {b} {name}: max excess {(err - bound).max():.3g}"


def test_hessian_lds_budget_changes_no_bit():
    """The wavefront solver keeps an env's Hessian in LDS when it fits the model's budget and in global memory
    otherwise; both stores give the same bits, so the budget only moves envs between them."""
    _need_gpu()
    w = W.pile_workload(n_boxes=8, n_envs=32, max_collision_pairs=200)
    M = compile_model(w.entities, w.options)
    q0 = torch.as_tensor(w.init_qpos(32, M.n_qs), dtype=torch.float32)
    out = []
    for H in (5120, 256, 0):
        S = RigidSolver(M, 32, "cuda", solver_variant=1, wave_h_floats=H)
        S.set_qpos(q0)
        for _ in range(80):
            S.substep()
        torch.cuda.synchronize()
        out.append((S.st.t["qpos"].clone(), S.st.t["vel"].clone()))
    for q, v in out[1:]:
        assert torch.equal(q, out[0][0]) and torch.equal(v, out[0][1])


@pytest.mark.parametrize("name,B,budgets", [
    # arm: 9 dofs, triangle 45 (48 with padding): 1024 stages 64 rows a pass, 128 stages 8, 118 leaves too little
    # room (the per-entity row walk)
    ("arm", 128, (1024, 128, 118)),
    # quadruped: 18 dofs, triangle 171 (172): 1024 stages 44 rows, 324 stages 8, 305 the row walk
    ("quadruped", 64, (1024, 324, 305)),
])
def test_dense_row_staging_changes_no_bit(name, B, budgets):
    """Models whose dofs all belong to one entity sum J^T f and J^T D J with lanes over dofs / Hessian entries over
    rows staged in LDS, a budget-dependent number of rows per pass; the per-entity row walk gives the same bits, over
    200 controlled steps with many contact rows. (A Hessian kept in global memory takes the chunked factorisation
    instead of the dense one, so its rounding differs; the parity tests cover it.)"""
    _need_gpu()
    w = W.workload_by_name(name, B)
    M = compile_model(w.entities, w.options)
    out, max_rows = [], 0
    for H in budgets:
        S = RigidSolver(M, B, "cuda", solver_variant=1, wave_h_floats=H)
        S.set_variants(1, 1, 1)
        W.init_state(w, M, S)
        ctrl = W.Control(w, M, S, seed=5)
        for step in range(200):
            ctrl(step)
            S.substep()
            if H == budgets[0] and step % 10 == 0:
                max_rows = max(max_rows, int(S.st.t["n_con"][0].max().item()))
        torch.cuda.synchronize()
        out.append((S.st.t["qpos"].clone(), S.st.t["vel"].clone()))
    for H, (q, v) in zip(budgets[1:], out[1:]):
        assert torch.equal(q, out[0][0]) and torch.equal(v, out[0][1]), f"budget {H}"
    # the budgets that stage 8 rows a pass make several passes
    assert max_rows > 8, max_rows


@pytest.mark.parametrize("variant", [0, 1])
def test_batched_links_joints_and_armature_match_per_env_oracles(variant):
    """batch_dofs_info = batch_links_info = batch_joints_info = True (options/solvers.py:373-375): per-env armature
    (whose inverse weights and mean inertia are recomputed per env, rigid_solver_decomp.py:354-515, 1964-1969), link
    masses, joint solver parameters and frictionloss. Joint targets past the limits keep limit rows active, and the
    frictionloss rows read the joint parameters every step. Env b is held over resynced steps to an oracle of the
    shared-parameter model carrying env b's tables and mean inertia; the env left at its build parameters keeps the
    build inverse weights bit for bit."""
    _need_gpu()
    import copy
    import dataclasses
    from genesis_amd.model import _init_invweight_and_meaninertia
    from oracle.oracle import Oracle
    B = 6
    w = W.arm_workload(n_envs=B)
    opts = dataclasses.replace(w.options, batch_dofs_info=True, batch_links_info=True, batch_joints_info=True)
    M = compile_model(w.entities, opts)
    S = RigidSolver(M, B, "cuda", solver_variant=variant)
    rng = np.random.default_rng(21)
    changed = [0, 1, 3, 4, 5]  # env 2 keeps its build parameters; env 5 is changed only through envs_idx
    arm7 = np.arange(7)
    # link masses first: the armature setter recomputes the inverse weights from each env's masses
    S.set_links_inertial_mass(rng.uniform(0.5, 4.0, (5, 2)), links_idx=[3, 5], envs_idx=changed)
    S.set_dofs_armature(rng.uniform(0.05, 0.6, (4, 7)), dofs_idx=arm7, envs_idx=changed[:4])
    S.set_dofs_armature(np.full(7, 0.3), dofs_idx=arm7, envs_idx=[5])
    S.set_dofs_frictionloss(rng.uniform(0.05, 0.5, (B, M.n_dofs)))
    sp = np.tile([0.0, 1.0, 0.9, 0.95, 0.001, 0.5, 2.0], (B, M.n_joints, 1))
    sp[:, :, 0] = rng.uniform(0.02, 0.08, (B, M.n_joints))
    sp[:, :, 1] = rng.uniform(0.5, 1.5, (B, M.n_joints))
    S.set_sol_params(sp, joints_idx=np.arange(M.n_joints))
    # getters: per-env shapes, the build values in the untouched env
    arm_ = S.get_dofs_armature()
    inv = S.get_dofs_invweight().cpu().numpy()
    linv = S.get_links_invweight().cpu().numpy()
    mi = S.get_meaninertia().cpu().numpy()
    assert tuple(arm_.shape) == (B, M.n_dofs) and tuple(linv.shape) == (B, M.n_links, 2)
    np.testing.assert_array_equal(inv[2], M.dof_invweight)
    np.testing.assert_array_equal(linv[2], M.link_invweight)
    assert mi[2] == np.float32(M.meaninertia) and (mi[changed] != mi[2]).all()
    np.testing.assert_allclose(S.get_sol_params(joints_idx=[2, 3]).cpu().numpy(), sp[:, 2:4], rtol=1e-6)
    with pytest.raises(GenesisException):
        S.set_sol_params(sp[0, :2], geoms_idx=[0, 1], envs_idx=[0])  # geom parameters are shared
    oracles = []
    for b in range(B):
        Mb = copy.copy(M)
        Mb.dof_armature = arm_[b].cpu().numpy().astype(np.float32)
        Mb.link_inertial_mass = S.get_links_inertial_mass()[b].cpu().numpy().astype(np.float32)
        Mb.dof_frictionloss = S.get_dofs_frictionloss()[b].cpu().numpy().astype(np.float32)
        Mb.joint_sol_params = S.get_sol_params(joints_idx=np.arange(M.n_joints))[b].cpu().numpy().astype(np.float32)
        ref = copy.copy(Mb)
        _init_invweight_and_meaninertia(ref)  # the host restatement on env b's armature and masses
        np.testing.assert_array_equal(inv[b], ref.dof_invweight)
        np.testing.assert_array_equal(linv[b], ref.link_invweight)
        assert mi[b] == np.float32(ref.meaninertia)
        Mb.dof_invweight, Mb.link_invweight, Mb.meaninertia = inv[b], linv[b], mi[b]
        oracles.append((Oracle(Mb, 1, 32), Oracle(Mb, 1, 64)))
    lo_, hi_ = (np.where(np.isfinite(M.dof_limit[:, k]), M.dof_limit[:, k], (-1, 1)[k]) for k in (0, 1))
    tgt = rng.uniform(lo_, hi_, (B, M.n_dofs))
    tgt[:, [1, 3, 5]] = hi_[[1, 3, 5]] + 0.4  # past the upper limits: limit rows
    S.control_dofs_position(torch.as_tensor(tgt, dtype=torch.float32))
    n_limit_steps = 0
    for step in range(50):
        torch.cuda.synchronize()
        for b, (O, O64) in enumerate(oracles):
            gpu_to_oracle(S, O, envs=slice(b, b + 1))
            gpu_to_oracle(S, O64, envs=slice(b, b + 1))
        S.substep()
        for O, O64 in oracles:
            O.run(1)
            O64.run(1)
        torch.cuda.synchronize()
        n_con = S.st.t["n_con"].cpu().numpy()
        qg, vg = S.st.t["qpos"].T.cpu().numpy(), S.st.t["vel"].T.cpu().numpy()
        for b, (O, O64) in enumerate(oracles):
            assert n_con[0, b] == O.get("n_con")[0], f"step {step} env {b}: row counts differ"
            for name, g_ in (("qpos", qg[b]), ("vel", vg[b])):
                o_, o64 = O.get(name)[0], O64.get(name)[0]
                err = np.abs(g_ - o_)
                bound = np.maximum(TOL + TOL * np.abs(o_), 4 * np.abs(o_ - o64).max() + TOL)
                assert (err <= bound).all(), f"step {step} env {b} {name}: max excess {(err - bound).max():.3g}"
        n_limit_steps += int((n_con[0] > n_con[2] + 4 * S.st.t["n_contacts"][0].cpu().numpy()).any())
    assert n_limit_steps > 10  # limit rows were active for most of the run


def test_selective_armature_needs_batched_links_and_dofs():
    """set_dofs_armature with envs_idx recomputes inverse weights per env, which needs both tables batched
    (rigid_solver_decomp.py:360-364)."""
    _need_gpu()
    import dataclasses
    w = W.arm_workload(n_envs=3)
    M = compile_model(w.entities, dataclasses.replace(w.options, batch_dofs_info=True))
    S = RigidSolver(M, 3, "cuda")
    with pytest.raises(GenesisException):
        S.set_dofs_armature([0.2], dofs_idx=[0], envs_idx=[1])
    # without envs_idx every env is recomputed; shared link inverse weights come from the first env
    S.set_dofs_armature(np.array([[0.1], [0.2], [0.3]]), dofs_idx=[0])
    mi = S.get_meaninertia().cpu().numpy()
    assert mi[0] < mi[1] < mi[2]
    inv = S.get_dofs_invweight().cpu().numpy()
    assert inv[0, 0] > inv[1, 0] > inv[2, 0]
