"""State snapshots and checkpoints on the GPU: RigidSolver.get_state / set_state (rigid_solver_decomp.py:1442-1512,
7165-7258), Scene.reset with a state (scene.py:857-885) and Scene.save_checkpoint / load_checkpoint
(scene.py:1283-1344). A checkpoint holds every array that outlives a substep, so a run resumed from it is bitwise the
uninterrupted run."""
import numpy as np
import pytest
import torch

import genesis_amd as gs

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _init():
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    gs.init(backend=gs.gpu, seed=0)
    yield
    gs.destroy()


def _box_scene(n_envs, variant, arm=False):
    scene = gs.Scene(sim_options=gs.options.SimOptions(dt=0.01),
                     rigid_options=gs.options.RigidOptions(enable_self_collision=False))
    scene.add_entity(gs.morphs.Plane())
    rng = np.random.default_rng(3)
    for i in range(6):
        scene.add_entity(gs.morphs.Box(size=tuple(rng.uniform(0.08, 0.2, 3)),
                                       pos=(0.3 * (i % 3) - 0.3, 0.3 * (i // 3) - 0.15, 0.2 + 0.25 * i)))
    robot = scene.add_entity(gs.morphs.MJCF(file="xml/franka_emika_panda/panda.xml", pos=(0, 1.0, 0))) if arm else None
    scene.build(n_envs=n_envs, solver_variant=variant)
    return scene, robot


def _qv(scene):
    t = scene.rigid_solver.st.t
    return t["qpos"].clone(), t["vel"].clone()


@pytest.mark.parametrize("variant", [0, 1])
def test_checkpoint_resume_is_bitwise(tmp_path, variant):
    a, _ = _box_scene(8, variant)
    for _ in range(30):
        a.step()
    path = tmp_path / "ckpt.npz"
    a.save_checkpoint(path)
    for _ in range(25):
        a.step()
    qa, va = _qv(a)
    b, _ = _box_scene(8, variant)
    b.load_checkpoint(path)
    assert b.t == 30
    for _ in range(25):
        b.step()
    qb, vb = _qv(b)
    assert torch.equal(qa, qb) and torch.equal(va, vb)
    ca, cb = a.rigid_solver.collider.get_contacts(), b.rigid_solver.collider.get_contacts()
    for k in ca:
        assert torch.equal(ca[k], cb[k]), k


def test_checkpoint_rejects_another_scene(tmp_path):
    a, _ = _box_scene(4, None)
    path = tmp_path / "ckpt.npz"
    a.save_checkpoint(path)
    b, _ = _box_scene(6, None)
    with pytest.raises(ValueError):
        b.load_checkpoint(path)


def test_get_state_set_state_and_reset():
    scene, arm = _box_scene(4, None, arm=True)
    S = scene.rigid_solver
    for _ in range(20):
        scene.step()
    st = scene.get_state()
    assert tuple(st.qpos.shape) == (4, S.model.n_qs) and tuple(st.links_quat.shape) == (4, S.model.n_links, 4)
    assert st.s_global == 20
    arm.control_dofs_position(torch.zeros((4, 7)), dofs_idx_local=np.arange(7))
    for _ in range(10):
        scene.step()
    moved = S.st.t["qpos"].clone()
    # partial reset: envs 1 and 3 go back to the snapshot, envs 0 and 2 keep their state
    scene.reset(st, envs_idx=[1, 3])
    q = S.st.t["qpos"]
    assert torch.equal(q[:, [1, 3]], st.qpos[[1, 3]].T)
    assert torch.equal(q[:, [0, 2]], moved[:, [0, 2]])
    assert torch.equal(S.st.t["vel"][:, [1, 3]], st.dofs_vel[[1, 3]].T)
    # kernel_set_state puts the reset envs' controls back to zero force
    d = torch.arange(arm.dof_start, arm.dof_start + 7)
    assert bool((S.st.t["ctrl_mode"][d][:, [1, 3]] == 0).all()) and bool((S.st.t["ctrl_mode"][d][:, [0, 2]] == 2).all())
    assert bool((S.st.t["ctrl_force"][:, [1, 3]] == 0).all())
    # link poses are recomputed from qpos by the kinematics
    lp = S.get_links_pos()[[1, 3]]
    torch.testing.assert_close(lp, st.links_pos[[1, 3]], atol=1e-6, rtol=0)
    assert scene.t == 0
    # the given state became the initial state: a plain reset returns every env to it
    for _ in range(5):
        scene.step()
    scene.reset()
    assert torch.equal(S.st.t["qpos"], st.qpos.T)
    # per-env shifts travel with the state
    S.set_links_mass_shift([0.5], links_idx=[2], envs_idx=[1])
    S.set_geoms_friction_ratio([0.25], geoms_idx=[1], envs_idx=[3])
    st2 = S.get_state()
    assert float(st2.mass_shift[1, 2]) == 0.5 and float(st2.friction_ratio[3, 1]) == 0.25
    scene.reset(st)
    assert float(S.get_links_mass_shift([2], envs_idx=[1])[0, 0]) == 0.0
    S.set_state(st2)
    assert float(S.get_links_mass_shift([2], envs_idx=[1])[0, 0]) == 0.5
    assert float(S.get_geoms_friction_ratio([1], envs_idx=[3])[0, 0]) == 0.25
