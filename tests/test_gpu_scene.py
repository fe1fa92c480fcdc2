"""The gs.Scene / RigidEntity surface on the GPU, written like the reference's own rigid tests."""
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


def test_free_fall_through_scene():
    """test_rigid_physics.py:3273-3307 through the Scene API: v = g dt after one step."""
    scene = gs.Scene(sim_options=gs.options.SimOptions(dt=0.01))
    box = scene.add_entity(gs.morphs.Box(size=(0.1, 0.1, 0.1), pos=(0.0, 0.0, 1.0)))
    scene.build(n_envs=4)
    scene.step()
    v = box.get_dofs_velocity().cpu().numpy()
    np.testing.assert_allclose(v[:, 2], -9.81 * 0.01, rtol=1e-5)
    np.testing.assert_allclose(v[:, [0, 1, 3, 4, 5]], 0.0, atol=1e-7)
    z = box.get_pos().cpu().numpy()[:, 2]
    np.testing.assert_allclose(z, 1.0 - 9.81 * 0.01 * 0.01, atol=1e-6)


def test_unbatched_shapes_drop_env_axis():
    scene = gs.Scene()
    scene.add_entity(gs.morphs.Plane())
    box = scene.add_entity(gs.morphs.Box(size=(0.2, 0.2, 0.2), pos=(0.0, 0.0, 0.5)))
    scene.build()  # n_envs = 0
    scene.step()
    assert tuple(box.get_dofs_velocity().shape) == (6,)
    assert tuple(box.get_dofs_position().shape) == (6,)


def test_arm_pd_tracking_and_local_indices():
    """Panda arm: PD targets through control_dofs_position reach the target (steady state) per env."""
    scene = gs.Scene(sim_options=gs.options.SimOptions(dt=0.01),
                     rigid_options=gs.options.RigidOptions(enable_self_collision=False))
    scene.add_entity(gs.morphs.Plane())
    arm = scene.add_entity(gs.morphs.MJCF(file="xml/franka_emika_panda/panda.xml"))
    scene.build(n_envs=8)
    assert arm.n_dofs == 9
    tgt = torch.tensor([0.0, -0.3, 0.0, -2.0, 0.0, 1.8, 0.7], dtype=torch.float32).repeat(8, 1)
    tgt[:, 0] = torch.linspace(-0.5, 0.5, 8)
    arm.control_dofs_position(tgt, dofs_idx_local=np.arange(7))
    arm.control_dofs_position(torch.full((8, 2), 0.02), dofs_idx_local=[7, 8])
    for _ in range(300):
        scene.step()
    q = arm.get_dofs_position(dofs_idx_local=np.arange(7)).cpu().numpy()
    np.testing.assert_allclose(q, tgt.numpy(), atol=5e-2)
    # env subset readback
    q2 = arm.get_dofs_position(dofs_idx_local=[0], envs_idx=[2, 5]).cpu().numpy()
    np.testing.assert_allclose(q2[:, 0], q[[2, 5], 0])


def test_reset_restores_build_state_and_caches():
    scene = gs.Scene()
    scene.add_entity(gs.morphs.Plane())
    box = scene.add_entity(gs.morphs.Box(size=(0.2, 0.2, 0.2), pos=(0.0, 0.0, 0.3)))
    scene.build(n_envs=3)
    q0 = box.get_qpos().clone()
    for _ in range(50):
        scene.step()
    assert not torch.allclose(box.get_qpos(), q0)
    scene.reset(envs_idx=[1])
    q = box.get_qpos()
    torch.testing.assert_close(q[1], q0[1])
    assert not torch.allclose(q[0], q0[0])
    # after reset the env steps exactly like a fresh one
    ref = gs.Scene()
    ref.add_entity(gs.morphs.Plane())
    rbox = ref.add_entity(gs.morphs.Box(size=(0.2, 0.2, 0.2), pos=(0.0, 0.0, 0.3)))
    ref.build(n_envs=3)
    for _ in range(20):
        scene.step()
        ref.step()
    torch.testing.assert_close(box.get_qpos()[1], rbox.get_qpos()[1])


def test_go2_urdf_standing():
    scene = gs.Scene(sim_options=gs.options.SimOptions(dt=0.005))
    scene.add_entity(gs.morphs.Plane())
    robot = scene.add_entity(gs.morphs.URDF(file="urdf/go2/urdf/go2.urdf", pos=(0.0, 0.0, 0.42)))
    scene.build(n_envs=16)
    stand = torch.tensor([0.0, 0.0, 0.0, 0.0, 0.8, 0.8, 1.0, 1.0, -1.5, -1.5, -1.5, -1.5]).repeat(16, 1)
    robot.set_qpos(torch.cat([torch.tensor([0.0, 0.0, 0.42, 1, 0, 0, 0]).repeat(16, 1), stand], 1))
    robot.control_dofs_position(stand, dofs_idx_local=np.arange(6, 18))
    for _ in range(200):
        scene.step()
    z = robot.get_pos().cpu().numpy()[:, 2]
    assert (z > 0.2).all() and (z < 0.45).all()


def test_mass_mat_of_identical_robots_and_its_factor():
    """test_rigid_physics.py:1788-1820: two identical arms have the same mass matrix, and L^T diag(1 / D_inv) L
    rebuilds it; the matrix also matches the oracle's at the same state."""
    from oracle.oracle import Oracle
    from tests.sync import gpu_to_oracle
    scene = gs.Scene(sim_options=gs.options.SimOptions(dt=0.01, substeps=1))
    scene.add_entity(gs.morphs.Plane())
    a1 = scene.add_entity(gs.morphs.MJCF(file="xml/franka_emika_panda/panda.xml", pos=(0, 0, 0)))
    a2 = scene.add_entity(gs.morphs.MJCF(file="xml/franka_emika_panda/panda.xml", pos=(0, 2, 0)))
    scene.build()
    m1, m2 = a1.get_mass_mat(), a2.get_mass_mat()
    assert tuple(m1.shape) == (a1.n_dofs, a1.n_dofs)
    torch.testing.assert_close(m1, m2, atol=1e-6, rtol=1e-6)
    L, dinv = a1.get_mass_mat(decompose=True)
    torch.testing.assert_close(L.T @ torch.diag(1.0 / dinv) @ L, m1, atol=1e-5, rtol=1e-5)
    S = scene.rigid_solver
    O = Oracle(S.model, 1, 64)
    gpu_to_oracle(S, O)
    O.run(stage=2)
    np.testing.assert_allclose(S.get_mass_mat().cpu().numpy(), O.get("mass_mat")[0], atol=1e-5, rtol=1e-5)
    # reading it between steps changes no result
    other = gs.Scene(sim_options=gs.options.SimOptions(dt=0.01, substeps=1))
    other.add_entity(gs.morphs.Plane())
    other.add_entity(gs.morphs.MJCF(file="xml/franka_emika_panda/panda.xml", pos=(0, 0, 0)))
    other.add_entity(gs.morphs.MJCF(file="xml/franka_emika_panda/panda.xml", pos=(0, 2, 0)))
    other.build()
    for _ in range(10):
        S.get_mass_mat()
        scene.step()
        other.step()
    assert torch.equal(S.st.t["qpos"], other.rigid_solver.st.t["qpos"])


def test_mass_mat_after_a_step_is_the_steps_own():
    """rigid_solver_decomp.py:2436-2449 returns the mass_mat field the last substep's dynamics filled: after a step,
    the matrix at the pose before that step's integration, not at the pose the step produced."""
    from oracle.oracle import Oracle
    from tests.sync import gpu_to_oracle
    scene = gs.Scene(sim_options=gs.options.SimOptions(dt=0.01, substeps=1))
    scene.add_entity(gs.morphs.Plane())
    a1 = scene.add_entity(gs.morphs.MJCF(file="xml/franka_emika_panda/panda.xml", pos=(0, 0, 0)))
    scene.build()
    S = scene.rigid_solver
    O = Oracle(S.model, 1, 64)
    a1.control_dofs_position(torch.full((a1.n_dofs,), 0.4))
    for _ in range(3):
        scene.step()
    torch.cuda.synchronize()
    gpu_to_oracle(S, O)
    O.run(stage=2)
    scene.step()
    m_step = S.get_mass_mat().cpu().numpy()
    np.testing.assert_allclose(m_step, O.get("mass_mat")[0], atol=1e-5, rtol=1e-5)
    gpu_to_oracle(S, O)  # M at the post-step pose is another matrix
    O.run(stage=2)
    assert np.abs(O.get("mass_mat")[0] - m_step).max() > 1e-6


def _xyz_to_quat(xyz):
    """Quaternion (w, x, y, z) of extrinsic x, then y, then z rotations."""
    out = []
    for a in ((xyz[0], (1, 0, 0)), (xyz[1], (0, 1, 0)), (xyz[2], (0, 0, 1))):
        out.append(np.array([np.cos(a[0] / 2), *(np.sin(a[0] / 2) * np.array(a[1]))]))

    def mul(u, v):
        w1, x1, y1, z1 = u
        w2, x2, y2, z2 = v
        return np.array([w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2, w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
                         w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2, w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2])
    return mul(out[2], mul(out[1], out[0]))


def test_box_on_box_settles():
    """test_rigid_physics.py:852-874: a 0.4 m cube dropped tilted onto another comes to rest with its centre at
    z = 0.6 (within 2 mm), velocities below 1e-2 after 100 steps; 8 envs with different drops."""
    B = 8
    scene = gs.Scene(sim_options=gs.options.SimOptions(dt=0.01))
    scene.add_entity(gs.morphs.Plane())
    c1 = scene.add_entity(gs.morphs.Box(size=(0.4, 0.4, 0.4), pos=(0.0, 0.0, 0.2)))
    c2 = scene.add_entity(gs.morphs.Box(size=(0.4, 0.4, 0.4), pos=(0.0, 0.0, 0.7)))
    scene.build(n_envs=B)
    rng = np.random.default_rng(0)
    q2 = np.array([np.concatenate([[0.0, 0.0, 0.65 + 0.1 * rng.random()],
                                   _xyz_to_quat([*(0.15 * rng.random(2)), np.pi * rng.random()])]) for _ in range(B)])
    c2.set_qpos(torch.as_tensor(q2, dtype=torch.float32))
    for i in range(110):
        scene.step()
        if i > 100:
            v = torch.cat([c1.get_dofs_velocity(), c2.get_dofs_velocity()], 1).cpu().numpy()
            np.testing.assert_allclose(v, 0.0, atol=1e-2)
    np.testing.assert_allclose(c2.get_dofs_position().cpu().numpy()[:, 2], 0.6, atol=2e-3)


def test_125_cubes_stay_on_their_grid():
    """test_rigid_physics.py:877-941 (static case): 125 unit cubes stacked 5 x 5 x 5 with 1 mm gaps stay where they
    are for 150 steps: positions within 5 cm, rotations within 0.03, mean speed below 0.05 over the last 50 steps."""
    scene = gs.Scene(sim_options=gs.options.SimOptions(dt=0.01),
                     rigid_options=gs.options.RigidOptions(max_collision_pairs=1000))
    scene.add_entity(gs.morphs.Plane())
    cubes = []
    for n in range(125):
        i, j, k = n // 25, (n // 5) % 5, n % 5
        cubes.append(scene.add_entity(gs.morphs.Box(pos=(i * (1.0 - 1e-3), j * (1.0 - 1e-3), k * (1.0 - 1e-3) + 0.5),
                                                    size=(1.0, 1.0, 1.0))))
    scene.build()
    for step in range(150):
        scene.step()
        if step > 100:
            v = scene.rigid_solver.get_dofs_velocity().reshape(-1, 6)
            assert float(torch.linalg.norm(v, dim=1).mean()) < 0.05, f"step {step}"
    for n, c in enumerate(cubes):
        i, j, k = n // 25, (n // 5) % 5, n % 5
        q = c.get_dofs_position().cpu().numpy()
        np.testing.assert_allclose(q[:3], (i * (1.0 - 1e-3), j * (1.0 - 1e-3), k * (1.0 - 1e-3) + 0.5), atol=0.05)
        np.testing.assert_allclose(q[3:], 0.0, atol=0.03)


@pytest.mark.parametrize("multi,count", [(True, 4), (False, 1)])
def test_box_on_plane_contact_count(multi, count):
    """test_gstaichi.py:95-124, 151: a box resting on the plane has 4 contacts with multi-contact and 1 without."""
    scene = gs.Scene(sim_options=gs.options.SimOptions(dt=0.01),
                     rigid_options=gs.options.RigidOptions(enable_multi_contact=multi))
    scene.add_entity(gs.morphs.Plane())
    box = scene.add_entity(gs.morphs.Box(size=(0.2, 0.2, 0.2), pos=(0.0, 0.0, 0.1)))
    scene.build(n_envs=2)
    for _ in range(20):
        scene.step()
    c = box.get_contacts()
    assert bool(c["valid_mask"].all()) and c["valid_mask"].shape == (2, count)
    if multi:  # one contact point cannot hold the box flat: it rocks on it
        np.testing.assert_allclose(box.get_pos().cpu().numpy()[:, 2], 0.1, atol=2e-3)
