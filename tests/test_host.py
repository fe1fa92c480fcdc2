"""Host-side checks that need no GPU: the C-ABI library loads and exports every declared symbol, the ctypes mirrors
match the header, model compilation follows the reference's build rules, the loaders parse the workload assets."""
import ctypes
import os
import re

import numpy as np
import pytest

from genesis_amd import _lib
from genesis_amd import workloads as W
from genesis_amd.model import CompileOptions, compile_model

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "genesis_rigid.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(?:int|int64_t)\s+(gsr_\w+)\s*\(", src)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(_lib.LIB_PATH):
        from genesis_amd import build
        build.build()
    return _lib.load()


def test_library_exports_every_declared_symbol(lib):
    names = declared_functions()
    assert {"gsr_step", "gsr_forward_kinematics", "gsr_reset_envs", "gsr_get_errno"} <= set(names)
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) <= set(_lib.EXPORTS) | {"gsr_abi_version", "gsr_model_size", "gsr_state_size"}
    assert lib.gsr_abi_version() == 1


def test_struct_layout_matches_library(lib):
    assert lib.gsr_model_size() == ctypes.sizeof(_lib.GsrModel)
    assert lib.gsr_state_size() == ctypes.sizeof(_lib.GsrState)


def test_null_arguments_are_rejected_without_gpu(lib):
    # argument validation happens before any HIP call
    assert lib.gsr_step(None, None, 1, None) == -1
    assert lib.gsr_get_errno(None, None, None) == -1
    assert lib.gsr_get_contacts(None, None, None, 0, 0, None, None, None) == -1
    assert lib.gsr_links_contact_force(None, None, None, 0, None, None) == -1
    assert lib.gsr_get_equalities(None, None, None, None, None) == -1


def test_checkpoint_skips_only_state_fields():
    from genesis_amd.rigid_solver import CHECKPOINT_SKIP
    assert CHECKPOINT_SKIP <= set(_lib.STATE_POINTERS)
    # the persistent state and caches are saved
    assert not {"qpos", "vel", "qacc_ws", "is_warmstart", "ctrl_mode", "ctrl_pos", "ctrl_vel", "ctrl_force",
                "cache_normal", "sort_value", "sort_ig", "first_time", "links_pos", "links_quat"} & CHECKPOINT_SKIP


def test_wave_hessian_budget_default():
    from genesis_amd.rigid_solver import WAVE_H_FLOATS, wave_h_floats_default
    pile64, pile20 = (compile_model(w.entities, w.options) for w in (W.pile_workload(n_boxes=64),
                                                                   W.pile_workload(n_boxes=20)))
    assert wave_h_floats_default(pile64) == WAVE_H_FLOATS
    # 20 free boxes: 20 x 21 diagonal floats, 4x that rounded up to 256
    assert wave_h_floats_default(pile20) == 1792
    arm = compile_model(W.arm_workload().entities, W.arm_workload().options)
    assert wave_h_floats_default(arm) == 1024


def test_state_size_rules_cover_every_pointer():
    from genesis_amd.rigid_solver import finalize_model, state_sizes
    for w in (W.arm_workload(), W.quadruped_workload()):
        M = compile_model(w.entities, w.options)
        finalize_model(M)
        sz = state_sizes(M)
        assert set(sz) == set(_lib.STATE_POINTERS)


def test_arm_model_matches_panda_description():
    w = W.arm_workload()
    M = compile_model(w.entities, w.options)
    # 9 dofs: 7 revolute + 2 prismatic fingers, 11 robot links + plane
    assert M.n_dofs == 9 and M.n_qs == 9 and M.n_links == 12
    assert list(M.joint_type[M.joint_type > 0]) == [1] * 7 + [2] * 2
    np.testing.assert_allclose(M.dof_kp, [4500, 4500, 3500, 3500, 2000, 2000, 2000, 100, 100])
    np.testing.assert_allclose(M.dof_kv, [450, 450, 350, 350, 200, 200, 200, 10, 10])
    np.testing.assert_allclose(M.dof_armature, 0.1)
    np.testing.assert_allclose(M.dof_damping, 1.0)
    np.testing.assert_allclose(M.dof_force_range[:4], [[-87, 87]] * 4)
    np.testing.assert_allclose(M.dof_force_range[4:7], [[-12, 12]] * 3)
    assert M.n_equalities == 1  # finger coupling (JOINT equality)
    # self-collision off: only plane-vs-robot pairs, none with the fixed base link
    G = M.n_geoms
    for a in range(G):
        for b in range(a + 1, G):
            if M.pair_idx[a, b] >= 0:
                assert M.geom_type[a] == 0 or M.geom_type[b] == 0
                assert not (M.link_is_fixed[M.geom_link[a]] and M.link_is_fixed[M.geom_link[b]])


def test_quadruped_model():
    w = W.quadruped_workload()
    M = compile_model(w.entities, w.options)
    assert M.n_dofs == 18 and M.n_qs == 19
    assert M.joint_type[M.link_joint_start[1]] == 4  # free base
    mass = M.link_inertial_mass[1:].sum()
    assert 12.0 < mass < 18.0  # ~15 kg
    # adjacent links never collide (enable_adjacent_collision=False)
    for a in range(M.n_geoms):
        for b in range(a + 1, M.n_geoms):
            if M.pair_idx[a, b] >= 0:
                la, lb = M.geom_link[a], M.geom_link[b]
                assert la != lb and M.link_parent[la] != lb and M.link_parent[lb] != la


def test_pile_capacities_follow_reference_rules():
    w = W.pile_workload(n_boxes=64, n_envs=4)
    M = compile_model(w.entities, w.options)
    assert M.n_dofs == 384 and M.n_qs == 448 and M.n_geoms == 65
    assert M.n_pairs == 65 * 64 // 2  # every pair of boxes + plane-box, no fixed-fixed pair
    assert M.max_contacts == min(1000, M.n_pairs) * 5
    assert M.max_rows == 4 * M.max_contacts + 0 + 384 + 6 * 8
    assert M.row_width == 12
    # box density 1000 kg/m^3
    s = np.array(w.entities[1].links[0].geoms[0].data)
    np.testing.assert_allclose(M.link_inertial_mass[1], 1000 * s.prod(), rtol=1e-5)


def test_box_vertex_order_and_init_aabb():
    M = compile_model([W.plane_entity(), W.box_entity((0.2, 0.4, 0.6), (0, 0, 1))], CompileOptions())
    v = M.verts[M.geom_vert_start[1]:M.geom_vert_end[1]]
    np.testing.assert_allclose(v[0], [-0.1, -0.2, -0.3])
    np.testing.assert_allclose(v[1], [-0.1, -0.2, 0.3])
    np.testing.assert_allclose(v[7], [0.1, 0.2, 0.3])
    a = M.geom_init_aabb[1]
    np.testing.assert_allclose(a[0], [-0.1, -0.2, -0.3])
    np.testing.assert_allclose(a[7], [0.1, 0.2, 0.3])
    np.testing.assert_allclose(a[5], [0.1, -0.2, 0.3])  # corner order (x hi, y lo, z hi)


def test_meaninertia_divides_inside_entity_loop():
    """rigid_solver_decomp.py:2866-2872 divides by n_dofs once per entity: plane (0 dofs) first, then the box."""
    M = compile_model([W.plane_entity(), W.box_entity((0.2, 0.2, 0.2), (0, 0, 1))], CompileOptions())
    from genesis_amd.model import _host_fk_crba
    *_, Mm = _host_fk_crba(M, M.qpos0.astype(np.float64))
    np.testing.assert_allclose(M.meaninertia, np.trace(Mm) / 6.0, rtol=1e-5)


NOALLOC_SCRATCH = ["jac_val", "jac_idx", "jac_nnz", "efc_diag", "efc_aref", "efc_D", "efc_fl", "efc_Jaref", "efc_jv",
                   "efc_quad", "efc_active", "qacc", "Ma", "grad", "Mgrad", "search", "mv", "qfrc_constraint",
                   "nt_vec", "nt_H", "solver_iters", "ent_row_ptr", "ent_row_list", "efc_ent", "ent_row_run", "gjk_ws",
                   "gjk_defer_n", "gjk_defer", "qpos_next", "vel_next"]


def test_noalloc_scratch_size_matches_state_layout(lib):
    """gsr_noalloc_scratch_bytes = the per-env sizes of the scratch arrays of the state layout, times n_envs, each
    array rounded up to 256 bytes; the entry point rejects null arguments before any HIP call."""
    from genesis_amd import workloads as W
    from genesis_amd.model import compile_model
    from genesis_amd.rigid_solver import finalize_model, state_sizes
    assert lib.gsr_step_noalloc(None, None, 1, None) == -1
    for w in (W.arm_workload(), W.workload_by_name("pile20", 4)):
        M = compile_model(w.entities, w.options)
        finalize_model(M)
        sz = state_sizes(M)
        d = _lib.GsrModel()
        d.n_dofs, d.n_qs, d.n_entities, d.max_rows, d.row_width, d.max_broad = (
            M.n_dofs, M.n_qs, M.n_entities, M.max_rows, M.row_width, M.max_broad)
        for n in (1, 37, 1000):
            want = sum(-(-sz[k] * n * 4 // 256) * 256 for k in NOALLOC_SCRATCH)
            assert lib.gsr_noalloc_scratch_bytes(ctypes.byref(d), n) == want
        assert lib.gsr_noalloc_scratch_bytes(ctypes.byref(d), 0) == -1
