"""Batching form on the GPU: the fused launch (and its hipGraph) gives bitwise the plain drop-in's result.

The fused kernel runs the plain path's per-env device code, so for the same solver variant every env's qpos/vel must
match the plain `gsr_step` bit for bit, whatever the tile size, the group order or the other groups in the launch.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _parts(counts, seed=0):
    from genesis_amd.model import compile_model
    from genesis_amd.rigid_solver import RigidSolver
    from genesis_amd.workloads import Control, init_state, workload_by_name
    out = []
    for i, (name, n) in enumerate(counts):
        w = workload_by_name(name, n, dt=0.01)
        M = compile_model(w.entities, w.options)
        S = RigidSolver(M, n, "cuda")
        init_state(w, M, S)
        out.append((w, M, S, Control(w, M, S, seed + i)))
    return out


def _state(parts):
    return [(S.st.t["qpos"].clone(), S.st.t["vel"].clone()) for _, _, S, _ in parts]


def _assert_same(a, b, what):
    for i, ((qa, va), (qb, vb)) in enumerate(zip(a, b)):
        assert torch.equal(qa, qb), f"{what}: group {i} qpos differs (max {float((qa - qb).abs().max()):.3g})"
        assert torch.equal(va, vb), f"{what}: group {i} vel differs (max {float((va - vb).abs().max()):.3g})"


COUNTS = [("arm", 70), ("quadruped", 90), ("pile20", 12)]


def _overlap_pile(parts, seed=11):
    """Random overlapping box poses for the pile part: MPR hands many pairs to GJK/EPA every step."""
    for w, M, S, _ in parts:
        if not w.name.startswith("pile"):
            continue
        rng = np.random.default_rng(seed)
        B, n = S._B, M.n_qs // 7
        q = np.zeros((B, n, 7))
        q[:, :, 0:2] = rng.uniform(-0.25, 0.25, (B, n, 2))
        q[:, :, 2] = rng.uniform(0.1, 0.6, (B, n))
        qq = rng.normal(size=(B, n, 4))
        q[:, :, 3:7] = qq / np.linalg.norm(qq, axis=-1, keepdims=True)
        S.set_qpos(torch.as_tensor(q.reshape(B, -1), dtype=torch.float32))


# (solver, collide, dyn) variants of every group: the all-lane path of the fused tile runner, each mixed branch, and
# the all-wavefront path; each with the single fused kernel (GJK/EPA inside or deferred to the resume launch) and the
# phased launches, with and without deep penetrations (the GJK/EPA pass)
VARIANTS = [(0, 0, 0), (0, 1, 0), (1, 0, 1), (0, 0, 1), (1, 1, 0), (1, 1, 1)]
LAUNCHES = [dict(persistent=True, waves_per_simd=4, gjk_inline=False),
            dict(persistent=False, waves_per_simd=2, gjk_inline=True),
            dict(phased=True)]


@pytest.mark.parametrize("variants", VARIANTS)
@pytest.mark.parametrize("li", range(len(LAUNCHES)))
def test_fused_equals_plain_bitwise(variants, li):
    """Every group and the plain solvers run the same (solver, collide, dyn) variants; the pile part starts from
    overlapping poses in every other case (deep: the GJK/EPA pass runs, deferred to the resume launch, inside the
    fused kernel or in its phase, while the plain step runs it in its own pass)."""
    from genesis_amd.batch import FusedBatch, Group
    launch = LAUNCHES[li]
    deep = (VARIANTS.index(variants) + li) % 2 == 1
    sv, cv, dv = variants
    tiles = (64, 64, 1) if sv == 0 else (4, 2, 1)
    if cv == 0 and sv == 0:
        tiles = (32, 64, 3)
    A, B = _parts(COUNTS), _parts(COUNTS)
    for parts in (A, B):
        for _, _, S, _ in parts:
            S.set_variants(sv, cv, dv)
    if deep:
        _overlap_pile(A)
        _overlap_pile(B)
    fb = FusedBatch([Group(S, envs_per_tile=t) for (_, _, S, _), t in zip(B, tiles)][::-1], **launch)
    for (_, _, SA, _), (_, _, SB, _) in zip(A, B):
        assert tuple(SA.dm.desc.__getattribute__(k) for k in ("solver_variant", "collide_variant", "dyn_variant")) == \
            tuple(SB.dm.desc.__getattribute__(k) for k in ("solver_variant", "collide_variant", "dyn_variant"))
    for step in range(10):
        for parts in (A, B):
            for _, _, S, c in parts:
                c(step)
        for _, _, S, _ in A:
            S.substep()
        fb.step()
        torch.cuda.synchronize()
        _assert_same(_state(A), _state(B), f"step {step} variants {variants} {launch}")
    if deep:
        for parts in (A, B):
            e = parts[2][2].st.t["errno_env"][0]
            assert bool(((e & 0x100) != 0).all()), "the GJK/EPA pass did not run in every pile env"
    for _, _, S, _ in B:
        S.check_errno()


@pytest.mark.parametrize("variants,tiles,solve_tiles,deep", [
    ((1, 0, 0), (64, 64, 16), (1, 2, 1), False),
    ((1, 0, 0), (32, 64, 12), (2, 4, 3), True),
    ((1, 1, 1), (8, 4, 2), (2, 1, 1), True),
    ((0, 0, 0), (64, 64, 3), (16, 8, 1), True),
])
def test_phased_solve_tiles_equal_plain_bitwise(variants, tiles, solve_tiles, deep):
    """The phased launch cuts each group twice: envs_per_tile for the per-env stages and solve tiles for the GJK/EPA
    pass and the Newton solve. Any pair of tilings gives the plain step's bits."""
    from genesis_amd.batch import FusedBatch, Group
    sv, cv, dv = variants
    A, B = _parts(COUNTS), _parts(COUNTS)
    for parts in (A, B):
        for _, _, S, _ in parts:
            S.set_variants(sv, cv, dv)
        if deep:
            _overlap_pile(parts)
    fb = FusedBatch([Group(S, envs_per_tile=t, wave_envs_per_tile=w)
                     for (_, _, S, _), t, w in zip(B, tiles, solve_tiles)], phased=True)
    assert fb.info.n_wave_tiles == sum(-(-S._B // w) for (_, _, S, _), w in zip(B, solve_tiles))
    for step in range(8):
        for parts in (A, B):
            for _, _, S, c in parts:
                c(step)
        for _, _, S, _ in A:
            S.substep()
        fb.step()
        torch.cuda.synchronize()
        _assert_same(_state(A), _state(B), f"step {step} variants {variants} tiles {tiles} solve tiles {solve_tiles}")


def test_batch_upload_rejects_conflicting_groups():
    """Groups over one state must agree on the variants (the row layout follows the solver variant) and must not
    share envs."""
    import ctypes
    from genesis_amd import _lib
    from genesis_amd.batch import FusedBatch, Group
    P = _parts([("quadruped", 40)])
    S = P[0][2]
    with pytest.raises(ValueError):
        FusedBatch([Group(S, 0, 20, 4, 1, 1), Group(S, 20, 20, 4, 0, 1)])
    with pytest.raises(RuntimeError):
        FusedBatch([Group(S, 0, 25, 4), Group(S, 20, 20, 4)])
    # the C-ABI itself refuses a variant mix on one state (the Python wrapper is not the only caller)
    lib = _lib.load()
    arr = (_lib.GsrGroup * 2)()
    for i, (b, n, v) in enumerate(((0, 20, 1), (20, 20, 0))):
        m = _lib.GsrModel.from_buffer_copy(S.dm.desc)
        m.solver_variant = v
        arr[i].model, arr[i].state = m, S.st.desc
        arr[i].env_begin, arr[i].n_envs, arr[i].envs_per_tile = b, n, 4
    nbytes = lib.gsr_batch_table_bytes(2)
    table = torch.zeros(nbytes, dtype=torch.uint8, device="cuda")
    info = _lib.GsrBatchInfo()
    torch.cuda.synchronize()
    assert lib.gsr_batch_upload(arr, 2, ctypes.c_void_p(table.data_ptr()), nbytes, ctypes.byref(info)) == -1


def test_fused_group_subrange_leaves_other_envs_alone():
    from genesis_amd.batch import FusedBatch, Group
    A, B = _parts([("quadruped", 50)]), _parts([("quadruped", 50)])
    SA, SB = A[0][2], B[0][2]
    before = _state(B)
    fb = FusedBatch([Group(SB, env_begin=10, n_envs=25, envs_per_tile=16)])
    for step in range(5):
        A[0][3](step)
        B[0][3](step)
        SA.substep()
        fb.step()
    torch.cuda.synchronize()
    qa, qb = SA.st.t["qpos"], SB.st.t["qpos"]
    assert torch.equal(qa[:, 10:35], qb[:, 10:35])
    assert torch.equal(qb[:, :10], before[0][0][:, :10]) and torch.equal(qb[:, 35:], before[0][0][:, 35:])


@pytest.mark.parametrize("launch", [dict(), dict(phased=True)])
def test_graph_replays_match_eager_fused(launch):
    from genesis_amd.batch import FusedBatch, Group
    A, B = _parts(COUNTS), _parts(COUNTS)
    fa = FusedBatch([Group(S) if S.dm.desc.solver_variant == 0 else Group(S, envs_per_tile=2) for _, _, S, _ in A],
                    **launch)
    fbb = FusedBatch([Group(S) if S.dm.desc.solver_variant == 0 else Group(S, envs_per_tile=2) for _, _, S, _ in B],
                     **launch)
    fbb.capture(1)
    for step in range(6):
        for parts in (A, B):
            for _, _, S, c in parts:
                c(step)
        fa.step()
        fbb.replay()
        # eager work and a synchronisation between replays
        x = torch.randn(1000, device="cuda").sum()
        torch.cuda.synchronize()
        _assert_same(_state(A), _state(B), f"replay {step}")
    assert torch.isfinite(x)


def test_autotune_returns_valid_plan():
    from genesis_amd.batch import FusedBatch, Group, autotune
    P = _parts([("arm", 64), ("pile20", 8)])
    before = _state(P)
    tuned, launch, report = autotune([Group(S) for _, _, S, _ in P], steps=1, warmup=1)
    _assert_same(before, _state(P), "autotune restores the state")
    assert len(tuned) == 2 and all(r["ms"] > 0 for r in report)
    fb = FusedBatch(tuned, **launch)
    fb.step(2)
    torch.cuda.synchronize()
    for _, _, S, _ in P:
        assert torch.isfinite(S.st.t["qpos"]).all()
