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


@pytest.mark.parametrize("tiles,launch,deep", [
    ((64, 64, 1), dict(persistent=True, waves_per_simd=4), False),
    ((32, 64, 3), dict(persistent=False, waves_per_simd=4), False),
    ((64, 16, 4), dict(persistent=True, waves_per_simd=2), False),
    ((64, 64, 1), dict(persistent=False, waves_per_simd=2, gjk_inline=False), True),
    ((32, 64, 2), dict(persistent=True, waves_per_simd=4, gjk_inline=True), True),
    ((64, 64, 1), dict(phased=True), False),
    ((32, 16, 2), dict(phased=True), True),
])
def test_fused_equals_plain_bitwise(tiles, launch, deep):
    """deep: overlapping boxes in the pile part, so the GJK/EPA pass runs (deferred to the resume launch, or inside
    the fused kernel) while the plain step runs it in its own pass."""
    from genesis_amd.batch import FusedBatch, Group
    A, B = _parts(COUNTS), _parts(COUNTS)
    if deep:
        _overlap_pile(A)
        _overlap_pile(B)
    # plain variants: lane solver for arm/quadruped, wavefront solver for the pile
    fb = FusedBatch([Group(S, envs_per_tile=t) for (_, _, S, _), t in zip(B, tiles)][::-1], **launch)
    for step in range(12):
        for parts in (A, B):
            for _, _, S, c in parts:
                c(step)
        for _, _, S, _ in A:
            S.substep()
        fb.step()
        torch.cuda.synchronize()
        _assert_same(_state(A), _state(B), f"step {step} tiles {tiles} {launch}")
    if deep:
        for parts in (A, B):
            e = parts[2][2].st.t["errno_env"][0]
            assert bool(((e & 0x100) != 0).all()), "the GJK/EPA pass did not run in every pile env"
    for _, _, S, _ in B:
        S.check_errno()


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
