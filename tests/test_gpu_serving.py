"""Serving form on the GPU (device-side scheduler): continuous batching gives each request the result it gets alone.

A request's envs run the same device code wherever its tiles are and whatever else is in the launch, and its
controls come from a hash of its own indices, so its final state must be bitwise the same alone, in a mix that
changes while it runs, in tiles another request has just left, and on the plain drop-in path (a RigidSolver of
exactly its envs driven with the same targets).
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _server(caps=(("arm", 64), ("quadruped", 64), ("pile20", 16)), budget=1 << 36, tiles=None, **kw):
    from genesis_amd.serving import Pool, Server
    pools = {s: Pool(s, c, dt=0.01) for s, c in caps}
    for s, t in (tiles or {}).items():
        pools[s].set_execution(envs_per_tile=t)
    return Server(pools, budget, **kw)


def _same(a, b, what):
    for k in ("qpos", "vel"):
        assert torch.equal(a[k], b[k]), f"{what} {k}: max diff {float((a[k] - b[k]).abs().max()):.3g}"


def _copy(r):
    from genesis_amd.serving import Request
    return Request(r.rid, r.scene, r.n_envs, r.horizon, 0.0, r.seed)


def test_request_same_alone_and_in_mix():
    from genesis_amd.serving import Request
    probes = [Request(100, "arm", 5, 30, 0.0, seed=7), Request(101, "quadruped", 6, 25, 0.0, seed=8),
              Request(102, "pile20", 3, 15, 0.0, seed=9)]
    alone = {}
    srv = _server()
    for r in probes:
        q = _copy(r)
        srv.run([q], record=True)
        alone[r.rid] = q.result
    # the mix: other requests take the first tiles of every pool, more arrive while the probes run
    mix = [Request(0, "arm", 7, 12, 0.0, seed=1), Request(1, "quadruped", 9, 40, 0.0, seed=2),
           Request(2, "pile20", 2, 8, 0.0, seed=3)]
    mix += [_copy(r) for r in probes]
    mix += [Request(3, "arm", 3, 20, 0.01, seed=4), Request(4, "pile20", 4, 10, 0.02, seed=5),
            Request(5, "quadruped", 11, 5, 0.03, seed=6)]
    srv = _server(tiles={"arm": 4, "quadruped": 2})  # other tile sizes than the lone runs
    st = srv.run(mix, record=True)
    assert st["completed"] == len(mix) and st["failed"] == 0
    got = {r.rid: r for r in mix}
    for r in probes:
        _same(alone[r.rid], got[r.rid].result, f"request {r.rid} ({r.scene}) alone vs in the mix")


def test_reused_tiles_give_the_alone_result():
    """A request admitted into tiles another request has just left gets the result it gets in a fresh pool."""
    from genesis_amd.serving import Request
    probe = Request(11, "pile20", 4, 12, 0.0, seed=5)
    fresh = _copy(probe)
    _server(caps=(("pile20", 4),)).run([fresh], record=True)
    srv = _server(caps=(("pile20", 4),))
    first = Request(10, "pile20", 4, 30, 0.0, seed=4)
    second = _copy(probe)
    st = srv.run([first, second], record=True)
    assert st["completed"] == 2 and second.done_tick == first.done_tick + probe.horizon
    _same(fresh.result, second.result, "request in reused tiles vs in a fresh pool")


def _mix32(h):
    h = np.uint32(h)
    h ^= h >> np.uint32(16)
    h = np.uint32((int(h) * 0x85EBCA6B) & 0xFFFFFFFF)
    h ^= h >> np.uint32(13)
    h = np.uint32((int(h) * 0xC2B2AE35) & 0xFFFFFFFF)
    h ^= h >> np.uint32(16)
    return h


def _hash_uniform(seed, env, step, dof):
    """Host restatement of the scheduler's control hash (four chained MurmurHash3-finaliser rounds)."""
    m = 0xFFFFFFFF
    h = _mix32(np.uint32(seed) ^ np.uint32(0x9E3779B9))
    h = _mix32(h ^ np.uint32((env * 0x27D4EB2F) & m))
    h = _mix32(h ^ np.uint32((step * 0x165667B1) & m))
    h = _mix32(h ^ np.uint32((dof * 0x61C88647) & m))
    return np.float32(int(h) >> 8) * np.float32(1.0 / 16777216.0)


@pytest.mark.parametrize("scene,n,h,seed", [("arm", 4, 120, 11), ("quadruped", 5, 20, 12), ("pile20", 2, 12, 13)])
def test_request_matches_plain_drop_in(scene, n, h, seed):
    """The served request equals a RigidSolver of exactly its envs from the same initial poses, driven with the same
    targets (the hash restated on the host; the arm redraws its targets at steps 0 and 100)."""
    from genesis_amd.model import compile_model
    from genesis_amd.rigid_solver import RigidSolver
    from genesis_amd.serving import Request
    from genesis_amd.workloads import GO2_STANDING, init_state, workload_by_name
    srv = _server()
    q = Request(0, scene, n, h, 0.0, seed)
    srv.run([Request(9, scene, 3, 5, 0.0, 99), q], record=True)  # q sits behind another request
    w = workload_by_name(scene, n, dt=0.01)
    M = compile_model(w.entities, w.options)
    S = RigidSolver(M, n, "cuda", solver_variant=int(srv.pools[scene].S.dm.desc.solver_variant))
    S.set_variants(None, int(srv.pools[scene].S.dm.desc.collide_variant),
                   int(srv.pools[scene].S.dm.desc.dyn_variant))
    init_state(w, M, S)
    lo, hi = M.dof_limit[:, 0].astype(np.float32), M.dof_limit[:, 1].astype(np.float32)
    for step in range(h):
        if scene == "arm" and step % 100 == 0:
            u = np.array([[_hash_uniform(seed, e, step, d) for d in range(M.n_dofs)] for e in range(n)], np.float32)
            S.control_dofs_position(torch.as_tensor(lo + (hi - lo) * u))
        elif scene == "quadruped":
            u = np.array([[_hash_uniform(seed, e, step, d) for d in range(12)] for e in range(n)], np.float32)
            tgt = GO2_STANDING.astype(np.float32) + (u * np.float32(0.6) - np.float32(0.3))
            S.control_dofs_position(torch.as_tensor(tgt), dofs_idx=np.arange(6, 18))
        S.substep()
    torch.cuda.synchronize()
    _same({k: S.st.t[k].T for k in ("qpos", "vel")}, q.result, f"{scene}: served vs plain")


def test_budgets_hold_back_and_release_requests():
    """With a memory budget of one request, requests run one after another, in arrival order."""
    from genesis_amd.serving import Request
    srv = _server(caps=(("arm", 64),))
    per = srv.pools["arm"].bytes_per_env * 10
    srv = _server(caps=(("arm", 64),), budget=per)
    reqs = [Request(i, "arm", 10, 6, 0.0, seed=i) for i in range(4)]
    st = srv.run(reqs, record=True)
    assert st["completed"] == 4
    ticks = [r.done_tick for r in reqs]
    assert ticks == sorted(ticks) and len(set(ticks)) == 4 and ticks[1] - ticks[0] == 6


def test_bench_sized_pools_large_request_alone_and_in_mix():
    """Pools of 8,192 envs per scene type (the serving benchmark's), a 4,096-env pile request alone and among
    others arriving while it runs."""
    from genesis_amd.serving import Request
    caps = (("arm", 8192), ("quadruped", 8192), ("pile20", 8192))
    srv = _server(caps=caps)
    big = Request(50, "pile20", 4096, 12, 0.0, seed=21)
    lone = _copy(big)
    srv.run([lone], record=True)
    mix = [Request(1, "quadruped", 3000, 20, 0.0, seed=1), Request(2, "pile20", 2500, 6, 0.0, seed=2), _copy(big),
           Request(3, "arm", 4096, 15, 0.001, seed=3), Request(4, "pile20", 1000, 9, 0.002, seed=4)]
    st = srv.run(mix, record=True)
    assert st["completed"] == len(mix) and st["failed"] == 0
    _same(lone.result, mix[2].result, "4,096-env pile request alone vs in the mix")
