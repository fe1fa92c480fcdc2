"""gsr_step_noalloc (the per-substep scratch in a static device array) gives gsr_step's results bitwise."""
import ctypes

import pytest
import torch

from genesis_amd import _lib
from genesis_amd import workloads as W
from genesis_amd.model import compile_model
from genesis_amd.rigid_solver import RigidSolver

from tests.test_host import NOALLOC_SCRATCH

# state arrays that outlive a substep or are read back after it
COMPARED = ["qpos", "vel", "acc", "acc_smooth", "qacc_ws", "is_warmstart", "dofs_pos", "links_pos", "links_quat",
            "geoms_pos", "geoms_quat", "cache_normal", "sort_value", "sort_ig", "errno_env", "n_contacts",
            "contact_geoms", "contact_pos", "contact_normal", "contact_pen", "contact_force", "efc_force", "n_con",
            "mass_blk", "mass_L_blk", "mass_D_inv"]


def _pair(w, n, variant):
    M = compile_model(w.entities, w.options)
    out = []
    for _ in range(2):
        S = RigidSolver(M, n, "cuda", solver_variant=variant)
        W.init_state(w, M, S)
        out.append(S)
    return M, out


def _noalloc_desc(S):
    d = _lib.GsrState.from_buffer_copy(S.st.desc)
    for k in NOALLOC_SCRATCH:
        setattr(d, k, None)
    return d


def _advance(S, steps, seed):
    c = W.Control(S._w, S.model, S, seed)
    for i in range(steps):
        c(i)
        S.substep()


@pytest.mark.gpu
@pytest.mark.parametrize("scene,variant,sizes,pre", [("arm", 1, (256, 100), 0), ("arm", 0, (512, 64), 0),
                                                       ("pile20", 1, (96, 40), 80), ("quadruped", 0, (256, 33), 20)])
def test_step_noalloc_equals_step(scene, variant, sizes, pre):
    lib = _lib.load()
    for call, n in enumerate(sizes):
        w = W.workload_by_name(scene, n)
        M, (A, B) = _pair(w, n, variant)
        need = lib.gsr_noalloc_scratch_bytes(A.dm.desc, n)
        print(f"{scene} variant {variant} call {call}: {n} envs, scratch {need / 2**20:.1f} MiB", flush=True)
        assert 0 < need <= (256 << 20)
        # both start from the same state a few steps into the scene (the piles in contact)
        for S in (A, B):
            S._w = w
            _advance(S, pre, seed=5)
        torch.cuda.synchronize()
        for k in COMPARED:
            assert torch.equal(A.st.t[k], B.st.t[k]), k
        ca, cb = W.Control(w, M, A, 7), W.Control(w, M, B, 7)
        dB = _noalloc_desc(B)
        for i in range(3):
            ca(i)
            cb(i)
            _lib.check(lib.gsr_step(A.dm.desc, A.st.desc, 1, A._stream()), "gsr_step")
            _lib.check(lib.gsr_step_noalloc(B.dm.desc, ctypes.byref(dB), 1, B._stream()), "gsr_step_noalloc")
        torch.cuda.synchronize()
        for k in COMPARED:
            assert torch.equal(A.st.t[k], B.st.t[k]), (call, k)


@pytest.mark.gpu
def test_step_noalloc_rejects_oversized_batches():
    lib = _lib.load()
    w = W.workload_by_name("pile20", 8)
    M = compile_model(w.entities, w.options)
    S = RigidSolver(M, 8, "cuda")
    per_env = lib.gsr_noalloc_scratch_bytes(S.dm.desc, 1)
    n_big = (256 << 20) // per_env + 64
    d = _lib.GsrState.from_buffer_copy(S.st.desc)
    d.n_envs = n_big  # never launched: the size check comes first
    assert lib.gsr_noalloc_scratch_bytes(S.dm.desc, n_big) > (256 << 20)
    assert lib.gsr_step_noalloc(S.dm.desc, ctypes.byref(d), 1, S._stream()) == -2
