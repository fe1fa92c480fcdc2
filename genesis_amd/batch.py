"""Batching form: heterogeneous envs (different robots and DoF counts) stepped by one fused launch per substep.

The reference batches only homogeneous envs (one model per scene; rigid_solver_decomp.py:80-352), so this form has
no reference counterpart. It changes execution, not arithmetic: every env runs the same device code as in the plain
drop-in (`gsr_step`), so with the same solver variant a fused result is bitwise the plain result.

* `FusedBatch(groups)` uploads one device table (`gsr_batch_upload`) and steps every group with `gsr_step_fused`:
  one-wavefront workgroups take tiles of `envs_per_tile` envs, heaviest group first, either in one launch that runs
  the whole substep per tile (from a persistent work queue or one workgroup per tile) or, `phased=True`, in one
  launch per substep stage (dynamics, collision, GJK/EPA pass, solve, integration), each over every tile.
* `capture()` records n substeps into a hipGraph (`gsr_graph_capture`); `replay()` launches it.
* `autotune()` times each group alone under each candidate (solver variant, collision variant, envs per tile) on a
  copy of its state, for the single-kernel and the phased launches, orders the winners by per-tile cost (longest
  first), then times the launch forms.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import List, Optional, Sequence, Tuple

import torch

from . import _lib
from .rigid_solver import WAVE_MAX_ENTITIES, RigidSolver


@dataclass
class Group:
    """Envs [env_begin, env_begin + n_envs) of one solver, with the execution choices for the fused launch."""
    solver: RigidSolver
    env_begin: int = 0
    n_envs: Optional[int] = None
    envs_per_tile: int = 64
    solver_variant: Optional[int] = None
    collide_variant: Optional[int] = None
    # phased launch only: envs per tile of the GJK/EPA pass and the solve (None: envs_per_tile)
    wave_envs_per_tile: Optional[int] = None

    def resolved(self) -> "Group":
        n = self.solver._B - self.env_begin if self.n_envs is None else self.n_envs
        d = self.solver.dm.desc
        v = int(d.solver_variant) if self.solver_variant is None else int(self.solver_variant)
        c = int(d.collide_variant) if self.collide_variant is None else int(self.collide_variant)
        return Group(self.solver, self.env_begin, n, self.envs_per_tile, v, c, self.wave_envs_per_tile)


class FusedBatch:
    """One device table over several groups; `step()` runs fused substeps, `capture()`/`replay()` a hipGraph."""

    def __init__(self, groups: Sequence[Group], persistent: bool = True, waves_per_simd: int = 4,
                 gjk_inline: bool = False, phased: bool = False):
        if not groups:
            raise ValueError("empty batch")
        self.lib = _lib.load()
        self.groups = [g.resolved() for g in groups]
        # the solver's own descriptor takes the group's variants, so plain launches and readbacks of the same state
        # (constraint rows and Hessian are laid out by the solver variant) see the layout the fused step wrote
        chosen = {}
        for g in self.groups:
            v = (g.solver_variant, g.collide_variant)
            if chosen.setdefault(id(g.solver), v) != v:
                raise ValueError("groups of one solver must use the same solver and collision variants")
        for g in self.groups:
            g.solver.set_variants(g.solver_variant, g.collide_variant)
        self.device = self.groups[0].solver.device
        n = len(self.groups)
        nbytes = int(self.lib.gsr_batch_table_bytes(n))
        if nbytes < 0:
            raise ValueError(f"a batch holds 1..64 groups, got {n}")
        self.table = torch.zeros(nbytes, dtype=torch.uint8, device=self.device)
        arr = (_lib.GsrGroup * n)()
        for i, g in enumerate(self.groups):
            arr[i].model = g.solver.dm.desc
            arr[i].state = g.solver.st.desc
            arr[i].env_begin, arr[i].n_envs, arr[i].envs_per_tile = g.env_begin, g.n_envs, g.envs_per_tile
            arr[i].wave_envs_per_tile = int(g.wave_envs_per_tile or 0)
        self._arr = arr
        self.info = _lib.GsrBatchInfo()
        self.info.persistent = int(bool(persistent))
        self.info.waves_per_simd = int(waves_per_simd)
        self.info.gjk_inline = int(bool(gjk_inline))
        self.info.phased = int(bool(phased))
        torch.cuda.synchronize(self.device)  # the table's zero fill precedes the synchronous upload
        _lib.check(self.lib.gsr_batch_upload(arr, n, ctypes.c_void_p(self.table.data_ptr()), nbytes,
                                             ctypes.byref(self.info)), "gsr_batch_upload")
        self._graphs = []
        self._side = None

    @property
    def n_envs(self) -> int:
        return sum(g.n_envs for g in self.groups)

    def _stream(self):
        return ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def step(self, n_substeps: int = 1):
        _lib.check(self.lib.gsr_step_fused(ctypes.c_void_p(self.table.data_ptr()), ctypes.byref(self.info),
                                           n_substeps, self._stream()), "gsr_step_fused")

    def kernel_names(self) -> List[str]:
        """The kernels of one fused substep, in launch order (`launch_kernel(k)` launches the k-th)."""
        i = self.info
        if i.phased:
            return [f"k_phase<{p}>" for p in range(5)]
        names = [f"k_fused<{i.waves_per_simd}, {i.gjk_inline}>"]
        if not i.gjk_inline:
            names += ["k_gjk_resume"] * i.n_types
        assert len(names) == self.lib.gsr_fused_kernel_count(ctypes.byref(i))
        return names

    def launch_kernel(self, k: int):
        _lib.check(self.lib.gsr_launch_fused_kernel(ctypes.c_void_p(self.table.data_ptr()), ctypes.byref(self.info),
                                                    k, self._stream()), "gsr_launch_fused_kernel")

    def capture(self, n_substeps: int = 1):
        """Record n_substeps fused launches into a hipGraph on a side stream; `replay()` launches the latest."""
        if self._side is None:
            self._side = torch.cuda.Stream(self.device)  # default priority
        self._side.wait_stream(torch.cuda.current_stream(self.device))
        exe = ctypes.c_void_p()
        rc = self.lib.gsr_graph_capture(ctypes.c_void_p(self.table.data_ptr()), ctypes.byref(self.info), n_substeps,
                                        ctypes.c_void_p(self._side.cuda_stream), ctypes.byref(exe))
        _lib.check(rc, "gsr_graph_capture")
        self._graphs.append(exe)
        return exe

    def replay(self, graph=None):
        g = graph if graph is not None else self._graphs[-1]
        _lib.check(self.lib.gsr_graph_launch(g, self._stream()), "gsr_graph_launch")

    def check_errno(self):
        for g in self.groups:
            g.solver.check_errno()

    def close(self):
        for g in self._graphs:
            self.lib.gsr_graph_destroy(g)
        self._graphs = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def candidates(solver: RigidSolver, phased: bool = False) -> List[Tuple[int, int, int, Optional[int]]]:
    """(solver_variant, collide_variant, envs_per_tile, wave_envs_per_tile) choices worth timing for one scene type.
    Under the phased launch the solve (and the GJK/EPA pass) has tiles of its own, so a lane-per-env collision with
    64 envs a tile can sit beside a wavefront solve with one to four."""
    M = solver.model
    out = [(0, 0, 64, None), (0, 0, 32, None), (0, 1, 8, None)]
    if M.n_entities <= WAVE_MAX_ENTITIES:
        out += [(1, 1, t, None) for t in (1, 2, 4, 8)] + [(1, 0, 4, None)]
        if phased:
            out += [(1, 0, 64, w) for w in (1, 2, 4)] + [(1, 1, 4, 1), (1, 1, 8, 2)]
    return out


def _snapshot(solver: RigidSolver):
    return {k: v.clone() for k, v in solver.st.t.items()}


def _restore(solver: RigidSolver, snap):
    for k, v in snap.items():
        solver.st.t[k].copy_(v)


def time_batch(groups: Sequence[Group], steps: int = 3, warmup: int = 1, **launch) -> float:
    """Milliseconds per fused substep of `groups` (their solvers' states are restored afterwards)."""
    solvers = list({id(g.solver): g.solver for g in groups}.values())
    snaps = [_snapshot(S) for S in solvers]
    fb = FusedBatch(groups, **launch)
    for _ in range(warmup):
        fb.step()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    fb.step(steps)
    e1.record()
    torch.cuda.synchronize(solvers[0].device)
    for S, sn in zip(solvers, snaps):
        _restore(S, sn)
    return e0.elapsed_time(e1) / steps


def time_group(group: Group, steps: int = 3, warmup: int = 1, **launch) -> float:
    return time_batch([group], steps, warmup, **launch)


# launch forms: persistent queue or one workgroup per tile; 128 or 256 VGPRs; GJK/EPA inside the fused kernel (one
# launch per substep, lower occupancy) or deferred to a per-type resume launch; or phased: one launch per substep
# stage over every tile, each stage with its own register allocation
LAUNCH_CANDIDATES = [dict(persistent=p, waves_per_simd=w, gjk_inline=g)
                     for g in (False, True) for w in (4, 2) for p in (True, False)] + [dict(phased=True)]


def _tune_groups(groups, launch, steps, warmup, verbose, log, report):
    """Pick (solver variant, collision variant, envs per tile, solve tile) per group by timing each group alone under
    `launch`, then order the groups by per-tile cost, descending, so the longest tiles start first. Under the phased
    launch the choices are then refined on the whole batch, one group at a time (a group timed alone shows its
    latency; in the batch its stages overlap the other groups' and what counts is the work it adds)."""
    phased = bool(launch.get("phased"))
    tuned, cands = [], []
    for gi, g in enumerate(groups):
        g = g.resolved()
        best = None
        cs = candidates(g.solver, phased)
        for v, c, t, w in cs:
            ms = time_group(Group(g.solver, g.env_begin, g.n_envs, t, v, c, w), steps, warmup, **launch)
            report.append({"group": gi, "solver": v, "collide": c, "envs_per_tile": t, "solve_tile": w, "ms": ms,
                           "phased": phased})
            if verbose:
                log(f"  group {gi} solver {v} collide {c} tile {t} solve tile {w} {launch}: {ms:.3f} ms")
            if best is None or ms < best[0]:
                best = (ms, v, c, t, w)
        ms, v, c, t, w = best
        n_tiles = (g.n_envs + t - 1) // t
        tuned.append((ms / n_tiles, Group(g.solver, g.env_begin, g.n_envs, t, v, c, w)))
        cands.append(cs)
    order = sorted(range(len(tuned)), key=lambda i: -tuned[i][0])
    plan = [tuned[i][1] for i in order]
    if not phased or len(plan) < 2:
        return plan
    best_ms = time_batch(plan, steps, warmup, **launch)
    for pos, gi in enumerate(order):
        for v, c, t, w in cands[gi]:
            g0 = plan[pos]
            if (v, c, t, w) == (g0.solver_variant, g0.collide_variant, g0.envs_per_tile, g0.wave_envs_per_tile):
                continue
            trial = list(plan)
            trial[pos] = Group(g0.solver, g0.env_begin, g0.n_envs, t, v, c, w)
            ms = time_batch(trial, steps, warmup, **launch)
            report.append({"group": gi, "solver": v, "collide": c, "envs_per_tile": t, "solve_tile": w, "ms": ms,
                           "phased": True, "in_batch": True})
            if verbose:
                log(f"  in batch: group {gi} solver {v} collide {c} tile {t} solve tile {w}: {ms:.3f} ms")
            if ms < best_ms:
                best_ms, plan = ms, trial
    return plan


def autotune(groups: Sequence[Group], steps: int = 3, warmup: int = 1, verbose: bool = False, log=print):
    """Tune the per-group choices twice, once under the single fused kernel and once under the phased launches (their
    best variants differ: the fused kernel's register budget is shared by every stage), then time the whole batch
    under every launch form with the matching plan and keep the fastest.

    Returns (tuned groups, launch kwargs for FusedBatch, report of every candidate's ms per substep)."""
    report = []
    plans = {False: _tune_groups(groups, LAUNCH_CANDIDATES[0], steps, warmup, verbose, log, report),
             True: _tune_groups(groups, dict(phased=True), steps, warmup, verbose, log, report)}
    best = None
    for lc in LAUNCH_CANDIDATES:
        plan = plans[bool(lc.get("phased"))]
        ms = time_batch(plan, steps, warmup, **lc)
        report.append(dict(group="all", launch=lc, ms=ms))
        if verbose:
            log(f"  batch {lc}: {ms:.3f} ms")
        if best is None or ms < best[0]:
            best = (ms, lc, plan)
    return best[2], best[1], report
