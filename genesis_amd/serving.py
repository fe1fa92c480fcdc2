"""Serving form: continuous batching of rollout requests on one GPU (no reference counterpart).

A request asks for `horizon` steps of `n_envs` envs of one scene type. The server keeps one resident state pool per
scene type in HBM (a RigidSolver whose envs are slots), admits queued requests in arrival order as slots and the
memory budget allow, and steps every admitted request together: each tick is one fused launch (`gsr_step_fused`)
over the active slot ranges, so requests that arrive while others run join the next tick and their work overlaps
on the device. Finished requests free their slots at once.

Per-env arithmetic does not depend on the slot, the pool size or the other requests in the launch, and a request's
controls come from its own seeded generator, so a request's result is the same alone and in any mix.
"""
from __future__ import annotations

import bisect
import collections
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

import numpy as np
import torch

from .batch import FusedBatch, Group
from .model import compile_model
from .rigid_solver import ERRNO_MESSAGES, RigidSolver
from .workloads import Control, init_state, workload_by_name


class SlotAllocator:
    """First-fit allocator of contiguous env ranges in [0, capacity); adjacent free ranges coalesce."""

    def __init__(self, capacity: int):
        self.capacity = capacity
        self.free: List[Tuple[int, int]] = [(0, capacity)]  # sorted (begin, length)

    def alloc(self, n: int) -> Optional[int]:
        for i, (b, l) in enumerate(self.free):
            if l >= n:
                if l == n:
                    self.free.pop(i)
                else:
                    self.free[i] = (b + n, l - n)
                return b
        return None

    def release(self, begin: int, n: int):
        i = bisect.bisect_left(self.free, (begin, 0))
        self.free.insert(i, (begin, n))
        merged = []
        for b, l in self.free:
            if merged and merged[-1][0] + merged[-1][1] == b:
                merged[-1] = (merged[-1][0], merged[-1][1] + l)
            else:
                merged.append((b, l))
        self.free = merged

    @property
    def n_free(self) -> int:
        return sum(l for _, l in self.free)


@dataclass
class Request:
    rid: int
    scene: str
    n_envs: int
    horizon: int
    arrival: float = 0.0  # seconds after the start of the run
    seed: int = 0
    # filled by the server
    slot: Optional[int] = None
    admitted: Optional[float] = None
    finished: Optional[float] = None
    steps_done: int = 0
    failed: bool = False
    result: Optional[Dict[str, torch.Tensor]] = None
    control: Optional[Control] = field(default=None, repr=False)


class Pool:
    """Resident state of one scene type: `capacity` env slots plus the tuned execution choices for it."""

    def __init__(self, scene: str, capacity: int, dt: float, device="cuda", solver_variant=None, envs_per_tile=None):
        self.scene = scene
        self.w = workload_by_name(scene, capacity, dt=dt)
        self.M = compile_model(self.w.entities, self.w.options)
        self.S = RigidSolver(self.M, capacity, device, solver_variant=solver_variant)
        self.slots = SlotAllocator(capacity)
        v = int(self.S.dm.desc.solver_variant)
        self.envs_per_tile = envs_per_tile or (64 if v == 0 else 1)
        self.bytes_per_env = sum(t.numel() * t.element_size() for t in self.S.st.t.values()) // capacity
        self.ms_per_env = 0.0  # GPU time one env adds to a tick (set from tuning; used by a tick-time budget)


class Server:
    """Continuous-batching request loop over per-scene-type pools under a device memory budget (bytes).

    With `tick_budget_ms`, admission also keeps the predicted tick time (sum over active requests of n_envs times
    the pool's `ms_per_env`) within the budget, so the step latency of admitted requests stays bounded while later
    arrivals wait; a request is always admitted when nothing is active, so a request larger than the budget still
    runs (alone)."""

    def __init__(self, pools: Dict[str, Pool], memory_budget: int, record_results: bool = False,
                 launch: Optional[dict] = None, tick_budget_ms: Optional[float] = None):
        self.pools = pools
        self.launch = launch or {}
        self.budget = memory_budget
        self.tick_budget = tick_budget_ms
        self.tick_load = 0.0
        self.used = 0
        self.record = record_results
        self.order = sorted(pools, key=lambda k: -pools[k].M.n_dofs)  # heavier scene types first in the launch

    def cost(self, r: Request) -> int:
        return r.n_envs * self.pools[r.scene].bytes_per_env

    def tick_cost(self, r: Request) -> float:
        return r.n_envs * self.pools[r.scene].ms_per_env

    def _try_admit(self, r: Request, now: float, n_active: int) -> bool:
        p = self.pools[r.scene]
        if self.used + self.cost(r) > self.budget or r.n_envs > p.slots.capacity:
            return False
        if self.tick_budget is not None and n_active and self.tick_load + self.tick_cost(r) > self.tick_budget:
            return False
        b = p.slots.alloc(r.n_envs)
        if b is None:
            return False
        r.slot, r.admitted, r.steps_done = b, now, 0
        self.used += self.cost(r)
        self.tick_load += self.tick_cost(r)
        init_state(p.w, p.M, p.S, b, r.n_envs)
        r.control = Control(p.w, p.M, p.S, r.seed, b, r.n_envs)
        return True

    def _finish(self, r: Request, now: float):
        p = self.pools[r.scene]
        if self.record:
            sl = slice(r.slot, r.slot + r.n_envs)
            r.result = {k: p.S.st.t[k][:, sl].T.clone() for k in ("qpos", "vel")}
        r.finished = now
        p.slots.release(r.slot, r.n_envs)
        self.used -= self.cost(r)
        self.tick_load -= self.tick_cost(r)
        r.control = None

    def _groups(self, active: List[Request]) -> List[Group]:
        """One group per maximal run of adjacent active slots of a pool (groups never span free slots)."""
        groups = []
        for scene in self.order:
            p = self.pools[scene]
            ranges = sorted((r.slot, r.n_envs) for r in active if r.scene == scene)
            runs = []
            for b, n in ranges:
                if runs and runs[-1][0] + runs[-1][1] == b:
                    runs[-1][1] += n
                else:
                    runs.append([b, n])
            groups += [Group(p.S, b, n, p.envs_per_tile) for b, n in runs]
        return groups

    def run(self, requests: List[Request], max_ticks: Optional[int] = None, idle_sleep: float = 2e-4, log=None):
        """Serve `requests` (arrival times relative to the start) to completion. Returns a stats dict.
        `log`, if given, receives a progress line every 2 s."""
        pending = collections.deque(sorted(requests, key=lambda r: (r.arrival, r.rid)))
        queue: collections.deque = collections.deque()
        active: List[Request] = []
        tick_lat, env_steps, ticks = [], 0, 0
        fb, dirty = None, True
        ev = torch.cuda.Event()
        t_start = time.perf_counter()
        t_log = 0.0
        while pending or queue or active:
            now = time.perf_counter() - t_start
            if log is not None and now - t_log > 2.0:
                t_log = now
                log(f"  t={now:.1f}s ticks {ticks} active {len(active)} queued {len(queue)} pending {len(pending)}")
            while pending and pending[0].arrival <= now:
                queue.append(pending.popleft())
            while queue and self._try_admit(queue[0], now, len(active)):
                active.append(queue.popleft())
                dirty = True
            if not active:
                if queue:
                    raise RuntimeError(f"request {queue[0].rid} can never be admitted (budget or pool too small)")
                time.sleep(min(idle_sleep, max(0.0, pending[0].arrival - now)) if pending else 0)
                continue
            if dirty:
                fb = FusedBatch(self._groups(active), **self.launch)
                dirty = False
            t0 = time.perf_counter()
            for r in active:
                r.control(r.steps_done)
            fb.step(1)
            ev.record()
            ev.synchronize()
            t1 = time.perf_counter()
            tick_lat.append(t1 - t0)
            ticks += 1
            env_steps += fb.n_envs
            done = []
            for r in active:
                r.steps_done += 1
                if r.steps_done >= r.horizon:
                    done.append(r)
            if ticks % 10 == 0 or done:
                self._check_errno(active)
            for r in done:
                active.remove(r)
                self._finish(r, t1 - t_start)
                dirty = True
            if max_ticks is not None and ticks >= max_ticks:
                break
        wall = time.perf_counter() - t_start
        lat = np.array(tick_lat) * 1e3
        req_lat = np.array([r.finished - r.arrival for r in requests if r.finished is not None]) * 1e3
        return {"ticks": ticks, "env_steps": env_steps, "wall_s": wall,
                "env_steps_per_s": env_steps / wall if wall > 0 else 0.0,
                "step_ms_p50": float(np.percentile(lat, 50)) if len(lat) else None,
                "step_ms_p99": float(np.percentile(lat, 99)) if len(lat) else None,
                "request_ms_p50": float(np.percentile(req_lat, 50)) if len(req_lat) else None,
                "request_ms_p99": float(np.percentile(req_lat, 99)) if len(req_lat) else None,
                "completed": int(sum(r.finished is not None for r in requests)),
                "failed": int(sum(r.failed for r in requests))}

    FATAL = sum(bit for bit, _ in ERRNO_MESSAGES)  # the bits RigidSolver.check_errno raises on

    def _check_errno(self, active: List[Request]):
        """Mark requests whose envs hit an error the reference raises on (overflow, NaN); a server fails the request
        instead of stopping every other request."""
        for r in active:
            e = self.pools[r.scene].S.st.t["errno_env"][0, r.slot:r.slot + r.n_envs]
            if bool(((e & self.FATAL) != 0).any()):
                r.failed = True


def poisson_requests(rate: float, duration: float, seed: int = 4, scenes=("arm", "quadruped", "pile20"),
                     max_envs: int = 4096, horizon=(50, 500), start_id: int = 0,
                     max_count: Optional[int] = None) -> List[Request]:
    """Poisson arrivals over [0, duration), at most max_count of them: scene uniform, n_envs log-uniform in
    [1, max_envs], horizon uniform."""
    rng = np.random.default_rng(seed)
    out, t, i = [], 0.0, start_id
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration or (max_count is not None and len(out) >= max_count):
            return out
        n = int(round(np.exp(rng.uniform(0.0, np.log(max_envs)))))
        out.append(Request(i, scenes[rng.integers(len(scenes))], max(1, n), int(rng.integers(horizon[0], horizon[1] + 1)),
                           t, seed=int(rng.integers(1 << 31))))
        i += 1
