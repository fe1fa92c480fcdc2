"""Serving form: continuous batching of rollout requests on one GPU with a device-side scheduler (no reference
counterpart).

A request asks for `horizon` steps of `n_envs` envs of one scene type. The server keeps one resident pool per scene
type in HBM (a RigidSolver whose envs are cut into tiles) and hands requests to the scheduler on the device
(`gsr_serve_*`, csrc/rigid_serve.hip): each tick, captured once in a hipGraph and replayed, admits arrived requests
in arrival order while request slots, free tiles and the memory and tick budgets allow, resets the admitted envs,
applies every active request's controls, runs one fused substep over the active tiles, advances every active
request and retires those at their horizon. The host only appends arrivals to a ring in pinned host memory, replays
the tick graph with a bounded number of ticks in flight, and reads completions and per-tick GPU timestamps from the
same memory: it never waits for the device inside the loop.

Per-env arithmetic does not depend on the tile, the pool size or the other requests in the launch, and a request's
controls come from a hash of (its seed, request-local env, request-local step, dof), so a request's result is the
same alone and in any mix.
"""
from __future__ import annotations

import ctypes
import time
from dataclasses import dataclass
from typing import Dict, List, Optional

import numpy as np
import torch

from . import _lib
from .model import compile_model
from .rigid_solver import RigidSolver
from .workloads import GO2_INIT_BASE, GO2_STANDING, workload_by_name

RING = 4096
MAX_REQUESTS = 1024


@dataclass
class Request:
    rid: int
    scene: str
    n_envs: int
    horizon: int
    arrival: float = 0.0  # seconds after the start of the run
    seed: int = 0
    # filled by the server
    finished: Optional[float] = None
    done_tick: Optional[int] = None
    failed: bool = False
    result: Optional[Dict[str, torch.Tensor]] = None


def _initial_qpos(w, M, n):
    """(n, n_qs) initial poses of request-local envs 0..n-1: the workload's seeded poses, the quadruped's standing
    pose, or qpos0."""
    if w.init_qpos is not None:
        return np.asarray(w.init_qpos(n, M.n_qs), np.float64)
    if w.name == "quadruped":
        return np.tile(np.concatenate([GO2_INIT_BASE, GO2_STANDING]), (n, 1))
    return np.tile(M.qpos0.astype(np.float64), (n, 1))


class Pool:
    """Resident state of one scene type: `capacity` env slots cut into tiles, the scene's initial poses and control
    program, and the execution choices for it."""

    def __init__(self, scene: str, capacity: int, dt: float, device="cuda", solver_variant=None,
                 envs_per_tile=None):
        self.scene = scene
        self.w = workload_by_name(scene, capacity, dt=dt)
        self.M = compile_model(self.w.entities, self.w.options)
        self.S = RigidSolver(self.M, capacity, device, solver_variant=solver_variant)
        self.capacity = capacity
        self.device = self.S.device
        v = int(self.S.dm.desc.solver_variant)
        self.envs_per_tile = envs_per_tile or (32 if v == 0 else 1)
        self.bytes_per_env = sum(t.numel() * t.element_size() for t in self.S.st.t.values()) // capacity
        self.ms_per_env = 0.0  # GPU time one env adds to a tick (set from tuning; used by the tick budget)
        q = _initial_qpos(self.w, self.M, capacity)
        self.init_qpos = torch.as_tensor(np.ascontiguousarray(q.T), dtype=torch.float32, device=self.device)
        # control program (workloads.Control's schedule): arm targets U(limits) every 100 steps, quadruped targets
        # standing pose + U(-0.3, 0.3) every step, piles uncontrolled
        self.ctrl = dict(kind=0, period=1, begin=0, end=0, amp=0.0)
        self._ctrl_tensors = {}
        if scene == "arm":
            D = self.M.n_dofs
            self.ctrl = dict(kind=1, period=100, begin=0, end=D, amp=0.0)
            self._ctrl_tensors = {"lo": torch.as_tensor(self.M.dof_limit[:, 0], dtype=torch.float32, device=self.device),
                                  "hi": torch.as_tensor(self.M.dof_limit[:, 1], dtype=torch.float32, device=self.device)}
        elif scene == "quadruped":
            self.ctrl = dict(kind=2, period=1, begin=6, end=18, amp=0.3)
            self._ctrl_tensors = {"base": torch.as_tensor(GO2_STANDING, dtype=torch.float32, device=self.device)}

    @property
    def n_tiles(self):
        return self.capacity // self.envs_per_tile

    @property
    def slots(self):
        return self.n_tiles * self.envs_per_tile

    def set_execution(self, envs_per_tile=None, solver_variant=None, collide_variant=None):
        if envs_per_tile is not None:
            self.envs_per_tile = int(envs_per_tile)
        self.S.set_variants(solver_variant, collide_variant)

    def descriptor(self) -> _lib.GsrServePool:
        p = _lib.GsrServePool()
        p.model = self.S.dm.desc
        p.state = self.S.st.desc
        p.envs_per_tile, p.n_tiles = self.envs_per_tile, self.n_tiles
        c = self.ctrl
        p.ctrl_kind, p.ctrl_period, p.ctrl_dof_begin, p.ctrl_dof_end, p.ctrl_amp = (c["kind"], c["period"], c["begin"],
                                                                                   c["end"], c["amp"])
        for k in ("lo", "hi", "base"):
            if k in self._ctrl_tensors:
                setattr(p, f"ctrl_{k}", self._ctrl_tensors[k].data_ptr())
        p.init_qpos = self.init_qpos.data_ptr()
        p.bytes_per_env = float(self.bytes_per_env)
        p.cost_per_env = float(self.ms_per_env)
        return p


class Server:
    """Continuous-batching rollout server over per-scene-type pools, scheduled on the device.

    memory_budget (bytes) bounds the sum over admitted requests of n_envs x the pool's per-env footprint;
    tick_budget_ms bounds the predicted tick time (sum over admitted requests of n_envs x the pool's ms_per_env),
    so the step latency of admitted requests stays bounded while later arrivals wait; a request is always admitted
    when nothing is active, so a request larger than the budget still runs (alone). `inflight` ticks may be queued on
    the GPU ahead of the host."""

    def __init__(self, pools: Dict[str, Pool], memory_budget: int, launch: Optional[dict] = None,
                 tick_budget_ms: Optional[float] = None, inflight: int = 3):
        self.lib = _lib.load()
        self.pools = pools
        self.order = sorted(pools, key=lambda k: -pools[k].M.n_dofs)  # heavier scene types first in the launch
        self.pool_index = {k: i for i, k in enumerate(self.order)}
        self.inflight = inflight
        self.device = pools[self.order[0]].device
        launch = dict(launch or {})
        self.descs = (_lib.GsrServePool * len(self.order))(*[pools[k].descriptor() for k in self.order])
        n_tiles = sum(pools[k].n_tiles for k in self.order)
        self.n_tiles = n_tiles
        self.dev_state = torch.zeros(int(self.lib.gsr_serve_state_bytes(n_tiles)), dtype=torch.uint8,
                                     device=self.device)
        self.table = torch.zeros(int(self.lib.gsr_serve_table_bytes(n_tiles)), dtype=torch.uint8, device=self.device)
        hp = ctypes.POINTER(_lib.GsrServeHost)()
        _lib.check(self.lib.gsr_serve_host_alloc(ctypes.byref(hp)), "gsr_serve_host_alloc")
        self._host_ptr = hp
        self.host = hp.contents
        self.info = _lib.GsrBatchInfo()
        self.info.persistent = int(bool(launch.get("persistent", False)))
        self.info.waves_per_simd = int(launch.get("waves_per_simd", 4))
        self.info.gjk_inline = int(bool(launch.get("gjk_inline", False)))
        self.info.phased = int(bool(launch.get("phased", False)))
        torch.cuda.synchronize(self.device)
        stream = torch.cuda.current_stream(self.device)
        _lib.check(self.lib.gsr_serve_init(self.descs, len(self.order), float(memory_budget),
                                           float(tick_budget_ms) if tick_budget_ms else 1e300,
                                           ctypes.c_void_p(self.dev_state.data_ptr()), self.dev_state.numel(),
                                           ctypes.c_void_p(self.table.data_ptr()), self.table.numel(),
                                           ctypes.byref(self.info), ctypes.c_void_p(stream.cuda_stream)),
                   "gsr_serve_init")
        self._side = torch.cuda.Stream(self.device)  # default priority
        self._side.wait_stream(stream)
        exe = ctypes.c_void_p()
        _lib.check(self.lib.gsr_serve_graph_capture(ctypes.c_void_p(self.dev_state.data_ptr()),
                                                    ctypes.c_void_p(self.table.data_ptr()), ctypes.byref(self.info),
                                                    hp, ctypes.c_void_p(self._side.cuda_stream), ctypes.byref(exe)),
                   "gsr_serve_graph_capture")
        self._graph = exe
        self.clock_khz = float(self.lib.gsr_serve_clock_khz())
        self._tile_base = {}
        t = 0
        for k in self.order:
            self._tile_base[k] = t
            t += pools[k].n_tiles

    def close(self):
        if self._graph is not None:
            self.lib.gsr_graph_destroy(self._graph)
            self._graph = None
        if self._host_ptr:
            torch.cuda.synchronize(self.device)
            self.lib.gsr_serve_host_free(self._host_ptr)
            self._host_ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _launch_tick(self):
        _lib.check(self.lib.gsr_graph_launch(self._graph, ctypes.c_void_p(self._side.cuda_stream)),
                   "gsr_graph_launch")

    def _tiles(self):
        """Host copy of the scheduler's tile arrays (req, first, n, init), valid between ticks."""
        torch.cuda.synchronize(self.device)
        off = self.dev_state.numel() - 4 * 4 * self.n_tiles
        a = self.dev_state[off:].view(torch.int32).cpu().numpy().reshape(4, self.n_tiles)
        return a

    def _result(self, r: Request, slot: int, tiles) -> Dict[str, torch.Tensor]:
        """Final qpos / vel of a retired request (its tiles remember the slot until they are taken again)."""
        p = self.pools[r.scene]
        t0 = self._tile_base[r.scene]
        req, first = tiles[0, t0:t0 + p.n_tiles], tiles[1, t0:t0 + p.n_tiles]
        mine = np.nonzero(req == -2 - slot)[0]
        mine = mine[np.argsort(first[mine])]
        envs = np.concatenate([np.arange(i * p.envs_per_tile, i * p.envs_per_tile + p.envs_per_tile) for i in mine])
        envs = torch.as_tensor(envs[:r.n_envs], device=self.device)
        return {k: p.S.st.t[k][:, envs].T.clone() for k in ("qpos", "vel")}

    def run(self, requests: List[Request], record: bool = False, log=None, idle_sleep: float = 2e-4):
        """Serve `requests` (arrival times in seconds relative to the start) to completion; returns a stats dict.
        record=True keeps every request's final qpos / vel (the host then waits for each tick, so the tiles of a
        retired request are read before they are reused). `log` receives a progress line every 2 s."""
        H = self.host
        reqs = sorted(requests, key=lambda r: (r.arrival, r.rid))
        by_rid = {r.rid: r for r in reqs}
        if len(by_rid) != len(reqs):
            raise ValueError("request ids must be unique")
        for r in reqs:
            if r.scene not in self.pools:
                raise ValueError(f"request {r.rid}: no pool for scene {r.scene}")
            if r.n_envs < 1 or r.n_envs > self.pools[r.scene].slots:
                raise ValueError(f"request {r.rid}: {r.n_envs} envs do not fit the {r.scene} pool "
                                 f"({self.pools[r.scene].slots} slots)")
            if r.horizon < 1:
                raise ValueError(f"request {r.rid}: horizon must be >= 1")
        inflight = 1 if record else self.inflight
        done0, tick0, steps0 = H.done_count, H.ticks_done, H.env_steps
        pushed, seen, launched = 0, done0, tick0
        n = len(reqs)
        completed = 0
        lat, drained = [], tick0
        ring_t0, ring_t1 = np.ctypeslib.as_array(H.tick_t0), np.ctypeslib.as_array(H.tick_t1)
        ring_envs = np.ctypeslib.as_array(H.tick_envs)

        def drain(upto):
            # GPU time of the finished ticks that stepped envs, before the timestamp ring wraps
            nonlocal drained
            if upto > drained:
                idx = np.arange(drained, upto) % RING
                dt = (ring_t1[idx].astype(np.int64) - ring_t0[idx].astype(np.int64)) / self.clock_khz
                lat.append(dt[ring_envs[idx] > 0])
                drained = upto

        t_start = time.perf_counter()
        t_log = 0.0
        while completed < n:
            if H.ticks_done - drained >= 256:
                drain(H.ticks_done)
            now = time.perf_counter() - t_start
            if log is not None and now - t_log > 2.0:
                t_log = now
                log(f"  t={now:.1f}s ticks {H.ticks_done - tick0} pushed {pushed} done {completed}")
            # arrivals: the entry first, then the head index (x86 keeps the two stores in order)
            while pushed < n and reqs[pushed].arrival <= now and H.arrivals_head - H.arrivals_tail < RING:
                r = reqs[pushed]
                e = H.arrivals[H.arrivals_head % RING]
                e.rid, e.pool, e.n_envs, e.horizon, e.seed = (r.rid, self.pool_index[r.scene], r.n_envs, r.horizon,
                                                              r.seed & 0xFFFFFFFF)
                H.arrivals_head += 1
                pushed += 1
            # completions
            dc = H.done_count
            if dc != seen:
                tiles = self._tiles() if record else None
                t_now = time.perf_counter() - t_start
                while seen < dc:
                    d = H.done[seen % RING]
                    r = by_rid[d.rid]
                    r.finished, r.done_tick, r.failed = t_now, int(d.tick), bool(d.failed)
                    if record:
                        r.result = self._result(r, int(d.slot), tiles)
                    seen += 1
                    completed += 1
            outstanding = pushed - (seen - done0)
            if outstanding > 0 and launched - H.ticks_done < inflight:
                self._launch_tick()
                launched += 1
                if record:
                    torch.cuda.synchronize(self.device)
            elif outstanding == 0 and pushed < n:
                time.sleep(min(idle_sleep, max(0.0, reqs[pushed].arrival - now)))
        torch.cuda.synchronize(self.device)
        wall = time.perf_counter() - t_start
        ticks = H.ticks_done - tick0
        drain(H.ticks_done)
        lat = np.concatenate(lat) if lat else np.zeros(0)  # ms of GPU time per tick that stepped envs
        env_steps = int(H.env_steps - steps0)
        req_lat = np.array([r.finished - r.arrival for r in reqs]) * 1e3
        return {"ticks": int(ticks), "env_steps": env_steps, "wall_s": wall,
                "env_steps_per_s": env_steps / wall if wall > 0 else 0.0,
                "step_ms_p50": float(np.percentile(lat, 50)) if len(lat) else None,
                "step_ms_p99": float(np.percentile(lat, 99)) if len(lat) else None,
                "request_ms_p50": float(np.percentile(req_lat, 50)) if len(req_lat) else None,
                "request_ms_p99": float(np.percentile(req_lat, 99)) if len(req_lat) else None,
                "completed": completed, "failed": int(sum(r.failed for r in reqs))}


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
