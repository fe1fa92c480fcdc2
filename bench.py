"""Benchmark of the batched rigid step on one MI355X.

Prints ONE JSON line. A step is one scene.step() (one substep) of every env of the workload, inputs resident in HBM.
Workloads (seeded synthetic scenes, random-init state; no datasets):
  mixed      [default] batching form: 16384 envs split (seed 3) among the 9-DoF arm, the go2 quadruped and a 20-box
             pile, dt = 0.01. From one snapshot of the contact-rich window (steps 150.. after the piles have landed)
             it times (1) the plain drop-in with the fused plan's per-type execution variants (one gsr_step per
             scene type: the like-for-like baseline), (2) the plain drop-in with each scene's default variants, and
             (3) the fused hipGraph-captured launch; `value` is (3), `speedup_vs_plain` is (3) / (1), and the fused
             run's final state is compared bitwise with (1)'s
  pile       64 free boxes per env (edges U[0.05, 0.2] m, 1000 kg/m^3, friction 1.0) falling in a 1 m column onto a
             plane, 4096 envs, dt = 0.01, plain drop-in
  arm        9-DoF Panda arm + plane, 64 envs, dt = 0.01, PD targets U(limits) resampled every 100 steps
  quadruped  go2 quadruped + plane, 32768 envs, dt = 0.005, PD targets standing + U(-0.3, 0.3) every step
  serve      serving form: Poisson rollout requests (scene uniform over arm / quadruped / 20-box pile, n_envs
             log-uniform in [1, 4096], horizon U{50..500}) served by continuous batching under a memory budget;
             capacity measured saturated, then offered load swept 20-100 % of it; p50/p99 step and request latency

Timed window: every workload first advances untimed from its seeded start to step WINDOW_START (150: the piles have
landed and settle under ~contact-rich load), snapshots, runs the warm-up steps, restores the snapshot and times
steps WINDOW_START .. WINDOW_START + steps - 1 (clipped to the workload's horizon; the whole horizon from step 0
when `--steps` covers it).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s
KERNELS = ["k_step1", "k_collide", "k_solve", "k_step2"]
HORIZON = {"pile": 500, "arm": 1000, "quadruped": 1000, "pile20": 500, "mixed": 500}
WINDOW_START = 150


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def window(name, steps):
    """(first timed step, timed steps) of a workload for --steps (0: its horizon)."""
    h = HORIZON[name]
    if steps <= 0 or steps >= h:
        return 0, steps if steps > 0 else h
    return min(WINDOW_START, h - steps), steps


def hbm_traffic(workload, kernel):
    """HBM bytes per launch of `kernel` in `workload` from the committed rocprofv3 counter summary of the same bench
    command (FETCH_SIZE and WRITE_SIZE per dispatch, each scaled by what the same counter reads for a device copy of
    known size), or None when no summary for that kernel and workload is committed."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        rec = json.load(f).get(workload)
    if not rec or rec.get("kernel") != kernel:
        return None
    return {"bytes": rec["fetch_bytes_per_launch"] + rec["write_bytes_per_launch"], "source": rec.get("source")}


def b_alg(M, n_ctrl):
    """Algorithmic HBM bytes per env-step: read+write qpos and qvel, read+write the warm start, read PD targets."""
    return 4 * (2 * M.n_qs + 4 * M.n_dofs + n_ctrl)


class Part:
    """One scene type of a workload: its model, solver, controls, and the snapshot the timed runs start from."""

    def __init__(self, w, n_envs, seed, solver_variant=None):
        import torch
        from genesis_amd.model import compile_model
        from genesis_amd.rigid_solver import RigidSolver
        from genesis_amd.workloads import Control, init_state
        self.w, self.name = w, w.name
        self.M = compile_model(w.entities, w.options)
        self.S = RigidSolver(self.M, n_envs, "cuda", solver_variant=solver_variant)
        init_state(w, self.M, self.S)
        self.ctrl = Control(w, self.M, self.S, seed)
        d = self.S.dm.desc
        self.default_variants = (int(d.solver_variant), int(d.collide_variant), int(d.dyn_variant))
        self.snap = None
        self.gen_state = None
        self.torch = torch
        self.n_reset = 0  # envs restarted after an errno bit the reference raises on

    @property
    def n(self):
        return self.S._B

    def balg(self):
        return b_alg(self.M, self.ctrl.n_ctrl)

    def save(self):
        self.snap = {k: v.clone() for k, v in self.S.st.t.items()}
        self.gen_state = self.ctrl.gen.get_state()

    def restore(self):
        for k, v in self.snap.items():
            self.S.st.t[k].copy_(v)
        self.ctrl.gen.set_state(self.gen_state)


class Runner:
    """Steps a list of parts through the workload's trajectory and times windows of it."""

    def __init__(self, parts):
        import torch
        self.parts = parts
        self.torch = torch
        self.step = 0

    def controls(self):
        for p in self.parts:
            p.ctrl(self.step)

    def plain_step(self):
        for p in self.parts:
            p.S._lib.gsr_step(p.S.dm.desc, p.S.st.desc, 1, p.S._stream())

    def advance(self, n):
        for _ in range(n):
            self.controls()
            self.plain_step()
            self.step += 1
            if self.step % 10 == 0:
                self.check()
            if self.step % 50 == 0:
                log(f"bench: advanced to step {self.step}")
        self.torch.cuda.synchronize()
        self.check()

    def check(self):
        """The errno poll of Simulator.step (every 10 steps), with the episode-termination rule instead of an
        exception: a diverged env (NaN, overflow) restarts from its initial state and is counted."""
        from genesis_amd.workloads import reset_failed_envs
        for p in self.parts:
            p.n_reset += reset_failed_envs(p.w, p.M, p.S)

    def snapshot(self):
        self.torch.cuda.synchronize()
        for p in self.parts:
            p.save()
        self.step0 = self.step

    def restore(self):
        self.torch.cuda.synchronize()
        for p in self.parts:
            p.restore()
        self.step = self.step0
        self.torch.cuda.synchronize()

    def timed(self, one_step, warmup, steps, label):
        """Warm-up steps from the snapshot, restore, then `steps` timed steps (controls, the step, an errno poll
        every 10 steps as Simulator.step does). Returns seconds."""
        torch = self.torch
        self.restore()
        for _ in range(warmup):
            self.controls()
            one_step(None)
            self.step += 1
        self.restore()
        resets0 = sum(p.n_reset for p in self.parts)
        t0 = time.perf_counter()
        for i in range(steps):
            self.controls()
            one_step(i)
            self.step += 1
            if self.step % 10 == 0:
                self.check()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        self.last_resets = sum(p.n_reset for p in self.parts) - resets0
        log(f"bench: {label}: {steps} steps in {dt:.3f} s, {self.last_resets} env(s) reset")
        return dt


def cpu_threads():
    return max(1, min(16, os.cpu_count() or 1))


def host_cpu():
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "model": model}


def oracle_rate(part, first_step, min_seconds, threads, n_envs):
    """Env-steps/s of the CPU oracle (scalar FP32 restatement of the reference step) on `n_envs` envs of the part's
    scene with `threads` host threads, over steps first_step.. of the scene's own trajectory (seeded start, the
    workload's control schedule drawn with numpy): untimed up to first_step, then timed until min_seconds pass."""
    from oracle.oracle import Oracle
    from genesis_amd.workloads import GO2_INIT_BASE, GO2_STANDING
    w, M = part.w, part.M
    O = Oracle(M, n_envs, 32)
    if w.init_qpos is not None:
        O.set("qpos", w.init_qpos(n_envs, M.n_qs))
    elif w.name == "quadruped":
        O.set("qpos", np.tile(np.concatenate([GO2_INIT_BASE, GO2_STANDING]), (n_envs, 1)))
    mode = np.zeros((n_envs, M.n_dofs), np.int32)
    rng = np.random.default_rng(0)
    if w.name == "arm":
        mode[:] = 2
    elif w.name == "quadruped":
        mode[:, 6:] = 2
    O.set("ctrl_mode", mode)

    def ctrl(step):
        if w.name == "arm" and step % 100 == 0:
            O.set("ctrl_pos", rng.uniform(M.dof_limit[:, 0], M.dof_limit[:, 1], (n_envs, M.n_dofs)))
        elif w.name == "quadruped":
            t = np.zeros((n_envs, M.n_dofs))
            t[:, 6:] = GO2_STANDING + rng.uniform(-0.3, 0.3, (n_envs, 12))
            O.set("ctrl_pos", t)

    for s in range(first_step):
        ctrl(s)
        O.run(1, n_threads=cpu_threads())
    t_total, done = 0.0, 0
    while t_total < min_seconds:
        ctrl(first_step + done)
        t0 = time.perf_counter()
        O.run(1, n_threads=threads)
        t_total += time.perf_counter() - t0
        done += 1
    return n_envs * done / t_total, done


def cpu_baseline(parts, first_step, seconds=4.0):
    """The oracle timed on the host: all cores (env-chunked threads, like the reference's GS_PARA_LEVEL=2 with
    TI_NUM_THREADS=N) and one thread (the reference's CPU default), per scene type over the same window of its
    trajectory as the GPU run; a mixed batch's rate is the split-weighted combination (time per mixed step =
    sum over scene types of envs / rate)."""
    T = cpu_threads()
    n_all, n_one = 8 * T, 8  # eight envs per thread, so thread start-up does not dominate a cheap scene's step
    per, t_all, t_one = {}, 0.0, 0.0
    for p in parts:
        r_all, _ = oracle_rate(p, first_step, seconds, T, n_all)
        r_one, _ = oracle_rate(p, first_step, seconds / 2, 1, n_one)
        per[p.name] = {"all_cores": round(r_all, 1), "one_thread": round(r_one, 1)}
        t_all += p.n / r_all
        t_one += p.n / r_one
    n = sum(p.n for p in parts)
    return {"value": round(n / t_all, 1), "unit": "env-steps/s", "cores": T, "kind": "port",
            "value_one_thread": round(n / t_one, 1), "per_scene": per, "host": host_cpu(),
            "sample": f"CPU oracle (FP32), per scene type {n_all} envs on {T} threads and {n_one} envs on 1 "
                      f"thread, steps {first_step}.. of each scene's seeded trajectory (untimed up to step "
                      f"{first_step}), timed >= {seconds:g} s (1 thread: >= {seconds / 2:g} s); a mixed batch's "
                      f"rate is the split-weighted combination"}


def plain_kernel_events(runner, K):
    """HIP events around every plain kernel of the timed steps, on the launch stream (torch's current stream)."""
    import torch
    ev = [[[(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in KERNELS]
           for _ in runner.parts] for _ in range(K)]

    def one_step(i):
        for j, p in enumerate(runner.parts):
            for k in range(4):
                if i is not None:
                    ev[i][j][k][0].record()
                p.S.launch_kernel(k)
                if i is not None:
                    ev[i][j][k][1].record()
            p.S._substep += 1

    def kernel_ms():
        out = {}
        for j, p in enumerate(runner.parts):
            for k in range(4):
                ms = sum(e[j][k][0].elapsed_time(e[j][k][1]) for e in ev) / K
                out[f"{p.name}:{p.S.kernel_name(k)}"] = round(ms, 4)
        return out

    return one_step, kernel_ms


def bench_mixed(args):
    import torch
    from genesis_amd import batch as FB
    from genesis_amd.workloads import mixed_workloads

    n_total = args.n_envs or 16384
    parts = [Part(w, n, args.seed + i) for i, (w, n) in enumerate(mixed_workloads(n_total)) if n > 0]
    runner = Runner(parts)
    first, K = window("mixed", args.steps)
    W = args.warmup
    log(f"bench: mixed {[(p.name, p.n) for p in parts]}, window steps {first}..{first + K - 1}")
    runner.advance(first)
    advance_resets = {p.name: p.n_reset for p in parts}
    runner.snapshot()

    # the fused plan: tiles and variants per scene type and the launch form, tuned on the snapshot state
    if args.plan:
        spec = json.loads(args.plan)
        plan = [FB.Group(next(p.S for p in parts if p.name == g["scene"]), 0, None, g["envs_per_tile"],
                         g["solver"], g["collide"], g.get("solve_tile")) for g in spec["groups"]]
        launch, report = spec["launch"], []
    else:
        plan, launch, report = FB.autotune([FB.Group(p.S) for p in parts])
    plan = [g.resolved() for g in plan]
    names = {id(p.S): p.name for p in parts}
    variants = {id(g.solver): (g.solver_variant, g.collide_variant) for g in plan}
    plan_json = {"groups": [{"scene": names[id(g.solver)], "solver": g.solver_variant, "collide": g.collide_variant,
                             "envs_per_tile": g.envs_per_tile, "solve_tile": g.wave_envs_per_tile} for g in plan],
                 "launch": launch}
    log("bench: plan " + json.dumps(plan_json))

    res = {}
    forms = args.forms.split(",")
    if "plain" in forms:
        for p in parts:
            p.S.set_variants(*variants[id(p.S)])
        one, kms = plain_kernel_events(runner, K)
        t = runner.timed(one, W, K, "plain (fused plan's variants)")
        res["plain"] = {"value": round(n_total * K / t, 1), "ms_per_step": round(1e3 * t / K, 4),
                        "kernel_ms": kms(), "envs_reset": runner.last_resets,
                        "variants": {p.name: ["lane", "wave"][variants[id(p.S)][0]] for p in parts}}
        final_plain = [(p.S.st.t["qpos"].clone(), p.S.st.t["vel"].clone()) for p in parts]
    if "default" in forms:
        for p in parts:
            p.S.set_variants(*p.default_variants)
        one, kms = plain_kernel_events(runner, K)
        t = runner.timed(one, W, K, "plain (default variants)")
        res["plain_default_variants"] = {"value": round(n_total * K / t, 1), "ms_per_step": round(1e3 * t / K, 4),
                                         "kernel_ms": kms(), "envs_reset": runner.last_resets,
                                         "variants": {p.name: ["lane", "wave"][p.default_variants[0]] for p in parts}}
    fb = FB.FusedBatch(plan, **launch)
    knames = fb.kernel_names()
    fb.capture(1)
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(K)]

    def fused_step(i):
        if i is not None:
            ev[i][0].record()
        fb.replay()
        if i is not None:
            ev[i][1].record()

    t_fused = runner.timed(fused_step, W, K, "fused graph")
    fused_resets = runner.last_resets
    substep_ms = sum(a.elapsed_time(b) for a, b in ev) / K
    bitwise = None
    if "plain" in forms:
        bitwise = all(torch.equal(q, p.S.st.t["qpos"]) and torch.equal(v, p.S.st.t["vel"])
                      for (q, v), p in zip(final_plain, parts))
    # per-kernel times of the fused substep: the graph's only kernel, or (several kernels) a pass that launches
    # each kernel of the same steps with events around it
    if len(knames) == 1:
        kernel_ms = {knames[0]: substep_ms}
    else:
        kev = [[(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in knames]
               for _ in range(K)]

        def fused_kernels(i):
            for k in range(len(knames)):
                if i is not None:
                    kev[i][k][0].record()
                fb.launch_kernel(k)
                if i is not None:
                    kev[i][k][1].record()

        runner.timed(fused_kernels, 0, K, "fused kernels (per-kernel events)")
        kernel_ms = {}
        for k, nm in enumerate(knames):
            kernel_ms[nm] = kernel_ms.get(nm, 0.0) + sum(e[k][0].elapsed_time(e[k][1]) for e in kev) / K
    dom = max(kernel_ms, key=kernel_ms.get)
    balg_launch = sum(p.balg() * p.n for p in parts)
    achieved = balg_launch / (kernel_ms[dom] * 1e-3) / 1e9
    value = n_total * K / t_fused
    tr = hbm_traffic("mixed", dom)
    out = {
        "metric": "env-steps/s (plain vs fused batched, 1 GPU)",
        "value": round(value, 1), "unit": "env-steps/s", "n_gpus": 1, "steps": K, "warmup": W,
        "ms_per_step": round(1e3 * t_fused / K, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded scenes and controls, random-init state)",
        "config": {"workload": "mixed batch: 9-DoF arm / go2 quadruped / 20-box pile", "n_envs": n_total,
                   "dt": 0.01, "form": "fused hipGraph", "window": [first, first + K - 1],
                   "split": {p.name: p.n for p in parts}, "plan": plan_json,
                   "grid": int(fb.info.grid), "lds_bytes": int(fb.info.lds_bytes)},
        "plain": res.get("plain"),
        "plain_default_variants": res.get("plain_default_variants"),
        "speedup_vs_plain": round(value / res["plain"]["value"], 3) if "plain" in res else None,
        "speedup_vs_plain_default_variants": (round(value / res["plain_default_variants"]["value"], 3)
                                              if "plain_default_variants" in res else None),
        "fused_bitwise_equals_plain": bitwise,
        "envs_reset": {"advance": advance_resets, "fused": fused_resets},
        "kernel_ms": {k: round(v, 4) for k, v in kernel_ms.items()},
        "fused_substep_gpu_ms": round(substep_ms, 4),
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": tr["bytes"] if tr else None, "traffic_source": tr["source"] if tr else None,
                     "b_alg_per_launch": balg_launch,
                     "b_alg_per_env_step": {p.name: p.balg() for p in parts},
                     "step_frac": balg_launch / n_total * value / 1e9 / HBM_PEAK_GBS},
        "autotune": [dict(r, scene=parts[r["group"]].name if r["group"] != "all" else "all", ms=round(r["ms"], 4))
                     for r in report],
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(parts, first)
    print(json.dumps(out))


def bench_single(args):
    import torch
    from genesis_amd import workloads as Wk
    n = args.n_envs or {"pile": 4096, "arm": 64, "quadruped": 32768}[args.workload]
    w = {"pile": lambda: Wk.pile_workload(n_envs=n), "arm": lambda: Wk.arm_workload(n_envs=n),
         "quadruped": lambda: Wk.quadruped_workload(n_envs=n)}[args.workload]()
    p = Part(w, n, args.seed, solver_variant=args.solver_variant)
    runner = Runner([p])
    first, K = window(w.name, args.steps)
    runner.advance(first)
    runner.snapshot()
    one, kms = plain_kernel_events(runner, K)
    t = runner.timed(one, args.warmup, K, "plain")
    value = n * K / t
    k_ms = kms()
    dom = max(k_ms, key=k_ms.get)
    balg = p.balg()
    achieved = balg * n / (k_ms[dom] * 1e-3) / 1e9
    kern = dom.split(":")[1]
    tr = hbm_traffic(w.name, kern)
    out = {
        "metric": "env-steps/s (plain drop-in, 1 GPU)", "value": round(value, 1), "unit": "env-steps/s",
        "n_gpus": 1, "steps": K, "warmup": args.warmup, "ms_per_step": round(1e3 * t / K, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded scene and controls, random-init state)",
        "config": {"workload": {"pile": "box pile 64 boxes/env", "arm": "9-DoF arm + plane",
                                "quadruped": "go2 quadruped + plane"}[w.name],
                   "n_envs": n, "dt": w.options.dt, "form": "plain", "window": [first, first + K - 1],
                   "n_dofs": p.M.n_dofs, "n_geoms": p.M.n_geoms,
                   "solver": ["lane-per-env", "wavefront-per-env"][p.S.dm.desc.solver_variant]},
        "kernel_ms": k_ms,
        "envs_reset": {"advance": p.n_reset - runner.last_resets, "timed": runner.last_resets},
        "roofline": {"bound": "hbm", "kernel": kern, "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": tr["bytes"] if tr else None, "traffic_source": tr["source"] if tr else None,
                     "b_alg_per_env_step": balg, "step_frac": balg * value / 1e9 / HBM_PEAK_GBS},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline([p], first)
    print(json.dumps(out))


def bench_serve(args):
    from genesis_amd.batch import FusedBatch, Group, autotune, time_group
    from genesis_amd.serving import Pool, Server, poisson_requests
    from genesis_amd.workloads import init_state

    cap_envs = args.n_envs or 8192
    pools = {s: Pool(s, cap_envs, dt=0.01) for s in ("arm", "quadruped", "pile20")}
    # execution choices per scene type, tuned once on a slice of each pool
    log("serve: tuning tiles")
    n_tune = min(512, cap_envs)
    for p in pools.values():
        init_state(p.w, p.M, p.S, 0, n_tune)
    # tuned on a contact-rich state: the slice first runs 60 substeps (the piles land)
    FusedBatch([Group(p.S, 0, n_tune) for p in pools.values()]).step(60)
    tuned, launch, _ = autotune([Group(p.S, 0, n_tune) for p in pools.values()], steps=2, warmup=1, verbose=True,
                                log=log)
    for g in tuned:
        for p in pools.values():
            if p.S is g.solver:
                p.set_execution(g.envs_per_tile, g.solver_variant, g.collide_variant)
                p.ms_per_env = time_group(g, steps=3, warmup=1, **launch) / n_tune
    log("serve: ms per env per tick " + ", ".join(f"{p.scene} {p.ms_per_env * 1e3:.3f} us" for p in pools.values()))
    budget = int(args.serve_budget_gb * 2**30)
    server = Server(pools, budget, launch=launch, tick_budget_ms=args.serve_tick_ms or None)
    sample = poisson_requests(1000.0, 100.0, seed=99, max_count=20000)
    mean_work = float(np.mean([r.n_envs * r.horizon for r in sample]))
    # capacity: a saturated queue (every request waiting at t = 0)
    sat = poisson_requests(1e9, 1.0, seed=5, max_count=args.serve_saturate)
    for r in sat:
        r.arrival = 0.0
    log(f"serve: capacity run with {len(sat)} queued requests")
    cap = server.run(sat, log=log)
    capacity = cap["env_steps_per_s"]
    log(f"serve: capacity {capacity:.0f} env-steps/s")
    sweep = []
    for load in (0.2, 0.4, 0.6, 0.8, 1.0):
        rate = load * capacity / mean_work
        reqs = poisson_requests(rate, args.serve_seconds, seed=4, start_id=0)
        st = server.run(reqs, log=log)
        st.update(load=load, offered_requests_per_s=rate, n_requests=len(reqs))
        log(f"serve: load {load:.1f}: {len(reqs)} requests, p50 {st['step_ms_p50']} ms, p99 {st['step_ms_p99']} ms")
        sweep.append({k: (round(v, 4) if isinstance(v, float) else v) for k, v in st.items()})
    at80 = next(x for x in sweep if x["load"] == 0.8)
    out = {
        "metric": "served p50/p99 step latency and env-steps/s (1 GPU)",
        "value": at80["env_steps_per_s"], "unit": "env-steps/s", "n_gpus": 1, "steps": at80["ticks"],
        "warmup": 0, "ms_per_step": at80["step_ms_p50"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded Poisson request stream, random-init state)",
        "config": {"workload": "serving arm/quadruped/20-box pile requests", "pool_envs_per_scene": cap_envs,
                   "memory_budget_gb": args.serve_budget_gb, "tick_budget_ms": args.serve_tick_ms or None,
                   "dt": 0.01, "form": "serving",
                   "tiles": {p.scene: {"envs_per_tile": p.envs_per_tile, "solver": int(p.S.dm.desc.solver_variant),
                                       "collide": int(p.S.dm.desc.collide_variant)}
                             for p in pools.values()}},
        "step_ms_p50": at80["step_ms_p50"], "step_ms_p99": at80["step_ms_p99"],
        "request_ms_p50": at80["request_ms_p50"], "request_ms_p99": at80["request_ms_p99"],
        "capacity_env_steps_per_s": round(capacity, 1), "mean_request_env_steps": round(mean_work, 1),
        "sweep": sweep,
    }
    print(json.dumps(out))


def calibrate_copy(mb):
    """A plain device copy of known size, profiled beside the step with the same HBM counters (FETCH_SIZE,
    WRITE_SIZE) to turn their readings into bytes: bytes per KiB reported = (mb MiB) / (counter reading)."""
    import torch
    n = int(mb * (1 << 20)) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "launches": 3}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100, help="timed steps (0: the workload's whole horizon)")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="mixed", choices=["mixed", "pile", "arm", "quadruped", "serve"])
    ap.add_argument("--forms", default="plain,default,fused",
                    help="mixed: the forms to time besides the fused graph (plain = fused plan's variants, "
                         "default = each scene's default variants)")
    ap.add_argument("--plan", default="", help="mixed: a fused plan (JSON as printed in config.plan) instead of "
                                               "autotuning")
    ap.add_argument("--serve-seconds", type=float, default=20.0, help="arrival window per offered-load level")
    ap.add_argument("--serve-budget-gb", type=float, default=12.0, help="device memory budget for admission")
    ap.add_argument("--serve-saturate", type=int, default=24, help="requests queued at t=0 to measure capacity")
    ap.add_argument("--serve-tick-ms", type=float, default=2.0,
                    help="admission keeps the predicted tick time within this (0: memory budget only)")
    ap.add_argument("--n-envs", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--solver-variant", type=int, default=None, help="0 lane/env, 1 wavefront/env (default: by model)")
    ap.add_argument("--calibrate-copy", type=float, default=0.0, help="only run a device copy of this many MiB")
    args = ap.parse_args()
    if args.calibrate_copy > 0:
        return calibrate_copy(args.calibrate_copy)
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1 (envs are independent: "
                                   "a multi-GPU run would shard the env axis, one process per GPU)"}))
        sys.exit(2)
    if args.workload == "mixed":
        return bench_mixed(args)
    if args.workload == "serve":
        return bench_serve(args)
    return bench_single(args)


if __name__ == "__main__":
    main()
