"""Benchmark of the batched rigid step on one MI355X (plain drop-in form).

Prints ONE JSON line. A step is one scene.step() (one substep) of every env of the workload, inputs resident in HBM.
Workloads (seeded synthetic scenes, random-init state; no datasets):
  pile       64 free boxes per env (edges U[0.05, 0.2] m, 1000 kg/m^3, friction 1.0) falling in a 1 m column onto a
             plane, 4096 envs, dt = 0.01, timed over its 500 steps           [default: the headline configuration]
  arm        9-DoF Panda arm + plane, 64 envs, dt = 0.01, PD targets U(limits) resampled every 100 steps
  quadruped  go2 quadruped + plane, 32768 envs, dt = 0.005, PD targets standing + U(-0.3, 0.3) every step
  mixed      batching form: 16384 envs split (seed 3) among arm / quadruped / 20-box pile, dt = 0.01; times the plain
             drop-in (one gsr_step per scene type) and the fused hipGraph-captured launch from the same state
  serve      serving form: Poisson rollout requests (scene uniform over arm / quadruped / 20-box pile, n_envs
             log-uniform in [1, 4096], horizon U{50..500}) served by continuous batching under a memory budget;
             capacity measured saturated, then offered load swept 20-100 % of it; p50/p99 step and request latency
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
HORIZON = {"pile": 500, "arm": 1000, "quadruped": 1000}  # steps of each configuration


def profiled_traffic(workload, kernel):
    """HBM bytes per launch of `kernel` on `workload` from the committed rocprofv3 counter summary (FETCH_SIZE and
    WRITE_SIZE per dispatch, each scaled by what the same counter reads for a device copy of known size), or None."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        rec = json.load(f).get(workload)
    if not rec or rec.get("kernel") != kernel:
        return None
    return rec["fetch_bytes_per_launch"] + rec["write_bytes_per_launch"]


def b_alg(M, n_ctrl):
    """Algorithmic HBM bytes per env-step: read+write qpos and qvel, read+write the warm start, read PD targets."""
    return 4 * (2 * M.n_qs + 4 * M.n_dofs + n_ctrl)


def make(workload, n_envs):
    from genesis_amd import workloads as W
    from genesis_amd.model import compile_model
    if workload == "pile":
        w = W.pile_workload(n_envs=n_envs or 4096)
    elif workload == "arm":
        w = W.arm_workload(n_envs=n_envs or 64)
    elif workload == "quadruped":
        w = W.quadruped_workload(n_envs=n_envs or 32768)
    else:
        raise SystemExit(f"unknown workload {workload}")
    return w, compile_model(w.entities, w.options)


def cpu_baseline(w, M, steps, seconds=30.0, min_seconds=10.0):
    """The CPU oracle (scalar restatement of the reference step, FP32) on the host cores, one env per thread, over
    the same trajectory as the timed GPU run: batches of `threads` envs (the workload's envs 0, 1, ... in order) each
    run from their seeded initial state for `steps` steps, until at least `min_seconds` of CPU time are sampled; a
    batch still running after `seconds` stops early."""
    from oracle.oracle import Oracle
    threads = max(1, min(16, os.cpu_count() or 1))
    n = threads
    env_steps, t_total, batches, done = 0, 0.0, 0, 0
    while t_total < min_seconds:
        O = Oracle(M, n, 32)
        if w.init_qpos is not None:
            O.set("qpos", w.init_qpos((batches + 1) * n, M.n_qs)[batches * n:])
        ctrl_mode = np.full((n, M.n_dofs), 2 if w.name != "pile" else 0, np.int32)
        if w.name == "quadruped":
            ctrl_mode[:, :6] = 0
        O.set("ctrl_mode", ctrl_mode)
        done, t0 = 0, time.time()
        while done < steps and t_total + time.time() - t0 < seconds:
            O.run(1, n_threads=threads)
            done += 1
        t_total += time.time() - t0
        env_steps += n * done
        batches += 1
        if done < steps:
            break
    return {"value": env_steps / t_total, "unit": "env-steps/s", "cores": threads, "kind": "port",
            "sample": f"{batches} batch(es) of {n} envs x steps 0-{steps - 1} of the same workload on the CPU oracle"
                      + (f" (the last batch to step {done - 1})" if done < steps else "") + f", {t_total:.1f} s"}


def bench_mixed(args):
    import torch
    from genesis_amd import batch as FB
    from genesis_amd.model import compile_model
    from genesis_amd.rigid_solver import RigidSolver
    from genesis_amd.workloads import Control, init_state, mixed_workloads

    n_total = args.n_envs or 16384
    parts = []
    for i, (w, n) in enumerate(mixed_workloads(n_total)):
        if n == 0:
            continue
        M = compile_model(w.entities, w.options)
        S = RigidSolver(M, n, "cuda")
        init_state(w, M, S)
        parts.append((w, M, S, Control(w, M, S, args.seed + i)))
    step = 0

    def controls():
        for w, M, S, c in parts:
            c(step)

    for _ in range(args.warmup):
        controls()
        for w, M, S, c in parts:
            S.substep()
        step += 1
    torch.cuda.synchronize()
    snap = [FB._snapshot(S) for _, _, S, _ in parts]
    step0 = step

    def timed(run):
        nonlocal step
        step = step0
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for it in range(args.steps):
            controls()
            run()
            step += 1
            if step % 10 == 0:
                for _, _, S, _ in parts:
                    S.check_errno()
        torch.cuda.synchronize()
        return time.perf_counter() - t0

    def plain():
        for _, _, S, _ in parts:
            S._lib.gsr_step(S.dm.desc, S.st.desc, 1, S._stream())

    t_plain = timed(plain)
    # tile and launch choices are tuned on the state the timed window reaches halfway (boxes landing), then the
    # fused run starts from the same state as the plain run
    for (_, _, S, _), sn in zip(parts, snap):
        FB._restore(S, sn)
    step = step0
    for _ in range(args.steps // 2):
        controls()
        plain()
        step += 1
    tuned, launch, report = FB.autotune([FB.Group(S) for _, _, S, _ in parts])
    for (_, _, S, _), sn in zip(parts, snap):
        FB._restore(S, sn)
    fb = FB.FusedBatch(tuned, **launch)
    fb.capture(1)
    t_fused = timed(fb.replay)
    B = sum(S._B for _, _, S, _ in parts)
    v_plain, v_fused = B * args.steps / t_plain, B * args.steps / t_fused
    names = {id(S): w.name for w, _, S, _ in parts}
    out = {
        "metric": "env-steps/s (fused batched vs plain drop-in, 1 GPU)",
        "value": round(v_fused, 1), "unit": "env-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * t_fused / args.steps, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded scenes, random-init state)",
        "config": {"workload": "mixed batch arm/quadruped/20-box pile", "n_envs": B, "dt": 0.01, "form": "fused",
                   "split": {w.name: S._B for w, _, S, _ in parts},
                   "tiles": [{"scene": names[id(g.solver)], "solver": ["lane-per-env", "wavefront-per-env"][g.solver_variant],
                              "envs_per_tile": g.envs_per_tile} for g in tuned],
                   "grid": fb.info.grid, "lds_bytes": fb.info.lds_bytes, "launch": launch},
        "plain": {"value": round(v_plain, 1), "ms_per_step": round(1e3 * t_plain / args.steps, 4)},
        "speedup_vs_plain": round(v_fused / v_plain, 3),
        "autotune": [dict(r, scene=parts[r["group"]][0].name if r["group"] != "all" else "all", ms=round(r["ms"], 4))
                     for r in report],
    }
    print(json.dumps(out))


def bench_serve(args):
    import numpy as np
    import torch
    from genesis_amd.batch import FusedBatch, Group, autotune, time_group
    from genesis_amd.serving import Pool, Server, poisson_requests
    from genesis_amd.workloads import init_state

    cap_envs = args.n_envs or 8192
    pools = {s: Pool(s, cap_envs, dt=0.01) for s in ("arm", "quadruped", "pile20")}
    # execution choices per scene type, tuned once on a 1024-env slice of each pool
    log = lambda *a: print(*a, file=sys.stderr, flush=True)  # noqa: E731
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
                p.envs_per_tile = g.envs_per_tile
                p.S.dm.desc.solver_variant = g.solver_variant
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
                   "tiles": {p.scene: {"envs_per_tile": p.envs_per_tile, "solver": int(p.S.dm.desc.solver_variant)}
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
    ap.add_argument("--steps", type=int, default=0, help="timed steps (default: the workload's horizon)")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="pile", choices=["pile", "arm", "quadruped", "mixed", "serve"])
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
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)

    if args.workload == "mixed":
        args.steps = args.steps or 100
        return bench_mixed(args)
    if args.workload == "serve":
        return bench_serve(args)

    import torch
    from genesis_amd.rigid_solver import RigidSolver
    from genesis_amd.workloads import Control, init_state

    w, M = make(args.workload, args.n_envs)
    # the timed steps are the workload's own trajectory from its seeded initial state (the pile: free fall and
    # settling); warm-up steps run first on the same state, which is then put back
    args.steps = args.steps or HORIZON[w.name]
    S = RigidSolver(M, w.n_envs if not args.n_envs else args.n_envs, "cuda", solver_variant=args.solver_variant)
    init_state(w, M, S)
    snap = {k: v.clone() for k, v in S.st.t.items()}
    ctrl = Control(w, M, S, args.seed)
    for step in range(args.warmup):
        ctrl(step)
        S.substep()
    torch.cuda.synchronize()
    for k, v in snap.items():
        S.st.t[k].copy_(v)
    del snap
    ctrl = Control(w, M, S, args.seed)
    step = 0
    torch.cuda.synchronize()

    # every kernel launch of the timed region is bracketed by events on the launch stream (torch's current stream)
    ev = [[(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in KERNELS]
          for _ in range(args.steps)]
    t0 = time.perf_counter()
    for it in range(args.steps):
        ctrl(step)
        for k in range(4):
            ev[it][k][0].record()
            S.launch_kernel(k)
            ev[it][k][1].record()
        S._substep += 1
        if S._substep % 10 == 0:
            S.check_errno()
        step += 1
        if step % 100 == 0:
            print(f"bench: {step} / {args.steps} steps, {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    B = S._B
    value = B * args.steps / elapsed
    k_avg = np.array([sum(e[k][0].elapsed_time(e[k][1]) for e in ev) for k in range(4)]) / args.steps
    dom = int(np.argmax(k_avg))
    balg = b_alg(M, ctrl.n_ctrl)
    achieved = balg * B / (k_avg[dom] * 1e-3) / 1e9
    out = {
        "metric": "env-steps/s (plain drop-in, 1 GPU)",
        "value": round(value, 1),
        "unit": "env-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded scene, random-init state)",
        "config": {"workload": {"pile": "box pile 64 boxes/env", "arm": "9-DoF arm + plane",
                                "quadruped": "go2 quadruped + plane"}[w.name],
                   "n_envs": B, "dt": w.options.dt, "form": "plain", "n_dofs": M.n_dofs, "n_geoms": M.n_geoms,
                   "solver": ["lane-per-env", "wavefront-per-env"][S.dm.desc.solver_variant]},
        "kernel_ms": {k: round(float(v), 4) for k, v in zip(KERNELS, k_avg)},
        "roofline": {"bound": "hbm", "kernel": KERNELS[dom], "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": profiled_traffic(w.name, S.kernel_name(dom)),
                     "step_frac": balg * value / 1e9 / HBM_PEAK_GBS, "b_alg_per_env_step": balg},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(w, M, args.steps)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
