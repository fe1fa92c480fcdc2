"""Benchmark of the batched rigid step on one MI355X (plain drop-in form).

Prints ONE JSON line. A step is one scene.step() (one substep) of every env of the workload, inputs resident in HBM.
Workloads (seeded synthetic scenes, random-init state; no datasets):
  pile       64 free boxes per env (edges U[0.05, 0.2] m, 1000 kg/m^3, friction 1.0) falling in a 1 m column onto a
             plane, 4096 envs, dt = 0.01                                    [default: the headline configuration]
  arm        9-DoF Panda arm + plane, 64 envs, dt = 0.01, PD targets U(limits) resampled every 100 steps
  quadruped  go2 quadruped + plane, 32768 envs, dt = 0.005, PD targets standing + U(-0.3, 0.3) every step
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


class Control:
    """Per-step PD targets of the workload, generated on the device from a seeded generator."""

    def __init__(self, w, M, S, seed):
        import torch
        self.w, self.M, self.S = w, M, S
        self.gen = torch.Generator(device=S.device).manual_seed(seed)
        self.n_ctrl = 0
        if w.name == "arm":
            self.lo = torch.as_tensor(M.dof_limit[:, 0], device=S.device)
            self.hi = torch.as_tensor(M.dof_limit[:, 1], device=S.device)
            self.n_ctrl = M.n_dofs
        elif w.name == "quadruped":
            from genesis_amd.workloads import GO2_STANDING
            self.stand = torch.as_tensor(GO2_STANDING, dtype=torch.float32, device=S.device)
            self.dofs = torch.arange(6, 18, device=S.device)
            self.n_ctrl = 12

    def __call__(self, step):
        import torch
        B = self.S._B
        if self.w.name == "arm" and step % 100 == 0:
            u = torch.rand((B, self.M.n_dofs), generator=self.gen, device=self.S.device)
            self.S.control_dofs_position(self.lo + (self.hi - self.lo) * u)
        elif self.w.name == "quadruped":
            u = torch.rand((B, 12), generator=self.gen, device=self.S.device)
            self.S.control_dofs_position(self.stand + (u * 0.6 - 0.3), dofs_idx=self.dofs)


def init_state(w, M, S):
    import torch
    from genesis_amd.workloads import GO2_INIT_BASE, GO2_STANDING
    if w.name == "pile":
        S.set_qpos(torch.as_tensor(w.init_qpos(S._B, M.n_qs), dtype=torch.float32))
    elif w.name == "quadruped":
        q = np.tile(np.concatenate([GO2_INIT_BASE, GO2_STANDING]), (S._B, 1))
        S.set_qpos(torch.as_tensor(q, dtype=torch.float32))


def cpu_baseline(w, M, seconds=15.0):
    """The CPU oracle (scalar restatement of the reference step, FP32) on the host cores, env-chunked threads."""
    from oracle.oracle import Oracle
    threads = max(1, min(16, os.cpu_count() or 1))
    n = max(threads, 16)
    O = Oracle(M, n, 32)
    if w.init_qpos is not None:
        O.set("qpos", w.init_qpos(n, M.n_qs))
    ctrl_mode = np.full((n, M.n_dofs), 2 if w.name != "pile" else 0, np.int32)
    if w.name == "quadruped":
        ctrl_mode[:, :6] = 0
    O.set("ctrl_mode", ctrl_mode)
    steps, t0 = 0, time.time()
    while time.time() - t0 < seconds:
        O.run(1, n_threads=threads)
        steps += 1
    dt = time.time() - t0
    return {"value": n * steps / dt, "unit": "env-steps/s", "cores": threads, "kind": "port",
            "sample": f"{n} envs x {steps} steps of the same workload from step 0 on the CPU oracle, {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="pile", choices=["pile", "arm", "quadruped"])
    ap.add_argument("--n-envs", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--solver-variant", type=int, default=None, help="0 lane/env, 1 wavefront/env (default: by model)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)

    import torch
    from genesis_amd.rigid_solver import RigidSolver

    w, M = make(args.workload, args.n_envs)
    S = RigidSolver(M, w.n_envs if not args.n_envs else args.n_envs, "cuda", solver_variant=args.solver_variant)
    init_state(w, M, S)
    ctrl = Control(w, M, S, args.seed)
    step = 0
    for _ in range(args.warmup):
        ctrl(step)
        S.substep()
        step += 1
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
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "step_frac": balg * value / 1e9 / HBM_PEAK_GBS, "b_alg_per_env_step": balg},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(w, M)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
