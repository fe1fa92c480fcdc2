"""Copy the persistent state between the HIP solver (env-minor torch tensors) and the CPU oracle (env-major)."""
import numpy as np
import torch

MAXB = 1 << 30


def gpu_fields(solver, envs=None):
    t = solver.st.t
    sel = slice(None) if envs is None else envs

    def g(k):
        return t[k][:, sel].T.detach().cpu().numpy()

    sort_ig = g("sort_ig")
    return dict(
        qpos=g("qpos").astype(np.float64), vel=g("vel").astype(np.float64), qacc_ws=g("qacc_ws").astype(np.float64),
        is_warmstart=g("is_warmstart")[:, 0], ctrl_mode=g("ctrl_mode"), ctrl_pos=g("ctrl_pos").astype(np.float64),
        ctrl_vel=g("ctrl_vel").astype(np.float64), ctrl_force=g("ctrl_force").astype(np.float64),
        cache_normal=g("cache_normal").astype(np.float64), sort_value=g("sort_value").astype(np.float64),
        sort_ig=(sort_ig & (MAXB - 1)).astype(np.int32), sort_ismax=((sort_ig & MAXB) != 0).astype(np.int32),
        first_time=g("first_time")[:, 0], mass_shift=g("mass_shift").astype(np.float64),
        i_pos_shift=g("i_pos_shift").astype(np.float64), friction_ratio=g("friction_ratio").astype(np.float64),
        eq_dyn_n=g("eq_dyn_n")[:, 0], eq_dyn_obj=g("eq_dyn_obj"), eq_dyn_data=g("eq_dyn_data").astype(np.float64),
    )


def gpu_to_oracle(solver, orc, envs=None):
    f = gpu_fields(solver, envs)
    idx = list(range(orc.B)) if envs is None else list(range(len(range(solver.st.B)[envs])))
    for k, v in f.items():
        orc.set(k, v, idx)
    orc.set("errno", np.zeros(orc.B, np.int32))


def oracle_to_gpu(orc, solver):
    t = solver.st.t
    dev = t["qpos"].device
    for k in ("qpos", "vel", "qacc_ws", "ctrl_pos", "ctrl_vel", "ctrl_force", "sort_value"):
        t[k][:] = torch.as_tensor(orc.get(k).T, dtype=torch.float32, device=dev)
    t["ctrl_mode"][:] = torch.as_tensor(orc.get("ctrl_mode").T, dtype=torch.int32, device=dev)
    ig, im = orc.get("sort_ig"), orc.get("sort_ismax")
    t["sort_ig"][:] = torch.as_tensor((ig | (im * MAXB)).T, dtype=torch.int32, device=dev)
    t["is_warmstart"][0] = torch.as_tensor(orc.get("is_warmstart"), dtype=torch.int32, device=dev)
    t["first_time"][0] = torch.as_tensor(orc.get("first_time"), dtype=torch.int32, device=dev)
    t["cache_normal"][:] = torch.as_tensor(orc.get("cache_normal").reshape(orc.B, -1).T, dtype=torch.float32, device=dev)
