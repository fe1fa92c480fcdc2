"""Host logic of the serving loop (no GPU): the synthetic arrival stream and the request checks."""
import numpy as np

from genesis_amd.serving import poisson_requests


def test_poisson_requests_match_the_stated_distribution():
    r = poisson_requests(rate=200.0, duration=20.0, seed=4)
    n = len(r)
    assert abs(n - 4000) < 4 * np.sqrt(4000)
    t = np.array([x.arrival for x in r])
    assert (np.diff(t) > 0).all() and t[-1] < 20.0
    envs = np.array([x.n_envs for x in r])
    assert envs.min() >= 1 and envs.max() <= 4096
    # log-uniform: the median of log(n) is about ln(4096) / 2
    assert abs(np.median(np.log(envs)) - np.log(4096) / 2) < 0.3
    h = np.array([x.horizon for x in r])
    assert h.min() >= 50 and h.max() <= 500 and abs(h.mean() - 275) < 10
    scenes = [x.scene for x in r]
    for s in ("arm", "quadruped", "pile20"):
        assert abs(scenes.count(s) / n - 1 / 3) < 0.05
