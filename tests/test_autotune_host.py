"""Host test of the autotuner's search and cache with stand-in plans."""
import numpy as np

from tvb_hip import autotune
from tvb_hip.plan import CSR, MonitorSpec


class _P:
    def __init__(self, ms):
        self.ms, self.n = ms, 0

    def run(self, n):
        self.n = n

    def sync(self):
        pass

    def last_run_ms(self):
        return self.ms * self.n

    def close(self):
        pass


def test_tune_picks_fastest_and_caches(tmp_path):
    from tvb_hip import _lib as L
    w = np.array([[0, 1.0], [1.0, 0]])
    kw = dict(model=L.MODEL_MPR, cfun=L.CFUN_LINEAR, scheme=L.HEUN_STO, precision=L.F32,
              noise=L.NOISE_PHILOX, conns=[CSR(w, np.array([[0, 3], [3, 0]]))], params=np.zeros((64, 12)),
              dt=0.05, obs_coef=np.eye(2), monitors=[MonitorSpec(L.MON_TAVG, 20, [0, 1])])
    speed = {(32, 16): 3.0, (32, 8): 2.0, (16, 16): 5.0, (16, 8): 4.0}
    calls = []

    def make(tile, waves):
        calls.append((tile, waves))
        if (tile, waves) not in speed:
            raise L.TvbkError(-4, "does not fit")
        return _P(speed[(tile, waves)])
    cache = str(tmp_path / "tune.json")
    t, wv, table = autotune.tune(kw, make_plan=make, cache_path=cache, device="x")
    assert (t, wv) == (32, 8) and table[(32, 8)] == 2.0 and len(table) == 4
    n = len(calls)
    t2, w2, _ = autotune.tune(kw, make_plan=make, cache_path=cache, device="x")
    assert (t2, w2) == (32, 8) and len(calls) == n          # served from the cache
