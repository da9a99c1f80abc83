"""Host logic of bench.py's error sample, without a GPU.

The device plan is replaced by a stand-in that runs the reference-faithful
oracle for each instance (what the MIXED kernel computes bit for bit), so
the aggregation must report: the MIXED path equal to the reference
precision element for element, its error within the bar on the
well-conditioned instances, the non-finite instances of the high-coupling
corner listed and flagged, and a line that serialises to JSON (numpy
scalars included).
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tvb-root_amd"))


class _OraclePlan:
    """Stand-in for the batched plan of bench.error_sample: each instance
    run on the faithful oracle from the history and noise it is given."""

    def __init__(self, ctx, G, sig, bold, precision):
        self.ctx, self.G, self.sig, self.bold, self.precision = ctx, G, sig, bold, precision
        self.nch, self.kernel, self.tile, self.waves = 2, 1, 32, 16
        self.hist = {}

    def set_history(self, b, hist, state, current_step):
        self.hist[b] = np.asarray(hist, float)

    def set_noise(self, z):
        self.z = z

    def run(self, n):
        from oracle.tvb_oracle import OracleSim
        self.out = []
        for b in range(len(self.G)):
            o = OracleSim(self.ctx["w"], self.ctx["Lmat"], 5.0, 0.05, "MontbrioPazoRoxin",
                          coupling_params=dict(a=float(self.G[b])), integrator="HeunStochastic",
                          nsig=np.array([self.sig[b]] * 2), initial_conditions=self.hist[b],
                          monitors=(("TemporalAverage", 1.0, None), ("Bold", self.bold, None)),
                          mode="faithful" if self.precision == 2 else "fp64")
            with np.errstate(all="ignore"):
                ((_, t), (_, bo)), st = o.run(n, noise=self.z[..., b], record_state=True)
            self.out.append((t[..., 0], bo[..., 0], np.isfinite(st).all()))

    def read(self, m):
        return [(None, o[m]) for o in self.out]

    def flags(self):
        return np.array([0 if o[2] else 1 for o in self.out], np.int32)

    def close(self):
        pass


def test_error_sample_aggregation_and_json(monkeypatch):
    import bench
    w, Lm, d, H, G, sig = bench.c2_setup(8192)
    ctx = dict(w=w, Lmat=Lm, d=d, H=H, G=G, sig=sig, skip=False)

    def fake_plan(w_, d_, Gs, ss, dt, n_steps, prec, noise, seed, bold=2000.0, **kw):
        return _OraclePlan(ctx, Gs, ss, bold, prec)

    monkeypatch.setattr(bench, "build_c2_plan", fake_plan)
    # G = 0 (well conditioned) and the G = 2, sigma = 1e-2 corner (diverges)
    r = bench.error_sample(ctx, ("mixed",), n_steps=3600, bold=20.0, grid=((0, 0), (127, 63)))
    m = r["mixed"]
    for scen in ("random_ic", "settled"):
        s = m[scen]
        assert s["tavg_bit_exact_vs_reference_precision"].startswith("1 of 1")
        assert s["bold_max_rel_vs_reference_precision"] == 0.0
        assert s["well_conditioned_instances"] == 1
        assert s["tavg_max_rel"] == s["reference_precision_tavg_max_rel"] <= 1e-4
        assert s["nonfinite_in_oracle"] == [8191]
        assert s["device_flags_agree_where_finiteness_is_stable"] is True
    assert m["meets_error_bar"] is True
    assert len(r["per_instance"]) == 4
    line = {"error_vs_fp64_oracle": r, "flag": np.bool_(True), "n": np.int64(3), "a": np.zeros(2)}
    back = json.loads(json.dumps(line, default=bench._json_default))
    assert back["flag"] is True and back["n"] == 3 and back["a"] == [0.0, 0.0]
