"""GPU parity of the Balloon-Windkessel BOLD kernel (tvbk_balloon) against
the restatement of BalloonModel.evaluate (oracle/balloon.py).

Tolerances: max |dev - ref| / max |ref| <= 1e-11 -- float64 in the
reference's operation order; the two `pow` evaluations per model step may
differ in the last bit between the device library and the host's libm.
Chained device calls are bit-identical to one call.
"""
import numpy as np
import pytest

from helpers import rel_err
from oracle import balloon as B

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


class _TS:
    def __init__(self, data, period, unit="ms"):
        self.data = data
        self.sample_period = period
        self.sample_period_unit = unit
        self.time = np.arange(data.shape[0]) * period


@pytest.mark.parametrize("kw", [dict(), dict(bold_model="linear"), dict(RBM=False),
                                dict(normalize_neural_input=True), dict(tau_s=1.2, alpha=0.4, E0=0.35)])
def test_balloon_model_matches_oracle(kw):
    _say("balloon", kw)
    from tvb_hip.analyzers import BalloonModel
    rng = np.random.default_rng(11)
    T, N = 4000, 37
    x = rng.uniform(0.0, 1.0, size=(T, 2, N, 1))                             # firing-rate-like input
    m = BalloonModel(_TS(x, 1.0), **kw)
    out = m.evaluate()
    par = {k: float(np.asarray(v).reshape(-1)[0]) for k, v in kw.items()
           if k in ("tau_s", "alpha", "E0")}
    ref, _ = B.evaluate(x[:, 0], 1.0, params=par, nonlinear=kw.get("bold_model", "nonlinear") == "nonlinear",
                        rbm=kw.get("RBM", True), normalize=kw.get("normalize_neural_input", False))
    assert out.data.shape == (T, 1, N, 1)
    e = rel_err(out.data[:, 0], ref)
    _say("   rel", e)
    assert e <= 1e-11
    np.testing.assert_allclose(out.time, np.arange(T) * 1e-3, rtol=0, atol=1e-15)


def test_balloon_chained_calls_bit_identical():
    import torch
    from tvb_hip.analyzers import BalloonModel, run_balloon
    rng = np.random.default_rng(5)
    T, n = 3000, 1000
    x = torch.from_numpy(rng.uniform(0.0, 1.0, (T, n))).cuda()
    par = BalloonModel().params()
    y1 = torch.empty((T, n), dtype=torch.float64, device="cuda")
    s1 = torch.empty((4, n), dtype=torch.float64, device="cuda")
    run_balloon(x.data_ptr(), 8, n, T, n, 1.0, par, 0, True, s1.data_ptr(), y1.data_ptr())
    y2 = torch.empty((T, n), dtype=torch.float64, device="cuda")
    s2 = torch.empty((4, n), dtype=torch.float64, device="cuda")
    for a, b in ((0, 700), (700, 2000), (2000, T)):
        run_balloon(x[a:].data_ptr(), 8, n, b - a, n, 1.0, par, 0, a == 0, s2.data_ptr(), y2[a:].data_ptr())
    torch.cuda.synchronize()
    assert torch.equal(y1, y2) and torch.equal(s1, s2)


@pytest.mark.parametrize("precision", ["mixed", "fp32"])
def test_balloon_stream_on_sweep_tavg(precision):
    """BalloonStream over a simulated sweep's TemporalAverage (two chunks,
    on the device) == the restatement applied to the TemporalAverage read
    back to the host."""
    _say("balloon stream", precision)
    import torch
    from tvb_hip import _lib as L
    from tvb_hip.analyzers import BalloonStream
    from tvb_hip.connectome import idelays, synthetic
    from tvb_hip.plan import CSR, MonitorSpec, Plan
    n, Bn, dt = 24, 64, 0.05
    w, Lm = synthetic(n, 0.3, 5.0, 60.0, seed=8)
    d, H = idelays(Lm, 3.0, dt)
    p = np.empty((Bn, 12))
    p[:, :8] = [1.0, 0.0, 1.0, 15.0, -5.0, 0.0, 1.0, 0.0]
    p[:, 8] = np.linspace(0.0, 0.6, Bn)
    p[:, 9] = 0.0
    p[:, 10:] = 1e-3
    prec = L.MIXED if precision == "mixed" else L.F32
    plan = Plan(model=L.MODEL_MPR, cfun=L.CFUN_LINEAR, scheme=L.HEUN_STO, precision=prec, noise=L.NOISE_PHILOX,
                conns=[CSR(w, d)], params=p, dt=dt, obs_coef=np.eye(2),
                monitors=[MonitorSpec(L.MON_TAVG, 20, [0, 1], period=1.0)], max_steps=30000, seed=3, tile=32)
    plan.init_uniform(4, [0.0, -2.0], [2.0, 1.5])
    stream = BalloonStream(plan, 0)
    ys, xs = [], []
    for steps in (30000, 20000):
        plan.run(steps)
        ys.append(stream.feed().cpu().numpy())
        xs.append(np.stack([x for _, x in plan.read(0)], axis=-1))    # (cnt, nvoi, n, B)
    torch.cuda.synchronize()
    y = np.concatenate(ys)                                            # (T, n * B)
    x = np.concatenate(xs)[:, 0].reshape(y.shape[0], n * Bn)           # first variable, same order
    ref, _ = B.evaluate(x.astype(np.float64), 1.0)
    assert y.shape == (2500, n * Bn)
    e = rel_err(y, ref)
    _say("   rel", e)
    assert e <= 1e-11
    plan.close()
