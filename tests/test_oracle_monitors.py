"""CPU checks that pin the oracle's SubSample, GlobalAverage,
SpatialAverage, AfferentCoupling(+TemporalAverage) monitors and the region
stimulus by what the definitions and closed forms fix:

* SubSample = Raw at every istep-th step; an identity spatial mask turns
  SpatialAverage into SubSample and a single-area mask into GlobalAverage;
* AfferentCoupling at step s = the brute-force delayed coupling of step s + 1
  recomputed from the Raw trace; its TemporalAverage = window means of it;
* stimulus: with the Sum model (dfun = X + c), zero coupling and a
  deterministic scheme, one step maps X to a X + b stim_k with a, b known in
  closed form, and the k-th step of a call takes temporal sample k.
"""
import numpy as np
import pytest

from helpers import make_case
from oracle.tvb_oracle import OracleSim


def _mpr(monitors, n=12, seed=2, **kw):
    orc, _, ic = make_case("MontbrioPazoRoxin", n=n, density=0.4, lengths=(5.0, 40.0), speed=3.0, dt=0.05,
                           scheme="HeunStochastic", cparams=dict(a=0.4, b=0.01), nsig=[0.01, 0.02], seed=seed,
                           monitors=monitors, **kw)
    return orc, ic


def test_subsample_and_area_averages_from_raw():
    mask = np.arange(12) % 3
    mons = (("Raw", None, None), ("SubSample", 0.5, None), ("GlobalAverage", 0.5, None),
            ("SpatialAverage", 0.5, None, dict(spatial_mask=mask)),
            ("SpatialAverage", 0.5, None, dict(spatial_mask=np.arange(12))),
            ("SpatialAverage", 0.5, None, dict(spatial_mask=np.zeros(12, int))))
    orc, _ = _mpr(mons)
    o = orc["fp64"]
    z = o.legacy_noise(400)
    out = o.run(400, noise=z)
    (tr, xr), (ts, xs), (tg, xg), (ta, xa), (ti, xi), (t1, x1) = out
    steps = o.current_step - 400 + 1 + np.arange(400)
    sel = steps % 10 == 0
    assert len(ts) == 40 == sel.sum()
    np.testing.assert_array_equal(ts, tr[sel])
    np.testing.assert_array_equal(xs, xr[sel])
    np.testing.assert_array_equal(xi, xs)                   # one area per node: the samples themselves
    np.testing.assert_allclose(x1, xg, rtol=1e-14, atol=0)   # one area: the global mean
    np.testing.assert_allclose(xg[:, :, 0], xs.mean(axis=2), rtol=1e-14)
    for a in range(3):
        np.testing.assert_allclose(xa[:, :, a], xs[:, :, mask == a].mean(axis=2), rtol=1e-13)
    assert xg.shape == (40, 2, 1, 1) and xa.shape == (40, 2, 3, 1)


def test_afferent_coupling_is_next_steps_delayed_coupling():
    mons = (("Raw", None, None), ("AfferentCoupling", None, None),
            ("AfferentCouplingTemporalAverage", 1.0, None))
    orc, ic = _mpr(mons)
    o = orc["fp64"]
    n = 600
    z = o.legacy_noise(n)
    s0 = o.current_step
    out = o.run(n, noise=z)
    (tr, xr), (tc, xc), (tt, xt) = out
    # state of every step, initial history included: x[s] for s <= s0 + n
    hist = {s0 - (ic.shape[0] - 1) + k: ic[k] for k in range(ic.shape[0])}
    for k in range(n):
        hist[s0 + 1 + k] = xr[k, :, :, 0]
    w, d = o.weights, o.idelays
    a, b = 0.4, 0.01
    for k in (0, 1, 57, n - 1):
        s = s0 + 1 + k                      # sample of step s = coupling of step s + 1
        c = np.zeros((2, o.N))
        for i in range(o.N):
            for j in range(o.N):
                if w[i, j] != 0:
                    c[:, i] += w[i, j] * hist[s - d[i, j]][:, j]     # slot (s + 1) - 1 - d
        np.testing.assert_allclose(xc[k, :, :, 0], a * c + b, rtol=1e-12, atol=1e-15)
    assert tc[0] == pytest.approx((s0 + 1) * 0.05)
    # TemporalAverage of the afferent samples: window means ending at multiples of istep
    ends = [k for k in range(n) if (s0 + 1 + k) % 20 == 0]
    for q, e in enumerate(ends[1:], 1):
        np.testing.assert_allclose(xt[q], xc[e - 19:e + 1].mean(axis=0), rtol=1e-13)


@pytest.mark.parametrize("scheme", ["EulerDeterministic", "HeunDeterministic"])
def test_stimulus_closed_form_with_sum_model(scheme):
    n, dt, nstep = 5, 0.1, 37
    w = np.zeros((n, n))
    L = np.ones((n, n))
    sp = np.linspace(0.5, 2.5, n)
    tp = np.sin(np.arange(60) * 0.3) + 2.0
    o = OracleSim(w, L, 1.0, dt, "Sum", coupling="Linear", integrator=scheme,
                  initial_conditions=np.ones((12, 1, n)), monitors=(("Raw", None, None),), mode="fp64",
                  stimulus=(sp, tp))
    (t, x), = o.run(nstep)
    if scheme == "EulerDeterministic":     # X' = X + dt (X + 0 + st)
        alpha, beta = 1 + dt, dt
    else:                                 # X' = X + dt/2 (X + X + dt (X + st)) + dt st
        alpha, beta = 1 + dt + dt * dt / 2, dt * dt / 2 + dt
    X = np.ones(n)
    for k in range(nstep):
        X = alpha * X + beta * sp * tp[k]
        np.testing.assert_allclose(x[k, 0, :, 0], X, rtol=1e-12)
