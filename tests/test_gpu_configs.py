"""GPU parity at the benchmark configurations' own shapes.

* C0 (76 nodes, 30 % density, lengths U[10, 200] mm at 3 mm/ms,
  Generic2dOscillator, Linear a = 0.0126, HeunDeterministic dt 0.1 ms,
  1000 ms = 10 000 steps, Raw + TemporalAverage 1 ms) in full: the
  reference-precision device mode (float32 history and sums, float64 state)
  against the oracle's reference-faithful mode, bit for bit.
* C1 (192 nodes, 20 %, lengths U[5, 250] mm at 4 mm/ms, MontbrioPazoRoxin,
  Linear a = 0.5, HeunStochastic dt 0.05 ms, nsig [0.01, 0.02], injected
  normals, TemporalAverage 1 ms + Bold 720 ms): the first 16 000 steps
  (800 ms, one Bold emission) against the faithful oracle -- TemporalAverage
  bit for bit, Bold to 1e-13; and at the full 200 000 steps (10 s) the
  size-independent property that one launch and four chained launches give
  identical outputs and state.
"""
import numpy as np
import pytest

from helpers import rel_err
from oracle.tvb_oracle import OracleSim, set_idelays

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _c0():
    from tvb_hip import sim as S
    from tvb_hip.connectome import synthetic
    n, dt, speed = 76, 0.1, 3.0
    w, L = synthetic(n, 0.30, 10.0, 200.0, seed=76)
    _, _, H = set_idelays(L, speed, dt)
    rng = np.random.default_rng(760)
    ic = rng.uniform(-1.0, 1.0, size=(H, 2, n))
    mons = (("Raw", None, None), ("TemporalAverage", 1.0, None))
    orc = OracleSim(w, L, speed, dt, "Generic2dOscillator", coupling_params=dict(a=0.0126),
                    integrator="HeunDeterministic", initial_conditions=ic, monitors=mons, mode="faithful")
    sim = S.Simulator(connectivity=S.Connectivity(weights=w, tract_lengths=L, speed=np.array([speed])),
                      model=S.Generic2dOscillator(), coupling=S.Linear(a=np.array([0.0126])),
                      integrator=S.HeunDeterministic(dt=dt),
                      monitors=[S.Raw(), S.TemporalAverage(period=1.0)], initial_conditions=ic).configure()
    return orc, sim


def _c1(n_steps_noise, bold=True):
    from tvb_hip import sim as S
    from tvb_hip.connectome import synthetic
    n, dt, speed = 192, 0.05, 4.0
    w, L = synthetic(n, 0.20, 5.0, 250.0, seed=192)
    _, _, H = set_idelays(L, speed, dt)
    rng = np.random.default_rng(1920)
    ic = np.empty((H, 2, n))
    ic[:, 0] = rng.uniform(0.0, 2.0, size=(H, n))
    ic[:, 1] = rng.uniform(-2.0, 1.5, size=(H, n))
    z = rng.standard_normal((n_steps_noise, 2, n))
    nsig = np.array([0.01, 0.02])
    mons = [("TemporalAverage", 1.0, None)] + ([("Bold", 720.0, None)] if bold else [])
    orc = OracleSim(w, L, speed, dt, "MontbrioPazoRoxin", coupling_params=dict(a=0.5),
                    integrator="HeunStochastic", nsig=nsig, initial_conditions=ic, monitors=tuple(mons),
                    mode="faithful")
    integ = S.HeunStochastic(dt=dt)
    integ.noise = S.Additive(nsig=nsig)
    smons = [S.TemporalAverage(period=1.0)] + ([S.Bold(period=720.0)] if bold else [])
    sim = S.Simulator(connectivity=S.Connectivity(weights=w, tract_lengths=L, speed=np.array([speed])),
                      model=S.MontbrioPazoRoxin(), coupling=S.Linear(a=np.array([0.5])), integrator=integ,
                      monitors=smons, initial_conditions=ic).configure()
    return orc, sim, z


def _backend():
    from tvb_hip.backend import HipBackend
    return HipBackend()


def test_c0_full_run_reference_precision_bit_exact():
    _say("C0 full run")
    orc, sim = _c0()
    nstep = 10000
    ref = orc.run(nstep)
    out = _backend().run_sim(sim, nstep=nstep)          # precision "auto" = the reference's mix
    for (tr, xr), (td, xd), name in zip(ref, out, ("raw", "tavg")):
        assert xd.shape == xr.shape == ((nstep, 1, 76, 1) if name == "raw" else (nstep // 10, 1, 76, 1))
        np.testing.assert_allclose(td, tr, rtol=0, atol=1e-9)
        np.testing.assert_array_equal(xd, xr)
    np.testing.assert_array_equal(np.asarray(sim.current_state)[..., 0], orc.current_state[..., 0])


def test_c1_prefix_reference_precision():
    _say("C1 prefix")
    nstep = 16000
    orc, sim, z = _c1(nstep)
    ref = orc.run(nstep, noise=z)
    out = _backend().run_sim(sim, nstep=nstep, noise=z)
    (tt, ta), (tb, bo) = ref
    (dt_, da), (db, dbo) = out
    assert da.shape == ta.shape and dbo.shape == bo.shape == (1, 2, 192, 1)
    np.testing.assert_allclose(dt_, tt, rtol=0, atol=1e-9)
    np.testing.assert_allclose(db, tb, rtol=0, atol=1e-9)
    np.testing.assert_array_equal(da, ta)
    e = rel_err(dbo, bo)
    _say("   bold rel", e)
    assert e <= 1e-13


def test_c1_full_length_one_launch_equals_chained_launches():
    """Size-independent property at C1's full 200 000 steps: the plan run in
    one launch and in four chained launches gives identical monitor outputs
    and final state (the kernel carries state, ring, monitor sums and step
    counters across launch boundaries exactly)."""
    _say("C1 full length chaining")
    from tvb_hip.backend import HipBackend
    nstep = 200000
    orc, sim, z = _c1(1, bold=True)
    be = HipBackend()
    res = []
    for chunks in ((nstep,), (50000,) * 4):
        plan = be.build_plan(sim, precision="mixed", noise="philox", max_steps=max(chunks), seed=1920)
        ta, bo = [], []
        for n in chunks:
            plan.run(n)
            (t0, x0), = plan.read(0)
            (t1, x1), = plan.read(1)
            ta.append(x0)
            bo.append(x1)
        hist, state, step = plan.export(0)
        assert step == sim.current_step + nstep
        res.append((np.concatenate(ta), np.concatenate(bo), state, hist))
        plan.close()
    (a0, b0, s0, h0), (a1, b1, s1, h1) = res
    assert a0.shape[0] == nstep // 20 and b0.shape[0] == nstep // 14400
    assert np.all(np.isfinite(a0))
    np.testing.assert_array_equal(a0, a1)
    np.testing.assert_array_equal(b0, b1)
    np.testing.assert_array_equal(s0, s1)
    np.testing.assert_array_equal(h0, h1)


def test_probe_gather_rates_are_plausible():
    """The gather-pattern calibration runs over a buffer beyond the caches
    and reports rates between 0.5 and 8 TB/s (HBM peak)."""
    from tvb_hip import _lib as L
    r = L.probe_gather(256, False)
    rw = L.probe_gather(256, True)
    print("probe", r / 1e9, rw / 1e9, flush=True)
    assert 5e11 < r < 8e12
    assert 5e11 < rw < 8e12
