"""GPU checks of checkpoint export / import (tvbk_export_state,
tvbk_export_monitor, tvbk_import_monitor, SimulationHistory).

Tolerances (stated per test):
  * device -> record -> new plan, continued: bit-identical to the plan that
    ran on, in mixed and float64 precision; float32 plans too except the
    TemporalAverage, whose float32 Kahan pair travels as one float64 sum
    (max rel <= 1e-6);
  * reference-format interop against the oracle (faithful mode vs the mixed
    device precision): TemporalAverage and Raw bit-exact, Bold max rel
    <= 1e-13 (the HRF sum order differs from numpy.dot's);
  * the oracle continuing from a record the device wrote: bit-identical to
    the oracle continuing from its own record.
"""
import numpy as np
import pytest

from helpers import make_case, oracle_fill, oracle_record, rel_err

pytestmark = pytest.mark.gpu

MONS = (("Raw", None, None), ("TemporalAverage", 1.0, None), ("Bold", 20.0, None))


def _say(*a):
    print(*a, flush=True)


def _case(n=14, seed=5):
    return make_case("MontbrioPazoRoxin", n=n, density=0.4, lengths=(5.0, 40.0), speed=3.0, dt=0.05,
                     scheme="HeunStochastic", cparams=dict(a=0.4), nsig=[0.01, 0.02], seed=seed, monitors=MONS)


def _sweep_plan(prec, kernel, B, hrf_taps=300):
    """MPR sweep plan (Philox noise) with Raw, TemporalAverage and a Bold
    monitor whose stock is kept on the device (HRF shortened to
    ``hrf_taps`` interim samples to keep the exported stock small)."""
    from tvb_hip import _lib as L
    from tvb_hip.connectome import idelays, synthetic
    from tvb_hip.plan import CSR, MonitorSpec, Plan
    from oracle.tvb_oracle import first_order_volterra
    n, dt = 18, 0.05
    w, Lm = synthetic(n, 0.3, 5.0, 50.0, seed=4)
    d, H = idelays(Lm, 3.0, dt)
    p = np.empty((B, 12))
    p[:, :8] = [1.0, 0.0, 1.0, 15.0, -5.0, 0.0, 1.0, 0.0]
    p[:, 8] = np.linspace(0.05, 0.6, B)
    p[:, 9] = 0.0
    p[:, 10:] = 1e-3
    taps = first_order_volterra(np.arange(hrf_taps) * 4e-3)
    prec_id = {"mixed": L.MIXED, "fp32": L.F32, "fp64": L.F64}[prec]
    kern = {"batch": L.KERNEL_BATCH_LANE, "edge": L.KERNEL_EDGE_LANE}[kernel]
    mons = [MonitorSpec(L.MON_RAW, 1, [0, 1]), MonitorSpec(L.MON_TAVG, 20, [0, 1], period=1.0),
            MonitorSpec(L.MON_BOLD, 200, [0], interim_istep=80, taps=taps, k1_v0=5.6 * 0.02, keep_stock=True)]
    plan = Plan(model=L.MODEL_MPR, cfun=L.CFUN_LINEAR, scheme=L.HEUN_STO, precision=prec_id, noise=L.NOISE_PHILOX,
                conns=[CSR(w, d)], params=p, dt=dt, obs_coef=np.eye(2), monitors=mons, max_steps=4000, seed=11,
                kernel=kern, tile=32 if kernel == "batch" else 0)
    plan.init_uniform(9, [0.0, -2.0], [2.0, 1.5])
    return plan


def _read(plan):
    return [[x.copy() for _, x in plan.read(m)] for m in range(len(plan.monitors))]


@pytest.mark.parametrize("prec,kernel,B", [("mixed", "batch", 40), ("fp32", "batch", 40), ("fp64", "batch", 40),
                                           ("mixed", "edge", 3)])
@pytest.mark.parametrize("s1", [1037, 1200])
def test_device_record_resumes_bit_exact(prec, kernel, B, s1):
    """Plan A runs s1 steps, every instance is recorded, A runs on; plan C
    (same description) loads the records and runs the same steps: identical
    outputs (Raw, TemporalAverage, Bold) and final state."""
    from tvb_hip.checkpoint import SimulationHistory
    _say("device resume", prec, kernel, B, s1)
    a = _sweep_plan(prec, kernel, B)
    a.run(s1)
    recs = [SimulationHistory.from_plan(a, b) for b in range(B)]
    assert recs[0].current_step == s1
    assert recs[0].integrator_noise_rng_state_algo == "Philox4x32-10"
    s2 = 1500
    a.run(s2)
    out_a = _read(a)
    c = _sweep_plan(prec, kernel, B)
    for b, h in enumerate(recs):
        h.to_plan(c, b)
    c.run(s2)
    out_c = _read(c)
    for m in range(3):
        for b in range(B):
            if prec == "fp32" and m == 1:
                e = rel_err(out_c[m][b], out_a[m][b])
                assert e <= 1e-6, (m, b, e)
            else:
                np.testing.assert_array_equal(out_c[m][b], out_a[m][b], err_msg="monitor %d inst %d" % (m, b))
    for b in (0, B - 1):
        ha, sa, ka = a.export(b)
        hc, sc, kc = c.export(b)
        assert ka == kc == s1 + s2
        np.testing.assert_array_equal(hc, ha)
        np.testing.assert_array_equal(sc, sa)
    a.close()
    c.close()


def test_import_into_every_instance_and_branch():
    """One record loaded into every instance (inst = -1) of a sweep over the
    coupling strength: the instance with the recorded parameters reproduces
    the recorded run, the others branch from the same state."""
    from tvb_hip.checkpoint import SimulationHistory
    _say("branch")
    B = 40
    a = _sweep_plan("mixed", "batch", B)
    a.run(1111)
    h = SimulationHistory.from_plan(a, 7)
    a.run(800)
    out_a = _read(a)
    c = _sweep_plan("mixed", "batch", B)
    h.to_plan(c, -1)
    c.run(800)
    out_c = _read(c)
    for m in range(3):
        np.testing.assert_array_equal(out_c[m][7], out_a[m][7])
    assert not np.array_equal(out_c[0][0], out_a[0][0])
    np.testing.assert_array_equal(c.export(0)[0].shape, a.export(0)[0].shape)
    a.close()
    c.close()


@pytest.mark.parametrize("s1", [1037, 1200])
def test_reference_record_continues_on_device(s1):
    """A record taken from the oracle (the reference's populate_from) filled
    into a simulator (fill_into: Bold interim window zero) and continued by
    HipBackend.run_sim == the oracle filled from the same record."""
    from tvb_hip.backend import HipBackend
    _say("reference record -> device", s1)
    orc, _, _ = _case()
    o = orc["faithful"]
    s2 = 1300
    z = o.legacy_noise(s1 + s2)
    o.run(s1, noise=z[:s1])
    rec = oracle_record(o)
    ref = oracle_fill(_case()[0]["faithful"], rec).run(s2, noise=z[s1:])
    _, sim, _ = _case()
    rec.fill_into(sim)
    out = HipBackend().run_sim(sim, nstep=s2, precision="mixed", noise=z[s1:])
    for m in range(3):
        t_r, x_r = ref[m]
        t, x = out[m]
        assert len(t) == len(t_r) > 0
        np.testing.assert_allclose(t, t_r, rtol=0, atol=1e-9)
        if m == 2:
            e = rel_err(x, x_r)
            _say("   bold rel", e)
            assert e <= 1e-13
        else:
            np.testing.assert_array_equal(x, x_r)
    assert sim.current_step == rec.current_step + s2


@pytest.mark.parametrize("s1", [1037, 1200])
def test_device_record_continues_on_oracle(s1):
    """HipBackend.run_sim (mixed precision, write-back) for s1 steps, then the
    reference's populate_from on the simulator; the oracle filled from that
    record continues exactly as it does from its own record at s1."""
    from tvb_hip.backend import HipBackend
    from tvb_hip.checkpoint import SimulationHistory
    _say("device record -> oracle", s1)
    orc, sim, _ = _case()
    o = orc["faithful"]
    s2 = 1300
    z = o.legacy_noise(s1 + s2)
    o.run(s1, noise=z[:s1])
    ref = oracle_fill(_case()[0]["faithful"], oracle_record(o)).run(s2, noise=z[s1:])
    HipBackend().run_sim(sim, nstep=s1, precision="mixed", noise=z[:s1])
    rec = SimulationHistory().populate_from(sim)
    assert rec.current_step == o.current_step
    assert rec.history.dtype == np.float32
    np.testing.assert_array_equal(rec.history[..., 0], o.ring)
    np.testing.assert_array_equal(rec.current_state, o.current_state)
    out = oracle_fill(_case()[0]["faithful"], rec).run(s2, noise=z[s1:])
    for m in range(3):
        np.testing.assert_array_equal(out[m][0], ref[m][0])
        np.testing.assert_array_equal(out[m][1], ref[m][1])


def test_run_sim_twice_equals_once():
    """Two run_sim calls on one simulator (state, history and monitor stocks
    written back in between) == one call over both spans."""
    from tvb_hip.backend import HipBackend
    _say("run_sim twice")
    _, sim1, _ = _case()
    _, sim2, _ = _case()
    z = np.random.default_rng(2).standard_normal((2500, 2, 14))
    be = HipBackend()
    one = be.run_sim(sim1, nstep=2500, precision="mixed", noise=z)
    p1 = be.run_sim(sim2, nstep=1037, precision="mixed", noise=z[:1037])
    p2 = be.run_sim(sim2, nstep=1463, precision="mixed", noise=z[1037:])
    for m in range(3):
        np.testing.assert_array_equal(np.concatenate([p1[m][1], p2[m][1]]), one[m][1])
        np.testing.assert_array_equal(np.concatenate([p1[m][0], p2[m][0]]), one[m][0])


def test_bold_stock_export_needs_keep_stock():
    from tvb_hip import _lib as L
    from tvb_hip.backend import HipBackend
    _, sim, _ = _case()
    plan = HipBackend().build_plan(sim, precision="mixed", noise="philox", max_steps=100)
    with pytest.raises(L.TvbkError):
        plan.export_monitor(0, 2, stock=True)
    st, it = plan.export_monitor(0, 2, stock=False, interim=True)
    assert st is None and it.shape == (80, 2, 14, 1)
    plan.close()
