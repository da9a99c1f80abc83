"""GPU parity: HIP step kernels through the C-ABI vs the numpy oracle.

Tolerances (stated per test):
  * integer KAT and slot arithmetic: bit-exact;
  * FP64 device mode vs the oracle's pure-fp64 mode: max relative error
    <= 1e-10 on state traces and every monitor output;
  * FP32 device mode vs the pure-fp64 oracle: TemporalAverage max-rel
    <= 1e-4, Bold max-rel <= 1e-6, per-region correlation >= 0.999.
"""
import numpy as np
import pytest

from helpers import make_case, per_region_corr, rel_err
from oracle.tvb_oracle import history_kat, set_idelays

pytestmark = pytest.mark.gpu


def _backend():
    from tvb_hip.backend import HipBackend
    return HipBackend()


def _say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_history_kat_on_device(precision):
    """history_test.py:67-112 exact-propagation table through the device
    ring (SUM model, IdCoupling, Identity scheme): bit-exact integers."""
    import tvb_hip._lib as L
    from tvb_hip.plan import CSR, MonitorSpec, Plan
    _say("history KAT", precision)
    sim, t_ref, raw_ref = history_kat()
    n = 4
    w = np.zeros((n, n))
    for i in range(n - 1):
        w[i, i + 1] = 1
    dist = np.r_[:n * n].reshape((n, n)).astype(float)
    _, idel, H = set_idelays(dist, 1.0, 1.0)
    plan = Plan(model=L.MODEL_SUM, cfun=L.CFUN_IDENTITY, scheme=L.IDENTITY,
                precision=L.F32 if precision == "fp32" else L.F64, noise=L.NOISE_NONE,
                conns=[CSR(w, idel)], params=np.zeros((1, 0)), dt=1.0, obs_coef=np.ones((1, 1)),
                monitors=[MonitorSpec(L.MON_RAW, 1, [0])], max_steps=10, kernel=L.KERNEL_EDGE_LANE)
    hist = np.ones((H, 1, n))
    plan.set_history(0, hist, np.ones((1, n)), H - 1)
    plan.run(10)
    (t, x), = plan.read(0)
    np.testing.assert_array_equal(np.asarray(x)[:, 0, :], raw_ref[:, 0, :, 0])
    np.testing.assert_array_equal(t, t_ref)
    plan.close()


def test_slot_arithmetic_bit_exact():
    """Device slot arithmetic == (step - 1 - d + H) % H (history.py:252)."""
    _say("slot arithmetic")
    orc, sim, _ = make_case("MontbrioPazoRoxin", n=24, lengths=(5.0, 80.0), speed=2.0, dt=0.05)
    be = _backend()
    plan = be.build_plan(sim, precision="fp32", max_steps=4)
    o = orc["fp64"]
    for step in (1, 2, o.H - 1, o.H, o.H + 1, 5 * o.H + 3, 123457):
        dev = plan.debug_slots(0, step)
        ref = ((step - 1 - o.nnz_idelays + o.H) % o.H).astype(np.int32)
        np.testing.assert_array_equal(dev, ref)
    plan.close()


@pytest.mark.parametrize("model,scheme,coupling,cparams,nsig,dt", [
    ("Generic2dOscillator", "HeunDeterministic", "Linear", dict(a=0.0126), None, 0.1),
    ("MontbrioPazoRoxin", "HeunStochastic", "Linear", dict(a=0.5), [0.01, 0.02], 0.05),
    ("JansenRit", "EulerStochastic", "SigmoidalJansenRit", {}, [0.0, 0.0, 0.0, 0.01, 0.01, 0.01], 0.1),
    ("MontbrioPazoRoxin", "EulerDeterministic", "Linear", dict(a=0.3), None, 0.05),
])
def test_single_instance_fp64_matches_oracle(model, scheme, coupling, cparams, nsig, dt):
    """FP64 device mode vs the pure-fp64 oracle, injected noise: max rel
    <= 1e-10 on Raw, TemporalAverage and Bold."""
    _say("single fp64", model, scheme)
    mons = (("Raw", None, None), ("TemporalAverage", 1.0, None), ("Bold", 20.0, None))
    orc, sim, _ = make_case(model, n=20, density=0.3, lengths=(5.0, 60.0), speed=3.0, dt=dt,
                            scheme=scheme, coupling=coupling, cparams=cparams, nsig=nsig, monitors=mons,
                            spread=0.02 if model == "JansenRit" else 1.0)
    nstep = 1200
    o = orc["fp64"]
    z = np.random.default_rng(7).standard_normal((nstep, len(sim.model.state_variables), o.N))
    ref = o.run(nstep, noise=z if nsig is not None else None)
    out = _backend().run_sim(sim, nstep=nstep, precision="fp64", noise=z)
    for (tr, xr), (td, xd), name in zip(ref, out, ("raw", "tavg", "bold")):
        assert xd.shape == xr.shape, (name, xd.shape, xr.shape)
        np.testing.assert_allclose(td, tr, rtol=0, atol=1e-9)
        e = rel_err(xd, xr)
        _say("  ", name, "rel", e)
        assert e <= 1e-10, (name, e)


def test_mpr_fp32_vs_fp64_oracle():
    """FP32 device vs pure-fp64 oracle (shared injected noise):
    TemporalAverage max-rel <= 1e-4, Bold max-rel <= 1e-6, corr >= 0.999."""
    _say("mpr fp32")
    mons = (("TemporalAverage", 1.0, None), ("Bold", 20.0, None))
    orc, sim, _ = make_case("MontbrioPazoRoxin", n=32, density=0.2, lengths=(5.0, 100.0), speed=4.0,
                            dt=0.05, scheme="HeunStochastic", cparams=dict(a=0.5), nsig=[0.01, 0.02],
                            monitors=mons)
    nstep = 4000
    o = orc["fp64"]
    z = np.random.default_rng(3).standard_normal((nstep, 2, o.N))
    ref = o.run(nstep, noise=z)
    out = _backend().run_sim(sim, nstep=nstep, precision="fp32", noise=z)
    (_, ta_r), (_, bo_r) = ref
    (_, ta_d), (_, bo_d) = out
    e_t, e_b = rel_err(ta_d, ta_r), rel_err(bo_d, bo_r)
    _say("   tavg rel", e_t, "bold rel", e_b, "corr", per_region_corr(ta_d, ta_r))
    assert e_t <= 1e-4
    assert e_b <= 1e-6
    assert per_region_corr(ta_d, ta_r) >= 0.999


@pytest.mark.parametrize("model,scheme,cparams,nsig,dt,exact", [
    ("Generic2dOscillator", "HeunDeterministic", dict(a=0.0126), None, 0.1, True),
    ("MontbrioPazoRoxin", "HeunStochastic", dict(a=0.5), [0.01, 0.02], 0.05, True),
    ("MontbrioPazoRoxin", "EulerStochastic", dict(a=0.3), [0.01, 0.01], 0.05, True),
    ("Generic2dOscillator", "EulerDeterministic", dict(a=0.02), None, 0.1, True),
    ("JansenRit", "EulerStochastic", dict(a=1e-3), [0.0, 0.0, 0.0, 0.01, 0.01, 0.01], 0.1, False),
])
def test_mixed_matches_faithful_oracle(model, scheme, cparams, nsig, dt, exact):
    """MIXED device mode (float32 history, weights and Linear sums; float64
    post, state, integration and monitors -- the reference's own precision)
    vs the oracle's reference-faithful mode, injected noise.  G2D and MPR
    (polynomial dfuns): bit-identical Raw, TemporalAverage, final state and
    float32 history; Bold max rel <= 1e-13 (its 5000-tap sum runs in another
    order than numpy.dot's).  JansenRit (exp in its sigmoid, whose last bit
    may differ between libraries): max rel <= 1e-10."""
    _say("mixed vs faithful", model, scheme)
    mons = (("Raw", None, None), ("TemporalAverage", 1.0, None), ("Bold", 20.0, None))
    orc, sim, _ = make_case(model, n=20, density=0.3, lengths=(5.0, 60.0), speed=3.0, dt=dt,
                            scheme=scheme, cparams=cparams, nsig=nsig, monitors=mons,
                            spread=0.02 if model == "JansenRit" else 1.0)
    nstep = 1500
    o = orc["faithful"]
    z = np.random.default_rng(11).standard_normal((nstep, len(sim.model.state_variables), o.N))
    ref = o.run(nstep, noise=z if nsig is not None else None)
    out = _backend().run_sim(sim, nstep=nstep, precision="mixed", noise=z)
    for (tr, xr), (td, xd), name in zip(ref, out, ("raw", "tavg", "bold")):
        assert xd.shape == xr.shape, (name, xd.shape, xr.shape)
        assert xd.dtype == np.float64
        np.testing.assert_allclose(td, tr, rtol=0, atol=1e-9)
        e = rel_err(xd, xr)
        _say("  ", name, "rel", e)
        if exact and name != "bold":
            np.testing.assert_array_equal(xd, xr)
        else:
            assert e <= (1e-13 if exact else 1e-10), (name, e)
    # the written-back simulator state is the oracle's: float64 state and
    # the float32 history ring in the reference's slot order
    st = np.asarray(sim.current_state)[..., 0]
    hb = np.asarray(sim.history.buffer)[..., 0]
    assert hb.dtype == np.float32
    if exact:
        np.testing.assert_array_equal(st, o.current_state[..., 0])
        np.testing.assert_array_equal(hb, o.ring)
    else:
        assert rel_err(st, o.current_state[..., 0]) <= 1e-10
        assert rel_err(hb, o.ring) <= 1e-6


def test_mixed_rejects_sigmoidal_coupling():
    """MIXED with SigmoidalJansenRit is refused (its reference pre and sum
    run in float64 on float32 history values, which the ring does not hold)."""
    import tvb_hip._lib as L
    orc, sim, _ = make_case("JansenRit", n=8, scheme="EulerStochastic", coupling="SigmoidalJansenRit",
                            nsig=[0.0, 0.0, 0.0, 0.01, 0.01, 0.01], spread=0.02)
    with pytest.raises(L.TvbkError):
        _backend().build_plan(sim, precision="mixed", max_steps=4)


def test_jansen_rit_auto_precision_is_fp64():
    """A drop-in run_sim of JansenRit + SigmoidalJansenRit with the default
    precision ("auto") integrates in float64 (the reference's pre and sum
    are float64, coupling.py:118-124, :370-373): max rel <= 1e-10 against
    the pure-fp64 oracle on Raw, TemporalAverage and Bold."""
    _say("jr auto precision")
    mons = (("Raw", None, None), ("TemporalAverage", 1.0, None), ("Bold", 20.0, None))
    nsig = [0.0, 0.0, 0.0, 0.01, 0.01, 0.01]
    orc, sim, _ = make_case("JansenRit", n=20, density=0.3, lengths=(5.0, 60.0), speed=3.0, dt=0.1,
                            scheme="EulerStochastic", coupling="SigmoidalJansenRit", nsig=nsig, monitors=mons,
                            spread=0.02)
    nstep = 1200
    o = orc["fp64"]
    z = np.random.default_rng(8).standard_normal((nstep, 6, o.N))
    ref = o.run(nstep, noise=z)
    out = _backend().run_sim(sim, nstep=nstep, noise=z)
    for (tr, xr), (td, xd), name in zip(ref, out, ("raw", "tavg", "bold")):
        assert xd.dtype == np.float64
        e = rel_err(xd, xr)
        _say("  ", name, "rel", e)
        assert e <= 1e-10, (name, e)
