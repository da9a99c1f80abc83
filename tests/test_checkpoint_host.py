"""Host-side checks of the checkpoint format (no GPU).

* the SimulationHistory mirror keeps the reference's field names and
  round-trips through its file form;
* populate_from / fill_into act on a simulator as the reference's do,
  including the noise stream position;
* the oracle pins the two facts the device export relies on: resuming from a
  record continues a run exactly (with the Bold interim window restored), and
  an averaging window written as "running sum in slot 0, zeros elsewhere"
  gives the reference's numpy.mean bit for bit.
"""
import numpy as np
import pytest

from helpers import make_case, oracle_fill, oracle_record
from tvb_hip.checkpoint import N_STOCKS, SimulationHistory

MONS = (("Raw", None, None), ("TemporalAverage", 1.0, None), ("Bold", 20.0, None))


def _case(n=12, seed=3):
    return make_case("MontbrioPazoRoxin", n=n, density=0.4, lengths=(5.0, 40.0), speed=3.0, dt=0.05,
                     scheme="HeunStochastic", cparams=dict(a=0.4), nsig=[0.01, 0.02], seed=seed, monitors=MONS)


def test_field_names_follow_the_reference():
    h = SimulationHistory()
    for f in ("history", "current_state", "current_step", "monitor_names", "integrator_noise_rng_state_algo",
              "integrator_noise_rng_state_keys", "integrator_noise_rng_state_pos",
              "integrator_noise_rng_state_has_gauss", "integrator_noise_rng_state_cached_gauss"):
        assert hasattr(h, f)
    for i in range(1, N_STOCKS + 1):
        assert getattr(h, "monitor_stock_%d" % i) is None
    assert not hasattr(h, "monitor_stock_%d" % (N_STOCKS + 1))
    with pytest.raises(TypeError):
        SimulationHistory(monitor_stock_16=np.zeros(1))


def test_save_load_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    rs = np.random.RandomState(7)
    rs.normal(size=5)
    st = rs.get_state()
    h = SimulationHistory(history=rng.standard_normal((9, 2, 4, 1)).astype(np.float32),
                          current_state=rng.standard_normal((2, 4, 1)), current_step=1234,
                          monitor_stock_1=rng.standard_normal((20, 2, 4, 1)),
                          monitor_stock_3=rng.standard_normal((50, 1, 4, 1)),
                          monitor_names=["TemporalAverage", "Raw", "Bold"],
                          integrator_noise_rng_state_algo=st[0], integrator_noise_rng_state_keys=st[1],
                          integrator_noise_rng_state_pos=st[2], integrator_noise_rng_state_has_gauss=st[3],
                          integrator_noise_rng_state_cached_gauss=st[4])
    h.monitor_interim[3] = rng.standard_normal((80, 1, 4, 1))
    p = tmp_path / "state.npz"
    h.save(p)
    g = SimulationHistory.load(p)
    assert g.history.dtype == np.float32
    np.testing.assert_array_equal(g.history, h.history)
    np.testing.assert_array_equal(g.current_state, h.current_state)
    assert g.current_step == 1234
    np.testing.assert_array_equal(g.monitor_stock_1, h.monitor_stock_1)
    assert g.monitor_stock_2 is None
    np.testing.assert_array_equal(g.monitor_stock_3, h.monitor_stock_3)
    np.testing.assert_array_equal(g.monitor_interim[3], h.monitor_interim[3])
    assert g.monitor_names == h.monitor_names
    assert g.integrator_noise_rng_state_algo == "MT19937"
    np.testing.assert_array_equal(g.integrator_noise_rng_state_keys, st[1])
    assert (g.integrator_noise_rng_state_pos, g.integrator_noise_rng_state_has_gauss) == (st[2], st[3])
    assert g.integrator_noise_rng_state_cached_gauss == st[4]


def test_populate_fill_on_simulator_restores_noise_stream():
    _, sim, _ = _case()
    sim.integrator.noise.random_stream.normal(size=17)
    sim.current_step = 77
    h = SimulationHistory().populate_from(sim)
    assert h.monitor_names == ["Raw", "TemporalAverage", "Bold"]
    assert h.monitor_stock_1 is None                       # Raw keeps no stock
    assert h.monitor_stock_2.shape == (20, 2, 12, 1)       # TemporalAverage (istep, nvoi, N, 1)
    assert h.monitor_stock_3.shape == (5000, 2, 12, 1)     # Bold (S, nvoi, N, 1)
    expect = sim.integrator.noise.random_stream.normal(size=9)
    _, sim2, _ = _case()
    h.fill_into(sim2)
    assert sim2.current_step == 77
    np.testing.assert_array_equal(sim2.integrator.noise.random_stream.normal(size=9), expect)
    np.testing.assert_array_equal(sim2.history.buffer, sim.history.buffer)


@pytest.mark.parametrize("s1", [1000, 1037, 1040])
def test_oracle_resume_is_exact_with_interim(s1):
    """Record at step s1 (mid-window or on a window edge), fill a fresh oracle
    and continue: every monitor output equals the uninterrupted run's when
    the Bold interim window travels with the record; without it only Bold's
    first HRF emissions differ (the reference's own fill_into behaviour)."""
    orc, _, _ = _case()
    o = orc["faithful"]
    s2 = 900
    z = o.legacy_noise(s1 + s2)
    ref = o.run(s1 + s2, noise=z)
    orc_a, _, _ = _case()
    a = orc_a["faithful"]
    a.run(s1, noise=z[:s1])
    rec = oracle_record(a)
    interim = {2: a.monitors[2]._interim_stock.copy()}
    b = oracle_fill(_case()[0]["faithful"], rec, interim)
    out = b.run(s2, noise=z[s1:])
    for m in range(3):
        t_ref, x_ref = ref[m]
        t, x = out[m]
        assert len(t) > 0
        k0 = len(t_ref) - len(t)
        np.testing.assert_array_equal(t, t_ref[k0:])
        np.testing.assert_array_equal(x, x_ref[k0:])
    c = oracle_fill(_case()[0]["faithful"], rec)            # interim window lost
    out_c = c.run(s2, noise=z[s1:])
    np.testing.assert_array_equal(out_c[1][1], out[1][1])   # TemporalAverage unaffected
    np.testing.assert_array_equal(out_c[0][1], out[0][1])   # Raw unaffected
    if s1 % 80:
        assert not np.array_equal(out_c[2][1], out[2][1])


@pytest.mark.parametrize("k", [0, 1, 7, 19])
def test_running_sum_in_slot0_gives_numpy_mean_bit_for_bit(k):
    """An averaging window whose first k samples are replaced by their
    sequential sum in slot 0 and zeros: numpy.mean over axis 0 (as
    TemporalAverage.sample and Bold.sample take it, monitors.py:389, :947)
    returns the same bits once the remaining slots are filled."""
    rng = np.random.default_rng(k)
    istep = 20
    for _ in range(50):
        x = rng.standard_normal((istep, 2, 7, 1)) * 10 ** rng.uniform(-3, 3)
        ref = np.mean(x, axis=0)
        y = np.zeros_like(x)
        s = np.zeros(x.shape[1:])
        for j in range(k):
            s = s + x[j]
        y[0] = s
        y[k:] = x[k:]
        np.testing.assert_array_equal(np.mean(y, axis=0), ref)
