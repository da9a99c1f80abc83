"""Host-side checks of the parameter-space adapter (no GPU): the grid order,
the ranges record and the values set on each point's simulator follow the
framework's SimulatorService (simulator_service.py:88-93, :136-143, :193-227)."""
import json

import numpy as np
import pytest

from helpers import make_case
from tvb_hip.pse import PSE, LinspaceRange, Range, RangeParameter, set_range_parameter


def _base():
    _, sim, _ = make_case("MontbrioPazoRoxin", n=10, density=0.4, lengths=(5.0, 40.0), speed=3.0, dt=0.05,
                          scheme="HeunStochastic", cparams=dict(a=0.2), nsig=[0.01, 0.02], seed=3,
                          monitors=(("TemporalAverage", 1.0, None),))
    return sim


def test_range_values_follow_numpy_arange_and_array_wrapping():
    r = RangeParameter("coupling.a", float, Range(0.1, 0.5, 0.1))
    np.testing.assert_array_equal(r.get_range_values(), np.arange(0.1, 0.5, 0.1))
    r2 = RangeParameter("coupling.a", float, Range(0.0, 1.0, 0.25), is_array=True)
    vals = r2.get_range_values()
    assert all(isinstance(v, np.ndarray) and v.shape == (1,) for v in vals)
    assert json.loads(r.to_json()) == ["coupling.a", {"lo": 0.1, "step": 0.1, "hi": 0.5}]
    np.testing.assert_array_equal(LinspaceRange(0, 1, 5).to_array(), np.linspace(0, 1, 5))


def test_grid_order_ranges_and_values():
    sim = _base()
    r1 = RangeParameter("coupling.a", float, Range(0.1, 0.4, 0.1), is_array=True)
    r2 = RangeParameter("model.eta", float, range_values=[-5.0, -4.5])
    pse = PSE(sim, r1, r2)
    a_vals = np.arange(0.1, 0.4, 0.1)              # 4 values: numpy's arange length rounds up
    assert len(pse.points) == len(a_vals) * 2
    expect = [(a, e) for a in a_vals for e in (-5.0, -4.5)]   # param1 outer
    for (ranges, s), (a, e) in zip(pse.points, expect):
        assert ranges == {"coupling.a": a, "model.eta": e}
        assert float(np.asarray(s.coupling.a).reshape(-1)[0]) == a
        assert float(np.asarray(s.model.eta).reshape(-1)[0]) == e
        assert s.current_step == sim.current_step      # configured from step 0 like the original
        assert s is not sim and s.coupling is not sim.coupling
    assert json.loads(pse.ranges[0]) == {"coupling.a": expect[0][0], "model.eta": -5.0}
    # the base simulator is untouched
    assert float(np.asarray(sim.coupling.a).reshape(-1)[0]) == 0.2


def test_conduction_speed_changes_the_delays():
    sim = _base()
    pse = PSE(sim, RangeParameter("conduction_speed", float, range_values=[2.0, 6.0]))
    d0 = pse.points[0][1].connectivity.idelays
    d1 = pse.points[1][1].connectivity.idelays
    assert d0.max() > d1.max()
    np.testing.assert_array_equal(d0, np.rint(pse.points[0][1].connectivity.tract_lengths / 2.0 / 0.05))


def test_each_point_restarts_its_noise_stream():
    sim = _base()
    pse = PSE(sim, RangeParameter("coupling.a", float, range_values=[0.1, 0.3]))
    z = [s.integrator.noise.random_stream.normal(size=5) for _, s in pse.points]
    np.testing.assert_array_equal(z[0], z[1])
    np.testing.assert_array_equal(z[0], np.random.RandomState(42).normal(size=5))


def test_dotted_setter():
    sim = _base()
    set_range_parameter(sim, "integrator.noise.nsig", np.array([0.5, 0.5]))
    np.testing.assert_array_equal(sim.integrator.noise.nsig, [0.5, 0.5])
    with pytest.raises(AttributeError):
        set_range_parameter(sim, "nope.x", 1.0)
