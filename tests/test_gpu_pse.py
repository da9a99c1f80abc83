"""GPU parity of the parameter-space adapter (tvb_hip.pse.PSE): every grid
point of one batched launch against the oracle run with that point's
parameters and that point's own noise stream (as its own framework operation
would draw it).

Tolerances: mixed device precision vs the reference-faithful oracle,
TemporalAverage bit-exact, Bold max rel <= 1e-13 (HRF sum order).
"""
import numpy as np
import pytest

from helpers import initial_history, make_case, rel_err
from tvb_hip.pse import PSE, Range, RangeParameter

pytestmark = pytest.mark.gpu

MONS = (("TemporalAverage", 1.0, None), ("Bold", 20.0, None))


def _say(*a):
    print(*a, flush=True)


# one initial history long enough for the slowest conduction speed tested:
# every simulator keeps its last `horizon` samples
IC = initial_history("MontbrioPazoRoxin", 450, 14, 77)


def _case(a=0.2, eta=-5.0, speed=3.0):
    return make_case("MontbrioPazoRoxin", n=14, density=0.4, lengths=(5.0, 40.0), speed=speed, dt=0.05,
                     scheme="HeunStochastic", cparams=dict(a=a), mparams=dict(eta=eta), nsig=[0.01, 0.02],
                     seed=3, monitors=MONS, initial_conditions=IC)


def _compare(res, orc, nstep):
    o = orc["faithful"]
    ref = o.run(nstep, noise=o.legacy_noise(nstep))
    (tt, xt), (tb, xb) = ref
    (dt_, dx), (bt, bx) = res
    np.testing.assert_allclose(dt_, tt, rtol=0, atol=1e-9)
    np.testing.assert_array_equal(dx, xt)
    e = rel_err(bx, xb)
    assert e <= 1e-13, e


@pytest.mark.parametrize("chunk", [2000, 700])
def test_pse_grid_one_launch_matches_per_point_oracle(chunk):
    _say("pse grid", chunk)
    _, base, _ = _case()
    r1 = RangeParameter("coupling.a", float, Range(0.1, 0.5, 0.1), is_array=True)
    r2 = RangeParameter("model.eta", float, range_values=[-5.0, -4.6, -4.2])
    pse = PSE(base, r1, r2)
    nstep = 2000
    out = pse.run(nstep=nstep, precision="mixed", chunk=chunk)
    assert len(out) == len(pse.points) == 12
    for k in (0, 5, 11):
        ranges, res = out[k]
        orc, _, _ = _case(a=float(ranges["coupling.a"]), eta=float(ranges["model.eta"]))
        _compare(res, orc, nstep)


def test_pse_conduction_speed_gives_per_point_connectomes():
    _say("pse speed")
    _, base, _ = _case()
    pse = PSE(base, RangeParameter("conduction_speed", float, range_values=[2.0, 3.0, 7.0]))
    nstep = 1600
    out = pse.run(nstep=nstep, precision="mixed")
    for (ranges, res) in out:
        orc, _, _ = _case(speed=float(ranges["conduction_speed"]))
        _compare(res, orc, nstep)


def test_pse_many_points_use_the_batch_lane_kernel():
    from tvb_hip import _lib as L
    _say("pse batch lane")
    _, base, _ = _case()
    pse = PSE(base, RangeParameter("coupling.a", float, Range(0.0, 0.8, 0.02), is_array=True))
    assert len(pse.points) == 40
    nstep = 800
    out = pse.run(nstep=nstep, precision="mixed")
    assert pse.last_plan_shape[0] == L.KERNEL_BATCH_LANE
    for k in (0, 23, 39):
        ranges, res = out[k]
        orc, _, _ = _case(a=float(ranges["coupling.a"]))
        _compare(res, orc, nstep)
