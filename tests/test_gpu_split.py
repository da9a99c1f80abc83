"""GPU checks of edge-lane instances split over several workgroups by node
range (blocks of K <= d_min steps, staged history committed between blocks).

The split changes which workgroup integrates a node and when history
reaches the ring, never the arithmetic: every output and final state must be
bit-identical to the same plan unsplit (TVBK_SPLIT=0), and the oracle
parity of the unsplit path carries over (checked directly on one case).
"""
import os

import numpy as np
import pytest

from helpers import make_case, rel_err

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _with_split(value, fn):
    old = os.environ.get("TVBK_SPLIT")
    os.environ["TVBK_SPLIT"] = str(value)
    try:
        return fn()
    finally:
        if old is None:
            del os.environ["TVBK_SPLIT"]
        else:
            os.environ["TVBK_SPLIT"] = old


def _jr_cases():
    specs = [(60, 400, 1), (150, 700, 2), (96, 250, 3), (220, 600, 4)]
    sims, zs = [], []
    for n, steps, sd in specs:
        _, sim, _ = make_case("JansenRit", n=n, density=0.2, lengths=(15.0, 60.0), speed=3.0, dt=0.1,
                              scheme="EulerStochastic", coupling="SigmoidalJansenRit",
                              nsig=[0, 0, 0, 1e-3, 1e-3, 1e-3], seed=sd, spread=0.05,
                              monitors=(("TemporalAverage", 2.0, None), ("Bold", 20.0, None),
                                        ("AfferentCouplingTemporalAverage", 1.0, None)))
        sims.append(sim)
        zs.append(np.random.default_rng(sd).standard_normal((700, 6, n)))
    return specs, sims, zs


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
@pytest.mark.parametrize("parts", [2, 3])
def test_split_uneven_batch_bit_identical(precision, parts):
    from tvb_hip import _lib as L
    from tvb_hip.backend import HipBackend
    _say("split uneven", precision, parts)
    outs, shapes = [], []
    for split in (0, parts):
        specs, sims, zs = _jr_cases()

        def run():
            be = HipBackend()
            plan = be.build_batch_plan(sims, precision=precision, noise="inject", max_steps=300, kernel="edge")
            shapes.append(plan.tiles())
            plan.set_steps(np.array([s for _, s, _ in specs], np.int64))
            res = [[] for _ in plan.monitors]
            done = 0
            while done < 700:                               # three calls of up to 300 steps
                n = min(300, 700 - done)
                z = np.zeros((n, 6, 220, len(sims)))
                for b, zz in enumerate(zs):
                    z[:, :, :zz.shape[2], b] = zz[done:done + n]
                plan.set_noise(z)
                plan.run(n)
                for m in range(len(plan.monitors)):
                    res[m].append([x.copy() for _, x in plan.read(m)])
                done += n
            states = [plan.export(b) for b in range(len(sims))]
            plan.close()
            return res, states
        outs.append(_with_split(split, run))
    assert shapes[0][1] == 0 and shapes[1][1] > 0, shapes
    assert shapes[1][0] > shapes[0][0]
    (r0, s0), (r1, s1) = outs
    for m in range(len(r0)):
        for c in range(len(r0[m])):
            for b in range(4):
                np.testing.assert_array_equal(r1[m][c][b], r0[m][c][b], err_msg="monitor %d call %d inst %d" % (m, c, b))
    for (h0, x0, k0), (h1, x1, k1) in zip(s0, s1):
        assert k0 == k1
        np.testing.assert_array_equal(h1, h0)
        np.testing.assert_array_equal(x1, x0)


def test_split_single_simulator_matches_oracle():
    """A split MPR simulator (run_sim, mixed precision, stimulus and afferent
    monitor) against the oracle: bit-exact as unsplit."""
    from tvb_hip import sim as S
    from tvb_hip.backend import HipBackend
    _say("split single vs oracle")
    n = 120
    teq = S.PulseTrain(parameters={"T": 15.0, "tau": 5.0, "amp": 0.5, "onset": 3.0})
    mons = (("TemporalAverage", 1.0, None), ("AfferentCoupling", None, None), ("Bold", 20.0, None))
    orc, sim, _ = make_case("MontbrioPazoRoxin", n=n, density=0.15, lengths=(10.0, 60.0), speed=2.0, dt=0.05,
                            scheme="HeunStochastic", cparams=dict(a=0.3), nsig=[0.01, 0.02], seed=12, monitors=mons,
                            stimulus=(np.where(np.arange(n) % 5 == 0, 1.0, 0.0), teq, 60.0))
    o = orc["faithful"]
    z = o.legacy_noise(1200)
    ref = o.run(1200, noise=z)
    be = HipBackend()
    out = _with_split(4, lambda: be.run_sim(sim, nstep=1200, precision="mixed", noise=z, chunk=500))
    for m, ((t, x), (tr, xr)) in enumerate(zip(out, ref)):
        np.testing.assert_allclose(t, tr, rtol=0, atol=1e-9)
        if m == 2:
            assert rel_err(x, xr) <= 1e-13
        else:
            np.testing.assert_array_equal(x, xr, err_msg="monitor %d" % m)


def test_automatic_split_of_a_large_instance():
    """Without the override, one large connectome among small ones is split
    automatically, with outputs equal to the unsplit plan."""
    from tvb_hip.backend import HipBackend
    _say("auto split")
    res = []
    for split in ("0", None):
        cases = [make_case("MontbrioPazoRoxin", n=n, density=0.3, lengths=(10.0, 80.0), speed=3.0, dt=0.05,
                           scheme="HeunDeterministic", cparams=dict(a=0.1), seed=20 + k,
                           monitors=(("TemporalAverage", 1.0, None),)) for k, n in enumerate([400, 40, 40, 40])]
        sims = [s for _, s, _ in cases]

        def run():
            be = HipBackend()
            out = be.run_batch(sims, 400, precision="mixed", kernel="edge")
            return out, be
        if split is None:
            old = os.environ.pop("TVBK_SPLIT", None)
            try:
                out, be = run()
            finally:
                if old is not None:
                    os.environ["TVBK_SPLIT"] = old
        else:
            out, be = _with_split(split, run)
        res.append(out)
    for b in range(4):
        np.testing.assert_array_equal(res[1][b][0][1], res[0][b][0][1])
