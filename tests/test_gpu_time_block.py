"""GPU checks of the time-blocked sequential loop: the coupling of a block
of up to d_min steps is gathered at once, then every node is integrated
through the block.  The blocking changes when sums are formed, never their
arithmetic: every output and final state must be bit-identical to the same
plan stepping one step at a time (TVBK_TIME_BLOCK=0), and the oracle parity
of that path carries over (checked directly on one case).
"""
import os

import numpy as np
import pytest

from helpers import make_case, rel_err
from tvb_hip import sim as S

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _with_tb(value, fn):
    old = os.environ.get("TVBK_TIME_BLOCK")
    if value is None:
        os.environ.pop("TVBK_TIME_BLOCK", None)
    else:
        os.environ["TVBK_TIME_BLOCK"] = str(value)
    try:
        return fn()
    finally:
        if old is None:
            os.environ.pop("TVBK_TIME_BLOCK", None)
        else:
            os.environ["TVBK_TIME_BLOCK"] = old


def _block_of(sim, **kw):
    from tvb_hip.backend import HipBackend
    plan = HipBackend().build_plan(sim, **kw)
    k = plan.time_block()
    plan.close()
    return k


MONS = (("TemporalAverage", 1.0, None), ("Bold", 20.0, None), ("SubSample", 0.5, None))


@pytest.mark.parametrize("precision", ["mixed", "fp32", "fp64"])
def test_single_instance_blocked_equals_stepwise_and_oracle(precision):
    """MPR with a region stimulus, three launches of uneven length (blocks
    cut by the launch ends)."""
    from tvb_hip.backend import HipBackend
    _say("time block single", precision)
    n = 40
    teq = S.PulseTrain(parameters={"T": 15.0, "tau": 5.0, "amp": 0.5, "onset": 3.0})
    stim = (np.where(np.arange(n) % 4 == 0, 1.0, 0.0), teq, 60.0)
    outs = []
    for tb in (None, 0):
        orc, sim, _ = make_case("MontbrioPazoRoxin", n=n, density=0.3, lengths=(10.0, 60.0), speed=3.0, dt=0.05,
                                scheme="HeunStochastic", cparams=dict(a=0.3), nsig=[0.01, 0.02], seed=31,
                                monitors=MONS, stimulus=stim)
        k = _with_tb(tb, lambda: _block_of(sim, precision=precision))
        assert (k > 1) if tb is None else (k == 0), k
        o = orc["faithful" if precision == "mixed" else "fp64"]
        z = o.legacy_noise(1000)
        out = _with_tb(tb, lambda: HipBackend().run_sim(sim, nstep=1000, precision=precision, noise=z, chunk=333))
        outs.append((out, sim.current_state.copy(), sim.history.buffer.copy()))
    (o1, x1, h1), (o0, x0, h0) = outs
    for m, ((t1, y1), (t0, y0)) in enumerate(zip(o1, o0)):
        np.testing.assert_array_equal(t1, t0)
        np.testing.assert_array_equal(y1, y0, err_msg="monitor %d" % m)
    np.testing.assert_array_equal(x1, x0)
    np.testing.assert_array_equal(h1, h0)
    ref = o.run(1000, noise=z)
    for m, ((t, y), (tr, yr)) in enumerate(zip(o1, ref)):
        np.testing.assert_allclose(t, tr, rtol=0, atol=1e-9)
        if precision == "mixed" and m != 1:
            np.testing.assert_array_equal(y, yr, err_msg="monitor %d" % m)
        elif precision == "mixed":
            assert rel_err(y, yr) <= 1e-13
        elif precision == "fp64":
            assert rel_err(y, yr) <= 1e-10, (m, rel_err(y, yr))


def test_area_monitors_blocked_match_oracle():
    """GlobalAverage and SpatialAverage (node samples reduced after the
    launch) through the time-blocked loop: bit-exact against the oracle's
    reference-faithful mode (SpatialAverage to 1e-14: the reference's dot
    product runs through BLAS)."""
    from tvb_hip.backend import HipBackend
    _say("time block area monitors")
    n = 24
    mons = (("GlobalAverage", 0.5, None), ("SpatialAverage", 1.0, None, dict(spatial_mask=np.arange(n) % 3)),
            ("TemporalAverage", 1.0, None))
    orc, sim, _ = make_case("MontbrioPazoRoxin", n=n, density=0.3, lengths=(10.0, 60.0), speed=3.0, dt=0.05,
                            scheme="HeunStochastic", cparams=dict(a=0.2), nsig=[0.01, 0.02], seed=41, monitors=mons)
    assert _block_of(sim, precision="mixed") > 1
    o = orc["faithful"]
    z = o.legacy_noise(900)
    ref = o.run(900, noise=z)
    out = HipBackend().run_sim(sim, nstep=900, precision="mixed", noise=z, chunk=400)
    for m, ((t, y), (tr, yr)) in enumerate(zip(out, ref)):
        np.testing.assert_allclose(t, tr, rtol=0, atol=1e-9)
        if m == 1:
            assert rel_err(y, yr) <= 1e-14
        else:
            np.testing.assert_array_equal(y, yr, err_msg="monitor %d" % m)


def test_short_delays_give_short_blocks():
    """Blocks are at most d_min steps long; d_min = 1 turns the blocking off."""
    _say("time block short delays")
    _, sim2, _ = make_case("Generic2dOscillator", n=12, density=0.5, lengths=(5.0, 40.0), speed=50.0, dt=0.05,
                           seed=3)
    d = sim2.connectivity.idelays[sim2.connectivity.weights != 0]
    assert 2 <= d.min() <= 3, d.min()
    assert _block_of(sim2, precision="mixed") == d.min()
    # the same connectome with one tract of 2.5 mm: a delay of one step
    w = sim2.connectivity.weights.copy()
    L = sim2.connectivity.tract_lengths.copy()
    i, j = np.argwhere(w != 0)[0]
    L[i, j] = L[j, i] = 2.5
    _, sim1, _ = make_case("Generic2dOscillator", weights=w, tract_lengths=L, speed=50.0, dt=0.05, seed=3)
    d = sim1.connectivity.idelays[sim1.connectivity.weights != 0]
    assert d.min() == 1, d.min()
    assert _block_of(sim1, precision="mixed") == 0
    from tvb_hip.backend import HipBackend
    res = []
    for tb in (None, 0):
        _, sim, _ = make_case("Generic2dOscillator", n=12, density=0.5, lengths=(5.0, 40.0), speed=50.0, dt=0.05,
                              seed=3, monitors=(("Raw", None, None), ("TemporalAverage", 0.5, None)))
        res.append(_with_tb(tb, lambda: HipBackend().run_sim(sim, nstep=700, precision="mixed", chunk=250)))
    for (t1, y1), (t0, y0) in zip(*res):
        np.testing.assert_array_equal(y1, y0)


def test_batch_lane_groups_blocked_equals_stepwise():
    """A batch-lane plan with lane groups (tile 16, one coupled variable:
    the CSR-record gather) and the uneven JansenRit batch with split
    instances, blocked vs stepwise."""
    from tvb_hip.backend import HipBackend
    _say("time block batch")
    for label in ("g2d-batch", "jr-split"):
        outs = []
        for tb in (None, 0):
            if label == "g2d-batch":
                cases = [make_case("Generic2dOscillator", n=20, density=0.3, lengths=(10.0, 60.0), speed=3.0,
                                   dt=0.05, scheme="HeunStochastic", cparams=dict(a=0.01 + 0.002 * b),
                                   nsig=[1e-3, 1e-3], seed=5,
                                   monitors=(("TemporalAverage", 1.0, None), ("Bold", 20.0, None)))
                         for b in range(20)]
                kw = dict(kernel="batch")
                nstep = 900
            else:
                cases = [make_case("JansenRit", n=n, density=0.2, lengths=(15.0, 60.0), speed=3.0, dt=0.1,
                                   scheme="EulerStochastic", coupling="SigmoidalJansenRit",
                                   nsig=[0, 0, 0, 1e-3, 1e-3, 1e-3], seed=sd, spread=0.05,
                                   monitors=(("TemporalAverage", 2.0, None), ("Bold", 20.0, None)))
                         for n, sd in ((60, 1), (150, 2), (96, 3), (220, 4))]
                kw = dict(kernel="edge")
                nstep = 700
            sims = [c[1] for c in cases]

            def run():
                os.environ["TVBK_SPLIT"] = "3" if label == "jr-split" else "0"
                try:
                    be = HipBackend()
                    plan = be.build_batch_plan(sims, precision="fp32", noise="philox", max_steps=300, **kw)
                    k = plan.time_block()
                    res = [[] for _ in plan.monitors]
                    done = 0
                    while done < nstep:
                        n = min(300, nstep - done)
                        plan.run(n)
                        for m in range(len(plan.monitors)):
                            res[m].append([x.copy() for _, x in plan.read(m)])
                        done += n
                    states = [plan.export(b) for b in range(len(sims))]
                    plan.close()
                    return k, res, states
                finally:
                    del os.environ["TVBK_SPLIT"]
            outs.append(_with_tb(tb, run))
        (k1, r1, s1), (k0, r0, s0) = outs
        _say("  ", label, "blocks", k1, k0)
        assert k1 > 1 and k0 == 0
        for m in range(len(r0)):
            for c in range(len(r0[m])):
                for b in range(len(r0[m][c])):
                    np.testing.assert_array_equal(r1[m][c][b], r0[m][c][b],
                                                  err_msg="%s monitor %d call %d inst %d" % (label, m, c, b))
        for (h1, x1, q1), (h0, x0, q0) in zip(s1, s0):
            assert q1 == q0
            np.testing.assert_array_equal(h1, h0)
            np.testing.assert_array_equal(x1, x0)


@pytest.mark.parametrize("case", ["one-node", "no-edges"])
def test_connectome_without_edges(case):
    """A one-node simulator and a connectome of zero weights (horizon 1):
    no time blocks, and the oracle's trajectory bit for bit."""
    from tvb_hip.backend import HipBackend
    _say("no edges", case)
    if case == "one-node":
        orc, sim, _ = make_case("MontbrioPazoRoxin", n=1, dt=0.05, scheme="HeunStochastic", nsig=[0.01, 0.02],
                                seed=3, monitors=(("TemporalAverage", 1.0, None), ("Raw", None, None)))
    else:
        w = np.zeros((6, 6))
        orc, sim, _ = make_case("Generic2dOscillator", weights=w, tract_lengths=w.copy(), dt=0.1, seed=3,
                                monitors=(("Raw", None, None), ("TemporalAverage", 0.5, None)))
    assert sim.connectivity.horizon == 1
    assert _block_of(sim, precision="mixed") == 0
    o = orc["faithful"]
    z = o.legacy_noise(300) if case == "one-node" else None
    ref = o.run(300, noise=z)
    out = HipBackend().run_sim(sim, nstep=300, precision="mixed", noise=z if z is not None else "legacy", chunk=120)
    for m, ((t, y), (tr, yr)) in enumerate(zip(out, ref)):
        np.testing.assert_allclose(t, tr, rtol=0, atol=1e-9)
        np.testing.assert_array_equal(y, yr, err_msg="monitor %d" % m)


def test_uniform_rows_blocked_equals_stepwise_and_oracle():
    """Tiles of 64 instances with two coupled variables (wave-uniform rows,
    too wide for the pipelined loop): the time-blocked loop gathers a block's
    rows for all its steps; bit-identical to stepping and, for the checked
    instances, to the oracle (MIXED)."""
    from tvb_hip import _lib as L
    from tvb_hip.backend import HipBackend
    _say("time block uniform rows")
    cases = [make_case("MontbrioPazoRoxin", n=20, density=0.3, lengths=(10.0, 60.0), speed=3.0, dt=0.05,
                       scheme="HeunStochastic", cparams=dict(a=0.1 + 0.005 * b), nsig=[0.01, 0.02], seed=8,
                       monitors=(("TemporalAverage", 1.0, None), ("Bold", 20.0, None))) for b in range(80)]
    sims = [c[1] for c in cases]
    zs = [c[0]["faithful"].legacy_noise(700) for c in cases]
    outs = []
    for tb in (None, 0):
        def run():
            plan = HipBackend().build_batch_plan(sims, precision="mixed", noise="inject", max_steps=350,
                                                 kernel="batch", tile=64)
            assert plan.kernel == L.KERNEL_BATCH_LANE and plan.tile == 64
            k = plan.time_block()
            res = [[] for _ in plan.monitors]
            for k0 in (0, 350):
                plan.set_noise(np.stack([zz[k0:k0 + 350] for zz in zs], axis=-1))
                plan.run(350)
                for m in range(len(plan.monitors)):
                    res[m].append(plan.read(m))
            states = [plan.export(b) for b in (0, 70)]
            plan.close()
            return k, res, states
        outs.append(_with_tb(tb, run))
    (k1, r1, s1), (k0, r0, s0) = outs
    assert k1 > 1 and k0 == 0, (k1, k0)
    for m in range(len(r0)):
        for c in range(2):
            for b in range(80):
                np.testing.assert_array_equal(r1[m][c][b][1], r0[m][c][b][1], err_msg="monitor %d inst %d" % (m, b))
    for (h1, x1, q1), (h0, x0, q0) in zip(s1, s0):
        np.testing.assert_array_equal(h1, h0)
        np.testing.assert_array_equal(x1, x0)
    for b in (0, 70):
        ref = cases[b][0]["faithful"].run(700, noise=zs[b])
        tavg = np.concatenate([r1[0][0][b][1], r1[0][1][b][1]])
        np.testing.assert_array_equal(tavg[..., np.newaxis], ref[0][1], err_msg="inst %d" % b)
        bold = np.concatenate([r1[1][0][b][1], r1[1][1][b][1]])
        assert rel_err(bold[..., np.newaxis], ref[1][1]) <= 1e-13
