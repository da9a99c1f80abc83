"""GPU checks of the two-step pipelined loop (TVBK_LOOP_PIPELINED2): rows of
32 values (n_channel x tile), every delay >= 3, up to two state variables;
one wave load gathers two consecutive steps of an edge, the upper half-wave
reading the next ring slot
(the slot after H-1 is a mirror of slot 0).  The loop changes which lanes
gather and when, never the arithmetic: every output and final state must be
bit-identical to the one-step pipelined loop (TVBK_PIPE2=0), over calls of
odd and even length (a pass that integrates one step), across ring wraps,
and, for checked instances, to the oracle's reference-precision mode.
"""
import os

import numpy as np
import pytest

from helpers import make_case, rel_err

pytestmark = pytest.mark.gpu

B = 64
CALLS = (351, 1, 2, 348)
MONS = (("TemporalAverage", 0.15, None), ("Bold", 20.0, None))


def _say(*a):
    print(*a, flush=True)


def _case(lengths, seed, a=0.3, eta=-5.0):
    return make_case("MontbrioPazoRoxin", n=20, density=0.3, lengths=lengths, speed=3.0, dt=0.05,
                     scheme="HeunStochastic", cparams=dict(a=a), mparams=dict(eta=eta), nsig=[0.01, 0.02],
                     seed=seed, monitors=MONS)


def _run(sim, overrides, precision, tile, skip, z, pipe2):
    from tvb_hip.backend import HipBackend
    old = os.environ.get("TVBK_PIPE2")
    if pipe2:
        os.environ.pop("TVBK_PIPE2", None)
    else:
        os.environ["TVBK_PIPE2"] = "0"
    try:
        plan = HipBackend().build_sweep_plan(sim, overrides, precision=precision, noise="inject",
                                             max_steps=max(CALLS), kernel="batch", tile=tile,
                                             skip_zero_coupling=skip)
    finally:
        if old is None:
            os.environ.pop("TVBK_PIPE2", None)
        else:
            os.environ["TVBK_PIPE2"] = old
    try:
        assert plan.tile == tile and plan.nch * tile == 32, (plan.tile, plan.nch)
        res = [[] for _ in plan.monitors]
        k0 = 0
        loops = set()
        for n in CALLS:
            plan.set_noise(z[k0:k0 + n])
            plan.run(n)
            loops.add(plan.loop())
            for m in range(len(plan.monitors)):
                res[m].append([y.copy() for _, y in plan.read(m)])
            k0 += n
        states = [plan.export(b) for b in (0, 37, B - 1)]
        assert plan.flags().sum() == 0
    finally:
        plan.close()
    return loops, res, states


@pytest.mark.parametrize("label,tile,skip,vary,precision,lengths", [
    ("mpr-1ch-plan-constant-model", 32, True, False, "mixed", (5.0, 30.0)),
    ("mpr-1ch-per-lane-model", 32, True, True, "mixed", (5.0, 30.0)),
    ("mpr-2ch-tile16", 16, False, False, "mixed", (5.0, 30.0)),
    ("mpr-1ch-fp32", 32, True, False, "fp32", (5.0, 30.0)),
    ("mpr-1ch-short-delays", 32, True, False, "mixed", (0.3, 3.0)),
])
def test_two_step_equals_one_step_and_oracle(label, tile, skip, vary, precision, lengths):
    orc, sim, _ = _case(lengths, seed=12)
    idl = sim.connectivity.idelays[sim.connectivity.weights != 0]
    dmin, H = int(idl.min()), sim.connectivity.horizon
    G = np.linspace(0.05, 0.6, B)
    eta = np.linspace(-5.5, -4.5, B) if vary else np.full(B, -5.0)
    overrides = {"coupling.a": G}
    if vary:
        overrides["model.eta"] = eta
    rng = np.random.default_rng(7)
    z = rng.standard_normal((sum(CALLS), 2, 20, B))
    _say("two-step loop:", label, "d_min", dmin, "H", H)
    l2, r2, s2 = _run(sim, overrides, precision, tile, skip, z, True)
    l1, r1, s1 = _run(sim, overrides, precision, tile, skip, z, False)
    assert l1 == {"pipelined"}, l1
    assert l2 == ({"pipelined-2"} if dmin >= 3 else {"pipelined"}), (l2, dmin)
    if dmin >= 3:
        assert sum(CALLS) > 2 * H            # the ring wraps: slot 0 and its mirror are rewritten
    for m in range(len(r1)):
        for c in range(len(CALLS)):
            for b in range(B):
                np.testing.assert_array_equal(r2[m][c][b], r1[m][c][b],
                                              err_msg="%s monitor %d call %d inst %d" % (label, m, c, b))
    for (h2, x2, q2), (h1, x1, q1) in zip(s2, s1):
        assert q2 == q1
        np.testing.assert_array_equal(h2, h1)
        np.testing.assert_array_equal(x2, x1)
    if precision != "mixed" or l2 != {"pipelined-2"}:
        return
    for b in (0, 37):
        o, _, _ = _case(lengths, seed=12, a=float(G[b]), eta=float(eta[b]))
        ref = o["faithful"].run(sum(CALLS), noise=z[..., b])
        tavg = np.concatenate([y for y in (r2[0][c][b] for c in range(len(CALLS))) if y.size])
        np.testing.assert_array_equal(tavg[..., np.newaxis], ref[0][1], err_msg="inst %d" % b)
        bold = np.concatenate([y for y in (r2[1][c][b] for c in range(len(CALLS))) if y.size])
        assert bold.shape[0] == ref[1][1].shape[0] >= 1
        assert rel_err(bold[..., np.newaxis], ref[1][1]) <= 1e-13
