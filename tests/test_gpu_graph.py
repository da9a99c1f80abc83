"""hipGraph replay of split calls: the block launches of a split plan's
call (step kernel + history commit per block) are captured when a call
repeats the previous call's length and arguments, and replayed while calls
repeat them; every call must give the bits of the same calls launched
directly (TVBK_GRAPH=0), with the host's own reads between the replays."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _plan(graph_env):
    import bench
    from tvb_hip import _lib as L
    from tvb_hip.plan import CSR, MonitorSpec, Plan
    w, Lm, d, H, ic, rng = bench.c1_setup()
    old = {k: os.environ.get(k) for k in ("TVBK_SPLIT", "TVBK_GRAPH")}
    os.environ["TVBK_SPLIT"] = "4"
    if graph_env is None:
        os.environ.pop("TVBK_GRAPH", None)
    else:
        os.environ["TVBK_GRAPH"] = graph_env
    try:
        p = np.array([[1.0, 0.0, 1.0, 15.0, -5.0, 0.0, 1.0, 0.0, 0.5, 0.0, 0.01, 0.02]])
        plan = Plan(model=L.MODEL_MPR, cfun=L.CFUN_LINEAR, scheme=L.HEUN_STO, precision=L.MIXED,
                    noise=L.NOISE_PHILOX, conns=[CSR(w, d)], params=p, dt=0.05, obs_coef=np.eye(2),
                    monitors=[MonitorSpec(L.MON_TAVG, 20, [0, 1], period=1.0), bench.make_bold_spec(0.05, 20.0)],
                    max_steps=2000, seed=77)
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    plan.set_history(0, ic, ic[-1], H - 1)
    return plan


def test_graph_replays_equal_direct_launches():
    g, dct = _plan(None), _plan("0")
    try:
        ntile, k = g.tiles()
        assert k > 0 and ntile >= 4
        _say("split plan: tiles", ntile, "block", k)
        for call, n in enumerate((1000, 1000, 1000, 700, 700, 1000)):
            outs = []
            for p in (g, dct):
                p.run(n)
                outs.append([p.read(m) for m in range(2)] + [p.export(0, np.float32)])
            for (a, b) in zip(outs[0][:2], outs[1][:2]):
                (ta, xa), = a
                (tb, xb), = b
                np.testing.assert_array_equal(ta, tb)
                np.testing.assert_array_equal(xa, xb)
            for u, v in zip(outs[0][2], outs[1][2]):
                np.testing.assert_array_equal(u, v)
            _say("  call", call, n, "steps: graph = direct")
        cap, launched = g.graphs()
        # direct, captured, replayed; 700 direct, captured; 1000 direct
        assert (cap, launched) == (2, 3), (cap, launched)
        assert dct.graphs() == (0, 0)
    finally:
        g.close()
        dct.close()
