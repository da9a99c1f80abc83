"""Batch-lane tiles split by node range: a tile of instances sharing one
connectome runs on several workgroups, each holding its node range in LDS,
stepping in time blocks with staged history commits.

* forced on a sweep that fits unsplit (TVBK_BATCH_SPLIT=3, against =0):
  bit-identical to the unsplit plan, over several calls, in both
  precisions and with the V-channel skip;
* automatic on a sweep of few tiles (64 instances): split, and
  bit-identical to the unsplit plan;
* automatic on a 400-node connectome whose tiles do not fit LDS: sampled
  instances bit-exact against the reference-faithful oracle fed the device's
  normals, and every instance bit-identical to the edge-lane kernel."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _env(key, value):
    old = os.environ.get(key)
    if value is None:
        os.environ.pop(key, None)
    else:
        os.environ[key] = value
    return old


@pytest.mark.parametrize("prec,skip", [("mixed", False), ("mixed", True), ("f32", False)])
def test_forced_batch_split_equals_unsplit(prec, skip):
    import bench
    from tvb_hip import _lib as L
    P = {"mixed": L.MIXED, "f32": L.F32}[prec]
    w, Lm, d, H, G, sig = bench.c2_setup(256)
    G, sig = G[::4][:64], sig[::4][:64]
    outs = []
    for env in ("0", "3", None):
        old = _env("TVBK_BATCH_SPLIT", env)
        try:
            plan = bench.build_c2_plan(w, d, G, sig, 0.05, 900, P, L.NOISE_PHILOX, 3, bold=20.0, skip=skip,
                                       kernel=L.KERNEL_BATCH_LANE)
        finally:
            _env("TVBK_BATCH_SPLIT", old)
        try:
            plan.init_uniform(11, [0.0, -2.0], [2.0, 1.5])
            ntile, k = plan.tiles()
            _say(prec, "skip" if skip else "", "split env", env, "tiles", ntile, "block", k, "tile", plan.tile)
            assert (k > 0) == (env != "0")
            res = []
            for n in (900, 900, 450):
                plan.run(n)
                res.append([plan.read_all(m)[0] for m in range(2)] + [plan.flags()])
            outs.append(res)
        finally:
            plan.close()
    for other in outs[1:]:
        for ca, cb in zip(outs[0], other):
            for a, b in zip(ca, cb):
                np.testing.assert_array_equal(a, b)


def test_large_connectome_batch_split_matches_oracle_and_edge_lane():
    from oracle.tvb_oracle import OracleSim
    from tvb_hip import _lib as L
    from tvb_hip.connectome import idelays, synthetic
    import bench
    from tvb_hip.plan import CSR, MonitorSpec, Plan
    n, B, n_steps = 400, 48, 1000
    w, Lm = synthetic(n, 0.2, 10.0, 180.0, 400)
    d, H = idelays(Lm, 5.0, 0.05)
    G = np.linspace(0.05, 0.6, B)
    sig = np.full(B, 1e-3)
    outs = {}
    for kern in (L.KERNEL_AUTO, L.KERNEL_EDGE_LANE):
        plan = Plan(model=L.MODEL_MPR, cfun=L.CFUN_LINEAR, scheme=L.HEUN_STO, precision=L.MIXED,
                    noise=L.NOISE_PHILOX, conns=[CSR(w, d)], params=bench.mpr_params(G, sig), dt=0.05,
                    obs_coef=np.eye(2), monitors=[MonitorSpec(L.MON_TAVG, 20, [0, 1], period=1.0),
                                                  bench.make_bold_spec(0.05, 20.0)],
                    max_steps=n_steps, seed=9, kernel=kern)
        try:
            plan.init_uniform(13, [0.0, -2.0], [2.0, 1.5])
            ntile, k = plan.tiles()
            _say("kernel", plan.kernel, "tile", plan.tile, "tiles", ntile, "block", k)
            if kern == L.KERNEL_AUTO:
                assert plan.kernel == L.KERNEL_BATCH_LANE and k > 0
                check = [0, 23, B - 1]
                init = {b: plan.export(b, np.float32) for b in check}
                z = {b: plan.debug_normals(b, 1, n_steps) for b in check}
            plan.run(n_steps)
            outs[kern] = ([plan.read(m) for m in range(2)], plan.flags())
        finally:
            plan.close()
    (ma, fa), (mb, fb) = outs[L.KERNEL_AUTO], outs[L.KERNEL_EDGE_LANE]
    np.testing.assert_array_equal(fa, fb)
    for m in range(2):
        for (ta, xa), (tb, xb) in zip(ma[m], mb[m]):
            np.testing.assert_array_equal(ta, tb)
            np.testing.assert_array_equal(xa, xb)
    for b in check:
        hist, state, cs = init[b]
        o = OracleSim(w, Lm, 5.0, 0.05, "MontbrioPazoRoxin", coupling_params=dict(a=float(G[b])),
                      integrator="HeunStochastic", nsig=np.array([sig[b], sig[b]]),
                      monitors=(("TemporalAverage", 1.0, None), ("Bold", 20.0, None)), mode="faithful")
        o.ring[...] = hist.astype(o.ring.dtype)
        o.current_state = state[..., np.newaxis].copy()
        o.current_step = cs
        (tt, xt), (tb_, xb_) = o.run(n_steps, noise=z[b])
        np.testing.assert_array_equal(ma[0][b][1], xt[..., 0])
        e = np.abs(ma[1][b][1] - xb_[..., 0]).max() / np.abs(xb_).max()
        assert e <= 1e-12, (b, e)
