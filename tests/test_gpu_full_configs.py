"""GPU parity at the benchmark configurations' real shapes.

* C2: the full 68-node x 8192-instance sweep (128 G x 64 sigma grid, tile of
  32 instances, pipelined batch-lane loop) in the reference's precision mix
  with device Philox noise.  Sampled instances, the high-coupling corner
  among them, are rerun on the reference-faithful oracle fed the device's
  own normals (tvbk_debug_normals) from the device's initial history
  (tvbk_export_state): TemporalAverage bit-exact, Bold <= 1e-12, and the
  device's non-finite flag equal to the oracle's non-finiteness.
* Long CSR rows: sweeps whose rows are longer than the register paths of
  the batch-lane loops (24 and 32 edges) -- the 192-node C1 connectome
  (rows up to 55) through the pipelined loop (tile 16) and a dense 48-node
  connectome (rows up to 46) through the time-blocked loop (tile 64) --
  bit-exact against the faithful oracle.
* C3: all 1024 uneven JansenRit instances (node counts 68 .. 998, own
  connectomes) at a short duration in FP64: the split plan bit-identical to
  the unsplit one, and a 998-node instance within 1e-10 of the fp64 oracle.
* C4: requests drawn from the serving distribution (1-512 instances,
  68-998 nodes) through the request loop; one instance of a served request
  bit-exact against the faithful oracle.

Mirrors the pattern of the reference's backend tests
(tests/library/simulator/backend/backendtestbase.py:125-134,
nbbackend_mpr_test.py:43-58): the same simulation run twice, compared per
monitor output.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _oracle_from_device(o, hist, state, current_step):
    """An oracle simulator continued from a device instance's exported
    history ring (H, nch, N), state (nsv, N) and step.  Channels a plan does
    not keep (zero coupling coefficient) are zero in the oracle's ring."""
    o.ring[...] = 0
    o.ring[:, :hist.shape[1]] = hist.astype(o.ring.dtype)
    o.current_state = np.asarray(state, np.float64)[..., np.newaxis].copy()
    o.current_step = int(current_step)
    return o


@pytest.mark.parametrize("skip", [False, True], ids=["all-channels", "skip-zero-coupling"])
def test_c2_full_shape_sampled_instances_bit_exact(skip):
    """skip: the V channel (cv = 0 everywhere) is not kept; the oracle's V
    history is then zero, which its zero coefficient makes irrelevant."""
    import bench
    from oracle.tvb_oracle import OracleSim
    from tvb_hip import _lib as L
    n_steps, B = 4000, 8192
    w, Lm, d, H, G, sig = bench.c2_setup(B)
    plan = bench.build_c2_plan(w, d, G, sig, 0.05, n_steps, L.MIXED, L.NOISE_PHILOX, 1234, bold=100.0, skip=skip)
    try:
        assert (plan.kernel, plan.tile) == (L.KERNEL_BATCH_LANE, 32)
        assert plan.nch == (1 if skip else 2)
        plan.init_uniform(99, [0.0, -2.0], [2.0, 1.5])
        # candidates: spread over the grid plus the high-G corner (G >= 1.8)
        cand = sorted(set([0, 63, 1000, 2600, 4100, 4800, 6000, 7000] +
                          [gi * 64 + si for gi in (116, 121, 125, 127) for si in (0, 30, 63)]))
        init = {b: plan.export(b, np.float32) for b in cand}
        _say("c2 full shape: run", n_steps, "steps, tile", plan.tile, "waves", plan.waves)
        plan.run(n_steps)
        assert plan.loop() == ("pipelined-2" if skip else "pipelined")
        (tav, bol) = (plan.read(0, cand), plan.read(1, cand))
        flags = plan.flags()
        _say("  non-finite instances:", int((flags != 0).sum()))
        bad = [b for b in cand if flags[b]][:3]
        good = [b for b in cand if not flags[b]]
        picked = good[:6] + [b for b in good if G[b] >= 1.8][:3] + bad
        assert any(G[b] >= 1.8 for b in picked)
        for b in dict.fromkeys(picked):
            k = cand.index(b)
            hist, state, cs = init[b]
            assert cs == 0
            z = plan.debug_normals(b, cs + 1, n_steps)
            o = OracleSim(w, Lm, 5.0, 0.05, "MontbrioPazoRoxin", coupling_params=dict(a=float(G[b])),
                          integrator="HeunStochastic", nsig=np.array([sig[b], sig[b]]),
                          monitors=(("TemporalAverage", 1.0, None), ("Bold", 100.0, None)), mode="faithful")
            _oracle_from_device(o, hist, state, cs)
            with np.errstate(all="ignore"):
                ((tt, xt), (tb, xb)), states = o.run(n_steps, noise=z, record_state=True)
            finite = bool(np.isfinite(states).all())
            _say("  instance", b, "G=%.3f sigma=%.2e" % (G[b], sig[b]), "device flag", int(flags[b]),
                 "oracle finite", finite)
            assert bool(flags[b]) == (not finite), b
            np.testing.assert_array_equal(tav[k][0], tt)
            np.testing.assert_array_equal(bol[k][0], tb)
            if finite or not skip:
                np.testing.assert_array_equal(tav[k][1], xt[..., 0])
            else:
                # past the divergence, 0 * c_V (NaN once V is) and the skip's
                # 0 * b differ: equal up to the first non-finite sample
                bad_d = ~np.isfinite(tav[k][1]).reshape(len(tt), -1).all(axis=1)
                bad_o = ~np.isfinite(xt).reshape(len(tt), -1).all(axis=1)
                assert bad_d.any() and bad_o.any()
                j = min(np.argmax(bad_d), np.argmax(bad_o))
                np.testing.assert_array_equal(tav[k][1][:j], xt[:j, ..., 0])
            if finite:
                e = np.abs(bol[k][1] - xb[..., 0]).max() / np.abs(xb).max()
                assert e <= 1e-12, (b, e)
            elif not skip:
                np.testing.assert_array_equal(np.isfinite(bol[k][1]), np.isfinite(xb[..., 0]))
    finally:
        plan.close()


def test_skip_zero_coupling_equals_all_channels_at_c2_shape():
    """The V-channel skip changes nothing the reference computes: the full
    8192-instance sweep with and without it, same initial draw and device
    noise, gives bit-identical TemporalAverage and Bold on every instance
    that stays finite, and the same non-finite flags."""
    import bench
    from tvb_hip import _lib as L
    n_steps, B = 2000, 8192
    w, Lm, d, H, G, sig = bench.c2_setup(B)
    out = {}
    for skip in (False, True):
        plan = bench.build_c2_plan(w, d, G, sig, 0.05, n_steps, L.MIXED, L.NOISE_PHILOX, 1234, bold=50.0, skip=skip)
        try:
            plan.init_uniform(99, [0.0, -2.0], [2.0, 1.5])
            plan.run(n_steps)
            out[skip] = (plan.read_all(0)[0], plan.read_all(1)[0], plan.flags(), plan.nch)
        finally:
            plan.close()
    (ta, ba, fa, na), (tb, bb, fb, nb) = out[False], out[True]
    assert (na, nb) == (2, 1)
    np.testing.assert_array_equal(fa, fb)
    ok = fa == 0
    _say("skip vs all channels: finite instances", int(ok.sum()), "of", B)
    np.testing.assert_array_equal(ta.reshape(-1, B)[:, ok], tb.reshape(-1, B)[:, ok])
    np.testing.assert_array_equal(ba.reshape(-1, B)[:, ok], bb.reshape(-1, B)[:, ok])


def _dense48():
    """48 nodes at 90 % density: rows of about 42 edges, and a tile of 64
    instances still fits LDS (TemporalAverage only)."""
    from tvb_hip.connectome import idelays, synthetic
    w, Lm = synthetic(48, 0.9, 5.0, 60.0, 48)
    d, H = idelays(Lm, 3.0, 0.05)
    rng = np.random.default_rng(480)
    ic = np.empty((H, 2, 48))
    ic[:, 0] = rng.uniform(0.0, 2.0, size=(H, 48))
    ic[:, 1] = rng.uniform(-2.0, 1.5, size=(H, 48))
    return w, Lm, d, H, ic, rng, 3.0, False


def _c1_conn():
    """The 192-node C1 connectome (rows up to 55 edges); a tile of 16
    instances fits LDS in the reference's mix with TemporalAverage."""
    import bench
    return bench.c1_setup() + (4.0, False)


@pytest.mark.parametrize("conn,tile,loop", [("c1_192", 16, "pipelined"), ("dense48", 64, "time-blocked")])
def test_long_rows_shared_connectome(conn, tile, loop):
    import bench
    from oracle.tvb_oracle import OracleSim
    from tvb_hip import _lib as L
    from tvb_hip.plan import CSR, MonitorSpec, Plan
    w, Lm, d, H, ic, rng, speed, with_bold = _c1_conn() if conn == "c1_192" else _dense48()
    n = w.shape[0]
    rows = (w != 0).sum(axis=1)
    assert rows.max() > 32                     # longer than the register paths (24 / 32 edges)
    B, n_steps = 64, 1000
    G = np.linspace(0.1, 0.8, B)
    p = np.empty((B, 12))
    p[:, :8] = [1.0, 0.0, 1.0, 15.0, -5.0, 0.0, 1.0, 0.0]
    p[:, 8], p[:, 9], p[:, 10], p[:, 11] = G, 0.0, 0.01, 0.02
    mons = [MonitorSpec(L.MON_TAVG, 20, [0, 1], period=1.0)]
    if with_bold:
        mons.append(bench.make_bold_spec(0.05, 20.0))
    plan = Plan(model=L.MODEL_MPR, cfun=L.CFUN_LINEAR, scheme=L.HEUN_STO, precision=L.MIXED,
                noise=L.NOISE_INJECTED, conns=[CSR(w, d)], params=p, dt=0.05, obs_coef=np.eye(2),
                monitors=mons, max_steps=n_steps, seed=0, kernel=L.KERNEL_BATCH_LANE, tile=tile)
    try:
        assert plan.tile == tile
        assert (plan.time_block() > 1) == (loop == "time-blocked")
        plan.set_history(-1, ic, ic[-1], H - 1)
        z = rng.standard_normal((n_steps, 2, n, B))
        plan.set_noise(z)
        _say("long rows:", loop, "tile", tile, "max row", int(rows.max()))
        plan.run(n_steps)
        assert plan.loop() == loop
        check = [0, 17, 40, B - 1]
        tav = plan.read(0, check)
        bol = plan.read(1, check) if with_bold else None
        omons = (("TemporalAverage", 1.0, None),) + ((("Bold", 20.0, None),) if with_bold else ())
        for k, b in enumerate(check):
            o = OracleSim(w, Lm, speed, 0.05, "MontbrioPazoRoxin", coupling_params=dict(a=float(G[b])),
                          integrator="HeunStochastic", nsig=np.array([0.01, 0.02]), initial_conditions=ic,
                          monitors=omons, mode="faithful")
            res = o.run(n_steps, noise=z[..., b])
            np.testing.assert_array_equal(tav[k][1], res[0][1][..., 0])
            if with_bold:
                xb = res[1][1]
                e = np.abs(bol[k][1] - xb[..., 0]).max() / np.abs(xb).max()
                assert e <= 1e-12, (b, e)
    finally:
        plan.close()


def test_c3_full_batch_split_unsplit_and_oracle():
    import bench
    from oracle.tvb_oracle import OracleSim
    from tvb_hip import _lib as L
    max_steps = 300
    old = os.environ.get("TVBK_SPLIT")
    plans = {}
    try:
        for label, env in (("split", None), ("unsplit", "0")):
            if env is None:
                os.environ.pop("TVBK_SPLIT", None)
            else:
                os.environ["TVBK_SPLIT"] = env
            plans[label] = bench.c3_plan(1024, duration_scale=0.001, max_steps=max_steps, precision=L.F64)
    finally:
        if old is None:
            os.environ.pop("TVBK_SPLIT", None)
        else:
            os.environ["TVBK_SPLIT"] = old
    try:
        (ps, steps, _, info), (pu, _, _, _) = plans["split"], plans["unsplit"]
        assert info["nodes_histogram"]["998"] > 0
        assert ps.tiles()[1] > 0 and pu.tiles()[1] == 0         # split blocks / none
        # one 998-node instance from a host-given history (H, 6, N): the
        # device ring holds the coupling's pre-synaptic sigmoid of y1 - y2
        big = next(b for b in range(1024) if ps.conns[b].n_node == 998)
        c = ps.conns[big]
        rng = np.random.default_rng(31)
        lo = np.array([-1.0, -500.0, -50.0, -6.0, -20.0, -500.0]) * 0.01
        ic = rng.uniform(lo, -lo, size=(c.horizon, c.n_node, 6)).transpose(0, 2, 1).copy()
        cmin, cmax, mid, r = 0.0, 0.005, 6.0, 0.56
        pre = (cmin + (cmax - cmin) / (1.0 + np.exp(r * (mid - (ic[:, 1] - ic[:, 2])))))[:, np.newaxis]
        for p in (ps, pu):
            p.set_history(big, pre, ic[-1], c.horizon - 1)
        _say("c3: 1024 instances, tiles split/unsplit", ps.tiles(), pu.tiles(), "998-node instance", big)
        for p in (ps, pu):
            p.run(max_steps)
        assert ps.loop() == pu.loop() == "time-blocked"
        for m in range(2):
            for (t1, x1), (t2, x2) in zip(ps.read(m), pu.read(m)):
                np.testing.assert_array_equal(t1, t2)
                np.testing.assert_array_equal(x1, x2)
        np.testing.assert_array_equal(ps.flags(), pu.flags())
        assert not ps.flags().any()
        # the 998-node instance against the fp64 oracle
        n = steps[big]
        z = ps.debug_normals(big, c.horizon, n)
        prm = ps.params[big]
        w = np.zeros((c.n_node, c.n_node))
        L_ = np.zeros_like(w)
        rows = np.repeat(np.arange(c.n_node), np.diff(c.row_ptr))
        w[rows, c.col] = c.weight
        # tract lengths giving the same integer delays at speed 3 mm/ms, dt 0.1
        L_[rows, c.col] = c.idelay * 0.3
        L_[0, 0] = (c.horizon - 1) * 0.3        # the horizon spans the full length matrix
        o = OracleSim(w, L_, 3.0, 0.1, "JansenRit", model_params=dict(a=prm[2], b=prm[3]),
                      coupling="SigmoidalJansenRit", integrator="EulerStochastic",
                      nsig=np.array([0.0, 0.0, 0.0, 1e-3, 1e-3, 1e-3]), initial_conditions=ic,
                      monitors=(("TemporalAverage", 2.0, None),), mode="fp64")
        assert o.H == c.horizon
        (tt, xt), = o.run(n, noise=z)
        (td, xd), = ps.read(0, [big])
        np.testing.assert_allclose(td, tt, atol=1e-9)
        e = np.abs(xd - xt[..., 0]).max() / np.abs(xt).max()
        _say("  998-node instance,", n, "steps: TemporalAverage max rel", e)
        assert e <= 1e-10
    finally:
        for p, *_ in plans.values():
            p.close()


def test_c4_served_request_matches_oracle():
    from oracle.tvb_oracle import OracleSim
    from tvb_hip import _lib as L
    from tvb_hip.connectome import idelays, synthetic
    from tvb_hip.serve import Server, synthetic_requests
    reqs = synthetic_requests(6, 50.0, seed=414, duration_scale=0.002)
    assert max(r.meta["B"] for r in reqs) > 64 and max(r.meta["n"] for r in reqs) > 300
    streams = [L.stream_create() for _ in range(4)]
    try:
        srv = Server(hbm_budget=64 << 30, streams=streams, chunk_steps=1000, keep_outputs=True)
        stats = srv.run(reqs)
        _say("c4:", stats["n_requests"], "requests, p50 %.3f s p99 %.3f s" % (stats["p50_latency_s"],
                                                                           stats["p99_latency_s"]))
        r = min(reqs, key=lambda q: q.meta["n"] * q.n_steps)
        b = r.meta["B"] // 2
        # the same request's plan, rebuilt: its initial history and normals
        plan = r.build(None, r.n_steps)
        try:
            hist, state, cs = plan.export(b, np.float32)
            z = plan.debug_normals(b, cs + 1, r.n_steps)
            G = float(plan.params[b, 8 if r.meta["model"] == "MPR" else 12])
            sig = float(plan.params[b, -1])
        finally:
            plan.close()
        m = r.meta
        w, Lm = synthetic(m["n"], 0.2, 10.0, 180.0, m["cseed"])
        model = "MontbrioPazoRoxin" if m["model"] == "MPR" else "Generic2dOscillator"
        o = OracleSim(w, Lm, m["speed"], m["dt"], model, coupling_params=dict(a=G), integrator="HeunStochastic",
                      nsig=np.array([sig, sig]), monitors=(("TemporalAverage", 1.0, [0]),), mode="faithful")
        _oracle_from_device(o, hist, state, cs)
        (tt, xt), = o.run(r.n_steps, noise=z)
        (td, xd), = r.outputs[b]
        _say("  request", r.rid, m, "instance", b)
        np.testing.assert_array_equal(td, tt)
        np.testing.assert_array_equal(xd, xt[..., 0])
    finally:
        for st in streams:
            L.stream_destroy(st)
