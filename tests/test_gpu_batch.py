"""GPU tests of batched plans: shared-connectome sweeps (batch-lane kernel),
per-instance connectomes with uneven durations (edge-lane kernel), the
Philox noise source, mix invariance, resume, zero delays, spatial
parameters and non-finite flags.  Tolerances as in test_gpu_parity.py:
FP64 device vs pure-fp64 oracle max rel <= 1e-10; FP32 device vs fp64
oracle TemporalAverage max rel <= 1e-4; kernel-family and mix invariance
bit-exact."""
import numpy as np
import pytest

from helpers import make_case, rel_err
from oracle import philox
from tvb_hip.connectome import synthetic

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _backend():
    from tvb_hip.backend import HipBackend
    return HipBackend()


MONS = (("Raw", None, None), ("TemporalAverage", 1.0, None), ("Bold", 10.0, None))


def _sweep_cases(B, n=20, seed=3, model="MontbrioPazoRoxin", scheme="HeunStochastic", monitors=MONS):
    """B cases on one shared connectome with distinct coupling, noise and ICs."""
    w, L = synthetic(n, 0.3, 5.0, 40.0, seed)
    rng = np.random.default_rng(seed)
    cases = []
    for b in range(B):
        a = float(rng.uniform(0.1, 0.8))
        sg = float(rng.uniform(1e-3, 2e-2))
        orc, sim, _ = make_case(model, weights=w, tract_lengths=L, speed=3.0, dt=0.05, scheme=scheme,
                                cparams=dict(a=a), nsig=[sg, sg] if scheme.endswith("Stochastic") else None,
                                seed=seed * 100 + b, monitors=monitors)
        cases.append((orc, sim))
    return cases


@pytest.mark.parametrize("tile", [16, 32, 64])
def test_batch_lane_fp64_matches_oracle(tile):
    """B = 40 instances (last tile partly padded) of MPR HeunStochastic on a
    shared connectome, FP64 batch-lane kernel vs the fp64 oracle per instance."""
    _say("batch fp64 tile", tile)
    B, nstep = 40, 800
    cases = _sweep_cases(B)
    N = cases[0][1].connectivity.number_of_regions
    z = np.random.default_rng(11).standard_normal((B, nstep, 2, N))
    out = _backend().run_batch([s for _, s in cases], nstep, precision="fp64", noise=list(z),
                               kernel="batch", tile=tile)
    for b in (0, 17, 39):
        ref = cases[b][0]["fp64"].run(nstep, noise=z[b])
        for (tr, xr), (td, xd) in zip(ref, out[b]):
            assert xd.shape == xr.shape
            assert rel_err(xd, xr) <= 1e-10, (b, rel_err(xd, xr))


def test_batch_and_edge_kernels_bit_identical():
    """The same instances through the batch-lane kernel (tile 32) and the
    edge-lane kernel (one instance per workgroup) give bit-identical FP32
    results: both sum rows in the same order per instance."""
    _say("batch vs edge bit identity")
    B, nstep = 33, 600
    cases = _sweep_cases(B, seed=5)
    N = cases[0][1].connectivity.number_of_regions
    z = list(np.random.default_rng(2).standard_normal((B, nstep, 2, N)))
    be = _backend()
    o1 = be.run_batch([s for _, s in cases], nstep, precision="fp32", noise=z, kernel="batch", tile=32)
    o2 = be.run_batch([s for _, s in cases], nstep, precision="fp32", noise=z, kernel="edge")
    for b in range(B):
        for (t1, x1), (t2, x2) in zip(o1[b], o2[b]):
            np.testing.assert_array_equal(x1, x2)
            np.testing.assert_array_equal(t1, t2)


def test_fp32_batch_vs_fp64_oracle_tavg():
    """FP32 batch-lane sweep vs the fp64 oracle: TemporalAverage max rel <= 1e-4."""
    _say("fp32 batch tavg")
    B, nstep = 32, 3000
    cases = _sweep_cases(B, seed=9, monitors=(("TemporalAverage", 1.0, None), ("Bold", 20.0, None)))
    N = cases[0][1].connectivity.number_of_regions
    z = np.random.default_rng(4).standard_normal((B, nstep, 2, N))
    out = _backend().run_batch([s for _, s in cases], nstep, precision="fp32", noise=list(z), kernel="batch")
    for b in (0, 31):
        (_, ta), (_, bo) = cases[b][0]["fp64"].run(nstep, noise=z[b])
        assert rel_err(out[b][0][1], ta) <= 1e-4
        assert rel_err(out[b][1][1], bo) <= 1e-6


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("fp32", 1e-4)])
def test_philox_noise_matches_oracle(precision, tol):
    """Device Philox4x32-10 normals (key = seed, counter = node|call, stream,
    step) equal the oracle's numpy restatement: the oracle integrates with
    those normals injected."""
    _say("philox", precision)
    B, nstep, seed = 32, 600, 77
    cases = _sweep_cases(B, seed=12, monitors=(("Raw", None, None),))
    sims = [s for _, s in cases]
    streams = np.arange(1000, 1000 + B, dtype=np.uint64)
    out = _backend().run_batch(sims, nstep, precision=precision, noise="philox", kernel="batch", seed=seed,
                               streams=streams)
    N = sims[0].connectivity.number_of_regions
    for b in (0, 21):
        o = cases[b][0]["fp64"]
        z = philox.normals(seed, int(streams[b]), nstep, 2, N, step0=o.current_step + 1,
                           precision="f64" if precision == "fp64" else "f32")
        (_, raw), = o.run(nstep, noise=z)
        assert rel_err(out[b][0][1], raw) <= tol, rel_err(out[b][0][1], raw)


def test_uneven_instances_own_connectomes():
    """Instances with their own connectome, node count, horizon and duration
    (edge-lane kernel, longest first) each match their own oracle run."""
    _say("uneven")
    specs = [(12, 300, 1), (23, 500, 2), (9, 420, 3), (17, 180, 4)]
    cases, sims, zs = [], [], []
    for n, steps, sd in specs:
        orc, sim, _ = make_case("JansenRit", n=n, density=0.25, lengths=(5.0, 45.0), speed=2.0, dt=0.1,
                                scheme="EulerStochastic", coupling="SigmoidalJansenRit",
                                nsig=[0, 0, 0, 1e-3, 1e-3, 1e-3], seed=sd, spread=0.02,
                                monitors=(("TemporalAverage", 2.0, None), ("Bold", 20.0, None)))
        cases.append(orc)
        sims.append(sim)
        zs.append(np.random.default_rng(sd).standard_normal((max(s for _, s, _ in specs), 6, n)))
    out = _backend().run_batch(sims, max(s for _, s, _ in specs), precision="fp64", noise=zs,
                               steps=[s for _, s, _ in specs])
    for (n, steps, sd), orc, o_dev, z in zip(specs, cases, out, zs):
        ref = orc["fp64"].run(steps, noise=z[:steps])
        for (tr, xr), (td, xd) in zip(ref, o_dev):
            assert xd.shape == xr.shape, (n, xd.shape, xr.shape)
            np.testing.assert_allclose(td, tr, atol=1e-9)
            assert rel_err(xd, xr) <= 1e-10


def test_mix_invariance_bit_exact():
    """An instance run alone equals the same instance inside a batch of
    others, bit for bit (FP32, Philox noise keyed by its stream id)."""
    _say("mix invariance")
    B, nstep = 24, 500
    cases = _sweep_cases(B, seed=21, monitors=(("TemporalAverage", 1.0, None),))
    sims = [s for _, s in cases]
    streams = np.arange(B, dtype=np.uint64) * 7 + 3
    be = _backend()
    mix = be.run_batch(sims, nstep, precision="fp32", noise="philox", seed=5, streams=streams, kernel="batch",
                       tile=16)
    k = 13
    cases2 = _sweep_cases(B, seed=21, monitors=(("TemporalAverage", 1.0, None),))
    alone = be.run_batch([cases2[k][1]], nstep, precision="fp32", noise="philox", seed=5,
                         streams=streams[k:k + 1])
    np.testing.assert_array_equal(alone[0][0][1], mix[k][0][1])


def test_resume_and_checkpoint_round_trip():
    """Two calls of 300 steps equal one of 600 (state, history and monitor
    accumulators persist on the device); exporting the history and state
    into a fresh plan continues the Raw trajectory bit-exactly."""
    _say("resume")
    orc, sim, _ = make_case("MontbrioPazoRoxin", n=18, dt=0.05, scheme="HeunDeterministic",
                            cparams=dict(a=0.4), monitors=(("Raw", None, None), ("TemporalAverage", 1.0, None)))
    be = _backend()
    p1 = be.build_plan(sim, precision="fp32", max_steps=600)
    p1.run(600)
    raw_a = p1.read(0)[0][1].copy()
    tav_a = p1.read(1)[0][1].copy()
    p2 = be.build_plan(sim, precision="fp32", max_steps=600)
    p2.run(300)
    r1, t1 = p2.read(0)[0][1].copy(), p2.read(1)[0][1].copy()
    hist, state, step = p2.export(0, np.float32)
    p2.run(300)
    r2, t2 = p2.read(0)[0][1], p2.read(1)[0][1]
    np.testing.assert_array_equal(np.concatenate([r1, r2]), raw_a)
    np.testing.assert_array_equal(np.concatenate([t1, t2]), tav_a)
    p3 = be.build_plan(sim, precision="fp32", max_steps=600)
    p3.set_history(0, hist, state, step)
    p3.run(300)
    np.testing.assert_array_equal(p3.read(0)[0][1], r2)
    for p in (p1, p2, p3):
        p.close()


def test_zero_delay_connectome():
    """All delays zero (speed -> infinity): coupling reads the state written
    in the same step (two barriers per step); FP64 vs oracle."""
    _say("zero delay")
    orc, sim, _ = make_case("MontbrioPazoRoxin", n=16, lengths=(5.0, 40.0), speed=1e9, dt=0.05,
                            scheme="HeunDeterministic", cparams=dict(a=0.3), monitors=(("Raw", None, None),))
    assert sim.connectivity.horizon == 1
    ref = orc["fp64"].run(400)
    out = _backend().run_sim(sim, nstep=400, precision="fp64")
    assert rel_err(out[0][1], ref[0][1]) <= 1e-10


def test_spatial_parameters():
    """Per-node eta (a spatialised model parameter, models/base.py:217-242)."""
    _say("spatial")
    eta = np.linspace(-5.5, -4.0, 14)
    orc, sim, _ = make_case("MontbrioPazoRoxin", n=14, dt=0.05, scheme="HeunStochastic",
                            cparams=dict(a=0.3), mparams=dict(eta=eta), nsig=[0.01, 0.01],
                            monitors=(("Raw", None, None),))
    z = np.random.default_rng(0).standard_normal((500, 2, 14))
    ref = orc["fp64"].run(500, noise=z)
    out = _backend().run_sim(sim, nstep=500, precision="fp64", noise=z)
    assert rel_err(out[0][1], ref[0][1]) <= 1e-10


def test_nonfinite_flag_marks_only_the_bad_instance():
    _say("nonfinite flag")
    import tvb_hip._lib as L
    cases = _sweep_cases(20, seed=31, scheme="HeunDeterministic", monitors=(("TemporalAverage", 1.0, None),))
    sims = [s for _, s in cases]
    sims[6].model.tau = np.array([1e-9])          # explodes within a few steps
    be = _backend()
    be.run_batch(sims, 200, precision="fp32")
    flags = be.last_flags
    assert flags[6] != 0
    assert np.all(np.delete(flags, 6) == 0)


def test_legacy_noise_drop_in():
    """run_sim(noise="legacy") draws the simulator's own RandomState stream
    after its random initial history, as the reference does; FP64 device vs
    the fp64 oracle built with the same seed."""
    _say("legacy noise")
    from oracle.tvb_oracle import OracleSim
    from tvb_hip import sim as S
    w, L = synthetic(10, 0.4, 5.0, 30.0, 1)
    s = S.Simulator(connectivity=S.Connectivity(weights=w, tract_lengths=L, speed=np.array([3.0])),
                    model=S.MontbrioPazoRoxin(), coupling=S.Linear(a=0.2),
                    integrator=S.HeunStochastic(dt=0.05, noise=S.Additive(nsig=np.array([0.01]), noise_seed=5)),
                    monitors=[S.TemporalAverage(period=1.0)]).configure()
    o = OracleSim(w, L, 3.0, 0.05, "MontbrioPazoRoxin", coupling_params=dict(a=0.2),
                  integrator="HeunStochastic", nsig=0.01, noise_seed=5, mode="fp64",
                  monitors=(("TemporalAverage", 1.0, None),))
    (t, x), = o.run(400, noise=o.legacy_noise(400))
    (td, xd), = _backend().run_sim(s, nstep=400, precision="fp64", noise="legacy")
    np.testing.assert_allclose(td, t)
    assert rel_err(xd, x) <= 1e-10
    assert s.current_step == 400


def test_dense_mfma_zero_delay_sweep():
    """All delays zero: the sweep runs on the matrix-core kernel (coupling as
    W X per channel).  FP32 vs the fp64 oracle: TemporalAverage max rel
    <= 1e-4; vs the sparse batch-lane kernel on the same instances (only the
    summation order differs): max rel <= 1e-5."""
    _say("dense mfma")
    import tvb_hip._lib as L
    B, nstep = 40, 2000
    w, Lm = synthetic(20, 0.3, 5.0, 40.0, 3)
    rng = np.random.default_rng(8)
    cases = []
    for b in range(B):
        orc, sim, _ = make_case("MontbrioPazoRoxin", weights=w, tract_lengths=Lm, speed=1e9, dt=0.05,
                                scheme="HeunStochastic", cparams=dict(a=float(rng.uniform(0.1, 0.6))),
                                nsig=[0.01, 0.01], seed=500 + b,
                                monitors=(("TemporalAverage", 1.0, None), ("Bold", 10.0, None)))
        cases.append((orc, sim))
    sims = [s for _, s in cases]
    z = list(np.random.default_rng(6).standard_normal((B, nstep, 2, 20)))
    be = _backend()
    dense = be.run_batch(sims, nstep, precision="fp32", noise=z)
    assert be.last_plan_shape[0] == L.KERNEL_DENSE_MFMA
    sparse = be.run_batch(sims, nstep, precision="fp32", noise=z, kernel="batch")
    for b in (0, 33, 39):
        (_, ta), (_, bo) = cases[b][0]["fp64"].run(nstep, noise=z[b])
        assert rel_err(dense[b][0][1], ta) <= 1e-4
        assert rel_err(dense[b][1][1], bo) <= 1e-6
        assert rel_err(dense[b][0][1], sparse[b][0][1]) <= 1e-5
