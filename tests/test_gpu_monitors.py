"""GPU parity of the SubSample, GlobalAverage, SpatialAverage,
AfferentCoupling and AfferentCouplingTemporalAverage monitors and of the
region stimulus, through HipBackend and the C-ABI, against the oracle.

Tolerances (stated per test):
  * mixed device precision vs the oracle's reference-faithful mode (G2D, MPR):
    SubSample, GlobalAverage (numpy's pairwise order), AfferentCoupling and
    its TemporalAverage bit-exact; SpatialAverage max rel <= 1e-14 (the
    reference's numpy.dot runs through BLAS, whose order is not pinned);
  * float64 device mode vs the fp64 oracle: max rel <= 1e-10;
  * a stimulated run is bit-exact (mixed vs faithful) and a run split into
    several launches equals one launch.
"""
import numpy as np
import pytest

from helpers import make_case, rel_err
from tvb_hip import sim as S

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


MASK = np.arange(16) % 4
# at most four monitors per plan (TVBK_MAX_MON)
MONS = (("SubSample", 0.5, None), ("GlobalAverage", 0.5, None),
        ("SpatialAverage", 1.0, None, dict(spatial_mask=MASK)), ("AfferentCoupling", None, None))
MONS_B = (("AfferentCouplingTemporalAverage", 1.0, None), ("SubSample", 0.25, None),
          ("TemporalAverage", 0.5, None), ("AfferentCoupling", None, [0]))


def _check(out, ref, exact, tol=1e-10, savg=(2,)):
    assert len(out) == len(ref)
    for m, ((t, x), (tr, xr)) in enumerate(zip(out, ref)):
        assert x.shape == xr.shape, (m, x.shape, xr.shape)
        assert len(t) > 0
        np.testing.assert_allclose(t, tr, rtol=0, atol=1e-9)
        if exact and m not in savg:
            np.testing.assert_array_equal(x, xr, err_msg="monitor %d" % m)
        else:
            e = rel_err(x, xr)
            _say("   monitor", m, "rel", e)
            assert e <= (1e-14 if exact else tol), (m, e)


@pytest.mark.parametrize("model,scheme,nsig", [("MontbrioPazoRoxin", "HeunStochastic", [0.01, 0.02]),
                                               ("Generic2dOscillator", "HeunDeterministic", None)])
@pytest.mark.parametrize("chunk", [1000, 333])
@pytest.mark.parametrize("mset", ["a", "b"])
def test_new_monitors_mixed_bit_exact(model, scheme, nsig, chunk, mset):
    from tvb_hip.backend import HipBackend
    _say("monitors mixed", model, scheme, chunk, mset)
    mons = MONS if mset == "a" else MONS_B
    orc, sim, _ = make_case(model, n=16, density=0.4, lengths=(5.0, 50.0), speed=3.0, dt=0.05, scheme=scheme,
                            cparams=dict(a=0.3, b=0.01), nsig=nsig, seed=4, monitors=mons)
    o = orc["faithful"]
    z = o.legacy_noise(1000) if nsig else None
    ref = o.run(1000, noise=z)
    out = HipBackend().run_sim(sim, nstep=1000, precision="mixed", noise=z if nsig else "legacy", chunk=chunk)
    _check(out, ref, exact=True, savg=(2,) if mset == "a" else ())


@pytest.mark.parametrize("mset", ["a", "b"])
def test_new_monitors_fp64(mset):
    from tvb_hip.backend import HipBackend
    _say("monitors fp64", mset)
    orc, sim, _ = make_case("MontbrioPazoRoxin", n=16, density=0.4, lengths=(5.0, 50.0), speed=3.0, dt=0.05,
                            scheme="EulerStochastic", cparams=dict(a=0.3), nsig=[0.01, 0.02], seed=6,
                            monitors=MONS if mset == "a" else MONS_B)
    o = orc["fp64"]
    z = o.legacy_noise(800)
    ref = o.run(800, noise=z)
    out = HipBackend().run_sim(sim, nstep=800, precision="fp64", noise=z, chunk=300)
    _check(out, ref, exact=False)


def test_afferent_jansen_rit_sigmoidal():
    """SigmoidalJansenRit: the reference's two coupling rows are equal
    (coupling.py:358-376); both afferent indices read the kernel's one term."""
    from tvb_hip.backend import HipBackend
    _say("afferent JR")
    mons = (("AfferentCoupling", None, None), ("AfferentCouplingTemporalAverage", 2.0, [1]),
            ("GlobalAverage", 1.0, None))
    orc, sim, _ = make_case("JansenRit", n=12, density=0.4, lengths=(5.0, 50.0), speed=3.0, dt=0.1,
                            scheme="EulerStochastic", coupling="SigmoidalJansenRit",
                            nsig=[0.0, 0.0, 0.0, 0.01, 0.01, 0.01], seed=7, monitors=mons)
    o = orc["fp64"]
    z = o.legacy_noise(600)
    ref = o.run(600, noise=z)
    out = HipBackend().run_sim(sim, nstep=600, precision="fp64", noise=z, chunk=250)
    _check(out, ref, exact=False)


def test_afferent_on_batch_lane_kernel():
    """Instances of a batch-lane plan (uniform rows: the sequential loop with
    the afferent pass) against the oracle, two launches."""
    from tvb_hip import _lib as L
    from tvb_hip.backend import HipBackend
    _say("afferent batch")
    cases = [make_case("MontbrioPazoRoxin", n=16, density=0.4, lengths=(5.0, 50.0), speed=3.0, dt=0.05,
                       scheme="HeunStochastic", cparams=dict(a=0.1 + 0.02 * b), nsig=[0.01, 0.02], seed=4,
                       monitors=MONS[1:3] + MONS_B[:1] + MONS[3:]) for b in range(40)]
    sims = [c[1] for c in cases]
    be = HipBackend()
    plan = be.build_batch_plan(sims, precision="mixed", noise="inject", max_steps=500, kernel="batch")
    assert plan.kernel == L.KERNEL_BATCH_LANE and plan.tile > 1
    zs = [c[0]["faithful"].legacy_noise(1000) for c in cases]
    outs = [[] for _ in plan.monitors]
    for k0 in (0, 500):
        z = np.stack([zz[k0:k0 + 500] for zz in zs], axis=-1)        # (n, nsv, N, B)
        plan.set_noise(z)
        plan.run(500)
        for m in range(len(plan.monitors)):
            outs[m].append(plan.read(m))
    for b in (0, 17, 39):
        ref = cases[b][0]["faithful"].run(1000, noise=zs[b])
        got = [(np.concatenate([outs[m][0][b][0], outs[m][1][b][0]]),
                np.concatenate([outs[m][0][b][1], outs[m][1][b][1]])[..., np.newaxis])
               for m in range(len(plan.monitors))]
        _check(got, ref, exact=True, savg=(1,))
    plan.close()


@pytest.mark.parametrize("model,scheme,nsig", [("Generic2dOscillator", "HeunDeterministic", None),
                                               ("MontbrioPazoRoxin", "HeunStochastic", [0.01, 0.02]),
                                               ("MontbrioPazoRoxin", "EulerDeterministic", None)])
def test_region_stimulus_bit_exact(model, scheme, nsig):
    """StimuliRegion (PulseTrain x per-region weights) on the model's stvar;
    three launches (the temporal index continues across them)."""
    from tvb_hip.backend import HipBackend
    _say("stimulus", model, scheme)
    n = 16
    weights = np.where(np.arange(n) % 3 == 0, 1.5, 0.0)
    teq = S.PulseTrain(parameters={"T": 12.0, "tau": 4.0, "amp": 0.8, "onset": 5.0})
    mons = (("Raw", None, None), ("TemporalAverage", 1.0, None))
    orc, sim, _ = make_case(model, n=n, density=0.4, lengths=(5.0, 50.0), speed=3.0, dt=0.05, scheme=scheme,
                            cparams=dict(a=0.2), nsig=nsig, seed=9, monitors=mons,
                            stimulus=(weights, teq, 60.0))
    o = orc["faithful"]
    nstep = 1200
    z = o.legacy_noise(nstep) if nsig else None
    ref = o.run(nstep, noise=z)
    out = HipBackend().run_sim(sim, nstep=nstep, precision="mixed", noise=z if nsig else "legacy", chunk=500)
    _check(out, ref, exact=True, savg=())
    # the same run without the stimulus differs
    orc2, sim2, _ = make_case(model, n=n, density=0.4, lengths=(5.0, 50.0), speed=3.0, dt=0.05, scheme=scheme,
                              cparams=dict(a=0.2), nsig=nsig, seed=9, monitors=mons)
    plain = HipBackend().run_sim(sim2, nstep=nstep, precision="mixed", noise=z if nsig else "legacy")
    assert not np.array_equal(plain[0][1], out[0][1])


def test_stimulus_time_grid_too_short_raises():
    from tvb_hip.backend import HipBackend
    orc, sim, _ = make_case("Generic2dOscillator", n=8, dt=0.1, monitors=(("Raw", None, None),),
                            stimulus=(np.ones(8), S.PulseTrain(), 10.0))
    with pytest.raises(ValueError):
        HipBackend().run_sim(sim, nstep=500, precision="mixed")
