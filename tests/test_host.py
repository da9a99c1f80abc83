"""Host-side tests (no GPU): the C-ABI library loads and exports every
declared symbol, fails loudly without a device, and the host logic that
derives a device plan from a configured simulator agrees with the oracle's
independent restatement."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from helpers import make_case
from oracle.tvb_oracle import OracleSim, set_idelays

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    src = open(os.path.join(ROOT, "include", "tvbk.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(tvbk_[a-z_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    from tvb_hip import _lib
    lib = _lib.load()
    declared = _declared_symbols()
    assert len(declared) >= 15
    for name in declared:
        assert hasattr(lib, name), name
    assert set(declared) == set(_lib.EXPORTED)
    assert lib.tvbk_abi_version() == _lib.ABI_VERSION


def test_model_info_is_host_only():
    from tvb_hip import _lib as L
    assert L.model_info(L.MODEL_MPR, L.CFUN_LINEAR) == (8, 2, 2, 2)
    assert L.model_info(L.MODEL_JR, L.CFUN_SIGMOID_JR) == (15, 5, 6, 1)
    assert L.model_info(L.MODEL_G2D, L.CFUN_LINEAR) == (12, 2, 2, 1)
    with pytest.raises(L.TvbkError):
        L.model_info(L.MODEL_G2D, L.CFUN_SIGMOID_JR)     # pre needs two cvars


def test_plan_fails_loudly_without_device():
    """No GPU in this container: plan creation returns an error, never a
    CPU fallback."""
    from tvb_hip import _lib as L
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    orc, sim, _ = make_case("MontbrioPazoRoxin", n=8)
    from tvb_hip.backend import HipBackend
    with pytest.raises(L.TvbkError):
        HipBackend().build_plan(sim)


def test_probe_gather_rejects_bad_arguments():
    """Argument checks of the calibration entry point run before any device
    call."""
    from tvb_hip import _lib as L
    for piece, write, nbytes in ((300, False, 1 << 30), (256, False, 1 << 20), (8192, True, 1 << 30)):
        with pytest.raises(L.TvbkError):
            L.probe_gather(piece, write, nbytes)


def test_missing_library_raises(monkeypatch, tmp_path):
    from tvb_hip import _lib
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "nope.so"))
    with pytest.raises(_lib.LibraryMissing):
        _lib.load()


def test_csr_order_and_horizon():
    """CSR rows follow SparseHistory's row-major nnz order (history.py:221-227);
    the horizon spans the full delay matrix (connectivity.py:323)."""
    from tvb_hip.plan import CSR
    w = np.array([[0, 2.0, 0, 1.0], [0, 0, 0, 0], [3.0, 0, 0, 4.0], [0, 5.0, 0, 0]])
    d = np.array([[0, 1, 9, 2], [1, 0, 3, 4], [5, 6, 0, 7], [8, 3, 2, 0]], np.int32)
    c = CSR(w, d)
    np.testing.assert_array_equal(c.row_ptr, [0, 2, 2, 4, 5])
    np.testing.assert_array_equal(c.col, [1, 3, 0, 3, 1])
    np.testing.assert_array_equal(c.idelay, [1, 2, 5, 7, 3])
    np.testing.assert_array_equal(c.weight, [2, 1, 3, 4, 5])
    assert c.horizon == 10                     # zero-weight pair (0, 2) has delay 9


@pytest.mark.parametrize("with_ic", [True, False])
def test_mirror_configure_matches_oracle(with_ic):
    """tvb_hip.sim.Simulator.configure and the oracle restate the reference's
    configure independently: delays, horizon, current step, state and ring agree."""
    from tvb_hip import sim as S
    from tvb_hip.connectome import synthetic
    w, L = synthetic(12, 0.3, 5.0, 50.0, seed=3)
    _, _, H = set_idelays(L, 2.5, 0.1)
    ic = np.random.default_rng(0).uniform(-1, 1, (H + 5, 2, 12)) if with_ic else None
    s = S.Simulator(connectivity=S.Connectivity(weights=w, tract_lengths=L, speed=np.array([2.5])),
                    model=S.MontbrioPazoRoxin(), coupling=S.Linear(a=0.1),
                    integrator=S.HeunStochastic(dt=0.1, noise=S.Additive(nsig=np.array([0.01]), noise_seed=9)),
                    monitors=[S.TemporalAverage(period=1.0)], initial_conditions=ic).configure()
    o = OracleSim(w, L, 2.5, 0.1, "MontbrioPazoRoxin", integrator="HeunStochastic", nsig=0.01,
                  initial_conditions=ic, noise_seed=9)
    np.testing.assert_array_equal(s.connectivity.idelays, o.idelays)
    assert s.connectivity.horizon == o.H
    assert s.current_step == o.current_step
    np.testing.assert_array_equal(s.current_state, o.current_state)
    np.testing.assert_array_equal(s.history.buffer[..., 0], o.ring)
    assert s.monitors[0].istep == 10
    # the noise stream continues identically after the initial-condition draw
    np.testing.assert_array_equal(s.integrator.noise.random_stream.normal(size=5), o.rng.normal(size=5))


def test_observables_and_parameter_vectors():
    from tvb_hip import sim as S
    from tvb_hip.backend import _param_vector, observables, MODELS, CFUNS
    m = S.Generic2dOscillator(variables_of_interest=("V", "W", "V + W", "V - W"))
    np.testing.assert_array_equal(observables(m), [[1, 0], [0, 1], [1, 1], [1, -1]])
    orc, sim, _ = make_case("MontbrioPazoRoxin", n=6, scheme="HeunStochastic", nsig=[0.01, 0.02],
                            mparams=dict(eta=np.linspace(-5, -4, 6)))
    vec, spatial = _param_vector(sim, MODELS["MontbrioPazoRoxin"], CFUNS["Linear"][1], True, 6)
    assert vec.shape == (8 + 2 + 2,)
    assert list(spatial) == [4]                # eta per node
    np.testing.assert_allclose(vec[-2:], [0.01, 0.02])


def test_bold_taps_are_forward_hrf():
    from tvb_hip.backend import monitor_spec
    orc, sim, _ = make_case("MontbrioPazoRoxin", n=4, dt=0.05, monitors=(("Bold", 720.0, None),))
    spec = monitor_spec(sim.monitors[0], 0.05)
    G = orc["fp64"].monitors[0].hrf[0, ::-1]
    np.testing.assert_array_equal(spec.taps, G)
    assert spec.istep == 14400 and spec.interim_istep == 80 and len(spec.taps) == 5000
    assert spec.k1_v0 == 5.6 * 0.02


def test_push_fir_schedule_matches_literal_bold():
    """The device Bold keeps one accumulator per pending emission instead of
    the 5000-sample stock: each interim mean m' is added with weight
    G[S-1] (age 0) or G[age-1] to every emission kk with
    kk*istep >= m'*ii and kk*istep < (m'+S)*ii, in a ring of
    Q = (S*ii - 1)//istep + 2 slots.  This restates that schedule on the CPU
    and checks it against the oracle's literal roll/dot Bold, including a
    partial first window and a run split in two calls."""
    from oracle.tvb_oracle import Monitor
    dt, N = 0.1, 2
    for period, length in ((200.0, 400.0), (520.0, 2000.0), (40.0, 400.0)):
        m = Monitor("Bold", period, [0], dt, N, hrf={"length": length})
        S, ii, istep = m._stock_steps, m._interim_istep, m.istep
        G = m.hrf[0, ::-1]
        Q = (S * ii - 1) // istep + 2
        acc = np.zeros((Q, N))
        live = set()
        bsum = np.zeros(N)
        rng = np.random.default_rng(1)
        n_out = 0
        for step in range(23, 23 + 30 * istep):
            x = rng.standard_normal((1, N, 1))
            ref = m.sample(step, x)
            bsum = bsum + x[0, :, 0]
            if step % ii == 0:
                mean, bsum, mp = bsum / ii, np.zeros(N), step // ii
                klo = (mp * ii + istep - 1) // istep
                khi = ((mp + S) * ii - 1) // istep
                for kk in range(klo, khi + 1):
                    age = (kk * istep) // ii - mp
                    assert 0 <= age < S
                    live.add(kk)
                    assert len({k % Q for k in live}) == len(live), "accumulator ring aliasing"
                    acc[kk % Q] += (G[S - 1] if age == 0 else G[age - 1]) * mean
            if step % istep == 0:
                kk = step // istep
                got = (acc[kk % Q] - 1.0) * m.k1_V0
                acc[kk % Q] = 0.0
                live.discard(kk)
                np.testing.assert_allclose(got, ref[1][0, :, 0], rtol=1e-12, atol=1e-14)
                n_out += 1
        assert n_out == 30


def test_synthetic_connectome_properties():
    from tvb_hip.connectome import synthetic, sweep_grid
    w, L = synthetic(68, 0.25, 10.0, 180.0, seed=11)
    assert np.allclose(w, w.T) and np.allclose(L, L.T)
    assert w.max() == 1.0 and np.all(np.diag(w) == 0) and np.all(np.diag(L) == 0)
    nnz_pairs = (np.triu(w, 1) > 0).sum()
    assert nnz_pairs == round(0.25 * 68 * 67 / 2)
    off = L[~np.eye(68, dtype=bool)]
    assert off.min() >= 10.0 and off.max() <= 180.0
    w2, _ = synthetic(68, 0.25, 10.0, 180.0, seed=11)
    np.testing.assert_array_equal(w, w2)
    G, s = sweep_grid()
    assert G.shape == (8192,) and G.min() == 0.0 and G.max() == 2.0
    assert np.isclose(s.min(), 1e-4) and np.isclose(s.max(), 1e-2)


def test_auto_precision_follows_the_reference_arithmetic():
    """"auto" is the reference's mix for Linear coupling and float64 for
    SigmoidalJansenRit (float32 only on request)."""
    import tvb_hip._lib as L
    from tvb_hip.backend import precision_id
    assert precision_id("auto", L.CFUN_LINEAR) == L.MIXED
    assert precision_id("auto", L.CFUN_SIGMOID_JR) == L.F64
    assert precision_id("fp32", L.CFUN_SIGMOID_JR) == L.F32
