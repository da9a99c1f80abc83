"""Pins of the oracle against the reference's own known-answer tests and
against what the mathematics fixes (closed forms, brute force), so that a
dropped term, a wrong sign or index, or a transposed operand fails here."""
import numpy as np
import pytest

from oracle import philox
from oracle.tvb_oracle import (MODELS, Monitor, OracleSim, first_order_volterra, history_kat, iround,
                               set_idelays, stationary_trajectory)


# ---------------------------------------------------------------- a2/a3/a9
def test_history_exact_propagation_table():
    """tests/library/simulator/history_test.py:94-112 (expected table :102-111)."""
    _, t, raw = history_kat()
    expected = np.array([[2., 2., 2., 1.], [3., 3., 3., 1.], [5., 4., 4., 1.], [8., 5., 5., 1.],
                         [12., 6., 6., 1.], [17., 7., 7., 1.], [23., 8., 8., 1.], [30., 10., 9., 1.],
                         [38., 13., 10., 1.], [48., 17., 11., 1.]])
    np.testing.assert_array_equal(raw[:, 0, :, 0], expected)
    # Raw time stamps are step * dt (monitors.py:188-190), steps start after current_step = H - 1
    np.testing.assert_array_equal(t, np.arange(16, 26, dtype=float))


def test_time_stamps_from_step_zero():
    """simulator_step_test.py:60-70: with current_step 0, Raw times are (i+1)*dt."""
    w = np.array([[0.0, 1.0], [1.0, 0.0]])
    L = np.array([[0.0, 3.0], [3.0, 0.0]])
    o = OracleSim(w, L, 3.0, 0.1, "Generic2dOscillator", initial_conditions=None)
    assert o.current_step == 0
    (t, x), = o.run(9)
    np.testing.assert_allclose(t, (np.arange(9) + 1) * 0.1)


# ---------------------------------------------------------------- a1
def test_idelays_round_half_even_and_full_horizon():
    """connectivity.py:275, :321-323: rint rounds half to even; horizon over
    the full matrix (zero-weight pairs included)."""
    L = np.array([[0.0, 1.25, 1.75], [1.25, 0.0, 2.25], [1.75, 2.25, 0.0]])
    _, idel, H = set_idelays(L, 1.0, 0.5)      # exact halves: 2.5 -> 2, 3.5 -> 4, 4.5 -> 4
    assert idel[0, 1] == 2 and idel[0, 2] == 4 and idel[1, 2] == 4
    assert H == 5
    w = np.array([[0.0, 1.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 0.0]])
    o = OracleSim(w, L, 1.0, 0.5, "Generic2dOscillator")
    assert o.H == 5 and set(o.nnz_idelays.tolist()) == {2}


def test_iround_kat():
    """backend/ref.py:104-121 docstring example and period/dt cases."""
    assert iround(4.999999999999999) == 5
    assert iround(1.0 / 0.05) == 20
    assert iround(0.9765625 / 0.1) == 10
    assert iround(720.0 / 0.05) == 14400


# ---------------------------------------------------------------- a4/a5
def test_difference_coupling_kat():
    """coupling_test.py:60-85: 2-node Difference coupling, a = 0.1."""
    w = np.array([[0, 1], [1, 0]], float)
    o = OracleSim(w, np.zeros((2, 2)), 1.0, 1.0, "Sum", coupling="Difference", integrator="Identity",
                  initial_conditions=np.array([[[[1.0]], [[2.0]]]]).reshape(1, 1, 2, 1), mode="fp64")
    res = o.coupling_at(1)[0, :, 0]
    np.testing.assert_allclose(res, [0.1 * (1 * (2 - 1)), 0.1 * (1 * (1 - 2))])


def _brute_linear(o, step):
    """Brute-force delayed sum c_i = a * sum_j w_ij x_j(step-1-d_ij) + b."""
    N = o.N
    out = np.zeros((len(o.cvar), N))
    for i in range(N):
        for j in range(N):
            if o.weights[i, j] != 0:
                slot = (step - 1 - o.idelays[i, j]) % o.H
                out[:, i] += o.weights[i, j] * o.ring[slot, :, j]
    return o.cp["a"] * out + o.cp["b"]


@pytest.mark.parametrize("step", [1, 7, 123])
def test_sparse_linear_coupling_vs_brute_force(step):
    """SparseCoupling + Linear (coupling.py:149-200) equals the explicit
    double loop over (target i, source j) with per-edge delays (fp64)."""
    rng = np.random.default_rng(5)
    N = 9
    w = rng.uniform(0, 1, (N, N)) * (rng.uniform(0, 1, (N, N)) < 0.5)
    np.fill_diagonal(w, 0)
    w[:, 0] = 0
    w[3] = 0                                     # a row without afferents gives b
    L = rng.uniform(1, 20, (N, N))
    ic = rng.standard_normal((40, 2, N))
    o = OracleSim(w, L, 2.0, 0.5, "MontbrioPazoRoxin", coupling_params=dict(a=0.7, b=0.1),
                  initial_conditions=ic, mode="fp64")
    o.ring[:] = rng.standard_normal(o.ring.shape)
    np.testing.assert_allclose(o.coupling_at(step)[..., 0], _brute_linear(o, step), rtol=1e-13, atol=1e-14)


def test_sigmoidal_jr_coupling_vs_brute_force():
    """Dense Coupling.__call__ with SigmoidalJansenRit.pre/post (coupling.py:118-124,
    :370-376): c_i = a sum_j w_ij S(y1_j - y2_j) at the delayed time, two equal rows."""
    rng = np.random.default_rng(2)
    N = 7
    w = rng.uniform(0, 1, (N, N)) * (rng.uniform(0, 1, (N, N)) < 0.6)
    np.fill_diagonal(w, 0)
    L = rng.uniform(1, 10, (N, N))
    ic = rng.standard_normal((30, 6, N))
    cp = dict(cmin=0.1, cmax=0.9, midpoint=0.3, r=1.7, a=2.5)
    o = OracleSim(w, L, 1.0, 0.5, "JansenRit", coupling="SigmoidalJansenRit", coupling_params=cp,
                  initial_conditions=ic, mode="fp64")
    step = 11
    got = o.coupling_at(step)[..., 0]
    ref = np.zeros(N)
    for i in range(N):
        for j in range(N):
            if w[i, j] != 0:
                slot = (step - 1 - o.idelays[i, j]) % o.H
                y1, y2 = o.ring[slot, 0, j], o.ring[slot, 1, j]
                ref[i] += w[i, j] * (cp["cmin"] + (cp["cmax"] - cp["cmin"]) / (1 + np.exp(cp["r"] * (cp["midpoint"] - (y1 - y2)))))
    np.testing.assert_allclose(got[0], cp["a"] * ref, rtol=1e-12)
    np.testing.assert_array_equal(got[0], got[1])


def test_faithful_mode_sums_in_float32():
    """history.py:49-51/:218 and coupling.py:161-166: the faithful Linear sum
    is float32 (differs from the fp64 sum by float32 rounding only)."""
    rng = np.random.default_rng(1)
    N = 12
    w = rng.lognormal(0, 1, (N, N))
    np.fill_diagonal(w, 0)
    L = rng.uniform(1, 10, (N, N))
    ic = rng.standard_normal((30, 2, N))
    f = OracleSim(w, L, 1.0, 0.5, "MontbrioPazoRoxin", initial_conditions=ic, coupling_params=dict(a=1.0))
    d = OracleSim(w, L, 1.0, 0.5, "MontbrioPazoRoxin", initial_conditions=ic, coupling_params=dict(a=1.0), mode="fp64")
    assert f.ring.dtype == np.float32 and d.ring.dtype == np.float64
    cf, cd = f.coupling_at(5), d.coupling_at(5)
    assert cf.dtype == np.float64
    assert 0 < np.max(np.abs(cf - cd)) < 1e-5 * np.max(np.abs(cd))


# ---------------------------------------------------------------- a6
def test_g2d_fixed_point():
    """models_test.py:196-226 via Model.stationary_trajectory (models/base.py:169-210)."""
    svr = MODELS["Generic2dOscillator"]["svr"]
    init = np.array([[(svr["V"][0] + svr["V"][1]) / 2], [(svr["W"][0] + svr["W"][1]) / 2]])
    y = stationary_trajectory("Generic2dOscillator", init)
    assert y.shape == (201, 2, 1, 1)
    np.testing.assert_allclose(y[-1, :, 0, 0], [1.048535, -4.507975], 1e-6, 1e-3)


def test_jr_fixed_point():
    """models_test.py:237-243."""
    svr = MODELS["JansenRit"]["svr"]
    init = np.array([[(lo + hi) / 2] for lo, hi in svr.values()])
    y = stationary_trajectory("JansenRit", init)
    fp = [1.105518e-02, 4.790847e+00, 8.826628e-01, 1.308187e-03, 2.367387e-01, 6.615113e-02]
    np.testing.assert_allclose(y[-1, :, 0, 0], fp, 1e-6, 1e-3)


def test_mpr_dfun_closed_form():
    """infinite_theta.py:171-207 against the model equations written out."""
    o = OracleSim(np.zeros((1, 1)), np.zeros((1, 1)), 1.0, 0.1, "MontbrioPazoRoxin", mode="fp64",
                  model_params=dict(tau=2.0, Delta=0.7, J=12.0, eta=-4.0, I=0.3, cr=0.8, cv=0.4))
    r, V, cr_, cv_ = 0.6, -1.3, 0.25, -0.5
    d = o.dfun(np.array([[[r]], [[V]]]), np.array([[[cr_]], [[cv_]]]))[:, 0, 0]
    tau, De, J, eta, I = 2.0, 0.7, 12.0, -4.0, 0.3
    assert np.isclose(d[0], (De / (np.pi * tau) + 2 * V * r) / tau)
    assert np.isclose(d[1], (V * V - (np.pi * tau * r) ** 2 + eta + J * tau * r + I + 0.8 * cr_ + 0.4 * cv_) / tau)


def _mpr_fixed_points(tau, Delta, J, eta, I):
    """Uncoupled MontbrioPazoRoxin fixed points from the model's equations
    alone: dr = 0 gives V = -Delta / (2 pi tau r); dV = 0 then gives the
    quartic -pi^2 tau^2 r^4 + J tau r^3 + (eta + I) r^2 + Delta^2 / (4 pi^2
    tau^2) = 0, solved as a polynomial (numpy.roots) -- not by the oracle."""
    roots = np.roots([-np.pi ** 2 * tau ** 2, J * tau, eta + I, 0.0, Delta ** 2 / (4 * np.pi ** 2 * tau ** 2)])
    rs = sorted(r.real for r in roots if abs(r.imag) < 1e-9 and r.real > 0)
    return [(r, -Delta / (2 * np.pi * tau * r)) for r in rs]


@pytest.mark.parametrize("params", [{}, dict(tau=2.0, Delta=0.7, J=20.0, eta=-4.0, I=0.3)])
def test_mpr_dfun_vanishes_at_independent_fixed_points(params):
    """The oracle's MPR dfun (infinite_theta.py:171-207) is zero at the
    fixed points the quartic fixes (two parameter sets), and, with the
    default parameters, Model.stationary_trajectory (models/base.py:169-210,
    as in models_test.py:174-243) converges to the two attracting ones from
    nearby starts while the middle one (a saddle) is left.  A dropped term, a wrong sign or factor moves the fixed points
    and fails both checks."""
    from oracle.tvb_oracle import MODELS
    p = dict(MODELS["MontbrioPazoRoxin"]["defaults"])
    p.update(params)
    fps = _mpr_fixed_points(p["tau"], p["Delta"], p["J"], p["eta"], p["I"])
    assert len(fps) == 3
    o = OracleSim(np.zeros((1, 1)), np.zeros((1, 1)), 1.0, 0.1, "MontbrioPazoRoxin", mode="fp64", model_params=params)
    zero_c = np.zeros((2, 1, 1))
    for r, V in fps:
        d = o.dfun(np.array([[[r]], [[V]]]), zero_c)[:, 0, 0]
        assert np.abs(d).max() < 1e-9 * max(1.0, abs(V) ** 2), (r, V, d)
    if params:          # the convergence checks use the default parameters
        return
    lo, mid, hi = fps
    for fp in (lo, hi):
        for f in (0.95, 1.05):
            y = stationary_trajectory("MontbrioPazoRoxin", [fp[0] * f, fp[1] * (2 - f)], coupling=np.zeros((2, 1)),
                                      n_step=40000, dt=0.001, params=params)
            np.testing.assert_allclose(y[-1].ravel(), fp, rtol=1e-4, atol=1e-5)
    ends = [stationary_trajectory("MontbrioPazoRoxin", [mid[0] * f, mid[1] * (2 - f)], coupling=np.zeros((2, 1)),
                                  n_step=40000, dt=0.001, params=params)[-1].ravel() for f in (0.9, 1.1)]
    assert all(np.abs(e - np.array(mid)).max() > 0.1 for e in ends)


def test_mpr_bound_applied():
    """r >= 0 after predictor and corrector (infinite_theta.py:129-134, integrators.py:355/360)."""
    ic = np.zeros((1, 2, 1))
    ic[0, 0, 0] = -1.0
    o = OracleSim(np.zeros((1, 1)), np.zeros((1, 1)), 1.0, 0.1, "MontbrioPazoRoxin", initial_conditions=ic)
    assert o.current_state[0, 0, 0] == 0.0   # history slices are bounded too (history.py:150-152)
    o.current_state[:] = np.array([[[-5.0]], [[0.0]]])
    x = o.scheme(o.current_state.copy(), np.zeros((2, 1, 1)), None)
    assert x[0, 0, 0] >= 0.0


# ---------------------------------------------------------------- a7/a8
def test_heun_and_euler_closed_forms():
    """integrators.py:300-318 / :377-391 on dx = x + c (the Sum model):
    Heun x' = x + dt/2 (f(x) + f(x + dt f(x))); Euler x' = x + dt f(x)."""
    for scheme in ("HeunDeterministic", "EulerDeterministic"):
        o = OracleSim(np.zeros((1, 1)), np.zeros((1, 1)), 1.0, 0.1, "Sum", coupling="IdCoupling",
                      integrator=scheme, mode="fp64")
        x, c, dt = 1.5, 0.25, 0.1
        got = o.scheme(np.array([[[x]]]), np.array([[[c]]]), None)[0, 0, 0]
        f = lambda y: y + c
        ref = x + dt / 2 * (f(x) + f(x + dt * f(x))) if scheme.startswith("Heun") else x + dt * f(x)
        assert np.isclose(got, ref, rtol=1e-15)


def test_stochastic_increment_closed_form():
    """HeunStochastic with zero drift adds xi twice to x... no: predictor and
    corrector both add the same xi = sqrt(dt) z sqrt(2 nsig) (integrators.py:344-360);
    EulerStochastic adds it once (:418-424)."""
    dt, nsig, z = 0.05, 0.02, 1.3
    xi = np.sqrt(dt) * z * np.sqrt(2 * nsig)
    for scheme in ("HeunStochastic", "EulerStochastic"):
        o = OracleSim(np.zeros((1, 1)), np.zeros((1, 1)), 1.0, dt, "Sum", coupling="IdCoupling",
                      integrator=scheme, nsig=nsig, mode="fp64")
        # Sum model with x = 0, c = 0 has zero drift at x=0; Heun's corrector sees f(xi) = xi
        got = o.scheme(np.zeros((1, 1, 1)), np.zeros((1, 1, 1)), np.array([[[z]]]))[0, 0, 0]
        ref = (0 + dt / 2 * (0 + xi)) + xi if scheme == "HeunStochastic" else xi
        assert np.isclose(got, ref, rtol=1e-14)


def test_legacy_noise_stream_equals_per_step_draws():
    """noise.py:195-198: drawing (n, nsv, N) at once equals n per-step draws."""
    a = np.random.RandomState(42)
    b = np.random.RandomState(42)
    whole = a.normal(size=(5, 2, 3))
    per = np.array([b.normal(size=(2, 3, 1))[..., 0] for _ in range(5)])
    np.testing.assert_array_equal(whole, per)


# ---------------------------------------------------------------- a11/a12
def test_numpy_orders_the_device_reproduces():
    """The device sums a CSR row as numpy add.reduceat does, x[0] +
    pairwise(x[1:]) with numpy's 8-accumulator blocks, and averages monitor
    windows sequentially in time as numpy.mean over axis 0 does."""
    def pw(x):
        n = len(x)
        if n < 8:
            r = 0.0
            for v in x:
                r = r + v
            return r
        if n <= 128:
            acc = list(x[:8])
            i = 8
            while i < n - n % 8:
                for k in range(8):
                    acc[k] = acc[k] + x[i + k]
                i += 8
            r = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]))
            for v in x[i:]:
                r = r + v
            return r
        n2 = n // 2
        n2 -= n2 % 8
        return pw(x[:n2]) + pw(x[n2:])
    rng = np.random.default_rng(0)
    for n in (1, 2, 5, 9, 17, 40, 129, 130, 300, 1025):
        for dt in (np.float32, np.float64):
            x = (rng.standard_normal(n) * rng.lognormal(0, 3, n)).astype(dt)
            ref = np.add.reduceat(x, [0])[0]
            got = x[0] if n == 1 else dt(x[0] + dt(pw([dt(v) for v in x[1:]])))
            assert got == ref, (n, dt)
    st = rng.standard_normal((20, 2, 5, 1))
    seq = st[0].copy()
    for k in range(1, 20):
        seq = seq + st[k]
    np.testing.assert_array_equal(np.mean(st, axis=0), seq / 20)


def test_tavg_constant_input_and_time_stamps():
    """monitors.py:380-391: mean of a constant is the constant; time (step - istep/2) dt."""
    m = Monitor("TemporalAverage", 1.0, [0], 0.05, 3)
    outs = [m.sample(s, np.full((1, 3, 1), 2.5)) for s in range(1, 41)]
    got = [o for o in outs if o is not None]
    assert len(got) == 2
    assert np.allclose(got[0][1], 2.5) and np.isclose(got[0][0], (20 - 10) * 0.05)


def test_bold_hrf_peak():
    """FirstOrderVolterra peaks at ~0.80 s (exploring_the_bold_monitor.ipynb:234);
    analytic peak of exp(-t/2ts) sin(qt): tan(qt) = 2 ts q -> 0.8017 s."""
    t = np.arange(0.0, 20.0, 20.0 / 5000)
    G = first_order_volterra(t)
    assert abs(t[np.argmax(G)] - 0.80) < 0.01
    q = np.sqrt(1 / 0.4 - 1 / (4 * 0.8 ** 2))
    assert abs(np.arctan(2 * 0.8 * q) / q - 0.8017) < 1e-3


def test_bold_roll_equals_age_weights():
    """Bold.sample's roll/dot (monitors.py:942-962) equals the age form the
    device uses: a sample of age a >= 1 (4 ms units) weighs G[a-1], the
    newest (age 0) weighs G[S-1]; unwritten stock slots are zero."""
    dt, N = 0.1, 3
    m = Monitor("Bold", 200.0, [0], dt, N, hrf={"length": 400.0})      # S = 100 samples
    S, ii, istep = m._stock_steps, m._interim_istep, m.istep
    G = m.hrf[0, ::-1]
    rng = np.random.default_rng(4)
    means = {}
    start = 37                                   # mid-window start: partial first mean
    acc = np.zeros((N,))
    checked = 0
    interim = np.zeros((N,))
    for step in range(start, start + 12000):
        x = rng.standard_normal((1, N, 1))
        out = m.sample(step, x)
        interim = interim + x[0, :, 0]
        if step % ii == 0:
            means[step // ii] = interim / ii
            interim = np.zeros((N,))
        if out is not None:
            M = step // ii
            ref = np.zeros(N)
            for mp, v in means.items():
                a = M - mp
                if 0 <= a < S:
                    ref += (G[S - 1] if a == 0 else G[a - 1]) * v
            np.testing.assert_allclose(out[1][0, :, 0], (ref - 1.0) * m.k1_V0, rtol=1e-12, atol=1e-14)
            checked += 1
    assert checked >= 5


# ---------------------------------------------------------------- Philox
@pytest.mark.parametrize("rounds,ctr,key,expected", [
    (10, [0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    (10, [0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    (10, [0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
    (7, [0, 0, 0, 0], [0, 0], [0x5f6fb709, 0x0d893f64, 0x4f121f81, 0x4f730a48]),
])
def test_philox_known_answers(rounds, ctr, key, expected):
    """Random123 known-answer vectors for Philox4x32 (Salmon et al., SC'11)."""
    out = philox.philox4x32(np.array(ctr, np.uint32), np.array(key, np.uint32), rounds)
    assert [int(v) for v in out] == expected


@pytest.mark.parametrize("precision,bits", [("f64", 32), ("f32", 24)])
def test_box_muller_closed_form_cases(precision, bits):
    """Box-Muller (Box and Muller 1958) on words that fix the answer:
    u1 = (w0 + 1) 2^-b, u2 = w1 2^-b with b = 32 (f64) or the top 24 bits
    (f32).  u1 = 1 gives radius 0; u2 = 0 gives (rad, 0), u2 = 1/4 gives
    (0, rad), u2 = 1/2 gives (-rad, 0); the smallest u1 gives the largest
    radius sqrt(2 b ln 2).  The same cases on the second pair (w2, w3); a
    swapped u1/u2, sin/cos or pair order fails them."""
    sh = 32 - bits
    top = (1 << 32) - 1                        # w0 = 2^32 - 1: u1 = 1
    quarter, half = 1 << 30, 1 << 31
    words = np.array([[top, 0, 0, 0], [0, 0, top, quarter], [0, half, 0, quarter]], np.uint32)
    z = philox.box_muller(words, precision)
    rmax = np.sqrt(2.0 * bits * np.log(2.0))
    np.testing.assert_allclose(z[0], [0.0, 0.0, rmax, 0.0], atol=1e-12)
    np.testing.assert_allclose(z[1], [rmax, 0.0, 0.0, 0.0], atol=1e-12)
    np.testing.assert_allclose(z[2], [-rmax, 0.0, 0.0, rmax], atol=1e-12)
    # a generic word: the published formulas evaluated directly
    w = np.array([[123456789, 987654321, 3141592653, 2718281828]], np.uint32)
    u = (w.astype(np.float64) if sh == 0 else (w >> np.uint32(sh)).astype(np.float64)) * 2.0 ** -bits
    r1, r2 = np.sqrt(-2 * np.log(u[0, 0] + 2.0 ** -bits)), np.sqrt(-2 * np.log(u[0, 2] + 2.0 ** -bits))
    ref = [r1 * np.cos(2 * np.pi * u[0, 1]), r1 * np.sin(2 * np.pi * u[0, 1]),
           r2 * np.cos(2 * np.pi * u[0, 3]), r2 * np.sin(2 * np.pi * u[0, 3])]
    np.testing.assert_allclose(philox.box_muller(w, precision)[0], ref, rtol=1e-14)


def test_philox_normals_statistics():
    z = philox.normals(seed=123, stream=7, n_steps=2000, nsvar=2, n_node=16)
    assert z.shape == (2000, 2, 16)
    assert abs(z.mean()) < 0.02 and abs(z.std() - 1) < 0.02
    z2 = philox.normals(seed=123, stream=8, n_steps=10, nsvar=2, n_node=16)
    assert not np.allclose(z[:10], z2)
