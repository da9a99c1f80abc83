"""Balloon-Windkessel BOLD restatement (test infrastructure only).

Restates ``BalloonModel.evaluate`` (tvb_library/tvb/analyzers/fmri_balloon.py:
198-275) for the default neural-input transformation "none" (first state
variable of the time series, :314-318) with the HeunDeterministic scheme it
uses by default (:91-99; simulator/integrators.py:296-318):

* integrator step dt = 1 / sample_rate in seconds (:219-221; sample_rate of a
  time series in ms is 1000 / sample_period, datatypes/time_series.py:84-93);
* state (s, f, v, q) starts at (0, 1, 1, 1) at the first sample (:208-213) and
  sample k >= 1 advances it with the input of sample k (:239-241);
* balloon_dfun (:330-357), the revised / classical coefficients (:277-303) and
  the non-linear / linear BOLD equations (:253-268).

Only tests/, the smoke check and the benchmark's CPU leg may import this.
"""
import numpy as np

DEFAULTS = dict(tau_s=1.54, tau_f=1.44, tau_o=0.98, alpha=0.32, TE=0.04, V0=4.0, E0=0.4, epsilon=0.5,
                nu_0=40.3, r_0=25.0)


def coefficients(p, rbm=True):
    """k1, k2, k3 (fmri_balloon.py:277-303)."""
    if not rbm:
        return 7.0 * p["E0"], 2.0 * p["E0"], 1.0 - p["epsilon"]
    return (4.3 * p["nu_0"] * p["E0"] * p["TE"], p["epsilon"] * p["r_0"] * p["E0"] * p["TE"],
            1 - p["epsilon"])


def dfun(state, x, p):
    """balloon_dfun (fmri_balloon.py:330-357); state (4, ...), x (...)."""
    s, f, v, q = state
    ds = x - (1.0 / p["tau_s"]) * s - (1.0 / p["tau_f"]) * (f - 1)
    df = s
    dv = (1.0 / p["tau_o"]) * (f - v ** (1.0 / p["alpha"]))
    dq = (1.0 / p["tau_o"]) * ((f * (1.0 - (1.0 - p["E0"]) ** (1.0 / f)) / p["E0"]) -
                              (v ** (1.0 / p["alpha"])) * (q / v))
    return np.array([ds, df, dv, dq])


def heun(state, x, dt, p):
    """HeunDeterministic.scheme with stimulus 0 and no bounds (integrators.py:309-318)."""
    m = dfun(state, x, p)
    inter = state + dt * (m + 0.0)
    dx = (m + dfun(inter, x, p)) * dt / 2.0
    return state + dx + dt * 0.0


def bold(state, p, nonlinear=True, rbm=True):
    """BOLD signal from (s, f, v, q) (fmri_balloon.py:253-268)."""
    k1, k2, k3 = coefficients(p, rbm)
    _, _, v, q = state
    if nonlinear:
        return p["V0"] * (k1 * (1. - q) + k2 * (1. - q / v) + k3 * (1. - v))
    return p["V0"] * ((k1 + k2) * (1. - q) + (k3 - k2) * (1. - v))


def evaluate(x, sample_period_ms, params=None, nonlinear=True, rbm=True, normalize=False, state=None):
    """BOLD of input series x (T, ...) sampled every ``sample_period_ms``.

    With ``state`` None the series starts the model (y[0] from the initial
    state); otherwise ``state`` (4, ...) continues a previous call and every
    sample advances it.  Returns (y (T, ...), final state)."""
    p = dict(DEFAULTS, **(params or {}))
    x = np.asarray(x, np.float64)
    if normalize:
        x = x - x.mean(axis=0)[np.newaxis]
    dt = 1.0 / (1000.0 / sample_period_ms)
    y = np.empty_like(x)
    if state is None:
        st = np.zeros((4,) + x.shape[1:])
        st[1:] = 1.0
        y[0] = bold(st, p, nonlinear, rbm)
        k0 = 1
    else:
        st = np.array(state, np.float64)
        k0 = 0
    for k in range(k0, x.shape[0]):
        st = heun(st, x[k], dt, p)
        y[k] = bold(st, p, nonlinear, rbm)
    return y, st


def steady_state(x, params=None):
    """Fixed point of balloon_dfun for a constant input x (closed form):
    s = 0, f = 1 + tau_f x, v = f^alpha, q = v (1 - (1 - E0)^(1/f)) / E0."""
    p = dict(DEFAULTS, **(params or {}))
    f = 1.0 + p["tau_f"] * x
    v = f ** p["alpha"]
    q = v * (1.0 - (1.0 - p["E0"]) ** (1.0 / f)) / p["E0"]
    return np.array([0.0 * np.asarray(x, float), f, v, q])
