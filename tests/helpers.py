"""Builders shared by the tests: one seeded case expressed twice, as the
oracle (numpy restatement) and as a configured simulator for HipBackend."""
import numpy as np

from oracle.tvb_oracle import MODELS, OracleSim, set_idelays
from tvb_hip import sim as S
from tvb_hip.connectome import synthetic

MIRROR_MODELS = {"Generic2dOscillator": S.Generic2dOscillator, "MontbrioPazoRoxin": S.MontbrioPazoRoxin,
                 "JansenRit": S.JansenRit}
MIRROR_SCHEMES = {"HeunDeterministic": S.HeunDeterministic, "HeunStochastic": S.HeunStochastic,
                  "EulerDeterministic": S.EulerDeterministic, "EulerStochastic": S.EulerStochastic}
MIRROR_MONS = {"Raw": S.Raw, "TemporalAverage": S.TemporalAverage, "Bold": S.Bold, "SubSample": S.SubSample,
               "GlobalAverage": S.GlobalAverage, "SpatialAverage": S.SpatialAverage,
               "AfferentCoupling": S.AfferentCoupling,
               "AfferentCouplingTemporalAverage": S.AfferentCouplingTemporalAverage}


def initial_history(model, H, n, seed, spread=1.0):
    """(H, nsv, n) float64 history drawn inside the model's state ranges."""
    rng = np.random.default_rng(seed)
    spec = MODELS[model]
    out = np.empty((H, len(spec["svars"]), n))
    for i, sv in enumerate(spec["svars"]):
        lo, hi = spec["svr"][sv]
        mid, half = (lo + hi) / 2, (hi - lo) / 2 * spread
        out[:, i] = rng.uniform(mid - half, mid + half, size=(H, n))
    return out


def make_case(model, n=16, density=0.3, lengths=(10.0, 60.0), speed=3.0, dt=0.1,
              scheme="HeunDeterministic", coupling="Linear", cparams=None, mparams=None,
              nsig=None, seed=0, monitors=(("Raw", None, None),), spread=1.0, weights=None,
              tract_lengths=None, stimulus=None, initial_conditions=None):
    """``stimulus``: (weights (N,), temporal equation with ``evaluate(t)``,
    simulation_length ms) for a region stimulus in both expressions."""
    if weights is None:
        w, L = synthetic(n, density, lengths[0], lengths[1], seed)
    else:
        w, L = np.asarray(weights, float), np.asarray(tract_lengths, float)
        n = w.shape[0]
    _, _, H = set_idelays(L, speed, dt)
    ic = initial_history(model, H, n, seed + 1000, spread) if initial_conditions is None else initial_conditions
    cparams = dict(cparams or {})
    mparams = dict(mparams or {})
    kw = dict(coupling=coupling, coupling_params=cparams, model_params=mparams,
              integrator=scheme, initial_conditions=ic, monitors=monitors)
    if nsig is not None:
        kw["nsig"] = nsig
    stim_obj = None
    if stimulus is not None:
        sw, teq, length = stimulus
        stim_obj = S.StimuliRegion(temporal=teq, weight=np.asarray(sw, float))
        stim_obj.configure_space()
        stim_obj.configure_time(np.r_[0.0:float(length):dt].reshape((1, -1)))
        kw["stimulus"] = (stim_obj.spatial_pattern[:, 0], stim_obj.temporal_pattern[0])
    orc = {mode: OracleSim(w, L, speed, dt, model, mode=mode, **kw) for mode in ("faithful", "fp64")}
    # mirror simulator
    integ = MIRROR_SCHEMES[scheme](dt=dt)
    if scheme.endswith("Stochastic"):
        integ.noise = S.Additive(nsig=np.asarray(nsig if nsig is not None else 1.0, float))
    mons = []
    for spec_m in monitors:
        kind, period, voi = spec_m[:3]
        kw2 = dict(spec_m[3]) if len(spec_m) > 3 else {}
        if period is not None:
            kw2["period"] = period
        if voi is not None:
            kw2["variables_of_interest"] = np.asarray(voi)
        mons.append(MIRROR_MONS[kind](**kw2))
    cfun = S.Linear(**cparams) if coupling == "Linear" else S.SigmoidalJansenRit(**cparams)
    sim = S.Simulator(connectivity=S.Connectivity(weights=w, tract_lengths=L, speed=np.array([speed])),
                      model=MIRROR_MODELS[model](**mparams), coupling=cfun, integrator=integ,
                      monitors=mons, initial_conditions=ic, stimulus=stim_obj).configure()
    if stimulus is not None:
        sim.simulation_length = float(stimulus[2])
    return orc, sim, ic


def rel_err(a, b):
    """max |a - b| / max |b| (per whole array)."""
    a = np.asarray(a, float)
    b = np.asarray(b, float)
    den = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / (den if den > 0 else 1.0))


def per_region_corr(a, b):
    """Minimum over (voi, node) of the Pearson correlation of two traces
    shaped (time, voi, node[, 1])."""
    a = np.asarray(a, float).reshape(a.shape[0], -1)
    b = np.asarray(b, float).reshape(b.shape[0], -1)
    cs = []
    for k in range(a.shape[1]):
        if np.std(b[:, k]) == 0 or np.std(a[:, k]) == 0:
            continue
        cs.append(np.corrcoef(a[:, k], b[:, k])[0, 1])
    return float(min(cs)) if cs else 1.0


def oracle_record(o):
    """The reference's SimulationHistory.populate_from applied to an oracle
    simulator (simulation_history_h5.py:76-88): ring as the (H, ncvar, N, 1)
    buffer, current state and step, each monitor's ``_stock``."""
    from tvb_hip.checkpoint import SimulationHistory
    h = SimulationHistory(history=o.ring[..., np.newaxis].copy(), current_state=o.current_state.copy(),
                          current_step=o.current_step)
    h.monitor_names = [m.kind for m in o.monitors]
    for i, m in enumerate(o.monitors):
        setattr(h, "monitor_stock_%d" % (i + 1), None if m.kind == "Raw" else m._stock.copy())
    return h


def oracle_fill(o, h, interim=None):
    """SimulationHistory.fill_into applied to a freshly built oracle
    simulator (simulation_history_h5.py:98-107): the interim windows of Bold
    monitors stay as configured (zeros) unless ``interim`` maps a monitor
    index to one."""
    o.ring[...] = np.asarray(h.history)[..., 0]
    o.current_step = int(h.current_step)
    o.current_state = np.array(h.current_state, copy=True)
    for i, m in enumerate(o.monitors):
        st = getattr(h, "monitor_stock_%d" % (i + 1))
        if st is not None:
            m._stock = np.array(st, copy=True)
        if interim and i in interim:
            m._interim_stock = np.array(interim[i], copy=True)
    return o
