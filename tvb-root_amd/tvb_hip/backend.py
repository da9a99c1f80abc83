"""HipBackend: the TVB backend plugin that runs Simulator.__call__ on gfx950.

Plugin surface as in tvb_library/tvb/simulator/backend/nb_mpr.py:79-121:
``check_compatibility(sim)`` raises NotImplementedError for unsupported
components, ``run_sim(sim, nstep=None, simulation_length=None, ...)`` returns
the ``Simulator.run()`` shaped list ``[(t, x), ...]`` (one entry per monitor,
``x.shape == (n, nvoi, N, 1)``).  ``run_batch`` runs many instances (a
parameter sweep over one connectome, or simulators with their own
connectomes and durations) in one launch.

The configured simulator may be a ``tvb.simulator.simulator.Simulator`` or
the mirror in :mod:`tvb_hip.sim`; components are recognised by class name.
"""
import numpy as np

from . import _lib as L
from .plan import CSR, MonitorSpec, Plan

MODELS = {"Generic2dOscillator": L.MODEL_G2D, "MontbrioPazoRoxin": L.MODEL_MPR, "JansenRit": L.MODEL_JR}
MODEL_PARAMS = {
    # raw parameter order of include/tvbk.h
    L.MODEL_G2D: ("tau", "I", "a", "b", "c", "d", "e", "f", "g", "alpha", "beta", "gamma"),
    L.MODEL_MPR: ("tau", "I", "Delta", "J", "eta", "Gamma", "cr", "cv"),
    L.MODEL_JR: ("A", "B", "a", "b", "v0", "nu_max", "r", "J", "a_1", "a_2", "a_3", "a_4", "p_min",
                 "p_max", "mu"),
}
CFUNS = {"Linear": (L.CFUN_LINEAR, ("a", "b")),
         "SigmoidalJansenRit": (L.CFUN_SIGMOID_JR, ("cmin", "cmax", "midpoint", "r", "a"))}
SCHEMES = {"HeunDeterministic": L.HEUN_DET, "HeunStochastic": L.HEUN_STO,
           "EulerDeterministic": L.EULER_DET, "EulerStochastic": L.EULER_STO}
MONITORS = {"Raw": L.MON_RAW, "TemporalAverage": L.MON_TAVG, "Bold": L.MON_BOLD,
            "RawVoi": L.MON_SUBSAMPLE, "SubSample": L.MON_SUBSAMPLE, "GlobalAverage": L.MON_GAVG,
            "SpatialAverage": L.MON_SAVG, "AfferentCoupling": L.MON_AFF,
            "AfferentCouplingTemporalAverage": L.MON_AFF_TAVG}
PRECISIONS = {"fp32": L.F32, "float32": L.F32, "fp64": L.F64, "float64": L.F64, "mixed": L.MIXED}


def precision_id(precision, cfun_id):
    """Plan precision for a name.  "auto" is the reference's own arithmetic:
    its mix (float32 history and sums, float64 state: "mixed") for Linear
    coupling; float64 for SigmoidalJansenRit, whose reference pre and sum run
    in float64 (coupling.py:118-124, :370-373) -- float32 stays an explicit
    choice ("fp32")."""
    if precision == "auto":
        return L.F64 if cfun_id == L.CFUN_SIGMOID_JR else L.MIXED
    return PRECISIONS[precision]


def _name(obj):
    return type(obj).__name__


def observables(model):
    """Linear coefficients of the model's variables_of_interest over its
    state variables (the observer of models/base.py:56-71)."""
    svars = list(model.state_variables)
    eye = np.eye(len(svars))
    ns = {sv: eye[i] for i, sv in enumerate(svars)}
    rows = []
    for expr in model.variables_of_interest:
        try:
            v = eval(expr, {"__builtins__": {}}, dict(ns))
        except Exception:
            raise NotImplementedError("variable of interest %r is not a linear combination of state variables" % expr)
        v = np.asarray(v, np.float64)
        if v.shape != (len(svars),):
            raise NotImplementedError("variable of interest %r is not linear" % expr)
        rows.append(v)
    return np.array(rows)


def monitor_spec(mon, dt, keep_stock=False):
    kind = MONITORS[_name(mon)]
    voi = [int(v) for v in np.atleast_1d(mon.voi)]
    if kind == L.MON_BOLD:
        hrf = np.asarray(mon.hemodynamic_response_function, np.float64).reshape(-1)
        taps = hrf[::-1].copy()                       # G in forward time order
        kern = mon.hrf_kernel
        if _name(kern) == "FirstOrderVolterra":
            k1v0, sub = kern.parameters["k_1"] * kern.parameters["V_0"], True
        else:
            k1v0, sub = 1.0, False
        return MonitorSpec(kind, mon.istep, voi, interim_istep=mon._interim_istep, taps=taps,
                           k1_v0=k1v0, subtract_one=sub, period=mon.period, name="Bold",
                           keep_stock=keep_stock)
    mask = None
    if kind == L.MON_SAVG:
        mask = np.asarray(mon.spatial_mask, np.int32)
    return MonitorSpec(kind, mon.istep, voi, period=getattr(mon, "period", None), name=_name(mon),
                       spatial_mask=mask)


def _param_vector(sim, model_id, cfun_params, stochastic, N):
    """Raw per-instance parameters (and spatial per-node values) of a sim."""
    names = MODEL_PARAMS[model_id]
    vals, spatial = [], {}
    for k, nm in enumerate(names):
        v = np.asarray(getattr(sim.model, nm), np.float64).reshape(-1)
        if v.size == 1:
            vals.append(float(v[0]))
        elif v.size == N:
            vals.append(float(v[0]))
            spatial[k] = v
        else:
            raise NotImplementedError("parameter %s has shape %s" % (nm, v.shape))
    for nm in cfun_params:
        v = np.asarray(getattr(sim.coupling, nm), np.float64).reshape(-1)
        if v.size != 1:
            raise NotImplementedError("coupling parameter %s must be scalar" % nm)
        vals.append(float(v[0]))
    if stochastic:
        nsv = len(sim.model.state_variables)
        nsig = np.asarray(sim.integrator.noise.nsig, np.float64)
        base = len(vals)
        if nsig.size == 1:
            vals += [float(nsig.reshape(-1)[0])] * nsv
        elif nsig.shape in ((nsv, 1, 1), (nsv,)):
            vals += [float(x) for x in nsig.reshape(-1)]
        elif nsig.shape in ((1, N, 1), (N,)):
            vals += [float(nsig.reshape(-1)[0])] * nsv
            for v in range(nsv):
                spatial[base + v] = nsig.reshape(-1)
        elif nsig.shape in ((nsv, N, 1), (nsv, N)):
            n2 = nsig.reshape((nsv, N))
            vals += [float(x) for x in n2[:, 0]]
            for v in range(nsv):
                spatial[base + v] = n2[v]
        else:
            raise NotImplementedError("noise nsig shape %s" % (nsig.shape,))
    return np.array(vals, np.float64), spatial


class HipBackend:
    """Runs configured simulators with the HIP step kernels."""

    # ---------------------------------------------------------------- checks
    def check_compatibility(self, sim):
        def need(cond, what):
            if not cond:
                raise NotImplementedError("Unsupported simulator component. " + what)
        need(getattr(sim, "surface", None) is None, "Surface simulation not supported.")
        stim = getattr(sim, "stimulus", None)
        if stim is not None:
            sp = getattr(stim, "spatial_pattern", None)
            need(sp is not None and np.size(sp) == sim.connectivity.number_of_regions,
                 "Only region stimuli (StimuliRegion with a configured spatial pattern) are supported.")
        need(_name(sim.model) in MODELS, "Given: %s\nExpected one of: %s" % (_name(sim.model), tuple(MODELS)))
        need(getattr(sim.model, "number_of_modes", 1) == 1, "Multi-mode models not supported.")
        need(_name(sim.coupling) in CFUNS, "Given: %s\nExpected one of: %s" % (_name(sim.coupling), tuple(CFUNS)))
        need(not (_name(sim.coupling) == "SigmoidalJansenRit" and _name(sim.model) != "JansenRit"),
             "SigmoidalJansenRit coupling needs a model with two coupling variables (JansenRit).")
        need(_name(sim.integrator) in SCHEMES, "Given: %s\nExpected one of: %s" % (_name(sim.integrator), tuple(SCHEMES)))
        if _name(sim.integrator).endswith("Stochastic"):
            need(_name(sim.integrator.noise) == "Additive", "Only Additive noise is supported.")
            need(float(getattr(sim.integrator.noise, "ntau", 0.0)) == 0.0, "Coloured noise not supported.")
        mons = sim.monitors if isinstance(sim.monitors, (list, tuple)) else [sim.monitors]
        need(1 <= len(mons) <= L.MAX_MON, "Configure with 1 to %d monitors." % L.MAX_MON)
        for m in mons:
            need(_name(m) in MONITORS, "Given: %s\nExpected one of: %s" % (_name(m), tuple(MONITORS)))
        observables(sim.model)

    # ------------------------------------------------------------ plan build
    def _components(self, sim, keep_stock=False):
        model_id = MODELS[_name(sim.model)]
        cfun_id, cfun_params = CFUNS[_name(sim.coupling)]
        scheme = SCHEMES[_name(sim.integrator)]
        stochastic = scheme in (L.HEUN_STO, L.EULER_STO)
        mons = sim.monitors if isinstance(sim.monitors, (list, tuple)) else [sim.monitors]
        specs = [monitor_spec(m, sim.integrator.dt, keep_stock) for m in mons]
        return model_id, cfun_id, cfun_params, scheme, stochastic, specs

    @staticmethod
    def _monitors(sim):
        return sim.monitors if isinstance(sim.monitors, (list, tuple)) else [sim.monitors]

    def _import_monitors(self, plan, inst, sim):
        """Carry the simulator's monitor stocks (TemporalAverage._stock,
        Bold._stock / _interim_stock, monitors.py:372-378, :927-935) into
        the plan at its current step, so that a run continues the averaging
        windows and HRF sums the reference would.  All-zero stocks (a fresh
        configure) match the plan's zeroed state and are skipped."""
        for m, mon in enumerate(self._monitors(sim)):
            stock = getattr(mon, "_stock", None)
            interim = getattr(mon, "_interim_stock", None)
            if stock is None or not plan.monitors[m].stateful:
                continue
            if not np.any(stock) and (interim is None or not np.any(interim)):
                continue
            plan.import_monitor(inst, m, stock, interim)

    def _export_monitors(self, plan, inst, sim):
        """The plan's monitor state back into the simulator's monitor stocks."""
        for m, mon in enumerate(self._monitors(sim)):
            ms = plan.monitors[m]
            if not ms.stateful:
                continue
            if ms.kind == L.MON_BOLD and not ms.keep_stock:
                continue
            stock, interim = plan.export_monitor(inst, m)
            mon._stock = stock
            if interim is not None:
                mon._interim_stock = interim

    def _initial(self, sim, precision, cfun_id, csr):
        """History channels (H, nch, N) and state from the configured sim."""
        chans = getattr(sim.history, "channels", None)
        if chans is not None:
            # the coupling channels a previous device run left (SigmoidalJansenRit:
            # the pre-synaptic transform, which the reference buffer cannot hold)
            return np.ascontiguousarray(chans), np.asarray(sim.current_state, np.float64)[..., 0]
        use64 = precision == L.F64 and getattr(sim.history, "initial_history", None) is not None
        if use64:
            hist = np.asarray(sim.history.initial_history, np.float64)[..., 0]       # (H, nsv, N)
            cv = hist[:, np.asarray(sim.model.cvar)]
        else:
            cv = np.asarray(sim.history.buffer)[..., 0]                              # (H, ncvar, N) f32
        if cfun_id == L.CFUN_SIGMOID_JR:
            c = sim.coupling
            diff = (cv[:, 0] - cv[:, 1])                       # float32 for the reference buffer
            pre = c.cmin + (c.cmax - c.cmin) / (1.0 + np.exp(c.r * (c.midpoint - diff)))
            cv = pre[:, np.newaxis, :].astype(np.float64)
        state = np.asarray(sim.current_state, np.float64)[..., 0]
        return np.ascontiguousarray(cv), state

    def build_plan(self, sim, precision="auto", noise="legacy", max_steps=65536, kernel="auto",
                   tile=0, waves=0, seed=None, stream=None, keep_stock=False):
        """A one-instance Plan from a configured simulator, with its history,
        state and monitor stocks uploaded.  keep_stock: Bold monitors also
        keep their interim-mean stock on the device (for export)."""
        self.check_compatibility(sim)
        model_id, cfun_id, cfun_params, scheme, stochastic, specs = self._components(sim, keep_stock)
        prec = precision_id(precision, cfun_id)
        N = sim.connectivity.number_of_regions
        csr = CSR(sim.connectivity.weights, sim.connectivity.idelays, sim.connectivity.horizon)
        params, spatial = _param_vector(sim, model_id, cfun_params, stochastic, N)
        node_params, mask = None, 0
        if spatial:
            node_params = np.repeat(params[:, None], N, axis=1)
            for k, v in spatial.items():
                node_params[k] = v
                mask |= 1 << k
        if not stochastic:
            noise_id = L.NOISE_NONE
        elif noise == "philox":
            noise_id = L.NOISE_PHILOX
        else:
            noise_id = L.NOISE_INJECTED
        if seed is None:
            seed = int(getattr(getattr(sim.integrator, "noise", None), "noise_seed", 42) or 0)
        kern = {"auto": L.KERNEL_AUTO, "batch": L.KERNEL_BATCH_LANE, "edge": L.KERNEL_EDGE_LANE}[kernel]
        plan = Plan(model=model_id, cfun=cfun_id, scheme=scheme, precision=prec, noise=noise_id,
                    conns=[csr], params=params[None, :], dt=sim.integrator.dt,
                    obs_coef=observables(sim.model), monitors=specs, max_steps=max_steps,
                    node_params=node_params, node_param_mask=mask, seed=seed, kernel=kern,
                    tile=tile, waves=waves, stream=stream)
        hist, state = self._initial(sim, prec, cfun_id, csr)
        plan.set_history(0, hist, state, sim.current_step)
        self._import_monitors(plan, 0, sim)
        return plan

    @staticmethod
    def _stimulus(sim, simulation_length, nstep):
        """(spatial (N,), temporal (>= nstep,), state-variable mask) of the
        simulator's region stimulus for one run, or None.  The temporal
        pattern is sampled on the run's own time grid and indexed by the
        run's step (Simulator._prepare_stimulus / _loop_update_stimulus,
        simulator.py:290-306)."""
        stim = getattr(sim, "stimulus", None)
        if stim is None:
            return None
        length = float(simulation_length if simulation_length is not None else sim.simulation_length)
        time = np.r_[0.0:length:sim.integrator.dt]
        stim.configure_time(time.reshape((1, -1)))
        tp = np.asarray(stim.temporal_pattern, np.float64).reshape(-1)
        sp = np.asarray(stim.spatial_pattern, np.float64).reshape(-1)
        if tp.size < nstep:
            raise ValueError("the stimulus time grid (%d samples of simulation_length %g ms) is shorter than the "
                             "run (%d steps)" % (tp.size, length, nstep))
        stvar = getattr(sim.model, "stvar", None)
        if stvar is None:
            stvar = sim.model.cvar
        mask = 0
        for v in np.atleast_1d(stvar):
            mask |= 1 << int(v)
        return sp, tp, mask

    # --------------------------------------------------------------- running
    def run_sim(self, sim, nstep=None, simulation_length=None, precision="auto", noise="legacy",
                chunk=65536, kernel="auto", write_back=True, **_ignored):
        """Run a configured simulator for nstep steps (or simulation_length ms).

        noise: "legacy" draws the standard normals from the simulator's own
        RandomState (noise.py:195-198) in the reference's order and injects
        them; an array (nstep, nsvar, N) injects those values; "philox" uses
        the device generator.  Returns [(t, x), ...] like Simulator.run."""
        if nstep is None:
            if simulation_length is None:
                simulation_length = sim.simulation_length
            nstep = int(np.ceil(simulation_length / sim.integrator.dt))   # simulator.py:351
        nstep = int(nstep)
        chunk = max(1, min(chunk, nstep))
        philox_noise = isinstance(noise, str) and noise == "philox"
        plan = self.build_plan(sim, precision=precision, noise="philox" if philox_noise else "inject",
                               max_steps=chunk, kernel=kernel, keep_stock=write_back)
        try:
            N = sim.connectivity.number_of_regions
            nsv = len(sim.model.state_variables)
            stochastic = plan.noise == L.NOISE_INJECTED
            ts = [[] for _ in plan.monitors]
            xs = [[] for _ in plan.monitors]
            stim = self._stimulus(sim, simulation_length, nstep)
            done = 0
            while done < nstep:
                n = min(chunk, nstep - done)
                if stim is not None:
                    plan.set_stimulus(stim[0], stim[1][done:done + n], stim[2])
                if stochastic:
                    if isinstance(noise, np.ndarray):
                        z = noise[done:done + n]
                    else:
                        z = sim.integrator.noise.random_stream.normal(size=(n, nsv, N))
                    plan.set_noise(np.ascontiguousarray(np.asarray(z, np.float64).reshape(n, nsv, N, 1)))
                plan.run(n)
                for m in range(len(plan.monitors)):
                    (t, x), = plan.read(m)
                    ts[m].append(t)
                    xs[m].append(np.asarray(x, np.float64))
                done += n
            flags = plan.flags()
            if write_back:
                sjr = _name(sim.coupling) == "SigmoidalJansenRit"
                hist, state, step = plan.export(0, np.float64 if sjr or plan.hsize == 8 else np.float32)
                sim.current_state = state[..., np.newaxis].copy()
                sim.current_step = step
                if sjr:
                    # the ring holds the coupling's pre-synaptic transform of
                    # y1 - y2, not y1 and y2: keep it beside the (now stale)
                    # reference buffer for the next device run
                    sim.history.channels = hist
                else:
                    sim.history.buffer = hist.astype(np.float32)[..., np.newaxis]
                    sim.history.channels = None
                    if plan.hsize == 8:
                        sim.history.channels = hist
                self._export_monitors(plan, 0, sim)
            self.last_flags = flags
            self.last_plan_shape = (plan.kernel, plan.tile, plan.waves, plan.lds_bytes)
        finally:
            plan.close()
        out = []
        for m in range(len(ts)):
            t = np.concatenate(ts[m]) if ts[m] else np.zeros(0)
            x = np.concatenate(xs[m], axis=0) if xs[m] else np.zeros((0,))
            out.append((t, x[..., np.newaxis] if x.ndim == 3 else x))
        return out

    # ---------------------------------------------------------------- batches
    def build_batch_plan(self, sims, precision="auto", noise="inject", max_steps=65536, kernel="auto",
                         tile=0, waves=0, seed=0, streams=None, stream=None):
        """One Plan over several configured simulators of the same model,
        coupling, integrator and monitors.  Simulators whose connectivity has
        identical weights and delays share one connectome (and can use the
        batch-lane kernel); others get their own.  Each instance starts from
        its simulator's history and current state."""
        sims = list(sims)
        if not sims:
            raise ValueError("no simulators")
        for s in sims:
            self.check_compatibility(s)
        model_id, cfun_id, cfun_params, scheme, stochastic, specs = self._components(sims[0])
        prec = precision_id(precision, cfun_id)
        dt = sims[0].integrator.dt
        key0 = (model_id, cfun_id, scheme, [(m.kind, m.istep, tuple(m.voi)) for m in specs])
        conns, inst_conn, sig = [], [], []
        params, spatial_any = [], False
        for s in sims:
            mi, ci, cpn, sc, st, sp = self._components(s)
            if (mi, ci, sc, [(m.kind, m.istep, tuple(m.voi)) for m in sp]) != key0 or s.integrator.dt != dt:
                raise ValueError("all simulators of a batch need the same components, monitors and dt")
            w, d = s.connectivity.weights, s.connectivity.idelays
            for k, (w2, d2) in enumerate(sig):
                if w2.shape == w.shape and np.array_equal(w2, w) and np.array_equal(d2, d):
                    inst_conn.append(k)
                    break
            else:
                sig.append((w, d))
                conns.append(CSR(w, d, s.connectivity.horizon))
                inst_conn.append(len(conns) - 1)
            pv, spatial = _param_vector(s, model_id, cfun_params, stochastic, s.connectivity.number_of_regions)
            spatial_any |= bool(spatial)
            params.append(pv)
        if spatial_any:
            raise NotImplementedError("spatial (per-node) parameters are supported for single simulators only")
        noise_id = L.NOISE_NONE if not stochastic else (L.NOISE_PHILOX if noise == "philox" else L.NOISE_INJECTED)
        kern = {"auto": L.KERNEL_AUTO, "batch": L.KERNEL_BATCH_LANE, "edge": L.KERNEL_EDGE_LANE}[kernel]
        plan = Plan(model=model_id, cfun=cfun_id, scheme=scheme, precision=prec, noise=noise_id, conns=conns,
                    params=np.array(params), dt=dt, obs_coef=observables(sims[0].model), monitors=specs,
                    max_steps=max_steps, inst_conn=np.array(inst_conn, np.int32) if len(conns) > 1 else None,
                    seed=seed, inst_stream=streams, kernel=kern, tile=tile, waves=waves, stream=stream)
        for b, s in enumerate(sims):
            hist, state = self._initial(s, prec, cfun_id, conns[inst_conn[b]])
            plan.set_history(b, hist, state, s.current_step)
            self._import_monitors(plan, b, s)
        return plan

    def run_batch(self, sims, nstep, precision="auto", noise="inject", steps=None, kernel="auto", tile=0,
                  seed=0, streams=None, chunk=65536):
        """Run several simulators in one launch per chunk of steps.  ``nstep``:
        steps of the longest; ``steps``: per-simulator step counts (uneven
        durations).  ``noise``: "inject" draws each simulator's normals from
        its own RandomState in the reference's order, "philox" uses the device
        generator, or a list with one array (n, nsvar, N) per simulator.
        Returns one ``[(t, x), ...]`` per simulator."""
        sims = list(sims)
        nstep = int(nstep)
        chunk = max(1, min(int(chunk), nstep))
        inject = not (isinstance(noise, str) and noise == "philox")
        plan = self.build_batch_plan(sims, precision=precision, noise="inject" if inject else "philox",
                                     max_steps=chunk, kernel=kernel, tile=tile, seed=seed, streams=streams)
        try:
            todo = np.full(len(sims), nstep, np.int64) if steps is None else np.asarray(steps, np.int64)
            if steps is not None:
                plan.set_steps(todo)
            nmax = max(c.n_node for c in plan.conns)
            nsv = len(sims[0].model.state_variables)
            parts = [[[] for _ in plan.monitors] for _ in sims]
            done = 0
            while done < nstep:
                n = min(chunk, nstep - done)
                if plan.noise == L.NOISE_INJECTED:
                    z = np.zeros((n, nsv, nmax, len(sims)))
                    for b, s in enumerate(sims):
                        nb = int(min(n, max(0, todo[b] - done)))   # steps this simulator still takes
                        if nb == 0:
                            continue
                        N = s.connectivity.number_of_regions
                        if isinstance(noise, str):
                            zb = s.integrator.noise.random_stream.normal(size=(nb, nsv, N))
                        else:
                            zb = np.asarray(noise[b][done:done + nb], np.float64)
                        z[:nb, :, :N, b] = zb
                    plan.set_noise(z)
                plan.run(n)
                for m in range(len(plan.monitors)):
                    for b, (t, x) in enumerate(plan.read(m)):
                        parts[b][m].append((t, np.asarray(x, np.float64)))
                done += n
            out = []
            for b in range(len(sims)):
                res = []
                for m in range(len(plan.monitors)):
                    t = np.concatenate([p[0] for p in parts[b][m]])
                    x = np.concatenate([p[1] for p in parts[b][m]], axis=0)
                    res.append((t, x[..., np.newaxis]))
                out.append(res)
            self.last_flags = plan.flags()
            self.last_plan_shape = (plan.kernel, plan.tile, plan.waves, plan.lds_bytes)
        finally:
            plan.close()
        return out

    def build_sweep_plan(self, sim, overrides, precision="auto", noise="philox", max_steps=65536,
                         kernel="auto", tile=0, waves=0, seed=0, init="sim", init_seed=1, stream=None,
                         skip_zero_coupling=False):
        """A Plan of B instances of one simulator with per-instance parameter
        overrides, e.g. {"coupling.a": G, "noise.nsig": sigma, "model.eta": eta}
        (arrays of length B; "noise.nsig" sets every state variable).
        init="sim" starts every instance from the simulator's history;
        init="uniform" draws each instance's history on the device uniformly
        in the model's state_variable_range (seeded by init_seed).
        skip_zero_coupling: coupling variables whose coefficient is zero in
        every instance are not kept or gathered (MontbrioPazoRoxin with
        cv = 0: the V channel, half the sweep's memory traffic); the plan's
        history then holds the remaining channels (``plan.nch``)."""
        self.check_compatibility(sim)
        model_id, cfun_id, cfun_params, scheme, stochastic, specs = self._components(sim)
        prec = precision_id(precision, cfun_id)
        N = sim.connectivity.number_of_regions
        base, spatial = _param_vector(sim, model_id, cfun_params, stochastic, N)
        if spatial:
            raise NotImplementedError("spatial parameters in sweeps are not supported")
        B = len(next(iter(overrides.values())))
        params = np.repeat(base[None, :], B, axis=0)
        names = {("model." + n): k for k, n in enumerate(MODEL_PARAMS[model_id])}
        names.update({("coupling." + n): len(MODEL_PARAMS[model_id]) + k for k, n in enumerate(cfun_params)})
        nsv = len(sim.model.state_variables)
        for key, vals in overrides.items():
            vals = np.asarray(vals, np.float64)
            if vals.shape != (B,):
                raise ValueError("override %s must have shape (%d,)" % (key, B))
            if key == "noise.nsig":
                if not stochastic:
                    raise ValueError("noise.nsig needs a stochastic integrator")
                params[:, -nsv:] = vals[:, None]
            elif key in names:
                params[:, names[key]] = vals
            else:
                raise KeyError("unknown sweep parameter %s" % key)
        noise_id = L.NOISE_NONE if not stochastic else (L.NOISE_PHILOX if noise == "philox" else L.NOISE_INJECTED)
        kern = {"auto": L.KERNEL_AUTO, "batch": L.KERNEL_BATCH_LANE, "edge": L.KERNEL_EDGE_LANE}[kernel]
        csr = CSR(sim.connectivity.weights, sim.connectivity.idelays, sim.connectivity.horizon)
        plan = Plan(model=model_id, cfun=cfun_id, scheme=scheme, precision=prec, noise=noise_id, conns=[csr],
                    params=params, dt=sim.integrator.dt, obs_coef=observables(sim.model), monitors=specs,
                    max_steps=max_steps, seed=seed, kernel=kern, tile=tile, waves=waves, stream=stream,
                    flags=L.PLAN_SKIP_ZERO_COUPLING if skip_zero_coupling else 0)
        if init == "uniform":
            svr = sim.model.state_variable_range
            lo = [float(svr[sv][0]) for sv in sim.model.state_variables]
            hi = [float(svr[sv][1]) for sv in sim.model.state_variables]
            plan.init_uniform(init_seed, lo, hi)
        else:
            hist, state = self._initial(sim, prec, cfun_id, csr)
            plan.set_history(-1, hist[:, :plan.nch], state, sim.current_step)
        return plan
