"""Simulation checkpoints in the reference's SimulationHistory format.

``SimulationHistory`` mirrors
tvb_framework/tvb/core/entities/file/simulator/simulation_history_h5.py:32-117
(same field names and meaning): the history ring, current state and step,
one ``monitor_stock_<i>`` per monitor with ``monitor_names``, and the noise
generator state.  ``populate_from`` / ``fill_into`` act on a simulator as the
reference's do; ``from_plan`` / ``to_plan`` act on one instance of a device
plan (a batched sweep or a serving plan), whose state never leaves the device
between runs otherwise.

Files: h5py is not part of this environment, so ``save`` / ``load`` write the
datasets of SimulationHistoryH5 (simulation_history_h5.py:120-138) under the
same names into a numpy ``.npz`` archive (no pickled objects; monitor_names
as JSON text, as the H5 file stores it).

Device monitor state: TemporalAverage and the Bold interim window are kept as
running sums, and export places the sum in slot 0 of the stock (the other
slots zero), which the reference's sequential mean over axis 0 turns into the
same value.  Bold's ``_stock`` of interim means is kept only by plans whose
Bold monitor was built with ``keep_stock``.  Noise: a Philox plan's stream is
fully determined by (seed, instance stream, step); it is recorded as algo
"Philox4x32-10", keys (seed low, seed high, stream low, stream high) and
pos = current step.
"""
import json

import numpy as np

from . import _lib as L

N_STOCKS = 15
RNG_FIELDS = ("integrator_noise_rng_state_algo", "integrator_noise_rng_state_keys",
              "integrator_noise_rng_state_pos", "integrator_noise_rng_state_has_gauss",
              "integrator_noise_rng_state_cached_gauss")
PHILOX = "Philox4x32-10"


def _name(obj):
    return type(obj).__name__


class SimulationHistory:
    """simulation_history_h5.py:32-117."""

    def __init__(self, **kwargs):
        self.history = None
        self.current_state = None
        self.current_step = 0
        for i in range(1, N_STOCKS + 1):
            setattr(self, "monitor_stock_%d" % i, None)
        self.monitor_names = []
        # interim windows of Bold monitors: not part of the reference's record
        # (fill_into leaves a freshly configured, zero interim stock); kept
        # here so that a device-to-device resume is exact
        self.monitor_interim = {}
        for f in RNG_FIELDS:
            setattr(self, f, None)
        for k, v in kwargs.items():
            if not hasattr(self, k):
                raise TypeError("unknown SimulationHistory field %r" % k)
            setattr(self, k, v)

    # -------------------------------------------------- simulator interface
    def populate_from(self, simulator_algorithm):
        """simulation_history_h5.py:76-96."""
        sim = simulator_algorithm
        if getattr(sim.history, "channels", None) is not None and _name(sim.coupling) == "SigmoidalJansenRit":
            raise NotImplementedError("after a device run the SigmoidalJansenRit history holds the coupling's "
                                      "pre-synaptic transform, not the y1 and y2 the reference buffer stores")
        self.history = np.array(sim.history.buffer, copy=True)
        self.current_step = int(sim.current_step)
        self.current_state = np.array(sim.current_state, copy=True)
        mons = sim.monitors if isinstance(sim.monitors, (list, tuple)) else [sim.monitors]
        if len(mons) > N_STOCKS:
            raise ValueError("at most %d monitors" % N_STOCKS)
        self.monitor_names = []
        for i, mon in enumerate(mons):
            setattr(self, "monitor_stock_%d" % (i + 1), getattr(mon, "_stock", None))
            self.monitor_names.append(_name(mon))
        integ = sim.integrator
        if _name(integ).endswith("Stochastic"):       # IntegratorStochastic (:90)
            rs = integ.noise.random_stream.get_state()
            self.integrator_noise_rng_state_algo = rs[0]
            self.integrator_noise_rng_state_keys = np.asarray(rs[1], np.uint32)
            self.integrator_noise_rng_state_pos = int(rs[2])
            self.integrator_noise_rng_state_has_gauss = int(rs[3])
            self.integrator_noise_rng_state_cached_gauss = float(rs[4])
        return self

    def fill_into(self, simulator_algorithm):
        """simulation_history_h5.py:98-117: history ring, step, state and
        monitor ``_stock`` (interim windows stay as configured), and the
        noise stream position."""
        sim = simulator_algorithm
        sim.history.buffer = np.array(self.history, copy=True)
        if hasattr(sim.history, "channels"):
            sim.history.channels = None
        sim.current_step = int(self.current_step)
        sim.current_state = np.array(self.current_state, copy=True)
        mons = sim.monitors if isinstance(sim.monitors, (list, tuple)) else [sim.monitors]
        for i, mon in enumerate(mons):
            mon._stock = getattr(self, "monitor_stock_%d" % (i + 1))
        if self.integrator_noise_rng_state_algo is not None:
            if self.integrator_noise_rng_state_algo == PHILOX:
                raise ValueError("a Philox record resumes device plans only (to_plan)")
            sim.integrator.noise.random_stream.set_state((
                self.integrator_noise_rng_state_algo,
                np.asarray(self.integrator_noise_rng_state_keys, np.uint32),
                int(self.integrator_noise_rng_state_pos),
                int(self.integrator_noise_rng_state_has_gauss),
                float(self.integrator_noise_rng_state_cached_gauss)))
        return sim

    # -------------------------------------------------------- plan interface
    @classmethod
    def from_plan(cls, plan, inst, monitor_names=None, interim=True):
        """State of instance ``inst`` of a device plan (synchronises the
        plan's stream).  history (H, ncvar, N, 1) in the ring's precision
        (float32 for float32 and mixed plans, the reference buffer's dtype)."""
        if plan.cfun == L.CFUN_SIGMOID_JR:
            raise NotImplementedError("the SigmoidalJansenRit ring holds the coupling's pre-synaptic transform, "
                                      "not the y1 and y2 the reference buffer stores")
        hist, state, step = plan.export(inst, np.float64 if plan.hsize == 8 else np.float32)
        h = cls(history=hist[..., np.newaxis], current_state=state[..., np.newaxis], current_step=step)
        names = list(monitor_names) if monitor_names else [ms.name or str(ms.kind) for ms in plan.monitors]
        h.monitor_names = names
        for m, ms in enumerate(plan.monitors):
            if not ms.stateful:
                continue
            if ms.kind == L.MON_BOLD and not ms.keep_stock:
                raise ValueError("monitor %d: a Bold stock is only exported by a plan built with keep_stock" % m)
            stock, it = plan.export_monitor(inst, m, stock=True, interim=interim)
            setattr(h, "monitor_stock_%d" % (m + 1), stock)
            if it is not None:
                h.monitor_interim[m + 1] = it
        if plan.noise == L.NOISE_PHILOX:
            seed = int(plan._desc.seed)
            stream = int(plan.inst_stream[inst]) if plan.inst_stream is not None else int(inst)
            h.integrator_noise_rng_state_algo = PHILOX
            h.integrator_noise_rng_state_keys = np.array(
                [seed & 0xFFFFFFFF, seed >> 32, stream & 0xFFFFFFFF, stream >> 32], np.uint32)
            h.integrator_noise_rng_state_pos = int(step)
            h.integrator_noise_rng_state_has_gauss = 0
            h.integrator_noise_rng_state_cached_gauss = 0.0
        return h

    def to_plan(self, plan, inst):
        """Load this record into instance ``inst`` (-1: every instance) of a
        plan: history, state, step, then each monitor's stock (and the Bold
        interim window when the record holds one).  A Philox plan continues
        its own stream at the recorded step; branching with another seed or
        another set of parameters is allowed."""
        hist = np.asarray(self.history)
        if hist.ndim == 4:
            hist = hist[..., 0]
        state = np.asarray(self.current_state, np.float64)
        if state.ndim == 3:
            state = state[..., 0]
        plan.set_history(inst, hist, state, int(self.current_step))
        for m, ms in enumerate(plan.monitors):
            stock = getattr(self, "monitor_stock_%d" % (m + 1))
            if not ms.stateful or stock is None:
                continue
            plan.import_monitor(inst, m, stock, self.monitor_interim.get(m + 1))
        return plan

    # ---------------------------------------------------------------- files
    def save(self, path):
        """Datasets of SimulationHistoryH5 into an .npz archive."""
        arrays = {"current_step": np.int64(self.current_step),
                  "monitor_names": np.array(json.dumps(list(self.monitor_names)))}
        for k in ("history", "current_state"):
            if getattr(self, k) is not None:
                arrays[k] = np.asarray(getattr(self, k))
        for i in range(1, N_STOCKS + 1):
            v = getattr(self, "monitor_stock_%d" % i)
            if v is not None:
                arrays["monitor_stock_%d" % i] = np.asarray(v)
        for i, v in self.monitor_interim.items():
            arrays["monitor_interim_%d" % i] = np.asarray(v)
        if self.integrator_noise_rng_state_algo is not None:
            arrays["integrator_noise_rng_state_algo"] = np.array(str(self.integrator_noise_rng_state_algo))
            arrays["integrator_noise_rng_state_keys"] = np.asarray(self.integrator_noise_rng_state_keys, np.uint32)
            arrays["integrator_noise_rng_state_pos"] = np.int64(self.integrator_noise_rng_state_pos)
            arrays["integrator_noise_rng_state_has_gauss"] = np.int64(self.integrator_noise_rng_state_has_gauss)
            arrays["integrator_noise_rng_state_cached_gauss"] = np.float64(
                self.integrator_noise_rng_state_cached_gauss)
        with open(path, "wb") as f:
            np.savez(f, **arrays)

    @classmethod
    def load(cls, path):
        h = cls()
        with np.load(path, allow_pickle=False) as z:
            for k in z.files:
                v = z[k]
                if k == "monitor_names":
                    h.monitor_names = json.loads(str(v))
                elif k.startswith("monitor_interim_"):
                    h.monitor_interim[int(k.rsplit("_", 1)[1])] = v
                elif k == "integrator_noise_rng_state_algo":
                    h.integrator_noise_rng_state_algo = str(v)
                elif k in ("current_step", "integrator_noise_rng_state_pos", "integrator_noise_rng_state_has_gauss"):
                    setattr(h, k, int(v))
                elif k == "integrator_noise_rng_state_cached_gauss":
                    h.integrator_noise_rng_state_cached_gauss = float(v)
                else:
                    setattr(h, k, v)
        return h
