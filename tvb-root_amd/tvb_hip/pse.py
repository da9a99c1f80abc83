"""Parameter space exploration as one batched launch.

The framework explores a parameter space by launching one simulation per
grid point (``SimulatorService._prepare_operations``,
tvb_framework/tvb/core/services/simulator_service.py:193-244): for every value
of the first range parameter (outer loop) and of the optional second (inner
loop) it deep-copies the simulator, sets the value on the attribute its dotted
name designates (:88-93) and records ``ranges`` = {name: value}.  ``PSE`` builds
the same grid, in the same order, and runs every point in one device plan:
points that share a connectome and delays share it (batch-lane kernel),
points whose connectivity or conduction speed differ get their own
(edge-lane kernel, uneven node counts included).

``RangeParameter`` and ``Range`` mirror
tvb_framework/tvb/core/entities/transient/range_parameter.py:33-80 and
tvb_library/tvb/basic/neotraits/_attr.py:619-634 (values numpy.arange(lo,
hi, step); float parameters of array attributes wrap each value as a
one-element array).
"""
import copy
import json

import numpy as np

from .backend import HipBackend


class Range:
    """neotraits Range: values numpy.arange(lo, hi, step)."""

    def __init__(self, lo, hi, step):
        self.lo, self.hi, self.step = lo, hi, step

    def to_array(self):
        return np.arange(self.lo, self.hi, self.step)


class LinspaceRange:
    """neotraits LinspaceRange: numpy.linspace(lo, hi, npoints)."""

    def __init__(self, lo, hi, npoints):
        self.lo, self.hi, self.npoints = lo, hi, npoints

    def to_array(self):
        return np.linspace(self.lo, self.hi, self.npoints)


class RangeParameter:
    """range_parameter.py:33-80: ``name`` is a dotted attribute path on the
    simulator ("coupling.a", "model.eta", "integrator.noise.nsig",
    "conduction_speed", "connectivity"); float parameters take their values
    from ``range_definition`` (a Range), others from ``range_values``."""

    def __init__(self, name, type=float, range_definition=None, is_array=False, range_values=None):
        self.name = name
        self.type = type
        self.is_array = is_array
        self.range_definition = range_definition
        self.range_values = range_values

    def get_range_values(self):
        if self.type is float and self.range_definition is not None:
            self.range_values = self.range_definition.to_array()
            if self.is_array:
                self.range_values = [np.array([v]) for v in self.range_values]
        return self.range_values

    def to_json(self):
        if self.type is float:
            return json.dumps([self.name, {"lo": self.range_definition.lo, "step": self.range_definition.step,
                                           "hi": self.range_definition.hi}])
        return json.dumps([self.name, self.range_definition])


def set_range_parameter(simulator, name, value):
    """SimulatorService._set_simulator_range_parameter (:88-93)."""
    parts = name.split(".")
    obj = simulator
    for p in parts[:-1]:
        obj = getattr(obj, p)
    setattr(obj, parts[-1], value)


def _range_value_in_dict(value):
    """SimulatorService._set_range_param_in_dict (:136-143)."""
    if isinstance(value, np.ndarray):
        return value[0]
    return value


class PSE:
    """The grid of ``range_param1`` x ``range_param2`` over ``simulator``
    (an unconfigured or configured simulator: each point is a deep copy,
    reset to step 0 and configured after its values are set)."""

    def __init__(self, simulator, range_param1, range_param2=None):
        self.simulator = simulator
        self.range_param1 = range_param1
        self.range_param2 = range_param2
        v2 = [None] if range_param2 is None else list(range_param2.get_range_values())
        self.points = []          # (ranges dict, simulator), operation order
        for a in range_param1.get_range_values():
            for b in v2:
                sim = copy.deepcopy(simulator)
                ranges = {range_param1.name: _range_value_in_dict(a)}
                self._set(sim, range_param1.name, a)
                if b is not None:
                    self._set(sim, range_param2.name, b)
                    ranges[range_param2.name] = _range_value_in_dict(b)
                # a fresh operation: step 0 and the noise stream at its seed
                # (the framework configures each point's own copy)
                sim.current_step = 0
                noise = getattr(sim.integrator, "noise", None)
                if noise is not None and hasattr(noise, "reset_random_stream"):
                    noise.reset_random_stream()
                sim.configure()
                self.points.append((ranges, sim))

    @staticmethod
    def _set(sim, name, value):
        if name == "conduction_speed":
            # the framework's conduction speed is the connectivity's speed
            set_range_parameter(sim, name, value)
            sim.connectivity.speed = np.atleast_1d(np.asarray(value, np.float64))
            return
        parts = name.split(".")
        obj = sim
        for p in parts[:-1]:
            obj = getattr(obj, p)
        old = getattr(obj, parts[-1], None)
        if isinstance(old, np.ndarray) and np.isscalar(value):
            value = np.array([value], dtype=old.dtype)     # array attributes take one-element arrays
        setattr(obj, parts[-1], value)

    @property
    def ranges(self):
        """The ``ranges`` JSON of each operation (simulator_service.py:220-227)."""
        return [json.dumps({k: (float(v) if isinstance(v, (np.floating, float)) else v)
                            for k, v in r.items()}) for r, _ in self.points]

    def run(self, simulation_length=None, nstep=None, precision="auto", noise="inject", kernel="auto", seed=0,
            chunk=65536, backend=None):
        """Every grid point in one batched plan; returns [(ranges, [(t, x), ...])]
        in operation order.  noise "inject" draws each point's normals from its
        own simulator's RandomState (as its own operation would), "philox"
        uses the device generator (stream = point index)."""
        be = backend or HipBackend()
        sims = [s for _, s in self.points]
        if nstep is None:
            length = simulation_length if simulation_length is not None else sims[0].simulation_length
            nstep = int(np.ceil(length / sims[0].integrator.dt))
        for s in sims:
            if getattr(s, "stimulus", None) is not None:
                raise NotImplementedError("stimulated simulators run one by one (HipBackend.run_sim)")
        out = be.run_batch(sims, nstep, precision=precision, noise=noise, kernel=kernel, seed=seed, chunk=chunk)
        self.last_plan_shape = be.last_plan_shape
        self.last_flags = be.last_flags
        return [(r, res) for (r, _), res in zip(self.points, out)]
