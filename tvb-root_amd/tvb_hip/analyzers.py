"""Balloon-Windkessel BOLD on the device.

``BalloonModel`` mirrors tvb_library/tvb/analyzers/fmri_balloon.py:73-364
(same attribute names and defaults) and evaluates on the HIP library's
``tvbk_balloon`` kernel; ``BalloonStream`` applies the same model to a
plan's TemporalAverage output chunk by chunk while it stays on the device
(the model state carried across chunks).  There is no CPU fallback.
"""
import ctypes as C

import numpy as np

from . import _lib as L

PARAM_ORDER = ("tau_s", "tau_f", "tau_o", "alpha", "TE", "V0", "E0", "epsilon", "nu_0", "r_0")


class BalloonTimeSeries:
    """Result of BalloonModel.evaluate (the fields of TimeSeriesRegion the
    reference fills, fmri_balloon.py:270-275): data (T, 1, N, M), time in s."""

    def __init__(self, data, time, sample_period):
        self.data = data
        self.time = time
        self.sample_period = sample_period
        self.sample_period_unit = "s"


class BalloonModel:
    """fmri_balloon.py:73-196 parameters; ``time_series`` needs ``data``
    (T, nvar, N, M), ``time`` and ``sample_period`` (+ ``sample_period_unit``,
    default "ms")."""

    _defaults = dict(tau_s=1.54, tau_f=1.44, tau_o=0.98, alpha=0.32, TE=0.04, V0=4.0, E0=0.4,
                     epsilon=np.array([0.5]), nu_0=40.3, r_0=25.0, bold_model="nonlinear", RBM=True,
                     normalize_neural_input=False, neural_input_transformation="none")

    def __init__(self, time_series=None, **kw):
        self.time_series = time_series
        for k, v in self._defaults.items():
            setattr(self, k, kw.pop(k, v))
        if kw:
            raise TypeError("unknown BalloonModel attributes: %s" % sorted(kw))

    def params(self):
        return np.array([float(np.asarray(getattr(self, k)).reshape(-1)[0]) for k in PARAM_ORDER], np.float64)

    def flags(self):
        bm = getattr(self.bold_model, "value", self.bold_model)
        return (1 if bm == "linear" else 0) | (0 if self.RBM else 2) | (4 if self.normalize_neural_input else 0)

    def evaluate(self):
        """BOLD of the first state variable of the time series on the device."""
        import torch
        mode = getattr(self.neural_input_transformation, "value", self.neural_input_transformation)
        if mode != "none":
            raise NotImplementedError("neural input transformation %r is not offered on the device" % (mode,))
        ts = self.time_series
        unit = getattr(ts, "sample_period_unit", "ms")
        sp = float(ts.sample_period)
        # TimeSeries.sample_rate in Hz (datatypes/time_series.py:84-93)
        sample_rate = {"s": 1.0, "sec": 1.0, "ms": 1000.0, "msec": 1000.0, "us": 1e6, "usec": 1e6}[unit] / sp
        sp_ms = 1000.0 / sample_rate                                # the kernel's dt = 1 / sample_rate
        if unit in ("ms", "msec"):
            sp_ms = sp
        data = np.asarray(ts.data)
        x = np.ascontiguousarray(data[:, 0], np.float64)         # (T, N, M), fmri_balloon.py:314-318
        T = x.shape[0]
        n = int(np.prod(x.shape[1:]))
        dev = torch.device("cuda")
        xd = torch.from_numpy(x.reshape(T, n)).to(dev)
        yd = torch.empty((T, n), dtype=torch.float64, device=dev)
        sd = torch.empty((4, n), dtype=torch.float64, device=dev)
        run_balloon(xd.data_ptr(), 8, n, T, n, sp_ms, self.params(), self.flags(), True, sd.data_ptr(),
                    yd.data_ptr(), torch.cuda.current_stream().cuda_stream)
        y = yd.cpu().numpy().reshape((T, 1) + x.shape[1:])
        t_int = np.asarray(ts.time) / (sample_rate * sp)             # seconds, fmri_balloon.py:312-317
        return BalloonTimeSeries(y, t_int, 1.0 / sample_rate)


def run_balloon(x_ptr, x_elem_bytes, x_stride, n_samples, n_series, sample_period_ms, params, flags, start,
                state_ptr, y_ptr, stream=None):
    """Queue the device kernel (pointers are device addresses)."""
    lib = L.load()
    par = np.ascontiguousarray(params, np.float64)
    L.check(lib.tvbk_balloon(C.c_void_p(x_ptr), int(x_elem_bytes), int(x_stride), int(n_samples),
                             int(n_series), float(sample_period_ms), L.ptr(par), int(flags), int(bool(start)),
                             C.c_void_p(state_ptr), C.c_void_p(y_ptr), C.c_void_p(stream)))


class BalloonStream:
    """Balloon BOLD of a plan's TemporalAverage (or Raw) output, fed after each
    ``plan.run`` while the output is still on the device: the first variable
    of interest of every (node, instance) series drives the model, whose state
    persists across calls.  ``feed()`` returns the chunk's BOLD as a device
    tensor (samples, nodes * instances) in the batch-lane series order."""

    def __init__(self, plan, monitor, model=None):
        import torch
        self.plan, self.m = plan, monitor
        self.model = model or BalloonModel()
        ms = plan.monitors[monitor]
        if ms.kind not in (L.MON_TAVG, L.MON_RAW):
            raise ValueError("the Balloon model takes a TemporalAverage or Raw output")
        self.sp_ms = ms.period if getattr(ms, "period", None) else ms.istep * plan.dt   # the monitor period
        self.nvoi = len(ms.voi)
        self.N = plan.conns[0].n_node
        self.n = self.N * plan.n_inst
        self.elem = plan.esize
        self.state = torch.empty((4, self.n), dtype=torch.float64, device="cuda")
        self.started = False

    def feed(self, stream=None):
        import torch
        lay, tot, off, cnt, first = self.plan.layout(self.m)
        if lay != 0:
            raise NotImplementedError("BalloonStream needs the batch-lane output layout")
        T = int(cnt[0])
        y = torch.empty((T, self.n), dtype=torch.float64, device="cuda")
        if T == 0:
            return y
        ptr, _ = self.plan.device_output(self.m)
        self.plan.sync()
        run_balloon(ptr, self.elem, self.nvoi * self.n, T, self.n, self.sp_ms, self.model.params(),
                    self.model.flags() & 3, not self.started, self.state.data_ptr(), y.data_ptr(),
                    stream if stream is not None else torch.cuda.current_stream().cuda_stream)
        self.started = True
        return y
