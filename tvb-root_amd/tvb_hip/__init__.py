"""tvb_hip: MI355X (gfx950) backend for TVB's simulator time-stepping loop.

``HipBackend`` is the backend plugin (check_compatibility / run_sim /
run_batch); ``Plan`` is the device plan over the C-ABI of ``libtvbk.so``;
``tvb_hip.sim`` mirrors the reference Simulator's configuration surface.
"""
from . import _lib
from ._lib import LibraryMissing, TvbkError
from .plan import CSR, MonitorSpec, Plan


def __getattr__(name):
    if name == "HipBackend":
        from .backend import HipBackend
        return HipBackend
    raise AttributeError(name)
