"""ctypes binding of ``libtvbk.so`` (the C-ABI declared in include/tvbk.h).

The library is loaded from this package directory.  There is no fallback: if
the HIP library is missing or was built for another ABI version, importing the
backend raises :class:`LibraryMissing`.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtvbk.so")

ABI_VERSION = 3
MAX_OBS = 8
MAX_MON = 4
NOALLOC_MAX_BYTES = 32 * 1024 * 1024   # TVBK_NOALLOC_MAX_BYTES

# enums (mirror include/tvbk.h)
OK, EINVAL, ENOMEM, EHIP, EUNSUPPORTED, ESTATE = 0, -1, -2, -3, -4, -5
MODEL_G2D, MODEL_MPR, MODEL_JR, MODEL_SUM = 0, 1, 2, 3
CFUN_LINEAR, CFUN_SIGMOID_JR, CFUN_IDENTITY = 0, 1, 2
EULER_DET, EULER_STO, HEUN_DET, HEUN_STO, IDENTITY = 0, 1, 2, 3, 4
F32, F64, MIXED = 0, 1, 2   # MIXED: float32 history, weights and sums; float64 state
NOISE_NONE, NOISE_INJECTED, NOISE_PHILOX = 0, 1, 2
MON_RAW, MON_TAVG, MON_BOLD = 0, 1, 2
MON_SUBSAMPLE, MON_GAVG, MON_SAVG, MON_AFF, MON_AFF_TAVG = 3, 4, 5, 6, 7
KERNEL_AUTO, KERNEL_BATCH_LANE, KERNEL_EDGE_LANE, KERNEL_DENSE_MFMA = 0, 1, 2, 3
# step loops (TVBK_LOOP_*)
LOOP_NAMES = {0: "sequential", 1: "pipelined", 2: "time-blocked", 3: "dense", 4: "pipelined-2"}
PLAN_SKIP_ZERO_COUPLING = 1


class LibraryMissing(RuntimeError):
    """The HIP library is not built or cannot be loaded."""


class TvbkError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("tvbk error %d: %s" % (code, msg))
        self.code = code


class Connectome(C.Structure):
    _fields_ = [("n_node", C.c_int32), ("horizon", C.c_int32), ("nnz", C.c_int64),
                ("row_ptr", C.c_void_p), ("col", C.c_void_p), ("idelay", C.c_void_p),
                ("weight", C.c_void_p)]


class Monitor(C.Structure):
    _fields_ = [("kind", C.c_int32), ("istep", C.c_int32), ("n_voi", C.c_int32),
                ("voi", C.c_int32 * MAX_OBS), ("interim_istep", C.c_int32),
                ("n_taps", C.c_int32), ("taps", C.c_void_p), ("k1_v0", C.c_double),
                ("subtract_one", C.c_int32), ("keep_stock", C.c_int32),
                ("spatial_mask", C.c_void_p), ("n_area", C.c_int32)]


class PlanDesc(C.Structure):
    _fields_ = [("abi_version", C.c_int32), ("model", C.c_int32), ("cfun", C.c_int32),
                ("scheme", C.c_int32), ("precision", C.c_int32), ("noise", C.c_int32),
                ("kernel", C.c_int32), ("n_inst", C.c_int32), ("n_conn", C.c_int32),
                ("conns", C.POINTER(Connectome)), ("inst_conn", C.c_void_p),
                ("n_param", C.c_int32), ("params", C.c_void_p), ("node_params", C.c_void_p),
                ("node_param_mask", C.c_uint64), ("dt", C.c_double), ("n_obs", C.c_int32),
                ("obs_coef", C.c_void_p), ("n_mon", C.c_int32), ("mon", Monitor * MAX_MON),
                ("seed", C.c_uint64), ("inst_stream", C.c_void_p), ("max_steps", C.c_int64),
                ("tile", C.c_int32), ("waves", C.c_int32), ("flags", C.c_uint32)]


_lib = None


def _bind(lib):
    P = C.c_void_p
    sig = {
        "tvbk_abi_version": ([], C.c_int),
        "tvbk_model_info": ([C.c_int32, C.c_int32, P, P, P, P], C.c_int),
        "tvbk_plan_create": ([C.POINTER(PlanDesc), P, C.POINTER(C.c_void_p)], C.c_int),
        "tvbk_plan_destroy": ([P], C.c_int),
        "tvbk_plan_footprint": ([C.POINTER(PlanDesc), P], C.c_int),
        "tvbk_poll": ([P, P], C.c_int),
        "tvbk_stream_create": ([P], C.c_int),
        "tvbk_stream_destroy": ([P], C.c_int),
        "tvbk_set_history": ([P, C.c_int32, P, C.c_int32, P, C.c_int64], C.c_int),
        "tvbk_set_history_noalloc": ([P, C.c_int32, P, C.c_int32, P, C.c_int64], C.c_int),
        "tvbk_init_uniform": ([P, C.c_uint64, P, P], C.c_int),
        "tvbk_set_noise": ([P, P, C.c_int64], C.c_int),
        "tvbk_run_steps": ([P, C.c_int64], C.c_int),
        "tvbk_set_steps": ([P, P], C.c_int),
        "tvbk_set_stimulus": ([P, P, P, C.c_int64, C.c_uint32], C.c_int),
        "tvbk_sync": ([P], C.c_int),
        "tvbk_monitor_layout": ([P, C.c_int32, P, P, P, P, P], C.c_int),
        "tvbk_read_monitor": ([P, C.c_int32, P, C.c_int64], C.c_int),
        "tvbk_monitor_device_ptr": ([P, C.c_int32, P, P], C.c_int),
        "tvbk_export_state": ([P, C.c_int32, P, C.c_int32, P, P], C.c_int),
        "tvbk_export_monitor": ([P, C.c_int32, C.c_int32, P, P], C.c_int),
        "tvbk_import_monitor": ([P, C.c_int32, C.c_int32, P, P], C.c_int),
        "tvbk_read_flags": ([P, P], C.c_int),
        "tvbk_debug_slots": ([P, C.c_int32, C.c_int64, P], C.c_int),
        "tvbk_debug_normals": ([P, C.c_int32, C.c_int64, C.c_int64, P], C.c_int),
        "tvbk_plan_shape": ([P, P, P, P, P], C.c_int),
        "tvbk_plan_tiles": ([P, P, P], C.c_int),
        "tvbk_plan_time_block": ([P, P], C.c_int),
        "tvbk_plan_loop": ([P, P], C.c_int),
        "tvbk_plan_channels": ([P, P], C.c_int),
        "tvbk_plan_graphs": ([P, P, P], C.c_int),
        "tvbk_last_run_ms": ([P, P], C.c_int),
        "tvbk_balloon": ([P, C.c_int32, C.c_int64, C.c_int64, C.c_int64, C.c_double, P, C.c_int32, C.c_int32,
                          P, P, P], C.c_int),
        "tvbk_probe_gather": ([C.c_int64, C.c_int32, C.c_int32, P], C.c_int),
        "tvbk_last_error": ([P], C.c_char_p),
    }
    for name, (argtypes, restype) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = argtypes
        fn.restype = restype
    return lib


EXPORTED = (
    "tvbk_abi_version", "tvbk_model_info", "tvbk_plan_create", "tvbk_plan_destroy",
    "tvbk_plan_footprint", "tvbk_poll", "tvbk_stream_create", "tvbk_stream_destroy",
    "tvbk_set_history", "tvbk_set_history_noalloc", "tvbk_init_uniform", "tvbk_set_noise", "tvbk_run_steps",
    "tvbk_set_steps", "tvbk_set_stimulus", "tvbk_sync", "tvbk_monitor_layout", "tvbk_read_monitor",
    "tvbk_monitor_device_ptr", "tvbk_export_state", "tvbk_export_monitor", "tvbk_import_monitor",
    "tvbk_read_flags",
    "tvbk_debug_slots", "tvbk_debug_normals", "tvbk_plan_shape", "tvbk_plan_tiles", "tvbk_plan_time_block", "tvbk_plan_loop",
    "tvbk_plan_channels", "tvbk_plan_graphs", "tvbk_last_run_ms",
    "tvbk_balloon",
    "tvbk_probe_gather", "tvbk_last_error")


def load():
    """Load and bind the HIP library; raises LibraryMissing if unavailable."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise LibraryMissing("HIP library %s is not built (run __graft_entry__.build())" % LIB_PATH)
    # PyTorch-ROCm ships its own HIP runtime: load it first when present so
    # that this library binds to the same runtime (one device context per
    # process, torch streams and pointers valid here).
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    try:
        lib = C.CDLL(LIB_PATH)
    except OSError as exc:
        raise LibraryMissing("cannot load %s: %s" % (LIB_PATH, exc))
    _bind(lib)
    if lib.tvbk_abi_version() != ABI_VERSION:
        raise LibraryMissing("libtvbk ABI %d != %d" % (lib.tvbk_abi_version(), ABI_VERSION))
    _lib = lib
    return lib


def check(rc, plan=None):
    if rc != OK:
        lib = load()
        msg = lib.tvbk_last_error(plan)
        raise TvbkError(rc, msg.decode() if msg else "")
    return rc


def ptr(a):
    """Address of a contiguous numpy array (or None)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


def probe_gather(piece_bytes=256, write=False, nbytes=3200 << 20):
    """Read (+ write) bytes/s of the gather's access pattern on this device
    (tvbk_probe_gather): 256-byte pieces read in segments of ``piece_bytes``
    at random offsets of an ``nbytes`` buffer, with ``write`` one segment
    stored per 16 read; ``piece_bytes`` 128: two random 128-byte pieces per
    wave load."""
    lib = load()
    r = C.c_double(0.0)
    check(lib.tvbk_probe_gather(int(nbytes), int(piece_bytes), 1 if write else 0, C.byref(r)))
    return r.value


def stream_create():
    """A HIP stream owned by the caller (release with stream_destroy)."""
    lib = load()
    h = C.c_void_p()
    check(lib.tvbk_stream_create(C.byref(h)))
    return h.value


def stream_destroy(stream):
    check(load().tvbk_stream_destroy(stream))


def model_info(model, cfun):
    lib = load()
    out = [C.c_int32() for _ in range(4)]
    check(lib.tvbk_model_info(model, cfun, *[C.addressof(o) for o in out]))
    return tuple(o.value for o in out)
