"""Pythonic wrapper of a device plan (``tvbk_plan``).

A :class:`Plan` holds B simulation instances of one model / coupling /
integrator on one GPU: their connectome(s) in CSR form, per-instance
parameters, history rings, states and monitor accumulators.  It is the
object :class:`tvb_hip.backend.HipBackend` builds from a configured
Simulator.
"""
import ctypes as C

import numpy as np

from . import _lib as L


class CSR:
    """One connectome as CSR rows (targets) of nonzero weights, in the
    row-major ``argwhere`` order of ``SparseHistory`` (history.py:221-227)."""

    def __init__(self, weights, idelays, horizon=None):
        weights = np.asarray(weights, np.float64)
        idelays = np.asarray(idelays)
        n = weights.shape[0]
        mask = weights != 0.0
        rows, cols = np.nonzero(mask)                    # row-major order
        self.n_node = n
        # horizon spans the full matrix (connectivity.py:323)
        self.horizon = int(idelays.max()) + 1 if horizon is None else int(horizon)
        self.row_ptr = np.zeros(n + 1, np.int32)
        np.cumsum(np.bincount(rows, minlength=n), out=self.row_ptr[1:])
        self.col = np.ascontiguousarray(cols, np.int32)
        self.idelay = np.ascontiguousarray(idelays[mask], np.int32)
        self.weight = np.ascontiguousarray(weights[mask], np.float64)
        self.nnz = int(mask.sum())

    def struct(self):
        return L.Connectome(self.n_node, self.horizon, self.nnz, L.ptr(self.row_ptr),
                            L.ptr(self.col), L.ptr(self.idelay), L.ptr(self.weight))


class MonitorSpec:
    """Device monitor description (``tvbk_monitor``)."""

    def __init__(self, kind, istep, voi, interim_istep=0, taps=None, k1_v0=0.0,
                 subtract_one=True, period=None, name=None, keep_stock=False, spatial_mask=None):
        self.kind = kind
        self.istep = int(istep)
        self.voi = [int(v) for v in voi]
        self.interim_istep = int(interim_istep)
        self.taps = None if taps is None else np.ascontiguousarray(taps, np.float64)
        self.k1_v0 = float(k1_v0)
        self.subtract_one = bool(subtract_one)
        self.period = period
        self.name = name
        self.keep_stock = bool(keep_stock)
        # SpatialAverage: area of every node; GlobalAverage: one area
        self.spatial_mask = None if spatial_mask is None else np.ascontiguousarray(spatial_mask, np.int32)
        if kind == L.MON_SAVG:
            self.n_area = int(self.spatial_mask.max()) + 1
        elif kind == L.MON_GAVG:
            self.n_area = 1
        else:
            self.n_area = 0

    @property
    def averaging(self):
        """TemporalAverage-like: a window of istep samples (and its stock)."""
        return self.kind in (L.MON_TAVG, L.MON_AFF_TAVG)

    @property
    def stateful(self):
        """Keeps state between steps (a stock in the reference)."""
        return self.averaging or self.kind == L.MON_BOLD

    def times(self, first_step, count, dt):
        steps = first_step + self.istep * np.arange(count)
        if self.averaging:
            return (steps - self.istep / 2.0) * dt      # monitors.py:389-391
        return steps * dt                                # monitors.py:188-190, :953


class Plan:
    def __init__(self, *, model, cfun, scheme, precision, noise, conns, params, dt,
                 obs_coef, monitors, max_steps, inst_conn=None, node_params=None,
                 node_param_mask=0, seed=0, inst_stream=None, kernel=L.KERNEL_AUTO,
                 tile=0, waves=0, stream=None, flags=0, tune=False, _dry=False):
        """tune: pick the batch-lane kernel's (tile, waves) with the
        autotuner (trial launches, cached per workload signature) instead of
        the built-in heuristic; the choice and its table are in
        ``self.tuned``."""
        self._lib = L.load()
        self.tuned = None
        shared = inst_conn is None or len(set(np.asarray(inst_conn).tolist())) == 1
        if tune and not _dry and not tile and not waves and shared and len(conns) == 1 and \
                kernel in (L.KERNEL_AUTO, L.KERNEL_BATCH_LANE) and np.asarray(params).shape[0] >= 32:
            from . import autotune
            kw = dict(model=model, cfun=cfun, scheme=scheme, precision=precision, noise=noise, conns=conns,
                      params=params, dt=dt, obs_coef=obs_coef, monitors=monitors, inst_conn=inst_conn,
                      node_params=node_params, node_param_mask=node_param_mask, seed=seed,
                      inst_stream=inst_stream, stream=stream, flags=flags)
            tile, waves, table = autotune.tune(kw, device=autotune.device_name())
            kernel = L.KERNEL_BATCH_LANE
            self.tuned = {"tile": tile, "waves": waves,
                          "table_ms_per_step": {"%dx%d" % k: v for k, v in sorted(table.items())}}
        self.handle = None
        self.model, self.cfun, self.scheme = model, cfun, scheme
        self.precision, self.noise = precision, noise
        self.dt = float(dt)
        self.conns = list(conns)
        self.params = np.ascontiguousarray(params, np.float64)
        self.n_inst = self.params.shape[0]
        self.inst_conn = None if inst_conn is None else np.ascontiguousarray(inst_conn, np.int32)
        self.obs_coef = np.ascontiguousarray(obs_coef, np.float64)
        self.monitors = list(monitors)
        self.node_params = None if node_params is None else np.ascontiguousarray(node_params, np.float64)
        self.inst_stream = None if inst_stream is None else np.ascontiguousarray(inst_stream, np.uint64)
        self.max_steps = int(max_steps)
        nmp, ncp, nsv, nch = L.model_info(model, cfun)
        self.nsv, self.nch = nsv, nch
        self._conn_structs = (L.Connectome * len(self.conns))(*[c.struct() for c in self.conns])
        d = L.PlanDesc()
        d.abi_version = L.ABI_VERSION
        d.model, d.cfun, d.scheme, d.precision, d.noise, d.kernel = model, cfun, scheme, precision, noise, kernel
        d.n_inst = self.n_inst
        d.n_conn = len(self.conns)
        d.conns = C.cast(self._conn_structs, C.POINTER(L.Connectome))
        d.inst_conn = L.ptr(self.inst_conn)
        d.n_param = self.params.shape[1]
        d.params = L.ptr(self.params)
        d.node_params = L.ptr(self.node_params)
        d.node_param_mask = int(node_param_mask)
        d.dt = self.dt
        d.n_obs = self.obs_coef.shape[0]
        d.obs_coef = L.ptr(self.obs_coef)
        d.n_mon = len(self.monitors)
        for m, ms in enumerate(self.monitors):
            mo = d.mon[m]
            mo.kind, mo.istep, mo.n_voi = ms.kind, ms.istep, len(ms.voi)
            for v, o in enumerate(ms.voi):
                mo.voi[v] = o
            if ms.kind == L.MON_BOLD:
                mo.interim_istep = ms.interim_istep
                mo.n_taps = len(ms.taps)
                mo.taps = L.ptr(ms.taps)
                mo.k1_v0 = ms.k1_v0
                mo.subtract_one = int(ms.subtract_one)
                mo.keep_stock = int(ms.keep_stock)
            if ms.kind == L.MON_SAVG:
                mo.spatial_mask = L.ptr(ms.spatial_mask)
                mo.n_area = ms.n_area
        d.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
        d.inst_stream = L.ptr(self.inst_stream)
        d.max_steps = self.max_steps
        d.tile, d.waves = tile, waves
        d.flags = int(flags)
        self._desc = d
        self.esize = 8 if precision in (L.F64, L.MIXED) else 4   # state and Raw/TAvg outputs
        self.hsize = 8 if precision == L.F64 else 4                # history ring
        self._z_keep = None
        if _dry:
            nb = C.c_int64()
            L.check(self._lib.tvbk_plan_footprint(C.byref(d), C.byref(nb)))
            self.footprint_bytes = nb.value
            return
        h = C.c_void_p()
        L.check(self._lib.tvbk_plan_create(C.byref(d), stream, C.byref(h)))
        self.handle = h
        k, t, w, lds = (C.c_int32() for _ in range(4))
        L.check(self._lib.tvbk_plan_shape(h, C.byref(k), C.byref(t), C.byref(w), C.byref(lds)), h)
        self.kernel, self.tile, self.waves, self.lds_bytes = k.value, t.value, w.value, lds.value
        nc = C.c_int32()
        L.check(self._lib.tvbk_plan_channels(h, C.byref(nc)), h)
        self.nch = nc.value          # history channels kept (fewer with PLAN_SKIP_ZERO_COUPLING)

    @classmethod
    def footprint(cls, **kw):
        """Device bytes the plan would allocate (no allocation)."""
        return cls(_dry=True, **kw).footprint_bytes

    def done(self):
        """Non-blocking: True once the last run has finished on the device."""
        flag = C.c_int32()
        self._chk(self._lib.tvbk_poll(self.handle, C.byref(flag)))
        return bool(flag.value)

    # ------------------------------------------------------------------
    def _chk(self, rc):
        return L.check(rc, self.handle)

    def close(self):
        if self.handle is not None:
            self._lib.tvbk_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_history(self, inst, history, state, current_step, noalloc=False):
        """history (H, nch, N) channel values; state (nsv, N) float64.
        ``noalloc``: stage through the library's static device scratch
        (tvbk_set_history_noalloc, no device allocation)."""
        history = np.ascontiguousarray(history)
        if history.dtype not in (np.float32, np.float64):
            history = history.astype(np.float64)
        state = np.ascontiguousarray(state, np.float64)
        fn = self._lib.tvbk_set_history_noalloc if noalloc else self._lib.tvbk_set_history
        self._chk(fn(self.handle, int(inst), L.ptr(history), history.dtype.itemsize, L.ptr(state),
                     int(current_step)))

    def init_uniform(self, seed, lo, hi):
        lo = np.ascontiguousarray(lo, np.float64)
        hi = np.ascontiguousarray(hi, np.float64)
        self._chk(self._lib.tvbk_init_uniform(self.handle, int(seed) & 0xFFFFFFFFFFFFFFFF, L.ptr(lo), L.ptr(hi)))

    def set_noise(self, z, n_steps=None):
        """z: float64 standard normals (n_steps, nsv, max_N, B) as a numpy
        array, or a device pointer (int) with n_steps given."""
        if isinstance(z, int):
            self._chk(self._lib.tvbk_set_noise(self.handle, C.c_void_p(z), int(n_steps)))
            return
        z = np.ascontiguousarray(z, np.float64)
        self._z_keep = z
        self._chk(self._lib.tvbk_set_noise(self.handle, L.ptr(z), z.shape[0]))

    def set_stimulus(self, spatial, temporal, svar_mask):
        """Region stimulus of the next run: spatial (N,) x temporal[k] at the
        run's k-th step for the state variables in svar_mask (bits);
        temporal None removes it."""
        if temporal is None:
            self._chk(self._lib.tvbk_set_stimulus(self.handle, None, None, 0, 0))
            return
        sp = np.ascontiguousarray(np.asarray(spatial, np.float64).reshape(-1))
        tp = np.ascontiguousarray(np.asarray(temporal, np.float64).reshape(-1))
        self._chk(self._lib.tvbk_set_stimulus(self.handle, L.ptr(sp), L.ptr(tp), tp.size, int(svar_mask)))

    def set_steps(self, remaining):
        if remaining is None:
            self._chk(self._lib.tvbk_set_steps(self.handle, None))
        else:
            r = np.ascontiguousarray(remaining, np.int64)
            self._chk(self._lib.tvbk_set_steps(self.handle, L.ptr(r)))

    def run(self, n_steps):
        self._chk(self._lib.tvbk_run_steps(self.handle, int(n_steps)))

    def sync(self):
        self._chk(self._lib.tvbk_sync(self.handle))

    def last_run_ms(self):
        ms = C.c_double()
        self._chk(self._lib.tvbk_last_run_ms(self.handle, C.byref(ms)))
        return ms.value

    def layout(self, m):
        B = self.n_inst
        lay = C.c_int32()
        tot = C.c_int64()
        off = np.zeros(B, np.int64)
        cnt = np.zeros(B, np.int64)
        first = np.zeros(B, np.int64)
        self._chk(self._lib.tvbk_monitor_layout(self.handle, m, C.byref(lay), C.byref(tot),
                                                L.ptr(off), L.ptr(cnt), L.ptr(first)))
        return lay.value, tot.value, off, cnt, first

    def device_output(self, m):
        """(device pointer, bytes) of monitor m's output buffer; valid until
        the next run or close.  Batch-lane layout [sample][voi][node][instance]."""
        ptr = C.c_void_p()
        nb = C.c_int64()
        self._chk(self._lib.tvbk_monitor_device_ptr(self.handle, m, C.byref(ptr), C.byref(nb)))
        return ptr.value, nb.value

    def read(self, m, instances=None):
        """Samples of monitor m from the last run: list over instances of
        (times, data (count, nvoi, N))."""
        ms = self.monitors[m]
        lay, tot, off, cnt, first = self.layout(m)
        dtype = np.float64 if (ms.kind == L.MON_BOLD or self.esize == 8) else np.float32
        buf = np.empty(tot // np.dtype(dtype).itemsize, dtype)
        self._chk(self._lib.tvbk_read_monitor(self.handle, m, L.ptr(buf), buf.nbytes))
        nv = len(ms.voi)
        B = self.n_inst
        insts = range(B) if instances is None else instances
        out = []
        if lay == 0:
            N = ms.n_area or self.conns[0].n_node
            data = buf.reshape((int(cnt[0]), nv, N, B))
            for b in insts:
                out.append((ms.times(first[b], cnt[b], self.dt), data[..., b]))
        else:
            for b in insts:
                N = ms.n_area or self.conns[0 if self.inst_conn is None else self.inst_conn[b]].n_node
                n = int(cnt[b]) * nv * N
                out.append((ms.times(first[b], cnt[b], self.dt),
                            buf[off[b]:off[b] + n].reshape((int(cnt[b]), nv, N))))
        return out

    def read_all(self, m):
        """Raw buffer and layout of monitor m (no per-instance split)."""
        ms = self.monitors[m]
        lay, tot, off, cnt, first = self.layout(m)
        dtype = np.float64 if (ms.kind == L.MON_BOLD or self.esize == 8) else np.float32
        buf = np.empty(tot // np.dtype(dtype).itemsize, dtype)
        self._chk(self._lib.tvbk_read_monitor(self.handle, m, L.ptr(buf), buf.nbytes))
        return buf, lay, off, cnt, first

    def export(self, inst, dtype=np.float64):
        c = self.conns[0 if self.inst_conn is None else self.inst_conn[inst]]
        hist = np.empty((c.horizon, self.nch, c.n_node), dtype)
        state = np.empty((self.nsv, c.n_node), np.float64)
        step = C.c_int64()
        self._chk(self._lib.tvbk_export_state(self.handle, inst, L.ptr(hist), hist.dtype.itemsize,
                                              L.ptr(state), C.byref(step)))
        return hist, state, step.value

    def _n_node(self, inst):
        return self.conns[0 if self.inst_conn is None or inst < 0 else self.inst_conn[inst]].n_node

    def export_monitor(self, inst, m, stock=True, interim=True):
        """Monitor m's state of instance ``inst`` as the reference's stock
        arrays (float64, trailing mode axis): TemporalAverage ``_stock``
        (istep, nvoi, N, 1) -> (stock, None); Bold ``_stock`` (S, nvoi, N, 1)
        (monitor built with keep_stock) and ``_interim_stock``
        (interim_istep, nvoi, N, 1); Raw -> (None, None)."""
        ms = self.monitors[m]
        if not ms.stateful:
            return None, None
        N, nv = self._n_node(inst), len(ms.voi)
        if ms.averaging:
            st, it = np.empty((ms.istep, nv, N)), None
        else:
            st = np.empty((len(ms.taps), nv, N)) if stock else None
            it = np.empty((ms.interim_istep, nv, N)) if interim else None
        self._chk(self._lib.tvbk_export_monitor(self.handle, int(inst), int(m), L.ptr(st), L.ptr(it)))
        return (None if st is None else st[..., np.newaxis]), (None if it is None else it[..., np.newaxis])

    def import_monitor(self, inst, m, stock, interim=None):
        """Inverse of export_monitor at the instance's current step (after
        set_history); ``interim`` None = zeros, the state a reference Bold
        monitor has after SimulationHistory.fill_into.  inst -1 = every
        instance."""
        ms = self.monitors[m]
        if not ms.stateful:
            return
        N, nv = self._n_node(inst), len(ms.voi)
        rows = ms.istep if ms.averaging else len(ms.taps)
        st = np.ascontiguousarray(np.asarray(stock, np.float64).reshape((rows, nv, N)))
        it = None
        if interim is not None and ms.kind == L.MON_BOLD:
            it = np.ascontiguousarray(np.asarray(interim, np.float64).reshape((ms.interim_istep, nv, N)))
        self._chk(self._lib.tvbk_import_monitor(self.handle, int(inst), int(m), L.ptr(st), L.ptr(it)))

    def tiles(self):
        """(workgroups per launch, block length of split instances or 0)."""
        n, k = C.c_int32(), C.c_int32()
        self._chk(self._lib.tvbk_plan_tiles(self.handle, C.byref(n), C.byref(k)))
        return n.value, k.value

    def graphs(self):
        """(hipGraphs captured, graph launches) of this plan's split calls."""
        c, n = C.c_int64(), C.c_int64()
        self._chk(self._lib.tvbk_plan_graphs(self.handle, C.byref(c), C.byref(n)))
        return c.value, n.value

    def time_block(self):
        """Steps per block of the time-blocked sequential loop (0: off)."""
        k = C.c_int32()
        self._chk(self._lib.tvbk_plan_time_block(self.handle, C.byref(k)))
        return k.value

    def loop(self):
        """Step loop of the last run (_lib.LOOP_NAMES), or None before it."""
        k = C.c_int32()
        self._chk(self._lib.tvbk_plan_loop(self.handle, C.byref(k)))
        return L.LOOP_NAMES.get(k.value)

    def flags(self):
        f = np.zeros(self.n_inst, np.int32)
        self._chk(self._lib.tvbk_read_flags(self.handle, L.ptr(f)))
        return f

    def debug_normals(self, inst, step0, n_steps):
        """Standard normals (n_steps, nsv, N) float64 the kernels of a Philox
        plan draw for instance ``inst`` at steps step0 .. step0+n_steps-1."""
        N = self._n_node(inst)
        z = np.empty((int(n_steps), self.nsv, N), np.float64)
        self._chk(self._lib.tvbk_debug_normals(self.handle, int(inst), int(step0), int(n_steps), L.ptr(z)))
        return z

    def debug_slots(self, conn, step):
        out = np.zeros(max(self.conns[conn].nnz, 1), np.int32)
        self._chk(self._lib.tvbk_debug_slots(self.handle, conn, int(step), L.ptr(out)))
        return out[:self.conns[conn].nnz]
