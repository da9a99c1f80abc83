"""Request loop: continuous admission of simulation requests on one GPU.

A request asks for ``n_instances`` runs of one model on one connectome
(its own seed and duration).  The loop

* admits waiting requests in arrival order as soon as their device
  footprint (``tvbk_plan_footprint``: history ring, state, monitor
  accumulators and output buffers) fits the free part of the HBM budget and
  an execution slot is free;
* keeps every admitted request's state on the device (one plan per
  request, bound to one of several HIP streams) and advances it in chunks
  of ``chunk_steps`` steps, so requests that arrive while others run start
  as soon as memory and a slot free up and their launches overlap on the
  device with the running ones;
* copies each chunk's monitor samples to the host when the chunk finishes
  (non-blocking poll), and retires a request when its last chunk is read:
  latency = result available - arrival.

The reference has no serving layer; its nearest counterpart is the
framework's one-process-per-operation launcher
(tvb_framework/tvb/core/services/backend_clients/standalone_client.py:56-130)
and the memory estimate of Simulator._guesstimate_memory_requirement
(simulator.py:474-583).
"""
import collections
import time

import numpy as np


class Request:
    """One simulation request.  ``build(stream, max_steps)`` returns a Plan
    with its history set; ``footprint(max_steps)`` its device bytes."""

    def __init__(self, rid, arrival, n_steps, build, footprint, region_steps, meta=None):
        self.rid = rid
        self.arrival = float(arrival)
        self.n_steps = int(n_steps)
        self.build = build
        self.footprint = footprint
        self.region_steps = float(region_steps)
        self.meta = meta or {}
        self.admitted = None
        self.finished = None
        self.outputs = None
        self.flags = None

    @property
    def latency(self):
        return None if self.finished is None else self.finished - self.arrival


class _Active:
    __slots__ = ("req", "plan", "stream", "left", "launched", "bytes", "chunks")

    def __init__(self, req, plan, stream, nbytes):
        self.req, self.plan, self.stream, self.bytes = req, plan, stream, nbytes
        self.left = req.n_steps
        self.launched = False
        self.chunks = []


class Server:
    """Runs a list of requests with Poisson (or any) arrival times."""

    def __init__(self, hbm_budget, streams, chunk_steps=20000, max_active=16, clock=time.perf_counter,
                 idle_sleep=2e-4, keep_outputs=True, log=None):
        self.budget = int(hbm_budget)
        self.log = log          # optional callable(str): a line per admission and retirement
        self.streams = list(streams)
        self.chunk = int(chunk_steps)
        self.max_active = int(max_active)
        self.clock = clock
        self.idle_sleep = idle_sleep
        self.keep_outputs = keep_outputs
        self.peak_bytes = 0
        self.max_concurrent = 0

    def run(self, requests):
        reqs = sorted(requests, key=lambda r: r.arrival)
        for r in reqs:
            fp = r.footprint(min(self.chunk, r.n_steps))
            if fp > self.budget:
                raise ValueError("request %s needs %d bytes, more than the budget %d" % (r.rid, fp, self.budget))
        pending = collections.deque(reqs)
        waiting = collections.deque()
        active = []
        used = 0
        stream_load = {i: 0 for i in range(len(self.streams))}
        t0 = self.clock()
        while pending or waiting or active:
            now = self.clock() - t0
            while pending and pending[0].arrival <= now:
                waiting.append(pending.popleft())
            # admission, arrival order (no overtaking: no starvation)
            while waiting and len(active) < self.max_active:
                r = waiting[0]
                fp = r.footprint(min(self.chunk, r.n_steps))
                if used + fp > self.budget:
                    break
                waiting.popleft()
                si = min(stream_load, key=stream_load.get)
                plan = r.build(self.streams[si], min(self.chunk, r.n_steps))
                r.admitted = self.clock() - t0
                if self.log:
                    self.log("request %s admitted at %.2f s (arrived %.2f s)" % (r.rid, r.admitted, r.arrival))
                a = _Active(r, plan, si, fp)
                active.append(a)
                stream_load[si] += 1
                used += fp
                self.peak_bytes = max(self.peak_bytes, used)
            self.max_concurrent = max(self.max_concurrent, len(active))
            progressed = False
            for a in list(active):
                if a.launched and not a.plan.done():
                    continue
                if a.launched:
                    if self.keep_outputs:
                        a.chunks.append([a.plan.read(m) for m in range(len(a.plan.monitors))])
                    a.launched = False
                    progressed = True
                if a.left == 0:
                    r = a.req
                    r.flags = a.plan.flags()
                    r.finished = self.clock() - t0
                    if self.log:
                        self.log("request %s done at %.2f s, latency %.2f s" % (r.rid, r.finished, r.latency))
                    if self.keep_outputs:
                        r.outputs = _concat_chunks(a.chunks, a.plan.n_inst, len(a.plan.monitors))
                    a.plan.close()
                    active.remove(a)
                    stream_load[a.stream] -= 1
                    used -= a.bytes
                    progressed = True
                    continue
                n = min(self.chunk, a.left)
                a.plan.run(n)
                a.left -= n
                a.launched = True
                progressed = True
            if not progressed:
                if pending and not waiting and not active:
                    time.sleep(max(0.0, min(pending[0].arrival - (self.clock() - t0), 0.01)))
                else:
                    time.sleep(self.idle_sleep)
        makespan = self.clock() - t0
        lat = np.array([r.latency for r in reqs])
        rs = sum(r.region_steps for r in reqs)
        return {
            "n_requests": len(reqs),
            "makespan_s": makespan,
            "p50_latency_s": float(np.percentile(lat, 50)),
            "p99_latency_s": float(np.percentile(lat, 99)),
            "mean_latency_s": float(lat.mean()),
            "region_steps": rs,
            "region_steps_per_s": rs / makespan,
            "peak_bytes": self.peak_bytes,
            "max_concurrent": self.max_concurrent,
        }


def _concat_chunks(chunks, n_inst, n_mon):
    """Per instance, per monitor (t, x) over all chunks."""
    out = []
    for b in range(n_inst):
        mons = []
        for m in range(n_mon):
            ts = [c[m][b][0] for c in chunks]
            xs = [np.asarray(c[m][b][1]) for c in chunks]
            mons.append((np.concatenate(ts), np.concatenate(xs, axis=0)))
        out.append(mons)
    return out


# ---------------------------------------------------------------------------
# workload of synthetic requests
# ---------------------------------------------------------------------------

def synthetic_requests(n_requests, rate_hz, seed=0, inst_range=(1, 512), node_range=(68, 998),
                       duration_range_ms=(5000.0, 60000.0), duration_scale=1.0, models=("G2D", "MPR"),
                       precision="mixed", skip_zero_coupling=True):
    """Poisson stream of sweep requests over seeded synthetic connectomes:
    N log-uniform in node_range, instances uniform in inst_range, model
    chosen uniformly, HeunStochastic, TemporalAverage(1 ms), per-request
    noise seed.  ``duration_scale`` shrinks the simulated durations (the
    benchmark states the factor).  ``skip_zero_coupling``: MontbrioPazoRoxin
    requests (cv = 0) keep only the r channel in their history ring
    (TVBK_PLAN_SKIP_ZERO_COUPLING)."""
    from . import _lib as L
    from .connectome import idelays, synthetic
    from .plan import CSR, MonitorSpec, Plan
    rng = np.random.default_rng(seed)
    t = 0.0
    reqs = []
    for k in range(n_requests):
        t += rng.exponential(1.0 / rate_hz)
        n = int(round(np.exp(rng.uniform(np.log(node_range[0]), np.log(node_range[1])))))
        B = int(rng.integers(inst_range[0], inst_range[1] + 1))
        model = models[int(rng.integers(len(models)))]
        dur = rng.uniform(*duration_range_ms) * duration_scale
        cseed = int(rng.integers(1 << 30))
        nseed = int(rng.integers(1 << 62))
        reqs.append(_make_request(k, t, n, B, model, dur, cseed, nseed, precision, skip_zero_coupling))
    return reqs


def _make_request(rid, arrival, n, B, model, dur_ms, cseed, nseed, precision, skip_zero_coupling=False):
    from . import _lib as L
    from .connectome import idelays, synthetic
    from .plan import CSR, MonitorSpec, Plan
    if model == "MPR":
        dt, speed, mid, sv = 0.05, 5.0, L.MODEL_MPR, ([0.0, -2.0], [2.0, 1.5])
        raw = [1.0, 0.0, 1.0, 15.0, -5.0, 0.0, 1.0, 0.0]
    else:
        dt, speed, mid, sv = 0.1, 3.0, L.MODEL_G2D, ([-2.0, -6.0], [4.0, 6.0])
        raw = [1.0, 0.0, -2.0, -10.0, 0.0, 0.02, 3.0, 1.0, 0.0, 1.0, 1.0, 1.0]
    density = 0.2
    w, lengths = synthetic(n, density, 10.0, 180.0, cseed)
    d, H = idelays(lengths, speed, dt)
    rng = np.random.default_rng(cseed)
    G = rng.uniform(0.0, 0.5, B)
    sig = rng.uniform(1e-4, 1e-3, B)
    params = np.empty((B, len(raw) + 2 + 2))
    params[:, :len(raw)] = raw
    params[:, len(raw)] = G
    params[:, len(raw) + 1] = 0.0
    params[:, len(raw) + 2:] = sig[:, None]
    n_steps = int(np.ceil(dur_ms / dt))
    prec = {"fp64": L.F64, "fp32": L.F32, "mixed": L.MIXED}[precision]
    csr = CSR(w, d, H)
    mons = [MonitorSpec(L.MON_TAVG, int(round(1.0 / dt)), [0], name="TemporalAverage")]
    obs = np.eye(2)[:1] if model == "G2D" else np.eye(2)[:1]
    kw = dict(model=mid, cfun=L.CFUN_LINEAR, scheme=L.HEUN_STO, precision=prec, noise=L.NOISE_PHILOX,
              conns=[csr], params=params, dt=dt, obs_coef=obs, monitors=mons, seed=nseed,
              flags=L.PLAN_SKIP_ZERO_COUPLING if skip_zero_coupling and model == "MPR" else 0)

    def footprint(max_steps):
        return Plan.footprint(max_steps=max_steps, **kw)

    def build(stream, max_steps):
        p = Plan(max_steps=max_steps, stream=stream, **kw)
        p.init_uniform(nseed ^ 0x5DEECE66D, sv[0], sv[1])
        return p

    return Request(rid, arrival, n_steps, build, footprint, region_steps=float(n) * B * n_steps,
                   meta=dict(n=n, B=B, model=model, duration_ms=dur_ms, dt=dt, speed=speed, cseed=cseed))
