#!/usr/bin/env python
"""Benchmark of the TVB simulator loop on one MI355X.

Workload (default, "c2"): batched parameter sweep on one shared 68-node
synthetic connectome (25 % density, log-normal weights, tract lengths
U[10, 180] mm, speed 5 mm/ms -> delays 40..720 steps at dt 0.05 ms),
8192 instances on a 128 x 64 grid of coupling gain G in [0, 2] (Linear a)
x noise sigma in [1e-4, 1e-2] (Additive nsig of both state variables),
MontbrioPazoRoxin, HeunStochastic dt = 0.05 ms, TemporalAverage(1 ms) +
Bold(2000 ms), FP32 state / history / coupling, FP64 Bold, device Philox
noise.  One "step" advances all 8192 x 68 region states by dt.

Prints one JSON line: region-steps/s of the batched FP32 sweep, the
step-kernel roofline (algorithmic HBM bytes per launch over the launch
time from HIP events), the numpy oracle timed on one host core on a sample
of the same workload, the plain single-instance FP64 device path on the
same connectome, and the sweep's TemporalAverage error against the FP64
oracle on a sample.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tvb-root_amd"))

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
FP32_MFMA_PEAK_TFLOPS = 157.3
TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "k_step_mpr_c2_hbm_traffic.md")


def measured_traffic_per_region_step(path=TRAFFIC_SUMMARY):
    """HBM bytes per region-step of the step kernel from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (line
    'traffic_bytes_per_region_step: X'), or None when there is none."""
    try:
        with open(path) as f:
            for ln in f:
                if ln.strip().startswith("traffic_bytes_per_region_step:"):
                    return float(ln.split(":", 1)[1].split()[0])
    except (OSError, ValueError):
        pass
    return None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def c2_setup(B=8192, seed=2024):
    from tvb_hip.connectome import idelays, sweep_grid, synthetic
    w, L = synthetic(68, 0.25, 10.0, 180.0, seed=seed)
    d, H = idelays(L, 5.0, 0.05)
    n_g, n_s = 128, max(1, B // 128)
    G, sig = sweep_grid(n_g, n_s)
    G, sig = G[:B], sig[:B]
    return w, L, d, H, G, sig


def mpr_params(G, sig):
    # raw MontbrioPazoRoxin defaults (tau I Delta J eta Gamma cr cv),
    # Linear (a=G, b=0), nsig for r and V
    B = len(G)
    p = np.empty((B, 8 + 2 + 2))
    p[:, :8] = [1.0, 0.0, 1.0, 15.0, -5.0, 0.0, 1.0, 0.0]
    p[:, 8] = G
    p[:, 9] = 0.0
    p[:, 10] = sig
    p[:, 11] = sig
    return p


def make_bold_spec(dt, period=2000.0):
    from tvb_hip import _lib as L
    from tvb_hip.plan import MonitorSpec
    from tvb_hip.sim import Bold, iround

    b = Bold(period=period)
    b.dt = dt
    b.istep = iround(period / dt)
    b.compute_hrf()
    taps = b.hemodynamic_response_function[0, ::-1].copy()
    k1v0 = b.hrf_kernel.parameters["k_1"] * b.hrf_kernel.parameters["V_0"]
    return MonitorSpec(L.MON_BOLD, b.istep, [0, 1], interim_istep=b._interim_istep, taps=taps,
                       k1_v0=k1v0, subtract_one=True, period=period, name="Bold")


def build_c2_plan(w, d, G, sig, dt, max_steps, precision, noise, seed, tile=0, waves=0, tavg=1.0, bold=2000.0,
                  stream=None):
    from tvb_hip import _lib as L
    from tvb_hip.plan import CSR, MonitorSpec, Plan
    from tvb_hip.sim import iround
    mons = [MonitorSpec(L.MON_TAVG, iround(tavg / dt), [0, 1], period=tavg, name="TemporalAverage")]
    if bold:
        mons.append(make_bold_spec(dt, bold))
    plan = Plan(model=L.MODEL_MPR, cfun=L.CFUN_LINEAR, scheme=L.HEUN_STO, precision=precision, noise=noise,
                conns=[CSR(w, d)], params=mpr_params(G, sig), dt=dt, obs_coef=np.eye(2), monitors=mons,
                max_steps=max_steps, seed=seed, tile=tile, waves=waves, stream=stream)
    return plan


def algorithmic_bytes_per_region_step(nnz, N, ncvar, s, nvoi_t, istep_t, nvoi_b, ii, S, istep_b, noise_bytes=0):
    """Survey model: gather + ring write + noise + monitors."""
    b = s * ncvar * (nnz / N) + s * ncvar + noise_bytes
    b += 4.0 * nvoi_t / istep_t if istep_t else 0.0
    if istep_b:
        b += 8.0 * nvoi_b * (1.0 / ii + S / istep_b)
    return b


def cpu_baseline(w, L, G, sig, n_steps):
    """Reference-faithful numpy oracle, one thread, one instance of the sweep."""
    os.environ.setdefault("OMP_NUM_THREADS", "1")
    from oracle.tvb_oracle import OracleSim
    o = OracleSim(w, L, 5.0, 0.05, "MontbrioPazoRoxin", coupling_params=dict(a=float(G)),
                  integrator="HeunStochastic", nsig=np.array([sig, sig]), initial_conditions=None,
                  monitors=(("TemporalAverage", 1.0, None), ("Bold", 2000.0, None)), mode="faithful")
    z = np.random.default_rng(0).standard_normal((n_steps, 2, w.shape[0]))
    t0 = time.perf_counter()
    o.run(n_steps, noise=z)
    el = time.perf_counter() - t0
    return w.shape[0] * n_steps / el, el


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


_CPU_CHILD = r"""
import sys, time
import numpy as np
sys.path.insert(0, sys.argv[1])
from oracle.tvb_oracle import OracleSim
w = np.load(sys.argv[2]); L = np.load(sys.argv[3]); n_steps = int(sys.argv[4]); seed = int(sys.argv[5])
o = OracleSim(w, L, 5.0, 0.05, "MontbrioPazoRoxin", coupling_params=dict(a=1.0), integrator="HeunStochastic",
              nsig=np.array([1e-3, 1e-3]), initial_conditions=None,
              monitors=(("TemporalAverage", 1.0, None), ("Bold", 2000.0, None)), mode="faithful")
z = np.random.default_rng(seed).standard_normal((n_steps, 2, w.shape[0]))
t0 = time.perf_counter()
o.run(n_steps, noise=z)
print(time.perf_counter() - t0)
"""


def cpu_baseline_all_cores(w, L, n_steps, procs=None):
    """The oracle on every host core the run may use: ``procs`` concurrent
    single-threaded processes (at most 16), one independent instance of the
    sweep each, as the framework runs one simulation per process; the rate
    is all region-steps over the wall time of the slowest process."""
    import subprocess
    import tempfile
    if procs is None:
        procs = max(1, min(len(os.sched_getaffinity(0)), 16))
    env = dict(os.environ, OMP_NUM_THREADS="1", OPENBLAS_NUM_THREADS="1", MKL_NUM_THREADS="1")
    with tempfile.TemporaryDirectory() as tmp:
        pw, pl = os.path.join(tmp, "w.npy"), os.path.join(tmp, "l.npy")
        np.save(pw, w)
        np.save(pl, L)
        root = os.path.dirname(os.path.abspath(__file__))
        t0 = time.perf_counter()
        ps = [subprocess.Popen([sys.executable, "-c", _CPU_CHILD, root, pw, pl, str(n_steps), str(k)],
                               stdout=subprocess.PIPE, stderr=subprocess.PIPE, env=env, text=True)
              for k in range(procs)]
        outs = [p.communicate() for p in ps]
        wall = time.perf_counter() - t0
    for p, (o, e) in zip(ps, outs):
        if p.returncode != 0:
            raise RuntimeError("oracle process failed: " + e[-400:])
    loop = max(float(o.split()[-1]) for o, _ in outs)
    return {"value": procs * w.shape[0] * n_steps / loop, "unit": "region-steps/s", "cores": procs, "kind": "port",
            "sample": "%d processes x 1 instance x %d steps of the c2 workload, reference-faithful numpy oracle, "
                      "1 thread each, slowest loop %.1f s (%.1f s wall with start-up), %s"
                      % (procs, n_steps, loop, wall, cpu_model())}


def gather_pattern_ceiling(roof):
    """The device's rate for the step kernel's memory pattern (random
    256-byte pieces, about 1/17 of the bytes written; tvbk_probe_gather)
    and k_step's measured HBM rate (traffic per launch / launch time) as a
    share of it; the same pattern with 1 KiB pieces for comparison."""
    from tvb_hip import _lib as L
    r = {"unit": "GB/s",
         "reads_256B": L.probe_gather(256, False) / 1e9,
         "reads_256B_writes_1_17": L.probe_gather(256, True) / 1e9,
         "reads_1KiB_writes_1_17": L.probe_gather(1024, True) / 1e9}
    if roof.get("traffic"):
        rate = roof["traffic"] / (roof["avg_launch_ms"] * 1e-3) / 1e9
        r["k_step_hbm_rate"] = rate
        r["k_step_share"] = rate / r["reads_256B_writes_1_17"]
    return r


def c2_rate(w, d, G, sig, precision, n_steps=40000, chunk=20000, stream=None):
    """Throughput of the full sweep in another precision: launches of
    ``chunk`` steps after a warm-up, HIP-event launch times."""
    from tvb_hip import _lib as L
    plan = build_c2_plan(w, d, G, sig, 0.05, chunk, precision, L.NOISE_PHILOX, 1234, stream=stream)
    plan.init_uniform(99, [0.0, -2.0], [2.0, 1.5])
    plan.run(2000)
    plan.sync()
    t0 = time.perf_counter()
    ms = []
    for _ in range(max(1, n_steps // chunk)):
        plan.run(chunk)
        ms.append(plan.last_run_ms())
    plan.sync()
    el = time.perf_counter() - t0
    plan.close()
    return len(G) * w.shape[0] * chunk * len(ms) / el, float(np.mean(ms))


def fp64_single_instance(w, d, G, sig, n_steps):
    """Plain single-instance FP64 device path on the same connectome."""
    from tvb_hip import _lib as L
    plan = build_c2_plan(w, d, np.array([G]), np.array([sig]), 0.05, n_steps, L.F64, L.NOISE_PHILOX, 7)
    plan.init_uniform(11, [0.0, -2.0], [2.0, 1.5])
    plan.run(min(2000, n_steps))
    plan.sync()
    t0 = time.perf_counter()
    plan.run(n_steps)
    plan.sync()
    el = time.perf_counter() - t0
    shape = (plan.kernel, plan.tile, plan.waves)
    plan.close()
    return w.shape[0] * n_steps / el, el, shape


def error_sample(w, L, d, G, sig, precisions, n_steps=8000, B=32, check=(0, 17), bold=100.0):
    """The batched kernel with injected noise vs the FP64 oracle on a sample
    of the sweep (instances ``check`` of B spread over the grid),
    TemporalAverage(1 ms) and Bold(``bold`` ms): max |err| / max |ref| per
    monitor, in two scenarios of ``n_steps`` each:

    * "random_ic": random initial history, as the sweep starts.  Its first
      200 ms are a transient in which trajectories pass near the MPR saddle,
      where any rounding difference grows by about 8% per step;
      TemporalAverage is also reported after those 200 ms.
    * "settled": initial history = the last horizon of states of the FP64
      oracle's random_ic run, so the sample starts on settled dynamics."""
    from tvb_hip import _lib as Lb
    from oracle.tvb_oracle import OracleSim
    idx = np.linspace(0, len(G) - 1, B).astype(int)
    Gs, ss = G[idx], sig[idx]
    H = int(d.max()) + 1
    N = w.shape[0]
    rng = np.random.default_rng(5)
    ic = np.empty((H, 2, N))
    ic[:, 0] = rng.uniform(0.0, 2.0, (H, N))
    ic[:, 1] = rng.uniform(-2.0, 1.5, (H, N))
    z1 = rng.standard_normal((n_steps, 2, N, B))
    z2 = rng.standard_normal((n_steps, 2, N, B))

    def oracle(b, init, z):
        o = OracleSim(w, L, 5.0, 0.05, "MontbrioPazoRoxin", coupling_params=dict(a=float(Gs[b])),
                      integrator="HeunStochastic", nsig=np.array([ss[b], ss[b]]), initial_conditions=init,
                      monitors=(("TemporalAverage", 1.0, None), ("Bold", bold, None)), mode="fp64")
        return o.run(n_steps, noise=z[..., b], record_state=True)

    # per-instance initial histories of the settled scenario
    refs = {"random_ic": {}, "settled": {}}
    settled_ic = np.repeat(ic[..., np.newaxis], B, axis=-1)          # (H, 2, N, B)
    for b in check:
        res, states = oracle(b, ic, z1)
        refs["random_ic"][b] = res
        settled_ic[..., b] = states[-H:]
        refs["settled"][b] = oracle(b, states[-H:], z2)[0]
    out = {"instances": [int(i) for i in idx[list(check)]], "steps": n_steps, "bold_period_ms": bold}
    for name in precisions:
        prec = {"f32": Lb.F32, "mixed": Lb.MIXED}[name]
        res = {}
        for scen, z in (("random_ic", z1), ("settled", z2)):
            plan = build_c2_plan(w, d, Gs, ss, 0.05, n_steps, prec, Lb.NOISE_INJECTED, 0, bold=bold, tile=32)
            for b in range(B):
                init = ic if scen == "random_ic" else settled_ic[..., b]
                plan.set_history(b, init, init[-1], H - 1)
            plan.set_noise(z)
            plan.run(n_steps)
            got_t, got_b = plan.read(0), plan.read(1)
            plan.close()
            et, et_late, eb = [], [], []
            for b in check:
                (tt, xt), (tb, xb) = refs[scen][b]
                rt, rb = xt[..., 0], xb[..., 0]
                dt_ = np.abs(got_t[b][1] - rt)
                et.append(float(dt_.max() / np.abs(rt).max()))
                et_late.append(float(dt_[200:].max() / np.abs(rt).max()))
                eb.append(float(np.abs(got_b[b][1] - rb).max() / np.abs(rb).max()))
            res[scen] = {"tavg_max_rel": max(et), "bold_max_rel": max(eb)}
            if scen == "random_ic":
                res[scen]["tavg_max_rel_after_200ms"] = max(et_late)
        out[name] = res
    return out


def c3_plan(n_inst=256, duration_scale=0.02, seed=303, stream=None):
    """Uneven batch: JansenRit + SigmoidalJansenRit, EulerStochastic dt 0.1,
    node counts from {68, 96, 192, 400, 998}, a distinct seeded connectome
    per instance (density U[0.1, 0.3], lengths U[10, 240] mm at 3 mm/ms ->
    delays up to 800 steps), durations U[2, 30] s x duration_scale,
    per-instance a, b ~ U[0.9, 1.1] x defaults, TemporalAverage(2 ms) + Bold."""
    from tvb_hip import _lib as L
    from tvb_hip.connectome import idelays, synthetic
    from tvb_hip.plan import CSR, MonitorSpec, Plan
    from tvb_hip.sim import iround
    rng = np.random.default_rng(seed)
    dt = 0.1
    conns, steps, params = [], [], []
    base = [3.25, 22.0, 0.1, 0.05, 5.52, 0.0025, 0.56, 135.0, 1.0, 0.8, 0.25, 0.25, 0.12, 0.32, 0.22]
    for b in range(n_inst):
        n = int(rng.choice([68, 96, 192, 400, 998]))
        w, Lm = synthetic(n, float(rng.uniform(0.1, 0.3)), 10.0, 240.0, int(rng.integers(1 << 30)))
        d, H = idelays(Lm, 3.0, dt)
        conns.append(CSR(w, d, H))
        steps.append(int(np.ceil(rng.uniform(2000.0, 30000.0) * duration_scale / dt)))
        p = list(base)
        p[2] *= rng.uniform(0.9, 1.1)
        p[3] *= rng.uniform(0.9, 1.1)
        params.append(p + [0.0, 0.005, 6.0, 0.56, 1.0] + [0.0, 0.0, 0.0, 1e-3, 1e-3, 1e-3])
    mons = [MonitorSpec(L.MON_TAVG, iround(2.0 / dt), [0, 1, 2, 3], name="TemporalAverage"),
            make_bold_spec(dt, 2000.0)]
    mons[1].voi = [0, 1, 2, 3]
    plan = Plan(model=L.MODEL_JR, cfun=L.CFUN_SIGMOID_JR, scheme=L.EULER_STO, precision=L.F32,
                noise=L.NOISE_PHILOX, conns=conns, params=np.array(params), dt=dt, obs_coef=np.eye(6)[:4],
                monitors=mons, max_steps=max(steps), inst_conn=np.arange(n_inst, dtype=np.int32), seed=seed,
                stream=stream)
    plan.set_steps(np.array(steps, np.int64))
    lo = [-1.0, -500.0, -50.0, -6.0, -20.0, -500.0]
    plan.init_uniform(seed + 1, [x * 0.01 for x in lo], [-x * 0.01 for x in lo])
    rs = float(sum(c.n_node * s for c, s in zip(conns, steps)))
    return plan, max(steps), rs, dict(instances=n_inst, duration_scale=duration_scale,
                                      total_nnz=int(sum(c.nnz for c in conns)))


def run_c3(stream=None, n_inst=256, duration_scale=0.02):
    plan, nmax, rs, info = c3_plan(n_inst, duration_scale, stream=stream)
    t0 = time.perf_counter()
    plan.run(nmax)
    plan.sync()
    el = time.perf_counter() - t0
    flags = plan.flags()
    plan.close()
    info.update(value=rs / el, unit="region-steps/s", seconds=el, nonfinite=int((flags != 0).sum()))
    return info


def c1_setup(seed=1920):
    """Single stochastic instance: 192 nodes, 20 % density, lengths
    U[5, 250] mm at 4 mm/ms, MontbrioPazoRoxin, Linear a = 0.5,
    HeunStochastic dt 0.05 ms, nsig [0.01, 0.02], random initial history."""
    from tvb_hip.connectome import idelays, synthetic
    n, dt = 192, 0.05
    w, L = synthetic(n, 0.20, 5.0, 250.0, seed=192)
    d, H = idelays(L, 4.0, dt)
    rng = np.random.default_rng(seed)
    ic = np.empty((H, 2, n))
    ic[:, 0] = rng.uniform(0.0, 2.0, size=(H, n))
    ic[:, 1] = rng.uniform(-2.0, 1.5, size=(H, n))
    return w, L, d, H, ic, rng


def run_c1(n_steps=200000, check_steps=16000, precisions=("f32", "mixed", "f64")):
    """The single-instance configuration: the full 10 s (200 000 steps) with
    injected float64 normals resident on the device, TemporalAverage 1 ms +
    Bold 720 ms, per precision; and the error against the fp64 oracle on
    the first ``check_steps`` steps (one Bold emission)."""
    import torch
    from tvb_hip import _lib as L
    from tvb_hip.plan import CSR, MonitorSpec, Plan
    from oracle.tvb_oracle import OracleSim
    w, Lm, d, H, ic, rng = c1_setup()
    n = w.shape[0]
    z = rng.standard_normal((n_steps, 2, n))
    zd = torch.from_numpy(z.reshape(n_steps, 2, n, 1)).cuda()
    p = np.array([[1.0, 0.0, 1.0, 15.0, -5.0, 0.0, 1.0, 0.0, 0.5, 0.0, 0.01, 0.02]])
    o = OracleSim(w, Lm, 4.0, 0.05, "MontbrioPazoRoxin", coupling_params=dict(a=0.5), integrator="HeunStochastic",
                  nsig=np.array([0.01, 0.02]), initial_conditions=ic,
                  monitors=(("TemporalAverage", 1.0, None), ("Bold", 720.0, None)), mode="fp64")
    (_, tref), (_, bref) = o.run(check_steps, noise=z[:check_steps])
    out = {"nodes": n, "steps": n_steps, "check_steps": check_steps}
    for name in precisions:
        prec = {"f32": L.F32, "mixed": L.MIXED, "f64": L.F64}[name]
        plan = Plan(model=L.MODEL_MPR, cfun=L.CFUN_LINEAR, scheme=L.HEUN_STO, precision=prec,
                    noise=L.NOISE_INJECTED, conns=[CSR(w, d)], params=p, dt=0.05, obs_coef=np.eye(2),
                    monitors=[MonitorSpec(L.MON_TAVG, 20, [0, 1], period=1.0), make_bold_spec(0.05, 720.0)],
                    max_steps=n_steps, seed=0)
        plan.set_history(0, ic, ic[-1], H - 1)
        plan.set_noise(zd.data_ptr(), n_steps)
        plan.sync()
        t0 = time.perf_counter()
        plan.run(n_steps)
        plan.sync()
        el = time.perf_counter() - t0
        (_, ta), = plan.read(0)
        (_, bo), = plan.read(1)
        ntile, k = plan.tiles()
        plan.close()
        kc = check_steps // 20
        out[name] = {"value": n * n_steps / el, "unit": "region-steps/s", "seconds": el, "tiles": ntile,
                     "split_block_steps": k,
                     "tavg_max_rel_vs_fp64_oracle": float(np.abs(ta[:kc] - tref[..., 0]).max() / np.abs(tref).max()),
                     "bold_max_rel_vs_fp64_oracle": float(np.abs(bo[:1] - bref[..., 0]).max() / np.abs(bref).max())}
    return out


def run_c2_zero_delay(B=8192, n_steps=40000, chunk=10000, stream=None, warmup=1000, precision=None):
    """C2 with infinite conduction speed: every delay zero, horizon 1, so
    the coupling of a step is W . x(t) for all instances at once -- a
    (68 x 68) x (68 x B) product per coupling variable, run on the dense
    MFMA kernel.  Returns region-steps/s and the kernel's MFMA roofline
    (2 N ncvar flops per region-step)."""
    from tvb_hip import _lib as L
    w, Lmat, d, H, G, sig = c2_setup(B)
    dz = np.zeros_like(d)
    prec = L.MIXED if precision is None else precision
    plan = build_c2_plan(w, dz, G, sig, 0.05, max(chunk, warmup), prec, L.NOISE_PHILOX, 4321, stream=stream)
    plan.init_uniform(98, [0.0, -2.0], [2.0, 1.5])
    plan.run(warmup)
    plan.sync()
    t0 = time.perf_counter()
    done, ms = 0, []
    while done < n_steps:
        n = min(chunk, n_steps - done)
        plan.run(n)
        ms.append((n, plan.last_run_ms()))
        done += n
    plan.sync()
    el = time.perf_counter() - t0
    full = [m for n, m in ms if n == chunk] or [m for _, m in ms]
    med = float(np.median(full))
    flops = 2.0 * 68 * 2 * 68 * B * chunk
    tf = flops / (med * 1e-3) / 1e12
    out = {"value": 68.0 * B * n_steps / el, "unit": "region-steps/s", "steps": n_steps, "instances": B,
           "kernel": [plan.kernel, plan.tile, plan.waves],
           "roofline": {"bound": "mfma", "achieved": tf, "peak": FP32_MFMA_PEAK_TFLOPS, "unit": "TFLOP/s",
                        "frac": tf / FP32_MFMA_PEAK_TFLOPS, "traffic": None,
                        "flops_per_region_step": 2.0 * 68 * 2, "median_launch_ms": med},
           "nonfinite_instances": int((plan.flags() != 0).sum())}
    plan.close()
    return out


def run_balloon_on_c2(plan_chunk_steps=20000, B=8192):
    """Balloon-Windkessel BOLD (FP64, on the device) of one C2 chunk's
    TemporalAverage output: 1000 samples x 68 x 8192 series, timed with HIP
    events around the kernel."""
    import torch
    from tvb_hip import _lib as L
    from tvb_hip.analyzers import BalloonModel, run_balloon
    T, n = plan_chunk_steps // 20, 68 * B
    x = torch.rand((T, 2, n), dtype=torch.float64, device="cuda")
    y = torch.empty((T, n), dtype=torch.float64, device="cuda")
    st = torch.empty((4, n), dtype=torch.float64, device="cuda")
    par = BalloonModel().params()
    stream = torch.cuda.current_stream().cuda_stream
    run_balloon(x.data_ptr(), 8, 2 * n, T, n, 1.0, par, 0, True, st.data_ptr(), y.data_ptr(), stream)
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ev[0].record()
    run_balloon(x.data_ptr(), 8, 2 * n, T, n, 1.0, par, 0, False, st.data_ptr(), y.data_ptr(), stream)
    ev[1].record()
    torch.cuda.synchronize()
    ms = ev[0].elapsed_time(ev[1])
    return {"samples": T, "series": n, "ms": ms, "series_samples_per_s": T * n / (ms * 1e-3),
            "share_of_chunk": None}


def run_copy(nbytes=2 << 30, reps=10):
    """Plain device copy of a known size (HBM read + write of nbytes each),
    the calibration for FETCH_SIZE / WRITE_SIZE and the measured copy
    bandwidth of the device."""
    import torch
    src = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").uniform_()
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ev[0].record()
    for _ in range(reps):
        dst.copy_(src)
    ev[1].record()
    torch.cuda.synchronize()
    ms = ev[0].elapsed_time(ev[1]) / reps
    return {"bytes_read": nbytes, "bytes_written": nbytes, "ms": ms, "GB/s": 2.0 * nbytes / (ms * 1e-3) / 1e9,
            "launches": reps + 1}


def run_c4(n_requests=24, rate_hz=40.0, duration_scale=0.01, inst_range=(1, 64), seed=404):
    """Serving: Poisson stream of sweep requests (G2D or MPR, HeunStochastic,
    N log-uniform in [68, 998], instances in inst_range, durations
    U[5, 60] s x duration_scale), admitted by device footprint."""
    import torch
    from tvb_hip import _lib
    from tvb_hip.serve import Server, synthetic_requests
    reqs = synthetic_requests(n_requests, rate_hz, seed=seed, inst_range=inst_range,
                              duration_scale=duration_scale)
    streams = [_lib.stream_create() for _ in range(4)]
    free, total = torch.cuda.mem_get_info()
    srv = Server(hbm_budget=int(0.5 * free), streams=streams, chunk_steps=2000, keep_outputs=True)
    stats = srv.run(reqs)
    for st in streams:
        _lib.stream_destroy(st)
    stats.update(duration_scale=duration_scale, rate_hz=rate_hz, inst_range=list(inst_range))
    return stats


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400000, help="timed integration steps (default: the full 20 s)")
    ap.add_argument("--warmup", type=int, default=2000)
    ap.add_argument("--chunk", type=int, default=20000, help="steps per kernel launch")
    ap.add_argument("--instances", type=int, default=8192)
    ap.add_argument("--tile", type=int, default=0)
    ap.add_argument("--waves", type=int, default=0)
    ap.add_argument("--cpu-steps", type=int, default=200000, help="oracle sample length (about 10-20 s of one host core)")
    ap.add_argument("--no-extras", action="store_true", help="skip CPU baseline, FP64 single path and error sample")
    ap.add_argument("--workload", default="c2", choices=("c2", "c1", "c2z", "c3", "c4", "copy"))
    ap.add_argument("--precision", default="f32", choices=("f32", "mixed"),
                    help="f32: float32 state, history and coupling, float64 Bold (the sweep's precision); "
                         "mixed: float32 history and coupling sums, float64 state (the reference's own mix)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU: every instance of the sweep is independent and the "
              "workload fits one MI355X; multi-GPU runs are not part of this build (--gpus %d)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    import torch
    from tvb_hip import _lib as L
    assert torch.cuda.is_available(), "no GPU"
    stream = torch.cuda.current_stream().cuda_stream
    if args.workload == "c1":
        print(json.dumps({"metric": "region-steps/s (single stochastic instance)", "workload": "c1",
                          **run_c1()}), flush=True)
        return
    if args.workload == "c3":
        print(json.dumps({"metric": "region-steps/s (uneven JansenRit batch)", "workload": "c3",
                          **run_c3(stream)}), flush=True)
        return
    if args.workload == "c2z":
        print(json.dumps({"metric": "region-steps/s (zero-delay sweep, dense MFMA)", "workload": "c2z",
                          **run_c2_zero_delay(args.instances, stream=stream,
                                              precision=L.MIXED if args.precision == "mixed" else L.F32)}),
                         flush=True)
        return
    if args.workload == "copy":
        print(json.dumps({"metric": "device copy bandwidth", "workload": "copy", **run_copy()}), flush=True)
        return
    if args.workload == "c4":
        print(json.dumps({"metric": "p50/p99 request latency (serving)", "workload": "c4", **run_c4()}),
              flush=True)
        return
    B = args.instances
    dt = 0.05
    w, Lmat, d, H, G, sig = c2_setup(B)
    nnz = int((w != 0).sum())
    log("workload c2: N=68 nnz=%d H=%d B=%d steps=%d" % (nnz, H, B, args.steps))
    chunk = min(args.chunk, max(args.steps, 1))
    prec = L.MIXED if args.precision == "mixed" else L.F32
    plan = build_c2_plan(w, d, G, sig, dt, max(chunk, args.warmup), prec, L.NOISE_PHILOX, 1234,
                         tile=args.tile, waves=args.waves, stream=stream)
    log("plan: kernel=%d tile=%d waves=%d lds=%d" % (plan.kernel, plan.tile, plan.waves, plan.lds_bytes))
    plan.init_uniform(99, [0.0, -2.0], [2.0, 1.5])
    # warmup
    if args.warmup:
        plan.run(args.warmup)
    torch.cuda.synchronize()
    plan.sync()
    # timed region
    t0 = time.perf_counter()
    done, launches = 0, []
    while done < args.steps:
        n = min(chunk, args.steps - done)
        plan.run(n)
        ms = plan.last_run_ms()                   # waits for this launch (HIP events)
        launches.append((n, ms))
        done += n
        log("  %d / %d steps, launch %.1f ms" % (done, args.steps, ms))
    plan.sync()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    flags = plan.flags()
    rs = 68.0 * B * args.steps
    value = rs / el
    # roofline of the step kernel
    full = [(n, ms) for n, ms in launches if n == chunk] or launches
    avg_ms = sum(ms for _, ms in full) / len(full)
    med_ms = float(np.median([ms for _, ms in full]))
    bpr = algorithmic_bytes_per_region_step(nnz, 68, 2, 4, 2, 20, 2, 80, 5000, 40000)
    per_launch = bpr * 68 * B * full[0][0]
    achieved = per_launch / (avg_ms * 1e-3) / 1e9
    tpr = measured_traffic_per_region_step()
    traffic = tpr * 68 * B * full[0][0] if tpr else None
    line = {
        "metric": "region-steps/s batched FP32 vs plain FP64 at stated error; p50/p99 latency",
        "value": value, "unit": "region-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": el * 1e3 / max(args.steps, 1), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 history/coupling, f64 state" if args.precision == "mixed" else "f32",
        "data": "synthetic",
        "config": {"workload": "c2: MPR HeunStochastic dt=0.05 sweep, 68 nodes x %d instances, shared "
                               "25%%-density connectome (nnz %d, horizon %d), TemporalAverage 1 ms + Bold 2000 ms, "
                               "device Philox noise" % (B, nnz, H),
                   "instances": B, "nodes": 68, "precision": args.precision, "kernel_tile": plan.tile,
                   "kernel_waves": plan.waves,
                   "steps_per_launch": chunk},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes_per_launch": per_launch, "algorithmic_bytes_per_region_step": bpr,
                     "traffic_bytes_per_region_step": tpr, "avg_launch_ms": avg_ms, "median_launch_ms": med_ms,
                     "launches": len(full)},
        "nonfinite_instances": int((flags != 0).sum()),
    }
    plan.close()
    if not args.no_extras:
        line["gather_pattern_ceiling"] = gather_pattern_ceiling(line["roofline"])
        v64, el64, shape64 = fp64_single_instance(w, d, 1.0, 1e-3, 40000)
        line["fp64_single_instance"] = {"value": v64, "unit": "region-steps/s", "steps": 40000,
                                        "kernel": list(shape64)}
        line["batched_fp32_over_fp64_single"] = value / v64
        other = "mixed" if args.precision == "f32" else "f32"
        v2, ms2 = c2_rate(w, d, G, sig, L.MIXED if other == "mixed" else L.F32, chunk=chunk, stream=stream)
        line["other_precision"] = {"precision": other, "value": v2, "unit": "region-steps/s",
                                   "avg_launch_ms": ms2, "roofline_frac": bpr * 68 * B * chunk / (ms2 * 1e-3) / 1e9
                                   / HBM_PEAK_GBS, "steps": 40000}
        errs = error_sample(w, Lmat, d, G, sig, ("f32", "mixed"))
        line["error_vs_fp64_oracle"] = errs
        e = errs[args.precision]
        line["tavg_max_rel_err_vs_fp64_oracle"] = {
            "settled": e["settled"]["tavg_max_rel"], "random_ic": e["random_ic"]["tavg_max_rel"],
            "random_ic_after_200ms": e["random_ic"]["tavg_max_rel_after_200ms"]}
        line["bold_max_rel_err_vs_fp64_oracle"] = {"settled": e["settled"]["bold_max_rel"],
                                                   "random_ic": e["random_ic"]["bold_max_rel"]}
        line["serving"] = run_c4()
        line["p50_latency_s"] = line["serving"]["p50_latency_s"]
        line["p99_latency_s"] = line["serving"]["p99_latency_s"]
        line["uneven_batch"] = run_c3(stream)
        line["single_instance_c1"] = run_c1()
        line["zero_delay_mfma"] = run_c2_zero_delay(B, stream=stream, precision=prec)
        bal = run_balloon_on_c2(chunk, B)
        bal["share_of_chunk"] = bal["ms"] / avg_ms
        line["balloon_bold_fp64"] = bal
        cpu_rate, cpu_el = cpu_baseline(w, Lmat, 1.0, 1e-3, args.cpu_steps)
        line["cpu_baseline"] = {"value": cpu_rate, "unit": "region-steps/s", "cores": 1, "kind": "port",
                                "sample": "1 instance x %d steps (%.1f s simulated) of the c2 workload, "
                                          "reference-faithful numpy oracle, 1 thread, %.1f s wall, %s"
                                          % (args.cpu_steps, args.cpu_steps * dt / 1000.0, cpu_el, cpu_model())}
        line["cpu_baseline_all_cores"] = cpu_baseline_all_cores(w, Lmat, args.cpu_steps // 4)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
