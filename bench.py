#!/usr/bin/env python
"""Benchmark of the TVB simulator loop on one MI355X.

Headline workload ("c2"): batched parameter sweep on one shared 68-node
synthetic connectome (25 % density, log-normal weights, tract lengths
U[10, 180] mm, speed 5 mm/ms -> delays 40..720 steps at dt 0.05 ms),
8192 instances on a 128 x 64 grid of coupling gain G in [0, 2] (Linear a)
x noise sigma in [1e-4, 1e-2] (Additive nsig of both state variables),
MontbrioPazoRoxin, HeunStochastic dt = 0.05 ms, TemporalAverage(1 ms) +
Bold(2000 ms), device Philox noise, inputs resident in HBM.

One bench "step" is one kernel launch that advances all 8192 x 68 region
states by ``--launch-steps`` integration steps (default 20 000 = 1 s
simulated, 1000 TemporalAverage samples, 250 Bold interim pushes; every
second launch emits a Bold sample).  ``--steps 20`` times the sweep's full
20 s.  ``--precision`` picks the arithmetic: "mixed" (default; float32
history ring, weights and coupling sums, float64 state, integration and
monitors -- the reference's own mix, the precision that meets the stated
error bar) or "f32" (float32 state too).

Prints ONE JSON line: the headline (region-steps/s, the step kernel's
roofline from HIP events on the launch stream, the timed work), then the
extras, each run under its own guard so that a failing extra is recorded
in the line instead of losing it: the other precision, the plain
single-instance FP64 device path, the error against the FP64 oracle, the
gather-pattern ceiling, C1, C3, C4, the zero-delay MFMA variant, the
Balloon-Windkessel pass and the CPU baselines.
"""
import argparse
import hashlib
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
CSRC = os.path.join(ROOT, "tvb-root_amd", "csrc")
ERROR_BAR = {"tavg_max_rel": 1e-4, "bold_max_rel": 1e-6}
PREC_NAMES = ("f32", "f64", "mixed")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def kernel_source_sha16():
    """Hash of the HIP library's sources (every file under csrc/): a traffic
    profile taken on other sources does not describe the built kernel."""
    h = hashlib.sha256()
    for f in sorted(os.listdir(CSRC)):
        with open(os.path.join(CSRC, f), "rb") as fh:
            h.update(f.encode() + b"\0" + fh.read() + b"\0")
    return h.hexdigest()[:16]


def measured_traffic(precision, channels, path=TRAFFIC_SUMMARY):
    """HBM bytes per region-step of the step kernel from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary, if that summary was taken on
    the current sources for the same precision and history channels.
    Returns (bytes per region-step or None, note)."""
    keys = {}
    try:
        with open(path) as f:
            for ln in f:
                s = ln.strip().strip("`")
                for k in ("traffic_bytes_per_region_step", "kernel_source_sha16", "precision", "channels"):
                    if s.startswith(k + ":"):
                        keys[k] = s.split(":", 1)[1].split()[0]
    except OSError:
        return None, "no traffic summary"
    if keys.get("kernel_source_sha16") != kernel_source_sha16():
        return None, "traffic summary was measured on other kernel sources (%s, built %s)" % (
            keys.get("kernel_source_sha16"), kernel_source_sha16())
    if keys.get("precision") != precision or keys.get("channels") != str(channels):
        return None, "traffic summary is for precision %s / %s channels" % (keys.get("precision"),
                                                                             keys.get("channels"))
    try:
        return float(keys["traffic_bytes_per_region_step"]), "rocprofv3 FETCH_SIZE+WRITE_SIZE, " + path
    except (KeyError, ValueError):
        return None, "traffic summary without a value"


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


def make_bold_spec(dt, period=2000.0, voi=(0, 1)):
    from tvb_hip import _lib as L
    from tvb_hip.plan import MonitorSpec
    from tvb_hip.sim import Bold, iround

    b = Bold(period=period)
    b.dt = dt
    b.istep = iround(period / dt)
    b.compute_hrf()
    taps = b.hemodynamic_response_function[0, ::-1].copy()
    k1v0 = b.hrf_kernel.parameters["k_1"] * b.hrf_kernel.parameters["V_0"]
    return MonitorSpec(L.MON_BOLD, b.istep, list(voi), interim_istep=b._interim_istep, taps=taps,
                       k1_v0=k1v0, subtract_one=True, period=period, name="Bold")


def build_c2_plan(w, d, G, sig, dt, max_steps, precision, noise, seed, tile=0, waves=0, tavg=1.0, bold=2000.0,
                  stream=None, kernel=0, skip=False, tune=False):
    """The C2 sweep's plan; ``skip``: MontbrioPazoRoxin's V channel, whose
    coupling coefficient cv is 0 in every instance, is not kept or gathered
    (PLAN_SKIP_ZERO_COUPLING)."""
    from tvb_hip import _lib as L
    from tvb_hip.plan import CSR, MonitorSpec, Plan
    from tvb_hip.sim import iround
    mons = [MonitorSpec(L.MON_TAVG, iround(tavg / dt), [0, 1], period=tavg, name="TemporalAverage")]
    if bold:
        mons.append(make_bold_spec(dt, bold))
    plan = Plan(model=L.MODEL_MPR, cfun=L.CFUN_LINEAR, scheme=L.HEUN_STO, precision=precision, noise=noise,
                conns=[CSR(w, d)], params=mpr_params(G, sig), dt=dt, obs_coef=np.eye(2), monitors=mons,
                max_steps=max_steps, seed=seed, tile=tile, waves=waves, stream=stream, kernel=kernel,
                flags=L.PLAN_SKIP_ZERO_COUPLING if skip else 0, tune=tune)
    return plan


def algorithmic_bytes_per_region_step(nnz, N, ncvar, s, nvoi_t, istep_t, nvoi_b, ii, S, istep_b, noise_bytes=0,
                                      tavg_bytes=4):
    """Gather (one history value per edge and channel) + ring write + noise
    + monitors (TemporalAverage output, Bold interim push into the pending
    HRF sums and their read at emission)."""
    b = s * ncvar * (nnz / N) + s * ncvar + noise_bytes
    b += float(tavg_bytes) * nvoi_t / istep_t if istep_t else 0.0
    if istep_b:
        b += 8.0 * nvoi_b * (1.0 / ii + S / istep_b)
    return b


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


# ---------------------------------------------------------------------------
# CPU baselines: the reference-faithful numpy oracle on the host cores
# ---------------------------------------------------------------------------

def cpu_baseline(w, L, G, sig, n_steps):
    """Reference-faithful numpy oracle, one thread, one instance of the sweep."""
    from oracle.tvb_oracle import OracleSim
    o = OracleSim(w, L, 5.0, 0.05, "MontbrioPazoRoxin", coupling_params=dict(a=float(G)),
                  integrator="HeunStochastic", nsig=np.array([sig, sig]), initial_conditions=None,
                  monitors=(("TemporalAverage", 1.0, None), ("Bold", 2000.0, None)), mode="faithful")
    z = np.random.default_rng(0).standard_normal((n_steps, 2, w.shape[0]))
    t0 = time.perf_counter()
    o.run(n_steps, noise=z)
    el = time.perf_counter() - t0
    return w.shape[0] * n_steps / el, el


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


def host_procs():
    return max(1, min(len(os.sched_getaffinity(0)), 16))


def child_env():
    return dict(os.environ, OMP_NUM_THREADS="1", OPENBLAS_NUM_THREADS="1", MKL_NUM_THREADS="1")


def cpu_baseline_all_cores(w, L, n_steps, procs=None):
    """The oracle on every host core the run may use: ``procs`` concurrent
    single-threaded processes (at most 16), one independent instance of the
    sweep each, as the framework runs one simulation per process; the rate
    is all region-steps over the wall time of the slowest process."""
    import subprocess
    import tempfile
    procs = procs or host_procs()
    with tempfile.TemporaryDirectory() as tmp:
        pw, pl = os.path.join(tmp, "w.npy"), os.path.join(tmp, "l.npy")
        np.save(pw, w)
        np.save(pl, L)
        t0 = time.perf_counter()
        ps = [subprocess.Popen([sys.executable, "-c", _CPU_CHILD, ROOT, pw, pl, str(n_steps), str(k)],
                               stdout=subprocess.PIPE, stderr=subprocess.PIPE, env=child_env(), text=True)
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


# ---------------------------------------------------------------------------
# C2 headline and its relatives
# ---------------------------------------------------------------------------

def c2_headline(args, stream):
    """The timed region: ``args.steps`` launches of ``args.launch_steps``
    integration steps over the whole sweep after ``args.warmup`` untimed
    launches; launch durations from HIP events on the plan's stream."""
    import torch
    from tvb_hip import _lib as L
    B, dt, K = args.instances, 0.05, args.launch_steps
    w, Lmat, d, H, G, sig = c2_setup(B)
    nnz = int((w != 0).sum())
    prec = {"mixed": L.MIXED, "f32": L.F32}[args.precision]
    skip = not args.all_channels
    plan = build_c2_plan(w, d, G, sig, dt, K, prec, L.NOISE_PHILOX, 1234, tile=args.tile, waves=args.waves,
                         stream=stream, skip=skip, tune=not args.no_tune)
    tuned = plan.tuned
    log("c2: N=68 nnz=%d H=%d B=%d, %d launches x %d steps, plan kernel=%d tile=%d waves=%d lds=%d"
        % (nnz, H, B, args.steps, K, plan.kernel, plan.tile, plan.waves, plan.lds_bytes))
    plan.init_uniform(99, [0.0, -2.0], [2.0, 1.5])
    for k in range(args.warmup):
        plan.run(K)
        log("  warmup %d / %d, launch %.1f ms" % (k + 1, args.warmup, plan.last_run_ms()))
    plan.sync()
    torch.cuda.synchronize()
    emitted = {"tavg_samples": 0, "bold_samples": 0}
    t0 = time.perf_counter()
    launches = []
    for k in range(args.steps):
        plan.run(K)
        ms = plan.last_run_ms()                   # waits for this launch (HIP events on the plan's stream)
        launches.append(ms)
        emitted["tavg_samples"] += int(plan.layout(0)[3][0])
        emitted["bold_samples"] += int(plan.layout(1)[3][0])
        log("  %d / %d launches, %.1f ms" % (k + 1, args.steps, ms))
    plan.sync()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    flags = plan.flags()
    nch = plan.nch
    tile, waves, kernel, loop = plan.tile, plan.waves, plan.kernel, plan.loop()
    plan.close()
    rs = 68.0 * B * K * args.steps
    avg_ms = float(np.mean(launches))
    bpr = algorithmic_bytes_per_region_step(nnz, 68, nch, 4, 2, 20, 2, 80, 5000, 40000,
                                            tavg_bytes=8 if args.precision == "mixed" else 4)
    per_launch = bpr * 68 * B * K
    achieved = per_launch / (avg_ms * 1e-3) / 1e9
    tpr, tnote = measured_traffic(args.precision, nch)
    dtype = ("mixed: f32 history ring, weights and coupling sums; f64 state, integration and monitors"
             if args.precision == "mixed" else "f32 (state, history, coupling; f64 Bold sums)")
    line = {
        "metric": "region-steps/s batched FP32 vs plain FP64 at stated error; p50/p99 latency",
        "value": rs / el, "unit": "region-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": el * 1e3 / max(args.steps, 1), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": dtype,
        "data": "synthetic (seeded connectome, parameter grid and device Philox noise; random initial history)",
        "config": {"workload": "c2: MontbrioPazoRoxin HeunStochastic dt=0.05 ms sweep, 68 nodes x %d instances "
                               "(128 G x 64 sigma), shared 25%%-density connectome (nnz %d, horizon %d), "
                               "TemporalAverage 1 ms + Bold 2000 ms" % (B, nnz, H),
                   "instances": B, "nodes": 68, "precision": args.precision, "kernel": kernel,
                   "kernel_tile": tile, "kernel_waves": waves, "kernel_loop": loop, "history_channels": nch,
                   "kernel_shape_from": "autotuner (trial launches of 2000 steps)" if tuned else "built-in heuristic",
                   "autotune": tuned,
                   "skipped_channels": ["V (MontbrioPazoRoxin cv = 0: its coupling term is 0 * c_V)"]
                   if nch == 1 else []},
        "timed": {"what": "one step = one launch of %d integration steps (%.0f ms simulated) of all %d x 68 "
                          "region states" % (K, K * dt, B),
                  "integration_steps": K * args.steps, "simulated_ms": K * args.steps * dt,
                  "region_steps": rs, "seconds": el, **emitted},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": tpr * 68 * B * K if tpr else None, "traffic_note": tnote,
                     "kernel": "k_step (batch-lane, %s loop)" % loop,
                     "algorithmic_bytes_per_launch": per_launch, "algorithmic_bytes_per_region_step": bpr,
                     "traffic_bytes_per_region_step": tpr, "avg_launch_ms": avg_ms,
                     "median_launch_ms": float(np.median(launches)), "launches": len(launches)},
        "nonfinite_instances": int((flags != 0).sum()),
        "nonfinite_note": "instances whose state left the finite range (the high-G part of the grid diverges "
                          "in the FP64 oracle as well); their work is done and counted like the others",
    }
    ctx = dict(w=w, Lmat=Lmat, d=d, H=H, G=G, sig=sig, nnz=nnz, bpr=bpr, avg_ms=avg_ms, value=rs / el, nch=nch,
               skip=skip)
    return line, ctx


def c2_rate(ctx, precision, launch_steps, n_launch=4, stream=None, skip=None):
    """Throughput of the full sweep in another precision (or with all
    history channels): ``n_launch`` launches of ``launch_steps`` after one
    warm-up launch."""
    from tvb_hip import _lib as L
    skip = ctx["skip"] if skip is None else skip
    plan = build_c2_plan(ctx["w"], ctx["d"], ctx["G"], ctx["sig"], 0.05, launch_steps, precision, L.NOISE_PHILOX,
                         1234, stream=stream, skip=skip)
    plan.init_uniform(99, [0.0, -2.0], [2.0, 1.5])
    plan.run(launch_steps)
    plan.sync()
    t0 = time.perf_counter()
    ms = []
    for _ in range(n_launch):
        plan.run(launch_steps)
        ms.append(plan.last_run_ms())
    plan.sync()
    el = time.perf_counter() - t0
    nch = plan.nch
    plan.close()
    B = len(ctx["G"])
    bpr = algorithmic_bytes_per_region_step(ctx["nnz"], 68, nch, 4, 2, 20, 2, 80, 5000, 40000,
                                            tavg_bytes=8 if precision == L.MIXED else 4)
    avg = float(np.mean(ms))
    return {"precision": PREC_NAMES[precision], "history_channels": nch,
            "value": B * 68.0 * launch_steps * n_launch / el,
            "unit": "region-steps/s", "avg_launch_ms": avg, "launches": n_launch, "launch_steps": launch_steps,
            "roofline_frac": bpr * 68 * B * launch_steps / (avg * 1e-3) / 1e9 / HBM_PEAK_GBS}


def fp64_single_instance(ctx, G=1.0, sig=1e-3, n_steps=400000):
    """Plain single-instance FP64 device path on the same connectome and
    duration (20 s = 400 000 steps), the denominator of the batched ratio."""
    from tvb_hip import _lib as L
    plan = build_c2_plan(ctx["w"], ctx["d"], np.array([G]), np.array([sig]), 0.05, n_steps, L.F64,
                         L.NOISE_PHILOX, 7, skip=ctx["skip"])
    plan.init_uniform(11, [0.0, -2.0], [2.0, 1.5])
    plan.run(2000)
    plan.sync()
    t0 = time.perf_counter()
    plan.run(n_steps)
    plan.sync()
    el = time.perf_counter() - t0
    shape = [plan.kernel, plan.tile, plan.waves]
    ntile, k = plan.tiles()
    nch = plan.nch
    plan.close()
    v = 68.0 * n_steps / el
    return {"value": v, "unit": "region-steps/s", "steps": n_steps, "seconds": el, "kernel": shape,
            "tiles": ntile, "split_block_steps": k, "history_channels": nch,
            "batched_over_fp64_single": ctx["value"] / v}


def gather_pattern_ceiling(line):
    """The device's rate for the step kernel's memory pattern (random
    256-byte pieces, about 1/17 of the bytes written; tvbk_probe_gather)."""
    from tvb_hip import _lib as L
    r = {"unit": "GB/s",
         "reads_128B_pairs": L.probe_gather(128, False) / 1e9,
         "reads_128B_pairs_writes_1_17": L.probe_gather(128, True) / 1e9,
         "reads_256B": L.probe_gather(256, False) / 1e9,
         "reads_256B_writes_1_17": L.probe_gather(256, True) / 1e9,
         "reads_1KiB_writes_1_17": L.probe_gather(1024, True) / 1e9}
    roof = line["roofline"]
    # the headline's pieces: 128 bytes with one history channel (tile of
    # 32), 256 with two
    ceil = r["reads_128B_pairs_writes_1_17"] if line["config"]["history_channels"] == 1 else \
        r["reads_256B_writes_1_17"]
    r["pattern_ceiling_of_headline"] = ceil
    r["algorithmic_share"] = roof["achieved"] / ceil
    if roof.get("traffic"):
        rate = roof["traffic"] / (roof["avg_launch_ms"] * 1e-3) / 1e9
        r["k_step_hbm_rate"] = rate
        r["k_step_share"] = rate / ceil
    return r


# ---------------------------------------------------------------------------
# error of the sweep against the FP64 oracle
# ---------------------------------------------------------------------------

# (G index, sigma index) of the checked instances on the 128 x 64 grid,
# corners and interior, with the high-G / high-sigma corner
ERROR_GRID = ((0, 0), (0, 63), (20, 40), (32, 10), (48, 50), (64, 0), (64, 63), (80, 25), (96, 63), (100, 5),
              (115, 63), (118, 30), (121, 63), (124, 0), (127, 40), (127, 63))


def _noise(seed, b, n_steps, N):
    return np.random.default_rng([seed, b]).standard_normal((n_steps, 2, N))


def _half_ulp_perturbation(shape, seed=11):
    """Relative moves of +-2^-24 (half a float32 ulp) with a fixed sign
    pattern: the size of the rounding a float32 history applies once."""
    return 1.0 + np.random.default_rng(seed).choice([-1.0, 1.0], size=shape) * 2.0 ** -24


def _oracle_error_child(inp, outp):
    """Child process of the error sample, one instance, two scenarios
    (random initial history; then settled, starting from the last horizon of
    the first FP64 run's states), three oracle runs each on the same
    normals: FP64 (the reference of the error), FP64 from initial values
    moved by half a float32 ulp (the conditioning floor: what one float32
    rounding of the inputs alone changes), and the reference-faithful
    precision mix (float32 history ring, float64 state; what the reference
    Simulator computes)."""
    from oracle.tvb_oracle import OracleSim
    a = np.load(inp)
    w, L, ic = a["w"], a["L"], a["ic"]
    G, s, n_steps, b, bold, H = float(a["G"]), float(a["sig"]), int(a["n_steps"]), int(a["b"]), float(a["bold"]), \
        int(a["H"])
    N = w.shape[0]

    def run(init, z, mode):
        o = OracleSim(w, L, 5.0, 0.05, "MontbrioPazoRoxin", coupling_params=dict(a=G), integrator="HeunStochastic",
                      nsig=np.array([s, s]), initial_conditions=init,
                      monitors=(("TemporalAverage", 1.0, None), ("Bold", bold, None)), mode=mode)
        with np.errstate(all="ignore"):
            ((_, t), (_, bo)), st = o.run(n_steps, noise=z, record_state=True)
        return t[..., 0], bo[..., 0], st

    out = {}
    init = ic
    for key, seed in (("1", int(a["seed1"])), ("2", int(a["seed2"]))):
        z = _noise(seed, b, n_steps, N)
        t, bo, st = run(init, z, "fp64")
        tp, bp, stp = run(init * _half_ulp_perturbation(init.shape), z, "fp64")
        tf, bf, stf = run(init, z, "faithful")
        out.update({"tavg" + key: t, "bold" + key: bo, "finite" + key: np.isfinite(st).all(),
                    "tavg_p" + key: tp, "bold_p" + key: bp, "finite_p" + key: np.isfinite(stp).all(),
                    "tavg_f" + key: tf, "bold_f" + key: bf, "finite_f" + key: np.isfinite(stf).all(),
                    "init" + key: init})
        if key == "1":
            init = st[-H:]
    np.savez(outp, **out)


def _rel(x, ref):
    """max |x - ref| / max |ref|; inf when x is not finite where ref is."""
    with np.errstate(all="ignore"):
        e = float(np.abs(x - ref).max() / np.abs(ref).max())
    return e if np.isfinite(e) else float("inf")


def error_sample(ctx, precisions, n_steps=42000, bold=2000.0, grid=ERROR_GRID, seeds=(7001, 7002)):
    """The batched kernel (tile of 32, the headline's pipelined loop) with
    injected noise vs the FP64 oracle on ``len(grid)`` instances of the
    sweep, TemporalAverage(1 ms) and Bold(``bold`` ms, >= one emission):
    max |err| / max |ref| per instance and monitor, in two scenarios of
    ``n_steps`` each: "random_ic" (random initial history, as the sweep
    starts) and "settled" (initial history = the last horizon of the FP64
    oracle's random_ic states).

    Beside each error stands the instance's conditioning floor: the same
    FP64 oracle from initial values moved by half a float32 ulp.  Where the
    floor exceeds the error bar (the high-coupling part of the grid, where
    noise-driven switching amplifies any rounding to O(1) within a second),
    no computation that rounds once to float32 -- the reference's own
    float32 history included -- can meet the bar; the error bar is judged on
    the instances whose floor is within it, and the others are listed with
    the reference precision's own error beside the device's.  The MIXED
    device path is also compared with the reference-faithful oracle (the
    reference's arithmetic) element for element."""
    import subprocess
    import tempfile
    from tvb_hip import _lib as Lb
    w, Lmat, d, H, G, sig = ctx["w"], ctx["Lmat"], ctx["d"], ctx["H"], ctx["G"], ctx["sig"]
    N = w.shape[0]
    n_s = 64
    idx = np.array([gi * n_s + si for gi, si in grid])
    Gs, ss = G[idx], sig[idx]
    B = len(idx)
    rng = np.random.default_rng(5)
    ic = np.empty((H, 2, N))
    ic[:, 0] = rng.uniform(0.0, 2.0, (H, N))
    ic[:, 1] = rng.uniform(-2.0, 1.5, (H, N))
    t0 = time.perf_counter()
    with tempfile.TemporaryDirectory() as tmp:
        procs = []
        for b in range(B):
            inp, outp = os.path.join(tmp, "in%d.npz" % b), os.path.join(tmp, "out%d.npz" % b)
            np.savez(inp, w=w, L=Lmat, ic=ic, G=Gs[b], sig=ss[b], n_steps=n_steps, b=b, bold=bold, H=H,
                     seed1=seeds[0], seed2=seeds[1])
            procs.append((subprocess.Popen([sys.executable, os.path.abspath(__file__), "--oracle-child", inp, outp],
                                           stdout=subprocess.DEVNULL, stderr=subprocess.PIPE, env=child_env(),
                                           text=True), outp))
            if len(procs) - sum(p.poll() is not None for p, _ in procs) >= host_procs():
                procs[-host_procs()][0].wait()
        # device runs of the random_ic scenario while the oracle children work
        z1 = np.empty((n_steps, 2, N, B))
        for b in range(B):
            z1[..., b] = _noise(seeds[0], b, n_steps, N)
        dev = {}

        def device(prec, inits, z):
            plan = build_c2_plan(w, d, Gs, ss, 0.05, n_steps, prec, Lb.NOISE_INJECTED, 0, bold=bold, tile=32,
                                 kernel=Lb.KERNEL_BATCH_LANE, skip=ctx["skip"])
            for b in range(B):
                plan.set_history(b, inits[b][:, :plan.nch], inits[b][-1], H - 1)
            plan.set_noise(z)
            plan.run(n_steps)
            out = (plan.read(0), plan.read(1), plan.flags(), [plan.kernel, plan.tile, plan.waves])
            plan.close()
            return out

        precs = {"f32": Lb.F32, "mixed": Lb.MIXED}
        for name in precisions:
            dev[(name, "random_ic")] = device(precs[name], [ic] * B, z1)
        del z1
        refs = []
        for p, outp in procs:
            _, err = p.communicate()
            if p.returncode != 0:
                raise RuntimeError("oracle child failed: " + err[-400:])
            refs.append(dict(np.load(outp)))
        z2 = np.empty((n_steps, 2, N, B))
        for b in range(B):
            z2[..., b] = _noise(seeds[1], b, n_steps, N)
        for name in precisions:
            dev[(name, "settled")] = device(precs[name], [r["init2"] for r in refs], z2)
        del z2
    bar_t, bar_b = ERROR_BAR["tavg_max_rel"], ERROR_BAR["bold_max_rel"]
    out = {"instances": [int(i) for i in idx], "G": [float(x) for x in Gs], "sigma": [float(x) for x in ss],
           "steps": n_steps, "bold_period_ms": bold, "tavg_period_ms": 1.0, "error_bar": ERROR_BAR,
           "oracle": "fp64 numpy restatement, one process per instance",
           "conditioning_floor": "fp64 oracle from initial values moved by half a float32 ulp (+-2^-24 relative), "
                                 "same normals; an instance is well conditioned when its floor is within the bar",
           "seconds": time.perf_counter() - t0}
    per = []
    for scen, key in (("random_ic", "1"), ("settled", "2")):
        for b in range(B):
            r = refs[b]
            fin = bool(r["finite" + key]) and bool(np.isfinite(r["tavg" + key]).all()) and \
                bool(np.isfinite(r["bold" + key]).all())
            rt, rb = r["tavg" + key], r["bold" + key]
            e = {"scenario": scen, "instance": int(idx[b]), "G": float(Gs[b]), "sigma": float(ss[b]),
                 "oracle_finite": fin, "perturbed_finite": bool(r["finite_p" + key]),
                 "faithful_finite": bool(r["finite_f" + key])}
            if fin:
                e["floor_tavg"], e["floor_bold"] = _rel(r["tavg_p" + key], rt), _rel(r["bold_p" + key], rb)
                e["faithful_tavg"], e["faithful_bold"] = _rel(r["tavg_f" + key], rt), _rel(r["bold_f" + key], rb)
                e["well_conditioned"] = e["floor_tavg"] <= bar_t and e["floor_bold"] <= bar_b
            for name in precisions:
                tav, bol, flags, _ = dev[(name, scen)]
                e[name + "_flag"] = int(flags[b])
                if fin:
                    e[name + "_tavg"], e[name + "_bold"] = _rel(tav[b][1], rt), _rel(bol[b][1], rb)
                if name == "mixed" and e["faithful_finite"]:
                    # the reference's arithmetic: TemporalAverage element for
                    # element, Bold to the order of its HRF sum (where it stays
                    # finite: past a divergence the skipped V channel's 0 * c_V
                    # is 0 on the device and NaN in the reference)
                    e["mixed_tavg_equals_faithful"] = bool(np.array_equal(tav[b][1], r["tavg_f" + key]))
                    e["mixed_bold_vs_faithful"] = _rel(bol[b][1], r["bold_f" + key])
            per.append(e)
    for name in precisions:
        res = {}
        for scen in ("random_ic", "settled"):
            rows = [e for e in per if e["scenario"] == scen]
            wc = [e for e in rows if e.get("well_conditioned")]
            ill = [e for e in rows if e["oracle_finite"] and not e.get("well_conditioned")]
            # finiteness that one float32 rounding does not change
            stable = [e for e in rows if e["oracle_finite"] == e["perturbed_finite"]]
            mx = lambda key, es: max((e[key] for e in es), default=None)  # noqa: E731
            res[scen] = {
                "well_conditioned_instances": len(wc),
                "tavg_max_rel": mx(name + "_tavg", wc), "bold_max_rel": mx(name + "_bold", wc),
                "reference_precision_tavg_max_rel": mx("faithful_tavg", wc),
                "reference_precision_bold_max_rel": mx("faithful_bold", wc),
                "ill_conditioned": [{"instance": e["instance"], "G": round(e["G"], 3),
                                     "floor_tavg": e["floor_tavg"], "faithful_tavg": e["faithful_tavg"],
                                     name + "_tavg": e[name + "_tavg"]} for e in ill],
                "nonfinite_in_oracle": [e["instance"] for e in rows if not e["oracle_finite"]],
                "device_flags_agree_where_finiteness_is_stable": all(
                    (e[name + "_flag"] != 0) == (not e["oracle_finite"]) for e in stable),
            }
            if name == "mixed":
                eq = [e["mixed_tavg_equals_faithful"] for e in rows if "mixed_tavg_equals_faithful" in e]
                res[scen]["tavg_bit_exact_vs_reference_precision"] = "%d of %d instances where it stays finite" % (
                    sum(eq), len(eq))
                res[scen]["bold_max_rel_vs_reference_precision"] = max(
                    (e["mixed_bold_vs_faithful"] for e in rows if "mixed_bold_vs_faithful" in e), default=None)
        res["meets_error_bar"] = all(
            res[s]["well_conditioned_instances"] > 0 and res[s]["tavg_max_rel"] <= bar_t and
            res[s]["bold_max_rel"] <= bar_b for s in ("random_ic", "settled"))
        out[name] = res
    for e in per:
        for k, v in e.items():
            if isinstance(v, float) and k not in ("G", "sigma"):
                e[k] = float("%.3g" % v)
    out["per_instance"] = per
    return out


# ---------------------------------------------------------------------------
# the other configurations
# ---------------------------------------------------------------------------

def c3_plan(n_inst=1024, duration_scale=1.0, seed=303, stream=None, max_steps=None, precision=None):
    """Uneven batch: JansenRit + SigmoidalJansenRit, EulerStochastic dt 0.1,
    node counts from {68, 96, 192, 400, 998}, a distinct seeded connectome
    per instance (density U[0.1, 0.3], lengths U[10, 240] mm at 3 mm/ms ->
    delays up to 800 steps), durations U[2, 30] s x duration_scale,
    per-instance a, b ~ U[0.9, 1.1] x defaults, TemporalAverage(2 ms) + Bold.
    Returns (plan, steps per instance, region-steps, info)."""
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
            make_bold_spec(dt, 2000.0, voi=(0, 1, 2, 3))]
    plan = Plan(model=L.MODEL_JR, cfun=L.CFUN_SIGMOID_JR, scheme=L.EULER_STO,
                precision=L.F32 if precision is None else precision,
                noise=L.NOISE_PHILOX, conns=conns, params=np.array(params), dt=dt, obs_coef=np.eye(6)[:4],
                monitors=mons, max_steps=max(steps) if max_steps is None else max_steps,
                inst_conn=np.arange(n_inst, dtype=np.int32), seed=seed, stream=stream)
    plan.set_steps(np.array(steps, np.int64))
    lo = [-1.0, -500.0, -50.0, -6.0, -20.0, -500.0]
    plan.init_uniform(seed + 1, [x * 0.01 for x in lo], [-x * 0.01 for x in lo])
    rs = float(sum(c.n_node * s for c, s in zip(conns, steps)))
    return plan, steps, rs, dict(instances=n_inst, duration_scale=duration_scale,
                                 total_nnz=int(sum(c.nnz for c in conns)),
                                 nodes_histogram={str(k): int(sum(c.n_node == k for c in conns))
                                                  for k in (68, 96, 192, 400, 998)},
                                 steps_min=int(min(steps)), steps_max=int(max(steps)))


def run_c3(stream=None, n_inst=1024, duration_scale=1.0, chunk=20000):
    """The uneven batch at its stated size: every instance runs its own
    duration (2-30 s); the plan advances in calls of ``chunk`` steps (each
    call's monitor samples stay on the device), instances stop at their
    budgets.  Timed from the first call to the end of the last."""
    t0 = time.perf_counter()
    plan, steps, rs, info = c3_plan(n_inst, duration_scale, stream=stream, max_steps=chunk)
    build_s = time.perf_counter() - t0
    ntile, k = plan.tiles()
    plan.sync()
    t0 = time.perf_counter()
    ms = []
    for _ in range((max(steps) + chunk - 1) // chunk):
        plan.run(chunk)
        ms.append(plan.last_run_ms())
        log("  c3 call %d: %.1f ms" % (len(ms), ms[-1]))
    plan.sync()
    el = time.perf_counter() - t0
    flags = plan.flags()
    shape = [plan.kernel, plan.tile, plan.waves]
    tb = plan.time_block()
    plan.close()
    info.update(value=rs / el, unit="region-steps/s", region_steps=rs, seconds=el, plan_build_s=build_s,
                calls=len(ms), call_ms=[round(m, 1) for m in ms], tiles=ntile, split_block_steps=k,
                time_block_steps=tb, kernel=shape, nonfinite=int((flags != 0).sum()))
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
           "precision": PREC_NAMES[prec], "kernel": [plan.kernel, plan.tile, plan.waves],
           "roofline": {"bound": "mfma", "achieved": tf, "peak": FP32_MFMA_PEAK_TFLOPS, "unit": "TFLOP/s",
                        "frac": tf / FP32_MFMA_PEAK_TFLOPS, "traffic": None,
                        "flops_per_region_step": 2.0 * 68 * 2, "median_launch_ms": med},
           "nonfinite_instances": int((plan.flags() != 0).sum())}
    plan.close()
    return out


def run_balloon_on_c2(ctx, chunk_steps=20000, B=8192):
    """Balloon-Windkessel BOLD (FP64, on the device) of one 20 000-step C2
    launch's TemporalAverage output: 1000 samples x 68 x 8192 series, timed
    with HIP events around the kernel, and its share of that launch."""
    import torch
    from tvb_hip.analyzers import BalloonModel, run_balloon
    T, n = chunk_steps // 20, 68 * B
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
    out = {"samples": T, "series": n, "ms": ms, "series_samples_per_s": T * n / (ms * 1e-3)}
    if ctx is not None and chunk_steps == 20000:
        out["share_of_launch"] = ms / ctx["avg_ms"]
    return out


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


def run_c4(n_requests=16, rate_hz=None, duration_scale=1.0, inst_range=(1, 512), seed=404, busy=0.75,
           chunk_steps=20000):
    """Serving: Poisson stream of sweep requests (G2D or MPR, HeunStochastic,
    N log-uniform in [68, 998], instances uniform in inst_range, durations
    U[5, 60] s x duration_scale), admitted by device footprint.  Without
    ``rate_hz`` the arrival rate is set from the device's serving capacity
    for this mix -- the same requests arriving together, served first as a
    burst (region-steps / makespan) -- so that the offered work keeps the
    GPU about ``busy`` occupied.  Each chunk's monitor samples stay on the
    device (copy-out is not in the latency)."""
    import torch
    from tvb_hip import _lib
    from tvb_hip.serve import Server, synthetic_requests
    streams = [_lib.stream_create() for _ in range(4)]
    free, total = torch.cuda.mem_get_info()

    def serve(reqs):
        srv = Server(hbm_budget=int(0.5 * free), streams=streams, chunk_steps=chunk_steps, keep_outputs=False,
                     log=log)
        return srv.run(reqs)

    kw = dict(seed=seed, inst_range=inst_range, duration_scale=duration_scale)
    calib = None
    if rate_hz is None:
        burst = synthetic_requests(n_requests, 1e9, **kw)       # the same requests, all at t = 0
        st0 = serve(burst)
        cap = st0["region_steps"] / st0["makespan_s"]
        mean_rs = float(np.mean([r.region_steps for r in burst]))
        rate_hz = busy * cap / mean_rs
        calib = {"capacity_region_steps_per_s": cap, "burst_makespan_s": st0["makespan_s"],
                 "burst_p50_latency_s": st0["p50_latency_s"], "burst_p99_latency_s": st0["p99_latency_s"]}
    reqs = synthetic_requests(n_requests, rate_hz, **kw)
    stats = serve(reqs)
    for st in streams:
        _lib.stream_destroy(st)
    mean_rs = float(np.mean([r.region_steps for r in reqs]))
    stats.update(duration_scale=duration_scale, rate_hz=rate_hz, inst_range=list(inst_range),
                 chunk_steps=chunk_steps,
                 requests=[dict(r.meta, rid=r.rid, latency_s=r.latency, region_steps=r.region_steps)
                           for r in reqs])
    if calib:
        stats.update(calib)
        stats["offered_load"] = rate_hz * mean_rs / calib["capacity_region_steps_per_s"]
        stats["arrival_rate_basis"] = ("%.2f x the serving capacity of this request mix (%.3g region-steps/s, the "
                                       "same requests served as a burst) over a mean request of %.3g region-steps"
                                       % (busy, calib["capacity_region_steps_per_s"], mean_rs))
    return stats


# ---------------------------------------------------------------------------

def _json_default(o):
    """numpy scalars and arrays in the line (bool_, int64, float64, ...)."""
    if isinstance(o, np.generic):
        return o.item()
    if isinstance(o, np.ndarray):
        return o.tolist()
    raise TypeError("not JSON serializable: %r" % type(o))


def guarded(name, fn, line):
    log("extra %s ..." % name)
    t0 = time.perf_counter()
    try:
        r = fn()
    except Exception as exc:   # recorded in the line, the headline is kept
        import traceback
        log(traceback.format_exc())
        r = {"error": "%s: %s" % (type(exc).__name__, exc)}
    if isinstance(r, dict):
        r.setdefault("extra_wall_s", time.perf_counter() - t0)
    line[name] = r
    log("extra %s done in %.1f s" % (name, time.perf_counter() - t0))
    return r


EXTRAS = ("other_precision", "all_channels", "fp64_single_instance", "error_vs_fp64_oracle", "gather_pattern_ceiling",
          "single_instance_c1", "uneven_batch_c3", "serving_c4", "zero_delay_mfma", "balloon_bold_fp64",
          "cpu_baseline", "cpu_baseline_all_cores")


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed launches (default 20: the sweep's full 20 s)")
    ap.add_argument("--warmup", type=int, default=2, help="untimed launches before the timed ones")
    ap.add_argument("--launch-steps", type=int, default=20000, help="integration steps per launch (one bench step)")
    ap.add_argument("--instances", type=int, default=8192)
    ap.add_argument("--tile", type=int, default=0)
    ap.add_argument("--waves", type=int, default=0)
    ap.add_argument("--cpu-steps", type=int, default=200000, help="oracle sample length (about 10-20 s of one core)")
    ap.add_argument("--no-extras", action="store_true", help="the headline only")
    ap.add_argument("--extras", default=",".join(EXTRAS), help="comma-separated subset of: " + ",".join(EXTRAS))
    ap.add_argument("--workload", default="c2", choices=("c2", "c1", "c2z", "c3", "c4", "copy"))
    ap.add_argument("--c3-scale", type=float, default=1.0,
                    help="duration factor of the uneven batch for --workload c3 (1: the stated 2-30 s)")
    ap.add_argument("--precision", default="mixed", choices=("mixed", "f32"))
    ap.add_argument("--no-tune", action="store_true", help="the built-in kernel shape instead of the autotuner's")
    ap.add_argument("--all-channels", action="store_true",
                    help="keep MontbrioPazoRoxin's V channel (coupling coefficient cv = 0) in the history ring")
    ap.add_argument("--oracle-child", nargs=2, help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.oracle_child:
        _oracle_error_child(*args.oracle_child)
        return
    if args.gpus != 1:
        print("bench.py runs on one GPU: every instance of the sweep is independent and the "
              "workload fits one MI355X; multi-GPU runs are not part of this build (--gpus %d)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)
    if args.steps < 1 or args.launch_steps < 1:
        sys.exit("--steps and --launch-steps must be positive")

    import torch
    from tvb_hip import _lib as L
    assert torch.cuda.is_available(), "no GPU"
    stream = torch.cuda.current_stream().cuda_stream
    if args.workload != "c2":
        fn = {"c1": lambda: run_c1(), "c3": lambda: run_c3(stream, duration_scale=args.c3_scale),
              "c2z": lambda: run_c2_zero_delay(args.instances, stream=stream,
                                               precision=L.MIXED if args.precision == "mixed" else L.F32),
              "copy": lambda: run_copy(), "c4": lambda: run_c4()}[args.workload]
        print(json.dumps({"workload": args.workload, **fn()}, default=_json_default), flush=True)
        return
    line, ctx = c2_headline(args, stream)
    log(json.dumps({k: line[k] for k in ("value", "ms_per_step", "roofline")}))
    extras = [] if args.no_extras else [e for e in args.extras.split(",") if e]
    prec = L.MIXED if args.precision == "mixed" else L.F32
    other = L.F32 if prec == L.MIXED else L.MIXED
    table = {
        "other_precision": lambda: c2_rate(ctx, other, args.launch_steps, stream=stream),
        "all_channels": lambda: c2_rate(ctx, prec, args.launch_steps, stream=stream, skip=False),
        "fp64_single_instance": lambda: fp64_single_instance(ctx),
        "error_vs_fp64_oracle": lambda: error_sample(ctx, ("f32", "mixed")),
        "gather_pattern_ceiling": lambda: gather_pattern_ceiling(line),
        "single_instance_c1": lambda: run_c1(),
        "uneven_batch_c3": lambda: run_c3(stream),
        "serving_c4": lambda: run_c4(),
        "zero_delay_mfma": lambda: run_c2_zero_delay(args.instances, stream=stream, precision=prec),
        "balloon_bold_fp64": lambda: run_balloon_on_c2(ctx),
        "cpu_baseline": lambda: cpu_baseline_dict(ctx, args.cpu_steps),
        "cpu_baseline_all_cores": lambda: cpu_baseline_all_cores(ctx["w"], ctx["Lmat"], args.cpu_steps // 4),
    }
    for name in extras:
        if name not in table:
            sys.exit("unknown extra %r" % name)
        guarded(name, table[name], line)
    # summary fields of the extras
    f64 = line.get("fp64_single_instance", {})
    if "value" in f64:
        line["batched_over_fp64_single"] = f64["batched_over_fp64_single"]
    err = line.get("error_vs_fp64_oracle", {})
    if args.precision in err:
        line["meets_error_bar"] = {p: err[p]["meets_error_bar"] for p in ("f32", "mixed") if p in err}
    srv = line.get("serving_c4", {})
    if "p50_latency_s" in srv:
        line["p50_latency_s"], line["p99_latency_s"] = srv["p50_latency_s"], srv["p99_latency_s"]
    print(json.dumps(line, default=_json_default), flush=True)


def cpu_baseline_dict(ctx, n_steps):
    rate, el = cpu_baseline(ctx["w"], ctx["Lmat"], 1.0, 1e-3, n_steps)
    return {"value": rate, "unit": "region-steps/s", "cores": 1, "kind": "port",
            "sample": "1 instance x %d steps (%.1f s simulated) of the c2 workload, reference-faithful numpy "
                      "oracle, 1 thread, %.1f s wall, %s" % (n_steps, n_steps * 0.05 / 1000.0, el, cpu_model())}


if __name__ == "__main__":
    main()
