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


def tavg_error_sample(w, L, d, G, sig, n_steps=4000, B=32, check=(0, 17)):
    """FP32 batched kernel with injected noise vs the FP64 oracle on a
    sample of the sweep: max |err| / max |ref| of TemporalAverage."""
    from tvb_hip import _lib as Lb
    from oracle.tvb_oracle import OracleSim
    idx = np.linspace(0, len(G) - 1, B).astype(int)
    Gs, ss = G[idx], sig[idx]
    plan = build_c2_plan(w, d, Gs, ss, 0.05, n_steps, Lb.F32, Lb.NOISE_INJECTED, 0, bold=0.0, tile=32)
    H = int(d.max()) + 1
    rng = np.random.default_rng(5)
    ic = np.empty((H, 2, w.shape[0]))
    ic[:, 0] = rng.uniform(0.0, 2.0, (H, w.shape[0]))
    ic[:, 1] = rng.uniform(-2.0, 1.5, (H, w.shape[0]))
    plan.set_history(-1, ic, ic[-1], H - 1)
    z = rng.standard_normal((n_steps, 2, w.shape[0], B))
    plan.set_noise(z)
    plan.run(n_steps)
    got = plan.read(0)
    errs = []
    for b in check:
        o = OracleSim(w, L, 5.0, 0.05, "MontbrioPazoRoxin", coupling_params=dict(a=float(Gs[b])),
                      integrator="HeunStochastic", nsig=np.array([ss[b], ss[b]]), initial_conditions=ic,
                      monitors=(("TemporalAverage", 1.0, None),), mode="fp64")
        (t, x), = o.run(n_steps, noise=z[..., b])
        ref = x[..., 0]
        dev = got[b][1]
        errs.append(float(np.max(np.abs(dev - ref)) / np.max(np.abs(ref))))
    plan.close()
    return max(errs)


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
    ap.add_argument("--cpu-steps", type=int, default=40000)
    ap.add_argument("--no-extras", action="store_true", help="skip CPU baseline, FP64 single path and error sample")
    ap.add_argument("--workload", default="c2", choices=("c2", "c3", "c4"))
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
    if args.workload == "c3":
        print(json.dumps({"metric": "region-steps/s (uneven JansenRit batch)", "workload": "c3",
                          **run_c3(stream)}), flush=True)
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
    plan = build_c2_plan(w, d, G, sig, dt, max(chunk, args.warmup), L.F32, L.NOISE_PHILOX, 1234,
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
    bpr = algorithmic_bytes_per_region_step(nnz, 68, 2, 4, 2, 20, 2, 80, 5000, 40000)
    per_launch = bpr * 68 * B * full[0][0]
    achieved = per_launch / (avg_ms * 1e-3) / 1e9
    line = {
        "metric": "region-steps/s batched FP32 vs plain FP64 at stated error; p50/p99 latency",
        "value": value, "unit": "region-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": el * 1e3 / max(args.steps, 1), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "c2: MPR HeunStochastic dt=0.05 sweep, 68 nodes x %d instances, shared "
                               "25%%-density connectome (nnz %d, horizon %d), TemporalAverage 1 ms + Bold 2000 ms, "
                               "device Philox noise" % (B, nnz, H),
                   "instances": B, "nodes": 68, "kernel_tile": plan.tile, "kernel_waves": plan.waves,
                   "steps_per_launch": chunk},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "algorithmic_bytes_per_region_step": bpr, "avg_launch_ms": avg_ms},
        "nonfinite_instances": int((flags != 0).sum()),
    }
    plan.close()
    if not args.no_extras:
        v64, el64, shape64 = fp64_single_instance(w, d, 1.0, 1e-3, 40000)
        line["fp64_single_instance"] = {"value": v64, "unit": "region-steps/s", "steps": 40000,
                                        "kernel": list(shape64)}
        line["batched_fp32_over_fp64_single"] = value / v64
        line["tavg_max_rel_err_vs_fp64_oracle"] = tavg_error_sample(w, Lmat, d, G, sig)
        line["serving"] = run_c4()
        line["p50_latency_s"] = line["serving"]["p50_latency_s"]
        line["p99_latency_s"] = line["serving"]["p99_latency_s"]
        line["uneven_batch"] = run_c3(stream)
        cpu_rate, cpu_el = cpu_baseline(w, Lmat, 1.0, 1e-3, args.cpu_steps)
        line["cpu_baseline"] = {"value": cpu_rate, "unit": "region-steps/s", "cores": 1, "kind": "port",
                                "sample": "1 instance x %d steps (%.1f s simulated) of the c2 workload, "
                                          "reference-faithful numpy oracle, 1 thread, %.1f s wall"
                                          % (args.cpu_steps, args.cpu_steps * dt / 1000.0, cpu_el)}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
