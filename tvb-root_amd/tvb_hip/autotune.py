"""Kernel-shape autotuner for the batch-lane step kernel.

For a workload signature (model, coupling, scheme, precision, node count,
nnz, horizon, history channels, instance count, monitors) it times short
trial launches of each candidate (tile, waves) shape with HIP events on a
throw-away plan, keeps the fastest, and caches the choice in a JSON file
keyed by the signature and the device name, so later plans of the same
shape skip the search.  Candidates that do not fit (LDS, registers) are
skipped.
"""
import json
import os

import numpy as np

DEFAULT_CANDIDATES = ((32, 16), (32, 8), (16, 16), (16, 8), (64, 16), (64, 8))
CACHE_PATH = os.environ.get("TVBK_TUNE_CACHE",
                            os.path.join(os.path.expanduser("~"), ".cache", "tvb_hip_tune.json"))


def signature(kw, device=""):
    c = kw["conns"][0]
    mons = [(m.kind, m.istep, len(m.voi)) for m in kw["monitors"]]
    return json.dumps(dict(dev=device, model=kw["model"], cfun=kw["cfun"], scheme=kw["scheme"],
                           prec=kw["precision"], noise=kw["noise"], n=c.n_node, nnz=c.nnz, h=c.horizon,
                           b=int(np.asarray(kw["params"]).shape[0]), mons=mons, flags=int(kw.get("flags", 0))),
                      sort_keys=True)


def device_name():
    """Name of the current GPU (part of the cache key)."""
    try:
        import torch
        return torch.cuda.get_device_name(0)
    except Exception:
        return ""


def load_cache(path=CACHE_PATH):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def save_cache(cache, path=CACHE_PATH):
    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(cache, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def tune(plan_kwargs, trial_steps=2000, candidates=DEFAULT_CANDIDATES, device="", cache_path=None,
         make_plan=None, init=None, use_cache=True):
    """Best (tile, waves) for the batch-lane kernel of this workload.

    plan_kwargs: keyword arguments of tvb_hip.plan.Plan (without tile, waves,
    max_steps, kernel).  make_plan(tile, waves) may be given to replace plan
    construction (tests); init(plan) prepares a trial plan's history.
    Returns (tile, waves, table) with table = {(tile, waves): ms per step}."""
    from . import _lib as L
    from .plan import Plan
    plan_kwargs = {k: v for k, v in plan_kwargs.items() if k not in ("tile", "waves", "max_steps", "kernel", "tune")}
    cache_path = CACHE_PATH if cache_path is None else cache_path
    key = signature(plan_kwargs, device)
    cache = load_cache(cache_path) if use_cache else {}
    if use_cache and key in cache:
        t, w = cache[key]["best"]
        return t, w, {tuple(k): v for k, v in cache[key]["table"]}
    if make_plan is None:
        def make_plan(tile, waves):
            return Plan(max_steps=trial_steps, kernel=L.KERNEL_BATCH_LANE, tile=tile, waves=waves, **plan_kwargs)
    table = {}
    for tile, waves in candidates:
        try:
            p = make_plan(tile, waves)
        except L.TvbkError:
            continue
        try:
            if init is not None:
                init(p)
            p.run(min(200, trial_steps))          # warm caches, first-launch costs
            p.sync()
            p.run(trial_steps)
            table[(tile, waves)] = p.last_run_ms() / trial_steps
        finally:
            p.close()
    if not table:
        raise RuntimeError("no candidate kernel shape fits this workload")
    best = min(table, key=table.get)
    if use_cache:
        cache[key] = {"best": list(best), "table": [[list(k), v] for k, v in table.items()]}
        try:
            save_cache(cache, cache_path)
        except OSError:          # a read-only home: the choice is still returned
            pass
    return best[0], best[1], table
