"""The allocation-free history upload (tvbk_set_history_noalloc) against
tvbk_set_history through the C-ABI: two plans built alike, one set with
each entry point, must hold bit-identical rings and states and produce
bit-identical monitor outputs; two calls in a row, the second with a
smaller input (so the static scratch holds stale bytes past it), and a
history larger than the scratch is refused before any device work."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _plans(n, B, seed, lengths=(5.0, 40.0)):
    from tvb_hip import _lib as L
    from tvb_hip.connectome import idelays, synthetic
    from tvb_hip.plan import CSR, MonitorSpec, Plan
    w, Lm = synthetic(n, 0.3, lengths[0], lengths[1], seed)
    d, H = idelays(Lm, 3.0, 0.05)
    rng = np.random.default_rng(seed)
    p = np.empty((B, 12))
    p[:, :8] = [1.0, 0.0, 1.0, 15.0, -5.0, 0.0, 1.0, 0.0]
    p[:, 8] = rng.uniform(0.0, 1.0, B)
    p[:, 9] = 0.0
    p[:, 10:] = 1e-3
    kw = dict(model=L.MODEL_MPR, cfun=L.CFUN_LINEAR, scheme=L.HEUN_STO, precision=L.MIXED,
              noise=L.NOISE_INJECTED, conns=[CSR(w, d)], params=p, dt=0.05, obs_coef=np.eye(2),
              monitors=[MonitorSpec(L.MON_TAVG, 20, [0, 1], period=1.0)], max_steps=400, seed=0)
    hist = np.empty((B, H, 2, n))
    hist[:, :, 0] = rng.uniform(0.0, 2.0, (B, H, n))
    hist[:, :, 1] = rng.uniform(-2.0, 1.5, (B, H, n))
    z = rng.standard_normal((400, 2, n, B))
    return Plan(**kw), Plan(**kw), hist, z, H


def _check_pair(a, b, hist, z, H, shared):
    B = hist.shape[0]
    if shared:     # every instance from the same arrays (inst = -1)
        a.set_history(-1, hist[0], hist[0][-1], H - 1)
        b.set_history(-1, hist[0], hist[0][-1], H - 1, noalloc=True)
    else:
        for k in range(B):
            a.set_history(k, hist[k], hist[k][-1], H - 1)
            b.set_history(k, hist[k], hist[k][-1], H - 1, noalloc=True)
    for k in range(B):
        ha, sa, ca = a.export(k, np.float32)
        hb, sb, cb = b.export(k, np.float32)
        np.testing.assert_array_equal(ha, hb)
        np.testing.assert_array_equal(sa, sb)
        assert ca == cb == H - 1
    for p in (a, b):
        p.set_noise(z)
        p.run(z.shape[0])
    ta, tb = a.read(0), b.read(0)
    for (t1, x1), (t2, x2) in zip(ta, tb):
        np.testing.assert_array_equal(t1, t2)
        np.testing.assert_array_equal(x1, x2)
    np.testing.assert_array_equal(a.flags(), b.flags())


def test_noalloc_history_equals_allocating_upload_twice():
    _say("noalloc call 1: N=48, 40 instances, batch-lane tiles, per-instance histories")
    a, b, hist, z, H = _plans(48, 40, 11, lengths=(5.0, 60.0))
    try:
        _check_pair(a, b, hist, z, H, shared=False)
    finally:
        a.close()
        b.close()
    _say("noalloc call 2: N=12, 3 instances (smaller input), one history for all")
    a, b, hist, z, H = _plans(12, 3, 12, lengths=(5.0, 20.0))
    try:
        _check_pair(a, b, hist, z, H, shared=True)
    finally:
        a.close()
        b.close()


def test_noalloc_refuses_a_history_larger_than_the_scratch():
    from tvb_hip import _lib as L
    from tvb_hip.plan import CSR, MonitorSpec, Plan
    n, Hd = 3000, 1499                       # ring 1500 x 2 x 3000 float32 = 36 MB > 32 MiB
    w = np.zeros((n, n))
    w[0, 1] = 1.0
    d = np.zeros((n, n), np.int32)
    d[0, 1] = Hd
    p = np.array([[1.0, 0.0, 1.0, 15.0, -5.0, 0.0, 1.0, 0.0, 0.1, 0.0]])
    plan = Plan(model=L.MODEL_MPR, cfun=L.CFUN_LINEAR, scheme=L.HEUN_DET, precision=L.MIXED, noise=L.NOISE_NONE,
                conns=[CSR(w, d)], params=p, dt=0.05, obs_coef=np.eye(2),
                monitors=[MonitorSpec(L.MON_TAVG, 20, [0], period=1.0)], max_steps=20, seed=0)
    try:
        hist = np.zeros((Hd + 1, 2, n), np.float32)
        assert hist.nbytes > L.NOALLOC_MAX_BYTES
        with pytest.raises(L.TvbkError, match="static scratch"):
            plan.set_history(0, hist, np.zeros((2, n)), Hd, noalloc=True)
        plan.set_history(0, hist, np.zeros((2, n)), Hd)          # the allocating upload takes it
    finally:
        plan.close()
