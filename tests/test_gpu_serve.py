"""GPU test of the request loop: requests that arrive while others run are
admitted and overlap on the device, every request completes, and each
request's result is the same alone and in the mix (bit for bit: its noise
is keyed by its own seed and instance, and its state lives in its own
device plan)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _streams(k):
    from tvb_hip import _lib
    ss = [_lib.stream_create() for _ in range(k)]
    return ss, ss


def test_serving_mix_equals_alone():
    print("serving mix", flush=True)
    from tvb_hip.serve import Server, _make_request
    specs = [(0, 0.000, 40, 12, "MPR", 150.0), (1, 0.002, 25, 3, "G2D", 300.0),
             (2, 0.004, 60, 40, "MPR", 100.0), (3, 0.006, 33, 1, "G2D", 250.0),
             (4, 0.008, 48, 20, "MPR", 120.0)]
    keep, streams = _streams(3)

    def make(spec):
        rid, t, n, B, model, dur = spec
        return _make_request(rid, t, n, B, model, dur, cseed=100 + rid, nseed=7000 + rid, precision="fp32")
    mix = [make(s) for s in specs]
    stats = Server(hbm_budget=8 << 30, streams=streams, chunk_steps=500).run(mix)
    assert stats["n_requests"] == 5 and stats["max_concurrent"] >= 2
    assert stats["p99_latency_s"] >= stats["p50_latency_s"] > 0
    for k in (0, 2, 3):
        alone = make(specs[k])
        alone.arrival = 0.0
        Server(hbm_budget=8 << 30, streams=streams[:1], chunk_steps=700).run([alone])
        a, b = alone.outputs, mix[k].outputs
        assert len(a) == len(b) == specs[k][3]
        for inst in range(len(a)):
            np.testing.assert_array_equal(a[inst][0][1], b[inst][0][1])
            np.testing.assert_array_equal(a[inst][0][0], b[inst][0][0])
        assert np.all(mix[k].flags == 0)
