"""Host-side tests of the request loop with stand-in plans (no GPU):
admission never exceeds the HBM budget, requests are admitted in arrival
order as memory frees up, several run concurrently, every request
finishes with its latency recorded, and an oversized request is refused."""
import numpy as np
import pytest

from tvb_hip.serve import Request, Server


class FakePlan:
    """Stands in for a Plan: a launch completes after `lag` polls."""

    def __init__(self, n_inst, lag, log):
        self.n_inst, self.lag, self.log = n_inst, lag, log
        self.monitors = [object()]
        self.polls = 0
        self.running = False
        self.steps = 0

    def run(self, n):
        assert not self.running
        self.running = True
        self.polls = 0
        self.steps += n
        self.log.append(("run", id(self), n))

    def done(self):
        self.polls += 1
        if self.polls >= self.lag:
            self.running = False
            return True
        return False

    def read(self, m):
        return [(np.arange(2.0), np.full((2, 1, 3), float(self.steps)))] * self.n_inst

    def flags(self):
        return np.zeros(self.n_inst, np.int32)

    def close(self):
        self.log.append(("close", id(self)))


def _req(rid, arrival, steps, nbytes, log, lag=2):
    return Request(rid, arrival, steps, build=lambda stream, ms: FakePlan(1, lag, log),
                   footprint=lambda ms: nbytes, region_steps=10.0 * steps)


def test_admission_respects_budget_and_order():
    log = []
    t = [0.0]

    def clock():
        t[0] += 1e-3
        return t[0]
    reqs = [_req(k, arrival=0.001 * k, steps=50 + 10 * k, nbytes=40, log=log) for k in range(8)]
    srv = Server(hbm_budget=100, streams=[None, None, None], chunk_steps=20, clock=clock, idle_sleep=0)
    stats = srv.run(reqs)
    assert stats["n_requests"] == 8
    assert srv.peak_bytes <= 100
    assert srv.max_concurrent == 2               # 2 x 40 bytes fit, a third does not
    admits = [r.admitted for r in sorted(reqs, key=lambda r: r.rid)]
    assert admits == sorted(admits)              # arrival order
    assert all(r.latency is not None and r.latency > 0 for r in reqs)
    for r in reqs:                               # every chunk of every request ran
        assert r.outputs[0][0][1][-1, 0, 0] == r.n_steps
    assert stats["p99_latency_s"] >= stats["p50_latency_s"]


def test_late_arrivals_overlap_running_requests():
    log = []
    t = [0.0]

    def clock():
        t[0] += 1e-3
        return t[0]
    long_req = _req(0, 0.0, 4000, 10, log, lag=3)
    late = [_req(k, 0.003 + 0.001 * k, 200, 10, log) for k in range(1, 5)]
    srv = Server(hbm_budget=1000, streams=[None] * 4, chunk_steps=40, clock=clock, idle_sleep=0)
    srv.run([long_req] + late)
    # the late requests were admitted and finished while the long one still ran
    assert all(r.admitted < long_req.finished for r in late)
    assert all(r.finished < long_req.finished for r in late)
    assert srv.max_concurrent >= 3


def test_oversized_request_is_refused():
    srv = Server(hbm_budget=10, streams=[None])
    with pytest.raises(ValueError):
        srv.run([_req(0, 0.0, 10, 11, [])])
