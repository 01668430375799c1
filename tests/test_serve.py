"""Serving loop: scheduler behaviour on the CPU (fake executor) and, on the GPU, that each
request's result is the same alone and in the mix."""
import threading
import time

import numpy as np
import pytest

from rgf_amd import serve


class FakeExecutor(serve.Executor):
    """Sleeps in proportion to the matrix size; tracks peak bytes and concurrency."""

    def __init__(self, budget):
        self.lock = threading.Lock()
        self.live = 0
        self.peak = 0
        self.budget = budget
        self.released = 0

    def footprint(self, req):
        return req.ns * req.nt

    def run(self, req, cached):
        req.reused_factor = cached is not None
        with self.lock:
            self.live += 1
            self.peak = max(self.peak, self.live)
        time.sleep(0.002 + req.ns * req.nt * 1e-7)
        req.logdet = float(req.ns * req.nt + req.nd)
        req.x = np.full(3, req.seed, dtype=float)
        with self.lock:
            self.live -= 1
        return ("handle", req.rid)

    def release(self, h):
        with self.lock:
            self.released += 1


def test_poisson_workload_shapes():
    reqs = serve.poisson_workload(500, rate=100.0, seed=3, followup_frac=0.2)
    t = np.array([r.arrival for r in reqs])
    assert np.all(np.diff(t) > 0)
    assert 3.0 < t[-1] < 7.5          # 500 arrivals at 100/s
    assert all(300 <= r.ns <= 3000 and 10 <= r.nt <= 120 and 0 <= r.nd <= 8 for r in reqs)
    fu = [r for r in reqs if r.followup_of is not None]
    assert fu and all((r.ns, r.nt, r.nd, r.seed) == (reqs[r.followup_of].ns, reqs[r.followup_of].nt,
                                                     reqs[r.followup_of].nd, reqs[r.followup_of].seed) for r in fu)


def test_scheduler_admission_budget_and_reuse():
    budget = 3_000_000
    ex = FakeExecutor(budget)
    srv = serve.Server(ex, slots=4, budget_bytes=budget)
    reqs = serve.poisson_workload(120, rate=2000.0, seed=1, ns_range=(300, 800), nt_range=(10, 120),
                                  followup_frac=0.3)
    stats = srv.run(reqs)
    assert stats["completed"] == 120 and stats["failed"] == 0
    assert srv.max_concurrent <= 4 and ex.peak <= 4
    assert srv.max_concurrent >= 2            # requests overlapped
    assert stats["factor_reuses"] > 0         # follow-ups found resident factors
    assert srv.in_use == 0                    # everything released at the end
    assert all(r.finish >= r.start >= r.arrival - 1e-6 for r in reqs)
    assert stats["p99_ms"] >= stats["p50_ms"] > 0


def test_scheduler_never_exceeds_budget_and_rejects_oversize():
    budget = 50_000
    ex = FakeExecutor(budget)
    srv = serve.Server(ex, slots=8, budget_bytes=budget)
    peak = {"v": 0}
    orig = srv._pick

    def pick():
        r = orig()
        peak["v"] = max(peak["v"], srv.in_use)
        return r
    srv._pick = pick
    reqs = [serve.Request(i, 0.001 * i, ns=300 + 10 * i, nt=20, nd=2, seed=i) for i in range(30)]
    reqs.append(serve.Request(30, 0.05, ns=3000, nt=100, nd=0, seed=30))   # 300k > budget
    stats = srv.run(reqs)
    assert peak["v"] <= budget
    assert stats["failed"] == 1 and reqs[30].error and "exceeds" in reqs[30].error
    assert stats["completed"] == 30


@pytest.mark.gpu
def test_results_same_alone_and_in_the_mix(gpu):
    import rgf_amd
    for prec in (rgf_amd.RGF_FP64, rgf_amd.RGF_MIXED_BF16X3):
        print(f"serve mix vs alone precision={prec}", flush=True)
        ex = serve.GPUExecutor(prec)
        reqs = serve.poisson_workload(10, rate=500.0, seed=7, ns_range=(60, 260), nt_range=(2, 6), nd_range=(0, 5))
        srv = serve.Server(ex, slots=5, budget_bytes=8 << 30)
        stats = srv.run(reqs)
        assert stats["completed"] == 10 and srv.max_concurrent >= 2
        for r in reqs:
            alone = serve.Request(r.rid, 0.0, r.ns, r.nt, r.nd, r.seed)
            h = ex.run(alone, None)
            ex.release(h)
            assert alone.logdet == r.logdet
            assert np.array_equal(alone.x, r.x)
