"""Continuous-batching request loop for RGF work on one GPU.

Each request asks for factorize + logdet + solve(1 RHS) of its own matrix
(block structure ns, nt, nd; values from the device generator with a per-request
seed), or for a follow-up solve on a matrix factorized earlier.  Requests arrive
over time (Poisson in the benchmark).  The scheduler admits a waiting request as
soon as a worker slot is free and its device footprint fits the HBM budget; every
admitted request runs on its own handle, i.e. its own HIP stream, so the work of
concurrent requests overlaps on the device.  Finished factors stay resident (LRU)
for follow-up solves and are evicted when the budget is needed.

The scheduler (admission, ordering, eviction, latency accounting) is independent
of the executor; tests drive it with a CPU executor, the benchmark with the GPU one.
Give the process at least one hardware queue per slot (GPU_MAX_HW_QUEUES, set before the
HIP runtime starts; HIP's default is 4): streams that share a hardware queue run their
kernels one after another.
"""
from __future__ import annotations

import collections
import dataclasses
import threading
import time
from typing import Callable, Optional

import numpy as np


@dataclasses.dataclass
class Request:
    rid: int
    arrival: float                 # seconds from the start of the run
    ns: int
    nt: int
    nd: int
    seed: int
    followup_of: Optional[int] = None   # solve again with the factor of request `followup_of`
    # filled by the server
    start: float = float("nan")
    finish: float = float("nan")
    logdet: float = float("nan")
    x: Optional[np.ndarray] = None
    error: Optional[str] = None
    reused_factor: bool = False

    @property
    def latency(self) -> float:
        return self.finish - self.arrival


def poisson_workload(n: int, rate: float, seed: int = 0, ns_range=(300, 3000), nt_range=(10, 120),
                     nd_range=(0, 8), followup_frac: float = 0.0) -> list[Request]:
    """Requests with exponential inter-arrival times (rate per second) and shapes
    ns ~ U{ns_range}, nt ~ U{nt_range}, nd ~ U{nd_range} (nd clipped to ns)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(n):
        t += rng.exponential(1.0 / rate)
        ns = int(rng.integers(ns_range[0], ns_range[1] + 1))
        nt = int(rng.integers(nt_range[0], nt_range[1] + 1))
        nd = min(int(rng.integers(nd_range[0], nd_range[1] + 1)), ns, 64)
        mseed = 1000003 * seed + i
        fu = None
        if followup_frac > 0 and i > 0 and rng.random() < followup_frac:
            fu = int(rng.integers(0, i))
            o = out[fu]
            ns, nt, nd, mseed = o.ns, o.nt, o.nd, o.seed   # same matrix as the original request
        out.append(Request(i, t, ns, nt, nd, seed=mseed, followup_of=fu))
    return out


class Executor:
    """Interface of the work performed for one request."""

    def footprint(self, req: Request) -> int:
        raise NotImplementedError

    def run(self, req: Request, cached) -> object:
        """Do the request's work; `cached` is a resident factor handle (follow-up) or
        None.  Returns the handle to keep resident (or None)."""
        raise NotImplementedError

    def release(self, handle) -> None:
        pass


class GPUExecutor(Executor):
    """librgf_hip.so: one handle (one stream) per request."""

    def __init__(self, precision: int, refine_tol: float = 1e-12, device_cache_gb: Optional[float] = None,
                 cholesky: Optional[int] = None):
        """device_cache_gb: idle device memory the library's arena may keep for reuse
        (None: the library's default).  A serving process sets it high: handles come and
        go continuously, and giving slabs back calls hipFree, which waits for the device.
        cholesky: the handles' Cholesky schedule (None: tile launches -- with many requests in
        flight the device is shared, and the task graph's single-matrix latency gain would come
        at the price of workgroups that wait on one matrix's chain while others could run)."""
        import rgf_amd
        self.rgf = rgf_amd
        if device_cache_gb is not None:
            rgf_amd.set_device_cache_limit(int(device_cache_gb * (1 << 30)))
        self.precision = precision
        self.keep = 1 if precision != rgf_amd.RGF_FP64 else 0   # refinement needs A
        self.refine_tol = refine_tol
        self.opts = rgf_amd.default_options(
            precision=precision, keep_matrix=self.keep, refine_tol=refine_tol,
            cholesky=rgf_amd.RGF_CHOLESKY_TILE_LAUNCHES if cholesky is None else cholesky)

    def footprint(self, req: Request) -> int:
        return self.rgf.footprint_bytes(req.ns, req.nt, req.nd, 1, self.opts)

    def rhs(self, req: Request) -> np.ndarray:
        n = req.ns * req.nt + req.nd
        return np.random.default_rng([req.seed, req.rid]).standard_normal(n)

    def run(self, req: Request, cached):
        rgf = self.rgf
        if cached is not None:
            h = cached
            req.reused_factor = True
        else:
            h = rgf.RGF.generated(req.ns, req.nt, req.nd, seed=req.seed, options=self.opts)
            h.factorize()
        req.logdet = h.logDet()
        req.x = h.solve(None, self.rhs(req), 1)
        return h

    def release(self, handle) -> None:
        handle.close()


class Server:
    """Continuous-batching scheduler over `slots` worker threads and an HBM budget."""

    def __init__(self, executor: Executor, slots: int = 8, budget_bytes: int = 160 << 30,
                 clock: Callable[[], float] = time.perf_counter,
                 progress: Optional[Callable[[int, int], None]] = None, progress_every: int = 100):
        self.progress = progress              # called as progress(done, total) every progress_every
        self.progress_every = progress_every
        self.total = 0
        self.ex = executor
        self.slots = slots
        self.budget = budget_bytes
        self.clock = clock
        self.lock = threading.Condition()
        self.in_use = 0                        # bytes of running requests + resident factors
        self.running = 0
        self.queue: collections.deque[Request] = collections.deque()
        self.resident: "collections.OrderedDict[int, tuple[object, int]]" = collections.OrderedDict()
        self.done: list[Request] = []
        self.max_concurrent = 0
        self.t0 = 0.0
        self._stop = False

    # -- admission ------------------------------------------------------------
    def _evict_for(self, need: int) -> bool:
        """Evict least-recently-used resident factors until `need` bytes fit."""
        while self.in_use + need > self.budget and self.resident:
            rid, (h, nb) = self.resident.popitem(last=False)
            self.in_use -= nb
            self.ex.release(h)
        return self.in_use + need <= self.budget

    def _pick(self) -> Optional[tuple[Request, object, int]]:
        """First waiting request (arrival order) that fits a slot and the budget."""
        if self.running >= self.slots:
            return None
        for i, req in enumerate(self.queue):
            cached = None
            if req.followup_of is not None and req.followup_of in self.resident:
                cached, nb = self.resident.pop(req.followup_of)
                self.in_use -= nb        # re-charged below as the running request's bytes
                need = nb
            else:
                need = self.ex.footprint(req)
            if self._evict_for(need):
                del self.queue[i]
                self.in_use += need
                self.running += 1
                self.max_concurrent = max(self.max_concurrent, self.running)
                return req, cached, need
            if cached is not None:       # put it back, nothing fits yet
                self.resident[req.followup_of] = (cached, need)
                self.in_use += need
            if i == 0 and need > self.budget:
                req.error = f"footprint {need} exceeds the budget {self.budget}"
                del self.queue[0]
                req.finish = req.start = self.clock() - self.t0
                self.done.append(req)
                self.lock.notify_all()
                return self._pick()
        return None

    def _worker(self) -> None:
        while True:
            with self.lock:
                item = None
                while item is None:
                    if self._stop and not self.queue:
                        return
                    item = self._pick()
                    if item is None:
                        self.lock.wait(0.05)
                req, cached, need = item
                req.start = self.clock() - self.t0
            handle = None
            try:
                handle = self.ex.run(req, cached)
            except Exception as e:   # a failing request fails alone
                req.error = f"{type(e).__name__}: {e}"
            with self.lock:
                req.finish = self.clock() - self.t0
                self.running -= 1
                if handle is not None and req.error is None:
                    self.resident[req.rid if cached is None else req.followup_of] = (handle, need)
                else:
                    self.in_use -= need
                    if handle is not None:
                        self.ex.release(handle)
                self.done.append(req)
                if self.progress and len(self.done) % self.progress_every == 0:
                    self.progress(len(self.done), self.total)
                self.lock.notify_all()

    def run(self, requests: list[Request]) -> dict:
        """Feed requests at their arrival times; returns throughput and latency stats."""
        self.t0 = self.clock()
        self.total = len(requests)
        workers = [threading.Thread(target=self._worker, daemon=True) for _ in range(self.slots)]
        for w in workers:
            w.start()
        for req in sorted(requests, key=lambda r: r.arrival):
            delay = req.arrival - (self.clock() - self.t0)
            if delay > 0:
                time.sleep(delay)
            with self.lock:
                self.queue.append(req)
                self.lock.notify_all()
        with self.lock:
            self._stop = True
            self.lock.notify_all()
        for w in workers:
            w.join()
        makespan = self.clock() - self.t0
        with self.lock:
            for h, nb in self.resident.values():
                self.ex.release(h)
                self.in_use -= nb
            self.resident.clear()
        ok = [r for r in self.done if r.error is None]
        lat = np.array([r.latency for r in ok]) if ok else np.array([float("nan")])
        return {
            "requests": len(requests), "completed": len(ok), "failed": len(self.done) - len(ok),
            "makespan_s": makespan, "req_per_s": len(ok) / makespan if makespan > 0 else 0.0,
            "p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "mean_ms": float(lat.mean() * 1e3), "max_concurrent": self.max_concurrent,
            "factor_reuses": sum(1 for r in ok if r.reused_factor),
        }
