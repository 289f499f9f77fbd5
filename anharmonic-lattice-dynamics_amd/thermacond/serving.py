"""A device-resident kappa-solve server with continuous batching and HBM-footprint admission.

Requests (a crystal with its force constants, a q-mesh, temperatures, V3 precision, delta
integration) arrive at any time.  The scheduler admits them in arrival order as soon as a
worker slot is free and their device footprint (alt3_footprint, computed on a host-only
context, so before any device memory is touched) fits in what the budget has left; a
request that does not fit waits while later, smaller ones may pass it (first fit), but no
request is passed more than `max_bypass` times.  Admitted requests run concurrently, one
host thread per slot, each on its own context and HIP stream, so their kernels overlap on
the device.

Contexts are kept device-resident across requests: a finished request leaves its context
(force constants, eigenvectors, |V3|^2 and scattering weights) in a cache keyed by crystal
and mesh, and a later request with the same key skips the eigen solve, V3 and FGR
(alt3_solve reuses what the context holds) and pays only the temperature-dependent part.
Cached contexts count against the budget and are evicted, least recently used first, when
an admission needs their memory.  Requests with the same key are serialised on that
context.

The |V3|^2 stage of admitted requests is batched: a worker whose context needs |V3|^2 (6 or
12 branches; larger cells fill the device on their own) queues it with the server's V3
batcher and waits; the batcher takes everything queued for one branch count and precision
and runs it as one shared persistent launch (capi.v3_batch), while requests that arrive
meanwhile queue for the next launch.  The rest of each solve (FGR, tau, kappa, CG) runs on
the request's own stream.

Every result is a pure function of its request: a context computes the same kappa whether
it runs alone or beside others, and a batched launch computes each context's |V3|^2 with the
same code as a lone one (tests/test_serving.py checks both on the device).
"""
from __future__ import annotations

import collections
import dataclasses
import itertools
import threading
import time
from typing import Callable, Dict, List, Optional, Sequence

import numpy as np

from . import capi
from .refio import InputSet


@dataclasses.dataclass
class Request:
    key: str                          # identifies crystal + force constants (+ mesh below)
    inputs: InputSet
    temps: Sequence[float]
    nc: Optional[Sequence[int]] = None
    precision: int = capi.V3_FP64
    delta: int = capi.DELTA_TETRA
    sigma: float = 0.0
    # filled by the server
    id: int = -1
    t_arrive: float = 0.0
    t_admit: float = 0.0
    t_done: float = 0.0
    footprint: int = 0
    nb: int = 0                       # branches of the cell (from the host-only footprint context)
    reused: bool = False
    result: Optional[dict] = None
    error: Optional[str] = None
    # the admission decision: the cached context this request was admitted on (its busy count
    # raised, its reserve grown), or None when it was charged its own footprint
    _entry: Optional["_Cached"] = dataclasses.field(default=None, repr=False, compare=False)
    # admitted with a place expected in the V3 batcher (settled once by run or unexpect)
    _expects_v3: bool = dataclasses.field(default=False, repr=False, compare=False)

    @property
    def mesh_key(self) -> str:
        nc = tuple(self.inputs.params.nc if self.nc is None else self.nc)
        return f"{self.key}|{nc[0]}x{nc[1]}x{nc[2]}"

    @property
    def latency(self) -> float:
        return self.t_done - self.t_arrive


class _Cached:
    def __init__(self, ctx: capi.Context, reserve: int, v3_precision: Optional[int] = None):
        self.ctx = ctx
        self.reserve = reserve          # bytes counted against the budget while cached
        self.lock = threading.Lock()
        self.busy = 0
        self.last_used = time.monotonic()
        self.v3_precision = v3_precision   # precision of the |V3|^2 the context holds


BATCHED_BRANCHES = (6, 12)


class V3Batcher:
    """Shared persistent |V3|^2 launches over the contexts queued by concurrent workers."""

    def __init__(self, window_s: float = 0.0):
        """window_s > 0: once a context is queued, wait up to that long for the admitted
        requests still on their way to the V3 stage (expect()) before launching; 0: launch
        at once with what is queued."""
        self._q: List[tuple] = []
        self._cv = threading.Condition()
        self._stop = False
        self._expected = 0
        self.window_s = window_s
        self.launches = 0               # batched launches run
        self.contexts = 0               # contexts they covered
        self.max_batch = 0
        self._t = threading.Thread(target=self._loop, daemon=True)
        self._t.start()

    def run(self, ctx: capi.Context, precision: int) -> None:
        """|V3|^2 of ctx (eigen data on the device) in the next shared launch; returns when it
        is done, raising the launch's error if it failed."""
        done, box = threading.Event(), {}
        with self._cv:
            self._expected = max(0, self._expected - 1)
            if self._stop:
                raise RuntimeError("V3 batcher closed")
            self._q.append((ctx, precision, done, box))
            self._cv.notify_all()
        done.wait()
        if "error" in box:
            raise box["error"]

    def expect(self) -> None:
        """An admitted request will queue its context (run) or withdraw (unexpect)."""
        with self._cv:
            self._expected += 1

    def unexpect(self) -> None:
        with self._cv:
            self._expected = max(0, self._expected - 1)
            self._cv.notify_all()

    def _loop(self) -> None:
        while True:
            with self._cv:
                while not self._q and not self._stop:
                    self._cv.wait()
                if not self._q:
                    return
                t_end = time.monotonic() + self.window_s
                while self._expected > 0 and not self._stop and time.monotonic() < t_end:
                    self._cv.wait(timeout=max(1e-4, t_end - time.monotonic()))
                key = (self._q[0][0].nb, self._q[0][1])
                group = [e for e in self._q if (e[0].nb, e[1]) == key]
                self._q = [e for e in self._q if (e[0].nb, e[1]) != key]
            try:
                capi.v3_batch([e[0] for e in group], key[1])
            except Exception as ex:   # each waiter sees it; its own solve then decides
                for e in group:
                    e[3]["error"] = ex
            self.launches += 1
            self.contexts += len(group)
            self.max_batch = max(self.max_batch, len(group))
            for e in group:
                e[2].set()

    def close(self) -> None:
        with self._cv:
            self._stop = True
            self._cv.notify_all()
        self._t.join()


def device_free_bytes() -> int:
    """Free device memory, from the library (alt3_device_free, hipMemGetInfo)."""
    return capi.device_free_bytes()


class KappaServer:
    """Continuous-batching scheduler over alt3 contexts (see the module docstring)."""

    def __init__(self, slots: int = 4, hbm_budget: Optional[int] = None, max_bypass: int = 8,
                 cache: bool = True, batch_v3: bool = True, batch_window_s: float = 0.0):
        capi.lib()                      # load the library once, before any worker thread
        self.slots = slots
        self.batcher = V3Batcher(batch_window_s) if batch_v3 else None
        self.budget = int(hbm_budget if hbm_budget is not None else 0.85 * device_free_bytes())
        self.max_bypass = max_bypass
        self.use_cache = cache
        self._pending: collections.deque = collections.deque()
        self._bypassed: Dict[int, int] = {}
        self._cache: "collections.OrderedDict[str, _Cached]" = collections.OrderedDict()
        self._in_use = 0                # bytes reserved by running requests and cached contexts
        self._running = 0
        self._cv = threading.Condition()
        self._ids = itertools.count()
        self._done: List[Request] = []
        self._stop = False
        self._threads: List[threading.Thread] = []
        self._sched = threading.Thread(target=self._schedule, daemon=True)
        self._sched.start()
        self.max_concurrent = 0

    # ------------------------------------------------------------------ API
    def submit(self, req: Request) -> Request:
        req.id = next(self._ids)
        req.t_arrive = time.perf_counter()
        req.footprint = self._estimate(req)
        with self._cv:
            if req.footprint > self.budget:
                req.error = f"footprint {req.footprint} B exceeds the HBM budget {self.budget} B"
                req.t_done = time.perf_counter()
                self._done.append(req)
                self._cv.notify_all()
                return req
            self._pending.append(req)
            self._cv.notify_all()
        return req

    def wait(self, reqs: Sequence[Request], timeout: Optional[float] = None) -> None:
        t_end = None if timeout is None else time.monotonic() + timeout
        with self._cv:
            while any(r.t_done == 0.0 for r in reqs):
                left = None if t_end is None else t_end - time.monotonic()
                if left is not None and left <= 0:
                    raise TimeoutError("requests still running")
                self._cv.wait(timeout=left if left is not None else 1.0)

    def close(self) -> None:
        with self._cv:
            self._stop = True
            self._cv.notify_all()
        self._sched.join()
        for t in self._threads:
            t.join()
        if self.batcher is not None:
            self.batcher.close()
        for c in self._cache.values():
            c.ctx.close()
        self._cache.clear()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ------------------------------------------------------------------ internals
    def _estimate(self, req: Request) -> int:
        with capi.Context(req.inputs, nc=req.nc, host_only=True) as h:
            req.nb = h.nb
            return h.footprint(len(req.temps), req.precision)

    def _evict_for(self, need: int) -> bool:
        """Drop idle cached contexts (least recently used first) until `need` bytes fit."""
        while self.budget - self._in_use < need:
            victim = next((k for k, c in self._cache.items() if c.busy == 0), None)
            if victim is None:
                return False
            c = self._cache.pop(victim)
            c.ctx.close()
            self._in_use -= c.reserve
        return True

    def _try_admit(self, req: Request) -> bool:
        cached = self._cache.get(req.mesh_key) if self.use_cache else None
        if cached is not None:
            extra = max(0, req.footprint - cached.reserve)
            if self.budget - self._in_use < extra and not self._evict_for(extra):
                return False
            cached.reserve += extra
            self._in_use += extra
            cached.busy += 1
            self._cache.move_to_end(req.mesh_key)
            req._entry = cached
            return True
        if self.budget - self._in_use < req.footprint and not self._evict_for(req.footprint):
            return False
        self._in_use += req.footprint
        req._entry = None
        return True

    def _schedule(self) -> None:
        with self._cv:
            while True:
                if self._stop and not self._pending:
                    return
                admitted = None
                if self._running < self.slots:
                    for i, req in enumerate(self._pending):
                        # a request passed over too often blocks everything behind it
                        if self._try_admit(req):
                            admitted = req
                            del self._pending[i]
                            for earlier in itertools.islice(self._pending, 0, i):
                                self._bypassed[earlier.id] = self._bypassed.get(earlier.id, 0) + 1
                            break
                        if self._bypassed.get(req.id, 0) >= self.max_bypass:
                            break
                if admitted is None:
                    self._cv.wait(timeout=0.5)
                    continue
                admitted.t_admit = time.perf_counter()
                if self.batcher is not None and admitted.nb in BATCHED_BRANCHES:
                    admitted._expects_v3 = True
                    self.batcher.expect()
                self._running += 1
                self.max_concurrent = max(self.max_concurrent, self._running)
                t = threading.Thread(target=self._work, args=(admitted,), daemon=True)
                self._threads.append(t)
                t.start()

    def _work(self, req: Request) -> None:
        ctx_entry: Optional[_Cached] = None
        own_ctx: Optional[capi.Context] = None
        try:
            # exactly what _try_admit decided: a cache entry inserted by another worker after
            # this request was charged its own footprint is not used (its busy count and
            # reserve were never raised for this request)
            ctx_entry = req._entry
            if ctx_entry is None:
                own_ctx = capi.Context(req.inputs, nc=req.nc)
                ctx, lock = own_ctx, threading.Lock()
            else:
                ctx, lock = ctx_entry.ctx, ctx_entry.lock
                req.reused = True
            with lock:
                ctx.set_delta(req.delta, req.sigma)
                held = ctx_entry is not None and ctx_entry.v3_precision == req.precision
                if req._expects_v3 and not held:
                    if ctx_entry is None:
                        ctx.eig(want=False)     # a cached context keeps its eigen data
                    req._expects_v3 = False     # run() settles the expectation
                    try:
                        self.batcher.run(ctx, req.precision)
                    except capi.Alt3Error:
                        pass    # the solve below recomputes what the launch did not leave
                out = ctx.solve([float(t) for t in req.temps], req.precision)
                if ctx_entry is not None:
                    ctx_entry.v3_precision = req.precision
            req.result = dict(kappa_rta=out["kappa_rta"], kappa_bte=out["kappa_bte"], cg_iters=out["cg_iters"],
                              stage_ms=out["stage_ms"])
        except Exception as e:  # the request fails, the server keeps serving
            req.error = f"{type(e).__name__}: {e}"
        finally:
            if req._expects_v3:
                req._expects_v3 = False
                self.batcher.unexpect()
            with self._cv:
                if ctx_entry is not None:
                    ctx_entry.busy -= 1
                    ctx_entry.last_used = time.monotonic()
                elif own_ctx is not None:
                    if self.use_cache and req.error is None and req.mesh_key not in self._cache:
                        self._cache[req.mesh_key] = _Cached(own_ctx, req.footprint, req.precision)
                    else:
                        own_ctx.close()
                        self._in_use -= req.footprint
                else:
                    self._in_use -= req.footprint
                self._running -= 1
                req.t_done = time.perf_counter()
                self._done.append(req)
                self._cv.notify_all()


def latency_summary(reqs: Sequence[Request]) -> dict:
    ok = [r for r in reqs if r.error is None]
    lat = np.array([r.latency for r in ok]) if ok else np.zeros(1)
    span = (max(r.t_done for r in reqs) - min(r.t_arrive for r in reqs)) if reqs else 0.0
    return {"requests": len(reqs), "completed": len(ok), "failed": len(reqs) - len(ok),
            "p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "mean_ms": float(lat.mean() * 1e3), "span_s": span,
            "solves_per_hour": len(ok) / span * 3600 if span > 0 else 0.0,
            "reused_contexts": sum(1 for r in ok if r.reused)}


def poisson_arrivals(n: int, rate_per_s: float, seed: int) -> np.ndarray:
    """Arrival times (s from the start) of a Poisson process."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate_per_s, size=n))


def replay(server: KappaServer, reqs: Sequence[Request], arrivals: Sequence[float],
           on_submit: Optional[Callable[[Request], None]] = None) -> None:
    """Submit each request at its arrival time (relative to now) and wait for all."""
    t0 = time.perf_counter()
    for req, ta in zip(reqs, arrivals):
        dt = t0 + ta - time.perf_counter()
        if dt > 0:
            time.sleep(dt)
        server.submit(req)
        if on_submit:
            on_submit(req)
    server.wait(reqs)
