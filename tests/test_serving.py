"""The kappa-solve server (thermacond/serving.py): footprint admission and continuous batching.

CPU: footprints come from host-only contexts and grow with the mesh, the branches and the
number of temperatures; a request larger than the budget is refused without touching the
device; the request stream draws meshes up to the lifted caps (16^3 for every cell whose
footprint fits, 8^3 for the 16-atom cell, whose 48 branches run the FP64 path: 348 GB at
10^3); a mixed request on that cell is charged the FP64 footprint.
GPU: a mixed stream of uneven requests (meshes 4^3-8^3, one to three temperatures, FP64 and
mixed V3, tetrahedron and Gaussian delta) run concurrently through the server gives, for
every request, the kappa the same request gives alone on a fresh context (lone_tol: 1e-9
relative, 1e-8 for a mixed request's kappa_BTE), the device held more than one
request at a time, and requests with a cached context reused it; requests queued together
share one persistent |V3|^2 launch and keep their lone results; the four cells (SW-Si, the
4-atom zincblende-type, the 8- and 16-atom rock-salt-type cells) at meshes up to 16^3 in
one stream, each result equal to its lone solve (lone_tol).
"""
import numpy as np
import pytest

from conftest import inputs_for
from thermacond import capi, serving, workloads


def lone_tol(req, k):
    """Relative tolerance of a request's kappa against its lone solve.  The FGR adds tetrahedron
    weights into FP64 sums with atomics whose order varies from run to run (1e-16); on the mixed
    path a last-bit change of a sum now and then flips the float rounding of a P element, and two
    lone solves of one mixed request differ by up to ~4e-11 (kappa_RTA) and ~3e-10 (kappa_BTE)."""
    if req.precision == capi.V3_FP64:
        return 1e-9
    return 1e-9 if k == "kappa_rta" else 1e-8


def test_footprint_grows_with_work(si_inputs):
    sizes = {}
    for n in (4, 8, 16):
        with capi.Context(si_inputs[n], host_only=True) as h:
            sizes[n] = (h.footprint(1, capi.V3_FP64), h.footprint(3, capi.V3_FP64), h.footprint(1, capi.V3_MIXED))
    assert sizes[4][0] < sizes[8][0] < sizes[16][0]
    for n in sizes:
        assert sizes[n][1] > sizes[n][0]
    assert sizes[16][2] < sizes[16][0]          # FP32 |V3|^2 halves the largest V3 array
    # 16^3 SW-Si: P1sm+P2sm alone are 3.9 GB
    assert sizes[16][0] > 2 * 244187136 * 8


def test_oversized_request_is_refused_on_the_host(si_inputs, monkeypatch):
    monkeypatch.setattr(serving, "device_free_bytes", lambda: 1 << 20)
    srv = serving.KappaServer.__new__(serving.KappaServer)
    # only the admission path: no scheduler thread, no device
    srv.budget = 1 << 20
    srv._cv = serving.threading.Condition()
    srv._done, srv._pending = [], serving.collections.deque()
    srv._ids = serving.itertools.count()
    req = srv.submit(serving.Request("si", si_inputs[8], [300.0]))
    assert req.error is not None and "exceeds the HBM budget" in req.error and req.t_done > 0


class _StubContext:
    """Stands in for capi.Context in the admission tests (no device)."""
    made = 0

    def __init__(self, *a, **k):
        _StubContext.made += 1
        self.closed = False

    def set_delta(self, *a):
        pass

    def solve(self, temps, precision):
        return {"kappa_rta": np.zeros((len(temps), 6, 3, 3)), "kappa_bte": np.zeros((len(temps), 6, 3, 3)),
                "cg_iters": np.zeros(len(temps), int), "stage_ms": np.zeros(8)}

    def close(self):
        self.closed = True


def _bare_server(budget):
    srv = serving.KappaServer.__new__(serving.KappaServer)
    srv.budget, srv.use_cache, srv.slots, srv.max_bypass = budget, True, 4, 8
    srv._cv = serving.threading.Condition()
    srv._done, srv._pending = [], serving.collections.deque()
    srv._cache = serving.collections.OrderedDict()
    srv._ids = serving.itertools.count()
    srv._in_use, srv._running, srv._bypassed = 0, 0, {}
    srv.batcher = None
    return srv


def test_worker_uses_the_admission_decision(si_inputs, monkeypatch):
    """A request admitted without a cached context (charged its own footprint) keeps to that
    decision even if another worker caches a context of the same crystal and mesh before this
    one starts: it builds its own context, leaves the cache entry's busy count alone and
    returns its footprint.  Using the new entry instead would decrement a busy count it never
    raised, so that a context still in use could be evicted and closed."""
    monkeypatch.setattr(serving.capi, "Context", _StubContext)
    srv = _bare_server(1 << 40)
    req = serving.Request("si", si_inputs[4], [300.0])
    req.footprint = 1000
    assert srv._try_admit(req) and req._entry is None and srv._in_use == 1000
    other = serving._Cached(_StubContext(), 700)          # another worker finished first
    srv._cache[req.mesh_key] = other
    srv._in_use += other.reserve
    srv._running = 1
    made = _StubContext.made
    srv._work(req)
    assert req.error is None and not req.reused
    assert _StubContext.made == made + 1                  # its own context
    assert other.busy == 0 and not other.ctx.closed
    assert srv._in_use == 700                             # only the cached context's reserve is left
    # a request admitted on the cached context does use it and releases its busy count
    req2 = serving.Request("si", si_inputs[4], [300.0])
    req2.footprint = 500
    srv._running = 1
    assert srv._try_admit(req2) and req2._entry is other and other.busy == 1
    srv._work(req2)
    assert req2.reused and other.busy == 0 and srv._in_use == 700


def _mix(si_inputs):
    S = si_inputs
    g = 0.1 * capi.THZ_TO_CM
    return [serving.Request("si", S[4], [300.0]),
            serving.Request("si", S[8], [100.0, 800.0]),
            serving.Request("si", S[4], [500.0], precision=capi.V3_MIXED),
            serving.Request("si", S[4], [200.0, 300.0, 400.0]),
            serving.Request("si", S[8], [300.0], delta=capi.DELTA_GAUSS, sigma=g),
            serving.Request("si", S[4], [300.0], nc=(6, 6, 6)),
            serving.Request("si", S[8], [250.0]),
            serving.Request("si", S[4], [700.0], nc=(6, 6, 6), precision=capi.V3_MIXED)]


@pytest.mark.gpu
def test_alone_equals_in_the_mix(si_inputs):
    print("case: serving mix of 8 requests", flush=True)
    alone = []
    with serving.KappaServer(slots=1, cache=False) as srv:
        for r in _mix(si_inputs):
            srv.submit(r)
            srv.wait([r])
            assert r.error is None, r.error
            alone.append(r)
    mix = _mix(si_inputs)
    with serving.KappaServer(slots=4, cache=True) as srv:
        serving.replay(srv, mix, serving.poisson_arrivals(len(mix), rate_per_s=200.0, seed=5))
        conc = srv.max_concurrent
    for a, m in zip(alone, mix):
        assert m.error is None, m.error
        for k in ("kappa_rta", "kappa_bte"):
            d = np.max(np.abs(m.result[k] - a.result[k])) / np.max(np.abs(a.result[k]))
            assert d < lone_tol(m, k), (m.mesh_key, m.temps, k, d)
        assert list(m.result["cg_iters"]) == list(a.result["cg_iters"])
    assert conc > 1
    assert any(m.reused for m in mix)
    s = serving.latency_summary(mix)
    print("serving mix:", s, flush=True)
    assert s["completed"] == len(mix)


@pytest.mark.gpu
def test_footprint_bounds_device_allocations(si_inputs):
    with capi.Context(si_inputs[8]) as c:
        est = c.footprint(3, capi.V3_FP64)
        c.solve([100.0, 300.0, 800.0], capi.V3_FP64)
        assert c.device_bytes() <= est


def test_mixed_footprint_of_a_48_branch_cell_is_the_fp64_one(cells, tmp_path_factory):
    inp = inputs_for(cells["rs16"], 2, tmp_path_factory, "fp_rs16")
    with capi.Context(inp, host_only=True) as h:
        assert h.footprint(1, capi.V3_MIXED) == h.footprint(1, capi.V3_FP64)
    inp = inputs_for(cells["rs8"], 8, tmp_path_factory, "fp_rs8")
    with capi.Context(inp, host_only=True) as h:   # 24 branches: FP32 |V3|^2 and P
        assert h.footprint(1, capi.V3_MIXED) < 0.6 * h.footprint(1, capi.V3_FP64)


def test_stream_draws_up_to_the_lifted_caps(monkeypatch):
    monkeypatch.setattr(workloads, "cell_inputs", lambda cell, n: (cell, n))
    monkeypatch.setattr(serving, "Request", lambda key, inp, temps, precision: inp)
    got = workloads.serving_requests(400, 0, capi.V3_MIXED)
    top = {c: max(n for cc, n in got if cc == c) for c in workloads.CELLS}
    assert top == {"sw-si": 16, "zb4": 16, "rs8": 16, "rs16": 8}
    got = workloads.serving_requests(400, 0, capi.V3_FP64)
    assert max(n for c, n in got if c == "rs16") == 8 and max(n for c, n in got if c == "rs8") == 16


@pytest.mark.gpu
def test_shared_v3_launch_keeps_lone_results(si_inputs, tmp_path_factory):
    print("case: serving, six requests in one shared V3 launch", flush=True)
    fc = si_inputs["fc"]
    inps = [inputs_for(fc, n, tmp_path_factory, f"sv{n}") for n in (4, 6, 8, 10)]

    def reqs():
        return [serving.Request("si", inp, [300.0], precision=p) for inp in inps[:3] for p in (capi.V3_MIXED,)] + \
               [serving.Request("si", inps[3], [200.0, 500.0], precision=capi.V3_MIXED),
                serving.Request("si", inps[1], [300.0], precision=capi.V3_FP64),
                serving.Request("si", inps[2], [300.0], precision=capi.V3_FP64)]
    alone = []
    with serving.KappaServer(slots=1, cache=False, batch_v3=False) as srv:
        for r in reqs():
            srv.submit(r)
            srv.wait([r])
            assert r.error is None, r.error
            alone.append(r)
    mix = reqs()
    with serving.KappaServer(slots=6, cache=False, batch_window_s=5.0) as srv:
        for r in mix:
            srv.submit(r)
        srv.wait(mix)
        b = srv.batcher
    assert b.contexts == len(mix) and b.launches < len(mix) and b.max_batch >= 2, (b.launches, b.max_batch)
    for a, m in zip(alone, mix):
        assert m.error is None, m.error
        for k in ("kappa_rta", "kappa_bte"):
            tol = 1e-12 if m.precision == capi.V3_FP64 else lone_tol(m, k)
            d = np.max(np.abs(m.result[k] - a.result[k])) / np.max(np.abs(a.result[k]))
            assert d < tol, (m.mesh_key, k, d)
        assert list(m.result["cg_iters"]) == list(a.result["cg_iters"])


def _four_cells(si_inputs, cells, tmp_path_factory):
    S = si_inputs
    I = lambda name, n: inputs_for(cells[name], n, tmp_path_factory, f"mix_{name}")
    M, F = capi.V3_MIXED, capi.V3_FP64
    zb16, zb8, rs16m, rs6, sc8, sc4 = I("zb4", 16), I("zb4", 8), I("rs8", 16), I("rs8", 6), I("rs16", 8), I("rs16", 4)
    return lambda: [serving.Request("si", S[16], [300.0], precision=M),
                    serving.Request("zb4", zb16, [250.0], precision=M),
                    serving.Request("rs8", rs16m, [300.0], precision=M),
                    serving.Request("rs16", sc8, [400.0], precision=M),
                    serving.Request("si", S[16], [100.0, 800.0], precision=M),
                    serving.Request("zb4", zb8, [700.0], precision=F),
                    serving.Request("rs8", rs6, [150.0], precision=F),
                    serving.Request("rs16", sc4, [600.0], precision=F),
                    serving.Request("si", S[8], [500.0], precision=F)]


@pytest.mark.gpu
def test_four_cell_mix_at_full_meshes(si_inputs, cells, tmp_path_factory):
    """SW-Si 16^3, the 4-atom cell at 16^3 and 8^3, the 8-atom cell at 16^3 (119 GB mixed)
    and 6^3, the 16-atom cell at 8^3 (60 GB) and 4^3, through one server: every result
    equals the request alone on a fresh context."""
    make = _four_cells(si_inputs, cells, tmp_path_factory)
    print("case: four-cell serving mix up to 16^3, alone", flush=True)
    alone = []
    with serving.KappaServer(slots=1, cache=False) as srv:
        for r in make():
            print(f"  alone {r.mesh_key} {r.temps} precision {r.precision}", flush=True)
            srv.submit(r)
            srv.wait([r])
            assert r.error is None, r.error
            alone.append(r)
    print("case: four-cell serving mix up to 16^3, in the mix", flush=True)
    mix = make()
    with serving.KappaServer(slots=4, cache=True) as srv:
        serving.replay(srv, mix, serving.poisson_arrivals(len(mix), rate_per_s=20.0, seed=11))
        conc = srv.max_concurrent
    for a, m in zip(alone, mix):
        assert m.error is None, m.error
        for k in ("kappa_rta", "kappa_bte"):
            d = np.max(np.abs(m.result[k] - a.result[k])) / np.max(np.abs(a.result[k]))
            assert d < lone_tol(m, k), (m.mesh_key, m.temps, k, d)
        assert list(m.result["cg_iters"]) == list(a.result["cg_iters"])
    assert conc > 1
    print("four-cell mix:", serving.latency_summary(mix), flush=True)
