"""The workload set beyond the SW-Si 16^3 step: the rock-salt-type 8-atom cell at 12^3, a
batched sweep of SW-Si instances, and a Poisson stream of kappa requests for the server.

Every workload builds its own synthetic inputs (model force constants written and read back
in the reference's file formats) and runs them through the C ABI on the device.
"""
from __future__ import annotations

import os
import tempfile
import time
from typing import Dict, List, Sequence

import numpy as np

from . import capi, models, refio, serving

CELLS = ("sw-si", "zb4", "rs8", "rs16")
# largest even mesh per cell in the request stream (16-atom cells beyond 6^3 need > 40 GB)
MESH_CAP = {"sw-si": 16, "zb4": 12, "rs8": 8, "rs16": 6}


_fc_done: Dict[tuple, models.ForceConstants] = {}


def cell_force_constants(name: str, a: float = 0.0, seed: int = 0):
    key = (name, a, seed)
    if key not in _fc_done:
        _fc_done[key] = _make_fc(name, a, seed)
    return _fc_done[key]


def _make_fc(name: str, a: float, seed: int):
    if name == "sw-si":
        return models.sw_silicon(a) if a else models.sw_silicon()
    if name == "zb4":
        return models.sw_zincblende4()
    if name == "rs8":
        return models.born_mayer_rocksalt(seed=seed)
    if name == "rs16":
        return models.born_mayer_supercell16(seed=seed)
    raise ValueError(name)


_inputs_cache: Dict[tuple, refio.InputSet] = {}


def cell_inputs(name: str, n: int, a: float = 0.0, seed: int = 0) -> refio.InputSet:
    key = (name, n, a, seed)
    if key not in _inputs_cache:
        d = tempfile.mkdtemp(prefix=f"alt3_{name}_{n}_")
        refio.write_inputs(d, cell_force_constants(name, a, seed), refio.RunParams(nc=(n, n, n)))
        _inputs_cache[key] = refio.read_inputs(d)
    return _inputs_cache[key]


def kiso(k: np.ndarray) -> float:
    return float(np.trace(k.sum(0)) / 3)


# ---------------------------------------------------------------------------- rock-salt 12^3
def run_rocksalt(n: int = 12, temps: Sequence[float] = (300.0,), log=print) -> dict:
    """8-atom rock-salt-type cell (24 branches): the whole path with FP64 and with mixed V3;
    |V3|^2 elements/s of each, and the kappa difference between the two."""
    inp = cell_inputs("rs8", n)
    out = {}
    with capi.Context(inp) as c:
        nel = c.nibz * c.nk * c.nb ** 3
        out["config"] = {"workload": f"rock-salt-type 8-atom cell, {n}^3 mesh, tetrahedron FGR + CG BTE at {list(temps)} K",
                         "nibz": c.nibz, "nk": c.nk, "nb": c.nb, "v3_elements": nel,
                         "footprint_bytes_fp64": c.footprint(len(temps), capi.V3_FP64)}
        t0 = time.perf_counter()
        c.eig(want=False)
        out["eig_ms"] = (time.perf_counter() - t0) * 1e3
        for name, p in (("fp64", capi.V3_FP64), ("mixed", capi.V3_MIXED)):
            c.eig(want=False)
            t0 = time.perf_counter()
            r = c.solve(list(temps), p)
            dt = time.perf_counter() - t0
            out[name] = {"ms": dt * 1e3, "elements_per_s": nel / dt,
                         "v3_kernel_elements_per_s": nel / (r["stage_ms"][1] * 1e-3),
                         "stage_ms": [round(float(x), 3) for x in r["stage_ms"]],
                         "kappa_rta": [kiso(k) for k in r["kappa_rta"]], "kappa_bte": [kiso(k) for k in r["kappa_bte"]]}
            log(f"rock-salt {n}^3 {name}: {dt * 1e3:.1f} ms, stages {out[name]['stage_ms']}")
    out["mixed_vs_fp64"] = {k: max(abs(a / b - 1) for a, b in zip(out["mixed"][k], out["fp64"][k]))
                            for k in ("kappa_rta", "kappa_bte")}
    return out


# ---------------------------------------------------------------------------- batched sweep
def sweep_requests(count: int = 64, seed: int = 0) -> List[serving.Request]:
    """SW-Si instances: lattice constant ~ U[5.4, 5.9] A per instance, mesh from
    {8,10,12,14}^3, temperatures from {100,200,300,500,800} K (all five per instance)."""
    rng = np.random.default_rng(seed)
    draws = [(round(float(rng.uniform(5.4, 5.9)), 4), int(rng.choice([8, 10, 12, 14]))) for _ in range(count)]
    for a, _ in draws:
        if ("sw-si", a, 0) not in _fc_done:
            _fc_done[("sw-si", a, 0)] = models.sw_silicon(a, batched=True)
    reqs = []
    for a, n in draws:
        reqs.append(serving.Request(f"sw-si-a{a}", cell_inputs("sw-si", n, a), [100.0, 200.0, 300.0, 500.0, 800.0],
                                    precision=capi.V3_MIXED))
    return reqs


def run_sweep(count: int = 64, seed: int = 0, slots: int = 4, log=print) -> dict:
    t0 = time.perf_counter()
    reqs = sweep_requests(count, seed)
    t_gen = time.perf_counter() - t0
    log(f"sweep: {count} instances generated in {t_gen:.1f} s")
    nel = 0
    with serving.KappaServer(slots=slots) as srv:
        t0 = time.perf_counter()
        for r in reqs:
            srv.submit(r)
        srv.wait(reqs)
        span = time.perf_counter() - t0
    for r in reqs:
        if r.error:
            raise RuntimeError(r.error)
        with capi.Context(r.inputs, host_only=True) as h:
            nel += h.nibz * h.nk * h.nb ** 3
    s = serving.latency_summary(reqs)
    return {"instances": count, "span_s": span, "v3_elements": nel, "elements_per_s": nel / span,
            "solves_per_hour": s["solves_per_hour"], "p50_ms": s["p50_ms"], "p99_ms": s["p99_ms"],
            "kappa_300K": {"rta_min": min(kiso(r.result["kappa_rta"][2]) for r in reqs),
                           "rta_max": max(kiso(r.result["kappa_rta"][2]) for r in reqs),
                           "bte_min": min(kiso(r.result["kappa_bte"][2]) for r in reqs),
                           "bte_max": max(kiso(r.result["kappa_bte"][2]) for r in reqs)},
            "generation_s": t_gen}


# ---------------------------------------------------------------------------- band-path self-energy
# The path of the reference's Ge example (THERMACOND/Ge-example/selfenergy/kpbs.params): 9
# directions of 40 points in primitive reciprocal units, L-G-K-X-W-L-K-W-X-G, 360 points.
FCC_PATH = """1
40
9
L 0.5 0.5 0.5
G 0. 0. 0
K 0.375 0.375 0.75
X 0.5 0.5 1
W 0.5 0.25 0.75
L 0.5 0.5 0.5
K 0.375 0.375 0.75
W 0.5 0.25 0.75
X 0.5 0.0  0.5
G 0. 0 0
"""


def run_selfenergy(n: int = 30, tempk: float = 300.0, eta: float = 30.0, log=print) -> dict:
    """Job 2 on SW-Si (diamond, 6 branches) at the shape of the reference's Ge example: an n^3
    mesh and the 360-point path above, every branch at its own frequency, T = 300 K, etaz = 30."""
    from . import selfenergy
    inp = cell_inputs("sw-si", n)
    lp = np.array(inp.latparams, dtype=float)
    lp[:3] *= inp.lattice_parameter
    kp, dk, _ = selfenergy.make_kp_bs(lp, np.asarray(inp.primlatt).reshape(3, 3, order="F"), FCC_PATH)
    with capi.Context(inp) as c:
        t0 = time.perf_counter()
        c.eig(want=False)
        t_eig = time.perf_counter() - t0
        c.self_w3(kp[:2], tempk, eta)   # warm-up (kernel loads)
        t0 = time.perf_counter()
        om, ns, us = c.self_w3(kp, tempk, eta)
        dt = time.perf_counter() - t0
        nk, nb = c.nk, c.nb
    nel = len(kp) * nk * nb ** 3
    lw = 2 * np.imag(ns + us)
    log(f"self-energy {n}^3, {len(kp)} points: {dt:.3f} s ({nel / dt:.3e} V3 elements/s)")
    return {"points": len(kp), "nk": nk, "nb": nb, "seconds": dt, "mesh_eig_s": t_eig,
            "points_per_s": len(kp) / dt, "v3_elements": nel, "v3_elements_per_s": nel / dt,
            "eigensolves": len(kp) * (nk + 1), "linewidth_cm1": {"min": float(lw.min()), "max": float(lw.max())}}


# ---------------------------------------------------------------------------- serving stream
def serving_requests(count: int, seed: int, precision: int) -> List[serving.Request]:
    """Cells uniform over CELLS, mesh an even size from 6^3 up to the cell's cap (4^3 where
    the cap is below 6), T ~ U[100, 800] K, one temperature per request."""
    rng = np.random.default_rng(seed)
    reqs = []
    for _ in range(count):
        cell = CELLS[int(rng.integers(len(CELLS)))]
        sizes = [m for m in range(6, MESH_CAP[cell] + 1, 2)] or [4]
        n = int(rng.choice(sizes))
        T = round(float(rng.uniform(100, 800)), 1)
        reqs.append(serving.Request(cell, cell_inputs(cell, n), [T], precision=precision))
    return reqs


def run_serving(count: int = 512, seed: int = 0, slots: int = 4, precision: int = capi.V3_MIXED,
                load: float = 0.8, probe: int = 32, log=print) -> dict:
    """Poisson arrivals at `load` times the rate one slot serves alone: the service time is
    measured first on `probe` requests of the same distribution run one at a time."""
    t0 = time.perf_counter()
    probe_reqs = serving_requests(probe, seed + 1000, precision)
    with serving.KappaServer(slots=1) as srv:
        for r in probe_reqs:
            srv.submit(r)
            srv.wait([r])
    service = float(np.mean([r.t_done - r.t_admit for r in probe_reqs]))
    rate = load / service
    log(f"serving: mean service time alone {service * 1e3:.1f} ms -> arrival rate {rate:.1f}/s")
    reqs = serving_requests(count, seed, precision)
    arrivals = serving.poisson_arrivals(count, rate, seed)
    with serving.KappaServer(slots=slots) as srv:
        serving.replay(srv, reqs, arrivals)
        conc = srv.max_concurrent
    s = serving.latency_summary(reqs)
    s.update({"arrival_rate_per_s": rate, "service_ms_alone": service * 1e3, "slots": slots,
              "max_concurrent": conc, "setup_s": time.perf_counter() - t0 - s["span_s"]})
    errs = [r.error for r in reqs if r.error]
    if errs:
        s["errors"] = errs[:3]
    return s
