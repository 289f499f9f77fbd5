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
# largest even mesh (at most 16^3) per cell in the request stream whose device footprint
# (alt3_footprint, one temperature) fits one 288 GB device: mixed V3 -- SW-Si 16^3 3.5 GB,
# zb4 16^3 38.5 GB, rs8 16^3 119 GB, rs16 8^3 108 GB (48 branches run the FP64 path, 10^3:
# 348 GB); FP64 -- rs8 16^3 233 GB, rs16 8^3 108 GB
MESH_CAP = {"sw-si": 16, "zb4": 16, "rs8": 16, "rs16": 8}
MESH_CAP_FP64 = {"sw-si": 16, "zb4": 16, "rs8": 16, "rs16": 8}


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
SWEEP_MESHES = (8, 10, 12, 14)
SWEEP_TEMPS = (100.0, 200.0, 300.0, 500.0, 800.0)


def sweep_draw(seed: int):
    """Instance `seed` of the SW-Si sweep: lattice constant ~ U[5.4, 5.9] A (rounded to
    1e-4 A) and mesh from {8,10,12,14}^3, drawn from numpy's default_rng(seed)."""
    rng = np.random.default_rng(seed)
    return round(float(rng.uniform(5.4, 5.9)), 4), int(rng.choice(SWEEP_MESHES))


def sweep_fc_pack(seeds: Sequence[int], fcs: Sequence[models.ForceConstants]) -> dict:
    """Force constants of sweep instances as flat arrays (npz): per instance its seed, a,
    mesh, atom list and FC2/FC3 term lists (indices int8, values float64)."""
    out = {"seeds": np.asarray(seeds, np.int64)}
    for s, fc in zip(seeds, fcs):
        a, n = sweep_draw(s)
        out[f"s{s}_a_n"] = np.array([a, n], np.float64)
        out[f"s{s}_atompos"] = fc.atompos
        out[f"s{s}_iatomcell0"] = fc.iatomcell0.astype(np.int8)
        out[f"s{s}_iatomcell"] = fc.iatomcell.astype(np.int8)
        out[f"s{s}_fc2_idx"] = fc.fc2[:, :4].astype(np.int8)
        out[f"s{s}_fc2_val"] = fc.fc2[:, 4]
        out[f"s{s}_fc3_idx"] = fc.fc3[:, :6].astype(np.int8)
        out[f"s{s}_fc3_val"] = fc.fc3[:, 6]
    return out


def sweep_fc_unpack(z, seed: int) -> models.ForceConstants:
    a, _ = z[f"s{seed}_a_n"]
    return models.ForceConstants(
        models.diamond(float(a)), z[f"s{seed}_atompos"], z[f"s{seed}_iatomcell0"].astype(np.int64),
        z[f"s{seed}_iatomcell"].astype(np.int64),
        np.column_stack([z[f"s{seed}_fc2_idx"].astype(np.float64), z[f"s{seed}_fc2_val"]]),
        np.column_stack([z[f"s{seed}_fc3_idx"].astype(np.float64), z[f"s{seed}_fc3_val"]]))


def sweep_instances(count: int = 64, seed: int = 0, fc_file: str = "") -> List[tuple]:
    """(seed, a, n, InputSet) of instances seed .. seed+count-1.  Force constants come from
    fc_file (sweep_fc_pack) where it holds the instance, else from models.sw_silicon(a)
    (about 7 s each on one core)."""
    z = np.load(fc_file) if fc_file and os.path.exists(fc_file) else None
    have = set(int(s) for s in z["seeds"]) if z is not None else set()
    out = []
    for s in range(seed, seed + count):
        a, n = sweep_draw(s)
        key = ("sw-si", a, 0)
        if key not in _fc_done:
            _fc_done[key] = sweep_fc_unpack(z, s) if s in have else models.sw_silicon(a, batched=True)
        out.append((s, a, n, cell_inputs("sw-si", n, a)))
    return out


def sweep_requests(count: int = 64, seed: int = 0, fc_file: str = "") -> List[serving.Request]:
    """SW-Si instances seed .. seed+count-1 (sweep_draw), all five temperatures of
    SWEEP_TEMPS each, mixed V3."""
    return [serving.Request(f"sw-si-s{s}-a{a}-{n}", inp, list(SWEEP_TEMPS), precision=capi.V3_MIXED)
            for s, a, n, inp in sweep_instances(count, seed, fc_file)]


def solve_packed(insts: Sequence[refio.InputSet], temps: Sequence[float], precision: int = capi.V3_MIXED,
                 tile: int = 0, log=None) -> dict:
    """Every instance on the device at once: eigen solves, then |V3|^2 of all instances in
    one shared persistent launch (capi.v3_batch), then per instance the rest of the path
    (alt3_solve reuses the |V3|^2 it holds).  Returns per-instance results, the tile used and
    the phase times."""
    ctxs = [capi.Context(inp) for inp in insts]
    try:
        t0 = time.perf_counter()
        for c in ctxs:
            c.eig(want=False)
        t1 = time.perf_counter()
        tile_used = capi.v3_batch(ctxs, precision, tile)
        t2 = time.perf_counter()
        res = [c.solve(list(temps), precision) for c in ctxs]
        t3 = time.perf_counter()
    finally:
        for c in ctxs:
            c.close()
    if log:
        log(f"packed: eig {1e3 * (t1 - t0):.1f} ms, V3 batch {1e3 * (t2 - t1):.1f} ms (tile {tile_used}), "
            f"rest {1e3 * (t3 - t2):.1f} ms")
    return {"results": res, "tile": tile_used, "eig_s": t1 - t0, "v3_s": t2 - t1, "rest_s": t3 - t2,
            "span_s": t3 - t0}


def run_sweep(count: int = 64, seed: int = 0, fc_file: str = "", log=print) -> dict:
    """The sweep packed: all instances resident, one shared V3 launch (tile autotuned, the
    tuning run done once before the timed pass), then each instance's FGR, tau, RTA and CG."""
    t0 = time.perf_counter()
    insts = sweep_instances(count, seed, fc_file)
    t_gen = time.perf_counter() - t0
    log(f"sweep: {count} instances ready in {t_gen:.1f} s")
    inps = [x[3] for x in insts]
    solve_packed(inps[:1], SWEEP_TEMPS[:1])     # kernel load + tile autotune, untimed
    r = solve_packed(inps, SWEEP_TEMPS, log=log)
    nel = 0
    for inp in inps:
        with capi.Context(inp, host_only=True) as h:
            nel += h.nibz * h.nk * h.nb ** 3
    res = r["results"]
    # the same instances one at a time (eig, per-context V3 launch, rest), for comparison
    lone_s = 0.0
    for inp in inps:
        with capi.Context(inp) as c:
            t0 = time.perf_counter()
            c.solve(list(SWEEP_TEMPS), capi.V3_MIXED)
            lone_s += time.perf_counter() - t0
    log(f"sweep: packed {r['span_s']:.2f} s, one at a time {lone_s:.2f} s")
    return {"instances": count, "span_s": r["span_s"], "lone_span_s": lone_s, "v3_batch_s": r["v3_s"], "tile": r["tile"],
            "v3_elements": nel, "elements_per_s": nel / r["span_s"], "v3_batch_elements_per_s": nel / r["v3_s"],
            "solves_per_hour": count / r["span_s"] * 3600,
            "kappa_300K": {"rta_min": min(kiso(o["kappa_rta"][2]) for o in res),
                           "rta_max": max(kiso(o["kappa_rta"][2]) for o in res),
                           "bte_min": min(kiso(o["kappa_bte"][2]) for o in res),
                           "bte_max": max(kiso(o["kappa_bte"][2]) for o in res)},
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
    """Cells uniform over CELLS, mesh an even size from 6^3 up to the cell's cap for the
    precision (MESH_CAP, MESH_CAP_FP64), T ~ U[100, 800] K, one temperature per request."""
    rng = np.random.default_rng(seed)
    caps = MESH_CAP if precision == capi.V3_MIXED else MESH_CAP_FP64
    reqs = []
    for _ in range(count):
        cell = CELLS[int(rng.integers(len(CELLS)))]
        sizes = [m for m in range(6, caps[cell] + 1, 2)]
        n = int(rng.choice(sizes))
        T = round(float(rng.uniform(100, 800)), 1)
        reqs.append(serving.Request(cell, cell_inputs(cell, n), [T], precision=precision))
    return reqs


def run_serving(count: int = 512, seed: int = 0, slots: int = 4, precision: int = capi.V3_MIXED,
                load: float = 0.8, calib: int = 64, log=print) -> dict:
    """Poisson arrivals at `load` times the server's capacity.  The capacity is measured first,
    closed loop, on the same server: `calib` requests of the same distribution (another seed)
    submitted at once, completions per second over their span.  The main stream then runs on
    that server (its context cache as the calibration left it), so that the offered rate is
    `load` of what the server sustains, and the latencies are those of a server at that load
    rather than of a growing queue."""
    t0 = time.perf_counter()
    cal_reqs = serving_requests(calib, seed + 1000, precision)
    reqs = serving_requests(count, seed, precision)
    t_gen = time.perf_counter() - t0
    log(f"serving: {calib} + {count} requests ready in {t_gen:.1f} s")
    with serving.KappaServer(slots=slots) as srv:
        t1 = time.perf_counter()
        for r in cal_reqs:
            srv.submit(r)
        srv.wait(cal_reqs)
        cal = serving.latency_summary(cal_reqs)
        capacity = cal["completed"] / cal["span_s"] if cal["span_s"] > 0 else 0.0
        rate = load * capacity
        log(f"serving: capacity {capacity:.2f} requests/s closed loop ({calib} requests, {cal['span_s']:.1f} s) "
            f"-> Poisson rate {rate:.2f}/s")
        arrivals = serving.poisson_arrivals(count, rate, seed)
        done = [0]

        def progress(req):
            done[0] += 1
            if done[0] % 64 == 0:
                log(f"serving: {done[0]}/{count} submitted at {time.perf_counter() - t1:.0f} s")

        serving.replay(srv, reqs, arrivals, on_submit=progress)
        conc = srv.max_concurrent
        b = srv.batcher
    s = serving.latency_summary(reqs)
    busy = sum(float(np.sum(r.result["stage_ms"])) for r in reqs if r.result) * 1e-3
    s.update({"arrival_rate_per_s": rate, "capacity_per_s": capacity, "load": load, "calibration": cal,
              "slots": slots, "max_concurrent": conc, "generation_s": t_gen,
              "device_stage_s_over_span": busy / s["span_s"] if s["span_s"] > 0 else None,
              "v3_shared_launches": b.launches, "v3_contexts_batched": b.contexts, "v3_max_batch": b.max_batch})
    errs = [r.error for r in reqs if r.error]
    if errs:
        s["errors"] = errs[:3]
    return s
