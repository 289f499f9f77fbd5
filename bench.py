#!/usr/bin/env python3
"""Benchmark of the three-phonon path on one MI355X.

Workload (the headline configuration): the 8-atom rock-salt-type cell (masses 72.63/78.97 amu,
a = 5.6 A, 24 branches, Born-Mayer-type force constants with a seeded 5 % cubic perturbation,
seed 0) on a Gamma-centred 12x12x12 q-mesh: 84 irreducible q1 x 1728 q2 x 13824 branch
triples = 2,006,581,248 |V3|^2 elements, tetrahedron delta integration and the full BTE by
conjugate gradients at T = 300 K.  One step = the whole path on the device: eigen solve on
the mesh, V3 over the irreducible zone, tetrahedron FGR, tau, kappa_RTA, collision matrix, CG
and kappa_BTE.

value = |V3|^2 elements per second of the whole step with the mixed-precision path (V3 on
FP16 matrix cores with split hi/lo operands and FP32 accumulation, FP32 tetrahedron weights
and P; FP64 eigen solve, scattering rates, CG and kappa); the same step with FP64 V3 is timed
beside it ("fp64"), together with the error of the mixed path against it ("mixed_vs_fp64":
kappa, per-mode tau^-1 and per-element |V3|^2).  The line also carries stage times, p50/p99
kappa-solve latency, the roofline of the dominant kernel ("roofline": the longer of the V3
contraction and the tetrahedron FGR, HIP-event times over the timed steps), both kernels'
rooflines ("roofline_v3" against the FP16 matrix peak, "roofline_fgr" against HBM), the CPU
baseline (the reference Fortran V3 routine, built from source, on one host core, over one
irreducible q1 of the same cell) and, under "sw_si_16cubed", the Stillinger-Weber Si 16^3 step
at 100/300/800 K.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import subprocess
import sys
import tempfile
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "anharmonic-lattice-dynamics_amd"))

PEAK_HBM_GBS = 8000.0
PEAK_FP32_MFMA_TFLOPS = 157.3
PEAK_FP16_MFMA_TFLOPS = 2500.0   # dense FP16 matrix peak (no sparsity)
STAGES = ["eig", "v3", "fgr", "tau", "kappa_rta", "collision", "cg", "kappa_bte"]


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--mesh", type=int, default=16)
    ap.add_argument("--precision", choices=["both", "mixed", "fp64", "fp64ref"], default="both")
    ap.add_argument("--temps", default="100,300,800")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-iter-bte", action="store_true", help="skip the iterative BTE (job 3) line")
    ap.add_argument("--no-iso", action="store_true", help="skip the isotope-scattering (iso = 1) line")
    ap.add_argument("--quiet", action="store_true")
    ap.add_argument("--workload", choices=["rocksalt", "step", "sweep", "serving", "selfenergy"], default="rocksalt",
                    help="rocksalt: the 8-atom rock-salt-type cell at 12^3, with the SW-Si 16^3 step beside it "
                         "(default); step: the SW-Si 16^3 step alone; sweep: 64 SW-Si instances through the "
                         "server; serving: a Poisson stream of 512 kappa requests over four cells")
    ap.add_argument("--no-sw-si", action="store_true", help="rocksalt: leave out the SW-Si 16^3 step")
    ap.add_argument("--requests", type=int, default=512)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB three times (reference for the "
                         "FETCH_SIZE/WRITE_SIZE counters) and exit")
    return ap.parse_args()


def copy_calibration(mib: int):
    import torch
    n = mib * 1024 * 1024 // 4
    x = torch.ones(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": n * 4, "launches": 3}))


def log(msg, quiet=False):
    if not quiet:
        print(msg, file=sys.stderr, flush=True)


def host_cores():
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count()
    return os.cpu_count(), usable


def cpu_baseline_reference(inputs_dir: str, quiet: bool, what: str):
    """Reference V3 routine (calculate_w3_ibz_split_sq, compiled from the reference
    sources into oracle/_ref/ref_probe) over IBZ q1 #1 of the inputs in inputs_dir, on one
    host core (the routine is serial; its cost per element does not depend on the mesh)."""
    probe = os.path.join(REPO, "oracle", "_ref", "ref_probe")
    if not os.path.exists(probe):
        return None
    env = dict(os.environ, REF_PROBE_V3_SAMPLE="1 1", OMP_NUM_THREADS="1")
    log(f"cpu baseline: reference V3 on one IBZ q1 of {what}", quiet)
    r = subprocess.run([probe], cwd=inputs_dir, env=env, capture_output=True, text=True, timeout=600)
    ncpu, usable = host_cores()
    for line in r.stdout.splitlines():
        m = re.match(r"\s*REF_PROBE_V3 elements=\s*(\d+)\s+seconds=\s*([0-9.EeDd+-]+)", line)
        if m:
            el, sec = int(m.group(1)), float(m.group(2).replace("D", "E").replace("d", "e"))
            return {"value": el / sec, "unit": "V3 elements/s", "cores": 1, "kind": "reference",
                    "host_cores": ncpu, "host_cores_usable": usable,
                    "sample": f"reference calculate_w3_ibz_split_sq over 1 irreducible q1 of {what} "
                              f"({el} elements, {sec:.1f} s) on 1 of the box's {usable} usable host cores"}
    return None


def timed(ctx, prec, temps, warmup, steps, label, quiet):
    """warmup + steps whole steps (eigen solve + alt3_solve); the step is synchronous, so the
    wall clock of each brackets its device work.  Also the HIP-event times of the V3
    contraction kernel and the FGR kernel of every timed step (context stream)."""
    lat, stage, out, kv, kf = [], None, None, [], []
    for it in range(warmup + steps):
        t0 = time.perf_counter()
        ctx.eig(want=False)
        out = ctx.solve(temps, prec)
        dt = time.perf_counter() - t0
        if it >= warmup:
            lat.append(dt)
            stage = out["stage_ms"].copy() if stage is None else stage + out["stage_ms"]
            v, f = ctx.kernel_ms()
            kv.append(v)
            kf.append(f)
        log(f"{label} step {it}: {dt * 1e3:.2f} ms  stages(ms)={[round(float(x), 3) for x in out['stage_ms']]}", quiet)
    stage = stage / steps
    lat_s = sorted(lat)
    return dict(ms=1e3 * sum(lat) / steps, p50=1e3 * lat_s[len(lat) // 2],
                p99=1e3 * lat_s[min(len(lat) - 1, int(round(0.99 * (len(lat) - 1))))], stage=stage, out=out,
                v3_kernel_ms=sum(kv) / len(kv), fgr_kernel_ms=sum(kf) / len(kf))


def timed_iter(ctx, prec, T, quiet):
    """The iterative BTE of job 3 at one temperature after a solve in the head precision: the
    first call builds the update matrix (S rows + star sums), the later ones reuse it."""
    ctx.eig(want=False)
    ctx.solve([T], prec)
    times, r = [], None
    for _ in range(3):
        t0 = time.perf_counter()
        r = ctx.kappa_iter(T)
        times.append((time.perf_counter() - t0) * 1e3)
    log(f"iterative BTE at {T} K: {r['iters']} iterations, {times[0]:.2f} ms first call, {times[-1]:.2f} ms after", quiet)
    return {"tempk": T, "iterations": r["iters"], "ms_first_call": times[0], "ms": times[-1],
            "ms_per_iteration": times[-1] / max(r["iters"], 1), "kappa": kappa_iso(r["kappa"]),
            "note": "kappa.f90 job 3: update_mpi as a device matrix, mixing 0.4, stop at normk < 1e-4"}


def timed_iso(ctx, prec, temps, giso, quiet):
    """Isotope scattering (iso = 1) on the same workload: the P_iso_split kernel alone (its
    HIP-event time), then whole steps with the isotope term in tau, collision and CG."""
    ctx.eig(want=False)
    ctx.set_iso(giso, want=False)
    t0 = time.perf_counter()
    ctx.set_iso(giso)                 # forms Piso now (the eigen data are current)
    piso_wall = (time.perf_counter() - t0) * 1e3
    piso_ms = ctx.last_ms()
    lat, out = [], None
    for it in range(3):
        t0 = time.perf_counter()
        ctx.eig(want=False)
        out = ctx.solve(temps, prec)
        lat.append((time.perf_counter() - t0) * 1e3)
    ctx.set_iso(None)
    ti = min(1, len(temps) - 1)
    log(f"iso: Piso kernel {piso_ms:.3f} ms, step with iso {min(lat[1:]):.2f} ms", quiet)
    return {"g_per_atom": [float(x) for x in giso], "piso_kernel_ms": piso_ms, "set_iso_wall_ms": piso_wall,
            "ms_per_step": sum(lat[1:]) / len(lat[1:]),
            "kappa_300K": {"rta": kappa_iso(out["kappa_rta"][ti]), "bte": kappa_iso(out["kappa_bte"][ti])},
            "note": "iso = 1: P_iso_split on the device (natural Si g), calculate_tauRTA_iso, CollisionMatrix_iso; "
                    "the step recomputes Piso after its eigen solve"}


def profiled_traffic(kernel_prefix: str, profile: str = "sw_si_16cubed_mixed_hbm_traffic.md"):
    """HBM bytes per launch of a kernel from a committed rocprofv3 FETCH_SIZE/WRITE_SIZE
    summary (calibrated against a plain copy), or None when it has no entry."""
    path = os.path.join(REPO, "profiles", profile)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch "):
            table = json.loads(line.split(" ", 1)[1])
            for name, b in table.items():
                if name.split("<")[0] == kernel_prefix:
                    return float(b)
    return None


def profiled_counter(kernel_prefix: str, counter: str, profile: str):
    """Per-launch value of a shader counter from a committed rocprofv3 --pmc summary (a
    markdown table with one column per counter), or None."""
    path = os.path.join(REPO, "profiles", profile)
    if not os.path.exists(path):
        return None
    cols = None
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if line.startswith("| kernel |"):
            cols = cells
        elif cols and line.startswith("| `") and counter in cols and cells[0].strip("`").split("<")[0] == kernel_prefix:
            return float(cells[cols.index(counter)])
    return None


# VALU issue rate of the chip: a wave64 FP32/FP64 vector instruction occupies a 16-lane SIMD
# for 4 cycles; 256 CUs x 4 SIMDs at 2.4 GHz
PEAK_VALU_WAVE_INSTS = 256 * 4 * 2.4e9 / 4


def fgr_roofline(nel: int, pb: int, fgr_ms: float, traffic_profile, sq_profile) -> dict:
    """The tetrahedron FGR kernel against HBM on the algorithmic bytes (read |V3|^2, write
    P1sm + P2sm: 3 pb bytes per element), with what it actually moves and issues beside it."""
    fgr_bytes = nel * 3 * pb
    fgr_gbs = fgr_bytes / (fgr_ms * 1e-3) / 1e9
    traffic = profiled_traffic("k_fgr", traffic_profile) if traffic_profile else None
    valu = profiled_counter("k_fgr", "SQ_INSTS_VALU", sq_profile) if sq_profile else None
    r = {"bound": "hbm", "achieved": fgr_gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": fgr_gbs / PEAK_HBM_GBS,
         "traffic": traffic, "kernel": "k_fgr", "kernel_ms": fgr_ms, "algorithmic_bytes_per_launch": fgr_bytes}
    if traffic is not None:
        r["measured_hbm_gbs"] = traffic / (fgr_ms * 1e-3) / 1e9
    if valu is not None:
        r["valu_issue"] = {"wave_insts_per_launch": valu, "peak_wave_insts_per_s": PEAK_VALU_WAVE_INSTS,
                           "frac": valu / (fgr_ms * 1e-3) / PEAK_VALU_WAVE_INSTS}
    if pb == 4:
        r["note"] = ("achieved counts the full |V3|^2 and P arrays; the FP32 path skips the rows whose l1 lies "
                     "outside an item's energy range (P row map), so the kernel moves 'traffic' bytes and is "
                     "bound by its vector ALU work (tetrahedron weights, list pushes, LDS atomics): "
                     "valu_issue.frac is SQ_INSTS_VALU over the chip's wave64 VALU issue rate")
    return r


def kappa_iso(k):
    return float(k.sum(0).trace() / 3)


def run_workload(args):
    """One JSON line for a workload other than the headline step (not timed by the driver)."""
    from thermacond import capi, workloads
    say = lambda m: log(m, args.quiet)
    if args.workload == "sweep":
        fc_file = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden", "sweep_fc.npz")
        out = workloads.run_sweep(64, 0, fc_file=fc_file, log=say)
        res = {"metric": "batched sweep: V3 elements/s over 64 SW-Si instances", "unit": "V3 elements/s",
               "value": out["elements_per_s"], "higher_is_better": True, "n_gpus": 1, "dtype": "fp32 V3 + f64",
               "data": "synthetic (SW-Si, seeds 0..63: a ~ U[5.4,5.9] A, mesh from {8,10,12,14}^3)",
               "config": {"workload": "64 SW-Si instances resident, |V3|^2 of all in one shared persistent launch "
                                      "(autotuned q2 tile), then FGR, tau, RTA and CG per instance at "
                                      "T = 100,200,300,500,800 K"}, **out}
    elif args.workload == "selfenergy":
        out = workloads.run_selfenergy(30, log=say)
        ref_s = 97889.0   # Ge-example/selfenergy/times.dat: self-energy stage of the same shape
        res = {"metric": "band-path self-energy: (q, branch) evaluations/s (job 2)", "unit": "evaluations/s",
               "value": out["points"] * out["nb"] / out["seconds"], "higher_is_better": True, "n_gpus": 1,
               "dtype": "f64", "data": "synthetic (SW-Si force constants; the reference's Ge path and mesh shape)",
               "config": {"workload": "SW-Si diamond, 30^3 mesh, 360-point L-G-K-X-W-L-K-W-X-G path, 6 branches, "
                                      "T = 300 K, etaz = 30 cm^-1"},
               "reference_published": {"seconds": ref_s, "what": "Ge (diamond, 6 branches), 30^3 mesh, the same "
                                       "360-point path, job 2 on CPU (Ge-example/selfenergy/times.dat); other "
                                       "force constants, so a shape comparison only"},
               **out}
    else:
        runs = {}
        for name, p in (("mixed", capi.V3_MIXED), ("fp64", capi.V3_FP64)):
            if args.precision in ("both", name):
                runs[name] = workloads.run_serving(args.requests, 0, precision=p, log=say)
                say(f"serving {name}: {runs[name]}")
        head = runs.get("mixed") or runs["fp64"]
        res = {"metric": "kappa solves/hour at p99 latency (Poisson stream at 80 % of the server's closed-loop "
                         "capacity)", "unit": "solves/hour",
               "value": head["solves_per_hour"], "higher_is_better": True, "n_gpus": 1,
               "dtype": "fp32 V3 + f64", "data": "synthetic (SW-Si, zincblende-type 4-atom, rock-salt-type 8- and "
               "16-atom cells)", "config": {"workload": f"{args.requests} requests, cells uniform, mesh even "
               f"6^3..cap {workloads.MESH_CAP}, T ~ U[100,800] K, 4 slots; arrival rate 0.8 x the capacity "
               "measured closed loop on 64 requests of the same distribution (another seed)"},
               "latency_ms": {"p50": head["p50_ms"], "p99": head["p99_ms"]}, **runs}
        if "fp64" in runs:
            res["fp64_solves_per_hour"] = runs["fp64"]["solves_per_hour"]
            res["fp64_latency_ms"] = {"p50": runs["fp64"]["p50_ms"], "p99": runs["fp64"]["p99_ms"]}
    print(json.dumps(res))


def bench_sw_si(args, with_cpu_baseline: bool) -> dict:
    """The SW-Si 16^3 step at the temperatures of --temps, in the precisions of --precision."""
    from thermacond import capi, models, refio

    temps = [float(t) for t in args.temps.split(",")]
    n = args.mesh
    tmp = tempfile.mkdtemp(prefix="alt3_bench_")
    refio.write_inputs(tmp, models.sw_silicon(), refio.RunParams(nc=(n, n, n)))
    inp = refio.read_inputs(tmp)
    log(f"bench: SW-Si {n}^3, precision={args.precision}, temps={temps}", args.quiet)
    ctx = capi.Context(inp)
    nel = ctx.nibz * ctx.nk * ctx.nb ** 3
    sizes = {"nibz": ctx.nibz, "nk": ctx.nk, "nb": ctx.nb}
    precs = {"both": [("mixed", capi.V3_MIXED), ("fp64", capi.V3_FP64)], "mixed": [("mixed", capi.V3_MIXED)],
             "fp64": [("fp64", capi.V3_FP64)], "fp64ref": [("fp64ref", capi.V3_FP64_REF)]}[args.precision]
    runs = {name: timed(ctx, p, temps, args.warmup, args.steps, name, args.quiet) for name, p in precs}
    head_name = precs[0][0]
    head = runs[head_name]
    ti = min(1, len(temps) - 1)
    iter_bte = None if args.no_iter_bte else timed_iter(ctx, precs[0][1], temps[ti], args.quiet)
    iso = None if args.no_iso else timed_iso(ctx, precs[0][1], temps, capi.giso_of(inp), args.quiet)
    ctx.close()

    def summary(r):
        return {"value": nel / (r["ms"] * 1e-3), "ms_per_step": r["ms"],
                "v3_kernel_elements_per_s": nel / (r["stage"][1] * 1e-3),
                "stage_ms": {k: round(float(v), 3) for k, v in zip(STAGES, r["stage"])},
                "latency_ms": {"p50": r["p50"], "p99": r["p99"]},
                "kappa_300K": {"rta": kappa_iso(r["out"]["kappa_rta"][ti]), "bte": kappa_iso(r["out"]["kappa_bte"][ti])}}

    # dominant kernel of the step: the tetrahedron FGR kernel (one launch per step)
    v3_bytes = 4 if head_name == "mixed" else 8   # |V3|^2 and P1sm/P2sm share the precision
    fgr_ms = float(head["stage"][2])
    prof = n == 16 and head_name == "mixed"
    res = {
        "metric": "3-ph V3 elements/s (mixed vs FP64) at kappa err <=0.5%; p50/p99 kappa-solve latency",
        "value": nel / (head["ms"] * 1e-3),
        "unit": "V3 elements/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": head["ms"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": {"mixed": "fp32 (V3, P) + f64", "fp64": "f64", "fp64ref": "f64"}[head_name],
        "data": "synthetic (Stillinger-Weber Si force constants by finite differences)",
        "config": {"workload": f"SW-Si 2-atom diamond, {n}^3 mesh, tetrahedron FGR + CG BTE at {temps} K",
                   **sizes, "v3_elements": nel, "precision": head_name},
        "latency_ms": {"p50": head["p50"], "p99": head["p99"]},
        "roofline": fgr_roofline(nel, v3_bytes, fgr_ms, "sw_si_16cubed_mixed_hbm_traffic.md" if prof else None,
                                 "sw_si_16cubed_mixed_sq_counters.md" if prof else None),
    }
    for name, r in runs.items():
        res[name] = summary(r)
    if iter_bte is not None:
        res["iter_bte"] = iter_bte
    if iso is not None:
        res["iso"] = iso
    if "mixed" in runs and "fp64" in runs:
        m, f = runs["mixed"], runs["fp64"]
        errs = {}
        for k in ("kappa_rta", "kappa_bte"):
            errs[k] = max(abs(kappa_iso(m["out"][k][i]) / kappa_iso(f["out"][k][i]) - 1) for i in range(len(temps)))
        res["mixed_vs_fp64"] = {"whole_step_speedup": f["ms"] / m["ms"],
                                "v3_kernel_speedup": float(f["stage"][1] / m["stage"][1]),
                                "kappa_rel_err_max": errs}
    if with_cpu_baseline and not args.no_cpu_baseline:
        cb = cpu_baseline_reference(tmp, args.quiet, f"the same {n}^3 SW-Si workload")
        if cb is not None:
            res["cpu_baseline"] = cb
    return res


def fc3_counts(inp):
    """Nt3 (FC3 terms with a home-cell first atom) and N_cp (distinct cell pairs (Rj, Rk) among
    them), the counts of the per-element flop figure 24 nb + 3 + (4 Nt3 + 6 N_cp) / nb^3."""
    a = np.asarray(inp.fc3_atoms) - 1
    ic0 = np.asarray(inp.iatomcell0) - 1
    home = a[ic0[a[:, 0]] == a[:, 0]]
    pos = np.asarray(inp.atompos)
    rj = pos[home[:, 1]] - pos[ic0[home[:, 1]]]
    rk = pos[home[:, 2]] - pos[ic0[home[:, 2]]]
    ncp = len({tuple(x) for x in np.round(np.hstack([rj, rk]) * 1e6).astype(np.int64)})
    ndr = len({tuple(x) for x in np.round((rj - rk) * 1e6).astype(np.int64)})
    return len(home), ncp, ndr


def gamma_acoustic_masks(ctx):
    """Per mesh index: Gamma, and the index of -q (for q3 = -q1-q2 = Gamma)."""
    nc = ctx.nc
    k = np.arange(ctx.nk)
    c2, c1, c0 = k % nc[2], (k // nc[2]) % nc[1], k // (nc[1] * nc[2])
    neg = ((nc[2] - c2) % nc[2]) + nc[2] * (((nc[1] - c1) % nc[1]) + nc[1] * ((nc[0] - c0) % nc[0]))
    gam = (nc[2] // 2) + nc[2] * ((nc[1] // 2) + nc[1] * (nc[0] // 2))
    return gam, neg, (c0, c1, c2)


def mixed_error_statement(ctx, T, nblock=4):
    """The mixed path against the FP64 path on the same context: per-mode tau^-1 (all irreducible
    modes but the three acoustic ones at Gamma, whose frequency is zero), and per-element
    |V3|^2 on the first nblock irreducible points, overall and on the triples with an acoustic
    mode within one mesh step of Gamma (the near-Gamma modes; those at Gamma itself are
    round-off of the acoustic sum rule in both precisions and masked by the FGR as q2 and q3)."""
    from thermacond import capi
    tau = {}
    for name, p in (("fp64", capi.V3_FP64), ("mixed", capi.V3_MIXED)):
        ctx.v3sq(0, ctx.nibz - 1, p, want=False)
        ctx.fgr(want=False)
        tau[name] = ctx.tau(T)["tautot"]
    gam, neg, (c0, c1, c2) = gamma_acoustic_masks(ctx)
    mapinv = ctx.mesh()["mapinv"]
    keep = np.ones(tau["fp64"].shape, bool)
    keep[[i for i in range(ctx.nibz) if mapinv[i] == gam], :3] = False
    rt = np.abs(tau["mixed"][keep] - tau["fp64"][keep]) / np.abs(tau["fp64"][keep])
    hi = min(nblock, ctx.nibz) - 1
    vf = ctx.v3sq(0, hi, capi.V3_FP64)
    vm = ctx.v3sq(0, hi, capi.V3_MIXED)
    nb = ctx.nb
    mx = float(np.abs(vf).max())
    d = np.abs(vm - vf)
    # triples with an acoustic mode at Gamma (q1, q2 or q3) and within one mesh step of it
    n = np.asarray(ctx.nc)
    h = n // 2
    near = (np.abs(c0 - h[0]) <= 1) & (np.abs(c1 - h[1]) <= 1) & (np.abs(c2 - h[2]) <= 1)
    at_g = np.zeros(vf.shape, bool)
    near_g = np.zeros(vf.shape, bool)
    ac = np.arange(nb) < 3
    kk = np.arange(ctx.nk)
    A1, A2, A3 = ac[None, :, None, None], ac[None, None, :, None], ac[None, None, None, :]   # (q2, l1, l2, l3)
    for s in range(hi + 1):
        k1 = mapinv[s]
        q3 = neg[k1]   # q3 = -q1-q2 is Gamma where q2 = -q1
        if k1 == gam:
            at_g[s] |= A1
        elif near[k1]:
            near_g[s] |= A1
        at_g[s] |= (kk == gam)[:, None, None, None] & A2
        near_g[s] |= (near & (kk != gam))[:, None, None, None] & A2
        at_g[s] |= (kk == q3)[:, None, None, None] & A3
        d0 = np.abs(((c0 - c0[q3] + h[0]) % n[0]) - h[0])
        d1 = np.abs(((c1 - c1[q3] + h[1]) % n[1]) - h[1])
        d2 = np.abs(((c2 - c2[q3] + h[2]) % n[2]) - h[2])
        nq3 = (d0 <= 1) & (d1 <= 1) & (d2 <= 1) & (kk != q3)
        near_g[s] |= nq3[:, None, None, None] & A3
    near_g &= ~at_g
    big = (np.abs(vf) > 1e-6 * mx) & ~at_g
    rel = d[big] / np.abs(vf[big])
    nb_ = (np.abs(vf) > 1e-6 * mx) & near_g
    reln = d[nb_] / np.abs(vf[nb_])
    return {"tauinv_rel_err": {"max": float(rt.max()), "p99": float(np.percentile(rt, 99)), "modes": int(rt.size)},
            "v3sq_rel_err": {"irreducible_points": hi + 1, "elements": int(big.sum()),
                             "max": float(rel.max()), "p99": float(np.percentile(rel, 99)),
                             "max_abs_over_largest": float(d[~at_g].max() / mx),
                             "selection": "elements above 1e-6 of the largest, triples with an acoustic mode at "
                                          "Gamma left out"},
            "v3sq_rel_err_near_gamma": {"elements": int(nb_.sum()),
                                        "max": float(reln.max()) if reln.size else None,
                                        "p99": float(np.percentile(reln, 99)) if reln.size else None,
                                        "selection": "triples with an acoustic mode within one mesh step of Gamma "
                                                     "(not at it), elements above 1e-6 of the largest"},
            "v3sq_at_gamma_acoustic": {"elements": int(at_g.sum()),
                                       "max_abs_over_largest": float(d[at_g].max() / mx) if at_g.any() else None,
                                       "note": "round-off of the acoustic sum rule in both precisions; the FGR "
                                               "masks these as q2 and q3"}}


def bench_rocksalt(args) -> dict:
    from thermacond import capi, models, refio, workloads
    T = 300.0
    n = 12
    inp = workloads.cell_inputs("rs8", n)
    log(f"bench: rock-salt-type 8-atom cell {n}^3, T = {T} K", args.quiet)
    ctx = capi.Context(inp)
    nb = ctx.nb
    nel = ctx.nibz * ctx.nk * nb ** 3
    sizes = {"nibz": ctx.nibz, "nk": ctx.nk, "nb": nb}
    precs = {"both": [("mixed", capi.V3_MIXED), ("fp64", capi.V3_FP64)], "mixed": [("mixed", capi.V3_MIXED)],
             "fp64": [("fp64", capi.V3_FP64)]}.get(args.precision, [("mixed", capi.V3_MIXED), ("fp64", capi.V3_FP64)])
    runs = {name: timed(ctx, p, [T], args.warmup, args.steps, f"rock-salt {name}", args.quiet) for name, p in precs}
    head_name = precs[0][0]
    head = runs[head_name]
    err = mixed_error_statement(ctx, T) if len(runs) == 2 else None
    ctx.close()
    nt3, ncp, ndr = fc3_counts(inp)
    f_elem = 24 * nb + 3 + (4 * nt3 + 6 * ncp) / nb ** 3     # SURVEY 8(d) per-element flops
    v3_ms = head["v3_kernel_ms"]
    tflops = f_elem * nel / (v3_ms * 1e-3) / 1e12
    fgr_ms = head["fgr_kernel_ms"]
    pb = 4 if head_name == "mixed" else 8
    if head_name == "mixed":
        # k_v3_chain16: FP16 MFMA on split (hi + lo) operands, three products per product.  Per
        # (q1, l1, tile of 16 points): 36 phase M-tiles x 2 k-steps x 3 x (Re, Im), per point U
        # 3 x 2 tiles x 2 k-steps x 3 and V 2 x 3 tiles x 2 k-steps x 3 v_mfma_f32_16x16x32_f16
        # (16384 flops each), over 16 x 576 elements
        mfma_per_tile = 36 * 2 * 3 * 2 + 16 * (36 + 36)
        f_exec = mfma_per_tile * 16384 / (16 * nb * nb)
        v3_roof = {"bound": "mfma", "achieved": tflops, "peak": PEAK_FP16_MFMA_TFLOPS, "unit": "TFLOP/s",
                   "frac": tflops / PEAK_FP16_MFMA_TFLOPS,
                   "traffic": profiled_traffic("k_v3_chain16", "rs8_12cubed_mixed_hbm_traffic.md"),
                   "kernel": "k_v3_chain16", "kernel_ms": v3_ms, "flops_per_element": f_elem,
                   "executed_fp16_flops_per_element": f_exec,
                   "executed_tflops": f_exec * nel / (v3_ms * 1e-3) / 1e12,
                   "executed_frac": f_exec * nel / (v3_ms * 1e-3) / 1e12 / PEAK_FP16_MFMA_TFLOPS,
                   "fp32_equivalent_frac": tflops / PEAK_FP32_MFMA_TFLOPS,
                   "note": "achieved = the algorithmic 582 flops per element (SURVEY 8(d) F_elem) over the kernel "
                           "time, against the dense FP16 matrix peak the kernel runs on; executed_* count the "
                           "split-operand FP16 products it issues (3 per product, padded tiles); "
                           "fp32_equivalent_frac = the algorithmic rate over the FP32 matrix peak",
                   "counts": {"nt3": nt3, "n_cp": ncp, "ndr": ndr},
                   "algorithmic_flops_per_launch": f_elem * nel}
    else:
        f_exec = 8 * ndr + 16 * nb + 3                        # flops the batched FP64 form executes
        v3_roof = {"bound": "fp64", "achieved": tflops, "peak": 78.6, "unit": "TFLOP/s", "frac": tflops / 78.6,
                   "traffic": None, "kernel": "k_v3_main", "kernel_ms": v3_ms, "flops_per_element": f_elem,
                   "executed_flops_per_element": f_exec, "counts": {"nt3": nt3, "n_cp": ncp, "ndr": ndr},
                   "algorithmic_flops_per_launch": f_elem * nel}
    fgr_roof = fgr_roofline(nel, pb, fgr_ms, "rs8_12cubed_mixed_hbm_traffic.md" if head_name == "mixed" else None,
                            "rs8_12cubed_mixed_sq_counters.md" if head_name == "mixed" else None)

    def summary(r):
        return {"value": nel / (r["ms"] * 1e-3), "ms_per_step": r["ms"],
                "v3_kernel_ms": r["v3_kernel_ms"], "v3_kernel_elements_per_s": nel / (r["v3_kernel_ms"] * 1e-3),
                "fgr_kernel_ms": r["fgr_kernel_ms"],
                "stage_ms": {k: round(float(v), 3) for k, v in zip(STAGES, r["stage"])},
                "latency_ms": {"p50": r["p50"], "p99": r["p99"]},
                "kappa_300K": {"rta": kappa_iso(r["out"]["kappa_rta"][0]), "bte": kappa_iso(r["out"]["kappa_bte"][0])},
                "cg_iterations": int(r["out"]["cg_iters"][0])}

    res = {
        "metric": "3-ph V3 elements/s (mixed vs FP64) at kappa err <=0.5%; p50/p99 kappa-solve latency",
        "value": nel / (head["ms"] * 1e-3),
        "unit": "V3 elements/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": head["ms"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": {"mixed": "f16x3 (V3 on FP16 matrix cores, split hi/lo operands, FP32 accumulation) + fp32 "
                           "(tetrahedron weights, P) + f64", "fp64": "f64"}[head_name],
        "data": "synthetic (Born-Mayer-type rock-salt force constants by finite differences, seeded 5 % cubic "
                "perturbation, seed 0)",
        "config": {"workload": f"rock-salt-type 8-atom cell (a = 5.6 A, masses 72.63/78.97, 24 branches), {n}^3 "
                               f"mesh, tetrahedron FGR + CG BTE at {T} K", **sizes, "v3_elements": nel,
                   "precision": head_name},
        "latency_ms": {"p50": head["p50"], "p99": head["p99"]},
        # the dominant kernel of the step (the longer of the V3 contraction and the FGR)
        "roofline": fgr_roof if fgr_ms >= v3_ms else v3_roof,
        "roofline_v3": v3_roof,
        "roofline_fgr": fgr_roof,
    }
    for name, r in runs.items():
        res[name] = summary(r)
    if len(runs) == 2:
        m, f = runs["mixed"], runs["fp64"]
        res["mixed_vs_fp64"] = {
            "whole_step_speedup": f["ms"] / m["ms"], "v3_kernel_speedup": f["v3_kernel_ms"] / m["v3_kernel_ms"],
            "kappa_rel_err": {k: abs(kappa_iso(m["out"][k][0]) / kappa_iso(f["out"][k][0]) - 1)
                              for k in ("kappa_rta", "kappa_bte")}, **err}
    if not args.no_cpu_baseline:
        tmp = tempfile.mkdtemp(prefix="alt3_bench_rs8_2_")
        refio.write_inputs(tmp, models.born_mayer_rocksalt(seed=0), refio.RunParams(nc=(2, 2, 2)))
        cb = cpu_baseline_reference(tmp, args.quiet, "the same rock-salt-type cell (24 branches, 2^3 mesh; the "
                                                     "routine's cost per element does not depend on the mesh)")
        if cb is not None:
            res["cpu_baseline"] = cb
    return res


def main():
    args = parse()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; sharding over q1 across GPUs is "
                                   "described in DESIGN.md and not launched", "n_gpus": args.gpus}))
        sys.exit(2)
    if args.workload == "step":
        print(json.dumps(bench_sw_si(args, True)))
        return
    if args.workload != "rocksalt":
        run_workload(args)
        return
    res = bench_rocksalt(args)
    if not args.no_sw_si:
        sw = bench_sw_si(args, False)
        res["sw_si_16cubed"] = {k: v for k, v in sw.items() if k not in ("metric", "n_gpus", "steps", "warmup",
                                                                         "higher_is_better", "scaling", "vs_baseline")}
    print(json.dumps(res))


if __name__ == "__main__":
    main()
