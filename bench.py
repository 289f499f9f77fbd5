#!/usr/bin/env python3
"""Benchmark of the three-phonon path on one MI355X.

Workload (the headline configuration): Stillinger-Weber Si, 2-atom diamond cell (a=5.43 A,
6 branches), Gamma-centred 16x16x16 q-mesh (276 irreducible q1 x 4096 q2 x 216 branch
triples = 244,187,136 |V3|^2 elements), tetrahedron delta integration and the full BTE by
conjugate gradients at T = 100, 300, 800 K.  One step = the whole path on the device:
eigen solve on the mesh, V3 over the irreducible zone, tetrahedron FGR, then per
temperature tau, kappa_RTA, collision matrix, CG and kappa_BTE.

value = |V3|^2 elements per second of the whole step (higher is better).  The JSON line
also carries the V3-kernel-only rate, stage times, p50/p99 kappa-solve latency, the
roofline of the dominant kernel and the CPU baseline (the reference Fortran V3 routine,
built from source, on one host core, over one irreducible q1).
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

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "anharmonic-lattice-dynamics_amd"))

PEAK_HBM_GBS = 8000.0
PEAK_FP64_TFLOPS = 78.6
PEAK_FP32_TFLOPS = 157.3


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--mesh", type=int, default=16)
    ap.add_argument("--precision", choices=["mixed", "fp64", "fp64ref"], default="fp64")
    ap.add_argument("--temps", default="100,300,800")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--quiet", action="store_true")
    return ap.parse_args()


def log(msg, quiet=False):
    if not quiet:
        print(msg, file=sys.stderr, flush=True)


def cpu_baseline_reference(inputs_dir: str, quiet: bool):
    """Reference V3 routine (calculate_w3_ibz_split_sq, compiled from the reference
    sources into oracle/_ref/ref_probe) over IBZ q1 #1 on one host core."""
    probe = os.path.join(REPO, "oracle", "_ref", "ref_probe")
    if not os.path.exists(probe):
        return None
    env = dict(os.environ, REF_PROBE_V3_SAMPLE="1 1", OMP_NUM_THREADS="1")
    log("cpu baseline: reference V3 on one IBZ q1 (about 45 s)", quiet)
    r = subprocess.run([probe], cwd=inputs_dir, env=env, capture_output=True, text=True, timeout=600)
    for line in r.stdout.splitlines():
        m = re.match(r"\s*REF_PROBE_V3 elements=\s*(\d+)\s+seconds=\s*([0-9.EeDd+-]+)", line)
        if m:
            el, sec = int(m.group(1)), float(m.group(2).replace("D", "E").replace("d", "e"))
            return {"value": el / sec, "unit": "V3 elements/s", "cores": 1, "kind": "reference",
                    "sample": f"reference calculate_w3_ibz_split_sq over 1 irreducible q1 of the same "
                              f"16^3 SW-Si workload ({el} elements, {sec:.1f} s)"}
    return None


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; multi-GPU sharding over q1 is "
                                   "described in DESIGN.md but not launched", "n_gpus": args.gpus}))
        sys.exit(2)
    from thermacond import capi, models, refio

    prec = {"mixed": capi.V3_MIXED, "fp64": capi.V3_FP64, "fp64ref": capi.V3_FP64_REF}[args.precision]
    temps = [float(t) for t in args.temps.split(",")]
    n = args.mesh
    tmp = tempfile.mkdtemp(prefix="alt3_bench_")
    fc = models.sw_silicon()
    refio.write_inputs(tmp, fc, refio.RunParams(nc=(n, n, n)))
    inp = refio.read_inputs(tmp)
    log(f"bench: SW-Si {n}^3, precision={args.precision}, temps={temps}", args.quiet)
    ctx = capi.Context(inp)
    nel = ctx.nibz * ctx.nk * ctx.nb ** 3
    lat = []
    stage = None
    for it in range(args.warmup + args.steps):
        t0 = time.perf_counter()
        ctx.eig(want=False)
        out = ctx.solve(temps, prec)
        dt = time.perf_counter() - t0
        if it >= args.warmup:
            lat.append(dt)
            stage = out["stage_ms"] if stage is None else stage + out["stage_ms"]
        log(f"step {it}: {dt * 1e3:.1f} ms  stages(ms)={[round(x, 2) for x in out['stage_ms']]} "
            f"kappa_RTA(300K)={out['kappa_rta'][min(1, len(temps) - 1)].sum(0).trace() / 3:.3f}", args.quiet)
    K = args.steps
    stage = stage / K
    ms_step = 1e3 * sum(lat) / K
    lat_sorted = sorted(lat)
    p50 = lat_sorted[len(lat) // 2] * 1e3
    p99 = lat_sorted[min(len(lat) - 1, int(round(0.99 * (len(lat) - 1))))] * 1e3
    # dominant kernel: the tetrahedron FGR kernel (one launch over the whole IBZ)
    fgr_ms = stage[2]
    fgr_bytes = nel * (8 + 16)   # read |V3|^2 (FP64), write P1sm + P2sm (FP64)
    fgr_gbs = fgr_bytes / (fgr_ms * 1e-3) / 1e9
    v3_ms = stage[1]
    res = {
        "metric": "3-ph V3 elements/s (mixed vs FP64) at kappa err <=0.5%; p50/p99 kappa-solve latency",
        "value": nel / (ms_step * 1e-3),
        "unit": "V3 elements/s",
        "n_gpus": 1, "steps": K, "warmup": args.warmup, "ms_per_step": ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": {"mixed": "fp32-mfma V3 / f64", "fp64": "f64", "fp64ref": "f64"}[args.precision],
        "data": "synthetic (Stillinger-Weber Si force constants by finite differences)",
        "config": {"workload": f"SW-Si 2-atom diamond, {n}^3 mesh, tetrahedron + CG BTE at {temps} K",
                   "nibz": ctx.nibz, "nk": ctx.nk, "nb": ctx.nb, "v3_elements": nel,
                   "precision": args.precision},
        "v3_kernel_elements_per_s": nel / (v3_ms * 1e-3),
        "stage_ms": {k: round(v, 3) for k, v in zip(
            ["eig", "v3", "fgr", "tau", "kappa_rta", "collision", "cg", "kappa_bte"], stage)},
        "latency_ms": {"p50": p50, "p99": p99},
        "roofline": {"bound": "hbm", "achieved": fgr_gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": fgr_gbs / PEAK_HBM_GBS, "traffic": None, "kernel": "k_fgr",
                     "algorithmic_bytes_per_launch": fgr_bytes, "kernel_ms": fgr_ms},
        "kappa_300K": {"rta": float(out["kappa_rta"][min(1, len(temps) - 1)].sum(0).trace() / 3),
                       "bte": float(out["kappa_bte"][min(1, len(temps) - 1)].sum(0).trace() / 3)},
    }
    ctx.close()
    if not args.no_cpu_baseline:
        cb = cpu_baseline_reference(tmp, args.quiet)
        if cb is not None:
            res["cpu_baseline"] = cb
    print(json.dumps(res))


if __name__ == "__main__":
    main()
