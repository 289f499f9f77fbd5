"""Summaries of rocprofv3 output for the kernels of this library.

    python -m thermacond.profiling stats  <trace_kernel_stats.csv>            # kernel time table
    python -m thermacond.profiling pmc    <counter_collection.csv> [...]      # counters per launch
    python -m thermacond.profiling traffic <fetch.csv> <write.csv> <copy_fetch.csv> <copy_write.csv> <copy_bytes>

`stats` prints a markdown table (calls, total ms, share, average per launch) of a
`--kernel-trace --stats` run.  `pmc` averages each counter over the dispatches of each kernel
(a `--pmc` run with `-f csv`).  `traffic` turns FETCH_SIZE / WRITE_SIZE (KiB per dispatch) into
HBM bytes per launch, each scaled by what the same counter reads for a plain device copy of
known size (bench.py --copy-calibration), and prints the `traffic_bytes_per_launch` JSON line
that bench.py reads back from the committed summary.
"""
from __future__ import annotations

import collections
import csv
import json
import sys
from typing import Dict, Tuple


def kernel_name(n: str) -> str:
    """Demangled name without namespaces' noise and argument list (template arguments kept).
    A name the profiler left mangled (its demangler gives up on _Float16 vector arguments) is
    reduced to the function's own name: _ZN12_GLOBAL__N_1<len><name>..."""
    if n.startswith("_ZN"):
        rest = n[3:]
        if rest.startswith("12_GLOBAL__N_1"):
            rest = rest[len("12_GLOBAL__N_1"):]
        digits = len(rest) - len(rest.lstrip("0123456789"))
        if digits:
            ln = int(rest[:digits])
            return rest[digits:digits + ln]
    n = n.replace("(anonymous namespace)::", "").replace("alt3::", "").replace("void ", "")
    depth, out = 0, ""
    for ch in n:
        if ch == "(" and depth == 0:
            break
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        out += ch
    return out.strip()


def stats_table(path: str, top: int = 20) -> str:
    rows = list(csv.DictReader(open(path)))
    tot = sum(float(r["TotalDurationNs"]) for r in rows) or 1.0
    out = ["| kernel | calls | total ms | % | avg us/launch |", "|---|---:|---:|---:|---:|"]
    for r in sorted(rows, key=lambda r: -float(r["TotalDurationNs"]))[:top]:
        t = float(r["TotalDurationNs"])
        out.append(f"| `{kernel_name(r['Name'])}` | {r['Calls']} | {t / 1e6:.3f} | {100 * t / tot:.2f} | "
                   f"{float(r['AverageNs']) / 1e3:.1f} |")
    return "\n".join(out)


def pmc_per_launch(path: str) -> Dict[str, Tuple[Dict[str, float], int, float]]:
    """kernel -> ({counter: mean per dispatch}, dispatches, mean duration s)."""
    acc = collections.defaultdict(lambda: collections.defaultdict(float))
    seen = collections.defaultdict(set)
    dur = collections.defaultdict(float)
    for r in csv.DictReader(open(path)):
        k = kernel_name(r["Kernel_Name"])
        acc[k][r["Counter_Name"]] += float(r["Counter_Value"])
        if r["Dispatch_Id"] not in seen[k]:
            dur[k] += (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-9
            seen[k].add(r["Dispatch_Id"])
    return {k: ({c: x / len(seen[k]) for c, x in v.items()}, len(seen[k]), dur[k] / len(seen[k]))
            for k, v in acc.items()}


def pmc_table(paths, min_ms: float = 0.1) -> str:
    merged: Dict[str, Dict[str, float]] = collections.defaultdict(dict)
    launch: Dict[str, Tuple[int, float]] = {}
    for p in paths:
        for k, (v, n, t) in pmc_per_launch(p).items():
            merged[k].update(v)
            launch[k] = (n, t)
    counters = sorted({c for v in merged.values() for c in v})
    out = ["| kernel | launches | ms/launch | " + " | ".join(counters) + " |",
           "|---|---:|---:|" + "---:|" * len(counters)]
    for k in sorted(merged, key=lambda k: -launch[k][1]):
        n, t = launch[k]
        if t * 1e3 < min_ms:
            continue
        out.append(f"| `{k}` | {n} | {t * 1e3:.3f} | " +
                   " | ".join(f"{merged[k].get(c, float('nan')):.4g}" for c in counters) + " |")
    return "\n".join(out)


def traffic(fetch: str, write: str, copy_fetch: str, copy_write: str, copy_bytes: int, min_ms: float = 0.1):
    """(table, {kernel: HBM bytes per launch}) with FETCH/WRITE scaled by the copy calibration."""
    F, W = pmc_per_launch(fetch), pmc_per_launch(write)
    # the copy kernel is the launch whose FETCH_SIZE is largest in the calibration run
    cfk = max(pmc_per_launch(copy_fetch).items(), key=lambda kv: kv[1][0].get("FETCH_SIZE", 0.0))
    cwk = max(pmc_per_launch(copy_write).items(), key=lambda kv: kv[1][0].get("WRITE_SIZE", 0.0))
    f_scale = copy_bytes / (cfk[1][0]["FETCH_SIZE"] * 1024.0)
    w_scale = copy_bytes / (cwk[1][0]["WRITE_SIZE"] * 1024.0)
    rows = ["| kernel | launches | ms/launch | FETCH KiB | WRITE KiB | read GB | written GB | HBM GB/launch | GB/s |",
            "|---|---:|---:|---:|---:|---:|---:|---:|---:|"]
    table = {}
    for k in sorted(F, key=lambda k: -F[k][2]):
        if k not in W or F[k][2] * 1e3 < min_ms:
            continue
        fk, wk = F[k][0].get("FETCH_SIZE", 0.0), W[k][0].get("WRITE_SIZE", 0.0)
        rb, wb = fk * 1024.0 * f_scale, wk * 1024.0 * w_scale
        t = 0.5 * (F[k][2] + W[k][2])
        table[k] = rb + wb
        rows.append(f"| `{k}` | {F[k][1]} | {t * 1e3:.3f} | {fk:.4g} | {wk:.4g} | {rb / 1e9:.3f} | {wb / 1e9:.3f} | "
                    f"{(rb + wb) / 1e9:.3f} | {(rb + wb) / t / 1e9:.0f} |")
    head = (f"Calibration: a {copy_bytes} B device copy reads FETCH_SIZE x 1024 x {f_scale:.3f} B and "
            f"WRITE_SIZE x 1024 x {w_scale:.3f} B.")
    return head + "\n\n" + "\n".join(rows), table


def main(argv):
    if len(argv) >= 2 and argv[0] == "stats":
        print(stats_table(argv[1]))
    elif len(argv) >= 2 and argv[0] == "pmc":
        print(pmc_table(argv[1:]))
    elif len(argv) == 6 and argv[0] == "traffic":
        text, table = traffic(argv[1], argv[2], argv[3], argv[4], int(argv[5]))
        print(text)
        print()
        print("traffic_bytes_per_launch " + json.dumps({k: round(v) for k, v in table.items()}))
    else:
        print(__doc__)
        return 2
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
