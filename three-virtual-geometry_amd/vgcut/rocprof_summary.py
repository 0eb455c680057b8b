"""Summaries of rocprofv3 results for the cut's kernels: per-kernel dispatch times and HBM traffic.

rocprofv3 (ROCm 7) writes a SQLite database (`<name>_results.db`) with a `kernels` view (one row per
dispatch: name, start, end) and, for `--pmc` runs, a `counters_collection` view (dispatch_id,
kernel_name, counter_name, value). This module reads both and writes the short summaries kept under
profiles/:

    python -m vgcut.rocprof_summary --trace T.db --fetch F.db --write W.db --copy-bytes N \\
        --command "..." --markdown out.md --json out.json

HBM traffic: FETCH_SIZE and WRITE_SIZE are reported in KiB per dispatch. Their scale is calibrated on a
plain device copy of `--copy-bytes` bytes read and the same written (the largest `copyBuffer` dispatch
of the run): factor = true bytes / counted bytes, per counter. A kernel's traffic per dispatch is
FETCH_SIZE * 1024 * fetch_factor + WRITE_SIZE * 1024 * write_factor, averaged over its busy dispatches
(those above 10% of its largest one, which leaves out the retry rounds that find nothing to redo).
"""
from __future__ import annotations

import argparse
import collections
import json
import sqlite3
from typing import Dict, List

COPY_KERNEL = "copyBuffer"


def short_name(name: str) -> str:
    """`vg::k_meshlets<6>(vg::Params)` -> `k_meshlets<6>`; other names are cut at the argument list."""
    base = name.split("(")[0]
    return base.replace("void ", "").replace("vg::", "").strip()


def kernel_times(db: str) -> Dict[str, List[int]]:
    """Dispatch durations in ns per kernel name."""
    con = sqlite3.connect(db)
    rows = con.execute("select name, start, end from kernels").fetchall()
    out: Dict[str, List[int]] = collections.defaultdict(list)
    for name, start, end in rows:
        out[name].append(int(end) - int(start))
    return out


def counter_values(db: str, counter: str) -> Dict[str, List[float]]:
    """Per kernel name: the counter's value of each dispatch (summed over its dimensions)."""
    con = sqlite3.connect(db)
    rows = con.execute("select dispatch_id, kernel_name, value from counters_collection where counter_name = ?",
                       (counter,)).fetchall()
    per_dispatch: Dict[tuple, float] = collections.defaultdict(float)
    for did, name, value in rows:
        per_dispatch[(did, name)] += float(value)
    out: Dict[str, List[float]] = collections.defaultdict(list)
    for (did, name), v in sorted(per_dispatch.items()):
        out[name].append(v)
    return out


def busy(values: List[float]) -> List[float]:
    top = max(values) if values else 0.0
    return [v for v in values if v > 0.1 * top]


def calibration(values: Dict[str, List[float]], true_bytes: float) -> float:
    copies = [v for name, vs in values.items() if COPY_KERNEL in name for v in vs]
    if not copies:
        raise ValueError("no copyBuffer dispatch to calibrate against")
    return true_bytes / (max(copies) * 1024.0)


def summarize(trace_db: str, fetch_db: str, write_db: str, copy_bytes: float) -> dict:
    times = kernel_times(trace_db)
    fetch = counter_values(fetch_db, "FETCH_SIZE")
    write = counter_values(write_db, "WRITE_SIZE")
    f_fetch = calibration(fetch, copy_bytes)
    f_write = calibration(write, copy_bytes)
    kernels = {}
    for name, durs in times.items():
        if not short_name(name).startswith("k_"):
            continue
        b = busy([float(d) for d in durs])
        fb, wb = busy(fetch.get(name, [])), busy(write.get(name, []))
        entry = {"dispatches": len(durs), "busy_dispatches": len(b), "avg_us": sum(b) / len(b) / 1e3 if b else 0.0,
                 "total_ms": sum(durs) / 1e6}
        if fb and wb:
            fetch_bytes = sum(fb) / len(fb) * 1024.0 * f_fetch
            write_bytes = sum(wb) / len(wb) * 1024.0 * f_write
            entry.update({"fetch_kib": sum(fb) / len(fb), "write_kib": sum(wb) / len(wb), "fetch_bytes": fetch_bytes,
                          "write_bytes": write_bytes, "hbm_bytes": fetch_bytes + write_bytes})
        kernels[short_name(name)] = entry
    return {"fetch_factor": f_fetch, "write_factor": f_write, "copy_bytes": copy_bytes, "kernels": kernels}


def add_counters(summary: dict, db: str) -> None:
    """Per kernel, the busy-dispatch average of every counter in a further --pmc database (instruction mix)."""
    con = sqlite3.connect(db)
    names = [r[0] for r in con.execute("select distinct counter_name from counters_collection").fetchall()]
    for counter in sorted(names):
        for name, values in counter_values(db, counter).items():
            k = summary["kernels"].get(short_name(name))
            if k is None or not values:
                continue
            b = busy(values) or values
            k.setdefault("counters", {})[counter] = sum(b) / len(b)


def markdown(summary: dict, title: str, commands: List[str]) -> str:
    lines = [f"# {title}", "", "Commands (from the repository root):", ""]
    lines += [f"    {c}" for c in commands]
    lines += ["", f"Counter calibration on a {summary['copy_bytes'] / 2**30:.2f} GiB device copy: FETCH_SIZE x "
                  f"{summary['fetch_factor']:.3f}, WRITE_SIZE x {summary['write_factor']:.3f}.", "",
              "| kernel | dispatches (busy) | avg busy dispatch (us) | FETCH_SIZE (KiB) | WRITE_SIZE (KiB) | HBM bytes / dispatch |",
              "| --- | --- | --- | --- | --- | --- |"]
    for name, k in sorted(summary["kernels"].items(), key=lambda kv: -kv[1]["total_ms"]):
        if "hbm_bytes" in k:
            lines.append(f"| {name} | {k['dispatches']} ({k['busy_dispatches']}) | {k['avg_us']:.1f} | {k['fetch_kib']:.0f} | "
                         f"{k['write_kib']:.0f} | {k['hbm_bytes'] / 1e6:.1f} MB |")
        else:
            lines.append(f"| {name} | {k['dispatches']} ({k['busy_dispatches']}) | {k['avg_us']:.1f} | - | - | - |")
    counters = sorted({c for k in summary["kernels"].values() for c in k.get("counters", {})})
    if counters:
        lines += ["", "Counters per busy dispatch (summed over the device; SQ_* per wave instruction or cycle, TA_/TCP_ summed over CUs):", "",
                  "| kernel | " + " | ".join(counters) + " |", "| --- |" + " --- |" * len(counters)]
        for name, k in sorted(summary["kernels"].items(), key=lambda kv: -kv[1]["total_ms"]):
            if "counters" in k:
                lines.append(f"| {name} | " + " | ".join(f"{k['counters'].get(c, 0):.3g}" for c in counters) + " |")
    return "\n".join(lines) + "\n"


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--trace", required=True, help="results db of the --kernel-trace --stats run")
    ap.add_argument("--fetch", required=True, help="results db of the --pmc FETCH_SIZE run")
    ap.add_argument("--write", required=True, help="results db of the --pmc WRITE_SIZE run")
    ap.add_argument("--copy-bytes", type=float, required=True, help="bytes read (= written) by the calibration copy")
    ap.add_argument("--counters", action="append", default=[], help="further --pmc databases (instruction mix)")
    ap.add_argument("--title", default="rocprofv3 summary")
    ap.add_argument("--command", action="append", default=[], help="the commands that produced the databases")
    ap.add_argument("--markdown", required=True)
    ap.add_argument("--json", required=True)
    args = ap.parse_args()
    s = summarize(args.trace, args.fetch, args.write, args.copy_bytes)
    for db in args.counters:
        add_counters(s, db)
    s["commands"] = args.command
    with open(args.markdown, "w") as f:
        f.write(markdown(s, args.title, args.command))
    with open(args.json, "w") as f:
        json.dump(s, f, indent=1, sort_keys=True)
        f.write("\n")
    print(markdown(s, args.title, args.command))


if __name__ == "__main__":
    main()
