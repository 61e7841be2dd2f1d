"""Summaries of rocprofv3 CSV output for the sweep kernels.

    python -m pscbfm_amd.profiling --stats DIR [--fetch DIR --write DIR]
        [--copy-bytes N] [--md OUT.md] [--json OUT.json] [--command "..."]

--stats   directory of a `rocprofv3 --kernel-trace --stats --output-format csv` run
--fetch   directory of a `rocprofv3 --pmc FETCH_SIZE --kernel-trace --output-format csv` run
--write   directory of a `rocprofv3 --pmc WRITE_SIZE --kernel-trace --output-format csv` run

HBM traffic per launch = FETCH_SIZE and WRITE_SIZE (KiB per dispatch) of a
kernel, each scaled by the factor that makes the same counter of the plain
copy kernel (k_copy_u4, which reads and writes exactly --copy-bytes) read
--copy-bytes.
"""
from __future__ import annotations

import argparse
import csv
import json
import re
from collections import defaultdict
from pathlib import Path
from typing import Dict, Optional

KERNELS = {"k_check_batch": "pass 1, batch (k_check_batch)", "k_perform_batch": "pass 2, batch (k_perform_batch)",
           "k_advance_batch": "counter advance, batch", "k_gather_state": "state gather, batch",
           "k_set_stops": "stop points, batch", "k_check": "pass 1 (k_check)", "k_perform": "pass 2 (k_perform)",
           "k_copy_u4": "copy reference", "k_advance": "counter advance", "k_fill_u32": "fill",
           "k_unfold": "coordinate decode (read-back)", "k_perform_lists": "pass 2, several lists (k_perform_lists)",
           "k_unfold_jobs": "coordinate decode, batch (read-back)"}


_NAME = re.compile(r"\b(k_[a-z0-9_]+?)(?=[<(]|$)")
_LATTICE = re.compile(r"\b(k_(?:check|perform)(?:_batch)?)<(\d+)")
SWEEP = ("k_check", "k_perform", "k_check_batch", "k_perform_batch")


def _short(name: str) -> str:
    """Kernel symbol without namespace, template arguments or parameters.
    The sweep kernels keep their lattice instance: `k_check` is the bit
    lattice, `k_check.u8` the byte lattice (first template argument 1 or 0),
    so one run that times both never mixes them."""
    m = _LATTICE.search(name)
    if m:
        return m.group(1) + ("" if m.group(2) == "1" else ".u8")
    m = _NAME.search(name)
    if m and m.group(1) in KERNELS:
        return m.group(1)
    return name.split("(")[0].split("::")[-1][-40:] or name[:40]


def label(key: str) -> str:
    base, _, kind = key.partition(".")
    text = KERNELS.get(base, key)
    return text + ", uint8 lattice" if kind == "u8" else text


def _find(d: Path, suffix: str) -> Path:
    hits = sorted(Path(d).rglob(f"*{suffix}"))
    if not hits:
        raise FileNotFoundError(f"no *{suffix} under {d}")
    return hits[0]


def _tail(v: list, last: Optional[int]) -> list:
    return v[-last:] if last else v


def kernel_durations(stats_dir: Path, last: Optional[int] = None) -> Dict[str, dict]:
    """Per kernel: calls, total and mean duration (us) from the kernel trace
    (mean over the last `last` dispatches of each kernel when given)."""
    rows = sorted(csv.DictReader(open(_find(stats_dir, "kernel_trace.csv"))), key=lambda r: int(r["Start_Timestamp"]))
    acc: Dict[str, list] = defaultdict(list)
    for r in rows:
        acc[_short(r["Kernel_Name"])].append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3)
    out = {}
    for k, v in acc.items():
        t = _tail(v, last if k not in ("k_copy_u4",) else None)
        out[k] = {"calls": len(v), "total_us": sum(v), "mean_us": sum(t) / len(t), "mean_over": len(t)}
    return out


def counter_means(pmc_dir: Path, counter: str, last: Optional[int] = None) -> Dict[str, float]:
    rows = sorted(csv.DictReader(open(_find(pmc_dir, "counter_collection.csv"))), key=lambda r: int(r["Dispatch_Id"]))
    acc: Dict[str, list] = defaultdict(list)
    for r in rows:
        if r["Counter_Name"] == counter:
            acc[_short(r["Kernel_Name"])].append(float(r["Counter_Value"]))
    return {k: (lambda t: sum(t) / len(t))(_tail(v, last if k != "k_copy_u4" else None)) for k, v in acc.items()}


def all_counters(pmc_dir: Path, last: Optional[int] = None) -> Dict[str, Dict[str, float]]:
    """Per kernel, the mean of every counter of one --pmc run."""
    rows = sorted(csv.DictReader(open(_find(pmc_dir, "counter_collection.csv"))), key=lambda r: int(r["Dispatch_Id"]))
    acc: Dict[str, Dict[str, list]] = defaultdict(lambda: defaultdict(list))
    for r in rows:
        acc[_short(r["Kernel_Name"])][r["Counter_Name"]].append(float(r["Counter_Value"]))
    return {k: {c: (lambda t: sum(t) / len(t))(_tail(v, last)) for c, v in cs.items()} for k, cs in acc.items()}


def traffic(fetch_dir: Path, write_dir: Path, copy_bytes: int, last: Optional[int] = None) -> dict:
    fetch = counter_means(fetch_dir, "FETCH_SIZE", last)
    write = counter_means(write_dir, "WRITE_SIZE", last)
    kib = copy_bytes / 1024.0
    f_r = kib / fetch["k_copy_u4"]
    f_w = kib / write["k_copy_u4"]
    out = {"copy_bytes": copy_bytes, "fetch_scale": f_r, "write_scale": f_w, "kernels": {}}
    for k in fetch:
        if k in write:
            r = fetch[k] * f_r * 1024.0
            w = write[k] * f_w * 1024.0
            out["kernels"][k] = {"fetch_kib": fetch[k], "write_kib": write[k], "read_bytes": r, "write_bytes": w,
                                 "traffic_bytes": r + w}
    return out


def main(argv: Optional[list] = None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--stats", type=Path, required=True)
    ap.add_argument("--fetch", type=Path)
    ap.add_argument("--write", type=Path)
    ap.add_argument("--copy-bytes", type=int, default=1 << 29)
    ap.add_argument("--counters", type=Path, nargs="*", default=[],
                    help="directories of further --pmc runs; their counters are tabulated per sweep kernel")
    ap.add_argument("--md", type=Path)
    ap.add_argument("--json", type=Path)
    ap.add_argument("--command", default="")
    ap.add_argument("--title", default="rocprofv3 summary")
    ap.add_argument("--last", type=int, default=None,
                    help="average only the last N dispatches of each sweep kernel (steady state)")
    args = ap.parse_args(argv)
    dur = kernel_durations(args.stats, args.last)
    tr = traffic(args.fetch, args.write, args.copy_bytes, args.last) if args.fetch and args.write else None
    lines = [f"# {args.title}", ""]
    if args.command:
        lines += ["Commands (run from the repository root on one MI355X):", "", "```", args.command, "```", ""]
    stats_csv = sorted(Path(args.stats).rglob("*kernel_stats.csv"))
    if stats_csv:
        lines += ["rocprofv3 --stats (all dispatches of the run):", "",
                  "| kernel | calls | average (us) | min (us) | max (us) | share % |", "|---|---|---|---|---|---|"]
        agg: Dict[str, dict] = {}
        for r in csv.DictReader(open(stats_csv[0])):  # template instances summed per kernel
            a = agg.setdefault(_short(r["Name"]), {"calls": 0, "total": 0.0, "min": float("inf"), "max": 0.0,
                                                   "share": 0.0})
            a["calls"] += int(r["Calls"])
            a["total"] += float(r["AverageNs"]) * int(r["Calls"])
            a["min"] = min(a["min"], float(r["MinNs"]))
            a["max"] = max(a["max"], float(r["MaxNs"]))
            a["share"] += float(r["Percentage"])
        for k, a in sorted(agg.items(), key=lambda kv: -kv[1]["share"]):
            lines.append(f"| {label(k)} | {a['calls']} | {a['total'] / a['calls'] / 1e3:.2f} | "
                         f"{a['min'] / 1e3:.2f} | {a['max'] / 1e3:.2f} | {a['share']:.1f} |")
        lines.append("")
    lines += ["Kernel trace, steady state (mean over the last dispatches of each sweep kernel):", "",
              "| kernel | calls | mean duration (us) | over last | total (us) |", "|---|---|---|---|---|"]
    for k, v in sorted(dur.items(), key=lambda kv: -kv[1]["total_us"]):
        lines.append(f"| {label(k)} | {v['calls']} | {v['mean_us']:.2f} | {v['mean_over']} | "
                     f"{v['total_us']:.1f} |")
    if tr:
        lines += ["", f"HBM traffic per launch (copy reference {tr['copy_bytes']} B: FETCH_SIZE scale "
                      f"{tr['fetch_scale']:.3f}, WRITE_SIZE scale {tr['write_scale']:.3f}):", "",
                  "| kernel | FETCH_SIZE KiB | WRITE_SIZE KiB | read MB | write MB | total MB |", "|---|---|---|---|---|---|"]
        for k, v in tr["kernels"].items():
            lines.append(f"| {label(k)} | {v['fetch_kib']:.0f} | {v['write_kib']:.0f} | "
                         f"{v['read_bytes'] / 1e6:.1f} | {v['write_bytes'] / 1e6:.1f} | {v['traffic_bytes'] / 1e6:.1f} |")
    counters: Dict[str, Dict[str, float]] = defaultdict(dict)
    for d in args.counters:
        for k, cs in all_counters(d, args.last).items():
            counters[k].update(cs)
    sweep = sorted((k for k in counters if k.partition(".")[0] in SWEEP), key=lambda k: (SWEEP.index(k.partition(".")[0]), k))
    if sweep:
        names = sorted({c for k in sweep for c in counters[k]})
        lines += ["", "Counters per dispatch (mean over the same dispatches):", "",
                  "| counter | " + " | ".join(label(k) for k in sweep) + " |",
                  "|---|" + "---|" * len(sweep)]
        for c in names:
            lines.append(f"| {c} | " + " | ".join(f"{counters[k].get(c, float('nan')):.4g}" for k in sweep) + " |")
        derived = []
        for k in sweep:
            cs = counters[k]
            row = {}
            if "TCC_HIT_sum" in cs and "TCC_MISS_sum" in cs and cs["TCC_HIT_sum"] + cs["TCC_MISS_sum"] > 0:
                row["L2 hit rate"] = cs["TCC_HIT_sum"] / (cs["TCC_HIT_sum"] + cs["TCC_MISS_sum"])
            if "SQ_WAIT_INST_ANY" in cs and cs.get("SQ_WAVE_CYCLES", 0) > 0:
                row["wave cycles waiting on memory"] = cs["SQ_WAIT_INST_ANY"] / cs["SQ_WAVE_CYCLES"]
            if "TCP_TCC_READ_REQ_LATENCY_sum" in cs and cs.get("TCP_TCC_READ_REQ_sum", 0) > 0:
                row["mean L1->L2 read latency (cycles)"] = cs["TCP_TCC_READ_REQ_LATENCY_sum"] / cs["TCP_TCC_READ_REQ_sum"]
            derived.append(row)
        keys = sorted({c for r in derived for c in r})
        if keys:
            lines += ["", "| derived | " + " | ".join(label(k) for k in sweep) + " |", "|---|" + "---|" * len(sweep)]
            for c in keys:
                lines.append(f"| {c} | " + " | ".join(f"{r.get(c, float('nan')):.3f}" for r in derived) + " |")
    text = "\n".join(lines) + "\n"
    if args.md:
        args.md.write_text(text)
    else:
        print(text)
    if args.json:
        args.json.write_text(json.dumps({"durations": dur, "traffic": tr, "counters": counters}, indent=1) + "\n")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
