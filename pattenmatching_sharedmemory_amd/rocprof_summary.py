"""Summaries of rocprofv3 CSV output for the pruning kernels.

    python -m pattenmatching_sharedmemory_amd.rocprof_summary stats DIR
        per-kernel launch statistics from --kernel-trace --stats
        (DIR/*_kernel_stats.csv, DIR/*_kernel_trace.csv)
    python -m pattenmatching_sharedmemory_amd.rocprof_summary traffic \\
        --fetch DIR --write DIR --copy-fetch DIR --copy-write DIR --copy-bytes N
        HBM bytes per launch from two --pmc passes (FETCH_SIZE, WRITE_SIZE;
        KiB per dispatch), each scaled by what the same counter reads for a
        plain device copy of N bytes (the copy passes)
    python -m pattenmatching_sharedmemory_amd.rocprof_summary counters DIR NAME...
        mean per dispatch of each named counter of one --pmc pass, per kernel
        (e.g. TCC_REQ_sum, TCC_HIT_sum, TCC_MISS_sum: L2 requests, hits and
        misses)

Kernels are reported under the names pj_kernel_stats uses (lcc_phase1,
nlcc_walk, ...), so the figures line up with the live per-kernel timing.
"""
from __future__ import annotations

import argparse
import csv
import glob
import os
import sys
from collections import defaultdict

_TEMPLATE_NAMES = {
    "k_lcc0": "lcc0",
    "k_build_bins": "build_bins",
    "k_lcc_phase1_all": "lcc_phase1",
    "k_lcc_phase1_hubs": "lcc_phase1",  # the bins and the hub chunks in one launch
    "k_lcc_hub_chunks": "lcc_phase1_hub_chunks",
    "k_lcc_hub_finish": "lcc_phase1_hub_finish",
    "k_phase2_all": "phase2",
    "k_phase2_hub_finish": "phase2_hub_finish",
    "k_triangles": "triangles",
    "k_triangles_mid": "triangles_mid",
    "k_triangles_wide": "triangles_wide",
    "k_triangles_memo": "triangles_memo",
    "k_memo_lds": "memo_walk",
    "k_view_count": "active_view",
    "k_view_hubs": "active_view",
    "k_view_fill": "active_view",
    "k_scan_spans": "active_view",
    "k_scan_totals": "active_view",
    "k_scan_apply": "active_view",
}


def short_name(full: str) -> str:
    """'void (anonymous namespace)::k_triangles<unsigned int, 32>(...)' -> 'k_triangles<unsigned int, 32>'."""
    s = full.strip().replace("(anonymous namespace)::", "")
    if s.startswith("void "):
        s = s[5:]
    depth = 0
    for i, ch in enumerate(s):  # cut the argument list: the first '(' outside template brackets
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif ch == "(" and depth == 0:
            return s[:i].strip()
    return s


def stat_name(full: str, workgroup: int | None = None) -> str:
    """Map a kernel symbol to its pj_kernel_stats name; other symbols keep their short name."""
    s = short_name(full)
    base = s.split("<", 1)[0]
    args = s[len(base) + 1:-1].split(",") if "<" in s else []
    args = [a.strip() for a in args]
    if base in ("k_lcc_phase1_all", "k_lcc_phase1_hubs", "k_lcc_hub_chunks") and args and args[-1] == "true":
        return _TEMPLATE_NAMES[base] + "_labels"  # the first LCC step after LCC0 (slot labels)
    if base in _TEMPLATE_NAMES:
        return _TEMPLATE_NAMES[base]
    if base == "k_nlcc":
        return "nlcc_walk_hubs" if workgroup == 64 else "nlcc_walk"
    return s


def _one(pattern: str) -> str:
    hits = sorted(glob.glob(pattern, recursive=True))
    if not hits:
        raise SystemExit(f"no file matches {pattern}")
    return hits[-1]


def _rows(path: str):
    with open(path, newline="") as f:
        yield from csv.DictReader(f)


def kernel_stats(d: str) -> list[dict]:
    """Per stat-name totals from the kernel trace (falls back to the stats file)."""
    traces = glob.glob(os.path.join(d, "**", "*_kernel_trace.csv"), recursive=True)
    acc: dict[str, dict] = defaultdict(lambda: {"calls": 0, "ns": 0, "min": None, "max": 0})
    if traces:
        for r in _rows(sorted(traces)[-1]):
            wg = int(r.get("Workgroup_Size_X") or r.get("Workgroup_Size") or 0)
            n = stat_name(r["Kernel_Name"], wg)
            ns = int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
            a = acc[n]
            a["calls"] += 1
            a["ns"] += ns
            a["min"] = ns if a["min"] is None else min(a["min"], ns)
            a["max"] = max(a["max"], ns)
    else:
        for r in _rows(_one(os.path.join(d, "**", "*_kernel_stats.csv"))):
            n = stat_name(r["Name"])
            a = acc[n]
            a["calls"] += int(r["Calls"])
            a["ns"] += int(r["TotalDurationNs"])
            a["min"] = int(r["MinNs"]) if a["min"] is None else min(a["min"], int(r["MinNs"]))
            a["max"] = max(a["max"], int(r["MaxNs"]))
    total = sum(a["ns"] for a in acc.values()) or 1
    out = [{"name": n, "calls": a["calls"], "total_ms": a["ns"] / 1e6, "avg_us": a["ns"] / a["calls"] / 1e3,
            "min_us": a["min"] / 1e3, "max_us": a["max"] / 1e3, "pct": 100.0 * a["ns"] / total}
           for n, a in acc.items()]
    return sorted(out, key=lambda x: -x["total_ms"])


def pmc(d: str, counter: str) -> dict[str, dict]:
    """Mean counter value per dispatch, per stat name."""
    path = _one(os.path.join(d, "**", "*counter_collection.csv"))
    per_dispatch: dict[tuple, float] = defaultdict(float)
    names: dict[tuple, str] = {}
    for r in _rows(path):
        if r.get("Counter_Name") != counter:
            continue
        wg = int(r.get("Workgroup_Size_X") or r.get("Workgroup_Size") or 0)
        key = (r.get("Dispatch_Id"), r.get("Agent_Id"))
        per_dispatch[key] += float(r["Counter_Value"])
        names[key] = stat_name(r["Kernel_Name"], wg)
    acc: dict[str, list] = defaultdict(list)
    for k, v in per_dispatch.items():
        acc[names[k]].append(v)
    return {n: {"dispatches": len(v), "mean": sum(v) / len(v), "total": sum(v)} for n, v in acc.items()}


def _copy_factor(d: str, counter: str, nbytes: int) -> float:
    got = pmc(d, counter)
    # the copy is the largest dispatch of its run
    best = max(got.values(), key=lambda x: x["mean"])
    return nbytes / (best["mean"] * 1024.0)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("dir")
    t = sub.add_parser("traffic")
    t.add_argument("--fetch", required=True)
    t.add_argument("--write", required=True)
    t.add_argument("--copy-fetch", required=True)
    t.add_argument("--copy-write", required=True)
    t.add_argument("--copy-bytes", type=int, required=True)
    c = sub.add_parser("counters")
    c.add_argument("dir")
    c.add_argument("names", nargs="+")
    a = ap.parse_args(argv)
    if a.cmd == "counters":
        got = {n: pmc(a.dir, n) for n in a.names}
        kernels = sorted({k for g in got.values() for k in g},
                         key=lambda k: -max(g.get(k, {"total": 0.0})["total"] for g in got.values()))
        print("| kernel | dispatches | " + " | ".join(f"{n} per dispatch" for n in a.names) + " |")
        print("|---|---:|" + "---:|" * len(a.names))
        for k in kernels:
            d = max(g.get(k, {"dispatches": 0})["dispatches"] for g in got.values())
            vals = " | ".join(f"{got[n][k]['mean']:.4g}" if k in got[n] else "-" for n in a.names)
            print(f"| {k} | {d} | {vals} |")
        return 0
    if a.cmd == "stats":
        print("| kernel | launches | total ms | avg us | min us | max us | % |")
        print("|---|---:|---:|---:|---:|---:|---:|")
        for k in kernel_stats(a.dir):
            print(f"| {k['name']} | {k['calls']} | {k['total_ms']:.3f} | {k['avg_us']:.1f} | {k['min_us']:.1f} | "
                  f"{k['max_us']:.1f} | {k['pct']:.1f} |")
        return 0
    ff = _copy_factor(a.copy_fetch, "FETCH_SIZE", a.copy_bytes)
    wf = _copy_factor(a.copy_write, "WRITE_SIZE", a.copy_bytes)
    fetch, write = pmc(a.fetch, "FETCH_SIZE"), pmc(a.write, "WRITE_SIZE")
    print(f"copy calibration: {a.copy_bytes} bytes; FETCH_SIZE scale {ff:.4f} bytes per reported byte, "
          f"WRITE_SIZE scale {wf:.4f}")
    print()
    print("| kernel | dispatches | FETCH_SIZE KiB/dispatch | WRITE_SIZE KiB/dispatch | HBM bytes/launch (scaled) |")
    print("|---|---:|---:|---:|---:|")
    rows = []
    for n in sorted(set(fetch) | set(write)):
        f = fetch.get(n, {"mean": 0.0, "dispatches": 0})
        w = write.get(n, {"mean": 0.0, "dispatches": 0})
        b = f["mean"] * 1024.0 * ff + w["mean"] * 1024.0 * wf
        rows.append((b, n, f, w))
    for b, n, f, w in sorted(rows, key=lambda x: -x[0]):
        print(f"| {n} | {f['dispatches']} | {f['mean']:.1f} | {w['mean']:.1f} | {b:.4g} |")
    print()
    for b, n, f, w in sorted(rows, key=lambda x: -x[0]):
        print(f"traffic: {n} {b:.6g}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
