"""Benchmark: plain PruneJuice pruning pipeline on one MI355X.

Workload (the single-GPU plain configuration): R-MAT scale 22, edge factor 16,
(a, b, c, d) = (0.57, 0.19, 0.19, 0.05), symmetrised / deduplicated / no
self-loops (4,194,304 vertices, 128,305,988 directed CSR slots), 32 labels
drawn from Zipf(s=1.2); template = labelled triangle t0-t1-t2 plus the tail
t2-t3-t4-t5 with labels [0, 1, 2, 3, 4, 3] (one cycle and one path
constraint).  One bench step = one whole pipeline solve on the HBM-resident
graph: state reset, LCC0, LCC to fixed point, CC, LCC, PC, LCC to the
pruning fixed point (enumeration is not part of the timed solve).

value = pruning TEPS = edge sweeps per second: every step of the solve works
on the active graph, so a solve sweeps the sum over its steps of the active
CSR slots at the start of the step; value = that sum / solve time.  The sum
comes from the step trace alone (it is the same for every implementation
that produces the same trace), so the CPU baseline uses the same definition.
teps_edge_sweeps repeats value under its own name; teps_slots_examined =
the CSR slots the kernels examined (a serial scan in the reference's loop
order, binary searches counting their probes) per second.  roofline =
the dominant kernel's algorithmic bytes per launch / its average launch time
(HIP events on the context stream) against 8 TB/s; traffic = HBM bytes per
launch of that kernel from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
summary of the same workload (profiles/), when there is one.  cpu_baseline =
the oracle (CPU restatement of the reference) solving the same scale-22
graph and template once, on every CPU the process may use (OMP_NUM_THREADS
when set), with the host's model and core counts.

The plain workload also reports the dynamic constraint order (the
reference's effectiveness-driven order, ORDER_USE_DYNAMIC): its solve time
and choices, and the label statistics pass it takes at each LCC fixed point
against the HBM roofline (label_stat).

--copy-calibration MIB runs only a device-to-device copy of MIB MiB three
times (the plain copy kernel the PMC byte counters are scaled by).

Prints one JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "pruning TEPS + HBM roofline %; batched/served queries/s, p50/p99 vs plain"
HBM_PEAK_GBS = 8000.0

GRAPH_SEED = 20221
LABEL_SEED = 7

# Named workloads.  "plain": the single-GPU plain pipeline (the metric's
# configuration).  "tds": the memory-scale configuration, whose 8-vertex
# template (two 4-cycles sharing the edge t1-t2, plus the chord t0-t2 and a
# tail) needs template-driven search.
WORKLOADS = {
    "plain": {"scale": 22, "labels": 32, "zipf": 1.2,
              "edges": [(0, 1), (1, 2), (2, 0), (2, 3), (3, 4), (4, 5)], "template_labels": [0, 1, 2, 3, 4, 3],
              "text": "triangle+tail template [0,1,2,3,4,3]"},
    "tds": {"scale": 27, "labels": 64, "zipf": 1.0,
            "edges": [(0, 1), (1, 2), (2, 3), (3, 0), (1, 4), (4, 5), (5, 2), (0, 2), (5, 6), (6, 7)],
            "template_labels": [8, 9, 10, 11, 12, 13, 14, 15],
            "text": "8-vertex template: 4-cycles t0-t1-t2-t3 and t1-t4-t5-t2, chord t0-t2, tail t5-t6-t7, "
                    "labels of Zipf ranks 8..15"},
}
TEMPLATE_EDGES = WORKLOADS["plain"]["edges"]
TEMPLATE_LABELS = WORKLOADS["plain"]["template_labels"]
STEP_NAMES = {2: "LCC0", 3: "LCC", 4: "CC", 5: "PC", 6: "TDS", 7: "ENUMERATION"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload(scale: int, name: str = "plain"):
    from pattenmatching_sharedmemory_amd import Constraints, Graph

    w = WORKLOADS[name]
    t0 = time.time()
    g = Graph.rmat(scale, 16, 0.57, 0.19, 0.19, seed=GRAPH_SEED).labels_zipf(w["labels"], w["zipf"], seed=LABEL_SEED)
    p = Graph.pattern(len(w["template_labels"]), w["edges"], w["template_labels"])
    c = Constraints(p)
    log(f"[bench] graph scale {scale}: V={g.vertex_count} E={g.edge_count} built in {time.time() - t0:.1f}s; "
        f"cycles={c.cycles()} paths={c.paths()} tds={len(c.tds())}")
    return g, p, c


def algorithmic_bytes(k: dict, vertices: int) -> float:
    """Bytes a kernel must move, counted once per access at natural width with
    no cache credit (W = candidate word bytes): offsets 8, ids 4, labels 1,
    flags 1, edge-active 1/8 per slot read or written.
      phase 1 (LCC): items*(9.25 + 2W) + slots*(5.25 + W) + 0.125*eliminated
      phase 1 of the first step after LCC0 (*_labels: it reads a slot's
                     label byte where the others read the neighbour id, 4,
                     and candidate word, W): items*(9.25 + 2W) + slots*1.25
                     + 0.125*eliminated
      phase 2:       items*(3W + 0.5) + 4.25*slots
      NLCC walks (nlcc_walk*, triangles*, memo_walk): 8*items + (4.25 + W)*slots
      bins:          vertices*(W + 9) + 4*listed"""
    W = k["word_bytes"]
    name = k["name"]
    if name.startswith("lcc_phase1") and name.endswith("_labels"):
        return k["items"] * (9.25 + 2 * W) + k["slots"] * 1.25 + 0.125 * k["edges_eliminated"]
    if name.startswith("lcc_phase1"):
        return k["items"] * (9.25 + 2 * W) + k["slots"] * (5.25 + W) + 0.125 * k["edges_eliminated"]
    if name.startswith("phase2"):
        return k["items"] * (3 * W + 0.5) + 4.25 * k["slots"]
    if name.startswith("nlcc") or name.startswith("triangles") or name == "memo_walk":
        return 8 * k["items"] + (4.25 + W) * k["slots"]
    if name == "build_bins":
        return k["launches"] * vertices * (W + 9)
    if name == "lcc0":
        return k["launches"] * vertices * (2 * W + 1.125)
    return 0.0


def roofline_from(stats: list[dict], vertices: int) -> dict | None:
    if not stats:
        return None
    dom = max(stats, key=lambda k: k["ms"])
    nbytes = algorithmic_bytes(dom, vertices)
    per_launch = nbytes / dom["launches"]
    avg_ms = dom["ms"] / dom["launches"]
    achieved = per_launch / (avg_ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": dom["name"],
            "bytes_per_launch": per_launch, "avg_launch_ms": avg_ms, "launches": dom["launches"]}


TRAFFIC_FILE = ROOT / "profiles" / "pruning_rmat22_hbm_traffic.md"


def measured_traffic(kernel: str) -> float | None:
    """HBM bytes per launch of `kernel` from the committed PMC summary."""
    if not TRAFFIC_FILE.exists():
        return None
    for line in TRAFFIC_FILE.read_text().splitlines():
        parts = line.split()
        if len(parts) == 3 and parts[0] == "traffic:" and parts[1] == kernel:
            return float(parts[2])
    return None


def copy_calibration(mib: int) -> int:
    import torch

    n = mib << 20
    src = torch.ones(n, dtype=torch.uint8, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": n, "copies": 3}), flush=True)
    return 0


def edge_sweeps(rows, edge_count: int) -> int:
    """Sum over the steps of the active CSR slots at the start of each step."""
    total, active = 0, edge_count
    for r in rows:
        total += active
        active = r.active_edges
    return total


def phases(rows, edge_count: int) -> dict:
    """Per step kind: steps, device ms, edge sweeps and sweeps per second."""
    out: dict[str, dict] = {}
    active = edge_count
    for r in rows:
        ph = out.setdefault(STEP_NAMES.get(r.step, str(r.step)), {"steps": 0, "ms": 0.0, "edge_sweeps": 0})
        ph["steps"] += 1
        ph["ms"] += r.ms
        ph["edge_sweeps"] += active
        active = r.active_edges
    for ph in out.values():
        ph["ms"] = round(ph["ms"], 4)
        ph["TEPS"] = ph["edge_sweeps"] / (ph["ms"] * 1e-3) if ph["ms"] else None
    return out


def label_stat_pass(pr, g, reps: int = 20) -> dict:
    """The label statistics pass (pj_graph_stat) of the dynamic constraint
    order on the state of its first decision (LCC0, then LCC to the fixed
    point): device time of the pass (bins of every active vertex, the
    counting launch and the hub finish) and its algorithmic bytes, counted
    once at natural width: per active vertex its bin entry 4, offsets 16,
    label code 1 and candidate word W; per slot of an active vertex its id
    4, edge-active bit 1/8, the neighbour's candidate word W and code 1."""
    import numpy as np

    from pattenmatching_sharedmemory_amd import STEP_LCC, STEP_LCC0

    pr.reset()
    pr.step(STEP_LCC0)
    while not pr.step(STEP_LCC).is_empty():
        pass
    st = pr.state()
    active = (st.vflags & 1).astype(bool)
    deg = np.diff(g.offsets().astype(np.int64))
    n_v, slots = int(active.sum()), int(deg[active].sum())
    pr.kernel_timing(True, only="label_stat")
    for _ in range(reps):
        pr.graph_stat()
    k = [x for x in pr.kernel_stats() if x["name"] == "label_stat"][0]
    pr.kernel_timing(False)
    W = k["word_bytes"] or 1
    nbytes = n_v * (4 + 16 + 1 + W) + slots * (4 + 0.125 + W + 1)
    avg = k["ms"] / k["launches"]
    gbs = nbytes / (avg * 1e-3) / 1e9
    log(f"[bench] label statistics: {n_v} active vertices, {slots} slots, {avg:.4f} ms per pass, "
        f"{gbs:.0f} GB/s algorithmic")
    return {"kernel": "label_stat", "active_vertices": n_v, "slots": slots, "avg_launch_ms": avg,
            "bytes_per_launch": nbytes, "achieved_GBs": gbs, "frac": gbs / HBM_PEAK_GBS, "launches": k["launches"]}


def dynamic_order(pr, rows_static, steps: int) -> dict:
    """The dynamic constraint order (PJH_ORDER_DYNAMIC) on the same graph and
    template: solve time, the constraints it chose, its statistics passes,
    and whether its pruning fixed point equals the static order's."""
    pr.run(enumerate_matches=False)
    static_sum = pr.checksum()
    times, dev = [], []
    for _ in range(steps):
        t0 = time.perf_counter()
        rows, dev_ms = pr.run(enumerate_matches=False, order="dynamic")
        times.append(time.perf_counter() - t0)
        dev.append(dev_ms)
    chosen = [(STEP_NAMES.get(r.step, str(r.step)), r.constraint, r.effectiveness) for r in rows
              if r.step in (4, 5, 6)]
    out = {"ms_per_solve": 1e3 * sum(times) / len(times), "device_ms_per_solve": sum(dev) / len(dev),
           "steps": len(rows), "static_steps": len(rows_static),
           "choices": [{"step": s, "constraint": i, "effectiveness": e} for s, i, e in chosen],
           "stat_passes": sum(r.stat_ms > 0 for r in rows), "stat_ms_per_solve": sum(r.stat_ms for r in rows),
           "same_fixed_point_as_static": pr.checksum() == static_sum}
    log(f"[bench] dynamic order: {out['ms_per_solve']:.2f} ms/solve, {len(rows)} steps, choices {chosen}, "
        f"statistics {out['stat_ms_per_solve']:.3f} ms/solve, same fixed point: {out['same_fixed_point_as_static']}")
    return out


def host_cpu() -> dict:
    """CPU model, sockets and logical CPUs of this host, and the CPUs this
    process may run on."""
    model, sockets = None, set()
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name") and model is None:
                model = line.split(":", 1)[1].strip()
            if line.startswith("physical id"):
                sockets.add(line.split(":", 1)[1].strip())
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return {"model": model, "sockets": len(sockets) or None, "logical_cpus": os.cpu_count(), "usable_cpus": usable}


def baseline_threads() -> int:
    """Every CPU the process may use: OMP_NUM_THREADS when the environment
    sets it (the host's share for this job), else the affinity mask."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    return int(env) if env.isdigit() and int(env) > 0 else host_cpu()["usable_cpus"]


def cpu_baseline(g, c, rows_gpu, gpu_s: float, threads: int, run_tds: bool = True) -> dict:
    """The oracle (CPU restatement of the reference) solving the same graph
    and template as the timed device solves.  LCC0, LCC and every phase 2 run
    the reference's loops; the cycle and path steps run the oracle's complete
    search, which decides exactly as the reference's memoised search on
    memo-inert constraints (checked here) and closes cycles by binary search
    instead of the reference's scan of the last row, so it is faster than
    the reference's own algorithm: a conservative baseline."""
    import oracle
    from pattenmatching_sharedmemory_amd import STEP_CC, STEP_PC

    inert = all(c.memo_inert(STEP_CC, i) == 1 for i in range(len(c.cycles()))) and \
        all(c.memo_inert(STEP_PC, i) == 1 for i in range(len(c.paths()))) and not (run_tds and c.tds())
    mode = oracle.COMPLETE if inert else oracle.FAITHFUL
    o = oracle.Oracle(g, c, threads=threads)
    t0 = time.time()
    rows, total_ms = o.run(mode, enumerate_matches=False, run_tds=run_tds)
    wall = time.time() - t0
    sweeps = edge_sweeps(rows, g.edge_count)
    trace_equal = [r.key() for r in rows] == [r.key() for r in rows_gpu]
    log(f"[bench] cpu baseline: {len(rows)} steps, {total_ms:.0f} ms of step time ({wall:.1f}s wall) on {threads} "
        f"threads, {sweeps} edge sweeps; trace equal to the device's: {trace_equal}")
    per_phase = {}
    active = g.edge_count
    for r in rows:
        ph = per_phase.setdefault(STEP_NAMES.get(r.step, str(r.step)), {"ms": 0.0, "edge_sweeps": 0})
        ph["ms"] = round(ph["ms"] + r.ms, 3)
        ph["edge_sweeps"] += active
        active = r.active_edges
    return {"value": sweeps / (total_ms / 1e3), "unit": "TEPS", "cores": threads, "kind": "port",
            "sample": f"the whole measured workload: one pruning solve of the same R-MAT scale-"
                      f"{int(round(math.log2(g.vertex_count)))} graph and template by the oracle (CPU restatement "
                      f"of the reference: its LCC loops and layout; "
                      f"{'complete-search cycle and path steps on memo-inert constraints' if inert else 'its memoised walks'}"
                      f"), OpenMP over {threads} threads; time = sum of step times as the reference measures",
            "host": host_cpu(), "seconds": total_ms / 1e3, "steps": len(rows), "phases": per_phase,
            "trace_equal": trace_equal, "speedup_gpu_vs_cpu": (total_ms / 1e3) / gpu_s}


QUERY_FORMS = {
    # batched form: templates of 3-10 vertices, 0-3 extra edges, on the plain graph
    "batch": {"scale": 22, "min_vertices": 3, "max_vertices": 10, "max_extra": 3, "queries": 256, "seed": 4242},
    # served form: templates of 3-12 vertices, 0-4 extra edges, Poisson arrivals
    "served": {"scale": 24, "min_vertices": 3, "max_vertices": 12, "max_extra": 4, "queries": 200, "seed": 777},
}
# Templates whose longest cycle constraint exceeds this are redrawn: the
# reference's cycle check is a per-source walk search, and 4-cycles through
# the hubs of a scale-22 R-MAT graph with frequent labels make single
# queries run for minutes (up to 10^12 slots examined per step).
MAX_CYCLE = 3


def query_templates(form: dict, n: int):
    """n random connected templates with distinct Zipf-drawn labels whose
    cycle constraints have at most MAX_CYCLE vertices."""
    from pattenmatching_sharedmemory_amd import Constraints, Graph
    from pattenmatching_sharedmemory_amd.templates import longest_cycle, random_templates

    out, specs, seed = [], [], form["seed"]
    while len(out) < n:
        for k, e, l in random_templates(4 * n, seed, form["min_vertices"], form["max_vertices"], form["max_extra"],
                                        labels=32, zipf=1.2, distinct_labels=True):
            c = Constraints(Graph.pattern(k, e, l))
            if longest_cycle(c) <= MAX_CYCLE:
                out.append(c)
                specs.append((k, len(e)))
                if len(out) == n:
                    break
        seed += 1
    return out, specs


def pct(xs, q):
    xs = sorted(xs)
    return xs[min(len(xs) - 1, int(round(q / 100.0 * (len(xs) - 1))))] if xs else None


def query_form(name: str, g=None, queries: int = 0, workers: int = 8, rates=(50.0, 200.0, 1000.0, 2000.0),
               budget_gb: float = 64.0, scale: int | None = None) -> dict:
    """Batched and served forms: many templates over one resident graph,
    each solved to the pruning fixed point before TDS.  Three ways are
    timed on the same templates: one at a time (one context), worker
    contexts sharing the graph (one stream and host thread each), and the
    shared sweeps (PJH_BATCH_LCC: the LCC0 and first LCC fixed point of up
    to 256 templates in one batch whose lanes take queries as they arrive,
    workers continuing each template from its parked state).  Every result
    is compared with the one-at-a-time result."""
    import numpy as np

    from pattenmatching_sharedmemory_amd import Graph, Pruner

    form = QUERY_FORMS[name]
    scale = scale or form["scale"]
    n = queries or form["queries"]
    if g is None:
        t0 = time.time()
        g = Graph.rmat(scale, 16, 0.57, 0.19, 0.19, seed=GRAPH_SEED).labels_zipf(32, 1.2, seed=LABEL_SEED)
        log(f"[bench] graph scale {scale}: V={g.vertex_count} E={g.edge_count} built in {time.time() - t0:.1f}s")
    cs, specs = query_templates(form, n)
    owner = Pruner(g, device=0)
    owner.run_queries(cs[: min(8, n)], workers=1, run_tds=False)  # warm the contexts' allocations
    seq, seq_ms = owner.run_queries(cs, workers=1, run_tds=False)
    log(f"[bench] {name}: one at a time {n} queries in {seq_ms:.0f} ms ({n / seq_ms * 1e3:.1f} q/s)")
    service = [r.end_ms - r.start_ms for r in seq]
    out = {"workload": f"{name} form: {n} random connected templates ({form['min_vertices']}-"
                       f"{form['max_vertices']} vertices, spanning tree + 0-{form['max_extra']} extra edges, "
                       f"distinct labels drawn from Zipf(1.2) over 32, cycles of at most {MAX_CYCLE} vertices) on "
                       f"R-MAT scale {scale} EF16, each solved to the pruning fixed point before TDS",
           "vertices": g.vertex_count, "directed_slots": g.edge_count, "queries": n, "workers": workers,
           "template_vertices_mean": float(np.mean([k for k, _ in specs])),
           "one_at_a_time": {"queries_per_s": n / seq_ms * 1e3, "wall_ms": seq_ms,
                             "service_ms_p50": pct(service, 50), "service_ms_p99": pct(service, 99)}}

    def equal(res):
        return sum(a.result() == b.result() for a, b in zip(seq, res))

    if name == "batch":
        con, con_ms = owner.run_queries(cs, workers=workers, run_tds=False)
        sh, sh_ms = owner.run_queries(cs, workers=workers, run_tds=False, batch_lcc=True)
        log(f"[bench] batch: worker contexts {n / con_ms * 1e3:.1f} q/s ({equal(con)}/{n} equal), shared sweeps "
            f"{n / sh_ms * 1e3:.1f} q/s ({equal(sh)}/{n} equal), one at a time {n / seq_ms * 1e3:.1f} q/s")
        out.update({"worker_contexts": {"queries_per_s": n / con_ms * 1e3, "wall_ms": con_ms,
                                        "results_equal": equal(con)},
                    "shared_sweeps": {"queries_per_s": n / sh_ms * 1e3, "wall_ms": sh_ms,
                                      "results_equal": equal(sh)},
                    "queries_per_s": n / sh_ms * 1e3, "speedup": seq_ms / sh_ms})
    else:
        rng = np.random.default_rng(form["seed"])
        budget = int(budget_gb * (1 << 30))
        table = []
        for rate in rates:
            arrivals = np.cumsum(rng.exponential(1e3 / rate, n))
            arrivals -= arrivals[0]
            row = {"rate_qps": rate}
            for key, kw in (("served", {"workers": workers, "batch_lcc": True, "budget_bytes": budget}),
                            ("worker_contexts", {"workers": workers, "budget_bytes": budget}),
                            ("one_at_a_time", {"workers": 1})):
                res, ms = owner.run_queries(cs, run_tds=False, arrivals_ms=arrivals, **kw)
                lat = [r.end_ms - r.arrival_ms for r in res]
                row[key] = {"queries_per_s": n / ms * 1e3, "p50_ms": pct(lat, 50), "p99_ms": pct(lat, 99),
                            "results_equal_alone": equal(res)}
            log(f"[bench] served rate {rate:.0f} q/s: " + " | ".join(
                f"{k} {row[k]['queries_per_s']:.1f} q/s p50 {row[k]['p50_ms']:.1f} p99 {row[k]['p99_ms']:.1f} ms "
                f"equal {row[k]['results_equal_alone']}/{n}" for k in ("served", "worker_contexts", "one_at_a_time")))
            table.append(row)
        top = table[-1]
        out.update({"budget_bytes": budget, "rates": table, "queries_per_s": top["served"]["queries_per_s"],
                    "speedup": top["served"]["queries_per_s"] / top["one_at_a_time"]["queries_per_s"],
                    "p99_ms": top["served"]["p99_ms"], "plain_solve_ms_p99": pct(service, 99)})
    owner.close()
    return out


def bench_queries(args) -> int:
    """The batched or served form as the bench line itself."""
    rates = tuple(float(x) for x in args.rates.split(","))
    q = query_form(args.workload, queries=args.queries, workers=args.workers, rates=rates, budget_gb=args.budget_gb,
                   scale=args.scale)
    result = {"metric": METRIC, "value": q["queries_per_s"], "unit": "queries/s", "n_gpus": 1, "steps": q["queries"],
              "warmup": 1, "ms_per_step": 1e3 / q["queries_per_s"], "higher_is_better": True, "scaling": "weak",
              "vs_baseline": None, "dtype": "u8", "data": "synthetic",
              "config": {k: v for k, v in q.items() if k in ("workload", "vertices", "directed_slots", "workers",
                                                              "template_vertices_mean", "budget_bytes")},
              "form": q}
    print(json.dumps(result), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=sorted(WORKLOADS) + sorted(QUERY_FORMS), default="plain")
    ap.add_argument("--queries", type=int, default=0, help="batch / served: number of templates")
    ap.add_argument("--workers", type=int, default=8, help="batch / served: worker contexts")
    ap.add_argument("--rates", default="50,200,1000,2000", help="served: arrival rates (queries/s)")
    ap.add_argument("--budget-gb", type=float, default=64.0, help="served: device bytes admitted queries may hold")
    ap.add_argument("--scale", type=int, default=None, help="default: the workload's own scale")
    ap.add_argument("--template-labels", default=None, help="comma-separated label ranks of the template vertices")
    ap.add_argument("--stop-before-tds", action="store_true",
                    help="end each solve when the first TDS step is due (phases up to TDS only)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-profile", action="store_true")
    ap.add_argument("--no-forms", action="store_true", help="plain workload: skip the batched and served forms")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: the pruning path is measured on a single MI355X "
              "(multi-GPU sharding is a design note, not a measured configuration)", file=sys.stderr)
        return 2
    if args.copy_calibration:
        return copy_calibration(args.copy_calibration)
    if args.workload in QUERY_FORMS:
        return bench_queries(args)

    from pattenmatching_sharedmemory_amd import Pruner

    wl = WORKLOADS[args.workload]
    if args.template_labels:
        wl["template_labels"] = [int(x) for x in args.template_labels.split(",")]
        wl["text"] += f" (labels overridden: {wl['template_labels']})"
    if args.scale is None:
        args.scale = wl["scale"]
    g, _, c = workload(args.scale, args.workload)
    t0 = time.perf_counter()
    pr = Pruner(g, c, device=0)
    log(f"[bench] upload and template: {time.perf_counter() - t0:.2f}s")
    for _ in range(args.warmup):
        t0 = time.perf_counter()
        rows, _ = pr.run(enumerate_matches=False, run_tds=not args.stop_before_tds)
        log(f"[bench] warmup solve: {len(rows)} steps in {time.perf_counter() - t0:.3f}s")

    # One solve with every kernel launch bracketed by HIP events on the context
    # stream gives the per-kernel table and the dominant kernel; in the timed
    # solves only that kernel's launches carry events (events around every
    # launch would add about a millisecond per solve).
    kstats_all: list[dict] = []
    if not args.no_profile:
        pr.kernel_timing(True)
        pr.run(enumerate_matches=False, run_tds=not args.stop_before_tds)
        kstats_all = pr.kernel_stats()
        dominant = max(kstats_all, key=lambda k: k["ms"])["name"]
        pr.kernel_timing(True, only=dominant)
    times, dev = [], []
    for k in range(args.steps):
        t0 = time.perf_counter()
        rows, dev_ms = pr.run(enumerate_matches=False, run_tds=not args.stop_before_tds)
        times.append(time.perf_counter() - t0)
        dev.append(dev_ms)
        log(f"[bench] solve {k}: {times[-1] * 1e3:.2f} ms wall, {dev_ms:.2f} ms device")
    slots = sum(r.slots for r in rows)
    sweeps = edge_sweeps(rows, g.edge_count)
    sec = sum(times) / len(times)
    log(f"[bench] {args.steps} solves: {sec * 1e3:.2f} ms/solve wall, {sum(dev) / len(dev):.2f} ms device, "
        f"{len(rows)} steps, {sweeps} edge sweeps, {slots} slots examined")
    for r in rows:
        log(f"    step {r.iteration:3d} {r.step} ve={r.vertex_eliminated} ee={r.edge_eliminated} "
            f"me={r.match_eliminated} en={r.enumeration} av={r.active_vertices} ae={r.active_edges} "
            f"slots={r.slots} ms={r.ms:.3f}")

    roofline = None
    kstats = kstats_all
    if not args.no_profile:
        timed = pr.kernel_stats()  # the dominant kernel over the timed solves
        pr.kernel_timing(False)
        total_k = sum(k["ms"] for k in kstats)
        log("    per-kernel table of one profiled solve:")
        for k in sorted(kstats, key=lambda k: -k["ms"]):
            gbs = algorithmic_bytes(k, g.vertex_count) / (k["ms"] * 1e-3) / 1e9 if k["ms"] else 0.0
            log(f"    kernel {k['name']:24s} launches {k['launches']:5d} ms {k['ms']:9.3f} "
                f"({100 * k['ms'] / max(total_k, 1e-9):5.1f}%) items {k['items']:11d} slots {k['slots']:12d} "
                f"algorithmic {gbs:8.1f} GB/s")
        roofline = roofline_from(timed, g.vertex_count)
        if roofline and args.workload == "plain" and args.scale == 22:
            roofline["traffic"] = measured_traffic(roofline["kernel"])

    result = {
        "metric": METRIC,
        "value": sweeps / sec,
        "unit": "TEPS",
        "teps_edge_sweeps": sweeps / sec,
        "teps_slots_examined": slots / sec,
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": f"{args.workload} pipeline, R-MAT scale {args.scale} EF16 (0.57,0.19,0.19,0.05), "
                               f"{wl['labels']} Zipf({wl['zipf']}) labels, {wl['text']}",
                   "vertices": g.vertex_count, "directed_slots": g.edge_count, "pipeline_steps": len(rows),
                   "edge_sweeps_per_solve": sweeps, "slots_examined_per_solve": slots,
                   "solve": "LCC0, LCC to fixed point, then each cycle / path / TDS step followed by LCC to fixed "
                            "point (pruning fixed point; no enumeration)",
                   "steps": [STEP_NAMES.get(r.step, str(r.step)) for r in rows]},
        "phases": phases(rows, g.edge_count),
        "roofline": roofline,
        "cpu_baseline": None,
        "kernels": [{"name": k["name"], "launches": k["launches"], "ms": round(k["ms"], 4),
                     "algorithmic_GBs": round(algorithmic_bytes(k, g.vertex_count) / (k["ms"] * 1e-3) / 1e9, 1)
                     if k["ms"] else 0.0} for k in sorted(kstats, key=lambda k: -k["ms"])],
    }
    if args.workload == "plain":
        result["dynamic_order"] = dynamic_order(pr, rows, max(1, min(args.steps, 5)))
        result["label_stat"] = label_stat_pass(pr, g)
    pr.close()
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(g, c, rows, sec, baseline_threads(), not args.stop_before_tds)
    if args.workload == "plain" and args.scale == 22 and not args.no_forms:
        # the batched form runs on this same graph; the served form on its own
        # scale-24 graph at the highest arrival rate
        result["batched_form"] = query_form("batch", g=g)
        del g
        result["served_form"] = query_form("served", rates=(50.0, 2000.0))
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
