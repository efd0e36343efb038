"""Benchmark: transitive closure on a deep sparse tree, one MI355X.

Workload: TC (path(x,z) :- path(x,y), edge(y,z), semi-naive, the reference's
TC program encoding) on a child -> parent tree of 200,000 u32 node ids whose
node i has its parent uniform in [max(0, i - 1000), i - 1] (SplitMix64 seed 0):
199,999 edges, about 41M path tuples, about 400 fixpoint iterations.

A step is one whole run of the program from device-resident edges: loading
the two relations (sort, unique, index) and the fixpoint.  `value` is derived
tuples per second (new unique tuples summed over iterations / step time).

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Prints one JSON line.  Also reports the roofline of the dominant kernel
(hipEvent-timed on the engine stream during the timed steps) and a CPU
baseline: the oracle (single-threaded C restatement of the reference
algorithm) timed on a bounded sample of the same workload.
"""
import argparse
import importlib.util
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, 8 TB/s

METRIC = "derived tuples/s + % HBM peak (1 GPU); batched queries/s vs serial at p99"


def _pkg():
    name = "gdlog_amd"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(
        name, os.path.join(REPO, "gdlog-multinode_amd", "__init__.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def cpu_baseline(edges, max_iteration):
    """Oracle (plain C, one thread) on the first iterations of the workload."""
    if REPO not in sys.path:
        sys.path.insert(0, REPO)
    import oracle
    t = time.perf_counter()
    r = oracle.tc(edges, max_iteration=max_iteration)
    dt = time.perf_counter() - t
    return dict(value=r["derived"] / dt, unit="derived tuples/s", cores=1, kind="port",
                sample=f"oracle TC on the same tree, first {r['iterations']} fixpoint "
                       f"iterations ({r['derived']} derived tuples, {dt:.1f} s)")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nodes", type=int, default=200_000)
    ap.add_argument("--window", type=int, default=1000)
    ap.add_argument("--cpu-iterations", type=int, default=100,
                    help="fixpoint iterations of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this engine runs on one GPU; --gpus must be 1 "
              "(multi-GPU sharding is a design, not a build)", file=sys.stderr)
        sys.exit(2)

    gdl = _pkg()
    edges = gdl.gen_tree(0, args.nodes, args.window)
    ctx = gdl.Context(device=0)
    d_edges = ctx.to_device(edges)

    res = None
    for _ in range(args.warmup):
        res = ctx.run_tc(d_edges, on_device=True, want_keys=False, want_trace=False)
    ctx.stats_reset()
    ctx.stats_enable(True)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        res = ctx.run_tc(d_edges, on_device=True, want_keys=False, want_trace=False)
    ctx.sync()
    elapsed = time.perf_counter() - t0
    ctx.stats_enable(False)
    stats = ctx.stats()
    ms_per_step = 1000.0 * elapsed / args.steps
    derived_per_s = res["derived"] / (elapsed / args.steps)

    # dominant kernel by hipEvent time
    timed = {k: v for k, v in stats.items() if v["launches"] and v["total_ms"] > 0}
    dom = max(timed.items(), key=lambda kv: kv[1]["total_ms"]) if timed else None
    roofline = None
    if dom:
        name, s = dom
        avg_ms = s["total_ms"] / s["launches"]
        avg_bytes = s["alg_bytes"] / s["launches"]
        achieved = avg_bytes / (avg_ms * 1e-3) / 1e9
        roofline = dict(bound="hbm", kernel=name, achieved=round(achieved, 2),
                        peak=HBM_PEAK_GBS, unit="GB/s", frac=round(achieved / HBM_PEAK_GBS, 5),
                        traffic=None, avg_launch_us=round(avg_ms * 1e3, 3),
                        avg_alg_bytes=round(avg_bytes))
    # whole-step HBM fraction: all algorithmic bytes over the step time
    tot_bytes = sum(v["alg_bytes"] for v in stats.values()) / args.steps
    kernel_ms = sum(v["total_ms"] for v in stats.values()) / args.steps
    phases = {k: dict(launches=v["launches"] // args.steps,
                      ms=round(v["total_ms"] / args.steps, 3))
              for k, v in sorted(stats.items(), key=lambda kv: -kv[1]["total_ms"])
              if v["launches"]}

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(edges, args.cpu_iterations)

    out = {
        "metric": METRIC,
        "value": round(derived_per_s, 1),
        "unit": "derived tuples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic (SplitMix64 seed 0)",
        "config": {"workload": f"tc_tree_{args.nodes // 1000}k: TC on a {args.nodes}-node "
                               f"child->parent tree, parent window {args.window}",
                   "edges": int(edges.size), "path_tuples": res["n_out"],
                   "iterations": res["iterations"], "derived_tuples": res["derived"]},
        "roofline": roofline,
        "step_hbm_frac": round(tot_bytes / (ms_per_step * 1e-3) / (HBM_PEAK_GBS * 1e9), 5),
        "kernel_busy_frac": round(kernel_ms / ms_per_step, 4),
        "phases": phases,
        "cpu_baseline": cpu,
    }
    d_edges.free()
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
