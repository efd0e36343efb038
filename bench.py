"""Benchmark: same generation (SG) on an R-MAT graph, one MI355X.

Workload: SG (sg(x,y) :- edge(p,x), edge(p,y), x != y;
sg(x,y) :- edge(a,x), sg(a,b), edge(b,y); the reference's SG program
encoding) on an R-MAT graph of scale 15 (32,768 u32 node ids), 40,000
distinct directed edges, a/b/c/d = 0.57/0.19/0.19/0.05, self-loops removed,
SplitMix64 seed 1: 67,916,723 sg tuples over 2 + 6 fixpoint iterations.

A step is one whole run of the program from device-resident edges: loading
the relation (sort, unique, index) and both fixpoints.  `value` is derived
tuples per second (new unique tuples summed over iterations / step time).

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Prints one JSON line.  Also reports:
  roofline      the dominant kernel's algorithmic GB/s against 8 TB/s, from
                hipEvents on the engine stream over the same K steps run
                again (the timed steps run without per-kernel events);
  phases        per kernel: launches, ms, algorithmic GB and GB/s per step;
  cpu_baseline  a restatement of Souffle's SG plan on the host's threads
                (oracle/cpu_souffle.c), with a one-thread sample beside it;
  beside        TC on a 200,000-node child->parent tree (parent window 1,000,
                ~41M tuples, 425 iterations) measured the same way, with
                gdlog's linear loop and Souffle's non-linear TC plan on the
                host beside it;
  batched       256 TC instances (log-uniform 1k-50k edges, ER or tree) in
                one continuous batch vs one program run at a time, with the
                autotuned launch / layout parameters, every instance checked
                against the oracle's goldens;
  served        Poisson streams of TC / SG / CSPA queries at several offered
                loads through the query server and through one-at-a-time
                service, both run for real: p50 / p99 latency per load and the
                highest throughput each sustains at p99 <= 1 s;
  datalog       tc.dl and cspa.dl through the Datalog front end vs the
                hand-written RA encodings, whole runs from host inputs to
                host results;
  path_length   the hop-count program (a dependent column) beside TC on the
                same tree, whole runs;
  micro         the primitives (radix sort, merge + set difference, set
                difference, prefix index build, join without load) on
                device-resident data, algorithmic GB/s against 8 TB/s.
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


def profiled_traffic(kernel, workload_tag):
    """HBM bytes per launch of `kernel` on this workload, from the committed
    rocprofv3 --pmc summary profiles/<kernel>_<workload>_traffic.md
    (FETCH_SIZE + WRITE_SIZE per dispatch, each scaled by what the counter reads
    for a plain device copy of known size), or None when there is none."""
    path = os.path.join(REPO, "profiles", f"{kernel}_{workload_tag}_traffic.md")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return int(line.split(":")[1])
    except (OSError, ValueError):
        pass
    return None


def phase_child(phase, extra):
    """Runs `bench.py --phase <phase>` as a child process (its progress
    lines pass through on stderr) and returns the JSON object it prints, or
    {"error": ...} when it fails (the headline line is still printed)."""
    import subprocess
    cmd = [sys.executable, "-u", os.path.abspath(__file__), "--phase", phase] + extra
    r = subprocess.run(cmd, stdout=subprocess.PIPE, text=True)
    lines = r.stdout.strip().splitlines()
    if r.returncode != 0 or not lines:
        return {"error": f"bench.py --phase {phase} exited with {r.returncode}"}
    try:
        return json.loads(lines[-1])
    except ValueError:
        return {"error": f"bench.py --phase {phase} printed no JSON line"}


def cpu_threads():
    """Host threads for the CPU baseline: this process's CPU share, at most 16
    (the share of one GPU on a shared box)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def host_info():
    """The host the CPU baselines ran on: CPU share and model."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        share = len(os.sched_getaffinity(0))
    except AttributeError:
        share = None
    return dict(nproc=os.cpu_count(), affinity=share, model=model)


def _best(fn, reps):
    best, r = None, None
    for _ in range(reps):
        t = time.perf_counter()
        r = fn()
        dt = time.perf_counter() - t
        best = dt if best is None else min(best, dt)
    return r, best


def cpu_baseline(kind, inp, max_iteration):
    """CPU baselines on the host's threads (this process's CPU share, at most
    16).  SG: the restatement of Souffle's SG plan (oracle/cpu_souffle.c,
    test/souffle/sg.cpp:474-645) on the whole workload, best of 2 after a
    warm-up on its first body, and on one thread over a bounded sample (the
    non-recursive rule and the first recursive body).  TC: gdlog's linear
    loop (oracle/cpu_tc_omp.c) on the whole workload, best of 3 after a
    warm-up; Souffle's non-linear TC plan (oracle/cpu_souffle.c,
    test/souffle/tc.cpp:379-489) on the whole workload once; the one-thread
    oracle on a bounded sample.  CSPA: Souffle's CSPA plan (oracle/cpu_souffle.c,
    test/souffle/cspa.cpp:344-815) on the whole workload, best of 2, with the
    one-thread oracle beside it."""
    if REPO not in sys.path:
        sys.path.insert(0, REPO)
    import oracle
    threads = cpu_threads()
    host = host_info()
    if kind == "sg":
        oracle.souffle_sg(inp, threads, max_iteration=0)  # warm-up: page-in, thread pool
        r, best = _best(lambda: oracle.souffle_sg(inp, threads), 2)
        s, dt1 = _best(lambda: oracle.souffle_sg(inp, 1, max_iteration=1), 1)
        return dict(value=round(r["n_out"] / best, 1), unit="derived tuples/s", cores=threads,
                    kind="port", host=host,
                    sample=f"oracle/cpu_souffle.c (Souffle's SG plan) on the whole workload "
                           f"({r['n_out']} derived tuples, {r['iterations']} recursive bodies, "
                           f"best of 2: {best:.2f} s, {threads} threads)",
                    single_thread=dict(value=round(s["n_out"] / dt1, 1), cores=1,
                                       sample=f"the same plan, non-recursive rule + first "
                                              f"recursive body ({s['n_out']} tuples, "
                                              f"{dt1:.1f} s)"))
    if kind == "tc":
        oracle.tc_omp(inp, threads)  # warm-up: page-in, thread pool
        r, best = _best(lambda: oracle.tc_omp(inp, threads), 3)
        q, dtq = _best(lambda: oracle.souffle_tc(inp, threads), 1)
        s, dt1 = _best(lambda: oracle.tc(inp, max_iteration=max_iteration), 1)
        return dict(value=round(r["derived"] / best, 1), unit="derived tuples/s", cores=threads,
                    kind="port", host=host,
                    sample=f"oracle/cpu_tc_omp.c (gdlog's linear loop) on the whole workload "
                           f"({r['derived']} derived tuples, {r['iterations']} iterations, "
                           f"best of 3: {best:.2f} s, {threads} threads)",
                    souffle_plan=dict(value=round((q["n_out"] - int(inp.size)) / dtq, 1),
                                      cores=threads,
                                      sample=f"oracle/cpu_souffle.c (Souffle's non-linear TC "
                                             f"plan) on the whole workload ({q['iterations']} "
                                             f"recursive bodies, {dtq:.2f} s)"),
                    single_thread=dict(value=round(s["derived"] / dt1, 1), cores=1,
                                       sample=f"oracle, first {s['iterations']} iterations "
                                              f"({s['derived']} derived tuples, {dt1:.1f} s)"))
    oracle.souffle_cspa(inp[0], inp[1], threads, max_iteration=1)  # warm-up
    r, best = _best(lambda: oracle.souffle_cspa(inp[0], inp[1], threads), 2)
    derived = sum(r["sizes"])
    t = time.perf_counter()
    s = oracle.cspa(*inp)
    dt = time.perf_counter() - t
    return dict(value=round(derived / best, 1), unit="derived tuples/s", cores=threads,
                kind="port", host=host,
                sample=f"oracle/cpu_souffle.c (Souffle's CSPA plan) on the whole workload "
                       f"({derived} tuples in ValueFlow, MemoryAlias, ValueAlias; "
                       f"{r['iterations']} recursive bodies; best of 2: {best:.2f} s, "
                       f"{threads} threads)",
                single_thread=dict(value=round(s["derived"] / dt, 1), cores=1,
                                   sample=f"oracle (gdlog's encoding) on the whole workload "
                                          f"({s['derived']} derived tuples, {dt:.1f} s)"))


_last_note = [0.0]


def note(msg, every=20.0):
    """Progress line on stderr, at most one per `every` seconds."""
    now = time.perf_counter()
    if now - _last_note[0] >= every:
        _last_note[0] = now
        print(f"bench: {msg}", file=sys.stderr, flush=True)


def pct(xs, q):
    xs = sorted(xs)
    return xs[min(len(xs) - 1, int(round(q / 100.0 * (len(xs) - 1))))]


def batched(gdl, ctx, n_inst, min_edges=1000, max_edges=50000, window=1000):
    """Batched TC: n_inst instances (seed q: an Erdos-Renyi digraph or a
    child->parent tree, log-uniform edge counts), all arriving at t = 0.
    Serial = one TC program run per instance in order; batched = one TcBatch
    admitting instances as slots and memory allow.  Latency of a query = time
    until its result (|path|, iterations; checksum for the batch) is on the
    host."""
    def make(seeds):
        out = []
        for q in seeds:
            _, _, n = gdl.gen_batched_shape(q, min_edges, max_edges)
            out.append((q, gdl.gen_batched_instance(q, min_edges, max_edges, window), n))
        return out

    inst = make(range(n_inst))
    n_edges = sum(int(e.size) for _, e, _ in inst)
    # autotune on a calibration sample of the same distribution (other seeds)
    calib = make(range(100_000, 100_016))

    def calib_run():
        with gdl.TcBatch(ctx, keep_keys=False) as b:
            b.run(calib)
        note("batched TC: autotuning")
    tune = importlib.import_module("gdlog_amd.tune") if "gdlog_amd.tune" in sys.modules else None
    if tune is None:
        spec = importlib.util.spec_from_file_location(
            "gdlog_amd.tune", os.path.join(REPO, "gdlog-multinode_amd", "tune.py"))
        tune = importlib.util.module_from_spec(spec)
        sys.modules["gdlog_amd.tune"] = tune
        spec.loader.exec_module(tune)
    tuned = tune.autotune(ctx, calib_run, key=f"tc_batched_{min_edges}_{max_edges}",
                          use_cache=False, space=tune.BATCH_SPACE, start=tune.BATCH_DEFAULT)
    # one at a time: the single-query defaults (each query has the device to
    # itself, so the sparser hash table of DEFAULT pays)
    ctx.set_tuning(**tune.DEFAULT)
    ctx.sync()
    t0 = time.perf_counter()
    lat_s, alone = [], {}
    for q, e, _ in inst:
        r = ctx.run_tc(e, want_keys=False, want_trace=False)
        lat_s.append(time.perf_counter() - t0)
        alone[q] = (r["n_out"], r["iterations"], r["derived"])
        note(f"batched TC: serial {q + 1}/{len(inst)}")
    serial_s = time.perf_counter() - t0
    ctx.set_tuning(**tuned["tuning"])
    t0 = time.perf_counter()
    lat_b, done = [], {}
    with gdl.TcBatch(ctx, keep_keys=False) as b:
        pending = list(inst)
        while len(done) < len(inst):
            while pending and b.admit(*pending[0]):
                pending.pop(0)
            b.step()
            for r in b.poll():
                done[r["qid"]] = r
                lat_b.append(time.perf_counter() - t0)
            note(f"batched TC: batch {len(done)}/{len(inst)} done")
        info = b.info()
    batch_s = time.perf_counter() - t0
    mism = [q for q, _, _ in inst
            if (done[q]["n_out"], done[q]["iterations"], done[q]["derived"]) != alone[q]]
    vs_oracle = None
    gold = os.path.join(REPO, "tests", "golden", "batch256.json")
    if n_inst == 256 and (min_edges, max_edges, window) == (1000, 50000, 1000) and \
            os.path.exists(gold):
        with open(gold) as f:
            g = {r["qid"]: r for r in json.load(f)["instances"]}
        vs_oracle = sum(1 for q, _, _ in inst
                        if (done[q]["n_out"], done[q]["iterations"], done[q]["derived"],
                            str(done[q]["checksum"]),
                            [[int(t[3]), int(t[4])] for t in done[q]["trace"]]) !=
                        (g[q]["n_out"], g[q]["iterations"], g[q]["derived"], g[q]["checksum"],
                         g[q]["trace"]))
    derived = sum(v[2] for v in alone.values())
    return {
        "workload": f"tc_batched_{n_inst}: {n_inst} TC instances, edges log-uniform "
                    f"{min_edges}-{max_edges}, ER (nodes = edges/2.5) or child->parent tree "
                    f"(window {window}) 50/50, seeds 0..{n_inst - 1}",
        "instances": n_inst, "edges": n_edges, "derived_tuples": derived,
        "queries_per_s": round(n_inst / batch_s, 3),
        "p50_ms": round(1e3 * pct(lat_b, 50), 2), "p99_ms": round(1e3 * pct(lat_b, 99), 2),
        "derived_tuples_per_s": round(derived / batch_s, 1),
        "serial": {"queries_per_s": round(n_inst / serial_s, 3),
                   "p50_ms": round(1e3 * pct(lat_s, 50), 2),
                   "p99_ms": round(1e3 * pct(lat_s, 99), 2),
                   "derived_tuples_per_s": round(derived / serial_s, 1)},
        "speedup_vs_serial": round(serial_s / batch_s, 3),
        "bodies": info["bodies"],
        "solo_instances": sum(1 for r in done.values() if r["slot"] == -1),
        "mismatches_vs_serial": len(mism),
        "mismatches_vs_oracle": vs_oracle,
        "tuning": dict(tuned["tuning"], calibration_s=tuned["seconds"],
                       default_s=tuned["default_seconds"], trials=tuned["trials"],
                       sample="16 instances of the same generator, seeds 100000-100015"),
    }


SERVED_MIX = (("tc_er_2000_5000", 0.6), ("tc_tree_200k", 0.1), ("sg_rmat15_40k", 0.1),
              ("cspa_stated_1000", 0.2))


SERVED_LOADS = (0.5, 1.0, 1.5, 2.5, 4.0)  # offered load, x the one-at-a-time capacity
SERVED_SLO_MS = 1000.0                   # the p99 latency both sides are held to


def served(gdl, ctx, n_req, seed=7, workers=1, loads=SERVED_LOADS, slo_ms=SERVED_SLO_MS):
    """Served mix: queries drawn from SERVED_MIX (TC / SG at the sizes of the
    single-query workloads, CSPA at the stated input ratio (3 assign and 1.5
    dereference facts per id) on 1,000 ids), offered as Poisson arrivals at
    several loads (multiples of the one-at-a-time engine's capacity).  At
    every load both sides run for real on the same arrivals: the query server
    (continuous batching, gdl_serve.h) and a one-at-a-time loop that takes the
    queries in arrival order and runs each alone.  Reported per load: the
    throughput each side achieved and its p50 / p99 latency (arrival to
    result on the host); and the highest throughput each side sustains with
    p99 <= slo_ms.  Every server result is checked against the query run
    alone."""
    import numpy as np
    rng = np.random.default_rng(seed)
    names = [m[0] for m in SERVED_MIX]
    per = max(20, n_req // len(loads))
    picks = rng.choice(len(SERVED_MIX), size=per, p=[m[1] for m in SERVED_MIX])
    reqs = []
    for i, k in enumerate(picks):
        s = 1000 + i
        if k == 0:
            reqs.append((i, gdl.QUERY_TC, gdl.gen_er(s, 2000, 5000), None))
        elif k == 1:
            reqs.append((i, gdl.QUERY_TC, gdl.gen_tree(s, 200_000, 1000), None))
        elif k == 2:
            reqs.append((i, gdl.QUERY_SG, gdl.gen_rmat(s, 15, 40000), None))
        else:
            a, d = gdl.gen_cspa(s, 1000, 1.1, 3000, 1500)
            reqs.append((i, gdl.QUERY_CSPA, a, d))
    run = {gdl.QUERY_TC: lambda a, b: ctx.run_tc(a, want_keys=False, want_trace=False),
           gdl.QUERY_SG: lambda a, b: ctx.run_sg(a, want_keys=False, want_trace=False),
           gdl.QUERY_CSPA: lambda a, b: ctx.run_cspa(a, b, want_keys=False, want_trace=False)}
    # an untimed pass over every request first: code objects, the allocator's
    # cache, and a device that an earlier process may have left with much of
    # its memory still being reclaimed
    for i, kind, a, b in reqs:
        run[kind](a, b)
        note(f"served mix: warm-up {i + 1}/{per}")
    ctx.sync()
    service, alone = [], {}
    for i, kind, a, b in reqs:
        ctx.sync()
        t0 = time.perf_counter()
        r = run[kind](a, b)
        service.append(time.perf_counter() - t0)
        alone[i] = (r["n_out"], r["iterations"], r["checksum"])
        note(f"served mix: alone {i + 1}/{per}")
    capacity = per / sum(service)

    def one_at_a_time(arrivals):
        t0, lat = time.perf_counter(), []
        for (i, kind, a, b), arr in zip(reqs, arrivals):
            now = time.perf_counter() - t0
            if now < arr:
                time.sleep(arr - now)
            run[kind](a, b)
            lat.append(time.perf_counter() - t0 - arr)
            note(f"served mix: one at a time {len(lat)}/{per}")
        return lat, time.perf_counter() - t0

    def server(arrivals):
        got = {}
        with gdl.Server(workers=workers, keep_keys=False) as srv:
            t0 = time.perf_counter()
            nxt = 0
            while len(got) < per:
                now = time.perf_counter() - t0
                while nxt < per and arrivals[nxt] <= now:
                    i, kind, a, b = reqs[nxt]
                    srv.submit(i, kind, a, b)
                    nxt += 1
                wait_ms = 1000.0 * (arrivals[nxt] - now) if nxt < per else 1000.0
                note(f"served mix: server {len(got)}/{per} done, {nxt} submitted")
                r = srv.wait(max(0.0, wait_ms))
                if r is not None:
                    got[r["qid"]] = (time.perf_counter() - t0, r)
            span = time.perf_counter() - t0
            st = srv.stats()
        lat = [got[i][0] - arrivals[i] for i in range(per)]
        bad = [i for i in range(per)
               if got[i][1]["status"] != 0 or
               (got[i][1]["n_out"], got[i][1]["iterations"], got[i][1]["checksum"]) != alone[i]]
        return lat, span, st, bad

    def summary(lat, span):
        return {"queries_per_s": round(per / span, 3), "p50_ms": round(1e3 * pct(lat, 50), 2),
                "p99_ms": round(1e3 * pct(lat, 99), 2)}

    points, mism, peak, batched_tc = [], 0, 0, 0
    for m in loads:
        rate = m * capacity
        arrivals = np.cumsum(rng.exponential(1.0 / rate, per))
        ctx.release_cached()  # the server's contexts get the memory this one cached
        lat_v, span_v, st, bad = server(arrivals)
        lat_s, span_s = one_at_a_time(arrivals)
        mism += len(bad)
        peak = max(peak, st["peak_reserved"])
        batched_tc += st["batched"]
        by_kind = {}
        for i, k in enumerate(picks):
            by_kind.setdefault(names[k], []).append(lat_v[i])
        points.append({"load": m, "offered_qps": round(rate, 2),
                       "server": dict(summary(lat_v, span_v),
                                      p99_ms_by_kind={k: round(1e3 * pct(v, 99), 2)
                                                      for k, v in sorted(by_kind.items())}),
                       "serial": summary(lat_s, span_s)})
        print(f"bench: served load {m}: {points[-1]}", file=sys.stderr, flush=True)

    def best(side):
        ok = [p[side]["queries_per_s"] for p in points if p[side]["p99_ms"] <= slo_ms]
        return max(ok) if ok else 0.0
    s_best, v_best = best("serial"), best("server")
    at = max((p for p in points if p["server"]["p99_ms"] <= slo_ms),
             key=lambda p: p["server"]["queries_per_s"], default=points[0])
    return {
        "workload": f"served_mix: " + ", ".join(f"{int(100 * p)}% {n}" for n, p in SERVED_MIX) +
                    f"; {per} queries per load, Poisson arrivals at {list(loads)} x the "
                    f"one-at-a-time capacity ({capacity:.1f} queries/s), seed {seed}",
        "slo_p99_ms": slo_ms,
        "capacity_qps": round(capacity, 3),
        "loads": points,
        "max_qps_at_slo": {"server": v_best, "serial": s_best},
        "speedup_vs_serial": round(v_best / s_best, 3) if s_best else None,
        "queries_per_s": at["server"]["queries_per_s"], "p50_ms": at["server"]["p50_ms"],
        "p99_ms": at["server"]["p99_ms"],
        "serial": max((p["serial"] for p in points if p["serial"]["p99_ms"] <= slo_ms),
                      key=lambda x: x["queries_per_s"], default=points[0]["serial"]),
        "workers": workers, "batched_tc": batched_tc,
        "peak_reserved_gb": round(peak / 1e9, 2),
        "mismatches_vs_alone": mism,
    }


def datalog_vs_hand(gdl, ctx, reps):
    """Whole runs, host inputs to host results: the package's tc.dl and
    cspa.dl lowered by the Datalog front end beside the hand-written RA
    encodings of the same rules (the TC and CSPA programs); best of `reps`
    after one warm-up run each."""
    import numpy as np
    dl_dir = os.path.join(REPO, "gdlog-multinode_amd", "programs", "dl")
    tree = np.unique(gdl.gen_tree(0, 200_000, 1000))
    a, d = gdl.gen_cspa(5, 500, 1.1, 750, 375)
    a, d = np.unique(a), np.unique(d)
    cases = {
        "tc_tree_200k": ("tc.dl", dict(edge=tree), "path",
                         lambda: ctx.run_tc(tree, want_trace=False)["out"]),
        "cspa_zipf_500": ("cspa.dl", dict(assign=a, dereference=d), "ValueAlias",
                          lambda: ctx.run_cspa(a, d, want_trace=False)["out"]),
    }

    def best(f):
        r = f()
        ts = []
        for _ in range(reps):
            t = time.perf_counter()
            r = f()
            ts.append(time.perf_counter() - t)
        return r, 1000.0 * min(ts)

    out = {}
    for name, (fn, inputs, rel, hand) in cases.items():
        with open(os.path.join(dl_dir, fn)) as f:
            prog = gdl.Datalog(f.read())
        got, dl_ms = best(lambda: prog.run(ctx, inputs)[rel])
        want, hand_ms = best(hand)
        out[name] = dict(datalog_ms=round(dl_ms, 2), hand_ms=round(hand_ms, 2),
                         equal=bool(np.array_equal(got, want)), tuples=int(got.size))
        prog.close()
        print(f"bench: datalog {name}: {out[name]}", file=sys.stderr, flush=True)
    return out


def path_length_vs_tc(gdl, ctx, reps):
    """The path_length program (hop count per pair, a dependent column) beside
    TC on the 200k-node tree: whole runs from host edges to host results,
    best of `reps` after a warm-up run each."""
    tree = gdl.gen_tree(0, 200_000, 1000)

    def best(f):
        r = f()
        ts = []
        for _ in range(reps):
            t = time.perf_counter()
            r = f()
            ts.append(time.perf_counter() - t)
        return r, 1000.0 * min(ts)

    plen, plen_ms = best(lambda: ctx.run_path_length(tree))
    tc, tc_ms = best(lambda: ctx.run_tc(tree, want_trace=False))
    out = {"tc_tree_200k": dict(path_length_ms=round(plen_ms, 2), tc_ms=round(tc_ms, 2),
                                pairs=int(plen["n_out"]), longest=int(plen["longest"]),
                                same_pairs=bool(plen["n_out"] == tc["n_out"] and
                                                plen["checksum"] == tc["checksum"]))}
    print(f"bench: path_length {out}", file=sys.stderr, flush=True)
    return out


def micro(gdl, ctx, reps):
    """The hot path's primitives on device-resident data at the headline's
    sizes (the reference's micro-benchmark shapes: test/datastructure.cu index
    build and join without load, test/merge.cu set difference and merge), each
    with its algorithmic bytes (w = 8 B per tuple) against the 8 TB/s peak."""
    import ctypes
    import numpy as np
    rng = np.random.default_rng(0)
    U = np.uint64
    eng, h = ctx.eng, ctx.h

    def rand(n, bits):
        return np.unique((rng.integers(0, 1 << bits, n, dtype=U) << U(32)) |
                         rng.integers(0, 1 << bits, n, dtype=U))

    def timed(f):
        f()
        ctx.sync()
        t = time.perf_counter()
        for _ in range(reps):
            f()
        ctx.sync()
        return (time.perf_counter() - t) / reps

    def entry(sec, nbytes, **kw):
        gbs = nbytes / sec / 1e9
        return dict(ms=round(sec * 1e3, 3), alg_bytes=int(nbytes), gbs=round(gbs, 1),
                    frac=round(gbs / HBM_PEAK_GBS, 4), **kw)

    out = {}
    # radix sort: 40.8M keys over 18-bit ids (the tree's output sort)
    keys = rand(41_000_000, 18)
    rng.shuffle(keys)
    n = keys.size
    d, o = ctx.to_device(keys), ctx.alloc(n)
    sec = timed(lambda: ctx._check(eng.gdl_sort_u64(h, d.ptr, n, 18, 18, o.ptr), "sort"))
    out["sort_u64"] = entry(sec, 16 * n, keys=n)
    d.free()
    # fused merge + set difference: FULL 40M sorted, newt 2M sorted (half new)
    full = o  # sorted above
    newt = np.unique(np.concatenate([keys[:1_000_000], rand(1_000_000, 18)]))
    dn = ctx.to_device(newt)
    u, dl = ctx.alloc(n + newt.size), ctx.alloc(newt.size)
    nu, nd = ctypes.c_uint64(), ctypes.c_uint64()
    sec = timed(lambda: ctx._check(eng.gdl_merge_diff_u64(
        h, full.ptr, n, dn.ptr, newt.size, u.ptr, ctypes.byref(nu), dl.ptr, ctypes.byref(nd)),
        "merge_diff"))
    out["merge_diff_u64"] = entry(sec, 8 * (n + newt.size + nu.value + nd.value), full=n,
                                  newt=int(newt.size), new=int(nd.value))
    sec = timed(lambda: ctx._check(eng.gdl_set_difference_u64(
        h, dn.ptr, newt.size, full.ptr, n, dl.ptr, ctypes.byref(nd)), "set_difference"))
    out["set_difference_u64"] = entry(sec, 8 * (n + newt.size + nd.value), full=n,
                                      newt=int(newt.size))
    for x in (u, dl):
        x.free()
    # prefix index over 40M sorted keys (one-column prefixes), then a join of 2M
    # outer tuples against it with the result dropped (disable_load)
    idx = gdl.Index()

    def build():
        if idx.slots:
            ctx.index_free(idx)
        ctx._check(eng.gdl_index_build(h, full.ptr, n, 1, ctypes.byref(idx)), "index_build")
    sec = timed(build)
    # read the keys; write one run word per slot (8 B dense, 16 B hash)
    out["index_build"] = entry(sec, 8 * n + (8 if idx.dense else 16) * idx.n_slots, keys=n,
                               dense=bool(idx.dense))
    outer = rng.integers(0, 1 << 18, 2_000_000, dtype=U) << U(32)
    do = ctx.to_device(outer)
    cm = gdl.ColMap(2, (gdl.COL_OUTER0, gdl.COL_INNER1))
    m = ctypes.c_uint64()

    def join():
        res = ctypes.c_void_p()
        ctx._check(eng.gdl_join(h, full.ptr, n, ctypes.byref(idx), do.ptr, do.n, cm,
                                gdl.Pred(gdl.PRED_NONE, 0, 1), ctypes.byref(res),
                                ctypes.byref(m)), "join")
        if res.value:
            eng.gdl_free(h, res.value)
    sec = timed(join)
    out["join"] = entry(sec, 8 * (do.n + 2 * m.value), outer=int(do.n), matches=int(m.value))
    ctx.index_free(idx)
    for x in (do, dn, full):
        x.free()
    print(f"bench: micro {out}", file=sys.stderr, flush=True)
    return out


def workloads(gdl, nodes, window):
    """name -> (kind, make_input, description)."""
    return {
        "tree": ("tc", lambda: gdl.gen_tree(0, nodes, window),
                 f"tc_tree_{nodes // 1000}k: TC on a {nodes}-node child->parent tree, "
                 f"parent window {window}"),
        "er": ("tc", lambda: gdl.gen_er(0, 2000, 5000),
               "tc_er_2000_5000: TC on a directed Erdos-Renyi graph, 2,000 nodes, 5,000 edges"),
        "sg": ("sg", lambda: gdl.gen_rmat(1, 15, 40000),
               "sg_rmat15_40k: same generation on an R-MAT graph, scale 15, 40,000 edges"),
        "cspa": ("cspa", lambda: gdl.gen_cspa(5, 500, 1.1, 750, 375),
                 "cspa_zipf_500: CSPA, 500 ids, 750 assign / 375 dereference (Zipf 1.1)"),
    }


def measure(gdl, ctx, kind, inp, desc, steps, warmup):
    """Times `steps` whole program runs from device-resident inputs (after
    `warmup` untimed ones), then runs the same steps again with a hipEvent
    pair around every launch on the engine stream: per-kernel times, their
    algorithmic bytes (GB/s per phase) and the dominant kernel's roofline."""
    if kind == "cspa":
        dev = (ctx.to_device(inp[0]), ctx.to_device(inp[1]))

        def step():
            return ctx.run_cspa(dev[0], dev[1], on_device=True, want_keys=False,
                                want_trace=False)
        n_in = int(inp[0].size + inp[1].size)
    else:
        dev = (ctx.to_device(inp),)
        run = ctx.run_tc if kind == "tc" else ctx.run_sg

        def step():
            return run(dev[0], on_device=True, want_keys=False, want_trace=False)
        n_in = int(inp.size)

    res = None
    for _ in range(warmup):
        res = step()
    # the timed steps run without per-kernel events (two hipEventRecords per
    # launch would add host time to ~1,000 launches a step)
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        res = step()
    ctx.sync()
    elapsed = time.perf_counter() - t0
    ms_per_step = 1000.0 * elapsed / steps
    ctx.stats_reset()
    ctx.stats_enable(True)
    ctx.sync()
    t1 = time.perf_counter()
    for _ in range(steps):
        step()
    ctx.sync()
    ms_instrumented = 1000.0 * (time.perf_counter() - t1) / steps
    ctx.stats_enable(False)
    stats = ctx.stats()
    for d in dev:
        d.free()

    timed = {k: v for k, v in stats.items() if v["launches"] and v["total_ms"] > 0}
    dom = max(timed.items(), key=lambda kv: kv[1]["total_ms"]) if timed else None
    roofline = None
    if dom:
        name, s = dom
        avg_ms = s["total_ms"] / s["launches"]
        avg_bytes = s["alg_bytes"] / s["launches"]
        achieved = avg_bytes / (avg_ms * 1e-3) / 1e9
        traffic = profiled_traffic(name, desc.split(":")[0])
        roofline = dict(bound="hbm", kernel=name, achieved=round(achieved, 2),
                        peak=HBM_PEAK_GBS, unit="GB/s", frac=round(achieved / HBM_PEAK_GBS, 5),
                        traffic=traffic, avg_launch_us=round(avg_ms * 1e3, 3),
                        avg_alg_bytes=round(avg_bytes))
        if traffic is not None:
            roofline["traffic_unit"] = "HBM bytes per launch (rocprofv3 --pmc, profiles/)"
            roofline["traffic_gbs"] = round(traffic / (avg_ms * 1e-3) / 1e9, 2)
    tot_bytes = sum(v["alg_bytes"] for v in stats.values()) / steps
    kernel_ms = sum(v["total_ms"] for v in stats.values()) / steps
    # per phase: launches, hipEvent ms, algorithmic bytes over that time
    phases = {}
    for k, v in sorted(stats.items(), key=lambda kv: -kv[1]["total_ms"]):
        if not v["launches"]:
            continue
        ms = v["total_ms"] / steps
        gbs = v["alg_bytes"] / steps / (ms * 1e-3) / 1e9 if ms > 0 else None
        phases[k] = dict(launches=v["launches"] // steps, ms=round(ms, 3),
                         gb=round(v["alg_bytes"] / steps / 1e9, 4),
                         gbs=None if gbs is None else round(gbs, 1),
                         frac=None if gbs is None else round(gbs / HBM_PEAK_GBS, 4))
    return dict(res=res, n_in=n_in, ms_per_step=ms_per_step,
                derived_per_s=res["derived"] / (elapsed / steps),
                ms_instrumented=round(ms_instrumented, 3), roofline=roofline, phases=phases,
                step_hbm_frac=round(tot_bytes / (ms_per_step * 1e-3) / (HBM_PEAK_GBS * 1e9), 5),
                kernel_busy_frac=round(min(1.0, kernel_ms / ms_per_step), 4))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="sg", choices=["tree", "er", "sg", "cspa"])
    ap.add_argument("--mode", default="auto", choices=["auto", "merge"],
                    help="FULL maintenance (merge = the reference's per-iteration merge)")
    ap.add_argument("--nodes", type=int, default=200_000)
    ap.add_argument("--window", type=int, default=1000)
    ap.add_argument("--cpu-iterations", type=int, default=100,
                    help="fixpoint iterations of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-beside", action="store_true",
                    help="skip the tree-TC measurement beside the SG headline")
    ap.add_argument("--batched-instances", type=int, default=256,
                    help="instances of the batched-TC measurement (0: skip it)")
    ap.add_argument("--served-requests", type=int, default=1000,
                    help="requests of the served-mix measurement (0: skip it)")
    ap.add_argument("--served-workers", type=int, default=4,
                    help="program workers of the query server beside its TC batch")
    ap.add_argument("--datalog-reps", type=int, default=3,
                    help="timed runs of the Datalog front-end and path-length "
                         "comparisons (0: skip them)")
    ap.add_argument("--micro-reps", type=int, default=10,
                    help="timed repeats of each primitive micro-benchmark (0: skip them)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a 256 MiB device copy (traffic-counter calibration)")
    ap.add_argument("--phase", choices=["batched", "served"],
                    help="only run that measurement and print its JSON object")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this engine runs on one GPU; --gpus must be 1 "
              "(multi-GPU sharding is a design, not a build)", file=sys.stderr)
        sys.exit(2)

    gdl = _pkg()
    if args.calibrate_copy:
        # a plain 256 MiB device-to-device copy: what FETCH_SIZE / WRITE_SIZE
        # read for a known number of bytes (the scale of the traffic counters)
        ctx = gdl.Context(device=0)
        n = 32 << 20
        a, b = ctx.alloc(n), ctx.alloc(n)
        ctx._check(ctx.eng.gdl_memcpy_d2d(ctx.h, b.ptr, a.ptr, n * 8), "copy")
        ctx.sync()
        print(json.dumps({"copied_bytes": n * 8}), flush=True)
        a.free(); b.free(); ctx.close()
        return
    if args.phase:
        ctx = gdl.Context(device=0)
        out = batched(gdl, ctx, args.batched_instances) if args.phase == "batched" else \
            served(gdl, ctx, args.served_requests, workers=args.served_workers)
        ctx.close()
        print(json.dumps(out), flush=True)
        return
    gdl.set_full_mode(gdl.FULL_MODE_MERGE if args.mode == "merge" else gdl.FULL_MODE_AUTO)
    wl = workloads(gdl, args.nodes, args.window)
    kind, make, desc = wl[args.workload]
    inp = make()
    ctx = gdl.Context(device=0)
    head = measure(gdl, ctx, kind, inp, desc, args.steps, args.warmup)
    res, ms_per_step = head["res"], head["ms_per_step"]
    derived_per_s = head["derived_per_s"]
    # the deep-tree TC workload beside the headline (its own roofline: the
    # stream-ordered linear loop of small bodies)
    beside = None
    if args.workload == "sg" and not args.no_beside:
        bk, bmake, bdesc = wl["tree"]
        binp = bmake()
        b = measure(gdl, ctx, bk, binp, bdesc, args.steps, args.warmup)
        beside = {"workload": bdesc, "value": round(b["derived_per_s"], 1),
                  "unit": "derived tuples/s", "ms_per_step": round(b["ms_per_step"], 3),
                  "output_tuples": b["res"]["n_out"], "iterations": b["res"]["iterations"],
                  "derived_tuples": b["res"]["derived"], "roofline": b["roofline"],
                  "step_hbm_frac": b["step_hbm_frac"], "phases": b["phases"],
                  "cpu_baseline": None if args.no_cpu else
                  cpu_baseline(bk, binp, args.cpu_iterations)}
    roofline, phases = head["roofline"], head["phases"]
    gdl.set_full_mode(gdl.FULL_MODE_AUTO)
    dlog = datalog_vs_hand(gdl, ctx, args.datalog_reps) if args.datalog_reps > 0 else None
    plen = path_length_vs_tc(gdl, ctx, args.datalog_reps) if args.datalog_reps > 0 else None
    micro_out = micro(gdl, ctx, args.micro_reps) if args.micro_reps > 0 else None
    # the batched and served measurements each run in a process of their own
    # (as a batch or a server would): device memory that one long run of
    # allocations and frees gave back comes back to the next one much more
    # slowly, so in one process the second measurement paid for the first
    ctx.release_cached()
    batch = serve = None
    if args.batched_instances > 0:
        batch = phase_child("batched", ["--batched-instances", str(args.batched_instances)])
    if args.served_requests > 0:
        serve = phase_child("served", ["--served-requests", str(args.served_requests),
                                       "--served-workers", str(args.served_workers)])

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(kind, inp, args.cpu_iterations)

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
        "data": "synthetic (SplitMix64 generators; R-MAT seed 1, tree seed 0)",
        "config": {"workload": desc, "full_mode": args.mode, "input_tuples": head["n_in"],
                   "output_tuples": res["n_out"], "iterations": res["iterations"],
                   "init_iterations": res["init_iterations"],
                   "derived_tuples": res["derived"]},
        "roofline": roofline,
        "step_hbm_frac": head["step_hbm_frac"],
        "kernel_busy_frac": head["kernel_busy_frac"],
        "ms_per_step_with_kernel_events": head["ms_instrumented"],
        "phases": phases,
        "cpu_baseline": cpu,
        "batched": batch,
        "served": serve,
        "beside": beside,
        "datalog": dlog,
        "path_length": plen,
        "micro": micro_out,
    }
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
