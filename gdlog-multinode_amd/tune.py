"""Autotuner of an engine context's launch and layout parameters.

The reference fixes them in code: hash load factor 0.8 (src/relation.cu:603-604),
512 threads x 32 blocks per SM for every kernel (test/tc.cu:224-225) and
Thrust's sort.  Here they are a `gdl_tuning` of the context
(include/gdl_hip.h): hash load factor, fused-join block size, radix sort digit
width and grid cap.  `autotune` times a calibration workload under candidate
settings by coordinate descent over

    load_factor   0.35, 0.5, 0.7, 0.8
    join_block    128, 256, 512
    radix_bits    6, 7, 8
    blocks_per_cu 4, 8, 16

(each candidate: best of `reps` runs), keeps the fastest, and caches the
choice per (device, workload key) in a JSON file, so a later run of the same
workload on the same device reuses it without timing anything.

A TC batch holds many instances' hash sets at once and keeps them at load 1/2
or more (memory, not probe length, limits it), so BATCH_SPACE / BATCH_DEFAULT
search load factors 0.5, 0.7 and 0.8 only.
"""
import json
import os
import time

SPACE = {
    "load_factor": [0.35, 0.5, 0.7, 0.8],
    "join_block": [128, 256, 512],
    "radix_bits": [6, 7, 8],
    "blocks_per_cu": [4, 8, 16],
}
DEFAULT = {"load_factor": 0.35, "join_block": 256, "radix_bits": 8, "blocks_per_cu": 8}
BATCH_SPACE = dict(SPACE, load_factor=[0.5, 0.7, 0.8])
BATCH_DEFAULT = dict(DEFAULT, load_factor=0.5)


def cache_path():
    base = os.environ.get("XDG_CACHE_HOME") or os.path.join(os.path.expanduser("~"), ".cache")
    return os.path.join(base, "gdlog_amd", "tuning.json")


def _load_cache(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store_cache(path, cache):
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass  # a read-only home only loses the cache


def autotune(ctx, workload, key, device_key="device0", reps=2, sweeps=1, use_cache=True,
             path=None, space=None, start=None):
    """Picks ctx's tuning for `workload` (a callable running it on ctx),
    searching `space` (default SPACE) from `start` (default DEFAULT).

    Returns {"tuning": {...}, "seconds": best time, "default_seconds": time
    with the defaults, "trials": n, "cached": bool}; ctx is left set to the
    chosen tuning."""
    path = path or cache_path()
    ck = f"{device_key}|{key}"
    cache = _load_cache(path) if use_cache else {}
    if ck in cache:
        ctx.set_tuning(**cache[ck]["tuning"])
        out = dict(cache[ck])
        out["cached"] = True
        return out

    timed = {}

    def run(t):
        sig = tuple(sorted(t.items()))
        if sig not in timed:
            ctx.set_tuning(**t)
            best = float("inf")
            for _ in range(reps):
                ctx.sync()
                t0 = time.perf_counter()
                workload()
                ctx.sync()
                best = min(best, time.perf_counter() - t0)
            timed[sig] = best
        return timed[sig]

    workload()  # warm-up: caches, code objects, allocator
    space = SPACE if space is None else space
    cur = dict(DEFAULT if start is None else start)
    cur_s = run(cur)
    default_s = cur_s
    for _ in range(sweeps):
        changed = False
        for name, values in space.items():
            for v in values:
                if v == cur[name]:
                    continue
                cand = dict(cur, **{name: v})
                s = run(cand)
                if s < cur_s * 0.98:  # a real gain, not timing noise
                    cur, cur_s, changed = cand, s, True
        if not changed:
            break
    ctx.set_tuning(**cur)
    out = {"tuning": cur, "seconds": round(cur_s, 5), "default_seconds": round(default_s, 5),
           "trials": len(timed)}
    if use_cache:
        cache[ck] = out
        _store_cache(path, cache)
    out = dict(out)
    out["cached"] = False
    return out
