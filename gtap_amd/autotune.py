"""Autotuner for the persistent scheduler: picks worker count, steal probe width, steal
batch width, deque capacity and the programs' split sizes by measurement.

The reference fixes these at compile time (GTAP_GRID_SIZE, GTAP_BLOCK_SIZE,
GTAP_MAX_TASKS_PER_WARP and the cutoffs of each program; runtime/detail/common/gtap_config.cuh
and evaluation/*/compare_*.sh). Here they are run-time settings (gtap_config and the
programs' arguments), so they can be searched on the GPU that runs the work.

Search: coordinate descent. Starting from the defaults, each parameter in turn is swept
over its candidate values with the others held at the best found so far; the sweep repeats
until a full pass changes nothing (or `passes` passes). Each candidate is timed `reps` times
(median of the launch times the runtime reports). A candidate whose result differs from the
first run's result is an error, not a slow point.

    python -m gtap_amd.autotune --workload nq14 --out tuned_nq14.json
    gtap_amd.autotune.apply(json.load(open("tuned_nq14.json"))["best"])
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time
from typing import Callable, Dict, List, Optional, Sequence, Tuple

# gtap_config fields the tuner may set (0 = the runtime's default / every co-resident workgroup)
CONFIG_KEYS = ("blocks", "block_workers", "probe_lanes", "steal_batch", "deque_capacity", "epaq_publish", "idle_sleep",
               "xcd_steal")

# The first value of each coordinate is the runtime's default and the search's start point.
DEFAULT_SPACE: Dict[str, Sequence[int]] = {
    "blocks": (0, 256, 512, 1024, 1536),
    "probe_lanes": (8, 1, 4, 16, 32),
    "steal_batch": (64, 8, 16, 32),
    "deque_capacity": (32768, 65536),
    "xcd_steal": (0, 1),
}

Measure = Callable[[Dict[str, int]], Tuple[float, object]]


def coordinate_descent(measure: Measure, space: Dict[str, Sequence[int]], start: Optional[Dict[str, int]] = None,
                       passes: int = 2, log: Callable[[str], None] = lambda s: None) -> Tuple[Dict[str, int], List[dict]]:
    """Minimises measure(params)[0] over the product space, one coordinate at a time.

    measure returns (time, result); every result must equal the first one. Returns the best
    parameters and the table of every point measured (params, time)."""
    best = {k: (start or {}).get(k, vals[0]) for k, vals in space.items()}
    cache: Dict[Tuple, float] = {}
    table: List[dict] = []
    ref_result: List[object] = []

    def run(p: Dict[str, int]) -> float:
        key = tuple(sorted(p.items()))
        if key in cache:
            return cache[key]
        t, res = measure(dict(p))
        if not ref_result:
            ref_result.append(res)
        elif res != ref_result[0]:
            raise RuntimeError(f"autotune: result changed at {p}: {res!r} != {ref_result[0]!r}")
        cache[key] = t
        table.append({"params": dict(p), "ms": t})
        log(f"  {p} -> {t:.4f} ms")
        return t

    best_t = run(best)
    for _ in range(passes):
        changed = False
        for k, vals in space.items():
            for v in vals:
                if v == best[k]:
                    continue
                cand = dict(best)
                cand[k] = v
                t = run(cand)
                if t < best_t:
                    best, best_t, changed = cand, t, True
        if not changed:
            break
    return best, table


# Best points found on an MI355X by `python -m gtap_amd.autotune --workload W`, one entry per
# workload: {"best": {...}, "best_ms": ..., "default_ms": ...}.
TUNED_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned_mi355x.json")


def tuned(name: str) -> Dict[str, int]:
    """The committed best parameters of a workload ({} when it was not tuned)."""
    try:
        with open(TUNED_PATH) as f:
            table = json.load(f)
    except (OSError, ValueError):
        return {}
    return dict(table.get(name, {}).get("best", {}))


def apply(params: Dict[str, int], **base: int) -> None:
    """Re-initialises the runtime with the gtap_config fields of `params` (over `base`, e.g.
    timeout_ms; fields in neither take the runtime's defaults)."""
    import gtap_amd
    cfg = dict(base)
    cfg.update({k: int(v) for k, v in params.items() if k in CONFIG_KEYS})
    gtap_amd.gtap_finalize()
    if gtap_amd.gtap_set_config(**cfg) != 0:
        raise RuntimeError(f"gtap_set_config({cfg}) failed")
    if gtap_amd.gtap_initialize() != 0:
        raise RuntimeError("gtap_initialize failed")


def _median_ms(fn: Callable[[], Tuple[object, float]], reps: int) -> Tuple[float, object]:
    times, res = [], None
    for _ in range(reps):
        res, ms = fn()
        times.append(ms)
    return statistics.median(times), res


def workload(name: str, reps: int = 3) -> Tuple[Measure, Dict[str, Sequence[int]]]:
    """A measure function and its search space for a built-in workload:
    fibN / nqN (thread mode, no cutoff), nqcutN (spawn cutoff row), cilksort (2^26 keys, block
    mode, leaf and merge cutoffs), bfs / bfs1 (R-MAT scale 22 with 4 EPAQ queues / one queue,
    split), spmv (row cutoff), stream (4096 jobs, sort jobs' merge leaf)."""
    import gtap_amd
    space = dict(DEFAULT_SPACE)

    if name.startswith("nqcut"):
        n = int(name[5:])
        space["cutoff"] = (7, 4, 5, 6, 8, 9, 10)  # rows run as tasks (nq.cu uses 7)

        def measure(p):
            apply(p)
            return _median_ms(lambda: (lambda r: (r[0], r[1]["kernel_ms"]))(
                gtap_amd.nqueens_cutoff(n, p.get("cutoff", 7))), reps)
        return measure, space

    if name == "spmv":
        rp, ci, va = gtap_amd.spmv_random(1 << 16, 300, 42)
        import numpy as np
        x = np.linspace(-1.0, 1.0, rp.size - 1)
        space["cutoff"] = (16, 4, 8, 32, 64, 128)  # rows per leaf task (spmv_thread.cu uses 16)

        def measure(p):
            apply(p)

            def one():
                y, st = gtap_amd.spmv(rp, ci, va, x, p.get("cutoff", 16))
                return hash(y.tobytes()), st["kernel_ms"]
            return _median_ms(one, reps)
        return measure, space

    if name.startswith("fib") or name.startswith("nq"):
        n = int(name[3:] if name.startswith("fib") else name[2:])
        fn = gtap_amd.fib if name.startswith("fib") else gtap_amd.nqueens

        def measure(p):
            apply(p)
            return _median_ms(lambda: (lambda r: (r[0], r[1]["kernel_ms"]))(fn(n)), reps)
        return measure, space

    if name == "cilksort":
        import numpy as np
        keys = gtap_amd.gen_keys(1 << 26, 123456789)
        space = {"block_workers": (0, 256, 512, 768), "merge_leaf": (32768, 16384, 65536, 131072),
                 "leaf": (4096, 2048, 1024)}

        def measure(p):
            apply(p)

            def one():
                out, st = gtap_amd.cilksort(keys, leaf=p.get("leaf", 4096), merge_leaf=p.get("merge_leaf", 32768))
                return hash(out.tobytes()), st["kernel_ms"]
            return _median_ms(one, reps)
        return measure, space

    if name in ("bfs", "bfs1"):
        row, col = gtap_amd.rmat_csr(22, 16, 1)
        queues = 4 if name == "bfs" else 1
        space = dict(DEFAULT_SPACE)
        space["split"] = (64, 16, 32, 128, 256)
        if queues > 1:
            space["epaq_publish"] = (0, 1)

        def measure(p):
            apply(p)

            def one():
                dist, st = gtap_amd.bfs(row, col, 0, queues, p.get("split", 64))
                return hash(dist.tobytes()), st["kernel_ms"]
            return _median_ms(one, reps)
        return measure, space

    if name == "stream":
        space = {k: DEFAULT_SPACE[k] for k in ("blocks", "probe_lanes", "steal_batch")}
        space["merge_leaf"] = (4096, 2048, 8192, 16384)  # sort jobs' one-wave merge size

        def measure(p):
            apply(p)
            s = gtap_amd.Stream(seed=2024, njobs=4096)
            s.set_merge_leaf(p.get("merge_leaf", 4096))
            try:
                def one():
                    res, _, st = s.run_batched()
                    return hash(res.tobytes()), st["kernel_ms"]
                return _median_ms(one, reps)
            finally:
                s.close()
        return measure, space

    raise ValueError(f"unknown workload {name!r}")


def main(argv: Optional[List[str]] = None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--workload", default="nq14", help="fibN, nqN, nqcutN, cilksort, bfs, bfs1, spmv or stream")
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--passes", type=int, default=2)
    ap.add_argument("--out", default=None, help="JSON file for the best parameters and the table")
    a = ap.parse_args(argv)
    import gtap_amd
    if gtap_amd.gtap_initialize() != 0:
        print("gtap_initialize failed", file=sys.stderr)
        return 1
    measure, space = workload(a.workload, a.reps)
    t0 = time.time()
    best, table = coordinate_descent(measure, space, passes=a.passes, log=lambda s: print(s, file=sys.stderr, flush=True))
    best_ms = min(r["ms"] for r in table if r["params"] == best)
    out = {"workload": a.workload, "best": best, "best_ms": best_ms, "default_ms": table[0]["ms"],
           "points": len(table), "seconds": time.time() - t0, "table": table}
    gtap_amd.gtap_finalize()
    text = json.dumps(out, indent=1)
    if a.out:
        with open(a.out, "w") as f:
            f.write(text + "\n")
    print(json.dumps({k: out[k] for k in ("workload", "best", "best_ms", "default_ms", "points")}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
