"""Worker-scaling sweep: launch time against the number of workers, for each scheduler.

The reference measures its batched work stealing (WS) against a global queue (GQ) and a
single-item Chase-Lev deque (CL) over worker counts, rebuilding each binary with a
different GTAP_GRID_SIZE / GTAP_BLOCK_SIZE (evaluation/1-worker_scalability/fib/
worker_scaling_fib.sh:44-172). Here the worker count (gtap_config.blocks) and the scheduler
(gtap_config.scheduler) are run-time settings, so one process sweeps them all. Workgroups
are always 256 threads (4 wavefront workers).

Each point is timed `runs` times; the CSV carries the median and the nearest-rank quartile
distances (worker_scaling_fib.sh:66-104), in the reference's column layout:

    block_size,total_threads,grid_size,max_tasks_per_warp,ws_med,ws_err_low,ws_err_high,
    gq_med,gq_err_low,gq_err_high,chaselev_med,chaselev_err_low,chaselev_err_high

    python -m gtap_amd.scaling --workload fib --n 30 --out fib_scaling_results.csv
"""
from __future__ import annotations

import argparse
import csv
import math
import sys
from typing import Dict, List, Optional, Sequence, Tuple

SCHEDULERS = (("ws", 1), ("gq", 2), ("chaselev", 3))
BLOCK_THREADS = 256


def median_iqr(times: Sequence[float]) -> Tuple[float, float, float]:
    """(median, median - q1, q3 - median), quartiles by nearest rank
    (worker_scaling_fib.sh:66-104)."""
    s = sorted(times)
    m = len(s)
    if m == 0:
        return (math.nan, math.nan, math.nan)
    med = s[m // 2] if m % 2 == 1 else (s[m // 2 - 1] + s[m // 2]) / 2
    q1 = s[min(max((m * 25 + 99) // 100 - 1, 0), m - 1)]
    q3 = s[min(max((m * 75 + 99) // 100 - 1, 0), m - 1)]
    return (med, med - q1, q3 - med)


def _workload(gtap_amd, name: str, n: int):
    if name == "fib":
        return lambda: gtap_amd.fib(n)
    if name == "nq":
        return lambda: gtap_amd.nqueens(n)
    raise ValueError(f"unknown workload {name!r} (fib, nq)")


def sweep(workload: str, n: int, grid_sizes: Sequence[int], runs: int = 20,
          schedulers: Sequence[Tuple[str, int]] = SCHEDULERS, timeout_ms: int = 60000,
          log=lambda s: None) -> List[Dict[str, object]]:
    """One row per grid size (workgroups); every scheduler timed `runs` times at that size.
    All runs of a workload must return the same value."""
    import gtap_amd
    rows: List[Dict[str, object]] = []
    ref: List[int] = []
    for g in grid_sizes:
        row: Dict[str, object] = {"block_size": BLOCK_THREADS, "total_threads": g * BLOCK_THREADS, "grid_size": g}
        for name, sched in schedulers:
            gtap_amd.gtap_finalize()
            if gtap_amd.gtap_set_config(timeout_ms=timeout_ms, blocks=g, scheduler=sched) != 0:
                raise RuntimeError("gtap_set_config failed")
            st = gtap_amd.gtap_initialize()
            if st != 0:
                raise RuntimeError(f"gtap_initialize failed: {st}")
            row["max_tasks_per_warp"] = gtap_amd.gtap_get_config()["deque_capacity"]
            run = _workload(gtap_amd, workload, n)
            times: List[float] = []
            workers = 0
            for _ in range(runs):
                v, s = run()
                if not ref:
                    ref.append(v)
                elif v != ref[0]:
                    raise RuntimeError(f"{workload}({n}) under {name} at {g} workgroups returned {v}, not {ref[0]}")
                times.append(s["kernel_ms"])
                workers = s["workers"]
            med, lo, hi = median_iqr(times)
            row[f"{name}_med"], row[f"{name}_err_low"], row[f"{name}_err_high"] = med, lo, hi
            row["grid_size"] = workers // 4  # the workgroups actually launched (capped at co-residency)
            row["total_threads"] = workers // 4 * BLOCK_THREADS
            log(f"{workload}({n}) {name:8s} workgroups {workers // 4:5d}: median {med:.3f} ms "
                f"(-{lo:.3f} / +{hi:.3f})")
        rows.append(row)
    gtap_amd.gtap_finalize()
    return rows


COLUMNS = ["block_size", "total_threads", "grid_size", "max_tasks_per_warp"] + [
    f"{s}_{k}" for s, _ in SCHEDULERS for k in ("med", "err_low", "err_high")]


def write_csv(rows: List[Dict[str, object]], path: str) -> None:
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=COLUMNS, extrasaction="ignore")
        w.writeheader()
        for r in rows:
            w.writerow({k: (f"{v:.6f}" if isinstance(v, float) else v) for k, v in r.items()})


def main(argv: Optional[Sequence[str]] = None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--workload", default="fib", choices=("fib", "nq"))
    ap.add_argument("--n", type=int, default=None, help="fib n (default 30) or n-queens N (default 12)")
    ap.add_argument("--grid", default="4,8,16,32,64,128,256,512,0",
                    help="workgroup counts, comma separated; 0 = every co-resident workgroup")
    ap.add_argument("--runs", type=int, default=20)
    ap.add_argument("--schedulers", default="ws,gq,chaselev")
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    n = a.n if a.n is not None else (30 if a.workload == "fib" else 12)
    grid = [int(x) for x in a.grid.split(",") if x.strip()]
    wanted = [s.strip() for s in a.schedulers.split(",") if s.strip()]
    scheds = [(s, v) for s, v in SCHEDULERS if s in wanted]
    rows = sweep(a.workload, n, grid, a.runs, scheds, log=lambda s: print(s, file=sys.stderr, flush=True))
    out = a.out or f"{a.workload}_scaling_results.csv"
    write_csv(rows, out)
    print(out)
    return 0


if __name__ == "__main__":
    sys.exit(main())
