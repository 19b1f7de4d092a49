#!/usr/bin/env python3
"""Benchmark of the MI355X-native GTaP runtime -- prints ONE JSON line on stdout.

Workload (the config the headline metric is quoted on): n-queens N=14, thread-executed,
one task per legal placement, taskwait + int64 sum, no cutoff -- 27,358,553 tasks per run.
The counts are summed by the additive join (each child's count rides on the atomic that
counts its parent's join down); the result-slot form is timed beside it in `extra`.
A step is one persistent launch (gtap_nqueens(14)) with everything resident on the GPU.
`value` = tasks x steps / wall time of the timed steps (host clock around the steps,
each step ends with a stream synchronize). The roofline uses the average kernel duration
from hipEvents recorded on the launch stream.

Also reported (outside `value`): fib(30) tasks/s; cilksort of 2^26 uint32 keys; BFS on an
R-MAT scale-22 graph; the 4096-job mixed stream (batched, per-job launches, served); the
measured HBM-copy and atomic peaks of this GPU and the cost of fence vs store-wait publication;
and a CPU baseline (the reference's own sequential n-queens, built from its sources into
oracle/_ref/, timed on this host).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import re
import statistics
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

NQ_N = 14
# Published counts used as the check (OEIS A000170; the task tree of one task per legal
# placement; F(30) and the 2F(31)-1 tasks of the fib tree).
NQ14_SOLUTIONS = 365_596
NQ14_TASKS = 27_358_553
FIB30 = 832_040
FIB40 = 102_334_155
FIB30_TASKS = 2_692_537


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def _ref_ms(binary: str, args: list, runs: int, env=None, timeout: int = 300) -> list:
    """Run one of the reference's CPU programs `runs` times; its own "Execution time: X ms"."""
    out = []
    for _ in range(runs):
        r = subprocess.run([binary, *map(str, args)], capture_output=True, text=True, timeout=timeout, env=env)
        m = re.search(r"Execution time: ([0-9.]+) ms", r.stdout)
        if m:
            out.append(float(m.group(1)))
    return out


def _median_iqr(xs: list) -> dict:
    xs = sorted(xs)
    n = len(xs)
    rank = lambda q: xs[min(n - 1, max(0, int(-(-q * n // 1)) - 1))]  # nearest rank
    return {"median_ms": statistics.median(xs), "q1_ms": rank(0.25), "q3_ms": rank(0.75), "runs": n}


def cpu_baseline(nq_tasks: int, headline_only: bool = False) -> dict:
    """CPU baselines timed on this host.

    value: sequential n-queens(14) -- the reference's seq_nq when it was built (oracle/_ref),
    else the oracle's restatement. `programs`: the reference's own sequential and OpenMP
    programs on the other configs (fib(30), cilksort of 2^26 keys) and the mixed job stream
    on a bounded sample, with the thread count used."""
    import numpy as np

    import oracle
    ref = oracle.ref_binary("seq_nq")
    times = []
    if ref:
        times = [t / 1e3 for t in _ref_ms(ref, [NQ_N], 3)]
        kind = "reference"
        sample = f"reference seq_nq {NQ_N} (evaluation/2-comparison/nq/seq_nq.c), full workload, median of 3 runs"
    if not times:
        for _ in range(3):
            t0 = time.perf_counter()
            oracle.nq(NQ_N)
            times.append(time.perf_counter() - t0)
        kind = "port"
        sample = f"oracle nq({NQ_N}) (restatement of seq_nq.c), full workload, median of 3 runs"
    t = statistics.median(times)
    out = {"value": nq_tasks / t, "unit": "tasks/s", "cores": 1, "kind": kind, "sample": sample, "seconds": t}
    if headline_only:
        return out

    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_STACKSIZE="64M")
    progs = {"threads_openmp": threads, "cpu_model": _cpu_model()}
    runs = [("seq_fib", [30], 5, None), ("omp_fib", [30], 3, env), ("seq_nq", [NQ_N], 3, None),
            ("omp_nq", [NQ_N, 7], 3, env)]
    for name, args, n, e in runs:
        b = oracle.ref_binary(name)
        if b:
            ms = _ref_ms(b, args, n, env=e)
            if ms:
                progs[f"{name} {' '.join(map(str, args))}"] = _median_iqr(ms)
    gen = oracle.ref_binary("gen_vector")
    if gen and oracle.ref_binary("seq_cilksort"):
        path = os.path.join(os.environ.get("TMPDIR", "/tmp"), f"gtap_keys_{os.getpid()}.bin")
        try:
            subprocess.run([gen, str(1 << 26), path], capture_output=True, timeout=300, check=True)
            for name, n, e in (("seq_cilksort", 1, None), ("omp_cilksort", 1, env)):
                b = oracle.ref_binary(name)
                ms = _ref_ms(b, [path], n, env=e) if b else []
                if ms:
                    progs[f"{name} 2^26"] = _median_iqr(ms)
        finally:
            if os.path.exists(path):
                os.remove(path)
    # the mixed job stream: all 4096 jobs with an OpenMP parallel-for over the jobs, and the
    # sequential loop over the first 512 (the same programs as the GPU's jobs, in C)
    os.environ.setdefault("OMP_NUM_THREADS", str(threads))
    dt, _ = oracle.stream_cpu(2024, 4096, True)
    progs["stream 4096 jobs, OpenMP parallel for schedule(dynamic) over jobs"] = {
        "seconds": dt, "jobs_per_s": 4096 / dt, "threads": threads}
    dt, _ = oracle.stream_cpu(2024, 512, False)
    progs["stream sequential, first 512 of 4096 jobs"] = {"seconds": dt, "jobs_per_s": 512 / dt, "threads": 1}
    # sequential queue BFS (examples/asynchronous_bfs/asynchronous_bfs_thread.cu:151-174) on
    # the benchmark's R-MAT 22 graph
    import gtap_amd
    row, col = gtap_amd.rmat_csr(22, 16, 1)
    bfs_s = []
    for _ in range(3):
        t0 = time.perf_counter()
        d = oracle.bfs(row, col, 0)
        bfs_s.append(time.perf_counter() - t0)
    reached = d < oracle.INF_DEPTH
    edges = int(np.diff(row)[reached].sum())
    progs["bfs sequential R-MAT 22"] = {"median_ms": statistics.median(bfs_s) * 1e3, "reached_edges": edges,
                                        "edges_per_s": edges / statistics.median(bfs_s), "threads": 1}
    out["programs"] = progs
    return out


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def bench_cilksort(gtap_amd, roofline, hbm_meas, reps: int = 3) -> dict:
    """cilksort of 2^26 gen_vector keys (block mode), inputs resident on the GPU; the
    unsorted keys are restored with a device copy before each run (outside the kernel time).
    Runs at the autotuner's committed settings for this workload."""
    import ctypes
    import numpy as np
    tuned = use_tuned("cilksort")
    leaf, merge_leaf = int(tuned.get("leaf", 4096)), int(tuned.get("merge_leaf", 0))
    n = 1 << 26
    keys = gtap_amd.gen_keys(n)
    src = gtap_amd.DeviceBuffer.from_array(keys)
    work = gtap_amd.DeviceBuffer(keys.nbytes)
    tmp = gtap_amd.DeviceBuffer(keys.nbytes)
    L = gtap_amd.lib()
    ms = []
    for _ in range(reps + 1):
        L.gtap_memcpy_dtod(work.ptr, src.ptr, keys.nbytes)
        s = gtap_amd.RunStats()
        st = L.gtap_cilksort(work.ptr, tmp.ptr, n, leaf, merge_leaf, ctypes.byref(s))
        if st != 0:
            raise RuntimeError(f"gtap_cilksort failed: {st}")
        ms.append(s.kernel_ms)
    out = work.to_array(np.empty(n, dtype=np.uint32))
    ok = bool(np.array_equal(out, np.sort(keys)))
    if not ok:
        raise RuntimeError("cilksort 2^26: the output is not the sorted input")
    for b in (src, work, tmp):
        b.free()
    med = statistics.median(ms[1:])
    gbps = roofline.cilksort_bytes(n) / (med / 1e3) / 1e9
    log(f"cilksort 2^26: {med:.3f} ms, {gbps:.0f} GB/s algorithmic, sorted={ok}")
    return {"kernel_ms_median": med, "keys_per_s": n / (med / 1e3), "bit_exact": ok,
            "tasks_per_s": s.tasks_executed / (med / 1e3), "tuned": tuned,
            "roofline": {"bound": "hbm", "achieved": gbps, "peak": roofline.MI355X_HBM_GBPS, "unit": "GB/s",
                         "frac": gbps / roofline.MI355X_HBM_GBPS, "measured_copy_peak": hbm_meas,
                         "traffic": load_traffic("cilksort_2e26")}}


def bench_bfs(gtap_amd, roofline, hbm_meas, rand_meas=None, reps: int = 3) -> dict:
    """Asynchronous BFS on the R-MAT scale-22 graph (edge factor 16, seed 1), EPAQ with 4
    degree-bucket queues vs one queue, each at its own tuned settings; CSR resident on the GPU."""
    import ctypes
    import numpy as np
    row, col = gtap_amd.rmat_csr(22, 16, 1)
    nv = row.size - 1
    drow = gtap_amd.DeviceBuffer.from_array(row)
    dcol = gtap_amd.DeviceBuffer.from_array(col)
    ddep = gtap_amd.DeviceBuffer(nv * 4)
    L = gtap_amd.lib()
    res = {}
    tuned = {}
    depth = np.empty(nv, dtype=np.int32)
    for q in (4, 1):
        tuned[q] = use_tuned("bfs" if q == 4 else "bfs1")
        split = int(tuned[q].get("split", 64))
        ms = []
        for _ in range(reps + 1):
            s = gtap_amd.RunStats()
            st = L.gtap_bfs(drow.ptr, dcol.ptr, nv, col.size, 0, ddep.ptr, q, split, ctypes.byref(s))
            if st != 0:
                raise RuntimeError(f"gtap_bfs failed: {st}")
            ms.append(s.kernel_ms)
        res[q] = (statistics.median(ms[1:]), s.tasks_executed, s.rounds, s.steal_attempts)
    ddep.to_array(depth)
    for b in (drow, dcol, ddep):
        b.free()
    reached = depth < gtap_amd.INF_DEPTH
    vr = int(reached.sum())
    er = int(np.diff(row)[reached].sum())
    # the distances' invariants (unique shortest-hop distances): depth[0] = 0, no edge from a
    # reached vertex improves its head, every other reached vertex has a tight in-edge
    src = np.repeat(np.arange(nv, dtype=np.int64), np.diff(row))
    dv, du = depth[src].astype(np.int64), depth[col].astype(np.int64)
    rs = dv < gtap_amd.INF_DEPTH
    tight = np.zeros(nv, dtype=bool)
    tight[col[rs & (du == dv + 1)]] = True
    tight[0] = True
    if depth[0] != 0 or np.any(du[rs] > dv[rs] + 1) or not np.all(tight[reached]):
        raise RuntimeError("bfs rmat22: the distances break the BFS invariants")
    med4, tasks4 = res[4][0], res[4][1]
    gbps = roofline.bfs_bytes(vr, er) / (med4 / 1e3) / 1e9
    log(f"bfs rmat22: epaq4 {med4:.3f} ms, 1 queue {res[1][0]:.3f} ms, {gbps:.0f} GB/s algorithmic")
    # request-rate roofline: the expansion's random accesses (a depth probe per reached edge,
    # the stale check and the row-offset pair per reached vertex) against the measured random
    # 4-byte load rate over a table the size of depth[]
    reqs = er + 2 * vr
    req_rate = reqs / (med4 / 1e3) / 1e9
    request_roofline = None
    if rand_meas:
        request_roofline = {"random_requests": reqs, "achieved_greq_s": req_rate,
                            "measured_random_load_greq_s": rand_meas["load_gps"],
                            "measured_random_fetch_min_gops": rand_meas["fetch_min_gps"],
                            "frac": req_rate / rand_meas["load_gps"]}
        log(f"bfs rmat22 requests: {req_rate:.1f} G random accesses/s of {rand_meas['load_gps']:.1f} measured "
            f"({req_rate / rand_meas['load_gps']:.2f})")
    return {"kernel_ms_median_epaq4": med4, "kernel_ms_median_1queue": res[1][0],
            "request_roofline": request_roofline,
            "epaq_speedup": res[1][0] / med4, "reached_vertices": vr, "reached_edges": er,
            "edges_per_s": er / (med4 / 1e3), "tasks_executed_epaq4": tasks4,
            "tuned": {"4_queues": tuned[4], "1_queue": tuned[1]},
            "counts": {f"{q}_queues": {"tasks_executed": r[1], "rounds": r[2], "steal_attempts": r[3]} for q, r in res.items()},
            "roofline": {"bound": "hbm", "achieved": gbps, "peak": roofline.MI355X_HBM_GBPS, "unit": "GB/s",
                         "frac": gbps / roofline.MI355X_HBM_GBPS, "measured_copy_peak": hbm_meas,
                         "algorithmic_bytes": roofline.bfs_bytes(vr, er), "traffic": load_traffic("bfs_rmat22")}}


def bench_spmv_tree(gtap_amd, reps: int = 3) -> dict:
    """The reference's other example workloads at their default sizes: SpMV of a random
    1,000,000 x 1,000,000 CSR matrix with about 300 entries per row (examples/spmv/
    spmv_thread.cu:123-126, range cutoff 16) and the binary tree of height 15 with 64 random
    reads and 512 fma steps per node (binary_tree/binary_tree_thread.cu:122-126). Matrix,
    vectors and input resident on the GPU; kernel time per launch."""
    import ctypes
    import numpy as np
    L = gtap_amd.lib()
    out = {}
    rp, ci, va = gtap_amd.spmv_random(1_000_000, 300, 42)
    rows = rp.size - 1
    x = np.ones(rows)
    bufs = [gtap_amd.DeviceBuffer.from_array(a) for a in (rp, ci, va, x)]
    dy = gtap_amd.DeviceBuffer(rows * 8)
    ms = []
    for _ in range(reps + 1):
        s = gtap_amd.RunStats()
        st = L.gtap_spmv(bufs[0].ptr, bufs[1].ptr, bufs[2].ptr, bufs[3].ptr, dy.ptr, rows, rows, 16, ctypes.byref(s))
        if st != 0:
            raise RuntimeError(f"gtap_spmv failed: {st}")
        ms.append(s.kernel_ms)
    y = dy.to_array(np.empty(rows))
    for b in bufs + [dy]:
        b.free()
    # x = 1: every row sum is the fma chain of the row's values, checked on a sample of rows
    sample = np.arange(0, rows, 9973)
    for i in sample:
        acc = 0.0
        for v in va[rp[i]:rp[i + 1]]:
            acc = acc + float(v)  # v * 1.0 + acc: fma with a unit factor rounds like the add
        if y[i] != acc:
            raise RuntimeError(f"spmv row {i}: {y[i]} != {acc}")
    med = statistics.median(ms[1:])
    nnz = int(rp[-1])
    out["spmv_1m_x300"] = {"kernel_ms_median": med, "rows": rows, "nnz": nnz, "nnz_per_s": nnz / (med / 1e3),
                           "tasks_executed": s.tasks_executed,
                           "algorithmic_gbps": (12 * nnz + 8 * nnz + 4 * (rows + 1) + 8 * rows) / (med / 1e3) / 1e9}
    log(f"spmv 1M x ~300: {med:.3f} ms, {nnz / (med / 1e3) / 1e9:.2f} G nnz/s")
    inp = gtap_amd.tree_input(1 << 20)
    ms = []
    for _ in range(reps + 1):
        o, ts = gtap_amd.tree(15, 64, 512, inp)
        ms.append(ts["kernel_ms"])
    med = statistics.median(ms[1:])
    nodes = (1 << 16) - 1
    out["tree_h15_m64_c512"] = {"kernel_ms_median": med, "nodes": nodes, "nodes_per_s": nodes / (med / 1e3)}
    log(f"tree height 15 (64 reads, 512 fma per node): {med:.3f} ms")
    return out


# The reference's published medians on 1x GH200 for the same programs at its own settings
# (evaluation/2-comparison/nq/nq_performance_results.csv:15-17, fib_performance_results.csv:16,
# 1-worker_scalability/fib/fib_scaling_results.csv:10).
REFERENCE_GH200_MS = {"nq14_cutoff7": 3.0135, "nq15_cutoff7": 17.175, "nq16_cutoff7": 115.6665,
                      "fib30": 0.958, "fib35": 7.778}


def bench_epaq_fib(gtap_amd, n: int = 40, cutoffs=(13, 16), reps: int = 5) -> dict:
    """EPAQ where the reference shows it paying off (evaluation/3-epaq/fib/fib_queue_3.cu:16-26,
    epaq_performance_results_fib.csv:13): fib(n) with a host-set cutoff below which the call
    recurses sequentially, lowered by gtap_amd.lower from gtap_amd/lower/examples/fib_epaq.gtap.cu
    (children bound for the sequential path on queue 1, the others on queue 0, parents resuming
    on queue 2) and fib_cut.gtap.cu (the same on one queue), launched by the host the way the
    reference's main() launches it; hipEvent time of the launch, median of `reps`."""
    import ctypes
    from gtap_amd import build as gbuild
    mods = {name: gtap_amd.EntryModule(gbuild.lowered_lib(name), name) for name in ("fib_cut", "fib_epaq")}
    ll = ctypes.c_longlong
    out = {"n": n, "reference_gh200_ms": {"1 queue": 6.036, "3 queues": 3.3495, "cutoff": 13, "n": 40}}
    expect = None
    for cut in cutoffs:
        row = {}
        for name, m in mods.items():
            m.set_symbol("d_cutoff", cut)
            ms = []
            for _ in range(reps):
                m.set_symbol("d_result", -1, ll)
                st, t = m.launch("fib_kernel", [ctypes.c_int], n, timed=True)
                if st != 0 or gtap_amd.gtap_get_runtime_error_code() != 0:
                    raise RuntimeError(f"{name}: launch failed")
                v = m.get_symbol("d_result", ll)
                expect = v if expect is None else expect
                if v != expect:
                    raise RuntimeError(f"{name}: fib({n}) = {v}, other runs gave {expect}")
                ms.append(t)
            row["3_queues_ms" if name == "fib_epaq" else "1_queue_ms"] = statistics.median(ms)
        row["epaq_speedup"] = row["1_queue_ms"] / row["3_queues_ms"]
        out[f"cutoff_{cut}"] = row
        log(f"EPAQ fib({n}) cutoff {cut}: 1 queue {row['1_queue_ms']:.3f} ms, 3 queues {row['3_queues_ms']:.3f} ms "
            f"({row['epaq_speedup']:.2f}x)")
    if expect != FIB40:
        raise RuntimeError(f"fib({n}) = {expect}, expected {FIB40}")
    out["result"] = expect
    return out


def bench_lowered_fib(gtap_amd, n: int = 30, reps: int = 5) -> dict:
    """The lowering's own output against the hand-lowered program: fib(n) from
    gtap_amd/lower/examples/fib.gtap.cu (written in the reference's dialect with a 64-bit
    result, lowered by gtap_amd.lower to result slots, launched by the host as
    fib_kernel<<<grid, 256>>>), from fib32.gtap.cu (the 32-bit result of the reference's
    examples/fib/fib.cu, which the lowering joins additively) and gtap_fib(n), all with no
    cutoff; hipEvent time of the launch, median of `reps`."""
    import ctypes
    from gtap_amd import build as gbuild

    def run_lowered(name, kernel, sym, ctype):
        m = gtap_amd.EntryModule(gbuild.lowered_lib(name), name)
        ms = []
        for _ in range(reps + 1):
            m.set_symbol(sym, -1, ctype)
            st, t = m.launch(kernel, [ctypes.c_int], n, timed=True)
            if st != 0 or m.get_symbol(sym, ctype) != FIB30:
                raise RuntimeError(f"lowered {name}: wrong result")
            ms.append(t)
        return ms

    lowered = run_lowered("fib", "fib_kernel", "d_fib_result", ctypes.c_longlong)
    lowered32 = run_lowered("fib32", "fib32_kernel", "d_fib32_result", ctypes.c_int)
    builtin = []
    for _ in range(reps + 1):
        v, s = gtap_amd.fib(n)
        if v != FIB30:
            raise RuntimeError("gtap_fib: wrong result")
        builtin.append(s["kernel_ms"])
    out = {"n": n, "lowered_host_launch_ms": statistics.median(lowered[1:]),
           "lowered_int32_additive_join_ms": statistics.median(lowered32[1:]),
           "hand_lowered_ms": statistics.median(builtin[1:])}
    out["lowered_over_hand"] = out["lowered_host_launch_ms"] / out["hand_lowered_ms"]
    out["lowered_int32_over_hand"] = out["lowered_int32_additive_join_ms"] / out["hand_lowered_ms"]
    log(f"fib({n}): lowered source, host launch {out['lowered_host_launch_ms']:.3f} ms (result slots), "
        f"{out['lowered_int32_additive_join_ms']:.3f} ms (32-bit, additive join); hand-lowered "
        f"{out['hand_lowered_ms']:.3f} ms")
    return out


def bench_round_phases(gtap_amd) -> dict:
    """Per-round breakdown of the thread-mode loop for n-queens(14) and fib(30): the profiling
    build of the kernel stamps the shader clock at each phase boundary of every round (acquire,
    record loads, task bodies, record writes issued, their completion, join atomics, publish)
    and sums per worker. This build waits for the record loads and the join atomics before
    stamping, so its rounds are slower than the plain kernel's; the shares show where a round
    goes. The clock rate is estimated from the total stamped cycles and the kernel time."""
    out = {}
    gtap_amd.gtap_set_profiling(True, 2)
    try:
        for name, fn, arg in (("nq14", gtap_amd.nqueens, NQ_N), ("fib30", gtap_amd.fib, 30)):
            fn(arg)  # warm-up
            v, st = fn(arg)
            ph = gtap_amd.round_phases()
            total = sum(ph["cycles"].values())
            ghz = total / max(1, st["workers"]) / (st["kernel_ms"] * 1e6)
            busy = total - ph["cycles"]["idle"]
            share = {k: ph["cycles"][k] / busy for k in gtap_amd.ROUND_PHASES[:7]} if busy else {}
            per_round_us = {k: c / ghz / 1e3 for k, c in ph["cycles_per_round"].items()} if ghz else {}
            out[name] = {"kernel_ms_profiling_build": st["kernel_ms"], "rounds": ph["rounds"],
                         "clock_ghz_estimate": ghz, "cycles_per_round": ph["cycles_per_round"],
                         "us_per_round": per_round_us, "share_of_busy_time": share,
                         "idle_share": ph["cycles"]["idle"] / total if total else None}
            log(f"round phases {name}: " + ", ".join(f"{k} {v * 100:.0f}%" for k, v in share.items()))
    finally:
        gtap_amd.gtap_set_profiling(False, 0)
    return out


def bench_reference_settings(gtap_amd, reps: int = 5) -> dict:
    """The reference's own benchmark programs at its own settings, timed here, beside the
    medians it publishes for 1x GH200: n-queens with tasks to row 7 and a sequential search
    below (no taskwait, one counter), and fib(35)."""
    out = {}
    for n, exp in ((14, 365596), (15, 2279184), (16, 14772512)):
        ms = []
        for _ in range(reps):
            v, s = gtap_amd.nqueens_cutoff(n, 7)
            if v != exp:
                raise RuntimeError(f"nqueens_cutoff({n}) = {v}")
            ms.append(s["kernel_ms"])
        key = f"nq{n}_cutoff7"
        med = statistics.median(ms)
        out[key] = {"kernel_ms_median": med, "reference_gh200_ms": REFERENCE_GH200_MS[key],
                    "speedup_vs_reference": REFERENCE_GH200_MS[key] / med}
    ms = []
    for _ in range(reps):
        v, s = gtap_amd.fib(35)
        if v != 9227465:
            raise RuntimeError(f"fib(35) = {v}")
        ms.append(s["kernel_ms"])
    med = statistics.median(ms)
    out["fib35"] = {"kernel_ms_median": med, "reference_gh200_ms": REFERENCE_GH200_MS["fib35"],
                    "speedup_vs_reference": REFERENCE_GH200_MS["fib35"] / med}
    log("reference settings: " + ", ".join(f"{k} {v['kernel_ms_median']:.3f} ms "
                                           f"({v['speedup_vs_reference']:.2f}x the GH200 median)" for k, v in out.items()))
    return out


def _rate_at_p99(points, form: str, p99_ms: float):
    """Largest arrival rate of the sweep at which `form` kept its p99 within p99_ms, with
    log-linear interpolation to the crossing between the last point inside and the first
    point outside; None if even the lowest rate exceeds it. ('lower_bound' when the form is
    still inside at the highest rate swept.)"""
    import math
    inside = None
    for i, p in enumerate(points):
        if form not in p:
            break
        if p[form]["p99_ms"] <= p99_ms:
            inside = i
            continue
        if inside is None:
            return None, False
        a, b = points[inside], p
        la, lb = math.log(a[form]["p99_ms"]), math.log(b[form]["p99_ms"])
        t = (math.log(p99_ms) - la) / (lb - la) if lb > la else 0.0
        ra, rb = math.log(a["rate_jobs_per_s"]), math.log(b["rate_jobs_per_s"])
        return math.exp(ra + t * (rb - ra)), False
    if inside is None:
        return None, False
    return points[inside]["rate_jobs_per_s"], True


def served_equal_p99(gtap_amd, seed: int, njobs: int, expected, lo: float = 1000.0, hi: float = 80000.0,
                     step: float = 1.25, merge_leaf: int = 4096) -> dict:
    """Poisson arrivals of the first `njobs` jobs of the stream at rates lo, lo*step, ... hi,
    served by one resident launch and by one launch per job; latency = arrival to the result
    seen by the host, for both. The per-job form is swept until it saturates (it completes
    under 80 % of the offered rate twice in a row). The target metric: at the p99 the per-job
    form keeps at its highest unsaturated rate, the rate the resident form sustains within the
    same p99, over the per-job form's rate."""
    import numpy as np
    points = []
    sv = gtap_amd.Stream(seed=seed, njobs=njobs)
    sv.set_merge_leaf(merge_leaf)
    saturated = 0
    try:
        rate = lo
        while rate <= hi * 1.0001:
            row = {"rate_jobs_per_s": rate}
            forms = (("resident", False),) if saturated >= 2 else (("resident", False), ("per_job_launch", True))
            for form, per_job in forms:
                r, lat, tot, st = sv.serve(rate=rate, arrival_seed=5, per_job_launch=per_job)
                if not np.array_equal(r, expected):
                    raise RuntimeError(f"stream: served ({form}) results differ from the batched run")
                ex = sv.exec_ms()
                row[form] = {"p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                             "exec_p50_ms": float(np.percentile(ex, 50)), "jobs_per_s": njobs / (tot / 1e3)}
                if not per_job:
                    row[form]["admission_holds"] = int(st.get("admission_holds", 0))
            if "per_job_launch" in row:
                saturated = saturated + 1 if row["per_job_launch"]["jobs_per_s"] < 0.8 * rate else 0
            points.append(row)
            pj = row.get("per_job_launch")
            log(f"stream served at {rate:.0f} jobs/s: resident p99 {row['resident']['p99_ms']:.3f} ms"
                + (f", per-job launch p99 {pj['p99_ms']:.3f} ms ({pj['jobs_per_s']:.0f} jobs/s)" if pj else ""))
            rate *= step
    finally:
        sv.close()
    # the per-job form's best operating point: its highest rate completed at >= 95 % of the offer
    ok = [p for p in points if "per_job_launch" in p and p["per_job_launch"]["jobs_per_s"] >= 0.95 * p["rate_jobs_per_s"]]
    out = {"jobs": njobs, "arrival_seed": 5, "rate_step": step, "latency": "arrival to result seen by the host",
           "points": points}
    if ok:
        ref = ok[-1]
        target = ref["per_job_launch"]["p99_ms"]
        r_res, bound = _rate_at_p99(points, "resident", target)
        r_pj, _ = _rate_at_p99(points, "per_job_launch", target)
        out["equal_p99"] = {
            "p99_ms": target,
            "per_job_launch_jobs_per_s": r_pj,
            "resident_jobs_per_s": r_res,
            "resident_is_lower_bound": bound,  # still within the p99 at the highest rate swept
            "resident_over_per_job": (r_res / r_pj) if (r_res and r_pj) else None,
        }
        log(f"stream at equal p99 {target:.2f} ms: resident {r_res:.0f} jobs/s vs one launch per job {r_pj:.0f} "
            f"jobs/s = {r_res / r_pj:.1f}x" + (" (lower bound)" if bound else ""))
    return out


def bench_stream(gtap_amd, njobs: int = 4096, seed: int = 2024, serve_jobs: int = 1024) -> dict:
    """The 4096-job mixed stream (fib / n-queens / sort jobs, fixed seed): one resident batched
    launch against one launch per job (the forms must agree job for job; the GPU tests check
    every result, sort jobs by the checksum of their sorted keys, against the oracle); then
    Poisson arrivals (the first `serve_jobs` jobs) swept in x1.25 steps, served by one resident
    launch and by one launch per job, with p50/p99 latency from arrival to the result seen by
    the host, and the rate each sustains at equal p99."""
    import numpy as np
    tuned = use_tuned("stream")
    merge_leaf = int(tuned.get("merge_leaf", 4096))
    st = gtap_amd.Stream(seed=seed, njobs=njobs)
    st.set_merge_leaf(merge_leaf)
    try:
        bat = []
        for _ in range(3):
            res, lat, s = st.run_batched()
            bat.append((s["kernel_ms"], res, lat))
        kms = statistics.median(b[0] for b in bat)
        res_b, lat_b = bat[-1][1], bat[-1][2]
        res_p, lat_p, tot_p = st.run_per_job()
        if not np.array_equal(res_b, res_p):
            raise RuntimeError("stream: batched and per-job results differ")
    finally:
        st.close()
    out = {
        "jobs": njobs,
        "tuned": tuned,
        "batched": {"kernel_ms_median": kms, "jobs_per_s": njobs / (kms / 1e3),
                    "p50_ms": float(np.percentile(lat_b, 50)), "p99_ms": float(np.percentile(lat_b, 99))},
        # every job arrives at the start in both forms: latency = its completion after the start
        "per_job_launch": {"total_ms": tot_p, "jobs_per_s": njobs / (tot_p / 1e3),
                           "p50_ms": float(np.percentile(lat_p, 50)), "p99_ms": float(np.percentile(lat_p, 99))},
    }
    out["batched_over_per_job"] = out["batched"]["jobs_per_s"] / out["per_job_launch"]["jobs_per_s"]
    out["served"] = served_equal_p99(gtap_amd, seed, serve_jobs, res_b[:serve_jobs], merge_leaf=merge_leaf)
    log(f"stream: batched {out['batched']['jobs_per_s']:.0f} jobs/s, per-job {out['per_job_launch']['jobs_per_s']:.0f}"
        f" jobs/s ({out['batched_over_per_job']:.1f}x)")
    return out


def use_tuned(name: str) -> dict:
    """Re-initialises the runtime with the autotuner's committed best point for `name` (its
    gtap_config fields; the program arguments among them are returned for the caller) and
    returns that point ({} = defaults)."""
    from gtap_amd import autotune
    params = autotune.tuned(name)
    autotune.apply(params, timeout_ms=30000)
    log(f"tuned settings for {name}: {params or 'defaults'}")
    return params


def load_traffic(name: str):
    """HBM bytes per launch from the committed rocprofv3 PMC summary, if present."""
    p = os.path.join(ROOT, "profiles", f"{name}_traffic.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        d = json.load(f)
    return d.get("traffic_bytes")


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--headline-only", action="store_true",
                    help="time the n-queens workload only (no fib/cilksort/BFS/stream extras); for profiler runs")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (the workload would shard by "
              f"independent subtrees of the task tree)", file=sys.stderr)
        return 2

    import gtap_amd
    from gtap_amd import roofline
    from gtap_amd.scaling import median_iqr

    nq_solutions, nq_tasks = NQ14_SOLUTIONS, NQ14_TASKS

    gtap_amd.gtap_set_config(timeout_ms=30000)
    st = gtap_amd.gtap_initialize()
    if st != 0:
        raise RuntimeError(f"gtap_initialize failed: {st}")
    tuned_nq = use_tuned("nq14")
    info = gtap_amd.device_info()
    log("device", info)
    hbm_peak_meas = gtap_amd.bench_hbm_copy(2 << 30, 3)
    atomic_peak_meas = gtap_amd.bench_atomics(3)
    atomic_contended = {str(w): gtap_amd.bench_atomics_contended(w, 3) for w in (1, 64, 4096)}
    rand_meas = gtap_amd.bench_random_access(16 << 20, 3)  # the size of depth[] at R-MAT scale 22
    log(f"measured peaks: hbm copy {hbm_peak_meas:.0f} GB/s, atomics {atomic_peak_meas:.2f} Gop/s "
        f"(contended over 1 / 64 / 4096 words: " + " / ".join(f"{v:.2f}" for v in atomic_contended.values()) + " Gop/s)")

    for i in range(args.warmup):
        v, s = gtap_amd.nqueens(NQ_N)
        assert v == nq_solutions, (v, nq_solutions)
        log(f"warmup {i}: {s['kernel_ms']:.3f} ms")
    kernel_ms = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        v, s = gtap_amd.nqueens(NQ_N)
        kernel_ms.append(s["kernel_ms"])
        if v != nq_solutions or s["tasks_spawned"] + 1 != nq_tasks:
            raise RuntimeError(f"nqueens({NQ_N}) returned {v} with {s['tasks_spawned'] + 1} tasks, "
                               f"expected {nq_solutions} with {nq_tasks}")
    wall = time.perf_counter() - t0
    workers = s["workers"]
    log(f"timed {args.steps} steps: wall {wall*1e3:.3f} ms, kernel median {statistics.median(kernel_ms):.3f} ms")

    value = nq_tasks * args.steps / wall
    avg_kernel_s = statistics.mean(kernel_ms) / 1e3
    achieved_gbps = roofline.nq_bytes(nq_tasks - 1) / avg_kernel_s / 1e9
    traffic = load_traffic("nq14")
    atomics_per_s = roofline.ATOMICS_PER_TASK * (nq_tasks - 1) / avg_kernel_s

    # fib(30): the other scheduler-bound config
    fib_tasks = FIB30_TASKS
    fib_ms = []
    tuned_fib = {} if args.headline_only else use_tuned("fib30")
    for _ in range(0 if args.headline_only else 5):
        fv, fs = gtap_amd.fib(30)
        if fv != FIB30 or fs["tasks_spawned"] + 1 != fib_tasks:
            raise RuntimeError(f"fib(30) returned {fv} with {fs['tasks_spawned'] + 1} tasks")
        fib_ms.append(fs["kernel_ms"])
    fib_med = statistics.median(fib_ms) if fib_ms else None

    # the same n-queens tasks with result slots (store + join atomic per child)
    slot_ms = []
    if not args.headline_only:
        use_tuned("nq14")
    for _ in range(0 if args.headline_only else 5):
        sv, ss = gtap_amd.nqueens_slots(NQ_N)
        if sv != nq_solutions or ss["tasks_spawned"] + 1 != nq_tasks:
            raise RuntimeError(f"nqueens_slots({NQ_N}) returned {sv} with {ss['tasks_spawned'] + 1} tasks")
        slot_ms.append(ss["kernel_ms"])
    slot_med = statistics.median(slot_ms) if slot_ms else None

    extra_sort = extra_bfs = extra_stream = fences = extra_more = extra_ref = extra_epaq = extra_phases = extra_lowered = None
    if not args.headline_only:
        extra_sort = bench_cilksort(gtap_amd, roofline, hbm_peak_meas)
        extra_bfs = bench_bfs(gtap_amd, roofline, hbm_peak_meas, rand_meas)
        extra_stream = bench_stream(gtap_amd)
        from gtap_amd import autotune
        autotune.apply({}, timeout_ms=30000)  # the remaining extras run at the runtime's defaults
        extra_more = bench_spmv_tree(gtap_amd)
        extra_ref = bench_reference_settings(gtap_amd)
        extra_epaq = bench_epaq_fib(gtap_amd)
        extra_phases = bench_round_phases(gtap_amd)
        extra_lowered = bench_lowered_fib(gtap_amd)
        fences = gtap_amd.bench_fences(max(4, workers // 4 * 4), 64)

    cpu = None if args.no_cpu else cpu_baseline(nq_tasks, args.headline_only)
    gtap_amd.gtap_finalize()

    line = {
        "metric": "tasks/sec (fib, nq, cilksort); batched-stream jobs/sec + p99 latency, 1 GPU",
        "value": value,
        "unit": "tasks/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64",
        "data": "synthetic",
        "config": {
            "workload": f"nqueens N={NQ_N} thread-executed, task per legal placement, taskwait + int64 sum "
                        f"(additive join), no cutoff ({nq_tasks} tasks per run)",
            "n": NQ_N,
            "tasks_per_step": nq_tasks,
            "workers": workers,
            "parallelism": "single GPU, wavefront workers",
            # the autotuner's best point on an MI355X for this workload (gtap_config fields;
            # {} = the runtime's defaults)
            "tuned": tuned_nq,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved_gbps,
            "peak": roofline.MI355X_HBM_GBPS,
            "unit": "GB/s",
            "frac": achieved_gbps / roofline.MI355X_HBM_GBPS,
            "traffic": traffic,
        },
        "cpu_baseline": None if cpu is None else {k: cpu[k] for k in ("value", "unit", "cores", "kind", "sample", "programs")
                                                  if k in cpu},
        "extra": {
            "kernel_ms_median": statistics.median(kernel_ms),
            # distances of the nearest-rank quartiles from the median over the timed steps
            # (evaluation/1-worker_scalability/fib/worker_scaling_fib.sh:66-104)
            "kernel_ms_err_low_high": list(median_iqr(kernel_ms)[1:]),
            "kernel_tasks_per_s": nq_tasks / (statistics.median(kernel_ms) / 1e3),
            "atomic_roofline": {
                "achieved_gops": atomics_per_s / 1e9,
                "measured_peak_gops": atomic_peak_meas,
                "contended_gops_by_words": atomic_contended,
                "frac": atomics_per_s / 1e9 / atomic_peak_meas if atomic_peak_meas else None,
            },
            "hbm_measured_copy_gbps": hbm_peak_meas,
            "random_access_measured": rand_meas,
            "nq14_slot_join": None if slot_med is None else {"kernel_ms_median": slot_med,
                                                             "tasks_per_s": nq_tasks / (slot_med / 1e3)},
            "fib30": None if fib_med is None else {"kernel_ms_median": fib_med, "tasks": fib_tasks, "tuned": tuned_fib,
                                                   "tasks_per_s": fib_tasks / (fib_med / 1e3),
                                                   "reference_gh200_ms": REFERENCE_GH200_MS["fib30"]},
            "cilksort_2e26": extra_sort,
            "bfs_rmat22": extra_bfs,
            "stream_4096": extra_stream,
            "spmv_tree": extra_more,
            "reference_settings": extra_ref,
            "epaq_fib": extra_epaq,
            "round_phases": extra_phases,
            "lowered_fib30": extra_lowered,
            "publication_order_ns_per_wave": fences,
            "device": info,
        },
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
