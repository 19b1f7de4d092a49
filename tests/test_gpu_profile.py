"""Timeline profiling: the four CSV files of the reference's profiler
(runtime/detail/thread/gtap_profile_thread.cuh:29-108, 188-268), same columns, and their
task totals agree with the scheduler's own counters."""
import csv
import os

import numpy as np
import pytest

import gtap_amd as G
import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def runtime():
    G.gtap_set_config(timeout_ms=15000)
    assert G.gtap_initialize() == 0
    yield
    assert G.gtap_finalize() == 0

HEADERS = {
    "warp_timeline_having_task": "warp_id,time_index,timestamp_ns,relative_time_ms,state,state_description",
    "warp_statistics_having_task": "warp_id,total_samples,first_activity_ms,last_activity_ms,duration_ms",
    "warp_timeline_working": "warp_id,time_index,timestamp_ns,relative_time_ms,state,state_description,tasks_in_batch",
    "warp_statistics_working": "warp_id,total_samples,first_activity_ms,last_activity_ms,duration_ms,total_tasks",
}


def _read(d, app, kind):
    with open(os.path.join(d, "profile", f"{app}_{kind}.csv")) as f:
        lines = f.read().splitlines()
    assert lines[0] == HEADERS[kind]
    return list(csv.DictReader(lines))


@pytest.mark.parametrize("max_samples", [1 << 14, 64])
def test_profile_fib(tmp_path, monkeypatch, max_samples):
    monkeypatch.chdir(tmp_path)
    n = 22
    print(f"case profile fib({n}) max_samples={max_samples}", flush=True)
    assert G.gtap_set_profiling(True, max_samples) == 0
    try:
        v, st = G.fib(n)
        assert v == O.fib(n)
        G.lib().gtap_visualize_profile(b"fib")
    finally:
        assert G.gtap_set_profiling(False, 0) == 0
    work = _read(tmp_path, "fib", "warp_statistics_working")
    assert len(work) == st["workers"]
    tl = _read(tmp_path, "fib", "warp_timeline_working")
    for r in tl:
        assert r["state_description"] == ("Working" if int(r["time_index"]) % 2 == 0 else "NotWorking")
        assert int(r["time_index"]) < max_samples
    total = sum(int(r["total_tasks"]) for r in work)
    if max_samples >= 1 << 14:  # no worker ran out of samples: every executed task is counted
        assert total == st["tasks_executed"] == 3 * O.fib(n + 1) - 2
    else:
        assert 0 < total <= st["tasks_executed"]
    hav = _read(tmp_path, "fib", "warp_timeline_having_task")
    assert hav and all(r["state_description"] in ("Having", "NotHaving") for r in hav)
    hs = _read(tmp_path, "fib", "warp_statistics_having_task")
    assert len(hs) == st["workers"]
    for r in hs:
        assert float(r["duration_ms"]) >= 0.0


def test_profile_off_writes_counters_only(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)
    assert G.gtap_set_profiling(False, 0) == 0
    v, st = G.fib(18)
    assert v == O.fib(18)
    G.lib().gtap_visualize_profile(b"fib_off")
    assert os.path.exists(tmp_path / "profile" / "fib_off_warp_counters.csv")
    assert not os.path.exists(tmp_path / "profile" / "fib_off_warp_timeline_working.csv")


@pytest.mark.parametrize("prog,arg,expect", [("nqueens", 10, 724), ("fib", 20, 6765), ("cilksort", 1 << 16, None)])
def test_worker_stats_sum_to_run_stats(prog, arg, expect):
    """The per-launch totals are summed on the device; the per-wave counters, copied on
    demand, must add up to them."""
    print(f"case worker stats {prog}({arg})", flush=True)
    if prog == "cilksort":
        keys = O.gen_keys_u32(arg)
        out, s = G.cilksort(keys)
        assert (out == np.sort(keys)).all()
    else:
        v, s = getattr(G, prog)(arg)
        assert v == expect
    ws = G.worker_stats()
    assert ws.shape == (s["workers"], 16)
    for i, f in enumerate(G.WORKER_STAT_FIELDS):
        assert int(ws[:, i].sum()) == s[f], f
    assert int(ws[:, 0].sum()) > 0


@pytest.mark.parametrize("view", ["having_task", "working"])
def test_single_view_prints_timeline_and_writes_its_files(tmp_path, monkeypatch, capfd, view):
    """visualize_having_task_time / visualize_working_time on their own: the text timeline of
    the first ten workers (80 columns, 'H'/'W' where a worker got work, 'N' where it ran out)
    and that view's two CSV files only, relative times starting at 0 on the view's first sample."""
    monkeypatch.chdir(tmp_path)
    print(f"case profile view {view}", flush=True)
    assert G.gtap_set_profiling(True, 1 << 12) == 0
    try:
        v, st = G.fib(20)
        assert v == O.fib(20)
        capfd.readouterr()
        getattr(G, f"visualize_{view}_time")("one")
    finally:
        assert G.gtap_set_profiling(False, 0) == 0
    out = capfd.readouterr().out.splitlines()
    title = "Having-Task" if view == "having_task" else "Working"
    assert f"=== Warp {title} Time Visualization ===" in out
    i = out.index("Warp  |" + "".join("|" if c % 10 == 0 else "-" for c in range(80)))
    rows = out[i + 1:i + 1 + min(10, st["workers"])]
    assert len(rows) == min(10, st["workers"])
    mark = "H" if view == "having_task" else "W"
    for w, r in enumerate(rows):
        assert r.startswith(f"{w:5d} |") and len(r) == 7 + 80
        assert set(r[7:]) <= {" ", mark, "N"}
    assert any(mark in r[7:] for r in rows)
    other = "working" if view == "having_task" else "having_task"
    assert not os.path.exists(tmp_path / "profile" / f"one_warp_timeline_{other}.csv")
    tl = _read(tmp_path, "one", f"warp_timeline_{view}")
    assert min(float(r["relative_time_ms"]) for r in tl) == 0.0
    assert len(_read(tmp_path, "one", f"warp_statistics_{view}")) == st["workers"]


@pytest.mark.parametrize("prog,arg", [("fib", 22), ("nqueens", 11)])
def test_round_phase_breakdown(prog, arg):
    """The profiling build times each round's phases with the shader clock: the phases of every
    round are counted (rounds equal the run statistics), and a round always runs its tasks,
    issues and waits for its writes and publishes; without profiling nothing is recorded."""
    print(f"case round phases {prog}({arg})", flush=True)
    assert G.gtap_set_profiling(True, 256) == 0
    try:
        v, st = getattr(G, prog)(arg)
        ph = G.round_phases()
    finally:
        assert G.gtap_set_profiling(False, 0) == 0
    assert v == (O.fib(arg) if prog == "fib" else O.nq(arg)[0])
    assert ph["rounds"] == st["rounds"] > 0
    for k in ("load", "run", "store_wait", "publish"):
        assert ph["cycles"][k] > 0, k
    per = ph["cycles_per_round"]
    assert all(x >= 0 for x in per.values())
    G.fib(15)  # a launch without profiling leaves the sums at 0
    assert G.round_phases()["rounds"] == 0
