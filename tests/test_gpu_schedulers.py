"""The alternative schedulers the reference compares its work stealing with (global queue:
runtime/detail/experimental/thread/gtap_runtime_thread_gq.cuh:484-621; single-item
Chase-Lev: gtap_runtime_thread_chaselev.cuh:516-712) give the same results and the same
task counts as batched work stealing, at full and at small worker counts."""
import pytest

import gtap_amd as G
import oracle as O

pytestmark = pytest.mark.gpu

SCHEDULERS = {"work_stealing": 1, "global_queue": 2, "chase_lev": 3}


@pytest.fixture()
def sched_runtime(request):
    sched, blocks = request.param
    G.gtap_finalize()
    assert G.gtap_set_config(timeout_ms=20000, scheduler=SCHEDULERS[sched], blocks=blocks) == 0
    assert G.gtap_initialize() == 0
    yield sched, blocks
    assert G.gtap_finalize() == 0


CASES = [(s, b) for s in SCHEDULERS for b in (0, 4)]


@pytest.mark.parametrize("sched_runtime", CASES, indirect=True, ids=[f"{s}-blocks{b}" for s, b in CASES])
def test_fib_and_nqueens_under_each_scheduler(sched_runtime):
    sched, blocks = sched_runtime
    for n in (0, 1, 2, 10, 20, 25):
        print(f"case {sched} blocks={blocks} fib({n})", flush=True)
        v, s = G.fib(n)
        assert v == O.fib(n)
        assert s["tasks_executed"] == (3 * O.fib(n + 1) - 2 if n >= 1 else 1)
    for n in (1, 6, 8, 10):
        print(f"case {sched} blocks={blocks} nqueens({n})", flush=True)
        v, s = G.nqueens(n)
        sols, tasks = O.nq(n)
        assert v == sols
        assert s["tasks_spawned"] + 1 == tasks
        assert s["tasks_executed"] == tasks + O.nq_internal(n)
    if blocks == 0:
        print(f"case {sched} nqueens(12)", flush=True)
        v, s = G.nqueens(12)
        assert v == O.nq(12)[0]
