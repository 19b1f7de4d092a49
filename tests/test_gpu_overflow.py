"""Deque overflow: with a small deque capacity, a deque holding more than half its capacity at
the end of a round hands its oldest published tasks to the per-queue overflow ring instead of
failing with GTAP_ERROR_QUEUE_OVERFLOW (the reference traps there:
runtime/detail/thread/gtap_runtime_thread.cuh:42-45). Results stay bit-exact against the
oracle; every task moved to a ring is taken from it again."""
import numpy as np
import pytest

import gtap_amd
import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def small_deques():
    gtap_amd.gtap_set_config(timeout_ms=30000, deque_capacity=4096)
    assert gtap_amd.gtap_initialize() == 0
    yield
    assert gtap_amd.gtap_finalize() == 0
    gtap_amd.gtap_set_config()


def ring_counts():
    ws = gtap_amd.worker_stats()
    return int(ws[:, gtap_amd.WORKER_RING_SPILLED].sum()), int(ws[:, gtap_amd.WORKER_RING_TAKEN].sum())


def test_bfs_rmat22_four_queues_capacity_4096():
    print("case bfs rmat scale=22, 4 queues, deque capacity 4096", flush=True)
    row, col = gtap_amd.rmat_csr(22, 16, 1)
    depth, st = gtap_amd.bfs(row, col, 0, num_queues=4)
    assert st["error_code"] == 0
    spilled, taken = ring_counts()
    print(f"ring: {spilled} tasks spilled, {taken} taken", flush=True)
    assert taken == spilled  # all-tasks termination: the rings end empty
    assert np.array_equal(depth, oracle.bfs(row, col, 0))


@pytest.mark.parametrize("queues", [1, 4])
def test_bfs_rmat18_repeated(queues):
    """Launches in a row on the same rings (positions carry over from launch to launch)."""
    row, col = gtap_amd.rmat_csr(18, 16, 5)
    ref = oracle.bfs(row, col, 0)
    for rep in range(3):
        print(f"case bfs rmat scale=18 queues={queues} rep={rep}", flush=True)
        depth, st = gtap_amd.bfs(row, col, 0, num_queues=queues)
        assert st["error_code"] == 0
        spilled, taken = ring_counts()
        assert taken == spilled
        assert np.array_equal(depth, ref)


@pytest.mark.parametrize("prog,arg", [("fib", 27), ("nqueens", 12)])
def test_joined_programs_with_small_deques(prog, arg):
    """Fork-join programs (joined children, first-task termination) at capacity 4096."""
    print(f"case {prog}({arg}) deque capacity 4096", flush=True)
    v, st = getattr(gtap_amd, prog)(arg)
    assert v == (oracle.fib(arg) if prog == "fib" else oracle.nq(arg)[0])
    assert st["error_code"] == 0


def test_nqueens_cutoff_detached_small_deques():
    print("case nqueens_cutoff(14, 7) deque capacity 4096", flush=True)
    v, st = gtap_amd.nqueens_cutoff(14, 7)
    assert v == 365596 and st["error_code"] == 0
    spilled, taken = ring_counts()
    assert taken == spilled
