"""The deque overflow ring at 4096-slot deques. A lowered program in which one task spawns n
detached children in one round (gtap_amd/lower/examples/burst.gtap.cu) publishes more than
half of its wave's deque at the round's end, so the wave hands its oldest tasks to the
overflow ring and idle waves take them back: every child runs once (the counters it adds to
are exact) and every task moved to the ring is taken again, launch after launch. A star
graph's BFS with 4 wavefront workers, where the deques fill and drain quickly, stays
bit-exact."""
import numpy as np
import pytest

import gtap_amd
import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def few_workers_small_deques():
    gtap_amd.gtap_set_config(timeout_ms=30000, deque_capacity=4096, blocks=1)
    assert gtap_amd.gtap_initialize() == 0
    yield
    assert gtap_amd.gtap_finalize() == 0
    gtap_amd.gtap_set_config()


@pytest.fixture(scope="module")
def burst_module():
    import ctypes
    from gtap_amd import build as gbuild
    return gtap_amd.EntryModule(gbuild.lowered_lib("burst"), "burst"), ctypes


def ring_counts():
    ws = gtap_amd.worker_stats()
    return int(ws[:, gtap_amd.WORKER_RING_SPILLED].sum()), int(ws[:, gtap_amd.WORKER_RING_TAKEN].sum())


def star(leaves):
    """Vertex 0 with an edge to each of `leaves` vertices that have no out-edges."""
    n = leaves + 1
    row = np.full(n + 1, leaves, dtype=np.int32)
    row[0] = 0
    col = np.arange(1, n, dtype=np.int32)
    return row, col


@pytest.mark.parametrize("queues", [1, 4])
@pytest.mark.parametrize("leaves", [1 << 18, 1 << 20])
def test_star_overflows_into_the_ring(queues, leaves):
    print(f"case bfs star leaves={leaves} queues={queues} split 8, deque capacity 4096, 4 workers", flush=True)
    row, col = star(leaves)
    depth, st = gtap_amd.bfs(row, col, 0, num_queues=queues, split=8)
    assert st["error_code"] == 0
    spilled, taken = ring_counts()
    print(f"ring: {spilled} tasks spilled, {taken} taken", flush=True)
    assert taken == spilled
    expect = np.ones(leaves + 1, dtype=np.int32)
    expect[0] = 0
    assert np.array_equal(depth, expect)
    assert np.array_equal(depth, oracle.bfs(row, col, 0))


def test_joined_program_after_ring_use():
    """A fork-join launch on the same rings after the star runs (positions carried over)."""
    print("case fib(25) after ring use", flush=True)
    v, st = gtap_amd.fib(25)
    assert v == oracle.fib(25) and st["error_code"] == 0


@pytest.mark.parametrize("n", [3050, 2000, 1])
def test_burst_spills_into_the_ring(burst_module, n):
    m, ctypes = burst_module
    u64 = ctypes.c_uint64
    total_spilled = 0
    for rep in range(3):
        print(f"case burst n={n} rep={rep} deque capacity 4096", flush=True)
        m.set_symbol("d_count", 0, u64)
        m.set_symbol("d_sum", 0, u64)
        assert m.launch("burst_kernel", [ctypes.c_int], n) == 0
        assert gtap_amd.gtap_get_runtime_error_code() == 0
        assert m.get_symbol("d_count", u64) == n
        assert m.get_symbol("d_sum", u64) == n * (n - 1) // 2
        spilled, taken = ring_counts()
        print(f"ring: {spilled} tasks spilled, {taken} taken", flush=True)
        assert taken == spilled
        total_spilled += spilled
    if n == 3050:
        # 2986 published at the round's end, past the 2048 half: thieves would have to take
        # 939 of them before the wave's own claim for a launch not to spill
        assert total_spilled > 0
