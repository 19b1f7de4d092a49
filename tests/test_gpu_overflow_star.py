"""The deque overflow ring with few workers: a star graph's source range splits into 8-edge
tasks that each spawn a task per leaf, so with 4 wavefront workers each round of a wave adds
up to 64 x 8 detached tasks to its deque while it runs only 64, and the 3 others take little.
A deque above half of its 4096 slots hands its oldest published tasks to the overflow ring
after the round (a round adds at most 576, which the other half holds), and idle waves take
them back; distances stay bit-exact and every task moved to a ring is taken again."""
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
    assert spilled > 0 and taken == spilled
    expect = np.ones(leaves + 1, dtype=np.int32)
    expect[0] = 0
    assert np.array_equal(depth, expect)
    assert np.array_equal(depth, oracle.bfs(row, col, 0))


def test_joined_program_after_ring_use():
    """A fork-join launch on the same rings after the star runs (positions carried over)."""
    print("case fib(25) after ring use", flush=True)
    v, st = gtap_amd.fib(25)
    assert v == oracle.fib(25) and st["error_code"] == 0
