"""XCD-local stealing (gtap_config.xcd_steal = 1: half of the steal probes go to victims on the
thief's own XCD) changes only where tasks run: results stay bit-exact against the oracle."""
import numpy as np
import pytest

import gtap_amd
import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def xcd_local():
    gtap_amd.gtap_set_config(timeout_ms=20000, xcd_steal=1)
    assert gtap_amd.gtap_initialize() == 0
    assert gtap_amd.gtap_get_config()["xcd_steal"] == 1
    yield
    assert gtap_amd.gtap_finalize() == 0
    gtap_amd.gtap_set_config()


@pytest.mark.parametrize("n", [2, 20, 27])
def test_fib(n):
    print(f"case fib({n}) xcd_steal=1", flush=True)
    v, st = gtap_amd.fib(n)
    assert v == oracle.fib(n) and st["error_code"] == 0
    assert st["tasks_executed"] == 3 * oracle.fib(n + 1) - 2


@pytest.mark.parametrize("n", [8, 12])
def test_nqueens(n):
    print(f"case nqueens({n}) xcd_steal=1", flush=True)
    v, st = gtap_amd.nqueens(n)
    assert v == oracle.nq(n)[0] and st["error_code"] == 0


@pytest.mark.parametrize("queues", [1, 4])
def test_bfs(queues):
    print(f"case bfs rmat 16 queues={queues} xcd_steal=1", flush=True)
    row, col = oracle.rmat_csr(16, 16, 7)
    depth, st = gtap_amd.bfs(row, col, 0, num_queues=queues)
    assert st["error_code"] == 0
    assert np.array_equal(depth, oracle.bfs(row, col, 0))
