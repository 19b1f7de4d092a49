"""EPAQ with every non-batch queue published at the end of each round (gtap_config.epaq_publish
= 1) instead of keeping its newest private entries: BFS distances stay bit-exact against the
oracle's queue BFS, with 4 degree-bucket queues and with one."""
import numpy as np
import pytest

import gtap_amd
import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def publish_all():
    gtap_amd.gtap_set_config(timeout_ms=20000, epaq_publish=1)
    assert gtap_amd.gtap_initialize() == 0
    yield
    assert gtap_amd.gtap_finalize() == 0
    gtap_amd.gtap_set_config()


@pytest.mark.parametrize("scale,seed", [(10, 3), (16, 7), (20, 1)])
@pytest.mark.parametrize("queues", [1, 4])
def test_bfs_publish_all(scale, seed, queues):
    print(f"case bfs epaq_publish=1 rmat scale={scale} queues={queues}", flush=True)
    row, col = oracle.rmat_csr(scale, 16, seed) if scale <= 16 else gtap_amd.rmat_csr(scale, 16, seed)
    depth, st = gtap_amd.bfs(row, col, 0, num_queues=queues)
    assert st["error_code"] == 0
    assert np.array_equal(depth, oracle.bfs(row, col, 0))


def test_stream_unaffected():
    """Two-queue programs (the stream's lane and wave queues) give the same results."""
    print("case stream 64 jobs epaq_publish=1", flush=True)
    s = gtap_amd.Stream(seed=3, njobs=64)
    try:
        res, _, st = s.run_batched()
        assert st["jobs_taken"] == 64
        for i in range(64):
            k, p = int(s.kind[i]), int(s.param[i])
            if k == 0:
                assert res[i] == oracle.fib(p)
            elif k == 1:
                assert res[i] == oracle.nq(p)[0]
            else:
                assert res[i] == oracle.sort_checksum(np.sort(oracle.gen_keys_u32(p, int(s.key_seed[i]))))
    finally:
        s.close()
