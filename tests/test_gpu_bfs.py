"""Asynchronous BFS (EPAQ by out-degree bucket, and one queue) through the C ABI on the
GPU: distances bit-exact against the oracle's queue BFS, plus the BFS invariants checked
directly at the benchmark scale."""
import numpy as np
import pytest

import gtap_amd
import oracle

pytestmark = pytest.mark.gpu
INF = oracle.INF_DEPTH


@pytest.fixture(scope="module", autouse=True)
def runtime():
    gtap_amd.gtap_set_config(timeout_ms=20000)
    assert gtap_amd.gtap_initialize() == 0
    yield
    assert gtap_amd.gtap_finalize() == 0


def check_invariants(row, col, depth, source):
    """depth[source] = 0; every edge v->u of a reached v has depth[u] <= depth[v] + 1; every
    reached u != source has an in-edge from depth[u] - 1; unreached vertices stay INF."""
    nv = row.size - 1
    src = np.repeat(np.arange(nv, dtype=np.int64), np.diff(row))
    dv = depth[src].astype(np.int64)
    du = depth[col].astype(np.int64)
    reached = dv < INF
    assert depth[source] == 0
    assert np.all((depth >= 0) & ((depth < INF) | (depth == INF)))
    assert np.all(du[reached] <= dv[reached] + 1)
    tight = reached & (du == dv + 1)
    has_parent = np.zeros(nv, dtype=bool)
    has_parent[col[tight]] = True
    r = (depth < INF)
    r[source] = False
    assert np.all(has_parent[r])


@pytest.mark.parametrize("scale,seed,source", [(8, 1, 0), (10, 3, 0), (12, 1, 5), (14, 1, 0), (16, 7, 0)])
@pytest.mark.parametrize("queues", [1, 4])
def test_bfs_rmat_parity(scale, seed, source, queues):
    print(f"case bfs rmat scale={scale} seed={seed} source={source} queues={queues}", flush=True)
    row, col = oracle.rmat_csr(scale, 16, seed)
    depth, st = gtap_amd.bfs(row, col, source, num_queues=queues)
    assert np.array_equal(depth, oracle.bfs(row, col, source))
    assert st["error_code"] == 0


def test_bfs_chain_star_isolated():
    n = 3000
    row = np.array(list(range(n)) + [n - 1], dtype=np.int32)
    col = np.arange(1, n, dtype=np.int32)
    for q in (1, 4):
        print(f"case bfs chain queues={q}", flush=True)
        d, _ = gtap_amd.bfs(row, col, 0, num_queues=q)
        assert np.array_equal(d, np.arange(n))
        d, _ = gtap_amd.bfs(row, col, 2500, num_queues=q)
        assert np.array_equal(d, oracle.bfs(row, col, 2500))
    # star with a hub of degree 100000 (range tasks split down to 64 edges) and self loops
    n = 100001
    col = np.concatenate([np.arange(1, n, dtype=np.int32), np.array([0], dtype=np.int32)])
    row = np.full(n + 1, col.size, dtype=np.int32)
    row[0] = 0
    for q in (1, 4):
        print(f"case bfs star queues={q}", flush=True)
        d, _ = gtap_amd.bfs(row, col, 0, num_queues=q)
        assert d[0] == 0 and np.all(d[1:] == 1)
    # a vertex with no out-edges as the source
    d, _ = gtap_amd.bfs(row, col, 5, num_queues=4)
    assert d[5] == 0 and np.all(np.delete(d, 5) == INF)


def test_bfs_split_thresholds_agree():
    row, col = oracle.rmat_csr(13, 16, 11)
    ref = oracle.bfs(row, col, 0)
    for split in (1, 8, 64, 4096):
        print(f"case bfs split={split}", flush=True)
        d, _ = gtap_amd.bfs(row, col, 0, num_queues=4, split=split)
        assert np.array_equal(d, ref)


def test_bfs_benchmark_scale_invariants_and_parity():
    print("case bfs rmat scale=22 (benchmark graph)", flush=True)
    row, col = gtap_amd.rmat_csr(22, 16, 1)
    depth, st = gtap_amd.bfs(row, col, 0, num_queues=4)
    check_invariants(row, col, depth, 0)
    assert np.array_equal(depth, oracle.bfs(row, col, 0))
