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
    # star with a hub of degree 100000 (range tasks cut down to 64 edges) and self loops
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


def hub_graph(deg, extra=3):
    """Vertex 0 with `deg` out-edges to vertices 1..deg; vertex i also points to i + 1 and back
    to 0, and `extra` more hubs of the same degree hang off vertex 1, so several wave tasks and
    their pieces are in flight at once."""
    n = deg + 1
    src, dst = [np.zeros(deg, np.int64)], [np.arange(1, n, dtype=np.int64)]
    src.append(np.arange(1, n - 1, dtype=np.int64))
    dst.append(np.arange(2, n, dtype=np.int64))
    src.append(np.arange(1, n, dtype=np.int64))
    dst.append(np.zeros(deg, np.int64))
    for h in range(extra):
        hub = n + h * (deg + 1)
        src.append(np.array([1], np.int64))
        dst.append(np.array([hub], np.int64))
        src.append(np.full(deg, hub, np.int64))
        dst.append(np.arange(hub + 1, hub + 1 + deg, dtype=np.int64))
    nv = n + extra * (deg + 1)
    s, d = np.concatenate(src), np.concatenate(dst)
    order = np.argsort(s, kind="stable")
    s, d = s[order], d[order]
    row = np.zeros(nv + 1, np.int64)
    np.add.at(row, s + 1, 1)
    return np.cumsum(row).astype(np.int32), d.astype(np.int32)


@pytest.mark.parametrize("deg", [511, 512, 513, 2048, 2049, 131072, 131137])
@pytest.mark.parametrize("queues", [1, 4])
def test_bfs_hub_degrees_around_the_wave_splits(deg, queues):
    """With 4 queues, ranges of 512 edges and more run as wave tasks, which cut ranges above
    2048 edges into up to 64 pieces (131137 edges: 64 pieces of more than 2048, split again);
    lane tasks cut ranges into up to 8 pieces."""
    print(f"case bfs hubs deg={deg} queues={queues}", flush=True)
    row, col = hub_graph(deg)
    for source in (0, 1):
        depth, st = gtap_amd.bfs(row, col, source, num_queues=queues)
        assert st["error_code"] == 0
        assert np.array_equal(depth, oracle.bfs(row, col, source))


@pytest.mark.parametrize("queues", [1, 4])
def test_bfs_bad_column_in_a_hub_is_reported(queues):
    """A column index past the vertex count inside a hub's list (a wave task with 4 queues)
    ends the launch with GTAP_ERROR_INVALID_ARGUMENT; the next launch runs normally."""
    print(f"case bfs bad column queues={queues}", flush=True)
    row, col = hub_graph(1000, extra=0)
    bad = col.copy()
    bad[700] = row.size + 5
    with pytest.raises(gtap_amd.GtapError) as e:
        gtap_amd.bfs(row, bad, 0, num_queues=queues)
    assert "INVALID_ARGUMENT" in str(e.value)
    depth, st = gtap_amd.bfs(row, col, 0, num_queues=queues)
    assert st["error_code"] == 0 and np.array_equal(depth, oracle.bfs(row, col, 0))


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
