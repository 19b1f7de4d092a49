"""SpMV (examples/spmv/spmv_thread.cu:34-66), the binary-tree workload
(examples/binary_tree/binary_tree_thread.cu:63-120) and the block-mode BFS
(examples/asynchronous_bfs/asynchronous_bfs_block.cu:17-37) on the GPU, bit-exact against
the oracle; a graph read from a file in the reference's CSR format through BFS."""
import numpy as np
import pytest

import gtap_amd as G
import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def runtime():
    G.gtap_set_config(timeout_ms=20000)
    assert G.gtap_initialize() == 0
    yield
    assert G.gtap_finalize() == 0


@pytest.mark.parametrize("rows,avg,cutoff", [(1, 3, 16), (17, 5, 16), (1000, 8, 16), (50000, 24, 16),
                                             (50000, 24, 1), (50000, 24, 200), (200000, 300, 16)])
def test_spmv_matches_oracle(rows, avg, cutoff):
    print(f"case spmv rows={rows} avg={avg} cutoff={cutoff}", flush=True)
    rp, ci, va = G.spmv_random(rows, avg, rows + avg)
    x = np.random.default_rng(rows).uniform(-2, 2, rows)
    y, s = G.spmv(rp, ci, va, x, cutoff)
    assert np.array_equal(y, O.spmv(rp, ci, va, x))
    # one task per row plus the range tasks (a binary split of [0, rows) down to <= cutoff rows)
    assert s["tasks_executed"] >= rows


def test_spmv_range_cutoff_beyond_deque_capacity_is_reported():
    # 64 lanes each spawning up to 1000 row tasks in one round need more deque slots than
    # gtap_config.deque_capacity provides beyond the published tasks: a round's children stay
    # private until the round ends, so the launch ends with a queue-overflow error (the
    # reference's GTAP_MAX_TASKS_PER_WARP trap), and the next launch runs normally
    print("case spmv cutoff 1000 (queue overflow reported)", flush=True)
    rp, ci, va = G.spmv_random(50000, 4, 1)
    x = np.ones(50000)
    with pytest.raises(G.GtapError) as e:
        G.spmv(rp, ci, va, x, 1000)
    assert "QUEUE_OVERFLOW" in str(e.value)
    y, _ = G.spmv(rp, ci, va, x, 16)
    assert np.array_equal(y, O.spmv(rp, ci, va, x))


def test_spmv_empty_rows_and_zero_matrix():
    print("case spmv empty rows", flush=True)
    rp = np.array([0, 0, 2, 2, 3, 3], dtype=np.int32)
    ci = np.array([0, 4, 2], dtype=np.int32)
    va = np.array([1.5, -2.0, 3.0])
    x = np.array([2.0, 9.0, 1.0, 9.0, 0.25])
    y, _ = G.spmv(rp, ci, va, x)
    assert y.tolist() == [0.0, 1.5 * 2.0 - 2.0 * 0.25, 0.0, 3.0, 0.0]
    y0, _ = G.spmv(np.zeros(8, dtype=np.int32), np.zeros(0, dtype=np.int32), np.zeros(0), np.ones(7))
    assert np.all(y0 == 0.0)


@pytest.mark.parametrize("height,mem_ops,iters", [(0, 4, 4), (1, 0, 0), (5, 64, 512), (10, 16, 32), (15, 64, 512),
                                                  (12, 0, 1000)])
def test_tree_matches_oracle(height, mem_ops, iters):
    print(f"case tree height={height} mem_ops={mem_ops} iters={iters}", flush=True)
    inp = G.tree_input(1 << 16, height + 1)
    out, s = G.tree(height, mem_ops, iters, inp)
    nodes = (1 << (height + 1)) - 1
    assert out.size == nodes
    assert np.array_equal(out, O.tree(height, mem_ops, iters, inp))
    assert s["tasks_spawned"] == nodes - 1


def test_bfs_on_graph_read_from_reference_format(tmp_path):
    print("case bfs from a CSR file", flush=True)
    row, col = G.rmat_csr(14, 8, 5)
    p = str(tmp_path / "graph.bin")
    G.write_csr(p, row, col)
    r2, c2 = G.read_csr(p)
    depth, _ = G.bfs(r2, c2, 0, 4, 64)
    assert np.array_equal(depth, O.bfs(row, col, 0))


@pytest.mark.parametrize("scale,ef", [(8, 4), (14, 8), (18, 16)])
def test_block_bfs_matches_oracle(scale, ef):
    print(f"case block bfs rmat scale={scale} ef={ef}", flush=True)
    row, col = G.rmat_csr(scale, ef, scale)
    depth, s = G.bfs_block(row, col, 0)
    assert np.array_equal(depth, O.bfs(row, col, 0))


def test_block_bfs_chain_and_star():
    print("case block bfs chain and star", flush=True)
    n = 3000
    row = np.concatenate([np.arange(n), [n - 1]]).astype(np.int32)  # chain 0 -> 1 -> ... (the reference's default graph)
    col = np.arange(1, n, dtype=np.int32)
    depth, _ = G.bfs_block(row, col, 0)
    assert depth.tolist() == list(range(n))
    star_row = np.array([0, 5000] + [5000] * 5000, dtype=np.int32)  # vertex 0 -> 1..5000
    star_col = np.arange(1, 5001, dtype=np.int32)
    depth, _ = G.bfs_block(star_row, star_col, 0)
    assert depth[0] == 0 and np.all(depth[1:] == 1)
