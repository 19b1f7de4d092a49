"""Algorithmic work per unit, per workload -- the constants every roofline figure uses.

Bytes are the minimum HBM traffic the algorithm needs (not what the hardware moved):

* fib, per spawned non-root task: 48 B = child frame written + read (args 4 + parent
  link 4 -> 16), int64 result written + read (16), task id enqueue + dequeue (8),
  join-counter read-modify-write (8).
* n-queens, per spawned non-root task: 72 B = frame (16 args + 4 link) written + read (40),
  result 16, id 8, join counter 8.
* cilksort of N uint32 keys: 15 passes (one leaf pass + 7 four-way levels x 2 merge
  passes) x (4N read + 4N write).
* BFS: 8 V_r (row-offset pair) + 4 V_r (depth read) + 4 E_r (column) + 8 E_r
  (atomicMin read-modify-write), V_r / E_r = reached vertices / their out-edges.

Atomics: one agent-scope read-modify-write (the join decrement) per spawned non-root task.
"""

FIB_BYTES_PER_TASK = 48
NQ_BYTES_PER_TASK = 72
CILKSORT_PASSES = 15
CILKSORT_BYTES_PER_KEY_PER_PASS = 8
BFS_BYTES_PER_VERTEX = 12
BFS_BYTES_PER_EDGE = 12
ATOMICS_PER_TASK = 1

# Published MI355X peaks (dense, no sparsity).
MI355X_HBM_GBPS = 8000.0


def fib_bytes(spawned_tasks: int) -> int:
    return FIB_BYTES_PER_TASK * spawned_tasks


def nq_bytes(spawned_tasks: int) -> int:
    return NQ_BYTES_PER_TASK * spawned_tasks


def cilksort_bytes(n_keys: int) -> int:
    return CILKSORT_PASSES * CILKSORT_BYTES_PER_KEY_PER_PASS * n_keys


def bfs_bytes(reached_vertices: int, reached_edges: int) -> int:
    return BFS_BYTES_PER_VERTEX * reached_vertices + BFS_BYTES_PER_EDGE * reached_edges
