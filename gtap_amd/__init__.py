"""MI355X-native GTaP runtime: Python host-side mirror of the GTaP host API.

The product is ``_lib/libgtap_amd.so`` (HIP for gfx950, built by ``gtap_amd.build``);
this module only binds its C ABI (``include/gtap.h``) with ctypes. Names follow the
reference host API -- ``gtap_initialize``, ``gtap_finalize``, ``gtap_reset``,
``gtap_synchronize``, ``gtap_get_runtime_error_code``, ``gtap_visualize_profile``
(runtime/detail/thread/gtap_runtime_thread.cuh:338-441,
runtime/detail/common/gtap_runtime_error.cuh:38-249) -- and return the same int status.
The task-program wrappers (``fib`` ...) raise :class:`GtapError` on a nonzero status.

There is no CPU fallback: if the library is missing, loading raises.
"""
from __future__ import annotations

import ctypes
import os
from typing import Optional

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "libgtap_amd.so")

# gtap_error_code (gtap_amd/csrc/gtap_types.h)
ERROR_NAMES = {
    0: "GTAP_ERROR_NONE",
    1: "GTAP_ERROR_INVALID_QUEUE_IDX",
    2: "GTAP_ERROR_INVALID_QUEUE_IDX_AFTER_JOIN",
    3: "GTAP_ERROR_QUEUE_OVERFLOW",
    4: "GTAP_ERROR_RESULT_HANDLE_OVERFLOW",
    5: "GTAP_ERROR_TASK_ID_POOL_SLOT_BUSY",
    6: "GTAP_ERROR_TASK_ID_POOL_LOW_HEADROOM",
    7: "GTAP_ERROR_GENERATED_TASK_ID_BUFFER_OVERFLOW",
    8: "GTAP_ERROR_INVALID_TASKWAIT",
    9: "GTAP_ERROR_WATCHDOG",
    10: "GTAP_ERROR_INVALID_FUNCTION",
    11: "GTAP_ERROR_INVALID_ARGUMENT",
    12: "GTAP_ERROR_NOT_INITIALIZED",
    13: "GTAP_ERROR_HIP",
}

TERM_JOBS, TERM_QUIESCE, TERM_SERVE = 0, 1, 2


class ErrorReport(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in
                ("valid", "code", "src_line", "block_idx", "thread_idx", "tid", "queue_idx", "value", "limit")]


class RunStats(ctypes.Structure):
    _fields_ = [
        ("kernel_ms", ctypes.c_double),
        ("tasks_executed", ctypes.c_uint64),
        ("tasks_spawned", ctypes.c_uint64),
        ("rounds", ctypes.c_uint64),
        ("pops", ctypes.c_uint64),
        ("steals", ctypes.c_uint64),
        ("steal_attempts", ctypes.c_uint64),
        ("jobs_taken", ctypes.c_uint64),
        ("pool_refills", ctypes.c_uint64),
        ("published", ctypes.c_uint64),
        ("overflowed", ctypes.c_uint64),
        ("suspended", ctypes.c_uint64),
        ("workers", ctypes.c_uint32),
        ("blocks", ctypes.c_uint32),
        ("error_code", ctypes.c_int32),
        ("admission_holds", ctypes.c_int32),
    ]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


class Config(ctypes.Structure):
    _fields_ = [
        ("pool_bytes", ctypes.c_uint64),
        ("record_stride", ctypes.c_uint32),
        ("deque_capacity", ctypes.c_uint32),
        ("num_queues", ctypes.c_uint32),
        ("blocks", ctypes.c_uint32),
        ("steal_batch", ctypes.c_uint32),
        ("probe_lanes", ctypes.c_uint32),
        ("max_jobs", ctypes.c_uint32),
        ("timeout_ms", ctypes.c_uint32),
        ("block_workers", ctypes.c_uint32),
        ("idle_sleep", ctypes.c_uint32),
        ("scheduler", ctypes.c_uint32),
        ("epaq_publish", ctypes.c_uint32),
        ("xcd_steal", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32 * 2),
    ]


class Job(ctypes.Structure):
    _fields_ = [
        ("func", ctypes.c_uint32),
        ("queue", ctypes.c_uint32),
        ("status", ctypes.c_uint32),
        ("kind", ctypes.c_uint32),
        ("result", ctypes.c_int64),
        ("t_submit", ctypes.c_uint64),
        ("t_start", ctypes.c_uint64),
        ("t_end", ctypes.c_uint64),
        ("t_done_host", ctypes.c_uint64),
        ("reserve_ids", ctypes.c_uint64),
        ("args", ctypes.c_uint8 * 64),
    ]


class GtapError(RuntimeError):
    def __init__(self, status: int, where: str):
        code = ctypes.c_int(0)
        rep = ErrorReport()
        try:
            lib().gtap_get_runtime_error_code(ctypes.byref(code))
            lib().gtap_get_runtime_error_report(ctypes.byref(rep))
        except Exception:  # pragma: no cover - library unusable
            pass
        self.status = status
        self.code = code.value
        self.report = {f: getattr(rep, f) for f, _ in rep._fields_}
        super().__init__(f"{where}: status {status}, GTaP error {self.code} "
                         f"({ERROR_NAMES.get(self.code, '?')}), report {self.report}")


# Every symbol include/gtap.h declares, with its ctypes signature.
_i, _u32, _u64, _i64, _vp = ctypes.c_int, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int64, ctypes.c_void_p
SIGNATURES: dict[str, tuple[list, object]] = {
    "gtap_initialize": ([], _i),
    "gtap_register_entry_module": ([_vp], _i),
    "gtap_finalize": ([], _i),
    "gtap_reset": ([], _i),
    "gtap_set_config": ([ctypes.POINTER(Config)], _i),
    "gtap_get_config": ([ctypes.POINTER(Config)], _i),
    "gtap_synchronize": ([], _i),
    "gtap_check_runtime_error": ([], _i),
    "gtap_get_runtime_error_code": ([ctypes.POINTER(_i)], _i),
    "gtap_print_runtime_error_report": ([], _i),
    "gtap_get_runtime_error_string": ([_i], ctypes.c_char_p),
    "gtap_get_runtime_error_report": ([ctypes.POINTER(ErrorReport)], _i),
    "gtap_visualize_profile": ([ctypes.c_char_p], None),
    "gtap_visualize_having_task_time": ([ctypes.c_char_p], None),
    "gtap_get_round_phases": ([_vp, _u32], _i),
    "gtap_visualize_working_time": ([ctypes.c_char_p], None),
    "gtap_set_profiling": ([_i, _u32], _i),
    "gtap_bench_fences": ([_u32, _i, ctypes.POINTER(ctypes.c_double)], _i),
    "gtap_get_last_stats": ([ctypes.POINTER(RunStats)], _i),
    "gtap_get_worker_stats": ([_vp, _u32, ctypes.POINTER(_u32)], _i),
    "gtap_device_info": ([ctypes.c_char_p, _i, ctypes.POINTER(_i), ctypes.POINTER(_u64)], _i),
    "gtap_device_malloc": ([ctypes.POINTER(_vp), _u64], _i),
    "gtap_device_free": ([_vp], _i),
    "gtap_memcpy_htod": ([_vp, _vp, _u64], _i),
    "gtap_memcpy_dtoh": ([_vp, _vp, _u64], _i),
    "gtap_memcpy_dtod": ([_vp, _vp, _u64], _i),
    "gtap_memset_device": ([_vp, _i, _u64], _i),
    "gtap_bench_hbm_copy": ([_u64, _i, ctypes.POINTER(ctypes.c_double)], _i),
    "gtap_bench_atomics": ([_i, ctypes.POINTER(ctypes.c_double)], _i),
    "gtap_bench_atomics_contended": ([_i, _u32, ctypes.POINTER(ctypes.c_double)], _i),
    "gtap_bench_random_access": ([_u64, _i, ctypes.POINTER(ctypes.c_double)], _i),
    "gtap_fib": ([_i, ctypes.POINTER(_i64), ctypes.POINTER(RunStats)], _i),
    "gtap_nqueens": ([_i, ctypes.POINTER(_i64), ctypes.POINTER(RunStats)], _i),
    "gtap_nqueens_slots": ([_i, ctypes.POINTER(_i64), ctypes.POINTER(RunStats)], _i),
    "gtap_bfs": ([_vp, _vp, ctypes.c_int32, _i64, ctypes.c_int32, _vp, ctypes.c_int32, ctypes.c_int32,
                  ctypes.POINTER(RunStats)], _i),
    "gtap_bfs_host": ([_vp, _vp, ctypes.c_int32, _i64, ctypes.c_int32, _vp, ctypes.c_int32, ctypes.c_int32,
                       ctypes.POINTER(RunStats)], _i),
    "gtap_cilksort": ([_vp, _vp, _i64, _i64, _i64, ctypes.POINTER(RunStats)], _i),
    "gtap_cilksort_host": ([_vp, _i64, _i64, _i64, ctypes.POINTER(RunStats)], _i),
    "gtap_cilksort_noalloc": ([_vp, _i64, _i64, _i64, ctypes.POINTER(RunStats)], _i),
    "gtap_jobs_generate": ([_u64, ctypes.c_int32, _vp, _vp, _vp], _i),
    "gtap_stream_create": ([_u64, ctypes.c_int32, ctypes.POINTER(_vp)], _i),
    "gtap_stream_destroy": ([_vp], _i),
    "gtap_stream_info": ([_vp, _vp, _vp, _vp, ctypes.POINTER(_i64)], _i),
    "gtap_stream_sorted_keys": ([_vp, ctypes.c_int32, _vp], _i),
    "gtap_stream_run_batched": ([_vp, _vp, _vp, ctypes.POINTER(RunStats)], _i),
    "gtap_stream_run_per_job": ([_vp, _vp, _vp, ctypes.POINTER(ctypes.c_double)], _i),
    "gtap_stream_exec_ms": ([_vp, _vp], _i),
    "gtap_stream_set_merge_leaf": ([_vp, _i], _i),
    "gtap_stream_serve": ([_vp, ctypes.c_double, _u64, ctypes.c_int32, ctypes.c_int32, _vp, _vp,
                           ctypes.POINTER(ctypes.c_double), ctypes.POINTER(RunStats)], _i),
    "gtap_rmat_edges": ([_i, _i64, _u64, _u32, _u32, _u32, _vp, _vp], _i),
    "gtap_csr_from_edges": ([ctypes.c_int32, _i64, _vp, _vp, _vp, _vp], _i),
    "gtap_rmat_csr": ([_i, _i, _u64, _u32, _u32, _u32, _vp, _vp], _i),
    "gtap_gen_keys": ([_u64, _u64, _vp], _i),
    "gtap_nqueens_cutoff": ([_i, _i, ctypes.POINTER(_i64), ctypes.POINTER(RunStats)], _i),
    "gtap_bfs_block": ([_vp, _vp, ctypes.c_int32, _i64, ctypes.c_int32, _vp, ctypes.POINTER(RunStats)], _i),
    "gtap_spmv": ([_vp, _vp, _vp, _vp, _vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(RunStats)], _i),
    "gtap_spmv_host": ([_vp, _vp, _vp, _vp, _vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                        ctypes.POINTER(RunStats)], _i),
    "gtap_spmv_random": ([ctypes.c_int32, ctypes.c_int32, _u64, _vp, _vp, _vp, ctypes.POINTER(_i64)], _i),
    "gtap_tree": ([ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _vp, _i64, _vp, ctypes.POINTER(RunStats)], _i),
    "gtap_tree_input": ([_u64, _i64, _vp], _i),
    "gtap_csr_file_info": ([ctypes.c_char_p, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(_i64)], _i),
    "gtap_csr_file_read": ([ctypes.c_char_p, _vp, _vp], _i),
    "gtap_csr_file_write": ([ctypes.c_char_p, ctypes.c_int32, _i64, _vp, _vp], _i),
    "gtap_keys_file_info": ([ctypes.c_char_p, ctypes.POINTER(_u64)], _i),
    "gtap_keys_file_read": ([ctypes.c_char_p, _vp], _i),
    "gtap_keys_file_write": ([ctypes.c_char_p, _u64, _vp], _i),
    "gtap_spmv_file_info": ([ctypes.c_char_p, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32),
                             ctypes.POINTER(ctypes.c_int32)], _i),
    "gtap_spmv_file_read": ([ctypes.c_char_p, _vp, _vp, _vp], _i),
    "gtap_spmv_file_write": ([ctypes.c_char_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _vp, _vp, _vp], _i),
}

_lib: Optional[ctypes.CDLL] = None


def lib() -> ctypes.CDLL:
    """Load libgtap_amd.so (raises if it has not been built -- there is no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m gtap_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (argtypes, restype) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = argtypes
            fn.restype = restype
        _lib = L
    return _lib


def _check(status: int, where: str) -> None:
    if status != 0:
        raise GtapError(status, where)


# ---------------------------------------------------------------- host API (reference names)

def gtap_set_config(**kw) -> int:
    c = Config()
    for k, v in kw.items():
        setattr(c, k, v)
    return lib().gtap_set_config(ctypes.byref(c))


WORKER_STAT_FIELDS = ("tasks_executed", "tasks_spawned", "rounds", "pops", "steals", "steal_attempts",
                      "jobs_taken", "pool_refills", "published", "overflowed", "suspended")
WORKER_START_TICKS = 11  # column of worker_stats(): 100 MHz wall-clock tick at which the worker started
WORKER_END_TICKS = 12    # and the tick at which it stopped
WORKER_RING_SPILLED = 13  # tasks the worker moved from a full deque to the overflow ring
WORKER_RING_TAKEN = 14    # tasks the worker took from the overflow ring


def worker_stats():
    """Per-wave counters of the last launch: a (workers, 16) uint64 array whose first columns
    follow WORKER_STAT_FIELDS."""
    import numpy as np
    nw = ctypes.c_uint32(0)
    _check(lib().gtap_get_worker_stats(None, 0, ctypes.byref(nw)), "gtap_get_worker_stats")
    out = np.zeros((nw.value, 16), dtype=np.uint64)
    if nw.value:
        _check(lib().gtap_get_worker_stats(out.ctypes.data, nw.value, ctypes.byref(nw)), "gtap_get_worker_stats")
    return out


def gtap_get_config() -> dict:
    c = Config()
    _check(lib().gtap_get_config(ctypes.byref(c)), "gtap_get_config")
    return {f: getattr(c, f) for f, _ in c._fields_ if f != "reserved"}


def gtap_initialize() -> int:
    return lib().gtap_initialize()


def gtap_finalize() -> int:
    return lib().gtap_finalize()


def gtap_reset() -> int:
    return lib().gtap_reset()


def gtap_synchronize() -> int:
    return lib().gtap_synchronize()


def gtap_get_runtime_error_code() -> int:
    c = ctypes.c_int(0)
    _check(lib().gtap_get_runtime_error_code(ctypes.byref(c)), "gtap_get_runtime_error_code")
    return c.value


def gtap_get_runtime_error_string(code: int) -> str:
    return lib().gtap_get_runtime_error_string(code).decode()


def gtap_visualize_profile(app_name: str) -> None:
    lib().gtap_visualize_profile(app_name.encode())


def visualize_having_task_time(app_name: str) -> None:
    """Text timeline + having-task CSVs of the last profiled launch."""
    lib().gtap_visualize_having_task_time(app_name.encode())


def visualize_working_time(app_name: str) -> None:
    """Text timeline + working CSVs of the last profiled launch."""
    lib().gtap_visualize_working_time(app_name.encode())


_LOWERED: dict = {}


def lowered(name: str) -> ctypes.CDLL:
    """A program lowered from `#pragma gtap` source by gtap_amd.lower and built by
    gtap_amd.build (gtap_amd/_lib/libgtap_lowered_<name>.so). Its entry points are
    `gtap_entry_<kernel>(<kernel params>, [<result>*,] gtap_run_stats*)` and
    `gtap_lowered_<name>_record_bytes()`; their argument types come from the source, so the
    caller sets `argtypes` itself."""
    if name not in _LOWERED:
        lib()  # the runtime library first: the lowered library links against it
        path = os.path.join(HERE, "_lib", f"libgtap_lowered_{name}.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing: run `python -m gtap_amd.build`")
        L = ctypes.CDLL(path)
        rb = getattr(L, f"gtap_lowered_{name}_record_bytes")
        rb.argtypes = []
        rb.restype = ctypes.c_uint32
        _LOWERED[name] = L
    return _LOWERED[name]


_HIP: Optional[ctypes.CDLL] = None


def hip_runtime() -> ctypes.CDLL:
    """The HIP runtime library libgtap_amd.so is linked against (the instance already loaded
    in this process)."""
    global _HIP
    if _HIP is None:
        lib()
        path = "libamdhip64.so"
        try:
            with open("/proc/self/maps") as f:
                for line in f:
                    if "libamdhip64.so" in line:
                        path = line.split()[-1]
                        break
        except OSError:
            pass
        H = ctypes.CDLL(path)
        H.hipMemcpyToSymbol.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_int]
        H.hipMemcpyToSymbol.restype = ctypes.c_int
        H.hipMemcpyFromSymbol.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_size_t,
                                          ctypes.c_int]
        H.hipMemcpyFromSymbol.restype = ctypes.c_int
        H.hipDeviceSynchronize.argtypes = []
        H.hipDeviceSynchronize.restype = ctypes.c_int
        H.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        H.hipEventCreate.restype = ctypes.c_int
        H.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        H.hipEventRecord.restype = ctypes.c_int
        H.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        H.hipEventSynchronize.restype = ctypes.c_int
        H.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p, ctypes.c_void_p]
        H.hipEventElapsedTime.restype = ctypes.c_int
        H.hipEventDestroy.argtypes = [ctypes.c_void_p]
        H.hipEventDestroy.restype = ctypes.c_int
        _HIP = H
    return _HIP


class EntryModule:
    """A module written by gtap_amd.lower, driven the way the reference's host programs drive
    theirs (examples/fib/fib.cu:26-57): device variables set and read with hipMemcpyToSymbol /
    hipMemcpyFromSymbol, an entry kernel launched with <<<grid, block>>> (its generated
    `gtap_launch_<kernel>`), then hipDeviceSynchronize. One launch may follow each
    gtap_initialize / gtap_reset; launch() calls gtap_reset first unless told not to."""

    def __init__(self, path: str, name: str):
        lib()  # the runtime first: the module links against it and registers with it on load
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing")
        self.L = ctypes.CDLL(path)
        self.name = name
        f = getattr(self.L, f"gtap_lowered_{name}_symbol")
        f.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_uint64)]
        f.restype = ctypes.c_void_p
        self._symbol = f
        d = getattr(self.L, f"gtap_lowered_{name}_launch_dims")
        d.argtypes = [ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32)]
        d.restype = ctypes.c_int
        self._dims = d
        rb = getattr(self.L, f"gtap_lowered_{name}_record_bytes")
        rb.argtypes = []
        rb.restype = ctypes.c_uint32
        self.record_bytes = rb()

    def symbol(self, sym: str):
        n = ctypes.c_uint64(0)
        addr = self._symbol(sym.encode(), ctypes.byref(n))
        if not addr:
            raise KeyError(f"{self.name}: no device variable {sym}")
        return addr, n.value

    def set_symbol(self, sym: str, value, ctype=ctypes.c_int) -> None:
        addr, n = self.symbol(sym)
        v = ctype(value)
        if ctypes.sizeof(v) != n:
            raise ValueError(f"{sym} is {n} bytes, {ctype.__name__} is {ctypes.sizeof(v)}")
        _check(hip_runtime().hipMemcpyToSymbol(addr, ctypes.byref(v), n, 0, 1), f"hipMemcpyToSymbol({sym})")

    def get_symbol(self, sym: str, ctype=ctypes.c_int):
        addr, n = self.symbol(sym)
        v = ctype()
        if ctypes.sizeof(v) != n:
            raise ValueError(f"{sym} is {n} bytes, {ctype.__name__} is {ctypes.sizeof(v)}")
        _check(hip_runtime().hipMemcpyFromSymbol(ctypes.byref(v), addr, n, 0, 2), f"hipMemcpyFromSymbol({sym})")
        return v.value

    def launch_dims(self) -> tuple:
        g, b = ctypes.c_uint32(0), ctypes.c_uint32(0)
        _check(self._dims(ctypes.byref(g), ctypes.byref(b)), "launch_dims")
        return g.value, b.value

    def launch(self, kernel: str, argtypes: list, *args, grid: Optional[int] = None, block: Optional[int] = None,
               reset: bool = True, timed: bool = False):
        """Launches `kernel` (its parameters' ctypes in `argtypes`) and synchronizes; returns the
        status of the synchronize, or (status, ms) with hipEvents around the launch."""
        if reset:
            _check(lib().gtap_reset(), "gtap_reset")
        g0, b0 = self.launch_dims()
        f = getattr(self.L, f"gtap_launch_{kernel}")
        f.argtypes = [ctypes.c_uint32, ctypes.c_uint32] + list(argtypes)
        f.restype = ctypes.c_int
        H = hip_runtime()
        ev = [ctypes.c_void_p(), ctypes.c_void_p()]
        if timed:
            for e in ev:
                _check(H.hipEventCreate(ctypes.byref(e)), "hipEventCreate")
            H.hipEventRecord(ev[0], None)
        st = f(grid or g0, block or b0, *args)
        if timed:
            H.hipEventRecord(ev[1], None)
        sync = H.hipDeviceSynchronize()
        st = st or sync
        if not timed:
            return st
        ms = ctypes.c_float(0)
        H.hipEventSynchronize(ev[1])
        H.hipEventElapsedTime(ctypes.byref(ms), ev[0], ev[1])
        for e in ev:
            H.hipEventDestroy(e)
        return st, ms.value


def gtap_set_profiling(enable: bool, max_samples: int = 4096) -> int:
    return lib().gtap_set_profiling(1 if enable else 0, max_samples)


ROUND_PHASES = ("acquire", "load", "run", "stage", "store_wait", "join", "publish", "idle")


def round_phases() -> dict:
    """Round-phase breakdown of the last thread-mode launch run with timeline profiling on:
    shader-clock cycles per phase summed over the workers, the rounds, and cycles per round
    of each phase (idle tries are not rounds; their total is reported as is)."""
    import numpy as np
    out = np.zeros(9, dtype=np.uint64)
    _check(lib().gtap_get_round_phases(out.ctypes.data, 9), "gtap_get_round_phases")
    rounds = int(out[8])
    res = {"rounds": rounds, "cycles": {k: int(out[i]) for i, k in enumerate(ROUND_PHASES)}}
    res["cycles_per_round"] = {k: (int(out[i]) / rounds if rounds else 0.0) for i, k in enumerate(ROUND_PHASES[:7])}
    return res


def last_stats() -> dict:
    s = RunStats()
    _check(lib().gtap_get_last_stats(ctypes.byref(s)), "gtap_get_last_stats")
    return s.as_dict()


def device_info() -> dict:
    name = ctypes.create_string_buffer(64)
    cus = ctypes.c_int(0)
    hbm = ctypes.c_uint64(0)
    _check(lib().gtap_device_info(name, 64, ctypes.byref(cus), ctypes.byref(hbm)), "gtap_device_info")
    return {"arch": name.value.decode(), "num_cus": cus.value, "hbm_bytes": hbm.value}


# ---------------------------------------------------------------- task programs

def fib(n: int) -> tuple[int, dict]:
    """fib(n) as one persistent launch; returns (value, run statistics)."""
    r = ctypes.c_int64(0)
    s = RunStats()
    _check(lib().gtap_fib(n, ctypes.byref(r), ctypes.byref(s)), f"gtap_fib({n})")
    return r.value, s.as_dict()


def nqueens(n: int) -> tuple[int, dict]:
    """Number of n-queens placements as one persistent launch; returns (count, stats)."""
    r = ctypes.c_int64(0)
    s = RunStats()
    _check(lib().gtap_nqueens(n, ctypes.byref(r), ctypes.byref(s)), f"gtap_nqueens({n})")
    return r.value, s.as_dict()


def nqueens_slots(n: int) -> tuple[int, dict]:
    """nqueens() with result slots instead of the additive join (each child stores its count
    into the parent's frame, then counts the join down); returns (count, stats)."""
    r = ctypes.c_int64(0)
    s = RunStats()
    _check(lib().gtap_nqueens_slots(n, ctypes.byref(r), ctypes.byref(s)), f"gtap_nqueens_slots({n})")
    return r.value, s.as_dict()


def bench_fences(waves: int = 8192, iters: int = 64) -> dict:
    """ns per publication per wave: release fence, acquire fence, store-completion wait."""
    out = (ctypes.c_double * 3)()
    _check(lib().gtap_bench_fences(waves, iters, out), "gtap_bench_fences")
    return {"waves": waves, "release_fence_ns": out[0], "acquire_fence_ns": out[1], "store_wait_ns": out[2]}


def bench_hbm_copy(nbytes: int = 2 << 30, reps: int = 3) -> float:
    """Measured streaming-copy bandwidth of this GPU in GB/s (read + write bytes)."""
    g = ctypes.c_double(0)
    _check(lib().gtap_bench_hbm_copy(nbytes, reps, ctypes.byref(g)), "gtap_bench_hbm_copy")
    return g.value


def bench_atomics_contended(words: int, reps: int = 3) -> float:
    """Agent-scope relaxed atomic adds with every lane on one of `words` words (Gop/s)."""
    g = ctypes.c_double(0)
    _check(lib().gtap_bench_atomics_contended(reps, words, ctypes.byref(g)), "gtap_bench_atomics_contended")
    return g.value


def bench_random_access(table_bytes: int = 16 << 20, reps: int = 3) -> dict:
    """G random 4-byte accesses/s over a table: plain loads and agent-scope fetch-min."""
    out = (ctypes.c_double * 2)()
    _check(lib().gtap_bench_random_access(table_bytes, reps, out), "gtap_bench_random_access")
    return {"load_gps": out[0], "fetch_min_gps": out[1]}


def bench_atomics(reps: int = 3) -> float:
    """Measured agent-scope atomic fetch-add throughput of this GPU in Gop/s."""
    g = ctypes.c_double(0)
    _check(lib().gtap_bench_atomics(reps, ctypes.byref(g)), "gtap_bench_atomics")
    return g.value


# ---------------------------------------------------------------- graphs and BFS

RMAT_THRESHOLDS = (2448131359, 3264175145, 4080218931)  # (a, a+b, a+b+c) * 2^32, a,b,c = .57,.19,.19
INF_DEPTH = 0x3F3F3F3F


def rmat_csr(scale: int, edge_factor: int = 16, seed: int = 1):
    """R-MAT graph (the library's own host generator) as CSR numpy arrays."""
    import numpy as np
    nv = 1 << scale
    row = np.empty(nv + 1, dtype=np.int32)
    col = np.empty(nv * edge_factor, dtype=np.int32)
    ta, tab, tabc = RMAT_THRESHOLDS
    _check(lib().gtap_rmat_csr(scale, edge_factor, seed, ta, tab, tabc, row.ctypes.data, col.ctypes.data),
           "gtap_rmat_csr")
    return row, col


def gen_keys(n: int, seed: int = 123456789):
    import numpy as np
    out = np.empty(n, dtype=np.uint32)
    _check(lib().gtap_gen_keys(seed, n, out.ctypes.data), "gtap_gen_keys")
    return out


def bfs(row, col, source: int = 0, num_queues: int = 4, split: int = 64):
    """BFS depths from host CSR arrays (copied to the GPU and back); returns (depth, stats)."""
    import numpy as np
    row = np.ascontiguousarray(row, dtype=np.int32)
    col = np.ascontiguousarray(col, dtype=np.int32)
    nv = row.size - 1
    depth = np.empty(nv, dtype=np.int32)
    s = RunStats()
    _check(lib().gtap_bfs_host(row.ctypes.data, col.ctypes.data, nv, col.size, source, depth.ctypes.data,
                               num_queues, split, ctypes.byref(s)), "gtap_bfs_host")
    return depth, s.as_dict()


class DeviceBuffer:
    """Device memory owned through the C ABI (no torch in the process)."""

    def __init__(self, nbytes: int):
        p = ctypes.c_void_p(0)
        _check(lib().gtap_device_malloc(ctypes.byref(p), max(1, nbytes)), "gtap_device_malloc")
        self.ptr = p.value
        self.nbytes = nbytes

    @classmethod
    def from_array(cls, a):
        b = cls(a.nbytes)
        _check(lib().gtap_memcpy_htod(b.ptr, a.ctypes.data, a.nbytes), "gtap_memcpy_htod")
        return b

    def to_array(self, out):
        _check(lib().gtap_memcpy_dtoh(out.ctypes.data, self.ptr, out.nbytes), "gtap_memcpy_dtoh")
        return out

    def free(self):
        if self.ptr:
            lib().gtap_device_free(self.ptr)
            self.ptr = None


def nqueens_cutoff(n: int, cutoff: int = 7) -> tuple[int, dict]:
    """n-queens at the reference's benchmark settings (tasks down to row `cutoff`, sequential
    search below, no taskwait); returns (solutions, stats)."""
    r = ctypes.c_int64(0)
    s = RunStats()
    _check(lib().gtap_nqueens_cutoff(n, cutoff, ctypes.byref(r), ctypes.byref(s)), "gtap_nqueens_cutoff")
    return r.value, s.as_dict()


def bfs_block(row, col, source: int = 0):
    """Block-mode BFS depths (one vertex task per workgroup); returns (depth, stats)."""
    import numpy as np
    row = np.ascontiguousarray(row, dtype=np.int32)
    col = np.ascontiguousarray(col, dtype=np.int32)
    nv = row.size - 1
    drow, dcol, ddep = DeviceBuffer.from_array(row), DeviceBuffer.from_array(col), DeviceBuffer(nv * 4)
    try:
        s = RunStats()
        _check(lib().gtap_bfs_block(drow.ptr, dcol.ptr, nv, col.size, source, ddep.ptr, ctypes.byref(s)),
               "gtap_bfs_block")
        depth = ddep.to_array(np.empty(nv, dtype=np.int32))
    finally:
        for b in (drow, dcol, ddep):
            b.free()
    return depth, s.as_dict()


# ---------------------------------------------------------------- SpMV and the binary tree

def spmv_random(rows: int, avg_nnz: int = 300, seed: int = 42):
    """(row_ptr, col_idx, val) of the library's random CSR matrix (rows x rows)."""
    import numpy as np
    row_ptr = np.empty(rows + 1, dtype=np.int32)
    nnz = ctypes.c_int64(0)
    _check(lib().gtap_spmv_random(rows, avg_nnz, seed, row_ptr.ctypes.data, None, None, ctypes.byref(nnz)),
           "gtap_spmv_random")
    col = np.empty(nnz.value, dtype=np.int32)
    val = np.empty(nnz.value, dtype=np.float64)
    _check(lib().gtap_spmv_random(rows, avg_nnz, seed, row_ptr.ctypes.data, col.ctypes.data, val.ctypes.data,
                                  ctypes.byref(nnz)), "gtap_spmv_random")
    return row_ptr, col, val


def spmv(row_ptr, col_idx, val, x, cutoff: int = 16):
    """y = A x on the GPU (task per row, divide-and-conquer over row ranges); returns (y, stats)."""
    import numpy as np
    row_ptr = np.ascontiguousarray(row_ptr, dtype=np.int32)
    col_idx = np.ascontiguousarray(col_idx, dtype=np.int32)
    val = np.ascontiguousarray(val, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float64)
    rows = row_ptr.size - 1
    y = np.empty(rows, dtype=np.float64)
    s = RunStats()
    _check(lib().gtap_spmv_host(row_ptr.ctypes.data, col_idx.ctypes.data if col_idx.size else None,
                                val.ctypes.data if val.size else None, x.ctypes.data if x.size else None,
                                y.ctypes.data if rows else None, rows, x.size, cutoff, ctypes.byref(s)),
           "gtap_spmv_host")
    return y, s.as_dict()


def tree_input(n: int, seed: int = 0xC0FFEE):
    """The binary-tree workload's input array (float64 in [-1, 1))."""
    import numpy as np
    out = np.empty(n, dtype=np.float64)
    _check(lib().gtap_tree_input(seed, n, out.ctypes.data), "gtap_tree_input")
    return out


def tree(height: int, mem_ops: int, compute_iters: int, inp):
    """The binary-tree workload on the GPU; returns (out: one float64 per node, stats)."""
    import numpy as np
    inp = np.ascontiguousarray(inp, dtype=np.float64)
    nodes = (1 << (height + 1)) - 1
    din = DeviceBuffer.from_array(inp)
    dout = DeviceBuffer(nodes * 8)
    try:
        s = RunStats()
        _check(lib().gtap_tree(height, mem_ops, compute_iters, din.ptr, inp.size, dout.ptr, ctypes.byref(s)),
               "gtap_tree")
        out = dout.to_array(np.empty(nodes, dtype=np.float64))
    finally:
        din.free()
        dout.free()
    return out, s.as_dict()


# ---------------------------------------------------------------- the reference's file formats

def read_csr(path: str):
    """(row, col) of a graph CSR file (int32 N, int32 M, int32 row[N+1], int32 col[M])."""
    import numpy as np
    nv, ne = ctypes.c_int32(0), ctypes.c_int64(0)
    _check(lib().gtap_csr_file_info(path.encode(), ctypes.byref(nv), ctypes.byref(ne)), "gtap_csr_file_info")
    row = np.empty(nv.value + 1, dtype=np.int32)
    col = np.empty(ne.value, dtype=np.int32)
    _check(lib().gtap_csr_file_read(path.encode(), row.ctypes.data, col.ctypes.data if ne.value else None),
           "gtap_csr_file_read")
    return row, col


def write_csr(path: str, row, col) -> None:
    import numpy as np
    row = np.ascontiguousarray(row, dtype=np.int32)
    col = np.ascontiguousarray(col, dtype=np.int32)
    _check(lib().gtap_csr_file_write(path.encode(), row.size - 1, col.size, row.ctypes.data,
                                     col.ctypes.data if col.size else None), "gtap_csr_file_write")


def read_keys(path: str):
    """uint32 bit patterns of a key-vector file (size_t n, int32 data[n])."""
    import numpy as np
    n = ctypes.c_uint64(0)
    _check(lib().gtap_keys_file_info(path.encode(), ctypes.byref(n)), "gtap_keys_file_info")
    out = np.empty(n.value, dtype=np.uint32)
    _check(lib().gtap_keys_file_read(path.encode(), out.ctypes.data if n.value else None), "gtap_keys_file_read")
    return out


def write_keys(path: str, keys) -> None:
    import numpy as np
    keys = np.ascontiguousarray(keys, dtype=np.uint32)
    _check(lib().gtap_keys_file_write(path.encode(), keys.size, keys.ctypes.data if keys.size else None),
           "gtap_keys_file_write")


def read_spmv(path: str):
    """(rows, cols, row_ptr, col_idx, val) of an SpMV CSR file."""
    import numpy as np
    r, c, z = ctypes.c_int32(0), ctypes.c_int32(0), ctypes.c_int32(0)
    _check(lib().gtap_spmv_file_info(path.encode(), ctypes.byref(r), ctypes.byref(c), ctypes.byref(z)),
           "gtap_spmv_file_info")
    row_ptr = np.empty(r.value + 1, dtype=np.int32)
    col = np.empty(z.value, dtype=np.int32)
    val = np.empty(z.value, dtype=np.float64)
    _check(lib().gtap_spmv_file_read(path.encode(), row_ptr.ctypes.data, col.ctypes.data if z.value else None,
                                     val.ctypes.data if z.value else None), "gtap_spmv_file_read")
    return r.value, c.value, row_ptr, col, val


def write_spmv(path: str, cols: int, row_ptr, col_idx, val) -> None:
    import numpy as np
    row_ptr = np.ascontiguousarray(row_ptr, dtype=np.int32)
    col_idx = np.ascontiguousarray(col_idx, dtype=np.int32)
    val = np.ascontiguousarray(val, dtype=np.float64)
    _check(lib().gtap_spmv_file_write(path.encode(), row_ptr.size - 1, cols, col_idx.size, row_ptr.ctypes.data,
                                      col_idx.ctypes.data if col_idx.size else None,
                                      val.ctypes.data if val.size else None), "gtap_spmv_file_write")


# ---------------------------------------------------------------- sorting

def cilksort(keys, leaf: int = 4096, merge_leaf: int = 32768):
    """Sorted copy of uint32 keys (block-cooperative cilksort on the GPU); returns (keys, stats)."""
    import numpy as np
    a = np.ascontiguousarray(keys, dtype=np.uint32).copy()
    s = RunStats()
    _check(lib().gtap_cilksort_host(a.ctypes.data if a.size else None, a.size, leaf, merge_leaf, ctypes.byref(s)),
           "gtap_cilksort_host")
    return a, s.as_dict()


# ---------------------------------------------------------------- job streams

def jobs_generate(seed: int, njobs: int):
    """(kind, param, key_seed) arrays of the mixed job stream (the library's generator)."""
    import numpy as np
    kind = np.empty(njobs, dtype=np.int32)
    param = np.empty(njobs, dtype=np.int32)
    kseed = np.empty(njobs, dtype=np.uint64)
    _check(lib().gtap_jobs_generate(seed, njobs, kind.ctypes.data, param.ctypes.data, kseed.ctypes.data),
           "gtap_jobs_generate")
    return kind, param, kseed


class Stream:
    """A generated job stream with device-resident sort inputs (see gtap_stream_* in gtap.h)."""

    def __init__(self, seed: int, njobs: int):
        h = ctypes.c_void_p(0)
        _check(lib().gtap_stream_create(seed, njobs, ctypes.byref(h)), "gtap_stream_create")
        self.h = h.value
        self.seed = seed
        self.njobs = njobs
        self.kind, self.param, self.key_seed = jobs_generate(seed, njobs)

    def close(self):
        if self.h:
            lib().gtap_stream_destroy(self.h)
            self.h = None

    def sorted_keys(self, job: int):
        import numpy as np
        out = np.empty(int(self.param[job]), dtype=np.uint32)
        _check(lib().gtap_stream_sorted_keys(self.h, job, out.ctypes.data), "gtap_stream_sorted_keys")
        return out

    def run_batched(self):
        import numpy as np
        res = np.empty(self.njobs, dtype=np.int64)
        lat = np.empty(self.njobs, dtype=np.float64)
        s = RunStats()
        _check(lib().gtap_stream_run_batched(self.h, res.ctypes.data, lat.ctypes.data, ctypes.byref(s)),
               "gtap_stream_run_batched")
        return res, lat, s.as_dict()

    def run_per_job(self):
        import numpy as np
        res = np.empty(self.njobs, dtype=np.int64)
        lat = np.empty(self.njobs, dtype=np.float64)
        tot = ctypes.c_double(0)
        _check(lib().gtap_stream_run_per_job(self.h, res.ctypes.data, lat.ctypes.data, ctypes.byref(tot)),
               "gtap_stream_run_per_job")
        return res, lat, tot.value

    def set_merge_leaf(self, merge_leaf: int) -> None:
        """Sort jobs merge runs of at most merge_leaf keys (512..16384) on one wave."""
        _check(lib().gtap_stream_set_merge_leaf(self.h, int(merge_leaf)), "gtap_stream_set_merge_leaf")

    def exec_ms(self):
        """Device time from admission to completion of each job of the last run, in ms."""
        import numpy as np
        out = np.empty(self.njobs, dtype=np.float64)
        _check(lib().gtap_stream_exec_ms(self.h, out.ctypes.data), "gtap_stream_exec_ms")
        return out

    def serve(self, rate: float, arrival_seed: int = 1, per_job_launch: bool = False, max_live: int = 0):
        import numpy as np
        res = np.empty(self.njobs, dtype=np.int64)
        lat = np.empty(self.njobs, dtype=np.float64)
        tot = ctypes.c_double(0)
        s = RunStats()
        _check(lib().gtap_stream_serve(self.h, rate, arrival_seed, 1 if per_job_launch else 0, max_live,
                                       res.ctypes.data, lat.ctypes.data, ctypes.byref(tot), ctypes.byref(s)),
               "gtap_stream_serve")
        return res, lat, tot.value, s.as_dict()
