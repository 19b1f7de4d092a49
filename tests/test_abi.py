"""The C-ABI library loads and exports every symbol include/gtap.h declares (CPU only,
no compute call)."""
import ctypes
import os
import re

import pytest

import gtap_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gtap.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|void|const char\s*\*)\s+(gtap_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for s in ("gtap_initialize", "gtap_finalize", "gtap_reset", "gtap_synchronize",
              "gtap_get_runtime_error_code", "gtap_visualize_profile", "gtap_fib"):
        assert s in syms


@pytest.mark.skipif(not os.path.exists(gtap_amd.LIB_PATH), reason="libgtap_amd.so not built")
def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(gtap_amd.LIB_PATH)
    for s in declared_symbols():
        assert hasattr(L, s), s
        assert s in gtap_amd.SIGNATURES, f"{s} has no ctypes signature in gtap_amd"


@pytest.mark.skipif(not os.path.exists(gtap_amd.LIB_PATH), reason="libgtap_amd.so not built")
def test_bindings_cover_only_declared_symbols():
    declared = set(declared_symbols())
    assert set(gtap_amd.SIGNATURES) <= declared


def test_error_strings_without_gpu():
    if not os.path.exists(gtap_amd.LIB_PATH):
        pytest.skip("libgtap_amd.so not built")
    assert gtap_amd.gtap_get_runtime_error_string(0) == "No error"
    assert gtap_amd.gtap_get_runtime_error_string(3) == "Queue overflow"
    assert gtap_amd.gtap_get_runtime_error_string(999) == "Unknown error"


def test_struct_sizes_match_c_layout():
    assert ctypes.sizeof(gtap_amd.Job) == 128
    assert ctypes.sizeof(gtap_amd.ErrorReport) == 36
    assert ctypes.sizeof(gtap_amd.RunStats) == 8 * 12 + 16
    assert ctypes.sizeof(gtap_amd.Config) == 8 + 4 * 16


def test_config_fields_and_scheduler_check_without_gpu():
    """gtap_set_config stores the configuration (no device call before gtap_initialize);
    gtap_get_config fills the defaults; an unknown scheduler is refused."""
    if not os.path.exists(gtap_amd.LIB_PATH):
        pytest.skip("libgtap_amd.so not built")
    names = [f for f, _ in gtap_amd.Config._fields_]
    assert names[-5:] == ["idle_sleep", "scheduler", "epaq_publish", "xcd_steal", "reserved"]
    assert gtap_amd.gtap_set_config(xcd_steal=3) != 0
    assert gtap_amd.gtap_set_config(scheduler=9) != 0
    assert gtap_amd.gtap_set_config(epaq_publish=2) != 0
    assert gtap_amd.gtap_set_config(epaq_publish=1) == 0
    assert gtap_amd.gtap_get_config()["epaq_publish"] == 1
    assert gtap_amd.gtap_set_config(scheduler=2, idle_sleep=4) == 0
    c = gtap_amd.gtap_get_config()
    assert c["scheduler"] == 2 and c["idle_sleep"] == 4
    assert gtap_amd.gtap_set_config() == 0
    c = gtap_amd.gtap_get_config()
    assert c["scheduler"] == 1 and c["idle_sleep"] == 32 and c["probe_lanes"] == 8


def test_initialize_rejects_bad_layout_without_gpu():
    """Record strides beyond 65520 bytes (result offsets must stay below the additive join's
    16-bit marker), odd deque capacities and queue counts are refused before any device call."""
    if not os.path.exists(gtap_amd.LIB_PATH):
        pytest.skip("libgtap_amd.so not built")
    try:
        for bad in ({"record_stride": 65536}, {"record_stride": 40}, {"deque_capacity": 5000}, {"num_queues": 9}):
            assert gtap_amd.gtap_set_config(**bad) == 0
            assert gtap_amd.gtap_initialize() == 1, bad  # hipErrorInvalidValue
            assert gtap_amd.gtap_get_runtime_error_code() == 11, bad  # GTAP_ERROR_INVALID_ARGUMENT
    finally:
        assert gtap_amd.gtap_set_config() == 0


def test_noalloc_entry_refuses_an_uninitialized_runtime_without_gpu():
    """gtap_cilksort_noalloc never initializes (initialization allocates): before
    gtap_initialize it returns GTAP_ERROR_NOT_INITIALIZED without any device call."""
    if not os.path.exists(gtap_amd.LIB_PATH):
        pytest.skip("libgtap_amd.so not built")
    assert gtap_amd.lib().gtap_cilksort_noalloc(None, 16, 0, 0, None) != 0
    assert gtap_amd.gtap_get_runtime_error_code() == 12  # GTAP_ERROR_NOT_INITIALIZED
