"""CPU model of the runtime's lock-free protocols (tests/protocol/gtap_protocol_model.cpp):
the deque word's owner pop / thief steal-half, the Vyukov id-ring cells, all-tasks
termination with job owners counted at the reset, and the additive-join word packing,
each checked over randomized interleavings under ThreadSanitizer (g++ -fsanitize=thread;
a data race or a broken invariant fails the run)."""
import os
import shutil
import subprocess

import pytest

SRC = os.path.join(os.path.dirname(os.path.abspath(__file__)), "protocol", "gtap_protocol_model.cpp")


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ is not available")
def test_protocol_model_under_thread_sanitizer(tmp_path):
    exe = str(tmp_path / "protocol_model")
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=thread", "-pthread", SRC, "-o", exe],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1 exitcode=66")
    r = subprocess.run([exe, "3"], capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-6000:])
    assert "protocol model ok" in r.stdout
