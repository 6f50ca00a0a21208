"""The resident serving ring's protocol on the CPU (tests/resident_ring_model.cpp), under
ThreadSanitizer: with the waves' relaxed read of the ring's tail (the kernel as it ran on the
device), nothing orders their reads of a request row after the host's writes of it -- a data race,
which on the device returned the slot's previous request.  With an acquire read, the protocol is
race-free and every request gets its own result."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "resident_ring_model.cpp")

pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not found")


def build_and_run(tmp_path, order):
    exe = os.path.join(str(tmp_path), f"ring_{order}")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-pthread", "-fsanitize=thread",
           f"-DTAIL_ORDER=std::memory_order_{order}", SRC, "-o", exe]
    b = subprocess.run(cmd, capture_output=True, text=True)
    if b.returncode != 0:
        pytest.skip(f"ThreadSanitizer build failed: {b.stderr[-400:]}")
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=0 report_signal_unsafe=0")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120, env=env)
    return r.returncode, r.stdout + r.stderr


def test_relaxed_tail_read_races_with_the_row_writes(tmp_path):
    rc, out = build_and_run(tmp_path, "relaxed")
    assert "WARNING: ThreadSanitizer: data race" in out
    # the racing read is the wave's read of the request (row, k or queue size)
    assert "in wave" in out or "wave " in out
    assert rc != 0


def test_acquire_tail_read_is_race_free_and_correct(tmp_path):
    rc, out = build_and_run(tmp_path, "acquire")
    assert "ThreadSanitizer" not in out, out[-2000:]
    assert "retired 4000 wrong 0" in out
    assert rc == 0
