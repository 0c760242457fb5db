"""The kernel-group timing bookkeeping of the library (csrc/mpm_events.h), built on the host with
a stand-in for the HIP event calls that fails where the test says (no GPU needed): failed event
creation or recording is reported and never leaves an event index that is read later."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_event_marks_with_failing_event_calls(tmp_path):
    exe = tmp_path / "event_marks_check"
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                    os.path.join(HERE, "native", "event_marks_check.cpp"), "-o", str(exe)], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "event marks ok" in r.stdout
