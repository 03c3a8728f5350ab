"""CPU tests of the C++ agent/pipe host (tests/cpp/ut_pipeline.cpp), plain and
under ThreadSanitizer.  The binaries are built by __graft_entry__.build()."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD = os.path.join(ROOT, "jenhine_amd", "host", "build")


@pytest.mark.parametrize("binary", ["ut_pipeline", "ut_pipeline_tsan"])
def test_pipe_and_agent_units(binary):
    path = os.path.join(BUILD, binary)
    if not os.path.exists(path):
        pytest.skip(f"{binary} not built")
    r = subprocess.run([path], capture_output=True, text=True, timeout=600)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ThreadSanitizer" not in r.stderr, r.stderr
    assert r.stdout.count("PASS") >= 8
