"""The CadR C++ facade (include/CadR/): its test program, built with the library
(tests/cpp/facade_test.cpp), run device-less on the CPU and on the GPU."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "cadr_amd", "bin", "facade_test")


def _run(mode, timeout):
    assert os.path.exists(BIN), "facade_test is not built (make -C cadr_amd/csrc)"
    p = subprocess.run([BIN, mode], capture_output=True, text=True, timeout=timeout)
    print(p.stdout[-2000:], p.stderr[-2000:])
    assert p.returncode == 0, p.stderr
    assert p.stdout.strip().endswith("OK")


def test_facade_host_bookkeeping():
    """Handle numbering and table levels (1 -> 2 at handle 2048, 2 -> 3 at
    4,194,304), allocator placement of tests/DataAllocationTest.cpp, StateSet
    swap-removal and DFS draw ranges."""
    _run("cpu", 120)


@pytest.mark.gpu
def test_facade_frames_on_gpu(gpu):
    """Renderer frames at handle levels 1, 2, 3: every indirect command and
    DrawablePointers record against the host bookkeeping."""
    _run("gpu", 110)
