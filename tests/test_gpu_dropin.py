"""The reference's own integration test (tests/test_traverse.cpp, 12 rays / 27 checks),
compiled unchanged and linked against the drop-in shim and librtx, runs on the GPU."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "oracle", "_ref", "test_traverse_rtx")

pytestmark = pytest.mark.gpu


@pytest.mark.skipif(not os.path.exists(BIN), reason="drop-in test binary not built (needs the reference checkout)")
def test_reference_test_traverse_passes_on_librtx(tmp_path):
    r = subprocess.run([BIN], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print(r.stdout[-3000:])
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "Results: 27/27 passed" in r.stdout
