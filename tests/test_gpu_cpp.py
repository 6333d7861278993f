"""Runs the C++ host-API test binary (tests/cpp/api_test.cpp, built by
__graft_entry__.build()) on the GPU."""
import os
import subprocess

import pytest

from helpers import ROOT

pytestmark = pytest.mark.gpu


def test_cpp_host_api():
    exe = os.path.join(ROOT, "tests", "cpp", "api_test")
    assert os.path.exists(exe), "tests/cpp/api_test is not built (run __graft_entry__.build())"
    res = subprocess.run([exe], capture_output=True, text=True, timeout=240)
    print(res.stdout)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "all ok" in res.stdout
    # the dense data square's root: SHA-256 of the cells column by column,
    # the order the reference hashes a square filled column by column
    import hashlib
    import numpy as np
    cols, rows, p = 7, 5, 2013265921
    cells = [(r * cols + c) * 2654435761 % p for c in range(cols) for r in range(rows)]
    want = hashlib.sha256(np.array(cells, dtype="<u4").tobytes()).hexdigest()
    assert f"data_square_root {want}" in res.stdout
