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
