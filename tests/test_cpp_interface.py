"""The femx-shaped C++ interface (femx_amd/include/femx_hip) through its
C++ test program, which restates femx's own unit tests: host tests on the
CPU, and every device test on a GPU."""
import os
import subprocess

import pytest

from femx_amd import build as B


def _run():
    assert os.path.exists(B.CPP_TEST), "C++ tests are not built (python -m femx_amd.build)"
    r = subprocess.run([B.CPP_TEST], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    return r


def test_cpp_interface_host():
    r = _run()
    assert r.returncode == 0, r.stdout + r.stderr
    assert "FAIL" not in r.stdout


@pytest.mark.gpu
def test_cpp_interface_device():
    r = _run()
    assert r.returncode == 0, r.stdout + r.stderr
    assert "SKIP" not in r.stdout and "FAIL" not in r.stdout
