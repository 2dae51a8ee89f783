"""The femx-shaped C++ interface (femx_amd/include/femx_hip) through its
C++ test program, which restates femx's own unit tests and checks the C++
contracts against femx's outputs: host tests on the CPU, and every device
test on a GPU.  The program reads the gls_constrained_n4 fixture (femx's
NavierStokesModel under ConstrainedTimeResidual) as raw arrays written
here."""
import os
import subprocess

import numpy as np
import pytest

from conftest import golden
from femx_amd import build as B


def _write_fixture(d):
    g = golden("gls_constrained_n4.npz")
    for k, v in g.items():
        a = np.asarray(v)
        if a.dtype.kind == "f":
            a.astype("<f8").tofile(os.path.join(d, k + ".f64"))
        else:
            a.astype("<i4").tofile(os.path.join(d, k + ".i32"))


def _run(tmp_path):
    assert os.path.exists(B.CPP_TEST), "C++ tests are not built (python -m femx_amd.build)"
    _write_fixture(str(tmp_path))
    r = subprocess.run([B.CPP_TEST, str(tmp_path)], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    return r


def test_cpp_interface_host(tmp_path):
    r = _run(tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "FAIL" not in r.stdout


@pytest.mark.gpu
def test_cpp_interface_device(tmp_path):
    r = _run(tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "SKIP" not in r.stdout and "FAIL" not in r.stdout
