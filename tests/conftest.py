"""Shared test set-up.  GPU tests are marked `gpu` and skip without a device.

The oracle (oracle/liboracle.so) is the checker; it is built here with gcc
when missing.  The product library must already be built
(`python -m femx_amd.build`); tests never fall back to anything else.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    if not os.path.exists(os.path.join(ROOT, "oracle", "liboracle.so")):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "oracle"], check=True,
                       capture_output=True)


def pytest_collection_modifyitems(config, items):
    import torch
    if torch.cuda.is_available():
        return
    skip = pytest.mark.skip(reason="no HIP device")
    for item in items:
        if "gpu" in item.keywords:
            item.add_marker(skip)


def golden(name):
    with np.load(os.path.join(GOLDEN, name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def close(a, b, rtol=1e-12):
    """Per-result FP64 bar: ||a-b||_2 <= rtol*||b||_2 and
    max|a-b| <= rtol*||b||_inf (mixed norm; pure entrywise relative error is
    ill-posed on entries that are mathematically zero)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    if a.shape != b.shape:
        return False
    if b.size == 0:
        return True
    nb2 = np.linalg.norm(b)
    nbi = np.abs(b).max()
    d = np.abs(a - b)
    return (np.linalg.norm(a - b) <= rtol * max(nb2, 1e-300)
            and d.max() <= rtol * max(nbi, 1e-300))
