"""Shared test setup: import paths, the `gpu` marker, problem helpers."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "simplemoc-kernel_amd")
for p in (PKG, os.path.join(ROOT, "oracle"), ROOT):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X) and libsmoc.so")


def bf16_round(x):
    """FP32 -> BF16 -> FP32, round to nearest even (finite inputs)."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def f16_round(x):
    return np.asarray(x, np.float32).astype(np.float16).astype(np.float32)


def quantized(prob, storage):
    """The problem as the device stores it: sigT and fine_source rounded."""
    import copy
    q = copy.copy(prob)
    if storage == "bf16":
        q.sigT, q.fine_source = bf16_round(prob.sigT), bf16_round(prob.fine_source)
    elif storage == "f16":
        q.sigT, q.fine_source = f16_round(prob.sigT), f16_round(prob.fine_source)
    return q


def rel(a, b):
    return abs(a - b) / abs(b)


def normwise(x, ref):
    x = np.asarray(x, np.float64).ravel()
    ref = np.asarray(ref, np.float64).ravel()
    return float(np.linalg.norm(x - ref) / np.linalg.norm(ref))


@pytest.fixture(scope="session")
def gpu_ready():
    import smoc
    n = smoc.device_count()
    if n < 1:
        pytest.fail("no HIP device visible to libsmoc (gpu tests need an MI355X)")
    return n
