"""The polynomial of the accurate exponential's series branch (r4 in
k_sweep): the FP32 constants in the kernel source are the ones
csrc/fit_r4.py computes, and evaluated in FP32 they keep 1 - e^-tau within
1e-7 absolute over [0, 1) (checked against long-double expm1, a library
routine independent of the fit)."""
import importlib.util
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "simplemoc-kernel_amd", "csrc")


def _fit():
    spec = importlib.util.spec_from_file_location("fit_r4", os.path.join(CSRC, "fit_r4.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def kernel_coefficients():
    src = open(os.path.join(CSRC, "smoc_kernels.hip")).read()
    body = re.search(r"float r4_series\(float tau\)\s*\{(.*?)\n\}", src, re.S).group(1)
    lits = re.findall(r"([+-]?)\s*(0x[0-9a-fA-F.]+p[+-]?\d+)f", body)
    # Horner order in the source: highest degree first; signs from the operators
    vals = [float.fromhex(v) * (-1.0 if sg == "-" else 1.0) for sg, v in lits]
    return np.array(vals[::-1], np.float32)


def test_kernel_constants_are_the_fit():
    fit = _fit()
    k = kernel_coefficients()
    assert k.size == fit.DEG + 1
    assert np.array_equal(k, fit.coefficients())


def test_series_error_bounds():
    rel, ev = _fit().errors(kernel_coefficients(), points=400_001)
    assert rel < 3e-7
    assert ev < 1e-7
    # a dropped or sign-flipped term is caught
    bad = kernel_coefficients().copy()
    bad[2] = -bad[2]
    assert _fit().errors(bad, points=20_001)[1] > 1e-4
