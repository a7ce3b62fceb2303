#!/usr/bin/env python3
"""Coefficients of the degree-4 polynomial that k_sweep's accurate exponential
uses for r4(tau) = sum_{n>=0} (-tau)^n / (n+4)! on [0, 1), so that
1 - e^-tau = tau - tau^2 (1/2 - tau/6 + tau^2 r4(tau)).

Method: Lawson's iteratively reweighted least squares toward the minimax
polynomial of the relative error, on 4000 Chebyshev nodes of [0, 1]; the
coefficients are then rounded to FP32.  Prints them as C hex-float literals
with the resulting errors: r4's relative error, and the absolute error of ev
computed in FP32 with fused multiply-adds (emulated) against e^-tau in long
double, over every FP32 tau in [0, 1) on a 2e6-point grid.
"""
from math import factorial

import numpy as np

DEG = 4


def r4_exact(t):
    t = np.asarray(t, np.longdouble)
    s = np.zeros_like(t)
    term = np.ones_like(t) / np.longdouble(factorial(4))
    for n in range(40):
        s += term
        term = term * (-t) / np.longdouble(n + 5)
    return s


def lawson(deg, iters=200):
    x = np.cos(np.pi * (np.arange(4000) + 0.5) / 4000)
    t = 0.5 + 0.5 * x
    y = np.asarray(r4_exact(t), np.float64)
    A = np.vander(t, deg + 1, increasing=True) / y[:, None]
    w = np.full(t.size, 1.0 / t.size)
    for _ in range(iters):
        sw = np.sqrt(w)
        c, *_ = np.linalg.lstsq(A * sw[:, None], sw, rcond=None)
        e = np.abs(A @ c - 1.0)
        w = w * e
        w /= w.sum()
    return c


def fma32(a, b, c):
    return (a.astype(np.float64) * b.astype(np.float64) + c.astype(np.float64)).astype(np.float32)


def coefficients():
    """FP32 coefficients, tau^0 first."""
    return lawson(DEG).astype(np.float32)


def errors(c32, points=2_000_001):
    """(max relative error of r4, max absolute error of ev) over [0, 1), both
    evaluated in FP32 as the kernel does."""
    c32 = np.asarray(c32, np.float32)
    t = np.linspace(0, 1, points).astype(np.float32)
    t = t[t < 1]
    r4 = np.full_like(t, c32[-1])
    for c in c32[-2::-1]:
        r4 = fma32(r4, t, np.full_like(t, c))
    exact = np.asarray(r4_exact(t), np.float64)
    t2 = (t * t).astype(np.float32)
    r = fma32(t2, r4, fma32(t, np.full_like(t, np.float32(-1 / 6)), np.full_like(t, np.float32(0.5))))
    ev = (t - (t2 * r).astype(np.float32)).astype(np.float32)
    ev_true = np.asarray(-np.expm1(-np.asarray(t, np.longdouble)), np.float64)
    return float((np.abs(r4 - exact) / exact).max()), float(np.abs(ev - ev_true).max())


def main():
    c32 = coefficients()
    print("coefficients (tau^0 .. tau^%d):" % DEG)
    for c in c32:
        print("   %sf" % float(c).hex().replace("0000000p", "p").replace("000p", "p"))
    rel, ev = errors(c32)
    print("r4 relative error (FP32 evaluation): %.3e" % rel)
    print("ev absolute error over [0, 1): %.3e" % ev)


if __name__ == "__main__":
    main()
