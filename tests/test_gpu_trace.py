"""The track producer on the GPU: ray tracing through the pin-cell lattice.

Parity bar: bit-exact against the C restatement of the tracer (FP64, no
contraction, correctly rounded division and square root on both sides), so
segment counts, lengths, regions and z are compared with array_equal.  A
traced stream then goes through the sweep like any other, against the FP64
oracle.
"""
import numpy as np
import pytest

import pyoracle
import smoc
from conftest import normwise

pytestmark = pytest.mark.gpu

GEOMETRIES = [
    (6, 5, 4, 1.26, 0.54, 7.0),
    (17, 17, 9, 1.26, 0.475, 30.0),   # a PWR assembly-like lattice
    (3, 40, 1, 0.8, 0.0, 0.5),        # no pins, one layer, flat box
    (1, 1, 50, 2.0, 0.9, 200.0),      # one tall cell
]


@pytest.mark.parametrize("geo", GEOMETRIES)
@pytest.mark.parametrize("seed", [1, 77])
def test_trace_bit_exact_against_restatement(gpu_ready, geo, seed):
    T = 2000
    print(f"case trace {geo} seed {seed}", flush=True)
    off, ln, rg, zz = smoc.trace(smoc.Geometry(*geo), T, seed)
    o_off, o_ln, o_rg, o_zz = pyoracle.trace(pyoracle.Geometry(*geo), seed, T)
    assert np.array_equal(off, o_off)
    assert np.array_equal(ln.view(np.uint32), o_ln.view(np.uint32))
    assert np.array_equal(rg, o_rg)
    assert np.array_equal(zz.view(np.uint32), o_zz.view(np.uint32))
    print(f"  {off[-1]} segments, {np.diff(off).mean():.1f} per track", flush=True)


def test_traced_stream_sweeps_like_any_other(gpu_ready):
    g = smoc.Geometry(17, 17, 9, 1.26, 0.475, 30.0)
    prob = smoc.traced_problem(g, 3000, G=100, F=5, seed=5)
    print(f"case sweep of a traced stream: R {prob.R}, S {prob.S}", flush=True)
    for tally in ("atomic-f32", "det-f64"):
        with smoc.Context(100, prob.R, 5, tally=tally) as ctx:
            ctx.load(prob)
            ctx.run()
            flux, psi = ctx.results()
        of, op = pyoracle.sweep_f64(prob, pyoracle.EXP_NATIVE, nthreads=16)
        e = (normwise(flux, of), normwise(psi, op))
        print(f"  {tally}: normwise {e[0]:.2e} / {e[1]:.2e}", flush=True)
        assert max(e) < 1e-5


def test_bad_geometry_rejected(gpu_ready):
    off = np.empty(11, np.int64)
    S = smoc._i64()
    for geo in ((0, 5, 4, 1.0, 0.3, 1.0), (5, 5, 4, 1.0, 0.5, 1.0), (5, 5, 4, 1.0, 0.3, 0.0)):
        rc = smoc.lib().smoc_trace_count(0, smoc.Geometry(*geo), 1, 10, off, S)
        assert rc == smoc.ERR_ARG, geo
