"""Outer (source) iteration through the C ABI: sweep, source update from the
track-length flux estimate, repack, sweep again.

The oracle iterates the FP64 sweep and the FP64 restatement of the update
(pyoracle.source_update) on the same inputs; the device keeps its source as
an FP32 master and, for 16-bit storage, quantises it at every repack, which
the oracle reproduces (conftest.quantized) before each of its sweeps.
"""
import numpy as np
import pytest

import pyoracle
import smoc
from conftest import normwise, quantized

pytestmark = pytest.mark.gpu
TOL = 1e-5


def oracle_iterate(prob, scatter, q_ext, n, storage="f32"):
    vol = pyoracle.track_volumes(prob)
    src = np.asarray(prob.fine_source, np.float32)
    flux = psi = None
    for _ in range(n):
        p = smoc.Problem(prob.G, prob.R, prob.F, prob.sigT, src, prob.len, prob.region,
                         prob.z, prob.track_off, prob.psi_in)
        flux, psi = pyoracle.sweep_f64(quantized(p, storage), pyoracle.EXP_NATIVE, nthreads=16)
        # the update uses the FP32 sigT master, as the device does
        q = pyoracle.source_update(flux, vol, prob.sigT, scatter, q_ext)
        src = q.astype(np.float32)
    return flux, psi, src


def scattering(prob, cmax, seed):
    rng = np.random.default_rng(seed)
    return (rng.random((prob.R, prob.G)) * cmax).astype(np.float32)


@pytest.mark.parametrize("storage,tally,sorted_stream", [
    ("f32", "atomic-f32", False), ("f32", "det-f64", False), ("bf16", "det-f64", False),
    ("bf16", "det-f64", True), ("f32", "atomic-f64", True)])
def test_iterations_match_oracle(gpu_ready, storage, tally, sorted_stream):
    prob = smoc.generate(200_000, G=100, R=500, F=5, seed=41, sorted_stream=sorted_stream)
    sc = scattering(prob, 0.8, 1)
    print(f"case outer iteration x3 {storage} {tally} sorted {sorted_stream}", flush=True)
    with smoc.Context(100, 500, 5, storage=storage, tally=tally) as ctx:
        ctx.load(prob)
        ctx.set_scattering(sc)  # q_ext = the loaded source
        tm, ums, chg = ctx.iterate(3)
        flux, psi = ctx.results()
        src = ctx.source()
    of, op, osrc = oracle_iterate(prob, sc, prob.fine_source, 3, storage)
    e = (normwise(flux, of), normwise(psi, op), normwise(src, osrc))
    print(f"  normwise fine_flux {e[0]:.2e} psi {e[1]:.2e} source {e[2]:.2e}; last change "
          f"{chg:.3e}; sweeps {tm.total_ms:.2f} ms, updates {ums:.3f} ms", flush=True)
    assert max(e) <= TOL and chg > 0


def test_zero_scattering_keeps_the_external_source(gpu_ready):
    prob = smoc.generate(50_000, G=64, R=200, F=5, seed=42)
    q_ext = (np.random.default_rng(2).random((200, 5, 64))).astype(np.float32)
    with smoc.Context(64, 200, 5, tally="det-f64") as ctx:
        ctx.load(prob)
        ctx.set_scattering(np.zeros((200, 64), np.float32), q_ext)
        ctx.run()
        chg = ctx.source_update()
        assert np.array_equal(ctx.source(), q_ext)
        assert chg > 0
        ctx.iterate(2)
        f_it, p_it = ctx.results()
        assert ctx.source_update() == 0.0
    plain = smoc.Problem(64, 200, 5, prob.sigT, q_ext, prob.len, prob.region, prob.z,
                         prob.track_off, prob.psi_in)
    with smoc.Context(64, 200, 5, tally="det-f64") as ctx:
        ctx.load(plain)
        ctx.run()
        f, p = ctx.results()
    assert np.array_equal(f, f_it) and np.array_equal(p, p_it)


def test_infinite_medium_fixed_point(gpu_ready):
    """One region, flat source q* = q_ext / (1 - c), tracks entering at
    psi = q* / sigT: the update returns q* (closed form)."""
    G, c, qext, sig = 32, 0.6, 0.25, 3.0
    qstar = qext / (1 - c)
    prob = smoc.generate(40_000, G=G, R=1, F=2, seed=43, track_len=200)
    prob.sigT = np.full((1, G), sig, np.float32)
    prob.fine_source = np.full((1, 2, G), qstar, np.float32)
    prob.psi_in = np.full((prob.T, G), qstar / sig, np.float32)
    with smoc.Context(G, 1, 2, tally="det-f64") as ctx:
        ctx.load(prob)
        ctx.set_scattering(np.full((1, G), c, np.float32), np.full((1, 2, G), qext, np.float32))
        ctx.run()
        chg = ctx.source_update()
        src = ctx.source()
    print(f"  fixed point: change {chg:.2e}, max rel dev {np.abs(src / qstar - 1).max():.2e}",
          flush=True)
    assert np.abs(src / qstar - 1).max() < 1e-5 and chg < 1e-5


def test_fixed_point_is_stable_over_iterations(gpu_ready):
    """Starting at the infinite-medium fixed point, repeated iterations (graph
    replays) stay there."""
    G, c, qext, sig = 16, 0.5, 0.2, 1.5
    qstar = qext / (1 - c)
    prob = smoc.generate(30_000, G=G, R=1, F=3, seed=44, track_len=150)
    prob.sigT = np.full((1, G), sig, np.float32)
    prob.fine_source = np.full((1, 3, G), qstar, np.float32)
    prob.psi_in = np.full((prob.T, G), qstar / sig, np.float32)
    with smoc.Context(G, 1, 3, tally="atomic-f64") as ctx:
        ctx.load(prob)
        ctx.set_scattering(np.full((1, G), c, np.float32), np.full((1, 3, G), qext, np.float32))
        changes = [ctx.iterate(1, use_graph=1)[2] for _ in range(5)]
        src = ctx.source()
    print("  changes " + " ".join(f"{x:.2e}" for x in changes), flush=True)
    assert max(changes) < 1e-5 and np.abs(src / qstar - 1).max() < 1e-5


def test_graph_iterations_match_eager(gpu_ready):
    prob = smoc.generate(80_000, G=100, R=300, F=5, seed=45)
    sc = scattering(prob, 0.7, 4)
    out = []
    for graph in (0, 1):
        with smoc.Context(100, 300, 5, tally="det-f64") as ctx:
            ctx.load(prob)
            ctx.set_scattering(sc)
            ctx.iterate(4, use_graph=graph)
            out.append((*ctx.results(), ctx.source()))
    for a, b in zip(*out):
        assert np.array_equal(a, b)


def test_state_errors(gpu_ready):
    prob = smoc.generate(1000, G=16, R=10, F=5, seed=46)
    with smoc.Context(16, 10, 5) as ctx:
        ctx.load(prob)
        with pytest.raises(smoc.SmocError) as e:
            ctx.source_update()
        assert e.value.code == smoc.ERR_STATE
        with pytest.raises(smoc.SmocError) as e:
            ctx.set_scattering(np.full((10, 16), 1.0, np.float32))
        assert e.value.code == smoc.ERR_ARG
        ctx.set_scattering(np.full((10, 16), 0.5, np.float32))
        with pytest.raises(smoc.SmocError) as e:
            ctx.source_update()  # no run yet
        assert e.value.code == smoc.ERR_STATE
        ctx.set_regions(prob.sigT, prob.fine_source)  # clears the scattering
        with pytest.raises(smoc.SmocError) as e:
            ctx.iterate(1)
        assert e.value.code == smoc.ERR_STATE


def test_fixed64_reports_a_diverging_iteration(gpu_ready):
    """On the benchmark distributions the iteration diverges (the reference's
    linear-source term dominates the flux), so within three iterations some
    FIXED64 run sums leave the fixed-point range: the run fails with
    SMOC_ERR_RANGE instead of returning wrapped sums."""
    prob = smoc.generate(200_000, G=100, R=500, F=5, seed=41, sorted_stream=True)
    with smoc.Context(100, 500, 5, tally="fixed64") as ctx:
        ctx.load(prob)
        ctx.set_scattering(scattering(prob, 0.8, 1))
        with pytest.raises(smoc.SmocError) as e:
            ctx.iterate(3)
        assert e.value.code == smoc.ERR_RANGE


@pytest.mark.parametrize("tally", ["det-f64", "det-kahan"])
def test_iterations_with_linear_chunks(gpu_ready, tally):
    """Outer iterations on a region-sorted stream of long tracks in chunks
    (each swept once and corrected per run after the scan), replayed from a
    hipGraph: against the iterated FP64 oracle."""
    prob = smoc.generate(0, G=100, R=60, F=5, seed=43, lognormal_tracks=32, median=1500.0,
                         sigma=1.0, lo=16, hi=8000, sorted_stream=True)
    sc = scattering(prob, 0.8, 2)
    print(f"case outer iteration x3, linear chunks, {tally}: S {prob.S}", flush=True)
    with smoc.Context(100, 60, 5, tally=tally) as ctx:
        ctx.load(prob)
        ctx.set_scattering(sc)
        tm, ums, chg = ctx.iterate(3, chunk=1024, persistent=1, use_graph=1)
        assert tm.used.chunk == 1024
        flux, psi = ctx.results()
        src = ctx.source()
    of, op, osrc = oracle_iterate(prob, sc, prob.fine_source, 3)
    e = (normwise(flux, of), normwise(psi, op), normwise(src, osrc))
    print(f"  normwise fine_flux {e[0]:.2e} psi {e[1]:.2e} source {e[2]:.2e}", flush=True)
    assert max(e) <= TOL and chg > 0
