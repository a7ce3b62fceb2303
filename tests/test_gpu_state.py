"""Context state across calls: a captured hipGraph never replays on buffers
that a later call freed or reallocated (new region data, a new segment stream,
a chunk plan for another chunk length), and the segment_flush switch."""
import numpy as np
import pytest

import pyoracle
import smoc
from conftest import normwise

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ragged():
    return smoc.generate(0, G=100, R=800, F=5, seed=31, lognormal_tracks=40, median=500.0,
                         sigma=1.0, lo=16, hi=6000)


def eager(prob, tally, init=None, **launch):
    with smoc.Context(prob.G, prob.R, prob.F, tally=tally) as ctx:
        ctx.load(prob, fine_flux_init=init)
        ctx.run(**launch)
        return ctx.results()


@pytest.mark.parametrize("launch", [{"chunk": 0}, {"chunk": 300, "persistent": 1}])
def test_graph_after_set_regions_with_init(gpu_ready, ragged, launch):
    """graph run, then set_regions with a fine_flux_init (every region buffer
    reallocated), then a graph run: equal to the eager run of the new data."""
    print(f"case graph / set_regions(init) / graph, {launch}", flush=True)
    rng = np.random.default_rng(3)
    init = rng.random((ragged.R, ragged.F, ragged.G), dtype=np.float32)
    sig2 = (ragged.sigT * np.float32(0.5)).astype(np.float32)
    with smoc.Context(ragged.G, ragged.R, ragged.F, tally="det-f64") as ctx:
        ctx.load(ragged)
        ctx.run(use_graph=1, **launch)
        ctx.set_regions(sig2, ragged.fine_source, init)
        ctx.run(use_graph=1, **launch)
        fg, pg = ctx.results()
    import copy
    p2 = copy.copy(ragged)
    p2.sigT = sig2
    fe, pe = eager(p2, "det-f64", init=init, **launch)
    assert np.array_equal(fg, fe) and np.array_equal(pg, pe)


def test_graph_plan_then_other_chunk_then_graph(gpu_ready, ragged):
    """run(chunk=512, graph), an eager run with another chunk length (the plan
    buffers are reallocated), run(chunk=512, graph) again: the replayed graph
    must not use the freed plan; equal to the eager chunk-512 run."""
    print("case graph chunk 512 / eager chunk 1024 / graph chunk 512", flush=True)
    fe, pe = eager(ragged, "det-f64", chunk=512, persistent=1)
    with smoc.Context(ragged.G, ragged.R, ragged.F, tally="det-f64") as ctx:
        ctx.load(ragged)
        ctx.run(chunk=512, persistent=1, use_graph=1)
        f0, p0 = ctx.results()
        ctx.run(chunk=1024, persistent=1)
        ctx.run(chunk=2048)
        for _ in range(2):
            ctx.run(chunk=512, persistent=1, use_graph=1)
            f1, p1 = ctx.results()
            assert np.array_equal(f1, fe) and np.array_equal(p1, pe)
    assert np.array_equal(f0, fe) and np.array_equal(p0, pe)


def test_graph_after_new_segments(gpu_ready, ragged):
    """a graph captured for one stream is dropped when a smaller stream is
    loaded into the same context."""
    print("case graph / set_segments(smaller) / graph", flush=True)
    small = smoc.generate(30_000, G=100, R=800, F=5, seed=32, track_len=70)
    fe, pe = eager(small, "det-f64")
    with smoc.Context(ragged.G, ragged.R, ragged.F, tally="det-f64") as ctx:
        ctx.load(ragged)
        ctx.run(use_graph=1)
        ctx.set_regions(small.sigT, small.fine_source)
        ctx.set_segments(small.len, small.region, small.z, small.track_off, small.psi_in)
        ctx.run(use_graph=1)
        f1, p1 = ctx.results()
    assert np.array_equal(f1, fe) and np.array_equal(p1, pe)


def test_failed_set_segments_leaves_no_stale_state(gpu_ready, ragged):
    """a rejected stream leaves the context without segments (run refuses)
    only when the upload itself started; argument errors keep the old data."""
    with smoc.Context(ragged.G, ragged.R, ragged.F) as ctx:
        ctx.load(ragged)
        ctx.run()
        bad = ragged.region.copy()
        bad[5] = ragged.R  # out of range: rejected before anything is freed
        with pytest.raises(smoc.SmocError):
            ctx.set_segments(ragged.len, bad, ragged.z, ragged.track_off, ragged.psi_in)
        ctx.run()  # the previous stream is intact
        ctx.results()


@pytest.fixture(scope="module")
def sorted_prob():
    return smoc.generate(400_000, G=100, R=500, F=5, seed=33, sorted_stream=True)


@pytest.fixture(scope="module")
def sorted_oracle(sorted_prob):
    return pyoracle.sweep_f64(sorted_prob, pyoracle.EXP_NATIVE, nthreads=8)


@pytest.mark.parametrize("tally", ["atomic-f32", "atomic-f64", "fixed64"])
def test_segment_flush_on_sorted_stream(gpu_ready, sorted_prob, sorted_oracle, tally):
    """the per-segment loop on a region-sorted stream: same physics, psi
    bit-identical to the run loop's only where the arithmetic is the same
    (it is not: the run loop hoists run constants), both at the oracle bar."""
    print(f"case segment_flush {tally}", flush=True)
    of, op = sorted_oracle
    out = {}
    with smoc.Context(sorted_prob.G, sorted_prob.R, sorted_prob.F, tally=tally) as ctx:
        ctx.load(sorted_prob)
        for sf in (0, 1):
            tm = ctx.run(segment_flush=sf)
            assert tm.used.segment_flush == sf
            out[sf] = ctx.results()
            e = (normwise(out[sf][0], of), normwise(out[sf][1], op))
            print(f"  segment_flush {sf}: {tm.sweep_ms:.3f} ms, normwise {e[0]:.2e} / {e[1]:.2e}",
                  flush=True)
            assert max(e) < 1e-5
        if tally == "fixed64":  # integer tally: bit-exact run to run per-segment as well
            ctx.run(segment_flush=1)
            assert np.array_equal(ctx.results()[0], out[1][0])


def test_segment_flush_refused_for_det_on_sorted(gpu_ready, sorted_prob):
    with smoc.Context(sorted_prob.G, sorted_prob.R, sorted_prob.F, tally="det-f64") as ctx:
        ctx.load(sorted_prob)
        with pytest.raises(smoc.SmocError) as e:
            ctx.run(segment_flush=1)
        assert e.value.code == smoc.ERR_ARG
        ctx.run()
