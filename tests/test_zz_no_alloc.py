"""smoc_run_noalloc: the sweep with its scratch (persistent work counters and
the chunk plan) in the library's fixed per-GPU array, checked through the C
ABI against smoc_run on the same context -- bit-exact with the deterministic
tally, psi_out bit-exact and fine_flux at rounding level with the atomic one --
on two calls in a row, the second with a smaller input, and with two contexts
taking turns on the one array."""
import numpy as np
import pytest

import smoc
from conftest import normwise

pytestmark = pytest.mark.gpu

LAUNCHES = [{"chunk": 256, "persistent": 1}, {"chunk": 700}, {"persistent": 1}, {}]


def big_problem():
    return smoc.generate(0, G=100, R=900, F=5, seed=41, lognormal_tracks=64, median=900.0,
                         sigma=1.0, lo=16, hi=9000)


def small_problem():
    return smoc.generate(0, G=100, R=900, F=5, seed=42, lognormal_tracks=17, median=400.0,
                         sigma=1.0, lo=16, hi=3000)


def check_pair(ctx, tally, launch):
    ref_tm = ctx.run(**launch)
    fr, pr = ctx.results()
    tm = ctx.run_noalloc(**launch)
    fn, pn = ctx.results()
    assert tm.used.chunk == ref_tm.used.chunk and tm.used.persistent == ref_tm.used.persistent
    # every chunk computes the same operations whatever the work order: psi_out
    # is bit-exact; the deterministic tally too, the atomic one to rounding
    assert np.array_equal(pn, pr)
    if tally == "atomic-f32":
        assert normwise(fn, fr) < 1e-6
    else:
        assert np.array_equal(fn, fr)
    return tm


@pytest.mark.parametrize("tally", ["det-f64", "atomic-f32", "fixed64"])
def test_noalloc_matches_run_then_smaller_input(gpu_ready, tally):
    big, small = big_problem(), small_problem()
    with smoc.Context(100, 900, 5, tally=tally) as ctx:
        ctx.load(big)
        for launch in LAUNCHES:
            print(f"case noalloc {tally} big {launch}", flush=True)
            check_pair(ctx, tally, launch)
        # a second call with a smaller input: the plan shrinks inside the array
        ctx.set_segments(small.len, small.region, small.z, small.track_off, small.psi_in)
        for launch in LAUNCHES:
            print(f"case noalloc {tally} small {launch}", flush=True)
            check_pair(ctx, tally, launch)


@pytest.mark.parametrize("tally", ["det-f64", "fixed64", "atomic-f32"])
def test_noalloc_sorted_chunked(gpu_ready, tally):
    """Region-sorted stream in chunks of a multiple of 1024: the plan built on
    the device takes each chunk's first run id from the run tables as the host
    plan does; then a smaller input."""
    sizes = ((40, 900.0, 9000), (13, 900.0, 4000))
    probs = [smoc.generate(0, G=100, R=40, F=5, seed=44 + i, lognormal_tracks=n, median=m,
                           sigma=1.0, lo=16, hi=h, sorted_stream=True)
             for i, (n, m, h) in enumerate(sizes)]
    with smoc.Context(100, 40, 5, tally=tally) as ctx:
        ctx.set_regions(probs[0].sigT, probs[0].fine_source)
        for i, prob in enumerate(probs):
            ctx.set_segments(prob.len, prob.region, prob.z, prob.track_off, prob.psi_in)
            for launch in ({"chunk": 1024, "persistent": 1}, {"chunk": 2048}):
                print(f"case noalloc sorted {tally} input {i} {launch}", flush=True)
                tm = check_pair(ctx, tally, launch)
                assert tm.used.chunk == launch["chunk"]


def test_noalloc_two_contexts_take_turns(gpu_ready):
    """Two contexts on one GPU share the array; calls one after another do
    not disturb each other (each rebuilds its plan in the array)."""
    big, small = big_problem(), small_problem()
    launch = {"chunk": 300, "persistent": 1}
    with smoc.Context(100, 900, 5, tally="det-f64") as a, \
            smoc.Context(100, 900, 5, tally="det-f64") as b:
        a.load(big)
        b.load(small)
        a.run(**launch)
        ref_a = a.results()
        b.run(**launch)
        ref_b = b.results()
        for _ in range(2):
            print("case noalloc two contexts", flush=True)
            a.run_noalloc(**launch)
            b.run_noalloc(**launch)
            fb, pb = b.results()
            fa, pa = a.results()
            assert np.array_equal(fa, ref_a[0]) and np.array_equal(pa, ref_a[1])
            assert np.array_equal(fb, ref_b[0]) and np.array_equal(pb, ref_b[1])


def test_noalloc_refusals(gpu_ready):
    prob = smoc.generate(1_000_000, G=100, R=900, F=5, seed=43, track_len=1000)
    with smoc.Context(100, 900, 5) as ctx:
        with pytest.raises(smoc.SmocError) as e:
            ctx.run_noalloc()
        assert e.value.code == smoc.ERR_STATE
        ctx.load(prob)
        for bad in ({"autotune": 1}, {"use_graph": 1}):
            with pytest.raises(smoc.SmocError) as e:
                ctx.run_noalloc(**bad)
            assert e.value.code == smoc.ERR_ARG
        # 1000 tracks of 1000 segments in chunks of 16: 63,000 chunks need
        # 63,000 x 128 x 12 bytes of psi and transfers, beyond the 64 MiB array
        with pytest.raises(smoc.SmocError) as e:
            ctx.run_noalloc(chunk=16)
        assert e.value.code == smoc.ERR_ARG and "scratch" in str(e.value)
        # a plan that fits still runs after the refusal
        ctx.run_noalloc(chunk=250, persistent=1)
        f1, p1 = ctx.results()
        ctx.run(chunk=250, persistent=1)
        f0, p0 = ctx.results()
        assert np.array_equal(p1, p0) and normwise(f1, f0) < 1e-6
