"""Batched sweeps: tracks of uneven length split into chunks joined by the
affine psi scan, persistent work-queue kernels, hipGraph replay and the
autotuner -- each against whole-track sweeps and the FP64 oracle."""
import numpy as np
import pytest

import pyoracle
import smoc
from conftest import normwise

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ragged():
    # lognormal batches, shorter than the benchmark's so the oracle is quick
    return smoc.generate(0, G=100, R=1000, F=5, seed=21, lognormal_tracks=48, median=600.0,
                         sigma=1.0, lo=16, hi=8000)


@pytest.fixture(scope="module")
def ragged_oracle(ragged):
    return pyoracle.sweep_f64(ragged, pyoracle.EXP_NATIVE, nthreads=8)


def sweep(prob, tally="atomic-f32", storage="f32", exp="accurate", ctx_out=None, **launch):
    with smoc.Context(prob.G, prob.R, prob.F, storage=storage, exp=exp, tally=tally) as ctx:
        ctx.load(prob)
        tm = ctx.run(**launch)
        flux, psi = ctx.results()
    return flux, psi, tm


@pytest.mark.parametrize("chunk", [0, 64, 500, 2048])
@pytest.mark.parametrize("persistent", [0, 1])
@pytest.mark.parametrize("tally", ["atomic-f32", "det-f64"])
def test_chunked_vs_oracle(gpu_ready, ragged, ragged_oracle, chunk, persistent, tally):
    print(f"case chunk={chunk} persistent={persistent} {tally} S={ragged.S}", flush=True)
    flux, psi, tm = sweep(ragged, tally, chunk=chunk, persistent=persistent)
    of, op = ragged_oracle
    ef = abs(flux.sum() - of.sum()) / np.abs(of).sum()
    print(f"  l1-scaled fine_flux {ef:.2e}, normwise {normwise(flux, of):.2e} / "
          f"{normwise(psi, op):.2e}; used chunk {tm.used.chunk}", flush=True)
    assert tm.used.chunk == chunk
    assert ef < 1e-5 and normwise(flux, of) < 1e-5 and normwise(psi, op) < 1e-5


def test_chunking_changes_only_rounding(gpu_ready, ragged):
    _, whole, _ = sweep(ragged, "det-f64")
    _, split, _ = sweep(ragged, "det-f64", chunk=100)
    assert normwise(split, whole) < 2e-6


def test_deterministic_chunked_bit_exact(gpu_ready, ragged):
    with smoc.Context(ragged.G, ragged.R, ragged.F, tally="det-f64") as ctx:
        ctx.load(ragged)
        ctx.run(chunk=256, persistent=1)
        f0, p0 = ctx.results()
        for _ in range(3):
            ctx.run(chunk=256, persistent=1)
            f1, p1 = ctx.results()
            assert np.array_equal(f0, f1) and np.array_equal(p0, p1)


@pytest.mark.parametrize("launch", [{"chunk": 300, "persistent": 1}, {"chunk": 0},
                                    {"chunk": 128, "persistent": 0, "groups_per_lane": 1}])
def test_graph_replays_match_eager(gpu_ready, ragged, launch):
    # every replay is checked against the eager run, with eager work and
    # synchronisation between replays
    with smoc.Context(ragged.G, ragged.R, ragged.F, tally="det-f64") as ctx:
        ctx.load(ragged)
        ctx.run(**launch)
        fe, pe = ctx.results()
        for _ in range(4):
            ctx.run(use_graph=1, **launch)
            fg, pg = ctx.results()
            assert np.array_equal(fg, fe) and np.array_equal(pg, pe)
            ctx.run(**launch)
            f2, p2 = ctx.results()
            assert np.array_equal(f2, fe) and np.array_equal(p2, pe)


def test_autotune_picks_valid_launch(gpu_ready, ragged, ragged_oracle):
    with smoc.Context(ragged.G, ragged.R, ragged.F, exp="accurate", tally="atomic-f32") as ctx:
        ctx.load(ragged)
        tm = ctx.run(autotune=1)
        flux, psi = ctx.results()
        u = tm.used
        print(f"  tuned: gpl {u.groups_per_lane} wpb {u.waves_per_block} chunk {u.chunk} "
              f"persistent {u.persistent}", flush=True)
        assert u.groups_per_lane in (1, 2) and 1 <= u.waves_per_block <= 16
        tm2 = ctx.run(autotune=1)  # cached choice
        assert (tm2.used.chunk, tm2.used.persistent) == (u.chunk, u.persistent)
    of, op = ragged_oracle
    assert normwise(flux, of) < 1e-5 and normwise(psi, op) < 1e-5


@pytest.fixture(scope="module")
def long_runs():
    # lognormal batches sorted by (region, interval) over few regions: runs of
    # a few hundred segments, many of them crossing a split point (a stream
    # index that is a multiple of 1024, where a run also ends)
    return smoc.generate(0, G=100, R=40, F=5, seed=24, lognormal_tracks=48, median=600.0,
                         sigma=1.0, lo=16, hi=8000, sorted_stream=True)


@pytest.fixture(scope="module")
def long_runs_oracle(long_runs):
    return pyoracle.sweep_f64(long_runs, pyoracle.EXP_NATIVE, nthreads=8)


@pytest.mark.parametrize("chunk", [0, 1024, 4096])
@pytest.mark.parametrize("persistent", [0, 1])
@pytest.mark.parametrize("tally", ["det-f64", "det-kahan", "fixed64", "atomic-f32",
                                   "atomic-f64"])
def test_sorted_chunked_vs_oracle(gpu_ready, long_runs, long_runs_oracle, chunk, persistent,
                                  tally):
    """Region-sorted stream, tracks cut at split points (the chunk length is a
    multiple of 1024, so every chunk begins a run): linear chunks for the
    deterministic and atomic tallies, the transfer pass for FIXED64; against
    the FP64 oracle."""
    print(f"case sorted chunk={chunk} persistent={persistent} {tally} S={long_runs.S}",
          flush=True)
    flux, psi, tm = sweep(long_runs, tally, chunk=chunk, persistent=persistent)
    of, op = long_runs_oracle
    print(f"  normwise {normwise(flux, of):.2e} / {normwise(psi, op):.2e}", flush=True)
    assert tm.used.chunk == chunk
    assert normwise(flux, of) < 1e-5 and normwise(psi, op) < 1e-5


@pytest.mark.parametrize("tally", ["det-f64", "fixed64"])
def test_sorted_chunked_bit_exact_across_shapes(gpu_ready, long_runs, tally):
    """For one chunk length every launch shape, the persistent kernel and a
    graph replay give the same bits; chunked and whole-track sweeps differ at
    rounding level only (chunks start from the psi of the transfer scan)."""
    print(f"case sorted chunked shapes {tally}", flush=True)
    with smoc.Context(long_runs.G, long_runs.R, long_runs.F, tally=tally) as ctx:
        ctx.load(long_runs)
        ref = None
        for launch in ({"chunk": 2048},
                       {"chunk": 2048, "persistent": 1, "waves_per_block": 4},
                       {"chunk": 2048, "groups_per_lane": 1, "waves_per_block": 2,
                        "reduce_waves": 8},
                       {"chunk": 2048, "persistent": 1, "use_graph": 1},
                       {"chunk": 2048, "persistent": 1, "use_graph": 1}):
            tm = ctx.run(**launch)
            assert tm.used.chunk == 2048
            out = ctx.results()
            if ref is None:
                ref = out
            assert np.array_equal(out[0], ref[0]) and np.array_equal(out[1], ref[1]), launch
        ctx.run()
        whole = ctx.results()
    assert normwise(ref[0], whole[0]) < 2e-6 and normwise(ref[1], whole[1]) < 2e-6


@pytest.mark.parametrize("tally", ["atomic-f32", "atomic-f64"])
def test_sorted_atomic_chunk_schemes(gpu_ready, long_runs, long_runs_oracle, tally):
    """Atomic tallies on a region-sorted stream: chunks of a multiple of 1024
    are linear, other lengths take the transfer pass, and a segment_flush
    launch at the same chunk length gets a plan of its own (its runs are
    single segments, so no linear chunks) -- each against the FP64 oracle, in
    an order that switches the plan back and forth."""
    of, op = long_runs_oracle
    with smoc.Context(long_runs.G, long_runs.R, long_runs.F, tally=tally) as ctx:
        ctx.load(long_runs)
        for launch in ({"chunk": 2048, "persistent": 1}, {"chunk": 1000},
                       {"chunk": 2048, "segment_flush": 1}, {"chunk": 2048},
                       {"chunk": 2048, "segment_flush": 1, "use_graph": 1},
                       {"chunk": 2048, "use_graph": 1}):
            print(f"case sorted {tally} {launch}", flush=True)
            tm = ctx.run(**launch)
            assert tm.used.chunk == launch["chunk"]
            flux, psi = ctx.results()
            e = (normwise(flux, of), normwise(psi, op))
            print(f"  normwise {e[0]:.2e} / {e[1]:.2e}", flush=True)
            assert max(e) < 1e-5, launch


def test_sorted_chunk_lengths_refused(gpu_ready, long_runs):
    """On a region-sorted stream the deterministic and FIXED64 tallies take
    chunk lengths that are multiples of 1024 only (others would cut runs, so
    results would depend on the chunk length); atomic tallies take any."""
    for tally in ("det-f64", "det-kahan", "fixed64"):
        with smoc.Context(long_runs.G, long_runs.R, long_runs.F, tally=tally) as ctx:
            ctx.load(long_runs)
            for bad in (100, 1000, 3000):
                with pytest.raises(smoc.SmocError) as e:
                    ctx.run(chunk=bad)
                assert e.value.code == smoc.ERR_ARG and "multiple of 1024" in str(e.value)
            assert ctx.run(chunk=3072).used.chunk == 3072
            ctx.run()  # whole tracks
    with smoc.Context(long_runs.G, long_runs.R, long_runs.F, tally="atomic-f32") as ctx:
        ctx.load(long_runs)
        assert ctx.run(chunk=100).used.chunk == 100


@pytest.mark.parametrize("gpl", [1, 2])
@pytest.mark.parametrize("wpb", [1, 4])
@pytest.mark.parametrize("chunk", [0, 1024])
@pytest.mark.parametrize("tally", ["atomic-f32", "det-f64"])
def test_launch_shapes_vs_oracle(gpu_ready, ragged, ragged_oracle, gpl, wpb, chunk, tally):
    flux, psi, tm = sweep(ragged, tally, groups_per_lane=gpl, waves_per_block=wpb, chunk=chunk)
    of, op = ragged_oracle
    ef, ep = normwise(flux, of), normwise(psi, op)
    bad = np.unravel_index(np.argmax(np.abs(flux - of)), flux.shape)
    print(f"  gpl {gpl} wpb {wpb} chunk {chunk} {tally}: normwise {ef:.2e} / {ep:.2e}; "
          f"worst flux entry {bad}", flush=True)
    assert ef < 1e-5 and ep < 1e-5


@pytest.mark.parametrize("storage", ["f32", "bf16"])
def test_fixed64_chunked_graph(gpu_ready, ragged, ragged_oracle, storage):
    """FIXED64 with chunked tracks, the persistent kernel and graph replays:
    bit-exact run to run (integer atomics), at the oracle bar."""
    print(f"case fixed64 chunked {storage}", flush=True)
    with smoc.Context(ragged.G, ragged.R, ragged.F, storage=storage, tally="fixed64") as ctx:
        ctx.load(ragged)
        ctx.run(chunk=512, persistent=1, use_graph=1)
        f1, p1 = ctx.results()
        ctx.run(chunk=512, persistent=1, use_graph=1)
        f2, p2 = ctx.results()
    assert np.array_equal(f1, f2) and np.array_equal(p1, p2)
    from conftest import quantized
    of, op = (ragged_oracle if storage == "f32" else
              pyoracle.sweep_f64(quantized(ragged, storage), pyoracle.EXP_NATIVE, nthreads=8))
    assert normwise(f1, of) < 1e-5 and normwise(p1, op) < 1e-5


@pytest.mark.parametrize("sorted_stream", [False, True])
def test_persistent_batched_pulls(gpu_ready, sorted_stream):
    """Enough short tracks that a persistent wave takes several consecutive
    items per update of the work counter (400,000 items over at most 8192
    resident waves at 16 groups: batches of 3): the deterministic tally is
    bit-identical to the grid launch's, and both match the FP64 oracle."""
    prob = smoc.generate(1_200_000, G=16, R=2000, F=5, seed=23, track_len=3,
                         sorted_stream=sorted_stream)
    assert prob.T == 400_000
    print(f"case persistent batched pulls, sorted {sorted_stream}", flush=True)
    out = []
    with smoc.Context(16, 2000, 5, tally="det-f64") as ctx:
        ctx.load(prob)
        for pers in (0, 1):
            ctx.run(persistent=pers)
            out.append(ctx.results())
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    of, op = pyoracle.sweep_f64(prob, pyoracle.EXP_NATIVE, nthreads=8)
    e = (normwise(out[1][0], of), normwise(out[1][1], op))
    print(f"  normwise {e[0]:.2e} / {e[1]:.2e}", flush=True)
    assert max(e) < 1e-5


@pytest.mark.parametrize("G", [37, 129, 200])
@pytest.mark.parametrize("tally", ["det-f64", "atomic-f32"])
def test_sorted_linear_chunks_group_counts(gpu_ready, G, tally):
    """Linear chunks with group counts that leave lanes of padding or need
    several group slices per track (the per-run W stores and the correction
    skip groups past G), at 1 and 2 groups per lane, against the FP64 oracle."""
    prob = smoc.generate(0, G=G, R=30, F=5, seed=G, lognormal_tracks=16, median=1500.0,
                         sigma=1.0, lo=16, hi=5000, sorted_stream=True)
    of, op = pyoracle.sweep_f64(prob, pyoracle.EXP_NATIVE, nthreads=8)
    for gpl in (1, 2):
        print(f"case linear chunks G={G} {tally} gpl {gpl}", flush=True)
        flux, psi, tm = sweep(prob, tally, chunk=1024, persistent=1, groups_per_lane=gpl)
        assert tm.used.chunk == 1024
        e = (normwise(flux, of), normwise(psi, op))
        print(f"  normwise {e[0]:.2e} / {e[1]:.2e}", flush=True)
        assert max(e) < 1e-5


@pytest.mark.parametrize("tally", ["det-f64", "atomic-f32"])
def test_sorted_linear_chunks_with_empty_tracks(gpu_ready, long_runs, tally):
    """Empty tracks between long ones: an empty track's single empty chunk
    holds the run id where the previous track's runs end, which bounds the
    last chunk's runs in the per-run correction; results against the FP64
    oracle, psi_out of the empty tracks equal to their psi_in."""
    import copy
    prob = copy.copy(long_runs)
    off = long_runs.track_off
    # an empty track before the first, after every fifth, and at the end
    cuts = [0] + [int(x) for x in off[1:-1:5]] + [int(off[-1])]
    prob.track_off = np.sort(np.concatenate([off, np.asarray(cuts, np.int64)]))
    T = prob.track_off.size - 1
    prob.psi_in = np.random.default_rng(4).random((T, prob.G), dtype=np.float32)
    of, op = pyoracle.sweep_f64(prob, pyoracle.EXP_NATIVE, nthreads=8)
    print(f"case linear chunks with empty tracks {tally}: T {T}", flush=True)
    flux, psi, tm = sweep(prob, tally, chunk=1024, persistent=1)
    assert tm.used.chunk == 1024
    e = (normwise(flux, of), normwise(psi, op))
    print(f"  normwise {e[0]:.2e} / {e[1]:.2e}", flush=True)
    assert max(e) < 1e-5
    empty = np.flatnonzero(np.diff(prob.track_off) == 0)
    assert empty.size >= 3 and np.array_equal(psi[empty], prob.psi_in[empty])


@pytest.mark.parametrize("storage", ["bf16", "f16"])
@pytest.mark.parametrize("tally", ["det-f64", "atomic-f32"])
def test_sorted_linear_chunks_16bit_storage(gpu_ready, long_runs, storage, tally):
    """Linear chunks with 16-bit sigT / fine_source, against the FP64 oracle
    on the dequantised inputs (the arithmetic error)."""
    from conftest import quantized
    of, op = pyoracle.sweep_f64(quantized(long_runs, storage), pyoracle.EXP_NATIVE, nthreads=8)
    print(f"case linear chunks {storage} {tally}", flush=True)
    flux, psi, tm = sweep(long_runs, tally, storage=storage, chunk=2048, persistent=1)
    assert tm.used.chunk == 2048
    e = (normwise(flux, of), normwise(psi, op))
    print(f"  normwise {e[0]:.2e} / {e[1]:.2e}", flush=True)
    assert max(e) < 1e-5


@pytest.mark.parametrize("F", [2, 3, 8])
def test_sorted_linear_chunks_fine_intervals(gpu_ready, F):
    """Linear chunks for F = 2 (every row a first or last interval, so every
    run takes the linear fit), 3 and 8, against the FP64 oracle."""
    prob = smoc.generate(0, G=100, R=40, F=F, seed=60 + F, lognormal_tracks=24, median=1500.0,
                         sigma=1.0, lo=16, hi=6000, sorted_stream=True)
    of, op = pyoracle.sweep_f64(prob, pyoracle.EXP_NATIVE, nthreads=8)
    print(f"case linear chunks F={F}", flush=True)
    flux, psi, tm = sweep(prob, "det-f64", chunk=1024, persistent=1)
    assert tm.used.chunk == 1024
    e = (normwise(flux, of), normwise(psi, op))
    print(f"  normwise {e[0]:.2e} / {e[1]:.2e}", flush=True)
    assert max(e) < 1e-5


@pytest.mark.parametrize("tally", ["det-f64", "atomic-f32", "atomic-f64"])
def test_sorted_linear_chunks_with_init(gpu_ready, long_runs, tally):
    """Linear chunks accumulating onto a fine_flux_init: the deterministic
    reduction adds it, the atomic tallies start from it and the per-run
    correction adds onto it; against the FP64 oracle with the same init."""
    init = np.random.default_rng(8).random((long_runs.R, long_runs.F, long_runs.G),
                                           dtype=np.float32)
    of, op = pyoracle.sweep_f64(long_runs, pyoracle.EXP_NATIVE, nthreads=8, fine_flux_init=init)
    print(f"case linear chunks with init {tally}", flush=True)
    with smoc.Context(long_runs.G, long_runs.R, long_runs.F, tally=tally) as ctx:
        ctx.load(long_runs, init)
        tm = ctx.run(chunk=1024, persistent=1, use_graph=1)
        assert tm.used.chunk == 1024
        flux, psi = ctx.results()
    e = (normwise(flux, of), normwise(psi, op))
    print(f"  normwise {e[0]:.2e} / {e[1]:.2e}", flush=True)
    assert max(e) < 1e-5
