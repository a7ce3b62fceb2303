"""Measurement probes on the GPU: they run and return a time.  The copy probe
calibrates the traffic counters; the tally probe times the unsorted sweep's
atomic flushes alone, which the sweep itself can only be slower than."""
import math

import pytest

import smoc

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("variant", range(4))
@pytest.mark.parametrize("nbytes", [64 << 20, (64 << 20) + 48])
def test_copy_probe_runs(gpu_ready, variant, nbytes):
    ms = smoc.probe_copy(nbytes, reps=3, variant=variant)
    print(f"  copy variant {variant}, {nbytes} B: {ms:.4f} ms, {2 * nbytes / ms / 1e6:.0f} GB/s",
          flush=True)
    assert math.isfinite(ms) and ms > 0


@pytest.mark.parametrize("G", [1, 64, 100, 129])
def test_tally_probe_runs(gpu_ready, G):
    print(f"case tally probe G={G}", flush=True)
    ms = smoc.probe_tally(1_000_003, 6750 * 5, G, reps=2)
    assert math.isfinite(ms) and ms > 0
    assert smoc.probe_tally(0, 10, G, reps=1) >= 0.0


def test_sweep_is_not_faster_than_its_tally_floor(gpu_ready):
    prob = smoc.generate(2_000_000, G=100, R=6750, F=5, seed=3)
    with smoc.Context(100, 6750, 5) as ctx:
        ctx.load(prob)
        ctx.run()
        sweep = min(ctx.run().sweep_ms for _ in range(3))
    floor = smoc.probe_tally(prob.S, 6750 * 5, 100, reps=3)
    print(f"  sweep {sweep:.3f} ms, atomic floor {floor:.3f} ms", flush=True)
    assert sweep > 0.8 * floor
