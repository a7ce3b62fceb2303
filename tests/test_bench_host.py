"""Host-side pieces of bench.py (no GPU): the byte model covers every tally,
the committed profile summaries it reads parse, and the CPU baseline leg
returns well-formed figures on a small sample."""
import importlib.util
import json
import os
import subprocess
import sys

import pytest

import smoc
import pyoracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_module", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_alg_bytes_cover_every_tally_and_storage():
    b = _bench()
    for tally in smoc.TALLY:
        for storage in smoc.STORAGE:
            v = b.alg_bytes_per_sg(storage, tally, 100, 100)
            assert v > 0
    # FP32 storage, FP32 atomic tally: sigT + 3 source values + tally read and write
    # + 12 B of stream per segment over 100 groups + 8 B of psi per track-group over 100
    assert b.alg_bytes_per_sg("f32", "atomic-f32", 100, 100) == pytest.approx(
        4 + 12 + 8 + 0.12 + 0.08)


def test_committed_profiles_parse():
    b = _bench()
    valu = b.mixed_valu_instructions()
    assert valu is not None and 1e9 < valu < 1e10
    traffic = b.measured_traffic()
    assert traffic is not None and traffic > 1e10


def test_cpu_leg_small_sample():
    # in a child process: the reference build is compiled with -Ofast, and loading
    # it switches the process to flush-to-zero arithmetic
    code = (
        "import importlib.util, json, os, sys\n"
        f"sys.path[:0] = [{os.path.join(ROOT, 'simplemoc-kernel_amd')!r}, "
        f"{os.path.join(ROOT, 'oracle')!r}, {os.path.join(ROOT, 'tests')!r}]\n"
        "import smoc\n"
        f"spec = importlib.util.spec_from_file_location('b', {os.path.join(ROOT, 'bench.py')!r})\n"
        "b = importlib.util.module_from_spec(spec); spec.loader.exec_module(b)\n"
        "prob = smoc.generate(60_000, G=32, R=200, F=5, seed=3)\n"
        "base, acc = b.cpu_leg(prob, 40_000, 10_000, None, None, False)\n"
        "print(json.dumps({'base': base, 'acc': acc}))\n")
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True,
                         timeout=300)
    assert out.returncode == 0, out.stderr
    res = json.loads(out.stdout.strip().splitlines()[-1])
    base = res["base"]
    assert res["acc"] is None
    assert base["value"] > 0 and base["cores"] >= 1
    assert base["unit"] == "segment-group attenuations/s"
    assert base["single_thread"]["value"] > 0 and base["single_thread"]["cores"] == 1
    if pyoracle.ref_available("exp"):
        assert base["kind"] == "reference"
        port = base["same_stream_port"]
        assert port["kind"] == "port" and port["value"] > 0
    else:
        assert base["kind"] == "port"
