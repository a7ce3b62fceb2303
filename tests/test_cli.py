"""The smoc driver keeps the reference command line (src/cuda/io.cu:107-162):
-s segments, -e groups, -p segments per block, -d device; unknown or
incomplete options print the usage and exit(1)."""
import os
import subprocess
import sys

import numpy as np
import pytest

import smoc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run(*args, timeout=120):
    return subprocess.run([smoc.DRIVER_PATH, *args], capture_output=True, text=True,
                          timeout=timeout)


def test_unknown_option_prints_usage():
    r = run("-x")
    assert r.returncode == 1
    assert "Usage:" in r.stdout and "-s <segments>" in r.stdout and "-p <segs per block>" in r.stdout


@pytest.mark.parametrize("args", [("-s",), ("-e",), ("-p", "0"), ("--tally", "bogus"),
                                  ("-s", "-5")])
def test_bad_values_exit_1(args):
    assert run(*args).returncode == 1


def test_no_device_is_an_error():
    if smoc.device_count() > 0:
        pytest.skip("a HIP device is present")
    r = run("-s", "1000", "-e", "8")
    assert r.returncode == 2
    assert "no usable HIP device" in r.stderr


def test_defaults_are_the_reference_problem():
    """With no options the driver sets up the reference's default problem
    (src/cpu/init.c:4-24, main.c:18-19): 50,000,000 segments x 128 groups over
    6750 derived source regions.  The summary is printed before the device is
    opened, so it is checked with or without a GPU; with one the run completes."""
    r = run(timeout=600)
    for line in ("Energy Groups:           128", "3D Source Regions:       6750",
                 "Segments:                50,000,000", "Fine Axial Intervals:    5"):
        assert line in r.stdout, (line, r.stdout[-1500:])
    if smoc.device_count() > 0:
        assert r.returncode == 0, r.stderr
        assert "Time per Intersection:" in r.stdout
    else:
        assert r.returncode == 2 and "no usable HIP device" in r.stderr


def test_bench_refuses_multi_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stdout


@pytest.mark.gpu
def test_driver_report_and_checksums_match_api(gpu_ready):
    r = run("-s", "200000", "-e", "100", "-p", "512", "--tally", "det-f64", "--exp", "accurate",
            "--steps", "2")
    assert r.returncode == 0, r.stderr
    out = r.stdout
    for line in ("INPUT SUMMARY", "3D Source Regions:       6750", "Segments:                200,000",
                 "Runtime:", "Time per Intersection:", "fine_flux checksum:",
                 "state_flux checksum:"):
        assert line in out, line
    ck = {l.split(":")[0]: float(l.split(":")[1]) for l in out.splitlines()
          if "checksum:" in l}
    prob = smoc.generate(200_000, G=100, R=6750, F=5, seed=42)
    with smoc.Context(100, 6750, 5, exp="accurate", tally="det-f64") as ctx:
        ctx.load(prob)
        ctx.run()
        flux, psi = ctx.results()
    f, p = smoc.checksums(flux, psi)
    # deterministic tally: the driver and the API produce identical sums
    assert ck["fine_flux checksum"] == float(f"{f:.17g}")
    assert ck["state_flux checksum"] == float(f"{p:.17g}")


@pytest.mark.gpu
def test_driver_checks_against_oracle_results_file(gpu_ready, tmp_path):
    """Parity through files: the problem and the FP64 oracle's results are
    written on the host, the driver loads the problem, sweeps it on the GPU
    and compares with the results file (exit 3 when over --tol)."""
    import pyoracle
    prob = smoc.generate(300_000, G=100, R=6750, F=5, seed=5, track_len=150)
    ppath, rpath = str(tmp_path / "p.smocprb"), str(tmp_path / "r.smocres")
    smoc.save_problem(ppath, prob)
    of, op = pyoracle.sweep_f64(prob, pyoracle.EXP_NATIVE, nthreads=16)
    smoc.save_results(rpath, of, op.astype(np.float32), exp="native", tally="det-f64")
    print("case: driver --load-problem 300k x 100 --check-results (FP64 oracle)", flush=True)
    r = run("--load-problem", ppath, "--check-results", rpath, "--tally", "det-f64")
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr
    assert "Check:                   passed" in r.stdout
    assert "Segments:                300,000" in r.stdout
    # a tolerance below the FP32 rounding of psi_out in the file fails with status 3
    r = run("--load-problem", ppath, "--check-results", rpath, "--tol", "1e-12")
    assert r.returncode == 3 and "FAILED" in r.stdout


@pytest.mark.gpu
def test_driver_traced_stream_and_outer_iterations(gpu_ready):
    """--trace builds the stream by ray tracing; the checksums equal the same
    problem built through the API; --outer runs the source iteration after."""
    geo = smoc.Geometry(12, 10, 6, 1.26, 0.54, 60.0)
    r = run("--trace", "4000", "--geometry", "12,10,6,1.26,0.54,60", "-e", "64",
            "--tally", "det-f64", "--seed", "9", "--outer", "2", "--scatter", "0.3")
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr
    assert "Traced geometry:" in r.stdout and "OUTER ITERATION" in r.stdout
    lines = [l for l in r.stdout.splitlines() if "checksum:" in l]
    drv = float(lines[0].split(":")[1])  # fine_flux of the first sweep
    prob = smoc.traced_problem(geo, 4000, G=64, F=5, seed=9)
    with smoc.Context(64, geo.regions, 5, tally="det-f64") as ctx:
        ctx.load(prob)
        ctx.run()
        f, _ = ctx.results()
    assert drv == smoc.checksums(f, np.zeros((0, 64), np.float32))[0]
