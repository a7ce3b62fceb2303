"""Fixture files: problems and results written and read through the C ABI.

A problem file carries the complete input of one attenuation sweep (region
tables, segment stream, tracks, incoming psi); a results file carries
fine_flux and psi_out with their checksums.  Both are hashed (FNV-1a over the
header and the arrays), so parity can move between hosts as data: write the
problem and the FP64 oracle's results on one machine, then run the GPU sweep
on another and compare.  These tests run on the CPU.
"""
import os
import subprocess

import numpy as np
import pytest

import pyoracle
import smoc
from conftest import normwise


def small(S=3000, G=24, R=40, F=5, **kw):
    return smoc.generate(S, G=G, R=R, F=F, seed=7, **kw)


def same_problem(a, b):
    for name in ("sigT", "fine_source", "len", "region", "z", "track_off", "psi_in"):
        x, y = getattr(a, name), getattr(b, name)
        assert x.dtype == y.dtype and x.shape == y.shape, name
        assert np.array_equal(x.view(np.uint8), y.view(np.uint8)), name  # bit-exact
    assert (a.G, a.R, a.F, a.seed) == (b.G, b.R, b.F, b.seed)


@pytest.mark.parametrize("kw", [{}, {"lognormal_tracks": 12}, {"track_len": 37}])
def test_problem_round_trip(tmp_path, kw):
    prob = small(**kw) if "lognormal_tracks" not in kw else small(S=0, **kw)
    path = str(tmp_path / "p.smocprb")
    smoc.save_problem(path, prob)
    same_problem(prob, smoc.load_problem(path))


def test_problem_sorted_flag_and_empty(tmp_path):
    prob = small(S=0)
    assert prob.S == 0 and prob.T == 0
    path = str(tmp_path / "empty.smocprb")
    smoc.save_problem(path, prob, sorted_stream=True)
    info = smoc.ProblemInfo()
    assert smoc.lib().smoc_problem_read_info(path.encode(), info) == 0
    assert (info.S, info.T, info.flags) == (0, 0, 1)
    same_problem(prob, smoc.load_problem(path))


def test_results_round_trip_and_compare(tmp_path):
    prob = small()
    flux, psi = pyoracle.sweep_f64(prob, pyoracle.EXP_NATIVE, nthreads=2)
    psi32 = psi.astype(np.float32)
    path = str(tmp_path / "r.smocres")
    smoc.save_results(path, flux, psi32, exp="native", tally="det-f64")
    info, f2, p2 = smoc.load_results(path)
    assert np.array_equal(f2, flux) and np.array_equal(p2, psi32)
    assert (info.G, info.R, info.F, info.T) == (prob.G, prob.R, prob.F, prob.T)
    assert (info.exp_mode, info.tally) == (smoc.EXP_NATIVE, smoc.TALLY_DET_F64)
    # the header checksums are the plain sums of the stored arrays
    assert info.fine_flux_checksum == smoc.lib().smoc_checksum_f64(flux.ravel(), flux.size)
    assert info.state_flux_checksum == smoc.lib().smoc_checksum_f32(psi32.ravel(), psi32.size)
    e = smoc.compare_results(path, flux, psi32)
    assert all(v == 0.0 for v in e.values()), e
    # a known perturbation shows up as its own normwise size
    bumped = flux * (1 + 3e-4)
    e = smoc.compare_results(path, bumped, psi32)
    assert abs(e["fine_normwise"] - 3e-4) < 1e-9 and abs(e["fine_checksum_rel"] - 3e-4) < 1e-9
    assert e["state_normwise"] == 0.0
    psi_b = psi32.copy()
    psi_b[0, 0] += 1.0
    e = smoc.compare_results(path, flux, psi_b)
    assert abs(e["state_normwise"] - normwise(psi_b, psi32)) < 1e-12


def _write_problem(tmp_path):
    path = str(tmp_path / "c.smocprb")
    smoc.save_problem(path, small(S=500))
    return path


def _expect_error(path, text):
    with pytest.raises(smoc.SmocError) as e:
        smoc.load_problem(path)
    assert e.value.code == smoc.ERR_ARG and text in str(e.value), str(e.value)


def test_corrupt_array_byte(tmp_path):
    path = _write_problem(tmp_path)
    b = bytearray(open(path, "rb").read())
    b[len(b) // 2] ^= 0x10
    open(path, "wb").write(bytes(b))
    _expect_error(path, "hash mismatch")


def test_corrupt_header_field(tmp_path):
    path = _write_problem(tmp_path)
    b = bytearray(open(path, "rb").read())
    b[40] ^= 0x01  # low byte of the seed: sizes still agree, the hash does not
    open(path, "wb").write(bytes(b))
    _expect_error(path, "hash mismatch")


def test_truncated_and_trailing(tmp_path):
    path = _write_problem(tmp_path)
    b = open(path, "rb").read()
    open(path, "wb").write(b[:-5])
    _expect_error(path, "truncated")
    open(path, "wb").write(b + b"\0")
    _expect_error(path, "trailing bytes")


def test_wrong_kind_and_missing(tmp_path):
    prob = small(S=200)
    rpath = str(tmp_path / "x.smocres")
    smoc.save_results(rpath, np.zeros((prob.R, prob.F, prob.G)),
                      np.zeros((prob.T, prob.G), np.float32))
    _expect_error(rpath, "not a smoc problem file")
    _expect_error(str(tmp_path / "nope"), "cannot open")
    with pytest.raises(smoc.SmocError):
        smoc.load_results(_write_problem(tmp_path))


def test_write_rejects_inconsistent_tracks(tmp_path):
    prob = small(S=300)
    prob.track_off = prob.track_off.copy()
    prob.track_off[-1] -= 1
    with pytest.raises(smoc.SmocError) as e:
        smoc.save_problem(str(tmp_path / "bad"), prob)
    assert e.value.code == smoc.ERR_ARG


def test_driver_writes_the_generated_problem(tmp_path):
    """--save-problem runs before any device call, so it works on the CPU; the
    file holds exactly what smoc.generate makes for the same options."""
    path = str(tmp_path / "drv.smocprb")
    r = subprocess.run([smoc.DRIVER_PATH, "-s", "4000", "-e", "16", "--regions", "90",
                        "--track-len", "50", "--seed", "11", "--save-problem", path],
                       capture_output=True, text=True, timeout=120)
    assert os.path.exists(path), r.stdout + r.stderr
    assert r.returncode in (0, 2)  # 2: no device here, after the file was written
    prob = smoc.generate(4000, G=16, R=90, F=5, seed=11, track_len=50)
    same_problem(prob, smoc.load_problem(path))


def test_header_claiming_huge_arrays_is_rejected_before_reading(tmp_path):
    """A header whose counts do not match the file size is refused by the
    info call itself, before a caller sizes buffers from it."""
    path = _write_problem(tmp_path)
    b = bytearray(open(path, "rb").read())
    b[32:40] = (1 << 39).to_bytes(8, "little")  # T: 2^39 tracks
    open(path, "wb").write(bytes(b))
    info = smoc.ProblemInfo()
    assert smoc.lib().smoc_problem_read_info(path.encode(), info) == smoc.ERR_ARG
    assert "truncated" in smoc.lib().smoc_last_error().decode()
