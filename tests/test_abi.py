"""The C-ABI library loads and exports every entry point include/smoc.h
declares; argument errors come back as codes (no GPU compute here)."""
import ctypes
import os
import re

import pytest

import smoc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "smoc.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(smoc_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_expected_surface():
    names = declared()
    for n in ("smoc_create", "smoc_destroy", "smoc_set_regions", "smoc_set_segments",
              "smoc_run", "smoc_get_results", "smoc_strerror", "smoc_last_error"):
        assert n in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(smoc.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python mirror binds exactly that surface
    assert sorted(smoc.SIGNATURES) == declared()


def test_abi_version_and_errors():
    L = smoc.lib()
    assert L.smoc_abi_version() == 2
    for code, word in ((0, b"ok"), (-1, b"invalid"), (-2, b"device"), (-3, b"HIP"),
                       (-4, b"order"), (-5, b"memory"), (-99, b"unknown")):
        assert word in L.smoc_strerror(code)


def test_create_rejects_bad_params_before_touching_a_device():
    for kw in ({"G": 0, "R": 10, "F": 5}, {"G": 10, "R": 0, "F": 5}, {"G": 10, "R": 10, "F": 1},
               {"G": 65537, "R": 10, "F": 5},
               {"G": 10, "R": 1 << 27, "F": 4}):  # R * F = 2^29: row codes need R * F < 2^29
        with pytest.raises(smoc.SmocError) as e:
            smoc.Context(kw["G"], kw["R"], kw["F"])
        assert e.value.code == smoc.ERR_ARG
    with pytest.raises(smoc.SmocError) as e:
        smoc.Context(10, 10, 5, storage=7)
    assert e.value.code == smoc.ERR_ARG


def test_create_without_device_fails_loudly():
    if smoc.device_count() > 0:
        pytest.skip("a HIP device is present")
    with pytest.raises(smoc.SmocError) as e:
        smoc.Context(100, 10, 5)
    assert e.value.code == smoc.ERR_NODEV
    assert "device" in str(e.value)


def test_null_context_calls():
    L = smoc.lib()
    assert L.smoc_destroy(None) == smoc.OK
    assert L.smoc_run(None, None, None) == smoc.ERR_ARG
    assert L.smoc_get_results(None, None, None) == smoc.ERR_ARG


def test_probes_reject_bad_arguments_before_touching_a_device():
    L = smoc.lib()
    ms = ctypes.c_double()
    for args in ((0, -1, 10, 100, 1), (0, 100, 0, 100, 1), (0, 100, 10, 0, 1),
                 (0, 100, 10, 65537, 1), (0, 100, 10, 100, 0),
                 (0, 400 * (1 << 31) * 100, 10, 100, 1)):  # grid beyond 2^31 - 1 blocks
        assert L.smoc_probe_tally(*args, ctypes.byref(ms)) == smoc.ERR_ARG, args
    assert L.smoc_probe_tally(0, 100, 10, 100, 1, None) == smoc.ERR_ARG
    for args in ((0, 0, 1, 0), (0, 24, 1, 0), (0, 1 << 20, 0, 0), (0, 1 << 20, 1, -1),
                 (0, 1 << 20, 1, 4)):
        assert L.smoc_probe_copy(*args, ctypes.byref(ms)) == smoc.ERR_ARG, args


def test_probe_without_device_fails_loudly():
    if smoc.device_count() > 0:
        pytest.skip("a HIP device is present")
    with pytest.raises(smoc.SmocError) as e:
        smoc.probe_tally(1000, 10, 100)
    assert e.value.code == smoc.ERR_NODEV
