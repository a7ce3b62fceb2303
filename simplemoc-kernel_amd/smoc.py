"""Python mirror of the libsmoc C ABI (include/smoc.h), over ctypes.

The host-side interface follows the reference driver's vocabulary
(ANL-CESAR/SimpleMOC-kernel): ``Input`` defaults and the derived region count
(src/cpu/init.c:4-24, src/cpu/main.c:18-19), ``initialize_sources`` ->
:meth:`Context.set_regions`, the segment draws of ``run_kernel``
(src/cpu/kernel.c:43-55) -> :meth:`Context.set_segments`, and ``run_kernel``
itself -> :meth:`Context.run`.  Every call goes through libsmoc.so; there is
no fallback path: when the library or a HIP device is missing the calls raise.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsmoc.so")
DRIVER_PATH = os.path.join(HERE, "smoc")

OK, ERR_ARG, ERR_NODEV, ERR_HIP, ERR_STATE, ERR_NOMEM, ERR_RANGE = 0, -1, -2, -3, -4, -5, -6
STORE_F32, STORE_BF16, STORE_F16 = 0, 1, 2
EXP_NATIVE, EXP_TABLE, EXP_ACCURATE = 0, 1, 2
TALLY_ATOMIC_F32, TALLY_ATOMIC_F64, TALLY_DET_F64, TALLY_DET_KAHAN, TALLY_FIXED64 = 0, 1, 2, 3, 4

STORAGE = {"f32": STORE_F32, "bf16": STORE_BF16, "f16": STORE_F16}
EXP = {"native": EXP_NATIVE, "table": EXP_TABLE, "accurate": EXP_ACCURATE}
TALLY = {"atomic-f32": TALLY_ATOMIC_F32, "atomic-f64": TALLY_ATOMIC_F64,
         "det-f64": TALLY_DET_F64, "det-kahan": TALLY_DET_KAHAN, "fixed64": TALLY_FIXED64}


class SmocError(RuntimeError):
    def __init__(self, what: str, code: int, message: str):
        super().__init__(f"{what}: {message} (code {code})")
        self.code = code


class Params(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int), ("egroups", ctypes.c_int),
                ("regions", ctypes.c_int), ("fine_axial", ctypes.c_int),
                ("storage", ctypes.c_int), ("exp_mode", ctypes.c_int),
                ("tally", ctypes.c_int)]


class Launch(ctypes.Structure):
    _fields_ = [("seg_per_block", ctypes.c_int), ("groups_per_lane", ctypes.c_int),
                ("waves_per_block", ctypes.c_int), ("chunk", ctypes.c_int),
                ("persistent", ctypes.c_int), ("use_graph", ctypes.c_int),
                ("autotune", ctypes.c_int), ("reduce_waves", ctypes.c_int),
                ("segment_flush", ctypes.c_int)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class Timing(ctypes.Structure):
    _fields_ = [("total_ms", ctypes.c_double), ("reset_ms", ctypes.c_double),
                ("xfer_ms", ctypes.c_double), ("sweep_ms", ctypes.c_double),
                ("reduce_ms", ctypes.c_double),
                ("segments", ctypes.c_int64), ("work_items", ctypes.c_int64),
                ("used", Launch)]


class Input(ctypes.Structure):
    _fields_ = [("source_2D_regions", ctypes.c_int), ("coarse_axial_intervals", ctypes.c_int),
                ("fine_axial_intervals", ctypes.c_int), ("decomp_assemblies_ax", ctypes.c_int),
                ("segments", ctypes.c_int64), ("egroups", ctypes.c_int),
                ("seg_per_block", ctypes.c_int), ("source_3D_regions", ctypes.c_int)]


class ProblemInfo(ctypes.Structure):
    _fields_ = [("G", ctypes.c_int), ("R", ctypes.c_int), ("F", ctypes.c_int),
                ("S", ctypes.c_int64), ("T", ctypes.c_int64), ("seed", ctypes.c_uint64),
                ("flags", ctypes.c_uint32)]


class ResultsInfo(ctypes.Structure):
    _fields_ = [("G", ctypes.c_int), ("R", ctypes.c_int), ("F", ctypes.c_int),
                ("T", ctypes.c_int64), ("storage", ctypes.c_int), ("exp_mode", ctypes.c_int),
                ("tally", ctypes.c_int), ("fine_flux_checksum", ctypes.c_double),
                ("state_flux_checksum", ctypes.c_double)]


class Geometry(ctypes.Structure):
    """Pin-cell lattice of the track producer (smoc_geometry)."""
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("pitch", ctypes.c_double), ("pin_radius", ctypes.c_double),
                ("height", ctypes.c_double)]

    @property
    def regions(self) -> int:
        return 2 * self.nx * self.ny * self.nz


_F32 = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_F64 = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_I32 = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_I64 = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_P = ctypes.c_void_p
_i, _i64, _u64, _f, _d = ctypes.c_int, ctypes.c_int64, ctypes.c_uint64, ctypes.c_float, ctypes.c_double

# name -> (restype, argtypes); every symbol declared in include/smoc.h
SIGNATURES = {
    "smoc_abi_version": (_i, []),
    "smoc_strerror": (ctypes.c_char_p, [_i]),
    "smoc_last_error": (ctypes.c_char_p, []),
    "smoc_default_input": (_i, [ctypes.POINTER(Input)]),
    "smoc_derive_regions": (_i, [ctypes.POINTER(Input)]),
    "smoc_build_exp_table": (_i, [_f, _f, _F32, _i, ctypes.POINTER(_i), ctypes.POINTER(_f),
                                  ctypes.POINTER(_f)]),
    "smoc_exp_table_lookup": (_f, [_F32, _f, _f, _f]),
    "smoc_decode_z": (None, [_f, _i, ctypes.POINTER(_i), ctypes.POINTER(_f)]),
    "smoc_generate_regions": (_i, [_u64, _i, _i, _i, _F32, _F32]),
    "smoc_generate_segments": (_i, [_u64, _i64, _i, _F32, _I32, _F32]),
    "smoc_generate_psi": (_i, [_u64, _i64, _i, _F32]),
    "smoc_tracks_fixed": (_i, [_i64, _i, _I64, ctypes.POINTER(_i64)]),
    "smoc_tracks_lognormal": (_i, [_u64, _i64, _d, _d, _i, _i, _I64, ctypes.POINTER(_i64)]),
    "smoc_sort_segments": (_i, [_i64, _i, _i, _F32, _I32, _F32]),
    "smoc_set_host_threads": (_i, [_i]),
    "smoc_checksum_f64": (_d, [_F64, _i64]),
    "smoc_checksum_f32": (_d, [_F32, _i64]),
    "smoc_problem_write": (_i, [ctypes.c_char_p, ctypes.POINTER(ProblemInfo), _F32, _F32, _P,
                                _P, _P, _I64, _P]),
    "smoc_problem_read_info": (_i, [ctypes.c_char_p, ctypes.POINTER(ProblemInfo)]),
    "smoc_problem_read": (_i, [ctypes.c_char_p, _F32, _F32, _P, _P, _P, _I64, _P]),
    "smoc_results_write": (_i, [ctypes.c_char_p, ctypes.POINTER(ResultsInfo), _F64, _P]),
    "smoc_results_read_info": (_i, [ctypes.c_char_p, ctypes.POINTER(ResultsInfo)]),
    "smoc_results_read": (_i, [ctypes.c_char_p, _F64, _P]),
    "smoc_results_compare": (_i, [ctypes.c_char_p, _F64, _P, ctypes.POINTER(_d),
                                  ctypes.POINTER(_d), ctypes.POINTER(_d), ctypes.POINTER(_d)]),
    "smoc_device_count": (_i, [ctypes.POINTER(_i)]),
    "smoc_device_info": (_i, [_i, ctypes.c_char_p, _i, ctypes.POINTER(_i)]),
    "smoc_probe_copy": (_i, [_i, _i64, _i, _i, ctypes.POINTER(_d)]),
    "smoc_probe_tally": (_i, [_i, _i64, _i, _i, _i, ctypes.POINTER(_d)]),
    "smoc_create": (_i, [ctypes.POINTER(_P), ctypes.POINTER(Params)]),
    "smoc_destroy": (_i, [_P]),
    "smoc_set_regions": (_i, [_P, _F32, _F32, _P]),
    "smoc_set_segments": (_i, [_P, _i64, _F32, _I32, _F32, _i64, _I64, _F32]),
    "smoc_run": (_i, [_P, ctypes.POINTER(Launch), ctypes.POINTER(Timing)]),
    "smoc_run_noalloc": (_i, [_P, ctypes.POINTER(Launch), ctypes.POINTER(Timing)]),
    "smoc_get_results": (_i, [_P, _P, _P]),
    "smoc_trace_count": (_i, [_i, ctypes.POINTER(Geometry), _u64, _i64, _I64,
                              ctypes.POINTER(_i64)]),
    "smoc_trace_segments": (_i, [_i, ctypes.POINTER(Geometry), _u64, _i64, _I64, _P, _P, _P]),
    "smoc_set_scattering": (_i, [_P, _F32, _P]),
    "smoc_generate_scatter": (_i, [_u64, _i, _i, _f, _F32]),
    "smoc_source_update": (_i, [_P, ctypes.POINTER(_d)]),
    "smoc_iterate": (_i, [_P, ctypes.POINTER(Launch), _i, ctypes.POINTER(Timing),
                          ctypes.POINTER(_d), ctypes.POINTER(_d)]),
    "smoc_get_source": (_i, [_P, _F32]),
}

_lib = None


def lib():
    """Load libsmoc.so (built in-tree by __graft_entry__.build()); raise if absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(
                f"{LIB_PATH} is not built; run `make -C {HERE}` or __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(rc: int, what: str):
    if rc != OK:
        msg = lib().smoc_last_error().decode(errors="replace")
        raise SmocError(what, rc, msg)


def default_input() -> Input:
    inp = Input()
    _check(lib().smoc_default_input(ctypes.byref(inp)), "smoc_default_input")
    return inp


def exp_table(precision: float = 0.01, max_val: float = 10.0):
    vals = np.zeros(1024, np.float32)
    n, dx, mx = _i(), _f(), _f()
    _check(lib().smoc_build_exp_table(precision, max_val, vals, vals.size, ctypes.byref(n),
                                      ctypes.byref(dx), ctypes.byref(mx)), "smoc_build_exp_table")
    return vals[: 2 * n.value].copy(), dx.value, mx.value


def decode_z(z: float, F: int):
    f, zin = _i(), _f()
    lib().smoc_decode_z(z, F, ctypes.byref(f), ctypes.byref(zin))
    return f.value, zin.value


def device_count() -> int:
    n = _i()
    rc = lib().smoc_device_count(ctypes.byref(n))
    return n.value if rc == OK else 0


def device_info(device: int = 0):
    name = ctypes.create_string_buffer(256)
    cus = _i()
    _check(lib().smoc_device_info(device, name, 256, ctypes.byref(cus)), "smoc_device_info")
    return name.value.decode(), cus.value


def probe_copy(nbytes: int, reps: int = 5, device: int = 0, variant: int = 0) -> float:
    """Mean time (ms) of the plain copy kernel over nbytes (read + write)."""
    ms = _d()
    _check(lib().smoc_probe_copy(device, nbytes, reps, variant, ctypes.byref(ms)),
           "smoc_probe_copy")
    return ms.value


def probe_tally(S: int, rows: int, G: int, reps: int = 5, device: int = 0) -> float:
    """Mean time (ms) of the tally-only probe: S segments' FP32 atomic flushes
    into pseudo-random rows of [rows][G], without gathers or arithmetic."""
    ms = _d()
    _check(lib().smoc_probe_tally(device, S, rows, G, reps, ctypes.byref(ms)), "smoc_probe_tally")
    return ms.value


@dataclass
class Problem:
    """One attenuation problem: region data, segment stream and tracks."""
    G: int
    R: int
    F: int
    sigT: np.ndarray         # [R][G] f32
    fine_source: np.ndarray  # [R][F][G] f32
    len: np.ndarray          # [S] f32
    region: np.ndarray       # [S] i32
    z: np.ndarray            # [S] f32
    track_off: np.ndarray    # [T+1] i64
    psi_in: np.ndarray       # [T][G] f32
    seed: int = 42

    @property
    def S(self) -> int:
        return int(self.len.size)

    @property
    def T(self) -> int:
        return int(self.track_off.size - 1)


def generate(S: int, G: int = 100, R: int = 6750, F: int = 5, seed: int = 42,
             track_len: int = 100, lognormal_tracks: int = 0, median: float = 4000.0,
             sigma: float = 1.0, lo: int = 16, hi: int = 200000,
             sorted_stream: bool = False) -> Problem:
    """Seeded inputs of the benchmark configurations (all FP32 masters):
    sigT ~ U(0.1, 10), fine_source ~ U(0,1), len ~ U(0, 0.1), region ~ U[0,R),
    z ~ U(0,1), psi_in ~ U(0,1); tracks of fixed length or lognormal batches."""
    L = lib()
    if lognormal_tracks > 0:
        T = int(lognormal_tracks)
        off = np.zeros(T + 1, np.int64)
        s_out = _i64()
        _check(L.smoc_tracks_lognormal(seed, T, median, sigma, lo, hi, off,
                                       ctypes.byref(s_out)), "smoc_tracks_lognormal")
        S = int(s_out.value)
    else:
        Tm = (S + track_len - 1) // track_len
        off = np.zeros(Tm + 1, np.int64)
        t_out = _i64()
        _check(L.smoc_tracks_fixed(S, track_len, off, ctypes.byref(t_out)), "smoc_tracks_fixed")
        T = int(t_out.value)
    sig = np.empty((R, G), np.float32)
    src = np.empty((R, F, G), np.float32)
    _check(L.smoc_generate_regions(seed, G, R, F, sig, src), "smoc_generate_regions")
    ln = np.empty(max(S, 1), np.float32)[:S]
    rg = np.empty(max(S, 1), np.int32)[:S]
    zz = np.empty(max(S, 1), np.float32)[:S]
    ln, rg, zz = np.ascontiguousarray(ln), np.ascontiguousarray(rg), np.ascontiguousarray(zz)
    _check(L.smoc_generate_segments(seed, S, R, ln, rg, zz), "smoc_generate_segments")
    if sorted_stream:
        _check(L.smoc_sort_segments(S, R, F, ln, rg, zz), "smoc_sort_segments")
    psi = np.empty((T, G), np.float32)
    _check(L.smoc_generate_psi(seed, T, G, psi), "smoc_generate_psi")
    return Problem(G, R, F, sig, src, ln, rg, zz, off, psi, seed)


def sort_stream(prob: Problem):
    """(copy of prob with the segment stream sorted by (region, fine interval),
    host seconds of smoc_sort_segments).  The sort is stable and the tracks
    stay the same index ranges of the stream, as generate(sorted_stream=True)."""
    import copy
    import time
    q = copy.copy(prob)
    q.len, q.region, q.z = prob.len.copy(), prob.region.copy(), prob.z.copy()
    t0 = time.perf_counter()
    _check(lib().smoc_sort_segments(q.S, q.R, q.F, q.len, q.region, q.z), "smoc_sort_segments")
    return q, time.perf_counter() - t0


class Context:
    """A device context: smoc_create ... smoc_destroy."""

    def __init__(self, G: int, R: int, F: int = 5, storage="f32", exp="accurate",
                 tally="atomic-f32", device: int = 0):
        self.G, self.R, self.F = G, R, F
        st = STORAGE[storage] if isinstance(storage, str) else storage
        ex = EXP[exp] if isinstance(exp, str) else exp
        ta = TALLY[tally] if isinstance(tally, str) else tally
        self.params = Params(device, G, R, F, st, ex, ta)
        self._h = _P()
        _check(lib().smoc_create(ctypes.byref(self._h), ctypes.byref(self.params)),
               "smoc_create")
        self.T = 0

    def close(self):
        if self._h:
            lib().smoc_destroy(self._h)
            self._h = _P()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_regions(self, sigT, fine_source, fine_flux_init=None):
        sigT = np.ascontiguousarray(sigT, np.float32)
        fine_source = np.ascontiguousarray(fine_source, np.float32)
        assert sigT.size == self.R * self.G and fine_source.size == self.R * self.F * self.G
        init = None
        if fine_flux_init is not None:
            init = np.ascontiguousarray(fine_flux_init, np.float32)
            assert init.size == fine_source.size
        _check(lib().smoc_set_regions(self._h, sigT.ravel(), fine_source.ravel(),
                                      None if init is None else init.ctypes.data),
               "smoc_set_regions")

    def set_segments(self, length, region, z, track_off, psi_in):
        length = np.ascontiguousarray(length, np.float32)
        region = np.ascontiguousarray(region, np.int32)
        z = np.ascontiguousarray(z, np.float32)
        track_off = np.ascontiguousarray(track_off, np.int64)
        psi_in = np.ascontiguousarray(psi_in, np.float32)
        S, T = length.size, track_off.size - 1
        assert region.size == S and z.size == S and psi_in.size == T * self.G
        _check(lib().smoc_set_segments(self._h, S, length, region, z, T, track_off,
                                       psi_in.ravel()), "smoc_set_segments")
        self.T = T

    def load(self, prob: Problem, fine_flux_init=None):
        self.set_regions(prob.sigT, prob.fine_source, fine_flux_init)
        self.set_segments(prob.len, prob.region, prob.z, prob.track_off, prob.psi_in)

    def run(self, **launch) -> Timing:
        ln = Launch(**launch)
        tm = Timing()
        _check(lib().smoc_run(self._h, ctypes.byref(ln), ctypes.byref(tm)), "smoc_run")
        return tm

    def run_noalloc(self, **launch) -> Timing:
        """smoc_run_noalloc: the sweep with its scratch in the library's fixed
        per-GPU array (no allocation in the call)."""
        ln = Launch(**launch)
        tm = Timing()
        _check(lib().smoc_run_noalloc(self._h, ctypes.byref(ln), ctypes.byref(tm)),
               "smoc_run_noalloc")
        return tm

    def results(self):
        flux = np.empty((self.R, self.F, self.G), np.float64)
        psi = np.empty((max(self.T, 0), self.G), np.float32)
        _check(lib().smoc_get_results(self._h, flux.ctypes.data, psi.ctypes.data),
               "smoc_get_results")
        return flux, psi

    # outer (source) iteration
    def set_scattering(self, scatter, q_ext=None):
        scatter = np.ascontiguousarray(scatter, np.float32).ravel()
        assert scatter.size == self.R * self.G
        q = None
        if q_ext is not None:
            q = np.ascontiguousarray(q_ext, np.float32).ravel()
            assert q.size == self.R * self.F * self.G
        _check(lib().smoc_set_scattering(self._h, scatter, q.ctypes.data if q is not None else None),
               "smoc_set_scattering")

    def source_update(self) -> float:
        chg = _d()
        _check(lib().smoc_source_update(self._h, ctypes.byref(chg)), "smoc_source_update")
        return chg.value

    def iterate(self, n_outer: int, **launch):
        ln, tm, ums, chg = Launch(**launch), Timing(), _d(), _d()
        _check(lib().smoc_iterate(self._h, ctypes.byref(ln), n_outer, ctypes.byref(tm),
                                  ctypes.byref(ums), ctypes.byref(chg)), "smoc_iterate")
        return tm, ums.value, chg.value

    def source(self):
        out = np.empty((self.R, self.F, self.G), np.float32)
        _check(lib().smoc_get_source(self._h, out.ravel()), "smoc_get_source")
        return out


def checksums(flux: np.ndarray, psi: np.ndarray):
    """Fixed-order FP64 checksums of fine_flux and state_flux (libsmoc's)."""
    L = lib()
    f = np.ascontiguousarray(flux, np.float64).ravel()
    p = np.ascontiguousarray(psi, np.float32).ravel()
    return L.smoc_checksum_f64(f, f.size), L.smoc_checksum_f32(p, p.size)


# ---- fixture files ----------------------------------------------------------
def _ptr(a):
    return a.ctypes.data if a.size else None


def save_problem(path: str, prob: Problem, sorted_stream: bool = False):
    info = ProblemInfo(prob.G, prob.R, prob.F, prob.S, prob.T, prob.seed, 1 if sorted_stream else 0)
    arrs = [np.ascontiguousarray(prob.len, np.float32), np.ascontiguousarray(prob.region, np.int32),
            np.ascontiguousarray(prob.z, np.float32), np.ascontiguousarray(prob.psi_in, np.float32)]
    _check(lib().smoc_problem_write(os.fsencode(path), ctypes.byref(info),
                                    np.ascontiguousarray(prob.sigT, np.float32).ravel(),
                                    np.ascontiguousarray(prob.fine_source, np.float32).ravel(),
                                    _ptr(arrs[0]), _ptr(arrs[1]), _ptr(arrs[2]),
                                    np.ascontiguousarray(prob.track_off, np.int64),
                                    _ptr(arrs[3])), "smoc_problem_write")


def load_problem(path: str) -> Problem:
    info = ProblemInfo()
    _check(lib().smoc_problem_read_info(os.fsencode(path), ctypes.byref(info)),
           "smoc_problem_read_info")
    G, R, F, S, T = info.G, info.R, info.F, info.S, info.T
    sig, src = np.empty((R, G), np.float32), np.empty((R, F, G), np.float32)
    ln, rg, zz = np.empty(S, np.float32), np.empty(S, np.int32), np.empty(S, np.float32)
    off, psi = np.empty(T + 1, np.int64), np.empty((T, G), np.float32)
    _check(lib().smoc_problem_read(os.fsencode(path), sig.ravel(), src.ravel(), _ptr(ln),
                                   _ptr(rg), _ptr(zz), off, _ptr(psi)), "smoc_problem_read")
    return Problem(G, R, F, sig, src, ln, rg, zz, off, psi, int(info.seed))


def save_results(path: str, flux, psi, storage="f32", exp="accurate", tally="atomic-f32"):
    flux = np.ascontiguousarray(flux, np.float64)
    psi = np.ascontiguousarray(psi, np.float32)
    R, F, G = flux.shape
    info = ResultsInfo(G, R, F, psi.shape[0], STORAGE.get(storage, storage), EXP.get(exp, exp),
                       TALLY.get(tally, tally), 0.0, 0.0)
    _check(lib().smoc_results_write(os.fsencode(path), ctypes.byref(info), flux.ravel(),
                                    _ptr(psi)), "smoc_results_write")


def load_results(path: str):
    info = ResultsInfo()
    _check(lib().smoc_results_read_info(os.fsencode(path), ctypes.byref(info)),
           "smoc_results_read_info")
    flux = np.empty((info.R, info.F, info.G), np.float64)
    psi = np.empty((info.T, info.G), np.float32)
    _check(lib().smoc_results_read(os.fsencode(path), flux.ravel(), _ptr(psi)),
           "smoc_results_read")
    return info, flux, psi


def compare_results(path: str, flux, psi):
    flux = np.ascontiguousarray(flux, np.float64)
    psi = np.ascontiguousarray(psi, np.float32)
    out = [_d() for _ in range(4)]
    _check(lib().smoc_results_compare(os.fsencode(path), flux.ravel(), _ptr(psi),
                                      *[ctypes.byref(o) for o in out]), "smoc_results_compare")
    return dict(zip(("fine_checksum_rel", "state_checksum_rel", "fine_normwise",
                     "state_normwise"), (o.value for o in out)))


def generate_scatter(R: int, G: int, cmax: float, seed: int = 42) -> np.ndarray:
    """Scattering ratios cmax * U(0,1) [R][G] from the library's generator."""
    out = np.empty((R, G), np.float32)
    _check(lib().smoc_generate_scatter(seed, R, G, cmax, out.ravel()), "smoc_generate_scatter")
    return out


def trace(geometry: Geometry, T: int, seed: int = 42, device: int = 0):
    """Ray-trace T tracks through the pin-cell lattice on the GPU:
    (track_off [T+1], len [S], region [S], z [S])."""
    off = np.empty(T + 1, np.int64)
    S = _i64()
    _check(lib().smoc_trace_count(device, ctypes.byref(geometry), seed, T, off,
                                  ctypes.byref(S)), "smoc_trace_count")
    n = S.value
    ln, rg, zz = np.empty(n, np.float32), np.empty(n, np.int32), np.empty(n, np.float32)
    _check(lib().smoc_trace_segments(device, ctypes.byref(geometry), seed, T, off, _ptr(ln),
                                     _ptr(rg), _ptr(zz)), "smoc_trace_segments")
    return off, ln, rg, zz


def traced_problem(geometry: Geometry, T: int, G: int = 100, F: int = 5, seed: int = 42,
                   device: int = 0) -> Problem:
    """A problem whose segment stream comes from the tracer; region data and
    psi_in from the library's generator (R = 2 nx ny nz)."""
    off, ln, rg, zz = trace(geometry, T, seed, device)
    R = geometry.regions
    sig, src = np.empty((R, G), np.float32), np.empty((R, F, G), np.float32)
    _check(lib().smoc_generate_regions(seed, G, R, F, sig.ravel(), src.ravel()),
           "smoc_generate_regions")
    psi = np.empty((T, G), np.float32)
    _check(lib().smoc_generate_psi(seed, T, G, psi.ravel()), "smoc_generate_psi")
    return Problem(G, R, F, sig, src, ln, rg, zz, off, psi, seed)
