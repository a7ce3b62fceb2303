"""ctypes bindings of the CPU checker (oracle/liboracle.so) and of the reference
build (oracle/_ref/libsmocref_*.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

_F32 = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_F64 = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_I32 = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_I64 = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
EXP_NATIVE, EXP_TABLE = 0, 1


class Table(ctypes.Structure):
    _fields_ = [("values", ctypes.c_float * 1024), ("dx", ctypes.c_float),
                ("maxVal", ctypes.c_float), ("N", ctypes.c_int)]


class Problem(ctypes.Structure):
    _fields_ = [("G", ctypes.c_int), ("R", ctypes.c_int), ("F", ctypes.c_int),
                ("exp_mode", ctypes.c_int),
                ("sigT", ctypes.c_void_p), ("fine_source", ctypes.c_void_p),
                ("S", ctypes.c_int64), ("len", ctypes.c_void_p), ("region", ctypes.c_void_p),
                ("z", ctypes.c_void_p), ("T", ctypes.c_int64), ("track_off", ctypes.c_void_p),
                ("psi_in", ctypes.c_void_p)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_LIB):
            raise FileNotFoundError(f"{ORACLE_LIB} not built; run `make -C {HERE}`")
        _lib = load(ORACLE_LIB)
    return _lib


def load(path):
    """A build of smoc_oracle.c at `path` with its signatures bound (the
    checker itself, or a mutant compiled by a test)."""
    L = ctypes.CDLL(path)
    TP = ctypes.POINTER(Table)
    L.oracle_build_table.argtypes = [ctypes.c_float, ctypes.c_float, TP]
    L.oracle_build_table.restype = ctypes.c_int
    L.oracle_table_lookup_f32.argtypes = [TP, ctypes.c_float]
    L.oracle_table_lookup_f32.restype = ctypes.c_float
    L.oracle_table_interval.argtypes = [TP, ctypes.c_float]
    L.oracle_table_interval.restype = ctypes.c_int
    L.oracle_decode_z.argtypes = [ctypes.c_float, ctypes.c_int,
                                  ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_float)]
    L.oracle_decode_z.restype = None
    L.oracle_segment_f32.argtypes = [ctypes.c_int, ctypes.c_int, _F32, _F32, ctypes.c_int,
                                     ctypes.c_float, ctypes.c_float, ctypes.c_int, TP,
                                     _F32, _F32]
    L.oracle_segment_f32.restype = None
    L.oracle_segment_f64.argtypes = [ctypes.c_int, ctypes.c_int, _F32, _F32, ctypes.c_int,
                                     ctypes.c_float, ctypes.c_float, ctypes.c_int, TP,
                                     _F64, _F64]
    L.oracle_segment_f64.restype = None
    PP = ctypes.POINTER(Problem)
    L.oracle_sweep_f64.argtypes = [PP, _F64, _F64, ctypes.c_int]
    L.oracle_sweep_f64.restype = None
    L.oracle_sweep_f32.argtypes = [PP, _F32, _F32]
    L.oracle_sweep_f32.restype = None
    L.oracle_baseline_f32.argtypes = [PP, _F32, _F32, ctypes.c_int]
    L.oracle_baseline_f32.restype = ctypes.c_double
    return L


def table(precision=0.01, max_val=10.0) -> Table:
    t = Table()
    assert lib().oracle_build_table(precision, max_val, ctypes.byref(t)) == 0
    return t


def decode_z(z: float, F: int):
    f, zin = ctypes.c_int(), ctypes.c_float()
    lib().oracle_decode_z(z, F, ctypes.byref(f), ctypes.byref(zin))
    return f.value, zin.value


class _Keep:
    """Holds the numpy arrays a Problem struct points into."""

    def __init__(self, prob, exp_mode):
        self.sig = np.ascontiguousarray(prob.sigT, np.float32)
        self.src = np.ascontiguousarray(prob.fine_source, np.float32)
        self.len = np.ascontiguousarray(prob.len, np.float32)
        self.reg = np.ascontiguousarray(prob.region, np.int32)
        self.z = np.ascontiguousarray(prob.z, np.float32)
        self.off = np.ascontiguousarray(prob.track_off, np.int64)
        self.psi = np.ascontiguousarray(prob.psi_in, np.float32)
        p = Problem()
        p.G, p.R, p.F, p.exp_mode = prob.G, prob.R, prob.F, exp_mode
        p.sigT, p.fine_source = self.sig.ctypes.data, self.src.ctypes.data
        p.S = self.len.size
        p.len, p.region, p.z = self.len.ctypes.data, self.reg.ctypes.data, self.z.ctypes.data
        p.T = self.off.size - 1
        p.track_off, p.psi_in = self.off.ctypes.data, self.psi.ctypes.data
        self.p = p


def sweep_f64(prob, exp_mode=EXP_NATIVE, nthreads=1, fine_flux_init=None):
    k = _Keep(prob, exp_mode)
    flux = np.zeros((prob.R, prob.F, prob.G), np.float64)
    if fine_flux_init is not None:
        flux += np.asarray(fine_flux_init, np.float64).reshape(flux.shape)
    psi = np.zeros((prob.T, prob.G), np.float64)
    lib().oracle_sweep_f64(ctypes.byref(k.p), flux, psi, nthreads)
    return flux, psi


def sweep_f32(prob, exp_mode=EXP_NATIVE):
    k = _Keep(prob, exp_mode)
    flux = np.zeros((prob.R, prob.F, prob.G), np.float32)
    psi = np.zeros((prob.T, prob.G), np.float32)
    lib().oracle_sweep_f32(ctypes.byref(k.p), flux, psi)
    return flux, psi


def baseline_f32(prob, exp_mode=EXP_NATIVE, nthreads=1):
    k = _Keep(prob, exp_mode)
    flux = np.zeros((prob.R, prob.F, prob.G), np.float32)
    psi = np.zeros((prob.T, prob.G), np.float32)
    sec = lib().oracle_baseline_f32(ctypes.byref(k.p), flux, psi, nthreads)
    return sec, flux, psi


# ---- the reference's own code (oracle/_ref) ------------------------------------
REF_VARIANTS = ("exp", "table", "exp_strict", "table_strict")


def ref_path(variant: str) -> str:
    return os.path.join(REF_DIR, f"libsmocref_{variant}.so")


def ref_available(variant: str = "exp") -> bool:
    return os.path.exists(ref_path(variant))


class Ref:
    """One loaded reference build (global state inside: one problem at a time)."""

    _loaded: dict = {}

    def __init__(self, variant: str):
        if variant not in Ref._loaded:
            L = ctypes.CDLL(ref_path(variant))
            L.ref_init.argtypes = [ctypes.c_int] * 3
            L.ref_init.restype = ctypes.c_int
            L.ref_set_regions.argtypes = [_F32, _F32, _F32]
            L.ref_set_regions.restype = None
            L.ref_get_flux.argtypes = [_F32]
            L.ref_get_flux.restype = None
            L.ref_run_segments.argtypes = [ctypes.c_long, _I32, _I32, _F32]
            L.ref_run_segments.restype = None
            L.ref_table.argtypes = [_F32, ctypes.POINTER(ctypes.c_int),
                                    ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]
            L.ref_table.restype = None
            L.ref_interp.argtypes = [ctypes.c_float]
            L.ref_interp.restype = ctypes.c_float
            L.ref_sweep.argtypes = [ctypes.c_long, _I64, _I32, _I32, _F32, _F32, ctypes.c_int]
            L.ref_sweep.restype = ctypes.c_double
            Ref._loaded[variant] = L
        self.L = Ref._loaded[variant]
        self.shape = None

    def init(self, G, R, F):
        self.L.ref_init(G, R, F)
        self.shape = (R, F, G)

    def set_regions(self, sigT, fine_source, fine_flux=None):
        R, F, G = self.shape
        ff = np.zeros((R, F, G), np.float32) if fine_flux is None else fine_flux
        self.L.ref_set_regions(np.ascontiguousarray(sigT, np.float32).ravel(),
                               np.ascontiguousarray(fine_source, np.float32).ravel(),
                               np.ascontiguousarray(ff, np.float32).ravel())

    def run_segments(self, region, fai, psi):
        psi = np.ascontiguousarray(psi, np.float32).copy()
        self.L.ref_run_segments(len(region), np.ascontiguousarray(region, np.int32),
                                np.ascontiguousarray(fai, np.int32), psi)
        return psi

    def flux(self):
        out = np.empty(self.shape, np.float32)
        self.L.ref_get_flux(out.ravel())
        return out

    def table(self):
        vals = np.empty(1024, np.float32)
        n, dx, mx = ctypes.c_int(), ctypes.c_float(), ctypes.c_float()
        self.L.ref_table(vals, ctypes.byref(n), ctypes.byref(dx), ctypes.byref(mx))
        return vals[: 2 * n.value].copy(), dx.value, mx.value

    def interp(self, x: float) -> float:
        return self.L.ref_interp(x)

    def sweep(self, track_off, region, fai, psi_in, nthreads):
        T = track_off.size - 1
        psi_out = np.empty_like(np.ascontiguousarray(psi_in, np.float32))
        sec = self.L.ref_sweep(T, np.ascontiguousarray(track_off, np.int64),
                               np.ascontiguousarray(region, np.int32),
                               np.ascontiguousarray(fai, np.int32),
                               np.ascontiguousarray(psi_in, np.float32).ravel(), psi_out.ravel(),
                               nthreads)
        return sec, psi_out


# ---- outer (source) iteration --------------------------------------------------
# A restatement of the library's source update (numpy, FP64).  The reference
# kernel has no source update; the update is the fixed-source, isotropic
# within-group scattering iteration of the application the kernel was cut
# from, with the track-length estimate of the angle-averaged flux.
def track_volumes(prob, weight=0.5):
    """Sum of weight * len over the segments of every (region, interval),
    FP64, shape [R][F]; the interval is decoded from z with the library's
    FP32 operations."""
    F = prob.F
    zf = (np.asarray(prob.z, np.float32) * np.float32(F)).astype(np.float32)
    f = np.clip(zf.astype(np.int32), 0, F - 1).astype(np.int64)
    vol = np.zeros(prob.R * F)
    np.add.at(vol, np.asarray(prob.region, np.int64) * F + f,
              weight * np.asarray(prob.len, np.float64))
    return vol.reshape(prob.R, F)


def source_update(fine_flux, vol, sigT, scatter, q_ext):
    """q_ext + scatter * sigT * fine_flux / vol per (region, interval, group),
    FP64; (region, interval) rows without segments keep q_ext."""
    vol = np.asarray(vol, np.float64)
    phi = np.where(vol[..., None] > 0,
                   np.asarray(fine_flux, np.float64) / np.where(vol > 0, vol, 1.0)[..., None],
                   0.0)
    return (np.asarray(q_ext, np.float64) + np.asarray(scatter, np.float64)[:, None, :] *
            np.asarray(sigT, np.float64)[:, None, :] * phi)


# ---- track producer restatement --------------------------------------------------
class Geometry(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("pitch", ctypes.c_double), ("pin_radius", ctypes.c_double),
                ("height", ctypes.c_double)]


class Ray(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("ex", "ey", "ez", "dx", "dy", "dz", "L")]


def _trace_lib():
    L = lib()
    if not getattr(L, "_trace_bound", False):
        L.oracle_trace_ray.argtypes = [ctypes.POINTER(Geometry), ctypes.c_uint64, ctypes.c_int64,
                                       ctypes.POINTER(Ray)]
        L.oracle_trace_ray.restype = None
        L.oracle_trace_track.argtypes = [ctypes.POINTER(Geometry), ctypes.c_uint64,
                                         ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p,
                                         ctypes.c_void_p, ctypes.c_void_p]
        L.oracle_trace_track.restype = ctypes.c_int64
        L._trace_bound = True
    return L


def trace_ray(g: Geometry, seed: int, t: int) -> Ray:
    r = Ray()
    _trace_lib().oracle_trace_ray(ctypes.byref(g), seed, t, ctypes.byref(r))
    return r


def trace(g: Geometry, seed: int, T: int):
    """All T tracks: (track_off [T+1], len [S], region [S], z [S])."""
    L = _trace_lib()
    counts = [L.oracle_trace_track(ctypes.byref(g), seed, t, 0, None, None, None)
              for t in range(T)]
    off = np.zeros(T + 1, np.int64)
    off[1:] = np.cumsum(counts)
    S = int(off[-1])
    ln, rg, zz = np.empty(S, np.float32), np.empty(S, np.int32), np.empty(S, np.float32)
    for t in range(T):
        b, n = int(off[t]), int(counts[t])
        if n:
            L.oracle_trace_track(ctypes.byref(g), seed, t, n, ln[b:].ctypes.data,
                                 rg[b:].ctypes.data, zz[b:].ctypes.data)
    return off, ln, rg, zz
