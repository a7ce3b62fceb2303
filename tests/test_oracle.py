"""Pins the CPU oracle (oracle/smoc_oracle.c) before it is trusted as checker.

* golden vectors written by the reference's own code (tests/golden/);
* the live reference build (oracle/_ref) on fresh random inputs, where built;
* closed forms the mathematics fixes: Beer-Lambert attenuation for a zero
  source, the flat-source MOC solution, affinity in the incoming flux;
* FP64 vs literal FP32 agreement, thread-count independence of the sweep.
"""
import os

import numpy as np
import pytest

import pyoracle

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = np.load(os.path.join(HERE, "golden", "ref_vectors.npz"))


def _segments_f32(d, exp_mode, ds=0.7, zin=0.3):
    """Run the fixture's 64 segments through oracle_segment_f32 with the
    reference constants (src/cpu/kernel.c:100,104)."""
    L = pyoracle.lib()
    tab = pyoracle.table()
    G, F = int(d["seg_G"]), int(d["seg_F"])
    flux = d["seg_fine_flux_in"].copy()
    psi = d["seg_psi_in"].copy()
    for r, f in zip(d["seg_region"], d["seg_fai"]):
        blk = np.ascontiguousarray(flux[r])
        L.oracle_segment_f32(G, F, np.ascontiguousarray(d["seg_sigT"][r]),
                             np.ascontiguousarray(d["seg_fine_source"][r]), int(f), ds, zin,
                             exp_mode, tab, psi, blk)
        flux[r] = blk
    return psi, flux


def test_table_bitwise_vs_reference_fixture():
    t = pyoracle.table(0.01, 10.0)
    n = int(GOLD["table_N"])
    assert t.N == n == 353
    vals = np.ctypeslib.as_array(t.values)[: 2 * n]
    assert np.array_equal(vals.view(np.uint32), GOLD["table_values"].view(np.uint32))
    assert np.float32(t.dx) == GOLD["table_dx"]
    assert np.float32(t.maxVal) == GOLD["table_maxVal"]


def test_table_lookup_bitwise_and_known_answers():
    t = pyoracle.table()
    L = pyoracle.lib()
    ys = np.array([L.oracle_table_lookup_f32(t, float(x)) for x in GOLD["interp_x"]], np.float32)
    assert np.array_equal(ys.view(np.uint32), GOLD["interp_y"].view(np.uint32))
    # known answers of the reference table (sign bug included), FP32 gcc -O2
    known = {0.0: 0.0, 0.05: 0.0068650, 0.3: 0.2341051, 0.7: 0.4831347, 1.0: 0.6258277,
             5.0: 0.9930691, 9.9: 0.9999486, 12.0: 1.0}
    for x, y in known.items():
        assert abs(L.oracle_table_lookup_f32(t, x) - y) < 1e-7, x
    # the slope is -exp(-n dx): the table's error against 1-e^-x is large
    x = np.float32(0.02792)
    assert abs(L.oracle_table_lookup_f32(t, float(x)) - (1 - np.exp(-0.02792))) > 0.05


@pytest.mark.parametrize("variant,mode", [("exp_strict", 0), ("table_strict", 1)])
def test_segment_f32_bitexact_vs_reference_strict(variant, mode):
    psi, flux = _segments_f32(GOLD, mode)
    assert np.array_equal(psi.view(np.uint32), GOLD[f"psi_out_{variant}"].view(np.uint32))
    assert np.array_equal(flux.view(np.uint32), GOLD[f"flux_{variant}"].view(np.uint32))


@pytest.mark.parametrize("variant,mode", [("exp", 0), ("table", 1)])
def test_segment_f32_vs_reference_fast_flags(variant, mode):
    # the reference's own -Ofast build reassociates; same arithmetic to ~1e-6
    psi, flux = _segments_f32(GOLD, mode)
    ref_psi, ref_flux = GOLD[f"psi_out_{variant}"], GOLD[f"flux_{variant}"]
    assert np.linalg.norm(psi - ref_psi) / np.linalg.norm(ref_psi) < 1e-6
    assert np.linalg.norm(flux - ref_flux) / np.linalg.norm(ref_flux) < 1e-6


def test_segment_f64_vs_f32_on_golden():
    L = pyoracle.lib()
    tab = pyoracle.table()
    d = GOLD
    G, F = int(d["seg_G"]), int(d["seg_F"])
    flux = d["seg_fine_flux_in"].astype(np.float64)
    psi = d["seg_psi_in"].astype(np.float64)
    for r, f in zip(d["seg_region"], d["seg_fai"]):
        blk = np.ascontiguousarray(flux[r])
        L.oracle_segment_f64(G, F, np.ascontiguousarray(d["seg_sigT"][r]),
                             np.ascontiguousarray(d["seg_fine_source"][r]), int(f), 0.7, 0.3, 0,
                             tab, psi, blk)
        flux[r] = blk
    ref = d["psi_out_exp_strict"].astype(np.float64)
    assert np.linalg.norm(psi - ref) / np.linalg.norm(ref) < 1e-5


def _tally_f64(L, d):
    """The FP64 oracle's fine_flux increments over the fixture's 64 segments
    (reference constants ds = 0.7, zin = 0.3: tau = sigT ds reaches 7, where
    the quadratic-source term of the flux integral is not negligible)."""
    tab = pyoracle.table()
    G, F = int(d["seg_G"]), int(d["seg_F"])
    flux0 = d["seg_fine_flux_in"].astype(np.float64)
    flux = flux0.copy()
    psi = d["seg_psi_in"].astype(np.float64)
    for r, f in zip(d["seg_region"], d["seg_fai"]):
        blk = np.ascontiguousarray(flux[r])
        L.oracle_segment_f64(G, F, np.ascontiguousarray(d["seg_sigT"][r]),
                             np.ascontiguousarray(d["seg_fine_source"][r]), int(f), 0.7, 0.3, 0,
                             tab, psi, blk)
        flux[r] = blk
    return flux - flux0


def _golden_tally(d):
    return d["flux_exp_strict"].astype(np.float64) - d["seg_fine_flux_in"].astype(np.float64)


# normwise bound of the FP64 tally against the reference's strict FP32 build:
# measured 5.8e-7 (the FP32 rounding of the reference); the q2 term of the flux
# integral (kernel.c:250-252) is 4.2e-4 of the tally here, so a dropped,
# doubled or sign-flipped q2 term is 700x beyond the bound
TALLY_PIN = 2e-6


def test_segment_f64_tally_vs_reference_golden():
    """The FP64 oracle's fine_flux tally -- every term of the flux integral,
    kernel.c:245-252, weighted as kernel.c:259-263 -- against the reference's
    own tallies on the golden vectors."""
    inc, ref = _tally_f64(pyoracle.lib(), GOLD), _golden_tally(GOLD)
    assert np.linalg.norm(inc - ref) / np.linalg.norm(ref) < TALLY_PIN


def test_tally_pin_catches_a_doubled_q2_term(tmp_path):
    """A mutant oracle whose q2 term is doubled (compiled here from the same
    source with -DORACLE_Q2_TERM_SCALE=2) fails the pin above."""
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler to build the mutant")
    src = os.path.join(os.path.dirname(pyoracle.__file__), "smoc_oracle.c")
    out = str(tmp_path / "liboracle_q2x2.so")
    subprocess.run([cc, "-std=gnu99", "-O2", "-ffp-contract=off", "-fno-fast-math", "-fopenmp",
                    "-fPIC", "-shared", "-DORACLE_Q2_TERM_SCALE=2.0", src, "-o", out, "-lm"],
                   check=True)
    inc, ref = _tally_f64(pyoracle.load(out), GOLD), _golden_tally(GOLD)
    err = np.linalg.norm(inc - ref) / np.linalg.norm(ref)
    assert err > 100 * TALLY_PIN, err


@pytest.mark.skipif(not pyoracle.ref_available("exp_strict"), reason="oracle/_ref not built")
@pytest.mark.parametrize("mode,variant", [(0, "exp_strict"), (1, "table_strict")])
def test_segment_f32_bitexact_vs_live_reference(mode, variant):
    rng = np.random.default_rng(99)
    G, R, F, n = 37, 9, 4, 200
    d = {"seg_G": G, "seg_R": R, "seg_F": F,
         "seg_sigT": rng.uniform(0.05, 12, (R, G)).astype(np.float32),
         "seg_fine_source": rng.uniform(-1, 2, (R, F, G)).astype(np.float32),
         "seg_fine_flux_in": rng.uniform(0, 1, (R, F, G)).astype(np.float32),
         "seg_region": rng.integers(0, R, n).astype(np.int32),
         "seg_fai": rng.integers(0, F, n).astype(np.int32),
         "seg_psi_in": rng.uniform(0, 1, G).astype(np.float32)}
    ref = pyoracle.Ref(variant)
    ref.init(G, R, F)
    ref.set_regions(d["seg_sigT"], d["seg_fine_source"], d["seg_fine_flux_in"])
    rpsi = ref.run_segments(d["seg_region"], d["seg_fai"], d["seg_psi_in"])
    rflux = ref.flux()
    psi, flux = _segments_f32(d, mode)
    assert np.array_equal(psi.view(np.uint32), rpsi.view(np.uint32))
    assert np.array_equal(flux.view(np.uint32), rflux.view(np.uint32))


def _one_f64(sig, src, f, ds, zin, psi, F=3):
    L = pyoracle.lib()
    G = sig.size
    flux = np.zeros((F, G), np.float64)
    p = psi.astype(np.float64).copy()
    L.oracle_segment_f64(G, F, sig.astype(np.float32), np.ascontiguousarray(src, np.float32), f,
                         ds, zin, 0, pyoracle.table(), p, flux)
    return p, flux[f]


def test_zero_source_is_beer_lambert():
    rng = np.random.default_rng(1)
    G = 64
    sig = rng.uniform(0.1, 10, G).astype(np.float32)
    psi = rng.uniform(0, 1, G)
    for f in (0, 1, 2):
        ds = np.float32(0.37)
        p, tally = _one_f64(sig, np.zeros((3, G)), f, ds, np.float32(0.01), psi)
        tau = sig.astype(np.float64) * np.float64(ds)
        np.testing.assert_allclose(p, psi * np.exp(-tau), rtol=1e-13)
        # tally = w * psi * (1 - e^-tau) / sigT
        np.testing.assert_allclose(tally, 0.5 * psi * -np.expm1(-tau) / sig, rtol=1e-12)


def test_flat_source_textbook_solution():
    rng = np.random.default_rng(2)
    G = 64
    sig = rng.uniform(0.1, 10, G).astype(np.float32)
    psi = rng.uniform(0, 1, G)
    q = np.float32(0.625)
    for f in (0, 1, 2):
        ds = np.float32(0.0831)
        p, tally = _one_f64(sig, np.full((3, G), q), f, ds, np.float32(-0.02), psi)
        s = sig.astype(np.float64)
        tau = s * np.float64(ds)
        ev = -np.expm1(-tau)
        # psi' = psi e^-tau + (q / sigT)(1 - e^-tau)
        np.testing.assert_allclose(p, psi * np.exp(-tau) + q / s * ev, rtol=1e-12)
        # integral: (q tau + (sigT psi - q) ev) / sigT^2
        np.testing.assert_allclose(tally, 0.5 * (q * tau + (s * psi - q) * ev) / s**2,
                                   rtol=1e-11)


def test_affine_in_incoming_flux():
    rng = np.random.default_rng(3)
    G = 50
    sig = rng.uniform(0.1, 10, G).astype(np.float32)
    src = rng.uniform(0, 1, (3, G))
    for f in (0, 1, 2):
        p0, t0 = _one_f64(sig, src, f, np.float32(0.05), np.float32(0.03), np.zeros(G))
        p1, t1 = _one_f64(sig, src, f, np.float32(0.05), np.float32(0.03), np.ones(G))
        p2, t2 = _one_f64(sig, src, f, np.float32(0.05), np.float32(0.03), 2 * np.ones(G))
        np.testing.assert_allclose(p2 - p1, p1 - p0, rtol=1e-9, atol=1e-12)
        np.testing.assert_allclose(t2 - t1, t1 - t0, rtol=1e-9, atol=1e-12)
        tau = sig.astype(np.float64) * np.float64(np.float32(0.05))
        np.testing.assert_allclose(p1 - p0, np.exp(-tau), rtol=1e-12)
        np.testing.assert_allclose(t1 - t0, 0.5 * -np.expm1(-tau) / sig, rtol=1e-10)


def _problem(S=4000, G=24, R=40, F=5, L=50, seed=0):
    rng = np.random.default_rng(seed)

    class P:
        pass
    p = P()
    p.G, p.R, p.F = G, R, F
    p.sigT = rng.uniform(0.1, 10, (R, G)).astype(np.float32)
    p.fine_source = rng.uniform(0, 1, (R, F, G)).astype(np.float32)
    p.len = (rng.random(S) * 0.1).astype(np.float32)
    p.region = rng.integers(0, R, S).astype(np.int32)
    p.z = rng.random(S).astype(np.float32)
    T = (S + L - 1) // L
    p.track_off = np.minimum(np.arange(T + 1) * L, S).astype(np.int64)
    p.psi_in = rng.random((T, G)).astype(np.float32)
    p.T = T
    return p


def test_sweep_threads_independent():
    p = _problem()
    f1, p1 = pyoracle.sweep_f64(p, 0, nthreads=1)
    f8, p8 = pyoracle.sweep_f64(p, 0, nthreads=8)
    assert np.array_equal(p1, p8)
    np.testing.assert_allclose(f8, f1, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("mode", [0, 1])
def test_sweep_f64_vs_literal_f32(mode):
    p = _problem(seed=4)
    f64, p64 = pyoracle.sweep_f64(p, mode)
    f32, p32 = pyoracle.sweep_f32(p, mode)
    assert np.linalg.norm(p32 - p64) / np.linalg.norm(p64) < 1e-5
    assert abs(f32.sum(dtype=np.float64) - f64.sum()) / np.abs(f64).sum() < 1e-5


def test_sweep_equals_segment_by_segment():
    # the sweep restates run_kernel's loop: psi carried along each track
    p = _problem(S=300, G=8, R=5, F=3, L=30, seed=5)
    f_sw, p_sw = pyoracle.sweep_f32(p, 0)
    L = pyoracle.lib()
    tab = pyoracle.table()
    flux = np.zeros((p.R, p.F, p.G), np.float32)
    for t in range(p.T):
        psi = p.psi_in[t].copy()
        for k in range(p.track_off[t], p.track_off[t + 1]):
            f, zin = pyoracle.decode_z(float(p.z[k]), p.F)
            r = p.region[k]
            blk = np.ascontiguousarray(flux[r])
            L.oracle_segment_f32(p.G, p.F, np.ascontiguousarray(p.sigT[r]),
                                 np.ascontiguousarray(p.fine_source[r]), f, float(p.len[k]), zin,
                                 0, tab, psi, blk)
            flux[r] = blk
        assert np.array_equal(psi, p_sw[t])
    assert np.array_equal(flux, f_sw)


def test_baseline_matches_sweep():
    p = _problem(seed=6)
    sec, fb, pb = pyoracle.baseline_f32(p, 0, nthreads=4)
    f32, p32 = pyoracle.sweep_f32(p, 0)
    assert sec > 0
    assert np.array_equal(pb, p32)  # psi does not depend on tally order
    np.testing.assert_allclose(fb, f32, rtol=1e-5, atol=1e-5)


# ---- outer iteration restatement ------------------------------------------------
def _uniform_problem(q, sig, psi_scale, S=4000, G=8, F=2, seed=3):
    """One region, the same source in every interval: a flat source with
    c1 = c2 = 0, tracks entering at psi_in = psi_scale * q / sigT."""
    prob = _problem(S=S, G=G, R=1, F=F, L=40, seed=seed)
    prob.sigT = np.full((1, G), sig, np.float32)
    prob.fine_source = np.full((1, F, G), q, np.float32)
    prob.psi_in = np.full((prob.T, G), psi_scale * q / sig, np.float32)
    return prob


def test_volumes_conserve_length():
    prob = _problem(S=5000, G=4, R=30, F=5, seed=4)
    vol = pyoracle.track_volumes(prob)
    assert vol.shape == (30, 5)
    assert abs(vol.sum() - 0.5 * prob.len.astype(np.float64).sum()) < 1e-9
    # brute force on the first rows
    for k in range(50):
        f = min(int(np.float32(prob.z[k]) * np.float32(5)), 4)
        assert vol[prob.region[k], f] > 0


def test_update_infinite_medium_fixed_point():
    """Flat source q* = q_ext / (1 - c) with tracks entering at equilibrium
    psi = q* / sigT: every segment keeps psi, the track-length flux estimate
    is q* / sigT exactly (in real arithmetic), so the update returns q*."""
    c, qext, sig = 0.4, 0.3, 2.5
    qstar = qext / (1 - c)
    prob = _uniform_problem(qstar, sig, 1.0)
    flux, _ = pyoracle.sweep_f64(prob, pyoracle.EXP_NATIVE)
    vol = pyoracle.track_volumes(prob)
    q = pyoracle.source_update(flux, vol, prob.sigT, np.full((1, 8), c),
                               np.full((1, 2, 8), qext))
    assert np.max(np.abs(q / qstar - 1)) < 1e-6  # the FP32 inputs limit it


def test_update_zero_scatter_and_linearity():
    prob = _problem(S=3000, G=6, R=200, F=3, seed=5)
    flux, _ = pyoracle.sweep_f64(prob, pyoracle.EXP_NATIVE)
    vol = pyoracle.track_volumes(prob)
    qext = np.random.default_rng(0).random((200, 3, 6))
    c = np.random.default_rng(1).random((200, 6)) * 0.9
    assert np.array_equal(pyoracle.source_update(flux, vol, prob.sigT, 0 * c, qext), qext)
    d1 = pyoracle.source_update(flux, vol, prob.sigT, c, qext) - qext
    d2 = pyoracle.source_update(flux, vol, prob.sigT, c / 2, qext) - qext
    np.testing.assert_allclose(d1, 2 * d2, rtol=1e-12)
    empty = vol == 0
    assert np.all(d1[empty] == 0)


# ---- tracer restatement ------------------------------------------------------------
def _geometry(nx=6, ny=5, nz=4, pitch=1.26, r=0.54, height=7.0):
    return pyoracle.Geometry(nx, ny, nz, pitch, r, height)


def test_trace_lengths_add_up_to_the_chord():
    g = _geometry()
    off, ln, rg, zz = pyoracle.trace(g, 7, 300)
    for t in range(300):
        ray = pyoracle.trace_ray(g, 7, t)
        # the chord ends on the box surface: an independent check of L
        ex, ey, ez = ray.ex + ray.L * ray.dx, ray.ey + ray.L * ray.dy, ray.ez + ray.L * ray.dz
        X, Y, Z = g.nx * g.pitch, g.ny * g.pitch, g.height
        on_face = [abs(v) < 1e-9 or abs(v - hi) < 1e-9
                   for v, hi in ((ex, X), (ey, Y), (ez, Z))]
        assert any(on_face)
        assert abs(ray.dx ** 2 + ray.dy ** 2 + ray.dz ** 2 - 1) < 1e-12
        seg = ln[off[t]:off[t + 1]].astype(np.float64)
        assert abs(seg.sum() - ray.L) <= 1e-6 * max(ray.L, 1e-3)


def test_trace_regions_and_z_by_brute_force():
    g = _geometry()
    off, ln, rg, zz = pyoracle.trace(g, 11, 200)
    R = 2 * g.nx * g.ny * g.nz
    assert rg.min() >= 0 and rg.max() < R and zz.min() >= 0 and zz.max() < 1
    dzl = g.height / g.nz
    for t in range(0, 200, 7):
        ray = pyoracle.trace_ray(g, 11, t)
        s = 0.0
        for k in range(off[t], off[t + 1]):
            m = s + 0.5 * float(ln[k])
            px, py, pz = ray.ex + m * ray.dx, ray.ey + m * ray.dy, ray.ez + m * ray.dz
            ix, iy = int(px // g.pitch), int(py // g.pitch)
            cx, cy = (ix + 0.5) * g.pitch, (iy + 0.5) * g.pitch
            outside = 0 if (px - cx) ** 2 + (py - cy) ** 2 < g.pin_radius ** 2 else 1
            layer = min(int(pz // dzl), g.nz - 1)
            if float(ln[k]) > 1e-6:  # skip slivers at faces, where rounding decides
                assert rg[k] == ((iy * g.nx + ix) * 2 + outside) * g.nz + layer, (t, k)
                assert abs(zz[k] - (pz - layer * dzl) / dzl) < 1e-5
            s += float(ln[k])


def test_trace_pin_fraction_is_the_volume_fraction():
    g = _geometry(nx=10, ny=10, nz=3, height=12.6)
    off, ln, rg, zz = pyoracle.trace(g, 3, 3000)
    inside = ((rg // g.nz) % 2) == 0
    frac = ln[inside].astype(np.float64).sum() / ln.astype(np.float64).sum()
    want = np.pi * g.pin_radius ** 2 / g.pitch ** 2
    assert abs(frac - want) < 0.02, (frac, want)


def test_glibc_expf_within_one_ulp():
    """The oracle's native exponential is glibc's expf (src/cpu/kernel.c:221
    calls expf).  Pinned against FP64 exp: at most 1 ulp over tau in [0, 20]."""
    import ctypes
    import ctypes.util
    libm = ctypes.CDLL(ctypes.util.find_library("m"))
    libm.expf.argtypes = [ctypes.c_float]
    libm.expf.restype = ctypes.c_float
    xs = np.linspace(-20.0, 0.0, 4001).astype(np.float32)
    got = np.array([libm.expf(float(x)) for x in xs], np.float32)
    want = np.exp(xs.astype(np.float64))
    ulp = np.spacing(want.astype(np.float32)).astype(np.float64)
    assert np.max(np.abs(got.astype(np.float64) - want) / ulp) <= 1.0
