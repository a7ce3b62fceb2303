"""Host-side logic of libsmoc (no GPU): reference defaults, the exponential
table, the z decode, the seeded generator (against an independent numpy
restatement of its definition), track builders, the sort, checksums."""
import ctypes
import os

import numpy as np
import pytest

import pyoracle
import smoc

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = np.load(os.path.join(HERE, "golden", "ref_vectors.npz"))

U64 = np.uint64
GOLDEN = U64(0x9E3779B97F4A7C15)


def mix64(z):
    z = np.asarray(z, np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> U64(30))) * U64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> U64(27))) * U64(0x94D049BB133111EB)
    return z ^ (z >> U64(31))


def key(seed, stream):
    with np.errstate(over="ignore"):
        return mix64(np.array([seed], np.uint64) * GOLDEN + U64(stream))[0]


def draws(k, n, start=0):
    i = np.arange(start, start + n, dtype=np.uint64)
    with np.errstate(over="ignore"):
        return mix64(U64(k) + (i + U64(1)) * GOLDEN)


def u01(x):
    return (x >> U64(40)).astype(np.float32) * np.float32(2.0**-24)


def uab(x, a, b):
    a, b = np.float32(a), np.float32(b)
    return (a + (b - a) * u01(x)).astype(np.float32)


def test_default_input_and_derived_regions():
    inp = smoc.default_input()
    assert (inp.source_2D_regions, inp.coarse_axial_intervals, inp.fine_axial_intervals,
            inp.decomp_assemblies_ax, inp.segments, inp.egroups, inp.seg_per_block) == \
        (5000, 27, 5, 20, 50_000_000, 128, 100)
    assert inp.source_3D_regions == 6750  # ceil(5000 * 27 / 20)
    inp.decomp_assemblies_ax = 7
    smoc.lib().smoc_derive_regions(ctypes.byref(inp))
    assert inp.source_3D_regions == int(np.ceil(5000 * 27 / 7))


def test_product_table_equals_reference_fixture():
    vals, dx, mx = smoc.exp_table(0.01, 10.0)
    assert np.array_equal(vals.view(np.uint32), GOLD["table_values"].view(np.uint32))
    assert np.float32(dx) == GOLD["table_dx"] and np.float32(mx) == GOLD["table_maxVal"]
    L = smoc.lib()
    ys = np.array([L.smoc_exp_table_lookup(vals, dx, mx, float(x)) for x in GOLD["interp_x"]],
                  np.float32)
    assert np.array_equal(ys.view(np.uint32), GOLD["interp_y"].view(np.uint32))


def test_decode_z_matches_oracle_and_numpy():
    rng = np.random.default_rng(0)
    zs = np.concatenate([rng.random(2000, dtype=np.float32),
                         np.array([0.0, 0.2, 0.4, 0.6, 0.8, 0.99999994, 1.0], np.float32),
                         np.nextafter(np.float32([0.2, 0.4, 0.6, 0.8]), np.float32(0))])
    for F in (2, 5, 7):
        for z in zs:
            f, zin = smoc.decode_z(float(z), F)
            assert (f, zin) == pyoracle.decode_z(float(z), F)
            zf = np.float32(z) * np.float32(F)
            fi = min(max(int(zf), 0), F - 1)
            assert f == fi
            assert np.float32(zin) == np.float32((zf - np.float32(fi)) - np.float32(0.5)) * \
                np.float32(0.1)
            assert -0.05 - 1e-7 <= zin <= 0.05 + 1e-6


def test_generator_matches_restatement():
    seed, G, R, F, S = 1234, 7, 11, 5, 5000
    p = smoc.generate(S, G=G, R=R, F=F, seed=seed, track_len=100)
    ks, kq = key(seed, 1), key(seed, 2)
    assert np.array_equal(p.sigT.ravel(), uab(draws(ks, R * G), 0.1, 10.0))
    assert np.array_equal(p.fine_source.ravel(), u01(draws(kq, R * F * G)))
    assert np.array_equal(p.len, uab(draws(key(seed, 3), S), 0.0, 0.1))
    reg = ((draws(key(seed, 4), S) >> U64(32)) * U64(R)) >> U64(32)
    assert np.array_equal(p.region, reg.astype(np.int32))
    assert np.array_equal(p.z, u01(draws(key(seed, 5), S)))
    assert np.array_equal(p.psi_in.ravel(), u01(draws(key(seed, 6), p.T * G)))


def test_generator_distributions():
    p = smoc.generate(400_000, G=100, R=6750, F=5, seed=42)
    assert 0.1 <= p.sigT.min() and p.sigT.max() < 10.0
    assert abs(p.sigT.mean() - 5.05) < 0.02
    assert 0.0 <= p.len.min() and p.len.max() < 0.1 and abs(p.len.mean() - 0.05) < 5e-4
    assert p.region.min() >= 0 and p.region.max() < 6750
    counts = np.bincount(p.region, minlength=6750)
    assert counts.min() > 20 and counts.max() < 110
    assert 0.0 <= p.z.min() and p.z.max() < 1.0 and abs(p.z.mean() - 0.5) < 3e-3
    assert abs(p.psi_in.mean() - 0.5) < 3e-3


def test_generator_thread_independent_and_prefix_stable():
    a = smoc.generate(10_000, G=10, R=100, seed=9)
    b = smoc.generate(20_000, G=10, R=100, seed=9)
    # element i depends only on (seed, array, i)
    assert np.array_equal(a.len, b.len[:10_000]) and np.array_equal(a.z, b.z[:10_000])
    assert np.array_equal(a.region, b.region[:10_000])
    c = smoc.generate(10_000, G=10, R=100, seed=10)
    assert not np.array_equal(a.len, c.len)


def test_tracks_fixed_tail():
    p = smoc.generate(1_000_001, G=4, R=10, seed=1, track_len=512)
    assert p.T == 1954
    assert p.track_off[-1] == 1_000_001 and p.track_off[-2] == 1953 * 512
    assert np.all(np.diff(p.track_off)[:-1] == 512) and p.track_off[-1] - p.track_off[-2] == 65


def test_tracks_lognormal():
    T = 8192
    p = smoc.generate(0, G=4, R=10, seed=42, lognormal_tracks=T)
    lens = np.diff(p.track_off)
    assert lens.size == T and lens.min() >= 16 and lens.max() <= 200_000
    assert abs(np.median(lens) - 4000) / 4000 < 0.05
    assert 45e6 < p.S < 60e6  # mean exp(mu + sigma^2/2) = 6595 per track
    # Box-Muller restatement of the same draws
    k = key(42, 7)
    x1, x2 = draws(k, T, 0)[::1], None
    d = draws(k, 2 * T)
    u1 = ((d[0::2] >> U64(11)) + U64(1)).astype(np.float64) * 2.0**-53
    u2 = (d[1::2] >> U64(11)).astype(np.float64) * 2.0**-53
    n = np.sqrt(-2.0 * np.log(u1)) * np.cos(6.283185307179586 * u2)
    ref = np.clip(np.floor(np.exp(np.log(4000.0) + n) + 0.5), 16, 200_000)
    assert np.abs(lens - ref).max() <= 1
    assert (lens == ref).mean() > 0.999
    del x1, x2


def test_sort_segments_stable_by_key():
    p = smoc.generate(50_000, G=4, R=300, F=5, seed=3)
    q = smoc.generate(50_000, G=4, R=300, F=5, seed=3, sorted_stream=True)
    F = 5
    key_of = lambda pr: pr.region.astype(np.int64) * F + np.clip(
        (pr.z * np.float32(F)).astype(np.int32), 0, F - 1)
    kp, kq = key_of(p), key_of(q)
    assert np.all(np.diff(kq) >= 0)
    order = np.argsort(kp, kind="stable")
    assert np.array_equal(q.len, p.len[order]) and np.array_equal(q.z, p.z[order])
    assert np.array_equal(q.region, p.region[order])


def test_checksums_fixed_order():
    rng = np.random.default_rng(0)
    x = rng.standard_normal(100_001)
    L = smoc.lib()
    s = 0.0
    for v in x:
        s += v
    assert L.smoc_checksum_f64(x, x.size) == s
    xf = x.astype(np.float32)
    s32 = 0.0
    for v in xf:
        s32 += float(v)
    assert L.smoc_checksum_f32(xf, xf.size) == s32


def test_generator_rejects_bad_arguments():
    L = smoc.lib()
    buf = np.zeros(10, np.float32)
    assert L.smoc_generate_regions(1, 0, 1, 1, buf, buf) == smoc.ERR_ARG
    off = np.zeros(3, np.int64)
    t = ctypes.c_int64()
    assert L.smoc_tracks_fixed(10, 0, off, ctypes.byref(t)) == smoc.ERR_ARG
    assert b"smoc_tracks_fixed" in L.smoc_last_error()


def test_generate_scatter():
    a = smoc.generate_scatter(40, 16, 0.7, seed=9)
    b = smoc.generate_scatter(40, 16, 0.7, seed=9)
    c = smoc.generate_scatter(40, 16, 0.7, seed=10)
    assert a.shape == (40, 16) and np.array_equal(a, b) and not np.array_equal(a, c)
    assert a.min() >= 0 and a.max() < 0.7 and abs(a.mean() - 0.35) < 0.03
    for bad in (1.0, -0.1):
        with pytest.raises(smoc.SmocError):
            smoc.generate_scatter(4, 4, bad)
