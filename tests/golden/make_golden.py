#!/usr/bin/env python3
"""Writes tests/golden/ref_vectors.npz from the reference's own CPU code.

Runs only where oracle/_ref has been built from the reference sources
(`make -C oracle ref`); the fixtures it writes are data -- inputs and the
reference's outputs -- and travel with the repository.

Contents:
  table_values/N/dx/maxVal  buildExponentialTable(0.01, 10) (src/cpu/init.c:81-117)
  interp_x / interp_y       interpolateTable at probe points (src/cpu/kernel.c:337-361)
  seg_*                     one carried state_flux vector through 64 segments of
                            attenuate_segment (src/cpu/kernel.c:75-333) with the
                            reference constants ds = 0.7, zin = 0.3, for the
                            exp and table builds, strict (-O2) and reference
                            (-Ofast) flags: psi_out_<variant>, flux_<variant>.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import pyoracle  # noqa: E402


def main():
    for v in pyoracle.REF_VARIANTS:
        if not pyoracle.ref_available(v):
            sys.exit(f"oracle/_ref/libsmocref_{v}.so missing: run `make -C oracle ref` first")
    rng = np.random.default_rng(20240601)
    G, R, F, n = 100, 16, 5, 64
    sig = rng.uniform(0.1, 10.0, (R, G)).astype(np.float32)
    src = rng.uniform(0.0, 1.0, (R, F, G)).astype(np.float32)
    flux0 = rng.uniform(0.0, 1.0, (R, F, G)).astype(np.float32)
    region = rng.integers(0, R, n).astype(np.int32)
    fai = rng.integers(0, F, n).astype(np.int32)
    fai[:5] = np.arange(5)  # every interval case appears
    psi0 = rng.uniform(0.0, 1.0, G).astype(np.float32)
    out = dict(seg_G=G, seg_R=R, seg_F=F, seg_sigT=sig, seg_fine_source=src,
               seg_fine_flux_in=flux0, seg_region=region, seg_fai=fai, seg_psi_in=psi0)
    for v in pyoracle.REF_VARIANTS:
        ref = pyoracle.Ref(v)
        ref.init(G, R, F)
        ref.set_regions(sig, src, flux0)
        out[f"psi_out_{v}"] = ref.run_segments(region, fai, psi0)
        out[f"flux_{v}"] = ref.flux()
    ref = pyoracle.Ref("table_strict")
    ref.init(G, R, F)
    vals, dx, mx = ref.table()
    xs = np.concatenate([np.array([0.0, 0.05, 0.3, 0.7, 1.0, 5.0, 9.9, 9.97, 9.98, 12.0],
                                  np.float32),
                         rng.uniform(0.0, 10.5, 200).astype(np.float32)]).astype(np.float32)
    ys = np.array([ref.interp(float(x)) for x in xs], np.float32)
    out.update(table_values=vals, table_N=vals.size // 2, table_dx=np.float32(dx),
               table_maxVal=np.float32(mx), interp_x=xs, interp_y=ys)
    path = os.path.join(HERE, "ref_vectors.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path}: {sorted(out)}")


if __name__ == "__main__":
    main()
