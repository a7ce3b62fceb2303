#!/usr/bin/env python3
"""Benchmark of the segment-attenuation sweep on one MI355X.

Workload (default): 50,000,000 segments x 100 energy groups, 6750 source
regions x 5 fine axial intervals, tracks of 100 segments, FP32, atomic
fine_flux tally -- the plain kernel at the reference's default scale.  A
step is one smoc_run over the whole stream (fine_flux reset, sweep); inputs
are resident in HBM before timing starts.

Prints one JSON line with the metric (segment-group attenuations/s), the
roofline of the sweep kernel (algorithmic bytes per segment-group / measured
kernel time, HIP events on the library's stream) and the reference's own CPU
attenuation code timed on this host's cores on a bounded sample of the same
workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "simplemoc-kernel_amd"))

import numpy as np  # noqa: E402

import smoc  # noqa: E402

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, 8 TB/s


def alg_bytes_per_sg(storage: str, tally: str, G: int, L: float) -> float:
    """Reference-formulation bytes per segment-group: sigT + 3 source values
    + tally read and write + segment stream (12 B / G) + psi in/out (8 B /
    (G L) per group... i.e. 8 / L per segment-group)."""
    b = 4 if storage == "f32" else 2
    bt = {"atomic-f32": 4, "atomic-f64": 8, "det-f64": 8, "det-kahan": 8}[tally]
    return b + 3 * b + 2 * bt + 12.0 / G + 8.0 / L


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--segments", type=int, default=50_000_000)
    ap.add_argument("--groups", type=int, default=100)
    ap.add_argument("--track-len", type=int, default=100)
    ap.add_argument("--storage", default="f32")
    ap.add_argument("--exp", default="accurate")
    ap.add_argument("--tally", default="atomic-f32")
    ap.add_argument("--gpl", type=int, default=0)
    ap.add_argument("--wpb", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-segments", type=int, default=4_000_000)
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run the plain copy kernel over this many MiB (counter "
                         "calibration) and print its bandwidth")
    return ap.parse_args()


def cpu_baseline(prob, sample_segments: int):
    """The reference's attenuate_segment (oracle/_ref, built from
    src/cpu/kernel.c + init.c with the reference's GNU flags) over the first
    tracks of the same stream, OpenMP over tracks on this host's cores."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count() or 1
    threads = min(threads, os.cpu_count() or threads)
    T = int(np.searchsorted(prob.track_off, sample_segments, side="right")) - 1
    T = max(1, min(T, prob.T))
    S = int(prob.track_off[T])
    F = prob.F
    zf = (prob.z[:S] * np.float32(F)).astype(np.float32)
    fai = np.clip(zf.astype(np.int32), 0, F - 1).astype(np.int32)
    if pyoracle.ref_available("exp"):
        ref = pyoracle.Ref("exp")
        ref.init(prob.G, prob.R, prob.F)
        ref.set_regions(prob.sigT, prob.fine_source)
        sec, _ = ref.sweep(prob.track_off[: T + 1], prob.region[:S], fai, prob.psi_in[:T], threads)
        kind = "reference"
    else:
        sub = smoc.Problem(prob.G, prob.R, prob.F, prob.sigT, prob.fine_source, prob.len[:S],
                           prob.region[:S], prob.z[:S], prob.track_off[: T + 1], prob.psi_in[:T])
        sec, _, _ = pyoracle.baseline_f32(sub, pyoracle.EXP_NATIVE, threads)
        kind = "port"
    rate = S * prob.G / sec
    return {"value": rate, "unit": "segment-group attenuations/s", "cores": threads,
            "kind": kind,
            "sample": f"first {T} tracks ({S} segments x {prob.G} groups) of the same stream; "
                      f"{'reference attenuate_segment (src/cpu/kernel.c, -Ofast -fopenmp), ds/zin fixed at its constants' if kind == 'reference' else 'C restatement, OpenMP'}; "
                      f"{sec:.3f} s wall"}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; --gpus must be 1 "
                          "(the sweep shards over tracks, see DESIGN.md)"}))
        sys.exit(2)
    if args.copy_probe_mib:
        nbytes = args.copy_probe_mib << 20
        ms = smoc.probe_copy(nbytes, reps=5)
        print(json.dumps({"copy_probe_bytes": nbytes, "mean_ms": ms,
                          "gbs_read_plus_write": 2 * nbytes / ms / 1e6}))
        return
    G, S, L = args.groups, args.segments, args.track_len
    t0 = time.time()
    prob = smoc.generate(S, G=G, R=6750, F=5, seed=42, track_len=L)
    gen_s = time.time() - t0
    dev_name, cus = smoc.device_info(0)
    launch = {}
    if args.gpl:
        launch["groups_per_lane"] = args.gpl
    if args.wpb:
        launch["waves_per_block"] = args.wpb

    with smoc.Context(G, prob.R, prob.F, storage=args.storage, exp=args.exp,
                      tally=args.tally) as ctx:
        ctx.load(prob)
        for _ in range(args.warmup):
            ctx.run(**launch)
        sweeps, totals = [], []
        t_start = time.perf_counter()
        for _ in range(args.steps):
            tm = ctx.run(**launch)  # returns after the run's last event completes
            sweeps.append(tm.sweep_ms)
            totals.append(tm.total_ms)
        wall = time.perf_counter() - t_start
        used = tm.used.as_dict()
        flux, psi = ctx.results()
    ck_f, ck_p = smoc.checksums(flux, psi)

    sg = S * G
    ms_per_step = wall / args.steps * 1e3
    value = sg / (wall / args.steps)
    sweep_ms = float(np.mean(sweeps))
    B = alg_bytes_per_sg(args.storage, args.tally, G, L)
    achieved = B * sg / (sweep_ms * 1e-3) / 1e9
    out = {
        "metric": "segment-group attenuations/s",
        "value": value,
        "unit": "segment-group attenuations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.storage == "f32" else f"{args.storage}-storage/f32-math",
        "data": "synthetic (seeded counter-based generator, seed 42)",
        "config": {"workload": f"plain sweep: {S} segments x {G} groups, 6750 regions x 5 fine "
                               f"intervals, tracks of {L} segments",
                   "segments": S, "groups": G, "regions": 6750, "fine_axial": 5,
                   "track_len": L, "storage": args.storage, "exp": args.exp,
                   "tally": args.tally, "launch": used},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_sweep", "kernel_ms": sweep_ms, "alg_bytes_per_sg": B},
        "hbm_gbs_alg": achieved,
        "checksums": {"fine_flux": ck_f, "state_flux": ck_p},
        "device": f"{dev_name}, {cus} CUs",
        "input_gen_s": gen_s,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(prob, args.cpu_sample_segments)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
