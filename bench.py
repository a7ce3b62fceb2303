#!/usr/bin/env python3
"""Benchmark of the segment-attenuation sweep on one MI355X.

Workload (the metric's configuration): 50,000,000 segments x 100 energy
groups, 6750 source regions x 5 fine axial intervals, tracks of 100
segments, FP32 storage and arithmetic, atomic FP32 fine_flux tally -- the
plain kernel at the reference's default scale.  A step is one smoc_run over
the whole stream (fine_flux reset + sweep); inputs are resident in HBM
before timing starts.  `value` is segment-group attenuations/s of that step.

Also measured, same shape (each line its own context, autotuned where it
says so, timed over the same number of steps):
  * `mixed`: the batched mixed-precision path -- the stream pre-sorted by
    (region, fine interval) and cut into tracks of 100 consecutive sorted
    segments, FP32 storage, FP32 math, runs of a track in one (region,
    interval) summed in registers, deterministic FP64 segmented reduction,
    launched as the autotuner picks it (during the warm-up steps) and
    replayed from a captured hipGraph; it meets the 1e-5 bar against FP64 on
    the FP32 masters.  `mixed_bf16_storage` / `mixed_f16_storage` are the
    same path with 16-bit sigT / fine_source (their errors against both
    oracle definitions are in `accuracy`);
  * attribution of the mixed-over-plain ratio, each on one change at a time:
    `plain_sorted_segment_flush` is the plain kernel (FP32, per-segment FP32
    atomics) on the sorted stream -- the stream order alone;
    `plain_sorted` adds the run sums (one flush per run, atomic FP32);
    `mixed` then swaps in the deterministic FP64 tally, autotuning and the
    graph; `mixed_track_order` is the mixed configuration on the plain
    problem's own track-order stream (every run one segment: per-segment FP32
    partials and the fixed-order reduction) -- the mixed kernel without
    sorting.  `sort` is the host time of the stable counting sort that
    produced the sorted stream, outside every timed region;
  * `roofline`: algorithmic bytes per segment-group of the dominant kernel
    (k_sweep) over its mean duration from HIP events on the library stream;
    `traffic` comes from the calibrated rocprofv3 FETCH_SIZE / WRITE_SIZE
    summary committed under profiles/ for this workload;
    `atomic_tally_floor` times the same stream's per-segment FP32 atomic
    flushes alone (k_tally_probe: no gathers, no arithmetic), the part of
    the sweep that no data layout removes;
  * `mixed.roofline`: the sorted sweep is bound by VALU issue, so its figure
    is VALU lane-operations per second (the committed profile's
    SQ_INSTS_VALU per launch x 64 over the step time) against the FP32 issue
    ceiling of 7.86e13 (157.3 TFLOP/s of FMAs), with `useful_frac` counting
    only the lanes that carry one of the G groups;
  * `batched_plain` / `batched_mixed`: 8192 track batches of lognormal length
    (median 4000, sigma 1.0, clipped to [16, 200000], about 5.3e7 segments),
    the plain kernel on them in generation order (atomic FP32) and the mixed
    path on the same batches sorted (deterministic FP64, long tracks in
    chunks), each autotuned and replayed from a hipGraph; `vs_plain` is the
    ratio of attenuations/s to the metric's plain line, and the batched mixed
    line's error against the FP64 oracle is in `accuracy`;
  * `handover`: the time smoc_set_regions + smoc_set_segments take to bring
    the host buffers in, and the rate of one step including it;
  * `cpu_baseline`: the reference's own attenuate_segment (oracle/_ref) timed
    on this host's cores on a bounded sample of the same stream (it fixes each
    segment's length and axial position at its constants), with
    `same_stream_port`: the C restatement on the same tracks with the stream's
    own geometry; plus the accuracy of the GPU paths against the FP64 oracle
    (the checker).
"""
import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "simplemoc-kernel_amd"))

import numpy as np  # noqa: E402

import smoc  # noqa: E402

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, 8 TB/s
TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "k_sweep_plain_50Mx100_traffic.txt")
MIXED_SUMMARY = os.path.join(ROOT, "profiles", "k_sweep_mixed_sorted_f32_50Mx100.txt")
# FP32 VALU issue ceiling: 157.3 TFLOP/s = 7.86e13 FMA lane-operations/s, which gfx950 reaches by
# issuing two VALU instructions per SIMD per quad-cycle (packed FP32 is not used here)
VALU_PEAK_LANE_OPS = 157.3e12 / 2


def log(msg: str):
    """Progress on stderr (stdout carries only the JSON line)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def alg_bytes_per_sg(storage: str, tally: str, G: int, L: float) -> float:
    """Reference-formulation bytes per segment-group: sigT + the 3-point source
    stencil + tally read and write + the segment stream (12 B per segment,
    shared by G groups) + psi in and out (8 B per track-group, shared by the
    L segments of a track)."""
    b = 4 if storage == "f32" else 2
    bt = {"atomic-f32": 4, "atomic-f64": 8, "det-f64": 8, "det-kahan": 8, "fixed64": 8}[tally]
    return b + 3 * b + 2 * bt + 12.0 / G + 8.0 / L


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--segments", type=int, default=50_000_000)
    ap.add_argument("--groups", type=int, default=100)
    ap.add_argument("--track-len", type=int, default=100)
    ap.add_argument("--storage", default="f32", choices=sorted(smoc.STORAGE))
    ap.add_argument("--exp", default="accurate", choices=sorted(smoc.EXP))
    ap.add_argument("--tally", default="atomic-f32", choices=sorted(smoc.TALLY))
    ap.add_argument("--gpl", type=int, default=0)
    ap.add_argument("--wpb", type=int, default=0)
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--cpu-sample-segments", type=int, default=16_000_000)
    ap.add_argument("--cpu-single-segments", type=int, default=1_000_000,
                    help="sample for the one-thread CPU run")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run the plain copy kernel over this many MiB (counter "
                         "calibration) and print its bandwidth")
    ap.add_argument("--copy-variant", type=int, default=0,
                    help="copy kernel variant for --copy-probe-mib (-1: every variant)")
    return ap.parse_args()


def timed_steps(ctx, steps, warmup, launch):
    for _ in range(warmup):
        ctx.run(**launch)
    sweeps, totals = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        tm = ctx.run(**launch)  # returns after the run's last event has completed
        sweeps.append(tm.sweep_ms)
        totals.append(tm.total_ms)
    wall = time.perf_counter() - t0
    return wall, float(np.mean(sweeps)), float(np.mean(totals)), tm.used.as_dict()


def measured_traffic():
    """HBM bytes per k_sweep launch from the committed rocprofv3 summary:
    FETCH_SIZE x its copy-kernel calibration factor + WRITE_SIZE x its."""
    try:
        with open(TRAFFIC_SUMMARY) as fh:
            txt = fh.read()
    except OSError:
        return None
    m = re.search(r"traffic_bytes_per_launch\s*=\s*([0-9.eE+]+)", txt)
    return float(m.group(1)) if m else None


def mixed_valu_instructions():
    """VALU instructions per k_sweep launch of the mixed workload from the
    committed rocprofv3 summary (SQ_INSTS_VALU; fixed for a given input)."""
    try:
        with open(MIXED_SUMMARY) as fh:
            txt = fh.read()
    except OSError:
        return None
    m = re.search(r"k_sweep\s+SQ_INSTS_VALU\s+([0-9.eE+]+)", txt)
    return float(m.group(1)) if m else None


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_leg(prob, sample_segments, single_segments, gpu_plain, mixed, accuracy, batched=None):
    """Reference CPU baseline on a bounded sample (all host threads, then one
    thread on a smaller sample), and the checker: FP64 oracle on the full
    stream for the accuracy of the GPU results (batched: the sorted batches
    and the batched mixed line's results, checked the same way)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count() or 1
    threads = min(threads, os.cpu_count() or threads)
    use_ref = pyoracle.ref_available("exp")
    ref = None
    if use_ref:
        ref = pyoracle.Ref("exp")
        ref.init(prob.G, prob.R, prob.F)
        ref.set_regions(prob.sigT, prob.fine_source)

    def timed(segments, nthreads, port=False):
        T = int(np.searchsorted(prob.track_off, segments, side="right")) - 1
        T = max(1, min(T, prob.T))
        S = int(prob.track_off[T])
        F = prob.F
        if use_ref and not port:
            zf = (prob.z[:S] * np.float32(F)).astype(np.float32)
            fai = np.clip(zf.astype(np.int32), 0, F - 1).astype(np.int32)
            sec, _ = ref.sweep(prob.track_off[: T + 1], prob.region[:S], fai, prob.psi_in[:T],
                               nthreads)
        else:
            sub = smoc.Problem(prob.G, prob.R, prob.F, prob.sigT, prob.fine_source,
                               prob.len[:S], prob.region[:S], prob.z[:S],
                               prob.track_off[: T + 1], prob.psi_in[:T])
            sec, _, _ = pyoracle.baseline_f32(sub, pyoracle.EXP_NATIVE, nthreads)
        return T, S, sec

    if use_ref:
        kind, what = "reference", ("the reference's attenuate_segment (src/cpu/kernel.c, "
                                   "-Ofast -fopenmp, dynamic schedule over tracks, per-(region, "
                                   "interval) locks); ds and zin fixed at its constants")
    else:
        kind, what = "port", "the C restatement of attenuate_segment, OpenMP over tracks"
    T, S, sec = timed(sample_segments, threads)
    T1, S1, sec1 = timed(single_segments, 1)
    base = {"value": S * prob.G / sec, "unit": "segment-group attenuations/s", "cores": threads,
            "kind": kind,
            "sample": f"first {T} tracks ({S} segments x {prob.G} groups) of the benchmark "
                      f"stream, {what}; {sec:.3f} s wall on {threads} threads "
                      f"({sec * threads:.1f} thread-seconds)",
            "cpu_model": cpu_model(),
            "single_thread": {"value": S1 * prob.G / sec1, "cores": 1,
                              "sample": f"first {T1} tracks ({S1} segments), same code, "
                                        f"{sec1:.3f} s"}}
    if use_ref:
        # the same tracks through the C restatement, which takes each segment's
        # own length and axial position (the reference fixes them)
        Tp, Sp, secp = timed(sample_segments, threads, port=True)
        base["same_stream_port"] = {
            "value": Sp * prob.G / secp, "cores": threads, "kind": "port",
            "sample": f"first {Tp} tracks ({Sp} segments), the C restatement of "
                      f"attenuate_segment (oracle_baseline_f32, OpenMP over tracks) with the "
                      f"stream's own lengths and axial positions; {secp:.3f} s"}
    acc = None
    if accuracy:
        acc = {}
        t0 = time.time()
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from conftest import quantized

        def errs(f, p, of, op):
            return {"fine_flux_checksum_rel": abs(f.sum() - of.sum()) / abs(of.sum()),
                    "state_flux_checksum_rel":
                        abs(p.astype(np.float64).sum() - op.sum()) / abs(op.sum()),
                    "fine_flux_normwise": float(np.linalg.norm(f - of) / np.linalg.norm(of)),
                    "state_flux_normwise": float(np.linalg.norm(p.astype(np.float64) - op) /
                                                 np.linalg.norm(op))}

        of, op = pyoracle.sweep_f64(prob, pyoracle.EXP_NATIVE, nthreads=threads)
        f, p = gpu_plain
        acc["plain_vs_fp64"] = errs(f, p, of, op)
        if mixed is not None:
            if mixed.get("track_order") is not None:
                f, p = mixed["track_order"]
                acc["mixed_track_order_vs_fp64"] = errs(f, p, of, op)
            mprob = mixed["prob"]
            of, op = pyoracle.sweep_f64(mprob, pyoracle.EXP_NATIVE, nthreads=threads)
            for name, (f, p, storage) in mixed["sorted"].items():
                acc[f"{name}_vs_fp64_on_fp32_masters"] = errs(f, p, of, op)
            for name, (f, p, storage) in mixed["sorted"].items():
                if storage != "f32":
                    qf, qp = pyoracle.sweep_f64(quantized(mprob, storage), pyoracle.EXP_NATIVE,
                                                nthreads=threads)
                    acc[f"{name}_vs_fp64_on_dequantised_inputs"] = errs(f, p, qf, qp)
        if batched is not None:
            bprob, (f, p) = batched
            of, op = pyoracle.sweep_f64(bprob, pyoracle.EXP_NATIVE, nthreads=threads)
            acc["batched_mixed_vs_fp64"] = errs(f, p, of, op)
        acc["oracle_s"] = time.time() - t0
    return base, acc


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU; --gpus must be 1 (a "
                          "multi-GPU run would shard the tracks, one context per GPU, and "
                          "all-reduce fine_flux; that path is not built)"}))
        sys.exit(2)
    if args.copy_probe_mib:
        nbytes = args.copy_probe_mib << 20
        for var in ([args.copy_variant] if args.copy_variant >= 0 else range(4)):
            ms = smoc.probe_copy(nbytes, reps=5, variant=var)
            print(json.dumps({"copy_probe_bytes": nbytes, "variant": var, "mean_ms": ms,
                              "gbs_read_plus_write": 2 * nbytes / ms / 1e6}), flush=True)
        return
    G, S, L = args.groups, args.segments, args.track_len
    R, F = 6750, 5
    t0 = time.time()
    prob = smoc.generate(S, G=G, R=R, F=F, seed=42, track_len=L)
    gen_s = time.time() - t0
    dev_name, cus = smoc.device_info(0)
    launch = {}
    if args.gpl:
        launch["groups_per_lane"] = args.gpl
    if args.wpb:
        launch["waves_per_block"] = args.wpb

    log(f"plain: {S} x {G}, {args.storage} storage, {args.tally}")
    with smoc.Context(G, R, F, storage=args.storage, exp=args.exp, tally=args.tally) as ctx:
        t0 = time.perf_counter()
        ctx.load(prob)  # host buffers -> HBM, stream decode; returns synchronised
        load_s = time.perf_counter() - t0
        wall, sweep_ms, total_ms, used = timed_steps(ctx, args.steps, args.warmup, launch)
        flux, psi = ctx.results()
    ck_f, ck_p = smoc.checksums(flux, psi)
    sg = S * G
    value = sg / (wall / args.steps)
    B = alg_bytes_per_sg(args.storage, args.tally, G, L)
    achieved = B * sg / (sweep_ms * 1e-3) / 1e9
    traffic = measured_traffic() if (args.storage, args.tally, S, G) == \
        ("f32", "atomic-f32", 50_000_000, 100) else None

    floor = None
    if args.tally == "atomic-f32":
        # the same stream's per-segment FP32 atomic flushes alone: the floor
        # that the tally sets for k_sweep on a stream in generation order
        fms = smoc.probe_tally(S, R * F, G, reps=5)
        floor = {"kernel": "k_tally_probe", "ms": fms,
                 "atomic_requests_64B": S * -(-G // 16),
                 "requests_per_s": S * -(-G // 16) / (fms * 1e-3),
                 "sweep_over_floor": sweep_ms / fms}

    out = {
        "metric": "segment-group attenuations/s",
        "value": value,
        "unit": "segment-group attenuations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.storage == "f32" else f"{args.storage}-storage/f32-math",
        "data": "synthetic (seeded counter-based generator, seed 42)",
        "config": {"workload": f"plain sweep: {S} segments x {G} energy groups, {R} source "
                               f"regions x {F} fine axial intervals, tracks of {L} segments "
                               f"in generation order, atomic fine_flux tally",
                   "segments": S, "groups": G, "regions": R, "fine_axial": F,
                   "track_len": L, "storage": args.storage, "exp": args.exp,
                   "tally": args.tally, "launch": used},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_sweep", "kernel_ms": sweep_ms, "alg_bytes_per_sg": B,
                     "atomic_tally_floor": floor},
        "hbm_gbs": achieved,
        "handover": {
            "load_ms": load_s * 1e3,
            "host_bytes": int(sum(a.nbytes for a in (prob.sigT, prob.fine_source, prob.len,
                                                     prob.region, prob.z, prob.track_off,
                                                     prob.psi_in))),
            "value_incl_load": sg / (load_s + wall / args.steps),
            "note": "one upload of pageable host buffers (PCIe), the stream decode and the "
                    "host-side track volumes, then one step; not the metric"},
        "checksums": {"fine_flux": ck_f, "state_flux": ck_p},
        "device": f"{dev_name}, {cus} CUs",
        "input_gen_s": gen_s,
    }

    mixed = batched = None
    if not args.no_mixed:
        tuned = {"autotune": 1, "use_graph": 1}
        mwarm = max(args.warmup, 1)

        def line(problem, storage, tally, launch, workload):
            """One timed configuration on its own context: the JSON entry and
            the results (fine_flux, psi_out)."""
            log(f"{storage} storage, {tally}, launch {launch}")
            with smoc.Context(G, R, F, storage=storage, exp=args.exp, tally=tally) as ctx:
                ctx.load(problem)
                lwall, lsweep, ltotal, lused = timed_steps(ctx, args.steps, mwarm, launch)
                lf, lp = ctx.results()
            v = sg / (lwall / args.steps)
            entry = {"workload": workload, "value": v, "ms_per_step": lwall / args.steps * 1e3,
                     "vs_plain": v / value, "sweep_ms": lsweep, "total_ms": ltotal,
                     "alg_bytes_per_sg": alg_bytes_per_sg(storage, tally, G, L),
                     "storage": storage, "tally": tally, "launch": lused,
                     "checksums": dict(zip(("fine_flux", "state_flux"),
                                           smoc.checksums(lf, lp)))}
            return entry, lf, lp

        mprob, sort_s = smoc.sort_stream(prob)
        out["sort"] = {"host_s": sort_s, "what": f"smoc_sort_segments of the {S}-segment "
                       "stream by (region, fine interval) on one host thread (stable counting "
                       "sort); tracks stay the same index ranges, so a sorted track is "
                       f"{L} consecutive sorted segments; outside every timed region"}
        sorted_runs = {}
        out["mixed"], mf, mp = line(
            mprob, "f32", "det-f64", tuned,
            f"same shape; stream pre-sorted by (region, fine interval), tracks of {L} "
            "consecutive sorted segments; FP32 storage and math, runs summed in registers, "
            "deterministic FP64 segmented reduction; autotuned launch replayed from a hipGraph")
        sorted_runs["mixed"] = (mf, mp, "f32")
        mused = out["mixed"]["launch"]
        mwall_step = out["mixed"]["ms_per_step"] * 1e-3
        # (the profile's launch has 2 groups per lane; the count holds for that shape)
        valu = mixed_valu_instructions() if ((S, G, L) == (50_000_000, 100, 100) and
                                             mused.get("groups_per_lane") == 2) else None
        if valu is not None:
            # lane-operations per second over the whole step (reset + sweep + reduce): a
            # lower bound for k_sweep's own rate; useful: the G of the 128 group slots a
            # wave carries at 2 groups per lane
            ach = valu * 64 / mwall_step
            slots = -(-G // 128) * 128
            out["mixed"]["roofline"] = {
                "bound": "valu", "achieved": ach, "peak": VALU_PEAK_LANE_OPS,
                "unit": "VALU lane-ops/s", "frac": ach / VALU_PEAK_LANE_OPS,
                "useful_frac": ach * G / slots / VALU_PEAK_LANE_OPS,
                "kernel": "k_sweep", "valu_instructions_per_launch": valu,
                "source": os.path.relpath(MIXED_SUMMARY, ROOT)}
        for st in ("bf16", "f16"):
            key = f"mixed_{st}_storage"
            out[key], f16f, f16p = line(mprob, st, "det-f64", tuned,
                                        f"as mixed, with {st} storage of sigT / fine_source")
            sorted_runs[key] = (f16f, f16p, st)
        out["plain_sorted"], f_, p_ = line(
            mprob, "f32", "atomic-f32", {},
            "the plain configuration (FP32 storage, atomic FP32 tally, default launch) on the "
            "sorted stream: runs summed in registers, one atomic flush per run")
        sorted_runs["plain_sorted"] = (f_, p_, "f32")
        out["plain_sorted_segment_flush"], f_, p_ = line(
            mprob, "f32", "atomic-f32", {"segment_flush": 1},
            "the plain kernel itself on the sorted stream: one atomic flush per segment "
            "(segment_flush), as on the track-order stream")
        sorted_runs["plain_sorted_segment_flush"] = (f_, p_, "f32")
        out["mixed_track_order"], tf, tp = line(
            prob, "f32", "det-f64", tuned,
            "the mixed configuration (FP32 storage, deterministic FP64 tally, autotuned graph) "
            "on the plain problem's own track-order stream: every run is one segment, so "
            "per-segment FP32 partials and the fixed-order segmented reduction")
        v = {k: out[k]["value"] for k in ("plain_sorted_segment_flush", "plain_sorted", "mixed",
                                          "mixed_track_order")}
        out["attribution"] = {
            "stream_order_alone": v["plain_sorted_segment_flush"] / value,
            "run_sums_on_sorted_stream": v["plain_sorted"] / v["plain_sorted_segment_flush"],
            "det_tally_graph_autotune": v["mixed"] / v["plain_sorted"],
            "mixed_over_plain": v["mixed"] / value,
            "mixed_without_sorting": v["mixed_track_order"] / value,
            "note": "ratios of attenuations/s, each changing one thing: plain -> plain kernel "
                    "on the sorted stream -> run sums -> deterministic FP64 tally + autotuned "
                    "graph (= mixed); and the mixed configuration on the unsorted stream"}
        mixed = {"prob": mprob, "sorted": sorted_runs, "track_order": (tf, tp)}

    if not args.no_batched:
        # 8192 track batches of lognormal length (median 4000, sigma 1.0, clipped to
        # [16, 200000]), about 5.3e7 segments: the plain kernel on them as generated
        # and the mixed path on the same batches sorted, each autotuned (chunk length,
        # persistence, groups per lane, waves per workgroup) and replayed from a graph
        t0 = time.time()
        bprob = smoc.generate(0, G=G, R=R, F=F, seed=42, lognormal_tracks=8192,
                              median=4000.0, sigma=1.0, lo=16, hi=200_000)
        bsorted, bsort_s = smoc.sort_stream(bprob)
        bgen_s = time.time() - t0
        tuned = {"autotune": 1, "use_graph": 1}
        bsg = bprob.S * G

        def bline(problem, tally, workload):
            log(f"batched: {problem.S} segments in {problem.T} tracks, {tally}")
            with smoc.Context(G, R, F, storage="f32", exp=args.exp, tally=tally) as ctx:
                ctx.load(problem)
                lwall, lsweep, ltotal, lused = timed_steps(ctx, args.steps,
                                                           max(args.warmup, 1), tuned)
                lf, lp = ctx.results()
            v = bsg / (lwall / args.steps)
            return {"workload": workload, "value": v, "ms_per_step": lwall / args.steps * 1e3,
                    "vs_plain": v / value, "sweep_ms": lsweep, "total_ms": ltotal,
                    "segments": problem.S, "tracks": problem.T, "tally": tally,
                    "launch": lused,
                    "checksums": dict(zip(("fine_flux", "state_flux"),
                                          smoc.checksums(lf, lp)))}, lf, lp

        out["batched_plain"], _, _ = bline(
            bprob, "atomic-f32",
            f"8192 lognormal track batches ({bprob.S} segments) x {G} groups, FP32, tracks "
            "in generation order, atomic FP32 tally; autotuned launch replayed from a hipGraph")
        out["batched_mixed"], bf, bp = bline(
            bsorted, "det-f64",
            "the same batches, stream pre-sorted by (region, fine interval); FP32 storage and "
            "math, runs summed in registers, deterministic FP64 segmented reduction; long "
            "tracks in chunks swept once each (linear in their incoming psi, corrected per "
            "run); autotuned launch replayed from a hipGraph")
        out["batched_mixed"]["vs_batched_plain"] = (out["batched_mixed"]["value"] /
                                                    out["batched_plain"]["value"])
        out["batched_input"] = {"generate_and_sort_s": bgen_s, "sort_host_s": bsort_s,
                                "note": "outside every timed region"}
        batched = (bsorted, (bf, bp))

    if not args.no_cpu_baseline:
        log("CPU baseline and FP64 oracle")
        base, acc = cpu_leg(prob, args.cpu_sample_segments, args.cpu_single_segments,
                            (flux, psi), mixed,
                            not args.no_accuracy, batched)
        out["cpu_baseline"] = base
        if acc is not None:
            out["accuracy"] = acc
    print(json.dumps(out))


if __name__ == "__main__":
    main()
