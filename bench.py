"""Benchmark: registration + dose filter + sum of 40-frame 4096x4096 movies on one MI355X.

One step = one movie through the whole alignImageStack path (forward spectra,
780-pair tri cross-correlation at fouCrop 3072, peaks, FP64 shift solver,
Lanczos shift + sums, even/odd FRC, dose filter), with the frames already
resident in HBM.  Prints ONE JSON line (rank 0).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline] [--inflight K]

The headline runs two movies in flight (two plans on two streams, each movie
still one pass through the path); single-movie time and latency are measured
first, in the same run, and reported under "single_movie".

Other workloads (each prints its own JSON line; the line above is the default):

    python bench.py --workload served    # Poisson arrivals of a 3-class movie mix, p50/p99 vs rate
    python bench.py --workload batched   # 48 uneven movies (20-60 frames, 4 detector shapes) at once
    python bench.py --workload hostinput # C2 from pinned host memory (float32 and 4-bit counts)
    python bench.py --workload configs   # the 20x1024^2 Fourier-shift movie and the 7676x7420 uint8
                                         # super-resolution counting movie (gain, binning, 0.5 A/px)
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = "movies/min (40x4096² fp32, tri-xcorr+dosefilt+sum); p50/p99 latency"
HBM_PEAK_GBS = 8000.0

WORKLOAD = dict(n=40, h=4096, w=4096, seed=2, dose=1.0, jump=3.0)
PARAMS = dict(triMode="tri", maxShift=40, Bmode="conv", Brad=150, subPixReg=16, shiftMethod="lanczos",
              weightMode="autologistic", originMode="centroid", suppressOrigin=True, fouCrop=(3072, 3072),
              shapePadded=(4096, 4096), filterMode="dose", pixelsize=0.1, voltage=300.0, doLazyFRC=True,
              frc_box=512, frc_overlap=0.5)


def k2_algorithmic_bytes(info):
    """Per launch of the pair transform: each pair reads two cropped half spectra
    and writes one 2m x 2m float32 window plus its peak record."""
    hs = info["Hc"] * (info["Wc"] // 2 + 1) * 8
    m = info["maxShift"]
    return info["pairs"] * (2 * hs + 4 * (2 * m) ** 2 + 16)


def k2_flops(info):
    """FFT-equivalent FLOPs of the pair transform per launch (the scope table's
    figure: 6 hs/8 for the conjugate products + 2.5 Hc Wc log2(Hc Wc) for the
    inverse transform, per pair)."""
    hc, wc = info["Hc"], info["Wc"]
    return info["pairs"] * (6.0 * hc * (wc // 2 + 1) + 2.5 * hc * wc * math.log2(hc * wc))


FP32_PEAK_TFLOPS = 157.3


def k2_compulsory_bytes(info):
    """Bytes the pair transform cannot avoid per launch: every cropped half
    spectrum read once, every lag window written once."""
    hs = info["Hc"] * (info["Wc"] // 2 + 1) * 8
    return info["N"] * hs + info["pairs"] * (4 * (2 * info["maxShift"]) ** 2 + 16)


def load_profile_json(name):
    """A committed rocprofv3 summary (profiles/<name>), if present."""
    p = os.path.join(HERE, "profiles", name)
    if os.path.exists(p):
        with open(p) as f:
            return json.load(f)
    return None


def load_traffic():
    """HBM bytes per K2 launch from the committed rocprofv3 PMC summary, if any."""
    d = load_profile_json("k2_pairs_traffic.json")
    return d.get("traffic_bytes") if d else None


PAIR_KERNELS = ("k2_cols", "k2_cols_r2", "k2_rows", "k2_rows_mfma")


def pmc_summary(fetch_csv, write_csv, out_json):
    """Per-kernel HBM bytes from two rocprofv3 --pmc runs (FETCH_SIZE, then
    WRITE_SIZE, KiB per dispatch, made with --copy-calibration).  The plain
    1 GiB device copy calibrates both counters; the pair transform's traffic
    per launch goes to out_json for the roofline line."""
    import csv
    import collections

    def load(path, counter):
        per = collections.defaultdict(list)
        copies = []
        for r in csv.DictReader(open(path)):
            if r["Counter_Name"] != counter:
                continue
            name = r["Kernel_Name"].split("(")[0].replace("void ", "")
            v = float(r["Counter_Value"]) * 1024.0
            per[name].append(v)
            if "copyBuffer" in name:
                copies.append(v)
        return per, sorted(copies)[-3:]

    F, fc = load(fetch_csv, "FETCH_SIZE")
    W, wc = load(write_csv, "WRITE_SIZE")
    gib = float(1 << 30)
    f_scale = gib / (sum(fc) / len(fc))
    w_scale = gib / (sum(wc) / len(wc))
    rows = []
    for k in sorted(F, key=lambda k: -sum(F[k]) / len(F[k])):
        f = sum(F[k]) / len(F[k]) * f_scale
        w = sum(W.get(k, [0.0])) / max(len(W.get(k, [])), 1) * w_scale
        rows.append((k, len(F[k]), f, w))
    pair = [r for r in rows if r[0].split("<")[0] in PAIR_KERNELS]
    traffic = sum(r[2] + r[3] for r in pair)
    res = dict(traffic_bytes=traffic, fetch_scale=f_scale, write_scale=w_scale,
               kernels={r[0]: dict(fetch_bytes=r[2], write_bytes=r[3]) for r in pair})
    with open(out_json, "w") as f:
        json.dump(res, f, indent=1)
    print(f"copy calibration: FETCH_SIZE x{f_scale:.3f}, WRITE_SIZE x{w_scale:.3f}")
    print("| kernel | dispatches | HBM read (MB) | HBM write (MB) |")
    print("|---|---|---|---|")
    for k, n, f, w in rows[:20]:
        print(f"| {k} | {n} | {f / 1e6:.1f} | {w / 1e6:.1f} |")
    print(f"pair transform traffic per launch: {traffic / 1e9:.3f} GB")


def counter_summary(paths, kernels=None):
    """Per-kernel mean of every counter in rocprofv3 --pmc CSV files (one
    counter pass each), printed as a markdown table; derived ratios for the
    VALU, LDS and L2 counters when present."""
    import csv
    import collections
    vals = collections.defaultdict(lambda: collections.defaultdict(list))
    for path in paths:
        for r in csv.DictReader(open(path)):
            name = r["Kernel_Name"].split("(")[0].replace("void ", "").strip()
            if kernels and not any(name.startswith(k) for k in kernels):
                continue
            vals[name][r["Counter_Name"]].append(float(r["Counter_Value"]))
    counters = sorted({c for k in vals for c in vals[k]})
    print("| kernel | dispatches | " + " | ".join(counters) + " |")
    print("|---|---|" + "---|" * len(counters))
    for k in sorted(vals, key=lambda k: -len(vals[k])):
        n = max(len(v) for v in vals[k].values())
        row = [f"{np.mean(vals[k][c]):.4g}" if c in vals[k] else "" for c in counters]
        print(f"| {k} | {n} | " + " | ".join(row) + " |")
    return {k: {c: float(np.mean(v)) for c, v in d.items()} for k, d in vals.items()}


def k2_counter_json(paths, out_json):
    """Executed FP32 work of the pair transform from rocprofv3 --pmc passes:
    64 lanes x (ADD + MUL + 2 FMA + TRANS) of the VALU plus 512 x the F32 MFMA
    MOPS counter, per launch, and the VALU busy percentage of k2_cols_r2."""
    stats = counter_summary(paths, kernels=["k2_cols", "k2_rows"])
    flop = 0.0
    for k, c in stats.items():
        flop += 64.0 * (c.get("SQ_INSTS_VALU_ADD_F32", 0.0) + c.get("SQ_INSTS_VALU_MUL_F32", 0.0) +
                        2.0 * c.get("SQ_INSTS_VALU_FMA_F32", 0.0) + c.get("SQ_INSTS_VALU_TRANS_F32", 0.0))
        flop += 512.0 * c.get("SQ_INSTS_VALU_MFMA_MOPS_F32", 0.0)
    cols = [v for k, v in stats.items() if k.startswith("k2_cols")]
    res = dict(executed_flop=flop, valu_busy=(cols[0].get("VALUBusy") if cols else None),
               lds_bank_conflict=(cols[0].get("LDSBankConflict") if cols else None),
               kernels={k: {c: round(v, 4) for c, v in d.items()} for k, d in stats.items()})
    with open(out_json, "w") as f:
        json.dump(res, f, indent=1)
    print(f"pair transform: executed {flop / 1e9:.1f} GFLOP per launch, VALU busy {res['valu_busy']}")
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run a plain 1 GiB device copy first (reference for rocprofv3 FETCH/WRITE counters)")
    ap.add_argument("--pmc-summary", nargs=3, metavar=("FETCH_CSV", "WRITE_CSV", "OUT_JSON"),
                    help="summarise two rocprofv3 --pmc counter_collection.csv files and exit (no GPU)")
    ap.add_argument("--counter-summary", nargs="+", metavar="CSV",
                    help="per-kernel means of rocprofv3 --pmc counter_collection.csv files and exit (no GPU)")
    ap.add_argument("--kernels", nargs="*", help="with --counter-summary: kernel name prefixes to keep")
    ap.add_argument("--k2-counters", nargs="+", metavar="CSV_OR_OUT",
                    help="counter CSVs then the output JSON: executed FP32 work of the pair transform (no GPU)")
    ap.add_argument("--workload", default="c2", choices=["c2", "served", "batched", "hostinput", "configs"])
    ap.add_argument("--config", default=None, help="configs: run only this entry of CONFIGS")
    ap.add_argument("--inflight", type=int, default=2,
                    help="C2 movies in flight at once for the headline (one plan and stream each); "
                         "single-movie figures are measured first and reported beside it")
    ap.add_argument("--seconds", type=float, default=6.0, help="served: arrival window per rate")
    ap.add_argument("--batch", type=int, default=4, help="batched: movies per batch plan")
    ap.add_argument("--no-concurrent", action="store_true",
                    help="one movie on the GPU at a time (profiling runs): the same as --inflight 1")
    ap.add_argument("--graph", type=int, default=1,
                    help="replay each movie's stage sequence as one captured HIP graph (1) or launch eagerly (0)")
    args = ap.parse_args()
    if args.pmc_summary:
        pmc_summary(*args.pmc_summary)
        return
    if args.counter_summary:
        counter_summary(args.counter_summary, args.kernels)
        return
    if args.k2_counters:
        k2_counter_json(args.k2_counters[:-1], args.k2_counters[-1])
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: there is no multi-GPU path (movies are independent: "
              "each GPU would run its own movies, one process per GPU).", file=sys.stderr)
        sys.exit(2)

    import torch
    if not torch.cuda.is_available():
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    if args.workload == "served":
        return bench_served(args)
    if args.workload == "batched":
        return bench_batched(args)
    if args.workload == "hostinput":
        return bench_hostinput(args)
    if args.workload == "configs":
        return bench_configs(args)
    return bench_c2(args)


def bench_c2(args):
    import torch
    from zorro_amd.engine import MoviePlan, make_params
    from zorro_amd.synth import make_movie_torch

    dev = torch.device("cuda:0")
    c = WORKLOAD
    print(f"[bench] generating {c['n']}x{c['h']}x{c['w']} movie on {dev}", file=sys.stderr, flush=True)
    frames, drift = make_movie_torch(c["n"], c["h"], c["w"], seed=c["seed"], dose=c["dose"], jump=c["jump"],
                                     device=dev)
    torch.cuda.synchronize()
    if args.copy_calibration:
        src = torch.empty(1 << 28, dtype=torch.float32, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    K = 1 if args.no_concurrent else max(1, args.inflight)
    plans = [MoviePlan(make_params(**PARAMS), c["n"], c["h"], c["w"]) for _ in range(max(K, 2))]
    for p in plans:
        p.set_frames(frames)
        if args.graph:
            p.use_graph(True)
    plan = plans[0]
    info = plan.info()
    # stage times from HIP events on plan 0's stream (inside its graph when graphs are on)
    plan.set_timing(True)

    print(f"[bench] plan ready {info}; {K} in flight; graph {args.graph}; warmup {args.warmup}", file=sys.stderr,
          flush=True)
    for i in range(max(args.warmup, len(plans) if args.graph else 0)):
        plans[i % len(plans)].align()
    for p in plans:
        p.synchronize()
    torch.cuda.synchronize()

    def timed(kin, steps):
        """kin plans on kin streams: movie s runs on plan s % kin once that
        plan's previous movie has finished; latency = enqueue -> finish."""
        lat = []
        start = [0.0] * kin
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for s in range(steps):
            k = s % kin
            if s >= kin:
                plans[k].synchronize()
                lat.append((time.perf_counter() - start[k]) * 1e3)
            start[k] = time.perf_counter()
            plans[k].align()
        for s in range(max(0, steps - kin), steps):
            k = s % kin
            plans[k].synchronize()
            lat.append((time.perf_counter() - start[k]) * 1e3)
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3 / steps, lat

    # single movie first: stage times on plan 0's stream (HIP events inside its
    # graph) with nothing else on the GPU, single-movie latency
    plan.stage_times(reset=True)
    ms1, lat1 = timed(1, args.steps)
    st = plan.stage_times()
    plan.set_timing(False)
    print(f"[bench] 1 in flight: {args.steps} movies, {ms1:.2f} ms/movie", file=sys.stderr, flush=True)
    single = dict(ms_per_movie=round(ms1, 3), value=round(60000.0 / ms1, 2),
                  latency_ms=dict(p50=round(float(np.percentile(lat1, 50)), 3),
                                  p99=round(float(np.percentile(lat1, 99)), 3)))
    if K > 1:
        # the headline: K movies in flight on K plans and streams (one movie's
        # latency-bound solver stages overlap another's bandwidth-bound ones)
        ms_step, lat = timed(K, args.steps)
        print(f"[bench] {K} in flight: {args.steps} movies, {ms_step:.2f} ms/movie", file=sys.stderr, flush=True)
    else:
        ms_step, lat = ms1, lat1

    movies_min = 60000.0 / ms_step
    sol = plan.solution()
    truth = -(drift - drift.mean(axis=0))
    rms = float(np.sqrt(np.mean((sol["translations"] - truth) ** 2)))

    k2_ms = st["pairs"][0] / max(st["pairs"][1], 1)
    sec = k2_ms * 1e-3
    flops = k2_flops(info)
    traffic = load_traffic()
    comp = k2_compulsory_bytes(info)
    counters = load_profile_json("k2_pairs_counters.json") or {}
    # The pair transform is fp32-compute bound (VALU busy and the FP32 instruction
    # counts in profiles/), so its roofline is the fp32 peak, with the scope
    # table's algorithmic FLOP count (a 2D inverse FFT per pair) as the work.
    roofline = dict(bound="mfma", achieved=round(flops / sec / 1e12, 2), peak=FP32_PEAK_TFLOPS, unit="TFLOP/s",
                    frac=round(flops / sec / (FP32_PEAK_TFLOPS * 1e12), 4), traffic=traffic,
                    kernel="pair transform (k2_cols_r2 + k2_rows_mfma), fp32 vector ALU",
                    launch_ms=round(k2_ms, 4), algorithmic_flop=flops,
                    hbm=dict(compulsory_bytes=comp, compulsory_frac=round(comp / sec / (HBM_PEAK_GBS * 1e9), 4),
                             traffic_frac=(round(traffic / sec / (HBM_PEAK_GBS * 1e9), 4) if traffic else None),
                             no_reuse_bytes=k2_algorithmic_bytes(info)),
                    executed_flop=counters.get("executed_flop"),
                    executed_frac=(round(counters["executed_flop"] / sec / (FP32_PEAK_TFLOPS * 1e12), 4)
                                   if counters.get("executed_flop") else None),
                    valu_busy=counters.get("valu_busy"))

    out = dict(metric=METRIC, value=round(movies_min, 2), unit="movies/min", n_gpus=1, steps=args.steps,
               warmup=args.warmup, ms_per_step=round(ms_step, 3), higher_is_better=True, scaling="weak",
               vs_baseline=None, dtype="f32", data="synthetic",
               config=dict(workload="40 frames x 4096x4096 float32 Poisson movie (1.0 e/px/frame, random-walk "
                                    "drift + 3 px exponential jump): 780-pair tri xcorr at fouCrop 3072, "
                                    "maxShift 40, Bmode conv Brad 150, subPixReg 16, autologistic weights, "
                                    "L1 shifts solver, Lanczos shift + sum, even/odd FRC (512 tiles), "
                                    "dose filter (1 A/px, 300 kV)", frames=c["n"], shape=[c["h"], c["w"]],
                           pairs=info["pairs"], fouCrop=[info["Hc"], info["Wc"]], inputs="resident in HBM"),
               inflight=K, hip_graph=bool(args.graph),
               frames_per_s=round(c["n"] * 1000.0 / ms_step, 1),
               latency_ms=dict(p50=round(float(np.percentile(lat, 50)), 3),
                               p99=round(float(np.percentile(lat, 99)), 3)),
               stages_ms={k: round(v[0] / max(v[1], 1), 4) for k, v in st.items() if v[1]},
               roofline=roofline,
               check=dict(rms_vs_true_drift_px=round(rms, 4), solver=plan.solver_info()))
    out["single_movie"] = single

    if not args.no_cpu_baseline:
        print("[bench] timing the CPU baseline sample", file=sys.stderr, flush=True)
        from oracle.cpu_reference_structure import time_movie
        v, d = time_movie(n_frames=c["n"], shape=c["h"], crop=3072)
        out["cpu_baseline"] = dict(value=round(v, 5), unit="movies/min", cores=d["cores"], kind="port",
                                   sample=d["sample"], stage_s={k: round(x, 2) for k, x in d["stage_s"].items()})
        out["speedup_vs_cpu"] = round(movies_min / v, 1)
    for p in plans:
        p.close()
    print(json.dumps(out), flush=True)


def bench_hostinput(args):
    """C2 movies whose frames start in pinned host memory: float32 frames
    (2.7 GB per movie) and 4-bit counting frames (two counts per byte, 0.34 GB,
    unpacked on the GPU).  Two plans on two streams; the host-to-device copy of
    the next movie runs on a copy stream while the other plan computes, so the
    rate is the slower of the PCIe copy and the path itself."""
    import torch
    from zorro_amd.engine import MoviePlan, make_params
    from zorro_amd.synth import make_movie_torch
    dev = torch.device("cuda:0")
    c = WORKLOAD
    frames, drift = make_movie_torch(c["n"], c["h"], c["w"], seed=c["seed"], dose=c["dose"], jump=c["jump"],
                                     device=dev)
    counts = torch.clamp(frames, max=15).to(torch.uint8)
    packed = (counts[..., 0::2] | (counts[..., 1::2] << 4)).contiguous()
    host = {"f32": torch.empty(frames.shape, dtype=torch.float32, pin_memory=True),
            "u4": torch.empty(packed.shape, dtype=torch.uint8, pin_memory=True)}
    host["f32"].copy_(frames)
    host["u4"].copy_(packed)
    del frames, counts, packed
    torch.cuda.synchronize()
    out = dict(metric="movies/min with host-resident frames (40x4096^2, copy overlapped)", unit="movies/min",
               n_gpus=1, data="synthetic", inflight=2, modes={})
    steps = max(args.steps, 6)
    for mode in ("f32", "u4"):
        h_src = host[mode]
        bufs = [torch.empty(h_src.shape, dtype=h_src.dtype, device=dev) for _ in range(2)]
        comp = [torch.cuda.Stream(device=dev) for _ in range(2)]
        copy_s = torch.cuda.Stream(device=dev)
        plans = [MoviePlan(make_params(**PARAMS), c["n"], c["h"], c["w"]) for _ in range(2)]
        for k, p in enumerate(plans):
            p.set_stream(comp[k].cuda_stream)
            if mode == "f32":
                p.set_frames(bufs[k])
            p.use_graph(True)
        done = [torch.cuda.Event() for _ in range(2)]
        copied = [torch.cuda.Event() for _ in range(2)]

        def movie(k):
            copy_s.wait_event(done[k])
            with torch.cuda.stream(copy_s):
                bufs[k].copy_(h_src, non_blocking=True)
            copied[k].record(copy_s)
            comp[k].wait_event(copied[k])
            if mode == "u4":
                plans[k].set_frames_raw(bufs[k], packed4=True)
            plans[k].align()
            done[k].record(comp[k])

        for s_ in range(4):                 # warm-up: graph capture on both plans
            movie(s_ % 2)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for s_ in range(steps):
            movie(s_ % 2)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / steps
        # the copy alone, for the PCIe rate
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        with torch.cuda.stream(copy_s):
            for _ in range(3):
                bufs[0].copy_(h_src, non_blocking=True)
        torch.cuda.synchronize()
        copy_ms = (time.perf_counter() - t1) * 1e3 / 3
        nbytes = h_src.numel() * h_src.element_size()
        sol = plans[0].solution()
        truth = -(drift - drift.mean(axis=0))
        out["modes"][mode] = dict(value=round(60000.0 / ms, 1), ms_per_movie=round(ms, 3),
                                  host_bytes_per_movie=nbytes, copy_ms=round(copy_ms, 3),
                                  h2d_gbs=round(nbytes / copy_ms / 1e6, 2),
                                  rms_vs_true_drift_px=round(float(np.sqrt(np.mean((sol["translations"] - truth) ** 2))), 4))
        print(f"[bench hostinput] {mode}: {ms:.2f} ms/movie, copy {copy_ms:.2f} ms ({nbytes / copy_ms / 1e6:.1f} GB/s)",
              file=sys.stderr, flush=True)
        for p in plans:
            p.close()
        del bufs
        torch.cuda.empty_cache()
    print(json.dumps(out), flush=True)


# ---------------------------------------------------------------- C1 and C4
CONFIGS = {
    # 20 frames of 1024^2 float32 (1.5 e/px/frame), no crop, Fourier applyShifts
    "c1_20x1024sq_fourier": dict(
        n=20, h=1024, w=1024, seed=21, dose=1.5, raw=False, scale=1.0,
        params=dict(PARAMS, fouCrop=(0, 0), shapePadded=(1024, 1024), shiftMethod="fourier")),
    # 50 frames of 7676x7420 uint8 counts (0.25 e/superres px/frame) times a float32
    # gain reference N(1, 0.05), Fourier-binned to 3838x3710, padded to 3840^2,
    # cropped to 3072 for the correlation, dose filter at 0.5 A/px
    "c4_50x7676x7420_u8_gain_bin3838x3710": dict(
        n=50, h=7676, w=7420, seed=76, dose=0.25, raw=True, scale=0.5,
        params=dict(PARAMS, fouCrop=(3072, 3072), shapePadded=(3840, 3840), shapeBinned=(3838, 3710),
                    pixelsize=0.025)),
}


def bench_configs(args):
    """The other two single-movie configurations, each with its input resident in
    HBM: a step is one movie through the path (for the counting movie: the gain
    product and the Fourier binning, then registration, dose filter and sums).
    One movie at a time first (latency, stage times), then two in flight."""
    import torch
    from zorro_amd.engine import MoviePlan, make_params
    from zorro_amd.synth import make_movie_torch
    dev = torch.device("cuda:0")
    out = dict(metric="movies/min per configuration (single GPU, inputs resident in HBM)", unit="movies/min",
               n_gpus=1, data="synthetic", configs={})
    for name, c in CONFIGS.items():
        if args.config and name != args.config:
            continue
        print(f"[bench configs] {name}: generating", file=sys.stderr, flush=True)
        frames, drift = make_movie_torch(c["n"], c["h"], c["w"], seed=c["seed"], dose=c["dose"], device=dev)
        gain = None
        if c["raw"]:
            frames = torch.clamp(frames, max=255).to(torch.uint8)
            g = torch.Generator(device=dev)
            g.manual_seed(c["seed"] + 1)
            gain = (1.0 + 0.05 * torch.randn((c["h"], c["w"]), generator=g, device=dev)).contiguous()
        torch.cuda.synchronize()
        plans = [MoviePlan(make_params(**c["params"]), c["n"], c["h"], c["w"]) for _ in range(2)]

        def movie(k):
            if c["raw"]:
                plans[k].set_frames_raw(frames, gain=gain)
            plans[k].align()

        for k, p in enumerate(plans):
            if not c["raw"]:
                p.set_frames(frames)
            p.use_graph(True)
            p.set_timing(k == 0)
            for _ in range(max(2, args.warmup)):
                movie(k)
            p.synchronize()
        plans[0].stage_times(reset=True)
        res = {}
        for kin in (1, 2):
            lat, start = [], [0.0, 0.0]
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for s_ in range(args.steps):
                k = s_ % kin
                if s_ >= kin:
                    plans[k].synchronize()
                    lat.append((time.perf_counter() - start[k]) * 1e3)
                start[k] = time.perf_counter()
                movie(k)
            for s_ in range(max(0, args.steps - kin), args.steps):
                k = s_ % kin
                plans[k].synchronize()
                lat.append((time.perf_counter() - start[k]) * 1e3)
            ms = (time.perf_counter() - t0) * 1e3 / args.steps
            res[kin] = dict(ms_per_movie=round(ms, 3), value=round(60000.0 / ms, 1),
                            frames_per_s=round(c["n"] * 1000.0 / ms, 1),
                            latency_ms=dict(p50=round(float(np.percentile(lat, 50)), 3),
                                            p99=round(float(np.percentile(lat, 99)), 3)))
            if kin == 1:
                st = plans[0].stage_times()
                plans[0].set_timing(False)
        sol = plans[0].solution()
        truth = -(drift - drift.mean(axis=0)) * c["scale"]
        rms = float(np.sqrt(np.mean((sol["translations"] - truth) ** 2)))
        info = plans[0].info()
        out["configs"][name] = dict(value=res[2]["value"], unit="movies/min", inflight=2, two_in_flight=res[2],
                                    single_movie=res[1], frames=c["n"], input_shape=[c["h"], c["w"]],
                                    registered_shape=[info["H"], info["W"]], padded=[info["Hp"], info["Wp"]],
                                    crop=[info["Hc"], info["Wc"]], pairs=info["pairs"],
                                    stages_ms={k: round(v[0] / max(v[1], 1), 4) for k, v in st.items() if v[1]},
                                    rms_vs_true_drift_px=round(rms, 4))
        print(f"[bench configs] {name}: {out['configs'][name]}", file=sys.stderr, flush=True)
        for p in plans:
            p.close()
        del frames, gain, plans
        torch.cuda.empty_cache()
    print(json.dumps(out), flush=True)


# ---------------------------------------------------------------- serving
SERVE_BASE = dict(triMode="tri", maxShift=40, Bmode="conv", Brad=150, subPixReg=16, shiftMethod="lanczos",
                  weightMode="autologistic", originMode="centroid", suppressOrigin=True, filterMode="dose",
                  pixelsize=0.1, voltage=300.0, doLazyFRC=True, frc_box=512, frc_overlap=0.5)
# detector shape -> (padded shape, fouCrop): padded to the smallest square 2^a3^b5^c size as
# findValidFFTWDim does (zorro/zorro_util.py:1068-1100), cropped to 3072 when larger
SHAPES = {(2048, 2048): ((2048, 2048), (0, 0)), (3838, 3710): ((3840, 3840), (3072, 3072)),
          (4096, 4096): ((4096, 4096), (3072, 3072)), (5760, 4092): ((5760, 5760), (3072, 3072))}


def _movie_pool(specs, seed0=100):
    """Device-resident synthetic movies, one per (n, h, w, dose) spec."""
    import torch
    from zorro_amd.synth import make_movie_torch
    pool = []
    for i, (n, h, w, dose) in enumerate(specs):
        f, d = make_movie_torch(n, h, w, seed=seed0 + i, dose=dose, device="cuda:0")
        pad, crop = SHAPES[(h, w)]
        pool.append((f, d, dict(SERVE_BASE, shapePadded=pad, fouCrop=crop)))
    torch.cuda.synchronize()
    return pool


def _rms_truth(res, drift):
    truth = -(drift - drift.mean(axis=0))
    return float(np.sqrt(np.mean((res.translations - truth) ** 2)))


def bench_served(args):
    """Serving stream: Poisson arrivals of a 30/40/30 mix of 2048^2 x 20,
    4096^2 x 40 and 5760x4092 x 60 movies at 0.2-2x the measured saturation
    rate; movies/min, p50/p99 latency and the error against the true drift at
    each rate."""
    from zorro_amd.serve import MovieServer, latency_summary
    classes = [((20, 2048, 2048, 1.0), 0.3), ((40, 4096, 4096, 1.0), 0.4), ((60, 5760, 4092, 1.0), 0.3)]
    print("[bench served] generating the movie pool", file=sys.stderr, flush=True)
    pool = _movie_pool([c for c, _ in classes])
    probs = np.array([p for _, p in classes])
    rng = np.random.default_rng(7)
    out = dict(metric="served movies/min and p50/p99 latency (30% 2048^2x20, 40% 4096^2x40, 30% 5760x4092x60)",
               unit="movies/min", n_gpus=1, data="synthetic", rates=[])
    inflight = max(2, args.inflight)
    out["inflight"] = inflight
    with MovieServer(max_inflight=inflight, return_sums="device") as srv:
        # warm-up: every in-flight slot of every class gets its plan before the clock
        # starts, then one movie of each class runs through it
        for f, d, kw in pool:
            srv.warm(f, count=inflight, **kw)
        for f, d, kw in pool:
            srv.submit(f, **kw).result()
        # saturation: a closed backlog in the mix proportions (90 movies, so that
        # the rate is not set by a few large ones)
        seq = rng.choice(len(pool), size=90, p=probs)
        t0 = time.perf_counter()
        futs = [srv.submit(pool[k][0], **pool[k][2]) for k in seq]
        res = [f.result() for f in futs]
        sat = len(seq) / (time.perf_counter() - t0)          # movies/s
        out["saturation_movies_min"] = round(sat * 60, 1)
        print(f"[bench served] saturation {sat * 60:.1f} movies/min", file=sys.stderr, flush=True)
        for frac in (0.2, 0.5, 0.8, 1.0, 1.5, 2.0):
            lam = sat * frac
            n = max(8, int(lam * args.seconds))
            gaps = rng.exponential(1.0 / lam, size=n)
            kinds = rng.choice(len(pool), size=n, p=probs)
            futs = []
            t0 = time.perf_counter()
            tnext = t0
            for g, k in zip(gaps, kinds):
                tnext += g
                while time.perf_counter() < tnext:
                    time.sleep(min(0.0005, max(0.0, tnext - time.perf_counter())))
                futs.append((k, srv.submit(pool[k][0], arrival=tnext, **pool[k][2])))
            res = [(k, f.result()) for k, f in futs]
            wall = max(r.done for _, r in res) - t0
            lat = latency_summary([r for _, r in res])
            err = max(_rms_truth(r, pool[k][1]) for k, r in res)
            # per class: the mix spans a 20x range of service times (2048^2 x 20 to
            # 5760x4092 x 60), so the pooled p99/p50 is set by the class mix
            per = {}
            for ci, (c, _) in enumerate(classes):
                rc = [r for k, r in res if k == ci]
                if rc:
                    lc = latency_summary(rc)
                    per[f"{c[0]}x{c[1]}x{c[2]}"] = dict(movies=len(rc), p50_ms=round(lc["p50_ms"], 2),
                                                        p99_ms=round(lc["p99_ms"], 2),
                                                        p99_over_p50=round(lc["p99_ms"] / lc["p50_ms"], 3))
            row = dict(rate_frac=frac, offered_movies_min=round(lam * 60, 1), movies=n,
                       achieved_movies_min=round(n / wall * 60, 1), p50_ms=round(lat["p50_ms"], 2),
                       p99_ms=round(lat["p99_ms"], 2), p99_over_p50=round(lat["p99_ms"] / lat["p50_ms"], 3),
                       max_rms_vs_true_drift_px=round(err, 4), per_class=per)
            print(f"[bench served] {row}", file=sys.stderr, flush=True)
            out["rates"].append(row)
        out["server"] = dict(srv.stats)
    print(json.dumps(out), flush=True)


def bench_batched(args):
    """C3: 48 uneven movies (frame counts U{20..60}, shapes cycling over 2048^2,
    3838x3710, 4096^2, 5760x4092, doses U(0.5, 2.0)).  The movies of each shape
    class are registered as batches of ``--batch`` movies (one ragged launch
    sequence per batch, captured as one graph); the batches of a class run on
    their own streams at once.  Plans and graphs are built before the clock
    starts; frames are resident in HBM.  Per class: wall time of all its
    batches; latency of a movie = the end of its batch."""
    import torch
    from zorro_amd.batch import BatchPlan
    from zorro_amd.engine import make_params
    from zorro_amd.synth import make_movie_torch
    rng = np.random.default_rng(11)
    shapes = list(SHAPES)
    specs = [(int(rng.integers(20, 61)), *shapes[i % 4], float(rng.uniform(0.5, 2.0))) for i in range(48)]
    from zorro_amd.serve import estimate_plan_bytes
    B = max(1, args.batch)
    wave_budget = 90e9          # device bytes of the batches resident at once (a wave)
    out = dict(metric="batched movies/min (48 movies, 20-60 frames, 4 detector shapes)", unit="movies/min",
               n_gpus=1, data="synthetic", batch_movies=B, classes=[])
    total_frames = sum(s[0] for s in specs)
    lat, errs = [], []
    t_all = 0.0
    for cls in range(4):
        group = [(i, s) for i, s in enumerate(specs) if i % 4 == cls]
        h, w = shapes[cls]
        pad, crop = SHAPES[(h, w)]
        params = make_params(**dict(SERVE_BASE, shapePadded=pad, fouCrop=crop))
        parts = [group[b0:b0 + B] for b0 in range(0, len(group), B)]
        # waves: consecutive batches whose estimated plans fit the wave budget together
        waves, cur, cur_b = [], [], 0.0
        for part in parts:
            pb = sum(estimate_plan_bytes(s[0], h, w, params) for _, s in part)
            if cur and cur_b + pb > wave_budget:
                waves.append(cur)
                cur, cur_b = [], 0.0
            cur.append(part)
            cur_b += pb
        waves.append(cur)
        wall_cls = 0.0
        for wave in waves:
            batches, drifts = [], []
            for part in wave:
                bp = BatchPlan(params, [s[0] for _, s in part], h, w)
                for k, (i, (n, _, _, dose)) in enumerate(part):
                    f, d = make_movie_torch(n, h, w, seed=300 + i, dose=dose, device="cuda:0")
                    bp.set_frames(k, f)
                    drifts.append(d)
                    del f
                bp.use_graph(True)
                bp.align()                      # capture + warm-up
                batches.append(bp)
            for bp in batches:
                bp.synchronize()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for bp in batches:
                bp.align()
            ends = []
            for bp in batches:
                bp.synchronize()
                ends.append(time.perf_counter() - t0)
            wall = time.perf_counter() - t0
            k = 0
            for bp, end in zip(batches, ends):
                for m in bp.movies:
                    lat.append((wall_cls + end) * 1e3)
                    truth = -(drifts[k] - drifts[k].mean(axis=0))
                    errs.append(float(np.sqrt(np.mean((m.solution()["translations"] - truth) ** 2))))
                    k += 1
            wall_cls += wall
            for bp in batches:
                bp.close()
            del batches
            torch.cuda.empty_cache()
        t_all += wall_cls
        out["classes"].append(dict(shape=[h, w], movies=len(group), frames=sum(s[0] for _, s in group),
                                   batches=len(parts), waves=len(waves), seconds=round(wall_cls, 4),
                                   movies_min=round(len(group) / wall_cls * 60, 1)))
        print(f"[bench batched] class {h}x{w}: {len(group)} movies in {len(parts)} batches ({len(waves)} waves), "
              f"{wall_cls * 1e3:.1f} ms", file=sys.stderr, flush=True)
    out.update(value=round(len(specs) / t_all * 60, 1), frames_per_s=round(total_frames / t_all, 1),
               seconds=round(t_all, 4), max_rms_vs_true_drift_px=round(max(errs), 4),
               latency_ms=dict(p50=round(float(np.percentile(lat, 50)), 2), p99=round(float(np.percentile(lat, 99)), 2)))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
