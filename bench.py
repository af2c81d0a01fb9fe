#!/usr/bin/env python3
"""Throughput of the dense-disparity path on one MI355X.

Workload (one step): a batch of synthetic 460x370 rectified pairs
(disparity 10-80 px, slanted planes + thin foreground boxes, sigma-2 noise,
uint8, resident in HBM) through the full plain pipeline:
predict_auto with the trained v5 StereoNet (probe + up to 3 passes, FP32
network) followed by refine_disparity(planes=False): census 9x7, band SGM
(R=10, P1=12, P2=150, 8 paths, both views), LR check (tol 2), 3x3 median.

The step is timed as individual launches, replayed as one HIP graph, and
replayed pipelined over batches (the network of batch k and the refinement of
batch k-1 as two branches of one graph); value is the fastest of the graph
replays, each checked equal to the individual launches' output.

Printed (one JSON line): pairs/s (value), the full-range SGM control on the
same pairs (eval_real.py:143-146 rule, one launch for all pairs, replayed as
a HIP graph like the headline), accuracy of both (bad2 / EPE vs the
generated ground truth on visible pixels), whether any band half-width /
weights setting reaches full-range bad2 (the sweep), the roofline of the
dominant kernel (band SGM aggregation + epilogue; live HIP-event timing of
its launches) and the CPU baseline (the float64 C oracle on this host's
cores, one process per core, on a bounded sample of the same workload;
one core beside it).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--batch B]
"""

import argparse
import gc
import importlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "pairs/s @460×370: net+band-SGM vs full-range SGM at equal bad2; HBM GB/s"
HBM_PEAK_GBS = 8000.0
W, H = 460, 370
R_BAND = 10


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# Measured HBM traffic of the band-SGM stage (k_sgm_ph2 + k_sgm_p2, both
# views) per 460x370 pair: rocprofv3 FETCH_SIZE x2.0 + WRITE_SIZE x1.0 (the
# factors a 256 MiB device copy reads under the same counters), from
# profiles/sgm_460x370_pmc.md at sgm_split=2, sgm_prefetch=8, sgm_htile=16, per
# workgroup order of the vertical/diagonal kernel (sgm_order 1: launch order;
# 2: one run of blocks per XCD, whose L2 then serves a pair's directions).
# A counter reading, not measured live: bench.py cannot read PMC counters.
SGM_TRAFFIC_PROFILE = ("profiles/sgm_460x370_pmc.md", {1: 5083.5e6 / 32, 2: 4056.6e6 / 32})


def canonical_bytes_per_pair(views, K, HW):
    """Algorithmic HBM bytes of the eight-path aggregation and its epilogue
    per pair (the canonical dataflow: per view, cell and path, read the u8
    cost and read+write a u16 path sum; the epilogue reads the u16 sums
    once): V*HW*K*(8*(1+2+2) + 2)."""
    return views * HW * K * (8 * (1 + 2 + 2) + 2)


def u8_dataflow_bytes_per_pair(views, Kp, HW, ndir=8):
    """Bytes the packed SGM kernels move per pair in their own u8 dataflow:
    per view, pixel and path, read Kp cost bytes and write Kp path-cost
    bytes (Kp = band length padded to 8)."""
    return views * HW * Kp * ndir * 2


V5_WEIGHTS = os.path.join(ROOT, "from-scratch-stereo-vision_amd", "weights", "stereonet_weights_v5.txt")
C5_WEIGHTS = os.path.join(ROOT, "from-scratch-stereo-vision_amd", "weights", "stereonet_c5.txt")


def read_stages(L, names, steps):
    """Per-step ms of the library's profiled stages since sv_profile_enable(1)."""
    import ctypes
    out = {}
    for name in names:
        ms, cnt = ctypes.c_double(0), ctypes.c_int(0)
        L.check(L.lib().sv_profile_read(name.encode(), ctypes.byref(ms), ctypes.byref(cnt)))
        out[name] = ms.value / max(steps, 1)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=None,
                    help="pairs per step (default: 64 for the headline pipeline, whose network "
                         "launches cost nearly the same for 32 or 64 pairs, else 32)")
    ap.add_argument("--cpu-pairs", type=int, default=24)
    ap.add_argument("--threads", type=int, default=16,
                    help="images workload: host decoding threads")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fullrange", action="store_true")
    ap.add_argument("--no-sweep", action="store_true",
                    help="skip the band half-width / weights sweep of the headline workload")
    ap.add_argument("--cpu-worker", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--copy-probe", action="store_true",
                    help="after the run, copy 256 MiB device-to-device once: a kernel of known "
                         "traffic to scale rocprofv3 FETCH_SIZE/WRITE_SIZE readings")
    ap.add_argument("--tune-file", default=None,
                    help="JSON of kernel variant choices: loaded when it exists (no autotuning), "
                         "else written after autotuning")
    ap.add_argument("--eager", action="store_true",
                    help="time the step as individual launches only (default: also capture it "
                         "in one HIP graph, check the replay equals the eager output, and report "
                         "the replay rate)")
    ap.add_argument("--streams", type=int, default=1,
                    help="graph replay: cut the batch into this many equal parts, each on its "
                         "own stream with its own network buffers, captured together so that one "
                         "part's kernels fill the other's gaps (1: the eager step as it is)")
    ap.add_argument("--no-pipeline", action="store_true",
                    help="graph replay: time only the sequential step (default: also replay a "
                         "pipelined step, the network of batch k beside the refinement of batch "
                         "k-1 as two graph branches, and report the faster of the two)")
    ap.add_argument("--tune-only", action="store_true",
                    help="autotune, write the choices to --tune-file and exit (the default "
                         "run tunes this way in a child process, then loads the file)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the library's built-in kernel variant heuristics")
    ap.add_argument("--tune", default=None,
                    help="apply these kernel variants instead of autotuning, e.g. "
                         "sgm_prefetch=8,sgm_htile=16,conv_rows=1,conv_tile_width=32")
    ap.add_argument("--workload", default="pipeline",
                    choices=["pipeline", "band", "ragged64", "widened", "consistent", "planes",
                             "images", "eval", "mesh", "train", "splat"],
                    help="pipeline: 460x370 pairs (the headline); ragged64: 64 pairs of uneven "
                         "size in one captured graph")
    args = ap.parse_args()
    if args.batch is None:
        args.batch = 64 if args.workload == "pipeline" else 32
    if args.cpu_worker:
        return cpu_worker(args.cpu_worker)
    if args.gpus != 1:
        print("bench.py measures one GPU: the library has no multi-GPU path "
              "(pairs are independent: several GPUs would each take their own share of a batch)", file=sys.stderr)
        sys.exit(2)
    if args.workload == "ragged64":
        return ragged_main(args)
    if args.workload == "widened":
        return widened_main(args)
    if args.workload == "consistent":
        return consistent_main(args)
    if args.workload == "planes":
        return planes_main(args)
    if args.workload == "images":
        return images_main(args)
    if args.workload == "eval":
        return eval_main(args)
    if args.workload == "mesh":
        return mesh_main(args)
    if args.workload == "splat":
        return splat_main(args)
    if args.workload == "band":
        return band_main(args)
    if args.workload == "train":
        return train_main(args)
    if args.tune_only and not args.tune_file:
        print("--tune-only needs --tune-file", file=sys.stderr)
        sys.exit(2)
    if not (args.tune_only or args.tune or args.no_autotune
            or (args.tune_file and os.path.exists(args.tune_file))):
        # tune in a child process and time in this one with the saved
        # choices: a graph replayed in the process that just ran the tuner
        # measured 2-3% slower than the same choices loaded in a fresh one
        import subprocess
        import tempfile
        path = args.tune_file or os.path.join(tempfile.mkdtemp(prefix="svtune"), "tune.json")
        cmd = [sys.executable, os.path.abspath(__file__), "--tune-only", "--tune-file", path,
               "--batch", str(args.batch), "--streams", str(args.streams)]
        if args.eager:
            cmd.append("--eager")
        log("autotuning in a child process: %s" % " ".join(cmd[1:]))
        subprocess.run(cmd, check=True, stdout=sys.stderr)
        args.tune_file = path

    import numpy as np
    import torch

    sv = importlib.import_module("from-scratch-stereo-vision_amd")
    L = importlib.import_module("from-scratch-stereo-vision_amd._lib")
    N = importlib.import_module("from-scratch-stereo-vision_amd.net")
    dev = L.device()
    B = args.batch
    log("generating %d synthetic pairs %dx%d" % (B, W, H))
    pairs = [sv.synth.pipeline_pair(s) for s in range(B)]
    L8 = torch.from_numpy(np.stack([p["left"] for p in pairs])).to(dev)
    R8 = torch.from_numpy(np.stack([p["right"] for p in pairs])).to(dev)
    gt = np.stack([p["disp"] for p in pairs]).reshape(B, -1)
    valid = np.stack([p["valid"] for p in pairs]).reshape(B, -1)
    wpath = V5_WEIGHTS
    net = N.load_model(wpath)
    d0 = torch.empty(B, H * W, dtype=torch.float64, device=dev)
    lv = N._Level([(H, W)] * B, dev)
    S = 1 if args.eager else args.streams
    if S < 1 or B % S:
        print("--streams must divide the batch", file=sys.stderr)
        sys.exit(2)
    hb = B // S

    t_net = [0.0]

    def step(events=True):
        if events:
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
        res = N.predict_auto_batch(net, L8, R8, W, H)
        N._call("sv_widen", N._p(res.full), N._p(lv.table), H * W, N._p(d0), H * W, None, B,
                H * W, L.stream())
        if events:
            e1.record()
        out, st = sv.refine_disparity(L8, R8, None, d0, W, H, planes=False)
        return out, st, res, ((e0, e1) if events else None)

    # the parted step: part i (pairs i*hb ...) on streams[i] with nets[i]
    # (activation buffers are per network object, scratch per stream)
    nets = []  # made after tuning, so their buffers sit in a settled allocator
    streams = [torch.cuda.Stream() for _ in range(S)] if S > 1 else []
    lvh = N._Level([(H, W)] * hb, dev)

    def part(i, nt=None):
        sl = slice(i * hb, (i + 1) * hb)
        res = N.predict_auto_batch(nt or nets[i], L8[sl], R8[sl], W, H)
        N._call("sv_widen", N._p(res.full), N._p(lvh.table), H * W, N._p(d0[sl]), H * W, None,
                hb, H * W, L.stream())
        return sv.refine_disparity(L8[sl], R8[sl], None, d0[sl], W, H, planes=False)

    def parted_step():
        cur = torch.cuda.current_stream()
        outs = []
        for i, s in enumerate(streams):
            s.wait_stream(cur)
            with torch.cuda.stream(s):
                outs.append(part(i))
        for s in streams:
            cur.wait_stream(s)
        return outs

    tuned = None
    A = importlib.import_module("from-scratch-stereo-vision_amd.autotune")
    if args.tune:
        tuned = {k: int(v) for k, v in (kv.split("=") for kv in args.tune.split(","))}
        A.apply(tuned)
    elif args.tune_file and os.path.exists(args.tune_file):
        tuned = A.load_choices(net, args.tune_file)
        log("kernel variants from %s" % args.tune_file)
    elif not args.no_autotune:
        # per-layer convolution variants first, then the global knobs of the
        # rest (every variant leaves the results unchanged)
        layers = A.tune_layers(net, lambda: step()[:2], reps=2)
        tuned = A.tune(lambda: step()[:2],
                       knobs=["sgm_split", "sgm_prefetch", "sgm_htile", "sgm_overlap", "wta_reads", "sgm_order", "sgm_queue"], reps=2,
                       log=log)
        tuned["conv_layers_tuned"] = len(layers)
    # the parts share the per-layer choices (keyed by launch shape, batch
    # included); shapes of the part batch get their own when missing,
    # tuned on a network object of their own
    if S > 1 and not args.no_autotune and not args.tune \
            and not any(k[2] == hb for k in net._conv_tune):
        tnet = N.load_model(wpath)
        tnet._conv_tune = net._conv_tune
        more = A.tune_layers(tnet, lambda: part(0, tnet), reps=2)
        del tnet
        # (tune_layers leaves every knob at its default)
        A.apply({k: v for k, v in (tuned or {}).items() if k in A.KNOBS})
        if tuned is not None:
            tuned["conv_layers_tuned"] = tuned.get("conv_layers_tuned", 0) + len(more)
        if args.tune_file and os.path.exists(args.tune_file):
            A.save_choices(net, {k: v for k, v in tuned.items() if k in A.KNOBS},
                           args.tune_file)
    if args.tune_file and not os.path.exists(args.tune_file) and tuned is not None \
            and not args.tune:
        A.save_choices(net, tuned, args.tune_file)
    if args.tune_only:
        log("kernel variants written to %s" % args.tune_file)
        return
    # return what tuning left cached before the parts' buffers are made
    gc.collect()
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    for _ in range(S if S > 1 else 0):
        nt = N.load_model(wpath)
        nt._conv_tune = net._conv_tune
        nets.append(nt)
    log("warmup %d" % args.warmup)
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    L.lib().sv_profile_enable(1)
    log("timing %d steps of %d pairs" % (args.steps, B))
    net_events = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out, st, res, ev = step()
        net_events.append(ev)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    import ctypes
    stages = {}
    for name in ("census", "costs", "sgm", "wta", "lr", "median"):
        ms, cnt = ctypes.c_double(0), ctypes.c_int(0)
        L.check(L.lib().sv_profile_read(name.encode(), ctypes.byref(ms), ctypes.byref(cnt)))
        stages[name] = (ms.value, cnt.value)
    L.lib().sv_profile_enable(0)
    net_ms = sum(a.elapsed_time(b) for a, b in net_events)
    eager_pairs_s = B * args.steps / dt
    pairs_s, execution = eager_pairs_s, "eager"
    seq_graph_pairs_s = pipe_pairs_s = None
    if not args.eager:
        # the same step replayed as one HIP graph (every launch recorded
        # once, the kernel variants chosen above baked in); its output must
        # equal the eager step's
        R = importlib.import_module("from-scratch-stereo-vision_amd.ragged")
        if S > 1:
            g = R.CapturedPipeline(lambda a, b: parted_step(), L8, R8, warmup=1)
            parts = g.replay()
            torch.cuda.synchronize()
            gout = torch.cat([o for o, _ in parts])
            gst = torch.cat([t for _, t in parts])
        else:
            g = R.CapturedPipeline(lambda a, b: step(events=False)[:2], L8, R8, warmup=1)
            gout, gst = g.replay()
            torch.cuda.synchronize()
        if not (torch.equal(gout, out) and torch.equal(gst, st)):
            raise RuntimeError("graph replay differs from the eager step")
        for _ in range(args.warmup):
            g.replay()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            g.replay()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        pairs_s = B * args.steps / dt
        execution = "hip graph" if S == 1 else "hip graph, %d streams x %d pairs" % (S, hb)
        log("graph replay: %.1f pairs/s (eager %.1f)" % (pairs_s, eager_pairs_s))
        seq_graph_pairs_s = pairs_s
        if S == 1 and not args.no_pipeline:
            pipe_pairs_s = pipelined_rate(args, sv, N, L, R, net, L8, R8, d0, lv, out, st)
            log("pipelined graph replay: %.1f pairs/s" % pipe_pairs_s)
            if pipe_pairs_s > pairs_s:
                pairs_s = pipe_pairs_s
                dt = B * args.steps / pipe_pairs_s
                execution = ("hip graph, pipelined: the network of batch k and the refinement "
                             "of batch k-1 as two branches of one graph per step")

    # accuracy of net + band refinement
    pred = out.reshape(B, -1).cpu().numpy()
    err = np.abs(pred - gt)
    bad2 = float((err > 2.0)[valid].mean() * 100.0)
    epe = float(err[valid].mean())
    net_err = np.abs(res.full.reshape(B, -1).double().cpu().numpy() - gt)
    net_bad2 = float((net_err > 2.0)[valid].mean() * 100.0)
    log("net+band: %.1f pairs/s, bad2 %.2f%%, EPE %.3f (network alone bad2 %.2f%%)"
        % (pairs_s, bad2, epe, net_bad2))

    # full-range SGM control (eval_real.py:143-146): each pair's own centre
    # R from its ground truth, K = 2R + 1, all pairs in one launch, replayed
    # as a HIP graph like the headline
    fullrange = None
    if not args.no_fullrange:
        fullrange = fullrange_control(args, sv, L, gt, valid, L8, R8)
        log("full-range: %.1f pairs/s, bad2 %.2f%%" % (fullrange["pairs_per_s"], fullrange["bad2_pct"]))

    # roofline of the dominant kernel: band SGM aggregation + its epilogue
    # (both views), against the canonical bytes; the two SGM kernels also
    # against their own u8 dataflow
    K = 2 * R_BAND + 1
    sgm_ms, wta_ms = stages["sgm"][0] / args.steps, stages["wta"][0] / args.steps
    canon = canonical_bytes_per_pair(2, K, W * H) * B
    u8flow = u8_dataflow_bytes_per_pair(2, 24, W * H) * B
    t_kern = (sgm_ms + wta_ms) / 1e3
    achieved = canon / t_kern / 1e9 if t_kern > 0 else 0.0
    A_ = importlib.import_module("from-scratch-stereo-vision_amd.autotune")
    split = A_.get_knob("sgm_split") == 2
    order = 2 if A_.get_knob("sgm_order") == 2 else 1
    traffic = SGM_TRAFFIC_PROFILE[1][order] * B if split else None
    roofline = {"bound": "hbm",
                "kernel": "band SGM aggregation + epilogue: %s + %s (8 paths, both views, %d pairs)"
                          % ("k_sgm_ph2 + k_sgm_p2" if split else "k_sgm_ph + k_sgm_p",
                             "k_wta_lds" if A_.get_knob("wta_reads") == 2 else "k_wta_k", B),
                "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": round(traffic) if traffic else None,
                "traffic_source": ("rocprofv3 FETCH_SIZE+WRITE_SIZE of k_sgm_ph2 + k_sgm_p2 from %s "
                                   "(sgm_split=2, sgm_prefetch=8, sgm_htile=16, sgm_order=%d), a "
                                   "constant: not read live" % (SGM_TRAFFIC_PROFILE[0], order))
                if traffic else None,
                "avg_launch_ms": round(t_kern * 1e3, 4),
                "algorithmic_bytes_per_launch": canon,
                "algorithmic_bytes": "canonical dataflow V*HW*K*(8*(1+2+2)+2), V=2, K=21",
                "sgm_kernels_u8_dataflow": {
                    "bytes_per_launch": u8flow, "avg_launch_ms": round(sgm_ms, 4),
                    "achieved": round(u8flow / (sgm_ms / 1e3) / 1e9, 1) if sgm_ms > 0 else None,
                    "frac": round(u8flow / (sgm_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4) if sgm_ms > 0 else None,
                    "traffic_frac": (round(traffic / (sgm_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4)
                                     if traffic and sgm_ms > 0 else None)}}

    # the same workload at other band half-widths and with the C5-trained
    # weights: is full-range bad2 reached at any setting?
    sweep = None
    if not args.no_sweep:
        sweep = band_sweep(args, sv, N, L, L8, R8, gt, valid, B, net._conv_tune)

    # CPU baseline: the float64 C oracle on this host, all cores (one
    # process per core) and one core
    cpu, cpu1 = None, None
    if not args.no_cpu and args.cpu_pairs > 0:
        cpu = cpu_baseline_parallel(args)
        cpu1 = cpu_baseline_one_core(args, sv, pairs)

    equal = None
    if fullrange is not None:
        cands = [s_ for s_ in (sweep or []) if s_["bad2_pct"] <= fullrange["bad2_pct"]]
        best = min((s_["bad2_pct"] for s_ in (sweep or [])), default=bad2)
        equal = {"reached": bool(cands) or bad2 <= fullrange["bad2_pct"],
                 "fullrange_bad2_pct": round(fullrange["bad2_pct"], 3),
                 "best_net_band_bad2_pct": round(min(best, bad2), 3),
                 "note": ("value is the fastest setting at bad2 <= full range" if cands else
                          "no setting of the sweep (R 10-24, v5 and C5-trained weights) reaches "
                          "full-range bad2 on these pairs: the errors are band misses on thin "
                          "foreground boxes (3-8 px wide at 68-80 px disparity) that the network "
                          "does not resolve at its working resolution; value is the reference "
                          "pipeline's own setting (R=10, v5)")}
        if cands:
            fastest = max(cands, key=lambda s_: s_["pairs_per_s"])
            equal["setting"] = fastest

    rec = {
        "metric": METRIC, "value": round(pairs_s, 2), "unit": "pairs/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(dt / args.steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 network / int SGM / f64 epilogue", "data": "synthetic",
        "config": {"workload": "net+band-SGM 460x370: predict_auto (v5 StereoNet, FP32) + "
                               "refine_disparity(planes=False), R=10 census 9x7",
                   "pairs_per_step": B, "parallelism": "single GPU", "execution": execution},
        "eager_pairs_per_s": round(eager_pairs_s, 2),
        "sequential_graph_pairs_per_s": round(seq_graph_pairs_s, 2) if seq_graph_pairs_s else None,
        "pipelined_graph_pairs_per_s": round(pipe_pairs_s, 2) if pipe_pairs_s else None,
        "accuracy": {"bad2_pct": round(bad2, 3), "epe_px": round(epe, 4),
                     "network_only_bad2_pct": round(net_bad2, 3)},
        "fullrange": fullrange,
        "speedup_vs_fullrange": (round(pairs_s / fullrange["pairs_per_s"], 3) if fullrange else None),
        "speedup_note": ("net+band (bad2 %.2f%%) against full range (bad2 %.2f%%): the accuracies "
                         "differ, see equal_bad2" % (bad2, fullrange["bad2_pct"])) if fullrange else None,
        "equal_bad2": equal,
        "sweep": sweep,
        "autotuned": ({k: v for k, v in tuned.items() if not k.startswith("_")} if tuned else None),
        "stage_ms_per_step": {k: round(v[0] / args.steps, 4) for k, v in stages.items()},
        "network_ms_per_step": round(net_ms / args.steps, 4),
        "roofline": roofline,
        "cpu_baseline": cpu,
        "cpu_baseline_one_core": cpu1,
        "gpu_over_cpu": (round(pairs_s / cpu["value"], 1) if cpu else None),
    }
    print(json.dumps(rec), flush=True)
    if args.copy_probe:
        src = torch.ones(64 << 20, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log("copy probe: 268435456 bytes read + 268435456 bytes written")


def pipelined_rate(args, sv, N, L, Rmod, net, L8, R8, d0, lv, out, st):
    """The step software-pipelined over batches: each replay runs the
    network (predict_auto + widening) of batch k on one stream and the
    refinement of batch k-1, whose disparity the previous replay left, on
    another; the two are branches of one HIP graph, so the network's small
    launches fill the gaps beside the SGM.  Two graphs alternate the pair of
    disparity buffers.  Every pair is still fully processed once per batch;
    each graph's refinement must equal the eager step's output."""
    import torch
    B = L8.shape[0]
    sN, sS = torch.cuda.Stream(), torch.cuda.Stream()
    bufs = [d0, torch.empty_like(d0)]

    def net_part(k):
        res = N.predict_auto_batch(net, L8, R8, W, H)
        N._call("sv_widen", N._p(res.full), N._p(lv.table), H * W, N._p(bufs[k % 2]), H * W,
                None, B, H * W, L.stream())

    def piped(k):
        cur = torch.cuda.current_stream()
        sN.wait_stream(cur)
        sS.wait_stream(cur)
        with torch.cuda.stream(sN):
            net_part(k)
        with torch.cuda.stream(sS):
            o = sv.refine_disparity(L8, R8, None, bufs[(k + 1) % 2], W, H, planes=False)
        cur.wait_stream(sN)
        cur.wait_stream(sS)
        return o

    net_part(1)     # the disparity of the batch before the first
    torch.cuda.synchronize()
    G = [Rmod.CapturedPipeline(lambda a, b, k=k: piped(k), L8, R8, warmup=1) for k in (0, 1)]
    for k in range(4):
        go, gs = G[k % 2].replay()
        torch.cuda.synchronize()
        if not (torch.equal(go, out) and torch.equal(gs, st)):
            raise RuntimeError("pipelined graph replay differs from the eager step")
    for k in range(args.warmup):
        G[k % 2].replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        G[k % 2].replay()
    torch.cuda.synchronize()
    return B * args.steps / (time.perf_counter() - t0)


def fullrange_control(args, sv, L, gt, valid, L8, R8):
    """The full-range SGM control of eval_real.py:143-146 on the same pairs:
    R = int(max visible truth + 2) // 2 + 1 per pair, K = 2R + 1, P1=12,
    P2=150, census 9x7, every pair in one launch (sv.full_range with a
    per-pair centre table).  Timed eager (stage split) and as a replayed HIP
    graph like the headline; the replay must equal the eager output."""
    import numpy as np
    import torch
    Rmod = importlib.import_module("from-scratch-stereo-vision_amd.ragged")
    B = L8.shape[0]
    Rs = [int(float(g[v].max()) + 2) // 2 + 1 for g, v in zip(gt, valid)]
    kw = dict(P1=12.0, P2=150.0, rx=4, ry=3)

    def f(a, b):
        return sv.full_range(a, b, W, H, Rs, **kw)

    out = f(L8, R8)
    for _ in range(args.warmup):
        f(L8, R8)
    torch.cuda.synchronize()
    L.lib().sv_profile_enable(1)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        f(L8, R8)
    torch.cuda.synchronize()
    eager_dt = (time.perf_counter() - t0) / args.steps
    st = read_stages(L, ("census", "costs", "sgm", "wta"), args.steps)
    L.lib().sv_profile_enable(0)
    g = Rmod.CapturedPipeline(f, L8, R8, warmup=1)
    gout = g.replay()
    torch.cuda.synchronize()
    if not torch.equal(gout, out):
        raise RuntimeError("full-range graph replay differs from the eager run")
    for _ in range(args.warmup):
        g.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        g.replay()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    fpred = out.reshape(B, -1).cpu().numpy()
    ferr = np.abs(fpred - gt)
    Kmax = 2 * max(Rs) + 1
    kp = 8 * ((Kmax + 1 + 7) // 8) if Kmax <= 125 else None
    for kb in (45, 65, 85, 125):
        if Kmax <= kb:
            kp = 8 * ((kb + 1 + 7) // 8)
            break
    canon = sum(canonical_bytes_per_pair(1, 2 * r + 1, W * H) for r in Rs)
    t_kern = (st["sgm"] + st["wta"]) / 1e3
    u8flow = u8_dataflow_bytes_per_pair(1, kp, W * H) * B if kp else None
    return {"pairs_per_s": B / dt, "eager_pairs_per_s": round(B / eager_dt, 2),
            "execution": "hip graph, one launch per stage for all pairs (per-pair centre table)",
            "bad2_pct": float((ferr > 2.0)[valid].mean() * 100.0),
            "epe_px": float(ferr[valid].mean()),
            "R": sorted(set(Rs)), "K": sorted(set(2 * r + 1 for r in Rs)),
            "kernels": "k_census8, k_costs_cb, k_sgm_ph2/k_sgm_p2 (constant centre), k_wta_lds",
            "stage_ms_per_step": {k: round(v, 4) for k, v in st.items()},
            "roofline": {"bound": "hbm", "kernel": "full-range SGM aggregation + epilogue",
                         "achieved": round(canon / t_kern / 1e9, 1) if t_kern > 0 else None,
                         "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(canon / t_kern / 1e9 / HBM_PEAK_GBS, 4) if t_kern > 0 else None,
                         "algorithmic_bytes_per_launch": canon,
                         "algorithmic_bytes": "canonical dataflow V*HW*K*(8*(1+2+2)+2), V=1, "
                                              "each pair's own K",
                         "avg_launch_ms": round(t_kern * 1e3, 4),
                         "sgm_kernels_u8_dataflow": {
                             "bytes_per_launch": u8flow, "avg_launch_ms": round(st["sgm"], 4),
                             "frac": (round(u8flow / (st["sgm"] / 1e3) / 1e9 / HBM_PEAK_GBS, 4)
                                      if u8flow and st["sgm"] > 0 else None)}}}


def band_sweep(args, sv, N, L, L8, R8, gt, valid, B, conv_tune):
    """net + band refinement (predict_auto + refine_disparity planes=False)
    at band half-widths R = 10, 14, 18, 24 with the v5 weights and with the
    C5-trained weights (the package's weights/): eager pairs/s (one stream)
    and bad2 / EPE of each on the same pairs, with the headline's tuned
    convolution variants.  Only R = 10 has the packed SGM kernels; the wider
    bands run on the general kernels."""
    import numpy as np
    import torch
    dev = L8.device
    lv = N._Level([(H, W)] * B, dev)
    d0 = torch.empty(B, H * W, dtype=torch.float64, device=dev)
    out = []
    for wname, path in (("v5", V5_WEIGHTS), ("c5", C5_WEIGHTS)):
        net = N.load_model(path)
        net._conv_tune = conv_tune
        for R in (10, 14, 18, 24):
            def step():
                res = N.predict_auto_batch(net, L8, R8, W, H)
                N._call("sv_widen", N._p(res.full), N._p(lv.table), H * W, N._p(d0), H * W, None,
                        B, H * W, L.stream())
                return sv.refine_disparity(L8, R8, None, d0, W, H, R=R, planes=False)
            step()
            torch.cuda.synchronize()
            n = max(2, args.steps // 4)
            t0 = time.perf_counter()
            for _ in range(n):
                o, _ = step()
            torch.cuda.synchronize()
            dt = (time.perf_counter() - t0) / n
            e = np.abs(o.reshape(B, -1).cpu().numpy() - gt)
            out.append({"weights": wname, "R": R, "pairs_per_s": round(B / dt, 1),
                        "bad2_pct": round(float((e > 2.0)[valid].mean() * 100.0), 3),
                        "epe_px": round(float(e[valid].mean()), 4),
                        "sgm_kernels": "packed" if R == 10 else "general"})
            log("sweep %s R=%d: %.1f pairs/s, bad2 %.3f%%" % (wname, R, B / dt, out[-1]["bad2_pct"]))
    return out


def _synth_module():
    """The package's synthetic-pair generator, loaded alone (numpy only, no
    GPU library), for CPU worker processes."""
    import importlib.util
    spec = importlib.util.spec_from_file_location(
        "sv_synth", os.path.join(ROOT, "from-scratch-stereo-vision_amd", "synth.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def cpu_worker(spec):
    """One CPU-baseline process: seeds "a,b,..." and a wall-clock start time
    after ';'.  Generates its pairs, waits for the start, runs the float64 C
    oracle (predict_auto + refine_disparity planes=False) on each and prints
    its start and end times."""
    import numpy as np
    import oracle
    seeds, start = spec.split(";")
    synth = _synth_module()
    pairs = [synth.pipeline_pair(int(s)) for s in seeds.split(",")]
    cfg, blocks = oracle.load_weights_txt(V5_WEIGHTS)
    arch = oracle.Arch(cfg["feat_C"], cfg["cost_C"], cfg["D"], cfg["refine_C"], cfg["refine_dil"],
                       cfg["target_peak"])
    P = np.concatenate(blocks)
    ins = [(synth.gray(p["left"]), synth.gray(p["right"])) for p in pairs]
    while time.time() < float(start):
        time.sleep(0.005)
    t0 = time.time()
    for gl, gr in ins:
        full, _, _, _, _ = oracle.predict_auto(arch, P, gl, gr, W, H)
        oracle.refine_disparity_plain(gl, gr, full, W, H)
    print(json.dumps({"t0": t0, "t1": time.time(), "n": len(ins)}), flush=True)


def cpu_baseline_parallel(args):
    """The float64 C oracle in one process per host core (at most 16, the
    box's CPU share for one GPU), two pairs each, all starting at the same
    wall-clock time: pairs/s = pairs / (last end - first start)."""
    import subprocess
    cores = max(1, min(16, os.cpu_count() or 1))
    per = 2
    start = time.time() + 8.0
    procs = []
    for c in range(cores):
        seeds = ",".join(str(c * per + i) for i in range(per))
        procs.append(subprocess.Popen([sys.executable, os.path.abspath(__file__), "--cpu-worker",
                                       "%s;%.3f" % (seeds, start)], stdout=subprocess.PIPE,
                                      stderr=subprocess.DEVNULL))
    res = []
    for p in procs:
        out, _ = p.communicate(timeout=300)
        if p.returncode != 0:
            raise RuntimeError("CPU baseline worker failed (%d)" % p.returncode)
        res.append(json.loads(out.decode().strip().splitlines()[-1]))
    t0 = min(r["t0"] for r in res)
    t1 = max(r["t1"] for r in res)
    n = sum(r["n"] for r in res)
    log("CPU baseline: %d pairs on %d processes in %.2f s" % (n, cores, t1 - t0))
    return {"value": round(n / (t1 - t0), 3), "unit": "pairs/s", "cores": cores, "kind": "port",
            "sample": "%d of the same 460x370 pairs (%d per process, one process per core) through "
                      "the float64 C oracle (predict_auto + refine_disparity planes=False), "
                      "%.2f s wall" % (n, per, t1 - t0)}


def cpu_baseline_one_core(args, sv, pairs):
    """The float64 C oracle on one core in this process (8 pairs)."""
    import numpy as np
    import oracle
    cfg, blocks = oracle.load_weights_txt(V5_WEIGHTS)
    arch = oracle.Arch(cfg["feat_C"], cfg["cost_C"], cfg["D"], cfg["refine_C"], cfg["refine_dil"],
                       cfg["target_peak"])
    P = np.concatenate(blocks)
    n = min(8, len(pairs))
    tc = time.perf_counter()
    for i in range(n):
        gl, gr = sv.synth.gray(pairs[i]["left"]), sv.synth.gray(pairs[i]["right"])
        full, _, _, _, _ = oracle.predict_auto(arch, P, gl, gr, W, H)
        oracle.refine_disparity_plain(gl, gr, full, W, H)
    tcpu = time.perf_counter() - tc
    return {"value": round(n / tcpu, 4), "unit": "pairs/s", "cores": 1, "kind": "port",
            "sample": "%d of the same pairs, one process, %.1f s" % (n, tcpu)}


def net_macs(net, H, W):
    """Multiply-adds of one StereoNet forward at H x W (the layer formulas of
    stereonet.py:459-1268: every conv output voxel costs ic * kd * 9)."""
    def co(n, s, d):
        return (n + 2 * d - (2 * d + 1)) // s + 1
    macs = 0
    h, w = H, W
    for c in net.feat:  # both views
        h, w = co(h, c.s, c.dil), co(w, c.s, c.dil)
        macs += 2 * h * w * c.oc * c.ic * 9
    for c in net.filt:
        macs += h * w * net.D * c.oc * c.ic * 27
    for r, (hh, ww) in enumerate(((H // 2, W // 2), (H, W))):
        for c in net.ref[r]:
            macs += hh * ww * c.oc * c.ic * 9
    return macs


def train_main(args):
    """Configuration 5's training step: StereoNet v5 at 260x192, batch 16,
    train_grads on the batch (forward, losses, hand-written backward) + one
    Adam step (train.train_net's step, jitter on), labels from full-range
    LR SGM on the same synthetic pairs.  value = training pairs/s;
    FLOP/step = 3 x the forward's multiply-adds x 2 (forward, data and
    weight gradients)."""
    import numpy as np
    import torch
    sv = importlib.import_module("from-scratch-stereo-vision_amd")
    L = importlib.import_module("from-scratch-stereo-vision_amd._lib")
    N = importlib.import_module("from-scratch-stereo-vision_amd.net")
    T = importlib.import_module("from-scratch-stereo-vision_amd.train")
    dev = L.device()
    Wt, Ht, Bt = 260, 192, 16
    log("generating 64 planar training pairs %dx%d" % (Wt, Ht))
    d = T.make_dataset([10001 + 2 * i for i in range(64)], Wt, Ht, kind="planar", workers=1)
    disp, valid = T.pseudo_labels(d["left"], d["right"], Wt, Ht, max_disp=42.0)
    left = torch.as_tensor(d["left"]).to(dev)
    right = torch.as_tensor(d["right"]).to(dev)
    net = N.load_model(V5_WEIGHTS)
    warm = max(args.warmup, 1)
    T.train_net(net, left, right, disp, valid, steps=warm, batch=Bt, lr=5e-4, log_every=10 ** 9)
    torch.cuda.synchronize()
    steps = max(args.steps, 10) * 5
    t0 = time.perf_counter()
    hist = T.train_net(net, left, right, disp, valid, steps=steps, batch=Bt, lr=5e-4, seed=1,
                       log_every=steps)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    flops = 3 * 2 * net_macs(net, Ht, Wt) * Bt
    rec = {"metric": "training pairs/s, StereoNet v5 260x192, batch 16 (train_grads + Adam)",
           "value": round(Bt / dt, 2), "unit": "pairs/s", "n_gpus": 1, "steps": steps,
           "warmup": warm, "ms_per_step": round(dt * 1e3, 3), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": "train_net step: 16 planar synthetic pairs 260x192 (peak "
                                  "disparity 22-40), full-range LR SGM pseudo-labels, brightness "
                                  "jitter, lr 5e-4", "pairs_per_step": Bt},
           "flop_per_step": flops, "tflops": round(flops / dt / 1e12, 2),
           "fp32_peak_frac": round(flops / dt / 1e12 / 157.3, 4),
           "loss": hist[-1][1]}
    print(json.dumps(rec), flush=True)


def band_main(args):
    """Plain band refinement only (refine_disparity planes=False: census 9x7,
    band SGM R=10 both views, LR check, median) on 192x144 pairs with the
    coarse estimate ground truth + U{-6..6}, 256 pairs per step; the float64
    C oracle on one core as the CPU baseline, and an exactness check of the
    GPU output against it."""
    import numpy as np
    import torch

    sv = importlib.import_module("from-scratch-stereo-vision_amd")
    L = importlib.import_module("from-scratch-stereo-vision_amd._lib")
    dev = L.device()
    B, Wb, Hb = 256, 192, 144
    pairs = [sv.synth.band_pair(s) for s in range(B)]
    l = torch.from_numpy(np.stack([p["left"] for p in pairs])).to(dev)
    r = torch.from_numpy(np.stack([p["right"] for p in pairs])).to(dev)
    d0 = torch.from_numpy(np.stack([p["disp0"] for p in pairs])).to(dev)
    for _ in range(args.warmup):
        sv.refine_disparity(l, r, None, d0, Wb, Hb, planes=False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out, st = sv.refine_disparity(l, r, None, d0, Wb, Hb, planes=False)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    o = out.reshape(B, -1).cpu().numpy()
    gt = np.stack([p["disp"] for p in pairs]).reshape(B, -1)
    va = np.stack([p["valid"] for p in pairs]).reshape(B, -1)
    rec = {"metric": "pairs/s @192x144, band refinement only (lr_refine + median, R=10)",
           "value": round(B / dt, 2), "unit": "pairs/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(dt * 1e3, 3), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "int SGM / f64 epilogue",
           "data": "synthetic",
           "config": {"workload": "band refinement 192x144, disparity 4-30 px, disp0 = truth + "
                                  "U{-6..6}, census 9x7, P1=12, P2=150", "pairs_per_step": B},
           "accuracy": {"bad2_pct": round(float((np.abs(o - gt) > 2)[va].mean() * 100), 3)}}
    if not args.no_cpu:
        import oracle
        n = min(max(args.cpu_pairs, 1) * 8, B)
        tc = time.perf_counter()
        exact = True
        for i in range(n):
            p = pairs[i]
            ref, rst = oracle.refine_disparity_plain(sv.synth.gray(p["left"]), sv.synth.gray(p["right"]),
                                                     p["disp0"], Wb, Hb)
            exact = exact and np.array_equal(np.asarray(ref).ravel(), o[i]) and np.array_equal(
                np.asarray(rst).ravel(), st.reshape(B, -1)[i].cpu().numpy())
        tcpu = time.perf_counter() - tc
        rec["cpu_baseline"] = {"value": round(n / tcpu, 3), "unit": "pairs/s", "cores": 1,
                               "kind": "port", "sample": "%d of the same pairs through the float64 "
                               "C oracle (refine_disparity planes=False), %.1f s" % (n, tcpu)}
        rec["gpu_equals_oracle"] = bool(exact)
    print(json.dumps(rec), flush=True)


def widened_main(args):
    """The precision variant: the v5 network widened 4x in channels (48/24/32,
    widen_net.py:18-41) on the same 460x370 pairs, once with every
    convolution in FP32 and once with the convolutions of 4k input and 8+
    output channels in BF16 on the matrix cores (FP32 accumulation), each
    followed by band SGM.  Reports both throughputs and the BF16 network's
    per-pixel error and EPE shift against the FP32 one."""
    import numpy as np
    import torch

    sv = importlib.import_module("from-scratch-stereo-vision_amd")
    L = importlib.import_module("from-scratch-stereo-vision_amd._lib")
    N = importlib.import_module("from-scratch-stereo-vision_amd.net")
    dev = L.device()
    B = args.batch
    pairs = [sv.synth.pipeline_pair(s) for s in range(B)]
    L8 = torch.from_numpy(np.stack([p["left"] for p in pairs])).to(dev)
    R8 = torch.from_numpy(np.stack([p["right"] for p in pairs])).to(dev)
    gt = np.stack([p["disp"] for p in pairs]).reshape(B, -1)
    valid = np.stack([p["valid"] for p in pairs]).reshape(B, -1)
    v5 = N.load_model(V5_WEIGHTS)
    net = N.widen(v5, 48, 24, 32)
    lv = N._Level([(H, W)] * B, dev)
    d0 = torch.empty(B, H * W, dtype=torch.float64, device=dev)

    def step():
        res = N.predict_auto_batch(net, L8, R8, W, H)
        N._call("sv_widen", N._p(res.full), N._p(lv.table), H * W, N._p(d0), H * W, None, B,
                H * W, L.stream())
        out, _ = sv.refine_disparity(L8, R8, None, d0, W, H, planes=False)
        return res.full, out, res

    rec = {}
    outs = {}
    traces = {}
    for mode in ("fp32", "bf16"):
        net.use_bf16(mode == "bf16")
        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            full, out, res = step()
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / args.steps
        traces[mode] = res.trace
        f = full.reshape(B, -1).double().cpu().numpy()
        o = out.reshape(B, -1).cpu().numpy()
        outs[mode] = f
        rec[mode] = {"pairs_per_s": round(B / dt, 2), "ms_per_step": round(dt * 1e3, 3),
                     "network_bad2_pct": round(float((np.abs(f - gt) > 2)[valid].mean() * 100), 3),
                     "network_epe_px": round(float(np.abs(f - gt)[valid].mean()), 4),
                     "refined_bad2_pct": round(float((np.abs(o - gt) > 2)[valid].mean() * 100), 3),
                     "refined_epe_px": round(float(np.abs(o - gt)[valid].mean()), 4)}
        log("%s: %.1f pairs/s" % (mode, B / dt))
    net.use_bf16(False)
    e = np.abs(outs["bf16"] - outs["fp32"])
    same = [a == b for a, b in zip(traces["bf16"], traces["fp32"])]
    es = e[np.array(same)] if any(same) else e[:0]
    line = {
        "metric": "pairs/s @460x370, 4x-widened StereoNet (48/24/32), BF16 MFMA vs FP32 network + band SGM",
        "value": rec["bf16"]["pairs_per_s"], "unit": "pairs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": rec["bf16"]["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "bf16 convolutions (fp32 accumulate)",
        "data": "synthetic",
        "config": {"workload": "widened net 48/24/32 (widen_net rule from v5), predict_auto + "
                               "refine_disparity(planes=False), 460x370", "pairs_per_step": B},
        "fp32": rec["fp32"], "bf16": rec["bf16"],
        "bf16_vs_fp32_network_px": {"max": round(float(e.max()), 4), "mean": round(float(e.mean()), 5),
                                    "epe_shift": round(rec["bf16"]["network_epe_px"] -
                                                       rec["fp32"]["network_epe_px"], 4),
                                    "same_width_decisions": "%d/%d" % (sum(same), len(same)),
                                    "max_where_same_widths": round(float(es.max()), 4) if es.size else None,
                                    "mean_where_same_widths": round(float(es.mean()), 5) if es.size else None},
    }
    print(json.dumps(line), flush=True)


def consistent_main(args):
    """predict_consistent (stereonet.py:5049-5076) on B synthetic 460x370
    pairs per step: predict_auto for all pairs, the mirrored pass of every
    pair at its own working width, the left-right check and the row fill
    (predict_consistent_batch).  Reports pairs/s, the three consistency
    kernels' HBM rate, accuracy against the ground truth, and the float64 C
    oracle on a few of the same pairs as the CPU baseline and the check."""
    import numpy as np
    import torch

    sv = importlib.import_module("from-scratch-stereo-vision_amd")
    L = importlib.import_module("from-scratch-stereo-vision_amd._lib")
    N = importlib.import_module("from-scratch-stereo-vision_amd.net")
    C = importlib.import_module("from-scratch-stereo-vision_amd.consistency")
    A = importlib.import_module("from-scratch-stereo-vision_amd.autotune")
    dev = L.device()
    B = args.batch
    pairs = [sv.synth.pipeline_pair(s) for s in range(B)]
    L8 = torch.from_numpy(np.stack([p["left"] for p in pairs])).to(dev)
    R8 = torch.from_numpy(np.stack([p["right"] for p in pairs])).to(dev)
    gt = np.stack([p["disp"] for p in pairs]).reshape(B, -1)
    gvalid = np.stack([p["valid"] for p in pairs]).reshape(B, -1)
    wpath = V5_WEIGHTS
    net = N.load_model(wpath)
    if args.tune_file and os.path.exists(args.tune_file):
        A.load_choices(net, args.tune_file)
    for _ in range(max(args.warmup, 1)):
        out, valid, cnt = C.predict_consistent_batch(net, L8, R8, W, H)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out, valid, cnt = C.predict_consistent_batch(net, L8, R8, W, H)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    # the consistency kernels alone (mirror of both views and of the right
    # map, the check, the fill), timed with HIP events on the stream
    d_l = out.reshape(B, -1).contiguous()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 20
    e0.record()
    for _ in range(reps):
        C._mirror_dev(L8, B * H, W)
        C._mirror_dev(R8, B * H, W)
        d_r = C._mirror_dev(d_l, B * H, W)
        v, _ = C._consistency_dev(d_l, d_r, W, H, 1.5, counts=True)
        C._fill_dev(d_l, v, W, H)
    e1.record()
    torch.cuda.synchronize()
    kms = e0.elapsed_time(e1) / reps
    # algorithmic bytes per pixel: mirrors 2 x (1 + 1) + (4 + 4), check
    # 4 + 4 + 1, fill 4 + 1 + 4 (reads of the row neighbours not counted)
    kbytes = B * W * H * (2 * 2 + 8 + 9 + 9)
    o = out.reshape(B, -1).double().cpu().numpy()
    frac = cnt.cpu().numpy() / float(W * H)
    err = np.abs(o - gt)
    log("predict_consistent: %.1f pairs/s, consistency kernels %.3f ms/step" % (B / dt, kms))
    line = {
        "metric": "pairs/s @460x370: predict_consistent (network LR check)",
        "value": round(B / dt, 2), "unit": "pairs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(dt * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 network / f64 check",
        "data": "synthetic",
        "config": {"workload": "predict_consistent: predict_auto (v5) + mirrored pass at the "
                               "working width + lr_consistency(tol=1.5) + fill_invalid, 460x370",
                   "pairs_per_step": B, "execution": "eager, one stream"},
        "valid_fraction": {"mean": round(float(frac.mean()), 4), "min": round(float(frac.min()), 4)},
        "accuracy": {"bad2_pct": round(float((err > 2.0)[gvalid].mean() * 100), 3),
                     "epe_px": round(float(err[gvalid].mean()), 4)},
        "roofline": {"bound": "hbm", "kernel": "sv_mirror_h x3 + sv_lr_consistency + sv_fill_invalid",
                     "achieved": round(kbytes / (kms * 1e-3) / 1e9, 1), "peak": 8000.0,
                     "unit": "GB/s", "frac": round(kbytes / (kms * 1e-3) / 8e12, 4),
                     "traffic": None, "avg_launch_ms": round(kms, 4),
                     "algorithmic_bytes_per_launch": kbytes},
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        import oracle
        cfg, blocks = oracle.load_weights_txt(wpath)
        arch = oracle.Arch(cfg["feat_C"], cfg["cost_C"], cfg["D"], cfg["refine_C"],
                           cfg["refine_dil"], cfg["target_peak"])
        P = np.concatenate(blocks)
        n = min(max(args.cpu_pairs // 4, 1), B)
        tc = time.perf_counter()
        worst, dfrac = 0.0, 0.0
        for i in range(n):
            p = pairs[i]
            rd, rfrac, _ = oracle.predict_consistent(arch, P, sv.synth.gray(p["left"]),
                                                     sv.synth.gray(p["right"]), W, H)
            worst = max(worst, float((np.abs(o[i] - rd) > 2e-3).mean()))
            dfrac = max(dfrac, abs(float(frac[i]) - rfrac))
        tcpu = time.perf_counter() - tc
        line["cpu_baseline"] = {"value": round(n / tcpu, 3), "unit": "pairs/s", "cores": 1,
                                "kind": "port",
                                "sample": "%d of the same pairs through the float64 C oracle "
                                          "(predict_consistent), %.1f s" % (n, tcpu)}
        line["gpu_over_cpu"] = round((B / dt) / (n / tcpu), 1)
        line["check_vs_oracle"] = {"pairs": n, "max_valid_fraction_diff": round(dfrac, 5),
                                   "max_share_beyond_2e-3_px": round(worst, 5)}
    print(json.dumps(line), flush=True)


def planes_main(args):
    """refine_disparity(planes=True) (stereonet.py:5865-5884) on B synthetic
    460x370 pairs per step with an RGB image whose colours follow the
    disparity's planar regions: lr_refine of both views, SLIC + RANSAC plane
    snapping of the status-0 pixels, one 3x3 median.  Reports pairs/s, the
    plane-snap stage time (HIP events), accuracy, and the float64 C oracle on
    a few of the same pairs as the CPU baseline and the exactness check."""
    import numpy as np
    import torch

    sv = importlib.import_module("from-scratch-stereo-vision_amd")
    L = importlib.import_module("from-scratch-stereo-vision_amd._lib")
    Rf = importlib.import_module("from-scratch-stereo-vision_amd.refine")
    dev = L.device()
    B = args.batch
    pairs = [sv.synth.band_pair(s, W=W, H=H) for s in range(B)]

    def rgb_of(p):
        g = p["left"].astype(np.float64) / 255.0
        t = np.clip(p["disp"] / 40.0, 0.0, 1.0)
        rgb = np.stack([0.6 * g + 0.4 * t, 0.3 * g + 0.7 * (1.0 - t), 0.5 * g + 0.5 * t * t], -1)
        return np.round(rgb * 255.0).astype(np.uint8)

    rgbs = [rgb_of(p) for p in pairs]
    L8 = torch.from_numpy(np.stack([p["left"] for p in pairs])).to(dev)
    R8 = torch.from_numpy(np.stack([p["right"] for p in pairs])).to(dev)
    C8 = torch.from_numpy(np.stack(rgbs)).to(dev)
    D0 = torch.from_numpy(np.stack([np.asarray(p["disp0"], dtype=np.float64) for p in pairs])).to(dev)
    gt = np.stack([p["disp"] for p in pairs]).reshape(B, -1)
    gvalid = np.stack([p["valid"] for p in pairs]).reshape(B, -1)

    def step():
        return sv.refine_disparity(L8, R8, C8, D0, W, H, planes=True)

    for _ in range(max(args.warmup, 1)):
        out, st = step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out, st = step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    # the plane-snap stage alone, on the lr_refine output of the last step
    d, st1 = sv.lr_refine(L8, R8, D0, W, H, tol=2.0, R=10, P1=12.0, P2=150.0, rx=4, ry=3)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 10
    e0.record()
    for _ in range(reps):
        Rf._snap_dev(d.reshape(-1).contiguous(), C8, L.SV_U8, B, W, H,
                     st1.reshape(-1).contiguous(), 1, 12, 1.0, 0.4, 40, 12345)
    e1.record()
    torch.cuda.synchronize()
    snap_ms = e0.elapsed_time(e1) / reps
    o = out.reshape(B, -1).cpu().numpy()
    plain, _ = sv.refine_disparity(L8, R8, None, D0, W, H, planes=False)
    po = plain.reshape(B, -1).cpu().numpy()
    err, perr = np.abs(o - gt), np.abs(po - gt)
    log("refine_disparity(planes=True): %.1f pairs/s, plane_snap %.3f ms/step" % (B / dt, snap_ms))
    line = {
        "metric": "pairs/s @460x370: refine_disparity(planes=True)",
        "value": round(B / dt, 2), "unit": "pairs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(dt * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "int SGM / f64 planes",
        "data": "synthetic",
        "config": {"workload": "refine_disparity(planes=True): lr_refine R=10 census 9x7 + "
                               "plane_snap(step 12, thr 1, min_inl 0.4) + median, 460x370, u8 RGB",
                   "pairs_per_step": B, "execution": "eager, one stream"},
        "plane_snap_ms_per_step": round(snap_ms, 3),
        "accuracy": {"bad2_pct": round(float((err > 2.0)[gvalid].mean() * 100), 3),
                     "epe_px": round(float(err[gvalid].mean()), 4),
                     "planes_false_bad2_pct": round(float((perr > 2.0)[gvalid].mean() * 100), 3),
                     "planes_false_epe_px": round(float(perr[gvalid].mean()), 4)},
        "roofline": None,
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        import oracle
        n = min(max(args.cpu_pairs // 6, 1), B)
        tc = time.perf_counter()
        same = 0
        for i in range(n):
            p = pairs[i]
            rd, rst = oracle.refine_disparity_planes(
                sv.synth.gray(p["left"]), sv.synth.gray(p["right"]),
                rgbs[i].astype(np.float64).ravel() * (1.0 / 255.0), p["disp0"], W, H)
            same += int(np.array_equal(rd, o[i]) and
                        np.array_equal(rst, st.reshape(B, -1)[i].cpu().numpy().astype(np.int32)))
        tcpu = time.perf_counter() - tc
        line["cpu_baseline"] = {"value": round(n / tcpu, 3), "unit": "pairs/s", "cores": 1,
                                "kind": "port",
                                "sample": "%d of the same pairs through the float64 C oracle "
                                          "(refine_disparity planes=True), %.1f s" % (n, tcpu)}
        line["gpu_over_cpu"] = round((B / dt) / (n / tcpu), 1)
        line["identical_to_oracle"] = "%d/%d pairs" % (same, n)
    print(json.dumps(line), flush=True)


def _png_rgb(rgb):
    """An 8-bit RGB PNG of rgb [H][W][3] uint8: rows filtered None, Sub, Up,
    Average, Paeth in turn (every unfilter path), zlib level 6."""
    import zlib
    import numpy as np
    Hh, Ww, _ = rgb.shape
    a = rgb.reshape(Hh, Ww * 3).astype(np.int64)
    out = bytearray()
    for y in range(Hh):
        f = y % 5
        cur = a[y]
        left = np.concatenate([np.zeros(3, np.int64), cur[:-3]])
        up = a[y - 1] if y else np.zeros_like(cur)
        ul = np.concatenate([np.zeros(3, np.int64), up[:-3]])
        if f == 4:
            p = left + up - ul
            pa, pb, pc = np.abs(p - left), np.abs(p - up), np.abs(p - ul)
            pred = np.where((pa <= pb) & (pa <= pc), left, np.where(pb <= pc, up, ul))
        else:
            pred = [np.zeros_like(cur), left, up, (left + up) >> 1][f]
        out.append(f)
        out += ((cur - pred) & 0xFF).astype(np.uint8).tobytes()

    def chunk(t, b):
        return len(b).to_bytes(4, "big") + t + b + (zlib.crc32(t + b) & 0xFFFFFFFF).to_bytes(4, "big")
    ihdr = Ww.to_bytes(4, "big") + Hh.to_bytes(4, "big") + bytes([8, 2, 0, 0, 0])
    return (b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", ihdr) + chunk(b"IDAT", zlib.compress(bytes(out), 6))
            + chunk(b"IEND", b""))


def images_main(args):
    """Input decoding (read_image, stereonet.py:3499-3613): B synthetic
    460x370 stereo pairs stored as 8-bit RGB PNG files (all five scanline
    filters, dynamic-Huffman DEFLATE) decoded per step into float64 gray
    and RGB planes on the GPU by read_images_gpu -- host inflate and
    unfilter on --threads threads, one pinned upload, the expand kernel.
    Reports pairs/s end to end from file bytes in host memory; the roofline
    is the expand kernel's (HIP events); the CPU baseline is the oracle's
    decoder (Python zlib + numpy) on one core, which also checks a sample
    of the outputs for exact equality."""
    import numpy as np
    import torch

    sv = importlib.import_module("from-scratch-stereo-vision_amd")
    L = importlib.import_module("from-scratch-stereo-vision_amd._lib")
    B = args.batch
    pairs = [sv.synth.band_pair(s, W=W, H=H) for s in range(B)]

    def rgb_of(img, disp):
        g = img.astype(np.float64) / 255.0
        t = np.clip(disp / 40.0, 0.0, 1.0)
        rgb = np.stack([0.6 * g + 0.4 * t, 0.3 * g + 0.7 * (1.0 - t), 0.5 * g + 0.5 * t * t], -1)
        return np.round(rgb * 255.0).astype(np.uint8)

    lf = [_png_rgb(rgb_of(p["left"], p["disp"])) for p in pairs]
    rf = [_png_rgb(rgb_of(p["right"], p["disp"])) for p in pairs]
    log("images: %d pairs, %.1f KiB per PNG" % (B, sum(map(len, lf + rf)) / 2048.0 / B))

    def step():
        gl, cl, _, _ = sv.read_images_gpu(lf, threads=args.threads)
        gr, cr, _, _ = sv.read_images_gpu(rf, threads=args.threads)
        return gl, cl, gr, cr

    for _ in range(max(args.warmup, 1)):
        out = step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out = step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    # the expand kernel alone on resident samples, HIP events on its stream
    npix = B * W * H
    samp = torch.randint(0, 256, (3 * npix,), dtype=torch.int16, device=L.device())
    gray = torch.empty(npix, dtype=torch.float64, device=L.device())
    rgb = torch.empty(3 * npix, dtype=torch.float64, device=L.device())
    st = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 20
    e0.record(st)
    for _ in range(reps):
        L.check(L.lib().sv_image_expand(L.ptr(samp), npix, 0, 255, L.ptr(gray), L.ptr(rgb),
                                        L.stream()))
    e1.record(st)
    torch.cuda.synchronize()
    k_ms = e0.elapsed_time(e1) / reps
    bytes_per_launch = 38.0 * npix  # 3 u16 samples in, 1 + 3 float64 out per pixel
    achieved = bytes_per_launch / (k_ms * 1e-3) / 1e9
    log("images: %.1f pairs/s, expand %.3f ms per %d px" % (B / dt, k_ms, npix))
    line = {
        "metric": "pairs/s @460x370: read_image (PNG RGB8) -> float64 gray + RGB on the GPU",
        "value": round(B / dt, 2), "unit": "pairs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(dt * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u8 -> f64", "data": "synthetic",
        "config": {"workload": "read_images_gpu: %d left + %d right 460x370 RGB8 PNG files "
                               "(filters 0-4, zlib level 6) per step" % (B, B),
                   "pairs_per_step": B, "host_threads": args.threads},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": 8000.0, "unit": "GB/s",
                     "frac": round(achieved / 8000.0, 4), "traffic": None,
                     "kernel": "k_image_expand", "kernel_ms": round(k_ms, 4),
                     "algorithmic_bytes_per_launch": int(bytes_per_launch)},
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        import oracle
        n = min(max(args.cpu_pairs // 4, 1), B)
        gl, cl, gr, cr = out
        tc = time.perf_counter()
        same = 0
        for i in range(n):
            g0, r0, _, _ = oracle.read_image_bytes(lf[i])
            g1, r1, _, _ = oracle.read_image_bytes(rf[i])
            same += int(np.array_equal(g0, gl[i].cpu().numpy().ravel())
                        and np.array_equal(r0, cl[i].cpu().numpy().ravel())
                        and np.array_equal(g1, gr[i].cpu().numpy().ravel())
                        and np.array_equal(r1, cr[i].cpu().numpy().ravel()))
        tcpu = time.perf_counter() - tc
        line["cpu_baseline"] = {"value": round(n / tcpu, 3), "unit": "pairs/s", "cores": 1,
                                "kind": "port",
                                "sample": "%d of the same pairs through the oracle's decoder "
                                          "(zlib + numpy unfilter), %.1f s incl. the comparison"
                                          % (n, tcpu)}
        line["identical_to_oracle"] = "%d/%d pairs" % (same, n)
    print(json.dumps(line), flush=True)


def eval_main(args):
    """Middlebury scoring (eval_real.py:78-127) of B synthetic 460x370 pairs
    per step: both views' ground-truth PNGs (grey level = 3 x disparity)
    decoded by read_images_gpu, ground truth and known masks, the
    non-occluded mask, then EPE / bad1 / bad2 / bad4 over the non-occluded
    and known masks and the constant baseline, for float32 predictions
    already on the GPU.  Reports pairs/s (decode included) and the scoring
    kernels alone; roofline on k_score (HIP events); CPU baseline the
    oracle's scoring on a few of the same pairs."""
    import numpy as np
    import torch

    sv = importlib.import_module("from-scratch-stereo-vision_amd")
    L = importlib.import_module("from-scratch-stereo-vision_amd._lib")
    ev = sv.evaluation
    B = args.batch
    pairs = [sv.synth.band_pair(s, W=W, H=H) for s in range(B)]

    def gt_png(d, valid):
        lv = np.where(valid, np.clip(np.round(d * 3.0), 0, 255), 0).astype(np.uint8)
        return _png_rgb(np.repeat(lv[:, :, None], 3, axis=2))

    fl, fr = [], []
    for p in pairs:
        d = np.asarray(p["disp"], np.float64).reshape(H, W)
        fl.append(gt_png(d, np.asarray(p["valid"], bool).reshape(H, W)))
        dr = np.asarray(sv.left_to_right(torch.from_numpy(d.ravel()).cuda(), W, H).cpu()).reshape(H, W)
        fr.append(gt_png(dr, dr > 0))
    rng = np.random.default_rng(1)
    pred = torch.from_numpy(np.stack([np.asarray(p["disp"], np.float64).reshape(H, W)
                                      + rng.normal(0, 1.5, (H, W)) for p in pairs])).float().cuda()

    def scoring(gl, gr):
        dl, kl = ev.gt_from_gray(gl, 3.0)
        dr, kr = ev.gt_from_gray(gr, 3.0)
        no = ev.nonocc_mask(dl, kl, dr, kr, W, H)
        s_no, g, m = ev.score_stats(pred, dl, no, B=B)
        s_all, _, _ = ev.score_stats(pred, dl, kl, B=B)
        ad = ev.abs_dev_sums(g, m, s_no, B=B)
        return dl, kl, no, s_no, s_all, ad

    def step():
        gl = sv.read_images_gpu(fl, threads=args.threads)[0]
        gr = sv.read_images_gpu(fr, threads=args.threads)[0]
        return scoring(gl, gr)

    for _ in range(max(args.warmup, 1)):
        out = step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out = step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    gl = sv.read_images_gpu(fl, threads=args.threads)[0]
    gr = sv.read_images_gpu(fr, threads=args.threads)[0]
    st = torch.cuda.current_stream()
    e0, e1, e2 = (torch.cuda.Event(enable_timing=True) for _ in range(3))
    reps = 20
    e0.record(st)
    for _ in range(reps):
        scoring(gl, gr)
    e1.record(st)
    dl, kl, no = out[0], out[1], out[2]
    for _ in range(reps):
        ev.score_stats(pred, dl, no, B=B)
    e2.record(st)
    torch.cuda.synchronize()
    sc_ms = e0.elapsed_time(e1) / reps
    k_ms = e1.elapsed_time(e2) / reps  # k_score + its 1-block partial sum
    npix = B * W * H
    bytes_per_launch = 13.0 * npix  # f32 prediction, float64 ground truth, u8 mask
    achieved = bytes_per_launch / (k_ms * 1e-3) / 1e9
    log("eval: %.1f pairs/s (scoring alone %.3f ms per step)" % (B / dt, sc_ms))
    line = {
        "metric": "pairs/s @460x370: Middlebury scoring (load_gt + masks + score + constant)",
        "value": round(B / dt, 2), "unit": "pairs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(dt * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "eval_real scoring: %d pairs of 460x370 ground-truth PNGs decoded, "
                               "masks, EPE/bad1/2/4 over nonocc and all, constant" % B,
                   "pairs_per_step": B, "host_threads": args.threads},
        "scoring_ms_per_step": round(sc_ms, 4),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": 8000.0, "unit": "GB/s",
                     "frac": round(achieved / 8000.0, 4), "traffic": None,
                     "kernel": "k_score<float> (+ k_sum_parts)", "kernel_ms": round(k_ms, 4),
                     "algorithmic_bytes_per_launch": int(bytes_per_launch)},
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        import oracle
        n = min(max(args.cpu_pairs // 4, 1), B)
        s_no, s_all, ad = out[3].cpu().numpy(), out[4].cpu().numpy(), out[5].cpu().numpy()
        tc = time.perf_counter()
        same = 0
        for i in range(n):
            g0, _, _, _ = oracle.read_image_bytes(fl[i])
            g1, _, _, _ = oracle.read_image_bytes(fr[i])
            d0, k0 = oracle.gt_from_gray(g0, 3.0)
            d1, k1 = oracle.gt_from_gray(g1, 3.0)
            n0 = oracle.nonocc(d0, k0, d1, k1, W, H)
            p = pred[i].double().cpu().numpy().ravel()
            r_no, r_all, c = oracle.score(p, d0, n0), oracle.score(p, d0, k0), oracle.constant(d0, n0)
            s, a = s_no[i], s_all[i]
            same += int(s[0] == int(n0.sum()) and abs(s[1] / s[0] - r_no[0]) <= 1e-12 * r_no[0]
                        and 100.0 * s[3] / s[0] == r_no[2] and 100.0 * a[3] / a[0] == r_all[2]
                        and abs(ad[i] / s[0] - c) <= 1e-12 * c)
        tcpu = time.perf_counter() - tc
        line["cpu_baseline"] = {"value": round(n / tcpu, 3), "unit": "pairs/s", "cores": 1,
                                "kind": "port",
                                "sample": "%d of the same pairs through the oracle (PNG decode, "
                                          "masks, sequential sums), %.1f s" % (n, tcpu)}
        line["identical_to_oracle"] = "%d/%d pairs (counts exact, sums within 1e-12)" % (same, n)
    print(json.dumps(line), flush=True)


def splat_main(args):
    """The Gaussian-splatting reconstruction of one 460x370 scene
    (reconstruct, stereonet.py:3139-3187, without its file output): the
    disparity unprojected, 520 mesh vertices selected, Delaunay + cull,
    one Gaussian per triangle, 60 photometric refinement iterations (render
    + loss + backward + CloudAdam), the original and five novel renders.
    value = refinement iterations per second (the dominant loop); the
    reconstruction's end-to-end time beside it.  CPU baseline: the oracle's
    numpy restatement of the renderer (one core), renders per second, on the
    same cloud; it must equal the GPU render."""
    import numpy as np
    import torch

    sv = importlib.import_module("from-scratch-stereo-vision_amd")
    p = sv.synth.pipeline_pair(0)
    disp = torch.from_numpy(np.asarray(p["disp"], np.float64).reshape(-1)).cuda()
    g = np.asarray(p["left"], np.float64).reshape(-1) / 255.0
    y, x = np.mgrid[0:H, 0:W]
    rgb_np = np.stack([g, np.clip(g * 0.8 + 0.2 * (x.ravel() / W), 0, 1),
                       np.clip(g * 0.7 + 0.3 * (y.ravel() / H), 0, 1)], -1).reshape(-1)
    rgb = torch.from_numpy(rgb_np).cuda()
    cam = sv.Camera.synthetic(W, H)
    iters = 60

    def build():
        pts3d = sv.geometry.disparity_to_points_gpu(disp, W, H, cam, min_disp=0.4)
        pts2d = sv.select_points(disp, W, H, None, target=520)
        tris = sv.cull_triangles(pts2d, sv.delaunay(pts2d), pts3d, W)
        return sv.gaussians_from_mesh(pts2d, tris, pts3d, rgb, W, H), len(pts2d), len(tris)

    def recon():
        gc, nv, nt = build()
        r = sv.SplatRenderer(cam, W, H)
        sv.refine_gaussians(gc, r, rgb, iters=iters, verbose=False)
        C0 = r.render(gc, sv.IDENTITY_R, sv.IDENTITY_T)
        for _, (Rm, t) in sv.orbit_views(gc, n_views=5):
            r.render(gc, Rm, t)
        return gc, C0, nv, nt

    for _ in range(max(args.warmup, 1)):
        recon()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    reps = max(1, min(args.steps, 5))
    for _ in range(reps):
        gc, C0, nv, nt = recon()
    torch.cuda.synchronize()
    t_recon = (time.perf_counter() - t0) / reps
    # the refinement loop alone
    gc, nv, nt = build()
    r = sv.SplatRenderer(cam, W, H)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    hist = sv.refine_gaussians(gc, r, rgb, iters=iters, verbose=False)
    torch.cuda.synchronize()
    t_ref = time.perf_counter() - t0
    # render alone (HIP events on the launching stream)
    st = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(20):
        C = r.render(gc, sv.IDENTITY_R, sv.IDENTITY_T)
    e1.record(st)
    torch.cuda.synchronize()
    r_ms = e0.elapsed_time(e1) / 20
    log("splat: %d vertices, %d triangles, %d Gaussians; reconstruction %.1f ms, refinement %.2f "
        "it/s, render %.3f ms" % (nv, nt, gc.count(), t_recon * 1e3, iters / t_ref, r_ms))
    line = {
        "metric": "refinement iterations/s @460x370: Gaussian splatting of the disparity "
                  "(render + loss + backward + Adam)",
        "value": round(iters / t_ref, 2), "unit": "iterations/s", "n_gpus": 1, "steps": iters,
        "warmup": args.warmup, "ms_per_step": round(t_ref / iters * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "reconstruct (no file output) of one 460x370 scene: 520 vertices, "
                               "Delaunay + cull, %d Gaussians, %d refinement iterations, 1 + 5 "
                               "renders" % (gc.count(), iters),
                   "gaussians": gc.count(), "triangles": nt},
        "reconstruction_ms": round(t_recon * 1e3, 2),
        "render_ms": round(r_ms, 4),
        "final_mse": round(float(hist[-1]), 6),
        "roofline": None,
        "roofline_note": "per-pixel front-to-back compositing over a depth-ordered list in "
                         "float64 with the reference's series exp: latency-bound serial work per "
                         "pixel, no HBM or matrix-core bound applies",
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        import oracle
        cl = gc.arrays()
        ident = [1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0]
        tc = time.perf_counter()
        n = 0
        while True:
            ref = oracle.splat_render(cl, (cam.fx, cam.cx, cam.cy), W, H, ident, [0.0, 0.0, 0.0])
            n += 1
            if time.perf_counter() - tc > 10.0:
                break
        tcpu = (time.perf_counter() - tc) / n
        line["cpu_baseline"] = {"value": round(1.0 / tcpu, 3), "unit": "renders/s", "cores": 1,
                                "kind": "port",
                                "sample": "%d renders of the refined cloud through the oracle's "
                                          "numpy renderer, %.1f s" % (n, tcpu * n)}
        line["gpu_render_over_cpu"] = round(tcpu / (r_ms * 1e-3), 1)
        line["identical_to_oracle"] = bool(np.array_equal(ref, C.cpu().numpy()))
    print(json.dumps(line))


def mesh_main(args):
    """The viewer's grid mesh (grid_mesh, stereonet.py:5770-5862; step 2 with
    the background fill) for B refined-quality 460x370 disparity maps per
    step (synthetic truth + noise): robust_min_disparity, the main grid,
    background_region, the fill grid and the pivot for all B maps with one
    launch per stage (grid_mesh_batch); the counts and pivots are read back
    once per step.  Roofline on the
    row min filter (k_minrow, radius W/16 = 28) timed alone with HIP events
    over all B maps in one launch; CPU baseline the oracle's grid_mesh."""
    import numpy as np
    import torch

    sv = importlib.import_module("from-scratch-stereo-vision_amd")
    L = importlib.import_module("from-scratch-stereo-vision_amd._lib")
    G = sv.geometry
    B = args.batch
    rng = np.random.default_rng(2)
    maps = np.stack([np.asarray(sv.synth.pipeline_pair(s)["disp"], np.float64).reshape(H, W)
                     + rng.normal(0, 0.1, (H, W)) for s in range(B)])
    D = torch.from_numpy(maps).cuda()
    cam = sv.Camera.synthetic(W, H)

    def step():
        return G.grid_mesh_batch(D, W, H, cam, step=2, fill=True)

    for _ in range(max(args.warmup, 1)):
        out = step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out = step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    flat = D.reshape(-1)
    mo = torch.empty_like(flat)
    mi = torch.empty(flat.numel(), dtype=torch.int32, device=flat.device)
    st = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 20
    r = int(W / 16)
    e0.record(st)
    for _ in range(reps):
        L.check(L.lib().sv_min_filter_arg(L.ptr(flat), B, W, H, r, 0, L.ptr(mo), L.ptr(mi), None,
                                          None, L.stream()))
    e1.record(st)
    torch.cuda.synchronize()
    k_ms = e0.elapsed_time(e1) / reps
    bytes_per_launch = 20.0 * B * W * H  # disparity in, minimum and source index out
    achieved = bytes_per_launch / (k_ms * 1e-3) / 1e9
    nv = sum(g["pos"].shape[0] for g in out) / B
    nt = sum(g["main_tris"] + g["fill_tris"] for g in out) / B
    log("mesh: %.1f maps/s, %.0f vertices and %.0f triangles per map; k_minrow %.3f ms"
        % (B / dt, nv, nt, k_ms))
    line = {
        "metric": "pairs/s @460x370: grid_mesh (step 2, background fill) of the refined disparity",
        "value": round(B / dt, 2), "unit": "pairs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(dt * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "grid_mesh(step=2, fill=True) of %d 460x370 disparity maps per "
                               "step in one batch (counts read back once)" % B,
                   "pairs_per_step": B, "vertices_per_map": round(nv), "triangles_per_map": round(nt)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": 8000.0, "unit": "GB/s",
                     "frac": round(achieved / 8000.0, 4), "traffic": None,
                     "kernel": "k_minrow (r = 28, %d maps per launch)" % B,
                     "kernel_ms": round(k_ms, 4),
                     "algorithmic_bytes_per_launch": int(bytes_per_launch)},
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        import oracle
        n = min(max(args.cpu_pairs // 8, 1), B)
        camt = (cam.fx, cam.cx, cam.cy, cam.baseline, cam.doffs)
        tc = time.perf_counter()
        same = 0
        for i in range(n):
            ref = oracle.grid_mesh(maps[i].ravel(), W, H, camt, step=2, fill=True)
            g = out[i]
            same += int(all(np.array_equal(g[f].cpu().numpy(), ref[f])
                            for f in ("pos", "uv", "disp", "fill", "idx"))
                        and g["pivot"] == ref["pivot"])
        tcpu = time.perf_counter() - tc
        line["cpu_baseline"] = {"value": round(n / tcpu, 3), "unit": "pairs/s", "cores": 1,
                                "kind": "port",
                                "sample": "%d of the same maps through the oracle's grid_mesh "
                                          "(numpy filters, Python vertex loops), %.1f s" % (n, tcpu)}
        line["identical_to_oracle"] = "%d/%d maps" % (same, n)
    print(json.dumps(line), flush=True)


def ragged_fullrange(args, sv, b, l, r, pairs, spec):
    """The full-range control of eval_real.py:143-146 on the same 64 uneven
    pairs: each pair's own R = int(max visible truth + 2) // 2 + 1 (K = 2R +
    1, up to 245 here), P1=12, P2=150, census 9x7, one sv.full_range call
    per pair (their sizes differ), the 64 calls captured in one HIP graph
    and replayed like the net + band step."""
    import numpy as np
    import torch
    RG = importlib.import_module("from-scratch-stereo-vision_amd.ragged")
    Rs = [int(float(np.asarray(p["disp"])[np.asarray(p["valid"])].max()) + 2) // 2 + 1
          for p in pairs]
    ls, rs = b.unpack(l), b.unpack(r)
    dims = [(w, h) for w, h, _ in spec]

    def f(a, c):
        A_, C_ = b.unpack(a), b.unpack(c)
        return [sv.full_range(A_[n], C_[n], w, h, Rs[n], P1=12.0, P2=150.0, rx=4, ry=3)
                for n, (w, h) in enumerate(dims)]

    log("full-range control: 64 calls, K up to %d" % (2 * max(Rs) + 1))
    eager = [o.clone() for o in f(l, r)]
    torch.cuda.synchronize()
    g = RG.CapturedPipeline(f, l, r, warmup=1)
    outs = g.replay()
    torch.cuda.synchronize()
    if not all(torch.equal(u, v) for u, v in zip(outs, eager)):
        raise RuntimeError("full-range graph replay differs from the eager run")
    steps = max(2, args.steps // 2)
    t0 = time.perf_counter()
    for _ in range(steps):
        g.replay()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    bad = []
    for o, p in zip(outs, pairs):
        e = np.abs(o.cpu().numpy().reshape(np.asarray(p["disp"]).shape) - p["disp"])
        bad.append(float((e > 2.0)[p["valid"]].mean() * 100.0))
    return {"pairs_per_s": round(64 / dt, 2), "ms_per_step": round(dt * 1e3, 3),
            "execution": "hip graph of 64 sv.full_range calls, one per pair",
            "K_max": 2 * max(Rs) + 1, "K_mean": round(float(np.mean([2 * v + 1 for v in Rs])), 1),
            "bad2_pct_mean": round(float(np.mean(bad)), 3),
            "bad2_pct_max": round(float(np.max(bad)), 3),
            "kernels": "constant-centre packed kernels for K <= 125, the general band kernels above"}


def ragged_fallback_sweep(args, RG, b, l, r, cap, pairs, spec):
    """Net + band with full-range fallback (RG.fallback_ragged) on the
    pairs whose share of occluded or mismatched pixels exceeds tau, whole
    pairs replaced, each pair's full range from its declared maximum
    disparity: bad2 and time per step (graph replay of net + band, then the
    host's choice of pairs and their full-range calls) for a range of tau."""
    import numpy as np
    import torch
    md = [d for _, _, d in spec]
    res = []
    for tau in (1.01, 0.3, 0.2, 0.15, 0.1, 0.05, -1.0):
        best = None
        for _ in range(2):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            out, st = cap.replay()
            picked = RG.fallback_ragged(b, l, r, out, st, md, tau=tau)
            torch.cuda.synchronize()
            t = time.perf_counter() - t0
            best = t if best is None or t < best else best
        outs = [o.cpu().numpy() for o in b.unpack(out)]
        bad = [float((np.abs(o - p["disp"]) > 2.0)[p["valid"]].mean() * 100.0)
               for o, p in zip(outs, pairs)]
        res.append({"tau": tau, "fallback_pairs": len(picked),
                    "fallback_px_frac": round(sum(spec[n][0] * spec[n][1] for n in picked) /
                                              float(sum(w * h for w, h, _ in spec)), 4),
                    "ms_per_step": round(best * 1e3, 3), "pairs_per_s": round(64 / best, 2),
                    "bad2_pct_mean": round(float(np.mean(bad)), 3)})
        log("fallback tau=%.2f: %d pairs, %.1f ms, bad2 %.2f%%" % (
            tau, len(picked), best * 1e3, res[-1]["bad2_pct_mean"]))
    return res


def ragged_main(args):
    """64 synthetic pairs of uneven size (200x150 up to 1390x1110, maximum
    disparity 16-240 px) through net + band SGM as one ragged batch: the
    whole pipeline (predict_auto, widening, refine_disparity planes=False)
    is captured once in a HIP graph, and a step is one replay."""
    import numpy as np
    from multiprocessing import Pool

    S = importlib.import_module("from-scratch-stereo-vision_amd.synth")
    spec = S.ragged_spec(64)
    log("generating 64 uneven pairs")
    with Pool(min(16, os.cpu_count() or 1)) as pool:  # before this process touches the GPU
        pairs = pool.map(S.ragged_pair, [(w, h, d, 1000 + i) for i, (w, h, d) in enumerate(spec)])
    import torch

    sv = importlib.import_module("from-scratch-stereo-vision_amd")
    L = importlib.import_module("from-scratch-stereo-vision_amd._lib")
    N = importlib.import_module("from-scratch-stereo-vision_amd.net")
    RG = importlib.import_module("from-scratch-stereo-vision_amd.ragged")
    b = RG.RaggedBatch([(w, h) for w, h, _ in spec])
    l = b.pack([p["left"] for p in pairs])
    r = b.pack([p["right"] for p in pairs])
    net = N.load_model(V5_WEIGHTS)
    A = importlib.import_module("from-scratch-stereo-vision_amd.autotune")

    def run():
        o, s_, _ = RG.net_band_ragged(net, b, l, r)
        return b.unpack(o) + b.unpack(s_)

    tuned = None
    if args.tune_file and os.path.exists(args.tune_file):
        tuned = A.load_choices(net, args.tune_file)
        log("kernel variants from %s" % args.tune_file)
    elif not args.no_autotune:
        layers = A.tune_layers(net, run, reps=2)
        tuned = A.tune(run, knobs=["sgm_split", "sgm_prefetch", "sgm_htile", "sgm_overlap", "wta_reads", "sgm_order", "sgm_queue"], reps=2, log=log)
        tuned["conv_layers_tuned"] = len(layers)
        if args.tune_file:
            A.save_choices(net, {k: v for k, v in tuned.items() if k in A.KNOBS}, args.tune_file)
            log("kernel variants written to %s" % args.tune_file)
    # eager reference run (stage profile of the refinement) and capture
    L.lib().sv_profile_enable(1)
    eout, est, eres = RG.net_band_ragged(net, b, l, r)
    torch.cuda.synchronize()
    import ctypes
    stages = {}
    for name in ("census", "costs", "sgm", "wta", "lr", "median"):
        ms, cnt = ctypes.c_double(0), ctypes.c_int(0)
        L.check(L.lib().sv_profile_read(name.encode(), ctypes.byref(ms), ctypes.byref(cnt)))
        stages[name] = round(ms.value, 4)
    L.lib().sv_profile_enable(0)
    t0 = time.perf_counter()
    for _ in range(2):
        RG.net_band_ragged(net, b, l, r)
    torch.cuda.synchronize()
    eager_s = (time.perf_counter() - t0) / 2
    cap = RG.CapturedPipeline(lambda a, c: RG.net_band_ragged(net, b, a, c)[:2], l, r,
                              warmup=args.warmup)
    for _ in range(args.warmup):
        cap.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        out, st = cap.replay()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    same = all(torch.equal(u, v) for u, v in zip(b.unpack(out), b.unpack(eout)))
    outs = [o.cpu().numpy() for o in b.unpack(out)]
    bad = [float((np.abs(o - p["disp"]) > 2.0)[p["valid"]].mean() * 100.0) for o, p in zip(outs, pairs)]
    fr = None if args.no_fullrange else ragged_fullrange(args, sv, b, l, r, pairs, spec)
    sweep = None if args.no_fullrange else ragged_fallback_sweep(args, RG, b, l, r, cap, pairs, spec)
    px = sum(w * h for w, h, _ in spec)
    K = 2 * R_BAND + 1
    canon = sum(2 * w * h * K * (1 + 8 * 5 + 2) + 2 * w * h * 10 + w * h for w, h, _ in spec)
    sgm_bytes = sum(canonical_bytes_per_pair(2, K, w * h) for w, h, _ in spec)
    t_kern = (stages["sgm"] + stages["wta"]) / 1e3
    rec = {
        "metric": "pairs/s, 64 uneven pairs (200x150..1390x1110), net+band-SGM in one graph replay",
        "value": round(64 / dt, 2), "unit": "pairs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(dt * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32 network / int SGM / f64 epilogue",
        "data": "synthetic",
        "config": {"workload": "64 uneven pairs, W 200-1390, H = 0.75-0.80 W, max disparity "
                               "16-240 px (<= W/3), predict_auto v5 + refine_disparity(planes=False) "
                               "as one ragged batch, HIP graph replay",
                   "pairs_per_step": 64, "megapixels": round(px / 1e6, 3)},
        "eager_ms_per_step": round(eager_s * 1e3, 3),
        "autotuned": ({k: v for k, v in tuned.items() if not k.startswith("_")} if tuned else None),
        "graph_equals_eager": bool(same),
        "accuracy": {"bad2_pct_mean": round(float(np.mean(bad)), 3),
                     "bad2_pct_max": round(float(np.max(bad)), 3)},
        "stage_ms_eager": stages,
        "fullrange": fr,
        "speedup_vs_fullrange": round(fr["ms_per_step"] / (dt * 1e3), 3) if fr else None,
        "fallback_sweep": sweep,
        "roofline": {"bound": "hbm", "unit": "GB/s", "peak": HBM_PEAK_GBS,
                     "pipeline_canonical_bytes": canon,
                     "pipeline_achieved": round(canon / dt / 1e9, 1),
                     "pipeline_frac": round(canon / dt / 1e9 / HBM_PEAK_GBS, 4),
                     "sgm_wta_canonical_bytes": sgm_bytes,
                     "sgm_wta_achieved": round(sgm_bytes / t_kern / 1e9, 1) if t_kern > 0 else None,
                     "sgm_wta_frac": (round(sgm_bytes / t_kern / 1e9 / HBM_PEAK_GBS, 4)
                                      if t_kern > 0 else None)},
    }
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
