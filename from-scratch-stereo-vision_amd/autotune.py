"""Autotuner for the library's kernel variants.

Some kernels come in variants that compute identical results at different
speeds: the prefetch depth of the vertical/diagonal SGM scan, the columns
per LDS tile of the horizontal SGM scan, whether the horizontal scan runs
beside the others on a second stream, and for the tiled convolution the
engine (vector ALUs or matrix cores), the output rows per thread, columns
per tile, weight source and pixel groups per wave.  The library reads the
choice from process-wide tuning knobs (sv_tune_set); this module picks them
for a workload by timing it on the GPU.

tune() does coordinate descent: for each knob in turn it times the workload
under every allowed value, the other knobs held at their best so far, and
keeps the fastest.  Every variant's output is compared with the first run's
and must be identical (a knob that changed results would be a bug, and tune
raises).  Choices can be cached in a JSON file keyed by device name and a
workload key, so a later process applies them without re-timing.
"""

import ctypes
import json
import os

import torch

from . import _lib as L

KNOBS = {"sgm_prefetch": (0, (2, 4, 8)),
         "sgm_htile": (1, (8, 16)),
         "conv_rows": (2, (1, 4)),
         "conv_tile_width": (3, (64, 32, 16)),
         "conv_weights": (4, (1, 2)),
         "sgm_overlap": (5, (2, 1)),
         "conv_engine": (6, (1, 2, 3, 4, 5, 6, 7)),
         "conv_mgroups": (7, (1, 2, 4)),
         "sgm_split": (8, (1, 2)),
         "wta_reads": (9, (1, 2)),
         "conv_dchunk": (10, (12, 6, 4)),
         "sgm_order": (11, (1, 2)),
         "sgm_queue": (12, (1, 2, 3))}


def set_knob(name, value):
    L.check(L.lib().sv_tune_set(KNOBS[name][0], int(value)))


def get_knob(name):
    v = ctypes.c_int(0)
    L.check(L.lib().sv_tune_get(KNOBS[name][0], ctypes.byref(v)))
    return v.value


def apply(settings):
    for k, v in settings.items():
        set_knob(k, v)


def reset():
    for k in KNOBS:
        set_knob(k, 0)


def _flat(out):
    if isinstance(out, torch.Tensor):
        return [out]
    res = []
    for o in out:
        res.extend(_flat(o))
    return res


def _time(run, reps):
    run()  # warm-up (first use of a variant)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    best, out = None, None
    for _ in range(reps):
        e0.record()
        out = run()
        e1.record()
        e1.synchronize()
        ms = e0.elapsed_time(e1)
        best = ms if best is None else min(best, ms)
    return best, out


def _cache_key(key):
    return "%s|%s" % (torch.cuda.get_device_name(L.device()), key)


def tune(run, key=None, knobs=None, reps=3, cache_path=None, log=None):
    """Pick the knob values that make `run` fastest.  run(): queues one
    workload on the current stream and returns its result tensor(s).
    Returns {knob: value} (also applied), plus "_ms" (best time) and
    "_trials" (every timing, for reporting)."""
    knobs = list(knobs or KNOBS)
    if cache_path and key and os.path.exists(cache_path):
        with open(cache_path) as f:
            cached = json.load(f).get(_cache_key(key))
        if cached:
            apply({k: v for k, v in cached.items() if k in KNOBS})
            return cached
    best = {k: KNOBS[k][1][0] for k in knobs}
    apply(best)
    ref_ms, ref = _time(run, reps)
    ref = [t.clone() for t in _flat(ref)]
    trials = {}
    best_ms = ref_ms
    for k in knobs:
        for v in KNOBS[k][1]:
            trial = dict(best)
            trial[k] = v
            apply(trial)
            ms, out = _time(run, reps)
            outs = _flat(out)
            if len(outs) != len(ref) or not all(torch.equal(a, b) for a, b in zip(outs, ref)):
                reset()
                raise RuntimeError("tuning knob %s=%d changed the results" % (k, v))
            trials[",".join("%s=%d" % kv for kv in sorted(trial.items()))] = ms
            if log:
                log("autotune %s: %.3f ms" % (trial, ms))
            if ms < best_ms:
                best_ms, best = ms, trial
    apply(best)
    res = dict(best, _ms=best_ms, _trials=trials)
    if cache_path and key:
        data = {}
        if os.path.exists(cache_path):
            with open(cache_path) as f:
                data = json.load(f)
        data[_cache_key(key)] = res
        os.makedirs(os.path.dirname(os.path.abspath(cache_path)), exist_ok=True)
        with open(cache_path, "w") as f:
            json.dump(data, f, indent=1)
    return res



def _conv_variants(kd=1, oc=16):
    """Every convolution variant as ((knob, value), ...): the vector kernel
    under each rows x tile width x weight source, its per-channel form under
    each rows x tile width, the matrix-core kernel
    (3D layers sliding over depth, or one depth per tile) under each pixel
    groups per wave x tile width, the 4-channel-block matrix-core
    kernel (2D layers) under 1 or 2 pixel groups x tile width, for 3D
    layers (kd = 3) the depth-column matrix-core kernel under each tile
    width x depth chunk, and for layers of at most 16 output channels the
    direct kernel."""
    K = {k: v[0] for k, v in KNOBS.items()}
    out = []
    for r in KNOBS["conv_rows"][1]:
        for t in KNOBS["conv_tile_width"][1]:
            for w in KNOBS["conv_weights"][1]:
                out.append(((K["conv_engine"], 1), (K["conv_rows"], r),
                            (K["conv_tile_width"], t), (K["conv_weights"], w)))
    for r in KNOBS["conv_rows"][1]:
        for t in KNOBS["conv_tile_width"][1]:
            out.append(((K["conv_engine"], 4), (K["conv_rows"], r), (K["conv_tile_width"], t)))
    for e in (2, 3):
        for g in KNOBS["conv_mgroups"][1]:
            for t in KNOBS["conv_tile_width"][1]:
                out.append(((K["conv_engine"], e), (K["conv_mgroups"], g), (K["conv_tile_width"], t)))
    for g in (1, 2):
        for t in KNOBS["conv_tile_width"][1]:
            out.append(((K["conv_engine"], 5), (K["conv_mgroups"], g), (K["conv_tile_width"], t)))
    if kd == 3:
        for t in KNOBS["conv_tile_width"][1]:
            for dc in KNOBS["conv_dchunk"][1]:
                out.append(((K["conv_engine"], 6), (K["conv_tile_width"], t), (K["conv_dchunk"], dc)))
    if oc <= 16:
        out.append(((K["conv_engine"], 7),))
    return out


def tune_layers(net, run, reps=3, log=None):
    """Per-layer variants of the tiled convolution.  Records every
    convolution launch of one run() (a workload on `net`), then for each
    distinct launch shape times every variant (the vector kernel's rows per
    thread, tile width and weight source; the matrix-core kernel's pixel
    groups per wave and tile width) on that launch alone, checks that the
    layer's output is identical under each, and keeps the fastest for that
    shape (net._conv_tune, applied at launch).  Returns {shape: (variant,
    ms)}."""
    from . import net as N

    calls = {}

    def hook(key, fn, args, out, tables):
        # the dims tables as they are at this launch (a later launch of the
        # same plan may refill them on the device)
        if fn == "sv_conv_fwd" and key not in calls:
            ti, to = tables[0].clone(), tables[1].clone()
            args = list(args)
            args[2], args[14] = N._p(ti), N._p(to)
            calls[key] = (tuple(args), out, (ti, to))

    net._conv_hook = hook
    try:
        run()
        torch.cuda.synchronize()
    finally:
        net._conv_hook = None
    chosen = {}
    try:
        for key, (args, out, _) in calls.items():
            best, ref = None, None
            for variant in _conv_variants(args[6], args[9]):
                reset()
                for knob, v in variant:
                    L.check(L.lib().sv_tune_set(knob, v))
                # the slots' padding past each sample is never written: clear
                # it so that stale contents (NaN patterns included) compare
                out.zero_()
                ms, _ = _time(lambda: N._call("sv_conv_fwd", *args), reps)
                if ref is None:
                    ref = out.clone()
                elif not torch.equal(out, ref):
                    raise RuntimeError("conv variant %s changed the output of layer %s"
                                       % (variant, key))
                if log:
                    log("autotune layer %s %s: %.4f ms" % (key, variant, ms))
                if best is None or ms < best[1]:
                    best = (variant, ms)
            chosen[key] = best
            net._conv_tune[key] = best[0]
    finally:
        reset()
    return chosen


def save_choices(net, knobs, path):
    """Write the process-wide knob choices and net's per-layer convolution
    variants (tune_layers) to a JSON file."""
    data = {"knobs": {k: int(v) for k, v in knobs.items() if k in KNOBS},
            "layers": [[list(key), [list(kv) for kv in variant]]
                       for key, variant in net._conv_tune.items()]}
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "w") as f:
        json.dump(data, f, indent=1)


def load_choices(net, path):
    """Apply choices written by save_choices; returns the knob settings
    plus the number of layers with a variant."""
    with open(path) as f:
        data = json.load(f)
    apply(data["knobs"])
    net._conv_tune = {tuple(key): tuple(tuple(kv) for kv in variant)
                      for key, variant in data["layers"]}
    return dict(data["knobs"], conv_layers_tuned=len(net._conv_tune))
