"""GPU training of the StereoNet proposal on synthetic scenes with
full-range-SGM pseudo-labels, and net + band-SGM inference with a
full-range-SGM fallback.

Training follows train_parallel.py (batch of samples -> mean gradient of
train_grads -> one Adam step; learning rate x0.35 at 60% and 85% of the run,
:139-142) with the batch computed in one GPU pass (train_grads_batch) instead
of a process pool.  Labels are never the ground truth: each training pair is
labelled by full-range SGM (refine_fullres with a constant centre R and
K = 2R + 1, eval_real.py:143-146) run on both views, and only pixels that
pass the left-right check (lr_refine status 0, stereonet.py:5343-5391) are
supervised.

Inference with fallback (not in the reference, which fills inconsistent
pixels from disp0 and their neighbours, :5372-5390): after net + band
refinement, a pair whose share of left-right-inconsistent pixels exceeds
`tau` is re-solved with full-range SGM, and its inconsistent pixels take the
full-range disparity.  The fallback fraction (pairs and pixels) is reported
with the timing of both stages.

    python -m from-scratch-stereo-vision_amd.train --steps 3000 --out weights.txt
"""

import argparse
import json
import os
import sys
import time
from multiprocessing import Pool

import numpy as np
import torch

from . import _lib as L
from . import net as N
from . import refine as RF
from . import scenes, synth

P1, P2, RX, RY = 12.0, 150.0, 4, 3


def full_range_R(max_disp):
    """Full-range half-width for disparities up to max_disp
    (eval_real.py:143-146): int(max_disp + 2) // 2 + 1."""
    return int(float(max_disp) + 2) // 2 + 1


# ------------------------------------------------------------------ data

def _rendered(args):
    seed, W, H, lo, hi = args
    return scenes.render_pair(seed, W, H, lo, hi)


def _planar(args):
    seed, W, H, lo, hi = args
    rng = np.random.default_rng(seed)
    peak = rng.uniform(lo, hi)
    p = synth.make_pair(W, H, seed, dmin=0.2 * peak, dmax=peak, n_boxes=int(rng.integers(0, 4)))
    p["peak"] = float(peak)
    return p


def make_dataset(seeds, W=260, H=192, peak=(22.0, 40.0), kind="mixed", workers=None):
    """Synthetic pairs for the given seeds, stacked: left/right uint8
    [N][H][W], disp float64 (0 where unknown), valid bool, peak per pair.
    kind "rendered" (ray-cast scenes), "planar" (warped textured planes
    and boxes) or "mixed" (even seeds rendered, odd seeds planar).
    workers > 1 renders in a process pool."""
    jobs = []
    for s in seeds:
        k = kind if kind != "mixed" else ("rendered" if s % 2 == 0 else "planar")
        jobs.append((k, (s, W, H, peak[0], peak[1])))
    workers = workers or min(16, os.cpu_count() or 1)
    ra = [a for k, a in jobs if k == "rendered"]
    pa = [a for k, a in jobs if k == "planar"]
    if workers > 1:
        # fork the pool only from a process that has not touched the GPU
        with Pool(workers) as pool:
            rend, plan = pool.map(_rendered, ra), pool.map(_planar, pa)
    else:
        rend, plan = [_rendered(a) for a in ra], [_planar(a) for a in pa]
    it_r, it_p = iter(rend), iter(plan)
    out = [next(it_r) if k == "rendered" else next(it_p) for k, _ in jobs]
    d = np.stack([o["disp"] for o in out])
    v = np.stack([o["valid"] for o in out]) & np.isfinite(d)
    return {"left": np.stack([o["left"] for o in out]), "right": np.stack([o["right"] for o in out]),
            "disp": np.where(v, d, 0.0), "valid": v, "peak": np.array([o["peak"] for o in out]),
            "W": W, "H": H}


def pseudo_labels(left, right, W, H, max_disp, tol=1.0, chunk=64):
    """Full-range SGM labels on both views (lr_refine with the centre R and
    K = 2R + 1 of eval_real.py:143-146, P1=12, P2=150, census 9x7, through
    the constant-centre kernels of refine.full_range): returns (disp float32
    [N][H][W], valid uint8) on the GPU, valid where the left-right check
    passes."""
    dev = L.device()
    R = full_range_R(max_disp)
    n = left.shape[0]
    disp = torch.empty(n, H, W, dtype=torch.float32, device=dev)
    valid = torch.empty(n, H, W, dtype=torch.uint8, device=dev)
    for a in range(0, n, chunk):
        b = min(n, a + chunk)
        l = torch.as_tensor(left[a:b]).to(dev)
        r = torch.as_tensor(right[a:b]).to(dev)
        d, st = RF.full_range(l, r, W, H, R, P1=P1, P2=P2, rx=RX, ry=RY, lr=True, tol=tol)
        disp[a:b] = d.reshape(b - a, H, W).float()
        valid[a:b] = (st.reshape(b - a, H, W) == 0).to(torch.uint8)
    return disp, valid


# -------------------------------------------------------------- training

def train_net(net, left, right, disp, valid, steps=3000, batch=16, lr=5.0e-4, seed=0,
              jitter=True, log_every=100, log=None):
    """Adam on the batch-mean train_grads (train_parallel.py:132-158): draws
    `batch` training pairs per step with a seeded generator, applies the
    brightness/contrast jitter of train_parallel.py:65-69 (gain U[0.88,
    1.12], offset U[-0.05, 0.05], clamped to [0, 1]) and schedules the
    learning rate x0.35 at 60% and 85% of the run.  left/right: uint8
    [N][H][W] on the GPU; disp float32, valid uint8.  Returns the mean loss
    per logging window."""
    H, W = left.shape[1], left.shape[2]
    rng = np.random.default_rng(seed)
    opt = N.Adam(net, lr=lr)
    dims = [(H, W)] * batch
    history = []
    run = torch.zeros((), dtype=torch.float64, device=left.device)
    t0 = time.time()
    for it in range(1, steps + 1):
        if it == int(steps * 0.6) or it == int(steps * 0.85):
            opt.lr *= 0.35
        idx = torch.as_tensor(rng.integers(0, left.shape[0], size=batch), device=left.device)
        lb = left[idx].float() * (1.0 / 255.0)
        rb = right[idx].float() * (1.0 / 255.0)
        if jitter:
            g = torch.as_tensor(rng.uniform(0.88, 1.12, size=(2, batch, 1, 1)), dtype=torch.float32,
                                device=left.device)
            o = torch.as_tensor(rng.uniform(-0.05, 0.05, size=(2, batch, 1, 1)), dtype=torch.float32,
                                device=left.device)
            lb = (lb * g[0] + o[0]).clamp_(0.0, 1.0)
            rb = (rb * g[1] + o[1]).clamp_(0.0, 1.0)
        losses, _ = N.train_grads_batch(net, lb, rb, disp[idx], valid[idx], dims, sync=False)
        opt.step()
        run += losses.mean()
        if it % log_every == 0 or it == steps:
            k = log_every if it % log_every == 0 else it % log_every
            m = float(run) / k
            history.append((it, m))
            run.zero_()
            if log:
                log("step %5d  loss %.4f  lr %.2e  (%.1f s)" % (it, m, opt.lr, time.time() - t0))
    return history


# ------------------------------------------------------------- inference

def _bad2(pred, gt, valid):
    err = np.abs(pred - gt)[valid]
    return float((err > 2.0).mean() * 100.0), float(err.mean())


def full_range(left, right, W, H, max_disps):
    """Full-range SGM control per pair (eval_real.py:143-146: each pair's R
    from its maximum disparity), all pairs in one launch per stage
    (refine.full_range with a per-pair centre table).  Returns float64
    [N][H][W] on the GPU."""
    n = left.shape[0]
    Rs = [full_range_R(m) for m in max_disps]
    out = RF.full_range(left, right, W, H, Rs, P1=P1, P2=P2, rx=RX, ry=RY)
    return out.reshape(n, H, W)


class FallbackResult:
    def __init__(self, disp, status, fallback_pairs, band_ms, fallback_ms):
        self.disp, self.status = disp, status
        self.fallback_pairs = fallback_pairs
        self.band_ms, self.fallback_ms = band_ms, fallback_ms


def infer_with_fallback(net, left, right, W, H, max_disps, tau=0.05, statuses=(1, 2)):
    """net + band SGM (predict_auto, refine_disparity planes=False), then
    full-range SGM for the pairs whose share of pixels with a status in
    `statuses` (1 occluded, 2 left-right mismatch) exceeds tau; those pixels
    take the full-range disparity."""
    dev = left.device
    n = left.shape[0]
    e = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    e[0].record()
    res = N.predict_auto_batch(net, left, right, W, H)
    d0 = res.full.reshape(n, H, W).double()
    out, st = RF.refine_disparity(left, right, None, d0, W, H, planes=False)
    out = out.reshape(n, H, W)
    st = st.reshape(n, H, W)
    flag = torch.zeros_like(st, dtype=torch.bool)
    for v in statuses:
        flag |= st == v
    frac = flag.float().mean(dim=(1, 2))
    e[1].record()
    need = (frac > tau).nonzero().flatten()
    fb = [int(i) for i in need.cpu()]
    if fb:
        it = torch.as_tensor(fb, device=dev)
        fr = full_range(left[it], right[it], W, H, [max_disps[i] for i in fb])
        sub = out[it]
        bad = flag[it]
        sub[bad] = fr[bad]
        out[it] = sub
    e[2].record()
    torch.cuda.synchronize()
    r = FallbackResult(out, st, fb, e[0].elapsed_time(e[1]), e[1].elapsed_time(e[2]))
    r.flag = flag
    return r


def evaluate(net, data, taus=(0.0, 0.05, 0.1, 0.2, 1.01), status_sets=((2,), (1, 2)), reps=3):
    """bad2 / EPE against the generated ground truth for: network alone,
    net + band, net + band with fallback for each status set and tau, and
    full-range SGM; with the wall time of each (best of `reps`)."""
    dev = L.device()
    W, H = data["W"], data["H"]
    left = torch.as_tensor(data["left"]).to(dev)
    right = torch.as_tensor(data["right"]).to(dev)
    gt, valid = data["disp"], data["valid"]
    md = [float(np.max(g[v])) for g, v in zip(gt, valid)]
    rep = {"pairs": int(left.shape[0]), "size": [W, H]}

    def timed(fn):
        best, out = None, None
        for _ in range(reps):
            torch.cuda.synchronize()
            t = time.perf_counter()
            out = fn()
            torch.cuda.synchronize()
            dt = time.perf_counter() - t
            best = dt if best is None else min(best, dt)
        return out, best

    fr, t_fr = timed(lambda: full_range(left, right, W, H, md))
    b2, epe = _bad2(fr.cpu().numpy(), gt, valid)
    rep["fullrange"] = {"bad2_pct": b2, "epe_px": epe, "s": t_fr,
                        "pairs_per_s": left.shape[0] / t_fr}
    net_only = N.predict_auto_batch(net, left, right, W, H).full.double().cpu().numpy()
    b2, epe = _bad2(net_only, gt, valid)
    rep["network"] = {"bad2_pct": b2, "epe_px": epe}
    for ss in status_sets:
        for tau in taus:
            if tau > 1.0 and ss != status_sets[0]:
                continue  # no fallback: same as the first set
            r, t = timed(lambda: infer_with_fallback(net, left, right, W, H, md, tau, ss))
            b2, epe = _bad2(r.disp.cpu().numpy(), gt, valid)
            fl = r.flag.cpu().numpy()
            fbp = float(np.mean([fl[i].mean() if i in r.fallback_pairs else 0.0
                                 for i in range(fl.shape[0])]))
            key = "band" if tau > 1.0 else "fallback%s_tau=%g" % ("".join(map(str, ss)), tau)
            rep[key] = {"bad2_pct": b2, "epe_px": epe, "s": t, "pairs_per_s": left.shape[0] / t,
                        "speedup_vs_fullrange": t_fr / t,
                        "fallback_pairs": len(r.fallback_pairs) / float(fl.shape[0]),
                        "fallback_pixels": fbp, "flagged_pixels": float(fl.mean())}
    return rep


def pipeline_data(n, seed0=0):
    """The 460x370 pipeline pairs (disparity 10-80 px, slanted planes and
    thin boxes) in make_dataset's layout."""
    ps = [synth.pipeline_pair(seed0 + s) for s in range(n)]
    return {"left": np.stack([p["left"] for p in ps]), "right": np.stack([p["right"] for p in ps]),
            "disp": np.stack([p["disp"] for p in ps]), "valid": np.stack([p["valid"] for p in ps]),
            "W": ps[0]["W"], "H": ps[0]["H"]}


# ------------------------------------------------------------------- CLI

def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--init", default=None, help="starting checkpoint (default: the v5 weights)")
    ap.add_argument("--out", required=True, help="where to write the trained checkpoint")
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--lr", type=float, default=5.0e-4)
    ap.add_argument("--train-pairs", type=int, default=1024)
    ap.add_argument("--eval-pairs", type=int, default=64)
    ap.add_argument("--width", type=int, default=260)
    ap.add_argument("--height", type=int, default=192)
    ap.add_argument("--kind", default="mixed", choices=["mixed", "rendered", "planar"])
    ap.add_argument("--pipeline-pairs", type=int, default=32,
                    help="also evaluate on this many 460x370 pipeline pairs (0: skip)")
    ap.add_argument("--report", default=None, help="write the JSON report here as well")
    a = ap.parse_args(argv)
    log = lambda s: print(s, file=sys.stderr, flush=True)
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    init = a.init or os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights", "stereonet_weights_v5.txt")
    W, H = a.width, a.height
    t = time.time()
    # the worker pool forks before this process touches the GPU
    tr = make_dataset([10000 + i for i in range(a.train_pairs)], W, H, kind=a.kind)
    ev = make_dataset([90000 + i for i in range(a.eval_pairs)], W, H, kind=a.kind)
    log("rendered %d train + %d eval pairs at %dx%d in %.1f s"
        % (a.train_pairs, a.eval_pairs, W, H, time.time() - t))
    dev = L.device()
    left = torch.as_tensor(tr["left"]).to(dev)
    right = torch.as_tensor(tr["right"]).to(dev)
    t = time.time()
    pd, pv = pseudo_labels(tr["left"], tr["right"], W, H, max_disp=40.0 + 2.0)
    torch.cuda.synchronize()
    gt = torch.as_tensor(tr["disp"]).to(dev)
    gv = torch.as_tensor(tr["valid"]).to(dev)
    pl_ok = (pv.bool() & gv)
    pl_err = (pd.double() - gt).abs()[pl_ok]
    label = {"s": time.time() - t, "valid_frac": float(pv.float().mean()),
             "bad2_vs_gt_pct": float((pl_err > 2.0).double().mean() * 100.0),
             "epe_vs_gt_px": float(pl_err.mean())}
    log("pseudo-labels: %s" % json.dumps(label))
    net = N.load_model(init)
    pipe = pipeline_data(a.pipeline_pairs) if a.pipeline_pairs > 0 else None
    before = evaluate(net, ev)
    before_pipe = evaluate(net, pipe) if pipe else None
    log("before: %s" % json.dumps(before))
    t = time.time()
    hist = train_net(net, left, right, pd, pv, steps=a.steps, batch=a.batch, lr=a.lr, log=log)
    torch.cuda.synchronize()
    t_train = time.time() - t
    N.save_weights(net, a.out)
    after = evaluate(net, ev)
    after_pipe = evaluate(net, pipe) if pipe else None
    rep = {"config": "train %d pairs %dx%d (%s), batch %d, %d Adam steps, lr %g"
                     % (a.train_pairs, W, H, a.kind, a.batch, a.steps, a.lr),
           "train_s": t_train, "ms_per_step": 1e3 * t_train / max(1, a.steps),
           "pseudo_labels": label, "loss": hist, "eval_before": before, "eval_after": after,
           "pipeline_before": before_pipe, "pipeline_after": after_pipe}
    print(json.dumps(rep), flush=True)
    if a.report:
        with open(a.report, "w") as f:
            json.dump(rep, f, indent=1)


if __name__ == "__main__":
    main()
