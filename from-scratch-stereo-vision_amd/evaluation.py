"""Middlebury evaluation harness on the GPU: drop-in counterparts of
eval_real.py's LAYOUT, scenes, load_gt, masks, score, constant, evaluate and
table (eval_real.py:48-196), with the same names, arguments and return
values.

Ground truth is decoded by the package's image reader (read_image_gpu) and
stays on the GPU: the ground-truth disparity and known mask
(sv_gt_from_gray), the non-occluded mask (sv_gt_nonocc) and the error sums
(sv_score, sv_abs_dev) are HIP kernels in libstereo_mi.so.  Only the final
statistics come back to the host.  The per-pixel rules are the reference's
float64 arithmetic, bit for bit; EPE and the constant baseline are float64
sums taken in a fixed parallel order, so they agree with the reference's
left-to-right sums to rounding (the bad-pixel counts exactly).

Lists in give lists out (load_gt, masks), as the reference; the *_gpu forms
return CUDA tensors.
"""

import ctypes
import os
import time

import numpy as np
import torch

from . import _lib as L
from . import images as I

_P, _I, _D, _LL, _SZ = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_longlong, \
    ctypes.c_size_t
L.extra_signatures({
    "sv_gt_from_gray": (_I, [_P, _LL, _D, _P, _P, _P]),
    "sv_gt_nonocc": (_I, [_P, _P, _P, _P, _I, _I, _I, _P, _P]),
    "sv_score_workspace_bytes": (_I, [_I, _LL, ctypes.POINTER(_SZ)]),
    "sv_score": (_I, [_P, _I, _P, _P, _I, _LL, _P, _P, _SZ, _P]),
    "sv_abs_dev": (_I, [_P, _P, _P, _I, _LL, _P, _P, _SZ, _P]),
})

# eval_real.py:48-52: left, right, left GT, right GT, GT scale per year
LAYOUT = {
    "2006": ("view1.png", "view5.png", "disp1.png", "disp5.png", 3.0),
    "2005": ("view1.png", "view5.png", "disp1.png", "disp5.png", 3.0),
    "2003": ("im2.png", "im6.png", "disp2.png", "disp6.png", 4.0),
}


def scenes(root):
    """scenes, eval_real.py:55-75: every scene with ground truth under root,
    tagged with its split (2006 sorted, index % 4 == 1 -> val, else train;
    2005 and 2003 -> test)."""
    out = []
    for year in ("2006", "2005", "2003"):
        ydir = os.path.join(root, year)
        if not os.path.isdir(ydir):
            continue
        L_, R_, DL, DR, scale = LAYOUT[year]
        names = sorted(n for n in os.listdir(ydir) if os.path.isdir(os.path.join(ydir, n)))
        for i, name in enumerate(names):
            d = os.path.join(ydir, name)
            if not os.path.exists(os.path.join(d, DL)):
                continue
            split = ("val" if i % 4 == 1 else "train") if year == "2006" else "test"
            out.append({"year": year, "name": name, "dir": d, "split": split,
                        "left": os.path.join(d, L_), "right": os.path.join(d, R_),
                        "gtl": os.path.join(d, DL), "gtr": os.path.join(d, DR), "scale": scale})
    return out


def gt_from_gray(gray, scale):
    """Device ground truth from a decoded disparity image's gray plane
    (float64 CUDA tensor, any shape): (disp float64, known uint8)."""
    g = gray.contiguous()
    if g.dtype != torch.float64 or not g.is_cuda:
        raise ValueError("gray must be a float64 CUDA tensor")
    disp = torch.empty_like(g)
    known = torch.empty(g.shape, dtype=torch.uint8, device=g.device)
    L.check(L.lib().sv_gt_from_gray(L.ptr(g), g.numel(), float(scale), L.ptr(disp), L.ptr(known),
                                    L.stream()))
    return disp, known


def load_gt_gpu(path, scale):
    """load_gt on the GPU: (disp [H][W] float64, known [H][W] uint8, W, H)."""
    g, _, W, H = I.read_image_gpu(path)
    d, k = gt_from_gray(g, scale)
    return d, k, W, H


def load_gt(path, scale):
    """load_gt, eval_real.py:78-87: gray level = gray * 255; pixels at level
    >= 0.5 are known with disparity level / scale, the rest 0 and unknown."""
    d, k, W, H = load_gt_gpu(path, scale)
    return (d.cpu().numpy().ravel().tolist(), k.cpu().numpy().astype(bool).ravel().tolist(), W, H)


def nonocc_mask(dl, kl, dr, kr, W, H):
    """Non-occluded mask from both views' device ground truth ([..., H, W]
    float64 disparities, uint8 known flags; any number of maps)."""
    n = dl.numel()
    if n == 0 or n % (W * H) or dr.numel() != n or kl.numel() != n or kr.numel() != n:
        raise ValueError("ground-truth maps do not hold W x H = %d x %d values" % (W, H))
    out = torch.empty(dl.shape, dtype=torch.uint8, device=dl.device)
    L.check(L.lib().sv_gt_nonocc(L.ptr(dl.contiguous()), L.ptr(kl.contiguous()),
                                 L.ptr(dr.contiguous()), L.ptr(kr.contiguous()), n // (W * H), W,
                                 H, L.ptr(out), L.stream()))
    return out


def masks_gpu(sc):
    """masks on the GPU: (gt, known, nonocc as [H][W] CUDA tensors, W, H)."""
    dl, kl, W, H = load_gt_gpu(sc["gtl"], sc["scale"])
    dr, kr, W2, H2 = load_gt_gpu(sc["gtr"], sc["scale"])
    if (W2, H2) != (W, H):
        raise ValueError("left and right ground truth differ in size")
    return dl, kl, nonocc_mask(dl, kl, dr, kr, W, H), W, H


def masks(sc):
    """masks, eval_real.py:90-102: (gt, known, nonocc, W, H) as lists; a
    known pixel is non-occluded when int(x - dL + 0.5) lands on a known right
    pixel within 1 px of dL."""
    dl, kl, no, W, H = masks_gpu(sc)
    tolist = lambda t: t.cpu().numpy().astype(bool).ravel().tolist()
    return dl.cpu().numpy().ravel().tolist(), tolist(kl), tolist(no), W, H


def _dev(a, dtype=None):
    if isinstance(a, torch.Tensor):
        t = a
    else:
        t = torch.from_numpy(np.ascontiguousarray(np.asarray(a, dtype=dtype)))
    return t.to(L.device()).contiguous()


def score_stats(pred, gt, mask, B=1):
    """Device sums for B maps: [B][6] float64 = (n, sum |pred - gt|, count
    > 1, count > 2, count > 4, sum gt), each over the mask's pixels."""
    p = _dev(pred, np.float64)
    if p.dtype not in (torch.float32, torch.float64):
        p = p.to(torch.float64)
    g = _dev(gt, np.float64).to(torch.float64)
    m = _dev(mask, np.bool_).reshape(-1).to(torch.uint8)
    n = p.numel()
    if n == 0 or n % B or g.numel() != n or m.numel() != n:
        raise ValueError("pred, gt and mask hold %d, %d and %d values for %d maps"
                         % (n, g.numel(), m.numel(), B))
    npix = n // B
    nb = ctypes.c_size_t(0)
    L.check(L.lib().sv_score_workspace_bytes(B, npix, ctypes.byref(nb)))
    ws = L.workspace.get(nb.value)
    out = torch.empty((B, 6), dtype=torch.float64, device=p.device)
    L.check(L.lib().sv_score(L.ptr(p), L.SV_F32 if p.dtype == torch.float32 else L.SV_F64,
                             L.ptr(g), L.ptr(m), B, npix, L.ptr(out), L.ptr(ws), nb.value,
                             L.stream()))
    return out, g, m


def _score_of(s):
    n = s[0]
    if n == 0:
        return None
    return (s[1] / n, 100.0 * s[2] / n, 100.0 * s[3] / n, 100.0 * s[4] / n)


def score(pred, gt, mask):
    """score, eval_real.py:105-121: (EPE, bad1 %, bad2 %, bad4 %) over the
    mask, or None when the mask is empty."""
    s, _, _ = score_stats(pred, gt, mask)
    return _score_of(s[0].cpu().tolist())


def abs_dev_sums(gt, mask, stats, B=1):
    """Device sums of |gt - mean| over the mask, the mean per map taken from
    score_stats' sums ([B] float64)."""
    npix = gt.numel() // B
    nb = ctypes.c_size_t(0)
    L.check(L.lib().sv_score_workspace_bytes(B, npix, ctypes.byref(nb)))
    ws = L.workspace.get(nb.value)
    out = torch.empty(B, dtype=torch.float64, device=gt.device)
    L.check(L.lib().sv_abs_dev(L.ptr(gt), L.ptr(mask), L.ptr(stats), B, npix, L.ptr(out),
                               L.ptr(ws), nb.value, L.stream()))
    return out


def constant(gt, mask):
    """constant, eval_real.py:124-127: mean absolute deviation of the masked
    ground truth from its mean (the error of the best constant prediction
    in the mean sense).  An empty mask divides by zero, as the reference."""
    g = _dev(gt, np.float64).to(torch.float64)
    s, g, m = score_stats(g, g, mask)
    n = float(s[0, 0])
    if n == 0:
        raise ZeroDivisionError("float division by zero")
    return float(abs_dev_sums(g, m, s)[0]) / n


def evaluate(net, sc, consistency=False, save_dir=None, mode="net", dump=None):
    """evaluate, eval_real.py:130-172: one scene through 'net' (predict_auto
    or predict_consistent), 'refine' (predict_auto + refine_disparity with
    the left RGB) or 'sgm' (full-range refine_fullres, R from the ground
    truth), scored on the GPU.  save_dir writes the prediction as a PGM
    scaled 0..max(gt) (write_pgm); dump writes the same pickled dict as the
    reference."""
    from . import net as N
    from . import refine as Rf
    from . import consistency as C
    lg, lrgb, W, H = I.read_image_gpu(sc["left"])
    rg, _, _, _ = I.read_image_gpu(sc["right"])
    gt, known, nonocc, Wg, Hg = masks_gpu(sc)
    torch.cuda.synchronize()
    t0 = time.time()
    if mode == "sgm":
        sel = gt[nonocc.bool()]
        R = int(float(sel.max()) + 2) // 2 + 1
        d0 = torch.full((H, W), float(R), dtype=torch.float64, device=lg.device)
        pred = Rf.refine_fullres(lg, rg, d0, W, H, R=R, P1=12.0, P2=150.0, rx=4, ry=3)
    elif consistency:
        pred, _ = C.predict_consistent(net, lg, rg, W, H, verbose=False)
    else:
        pred, _, _, _ = N.predict_auto(net, lg, rg, W, H, verbose=False)
        if mode == "refine":
            pred, _ = Rf.refine_disparity(lg, rg, lrgb, pred, W, H)
    torch.cuda.synchronize()
    dt = time.time() - t0
    pred = pred.reshape(-1)
    s_no, g, m_no = score_stats(pred, gt, nonocc)
    s_all, _, _ = score_stats(pred, gt, known)
    n_no = float(s_no[0, 0])
    const = float(abs_dev_sums(g, m_no, s_no)[0]) / n_no if n_no else float("nan")
    if dump:
        import pickle
        os.makedirs(dump, exist_ok=True)
        with open(os.path.join(dump, "%s_%s.pkl" % (sc["year"], sc["name"])), "wb") as f:
            pickle.dump({"scene": sc["name"], "year": sc["year"], "split": sc["split"], "W": W,
                         "H": H, "pred": pred.double().cpu().numpy().tolist(),
                         "gt": gt.cpu().numpy().ravel().tolist(),
                         "nonocc": nonocc.cpu().numpy().astype(bool).ravel().tolist(),
                         "known": known.cpu().numpy().astype(bool).ravel().tolist()}, f)
    if save_dir:
        os.makedirs(save_dir, exist_ok=True)
        I.write_pgm(os.path.join(save_dir, "%s_%s.pgm" % (sc["year"], sc["name"])),
                    pred.double().cpu().numpy(), W, H, 0.0, float(gt.max()))
    return {"nonocc": _score_of(s_no[0].cpu().tolist()), "all": _score_of(s_all[0].cpu().tolist()),
            "const": const, "seconds": dt, "size": "%dx%d" % (W, H)}


def table(rows, title):
    """table, eval_real.py:165-195: the Markdown results table with per-split
    means."""
    out = ["", "### " + title, "",
           "| scene | split | EPE | bad1 | bad2 | bad4 | all EPE | constant | s |",
           "|---|---|---|---|---|---|---|---|---|"]
    for sc, r in rows:
        n, a = r["nonocc"], r["all"]
        out.append("| %s %s | %s | %.2f | %.1f%% | %.1f%% | %.1f%% | %.2f | %.2f | %.0f |"
                   % (sc["year"], sc["name"], sc["split"], n[0], n[1], n[2], n[3], a[0],
                      r["const"], r["seconds"]))
    for split in ("train", "val", "test"):
        sub = [r for sc, r in rows if sc["split"] == split]
        if not sub:
            continue
        k = len(sub)
        out.append("| **mean %s (%d)** | | **%.2f** | **%.1f%%** | **%.1f%%** | **%.1f%%** "
                   "| **%.2f** | **%.2f** | |"
                   % (split, k, sum(r["nonocc"][0] for r in sub) / k,
                      sum(r["nonocc"][1] for r in sub) / k, sum(r["nonocc"][2] for r in sub) / k,
                      sum(r["nonocc"][3] for r in sub) / k, sum(r["all"][0] for r in sub) / k,
                      sum(r["const"] for r in sub) / k))
    return "\n".join(out)
