"""Network left-right consistency on the GPU: drop-in counterparts of
mirror_h, lr_consistency, fill_invalid and predict_consistent
(stereonet.py:4974-5076), plus a batched predict_consistent for many pairs
of one size.

Inputs may be Python lists (returned as lists, like the reference), numpy
arrays (returned as numpy) or torch tensors (returned as CUDA tensors).
The three helpers run hand-written HIP kernels (sv_mirror_h,
sv_lr_consistency, sv_fill_invalid in libstereo_mi.so); the network passes
are the package's own (net.py).  There is no CPU fallback.
"""

import ctypes

import numpy as np
import torch

from . import _lib as L
from . import net as N

_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_LL = ctypes.c_longlong

L.extra_signatures({
    "sv_mirror_h": (_I, [_P, _I, _LL, _I, _P, _P]),
    "sv_lr_consistency": (_I, [_P, _P, _I, _I, _I, _I, _D, _P, _P, _P]),
    "sv_fill_invalid": (_I, [_P, _P, _I, _I, _I, _I, _P, _P]),
})

_DT = {torch.uint8: L.SV_U8, torch.float32: L.SV_F32, torch.float64: L.SV_F64}


def _kind(a):
    if isinstance(a, torch.Tensor):
        return "torch"
    if isinstance(a, np.ndarray):
        return "numpy"
    return "list"


def _dev(a, float_only=False):
    """Device tensor of a (lists as float64, the reference's floats)."""
    if isinstance(a, torch.Tensor):
        t = a
    elif isinstance(a, np.ndarray):
        t = torch.from_numpy(np.ascontiguousarray(a))
    else:
        t = torch.tensor(a, dtype=torch.float64)
    allowed = (torch.float32, torch.float64) if float_only else tuple(_DT)
    if t.dtype not in allowed:
        t = t.to(torch.float64)
    return t.to(L.device()).contiguous()


def _out(t, kind, like=None):
    if kind == "torch":
        return t.reshape(like.shape) if like is not None else t
    a = t.cpu().numpy()
    if kind == "numpy":
        return a.reshape(like.shape) if like is not None else a
    return a.ravel().tolist()


def _count(t, W, H):
    n = t.numel()
    if W <= 0 or H <= 0 or n == 0 or n % (W * H):
        raise ValueError("array of %d values does not hold W x H = %d x %d maps" % (n, W, H))
    return n // (W * H)


def _mirror_dev(t, rows, W):
    out = torch.empty_like(t)
    L.check(L.lib().sv_mirror_h(L.ptr(t), _DT[t.dtype], rows, W, L.ptr(out), L.stream()))
    return out


def mirror_h(img, C, W, H):
    """mirror_h(img, C, W, H), stereonet.py:4974-4983: every row of every
    channel reversed (a batch of such images flips image by image)."""
    kind = _kind(img)
    t = _dev(img)
    n = t.numel()
    if n == 0 or n % (C * W * H):
        raise ValueError("array of %d values does not hold %d x %d x %d images" % (n, C, W, H))
    out = _mirror_dev(t, n // W, W)
    if kind == "list" and t.dtype == torch.uint8:
        out = out.to(torch.float64)
    return _out(out, kind, img if kind != "list" else None)


def _consistency_dev(dl, dr, W, H, tol, counts=False):
    B = _count(dl, W, H)
    if dr.numel() != dl.numel():
        raise ValueError("disparity maps of %d and %d values" % (dl.numel(), dr.numel()))
    if dr.dtype != dl.dtype:
        dr = dr.to(dl.dtype)
    valid = torch.empty(dl.numel(), dtype=torch.uint8, device=dl.device)
    cnt = torch.empty(B, dtype=torch.int32, device=dl.device) if counts else None
    L.check(L.lib().sv_lr_consistency(L.ptr(dl), L.ptr(dr), _DT[dl.dtype], B, W, H, float(tol),
                                      L.ptr(valid), L.ptr(cnt), L.stream()))
    return valid, cnt


def lr_consistency(disp_l, disp_r, W, H, tol=1.5):
    """lr_consistency(disp_l, disp_r, W, H, tol=1.5), stereonet.py:4986-5008:
    True where the left map's disparity lands inside the right view and
    agrees within tol with the right map (in the right view's own frame)
    interpolated there.  Returns a list of bools for lists, a bool array or
    tensor otherwise."""
    kind = _kind(disp_l)
    dl, dr = _dev(disp_l, float_only=True), _dev(disp_r, float_only=True)
    valid, _ = _consistency_dev(dl, dr, W, H, tol)
    v = valid.to(torch.bool)
    if kind == "torch":
        return v.reshape(disp_l.shape)
    a = v.cpu().numpy()
    return a.reshape(disp_l.shape) if kind == "numpy" else a.ravel().tolist()


def _fill_dev(d, valid, W, H):
    B = _count(d, W, H)
    out = torch.empty_like(d)
    L.check(L.lib().sv_fill_invalid(L.ptr(d), L.ptr(valid), _DT[d.dtype], B, W, H, L.ptr(out),
                                    L.stream()))
    return out


def fill_invalid(disp, valid, W, H):
    """fill_invalid(disp, valid, W, H), stereonet.py:5011-5046: each run of
    rejected pixels takes the smaller disparity of its nearest agreeing
    neighbours on the row."""
    kind = _kind(disp)
    d = _dev(disp, float_only=True)
    if isinstance(valid, torch.Tensor):
        v = valid.to(L.device())
    else:
        v = torch.from_numpy(np.ascontiguousarray(np.asarray(valid, dtype=bool))).to(L.device())
    v = v.reshape(-1).to(torch.uint8).contiguous()
    if v.numel() != d.numel():
        raise ValueError("%d validity flags for %d disparities" % (v.numel(), d.numel()))
    return _out(_fill_dev(d, v, W, H), kind, disp if kind != "list" else None)


def predict_consistent(net, left, right, W, H, work_width=None, tol=1.5, verbose=True):
    """predict_consistent(net, left, right, W, H, work_width=None, tol=1.5,
    verbose=True), stereonet.py:5049-5076: the left map from predict_auto
    (or predict_at_scale at work_width), the right-reference map from the
    mirrored, swapped pair at the same working width, mirrored back; pixels
    that disagree by more than tol are refilled from their row neighbours.
    Returns (disp, valid_fraction)."""
    kind = _kind(left)
    l = N._img(left).reshape(-1)
    r = N._img(right).reshape(-1)
    if work_width is None:
        d_l, _, Ww, _ = N.predict_auto(net, l, r, W, H, verbose=verbose)
    else:
        d_l, _, Ww, _ = N.predict_at_scale(net, l, r, W, H, work_width, verbose)
    ml, mr = _mirror_dev(l, H, W), _mirror_dev(r, H, W)
    d_rm, _, _, _ = N.predict_at_scale(net, mr, ml, W, H, Ww, verbose=False)
    d_r = _mirror_dev(d_rm.reshape(-1).contiguous(), H, W)
    d_l = d_l.reshape(-1).contiguous()
    valid, cnt = _consistency_dev(d_l, d_r, W, H, tol, counts=True)
    frac = int(cnt.item()) / float(W * H)
    if verbose:
        print("  left-right consistency: %.1f%% of pixels agree within %.1f px"
              % (100.0 * frac, tol))
    out = _fill_dev(d_l, valid, W, H)
    if kind == "torch":
        return out, frac
    return N._ret(out, kind), frac


def predict_consistent_batch(net, left, right, W, H, tol=1.5):
    """predict_consistent for N pairs of one size (uint8 or float32 [N][H][W]
    on the GPU): predict_auto_batch, then the mirrored pass of every pair at
    its own final working width in one batched pass, the consistency check
    and the fill.  Returns (disp [N][H][W] float32, valid [N][H][W] uint8,
    agreeing pixels per pair [N] int32).  Reading the working widths makes
    one host round trip."""
    Nn = left.shape[0]
    res = N.predict_auto_batch(net, left, right, W, H)
    d_l = res.full.reshape(Nn, H * W)
    sizes = [N._work_size(W, H, w) for w in res.works]
    ml = _mirror_dev(left.contiguous(), Nn * H, W).reshape(Nn, H, W)
    mr = _mirror_dev(right.contiguous(), Nn * H, W).reshape(Nn, H, W)
    # predict_at_scale at each pair's working width (the full size when it
    # reaches W), back to W x H with the values times 1/scale
    plw = net.forward_resized(mr, ml, W, H, [(h, w) for w, h, _ in sizes])
    inv = torch.tensor([1.0 / s for _, _, s in sizes], dtype=torch.float32, device=left.device)
    d_rm = torch.empty(Nn, H * W, dtype=torch.float32, device=left.device)
    lv = N._Level([(H, W)] * Nn, left.device)
    N._call("sv_bilinear_fwd", N._p(plw.d2), L.SV_F32, N._p(plw.full.table), plw.full.slot, 1,
            N._p(d_rm), N._p(lv.table), H * W, 0, 1.0, N._p(inv), Nn, H, W, L.stream())
    d_r = _mirror_dev(d_rm, Nn * H, W).reshape(Nn, H * W)
    valid, cnt = _consistency_dev(d_l.contiguous(), d_r, W, H, tol, counts=True)
    out = _fill_dev(d_l.contiguous(), valid, W, H)
    return out.reshape(Nn, H, W), valid.reshape(Nn, H, W), cnt
