"""Ragged batches: stereo pairs of different sizes through one launch per
stage.

A RaggedBatch fixes the layout every per-pixel buffer of the batch shares:
pair n (W_n x H_n) occupies pixels [n*slot, n*slot + W_n*H_n) with its rows
packed at its own width, slot being the largest pair's pixel count, and a
device table holds (H_n, W_n) per pair.  The network (predict_auto_batch)
and the refinement kernels (sv_refine_ragged) read that table, so a batch
of uneven pairs costs one launch per stage instead of one per pair, and the
whole sequence can be captured once in a HIP graph and replayed
(capture_pipeline).

Per pair the results equal those of the same pair run alone: every kernel
bounds its work by the pair's own size.
"""

import ctypes

import numpy as np
import torch

from . import _lib as L
from . import net as N

L.extra_signatures({
    "sv_refine_ragged_workspace_bytes": (ctypes.c_int, [ctypes.c_int, ctypes.c_int,
                                                        ctypes.c_longlong, ctypes.c_int,
                                                        ctypes.c_int, ctypes.c_int,
                                                        ctypes.c_double, ctypes.c_double,
                                                        ctypes.c_int, ctypes.c_int,
                                                        ctypes.POINTER(ctypes.c_size_t)]),
    "sv_refine_ragged": (ctypes.c_int, [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                        ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_longlong, ctypes.c_int, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]),
})

_DT = {torch.uint8: L.SV_U8, torch.float32: L.SV_F32, torch.float64: L.SV_F64}


class RaggedBatch:
    """Layout of a batch of pairs of sizes [(W, H)] (the reference's
    argument order)."""

    def __init__(self, sizes):
        self.sizes = [(int(w), int(h)) for w, h in sizes]
        if not self.sizes or any(w < 1 or h < 1 for w, h in self.sizes):
            raise ValueError("a ragged batch needs pairs of positive size: %s" % (sizes,))
        self.B = len(self.sizes)
        self.maxW = max(w for w, _ in self.sizes)
        self.maxH = max(h for _, h in self.sizes)
        self.slot = max(w * h for w, h in self.sizes)
        self.hw = [(h, w) for w, h in self.sizes]
        self.level = N._Level(self.hw, L.device())
        self.table = self.level.table

    def pack(self, images, dtype=None):
        """Per-pair images ((H, W) arrays or tensors, or flat W*H) into one
        [B][slot] device tensor."""
        if len(images) != self.B:
            raise ValueError("%d images for a batch of %d pairs" % (len(images), self.B))
        dev = L.device()
        first = images[0]
        if dtype is None:
            dtype = first.dtype if isinstance(first, torch.Tensor) else (
                torch.uint8 if np.asarray(first).dtype == np.uint8 else torch.float64)
        out = torch.zeros(self.B, self.slot, dtype=dtype, device=dev)
        for n, (img, (w, h)) in enumerate(zip(images, self.sizes)):
            t = img if isinstance(img, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(img))
            if t.numel() != w * h:
                raise ValueError("pair %d: %d values for %dx%d" % (n, t.numel(), w, h))
            out[n, :w * h] = t.reshape(-1).to(device=dev, dtype=dtype)
        return out

    def unpack(self, buf):
        """[B][slot] buffer -> list of (H, W) views."""
        b = buf.reshape(self.B, -1)
        return [b[n, :w * h].reshape(h, w) for n, (w, h) in enumerate(self.sizes)]


def refine_ragged(batch, left, right, disp0, mode=2, R=10, P1=12.0, P2=150.0, rx=4, ry=3,
                  tol=2.0, out=None, status=None):
    """refine_fullres (mode 0), lr_refine (1) or refine_disparity with
    planes=False (2) over a ragged batch, integer pipeline (stereonet.py:
    5244-5305, :5343-5391, :5865-5884).  left/right: packed [B][slot]
    uint8/float32/float64 on the GPU; disp0: packed float64.  Returns (out,
    status) packed like the inputs (status None in mode 0)."""
    if left.dtype not in _DT or right.dtype != left.dtype:
        raise ValueError("left/right must share one of uint8, float32, float64")
    for t in (left, right, disp0):
        if t.numel() != batch.B * batch.slot or not t.is_cuda:
            raise ValueError("inputs must be packed [B][slot] device tensors of the batch")
    d0 = disp0 if disp0.dtype == torch.float64 else disp0.double()
    dev = left.device
    if out is None:
        out = torch.empty(batch.B, batch.slot, dtype=torch.float64, device=dev)
    if status is None and mode != 0:
        status = torch.empty(batch.B, batch.slot, dtype=torch.uint8, device=dev)
    nb = ctypes.c_size_t(0)
    L.check(L.lib().sv_refine_ragged_workspace_bytes(mode, batch.B, batch.slot, batch.maxW,
                                                     batch.maxH, R, float(P1), float(P2), rx, ry,
                                                     ctypes.byref(nb)))
    ws = L.workspace.get(nb.value)
    L.check(L.lib().sv_refine_ragged(mode, L.ptr(left.contiguous()), L.ptr(right.contiguous()),
                                     _DT[left.dtype], L.ptr(d0.contiguous()), batch.B,
                                     L.ptr(batch.table), batch.maxW, batch.maxH, batch.slot, R,
                                     float(P1), float(P2), rx, ry, float(tol), L.ptr(out),
                                     L.ptr(status), L.ptr(ws), nb.value, L.stream()))
    return out, status


def net_band_ragged(net, batch, left, right, R=10, P1=12.0, P2=150.0, rx=4, ry=3, tol=2.0):
    """The plain pipeline over a ragged batch: predict_auto on every pair
    (one ragged network pass per stage), widening to float64, then
    refine_disparity(planes=False).  Returns (disparity float64, status,
    AutoResult), packed [B][slot]."""
    res = N.predict_auto_batch(net, left, right, None, None, batch=batch)
    d0 = torch.empty(batch.B, batch.slot, dtype=torch.float64, device=left.device)
    N._call("sv_widen", N._p(res.full), N._p(batch.table), batch.slot, N._p(d0), batch.slot, None,
            batch.B, batch.slot, L.stream())
    out, st = refine_ragged(batch, left, right, d0, 2, R, P1, P2, rx, ry, tol)
    return out, st, res


def fallback_ragged(batch, left, right, out, status, max_disps, tau=0.1, statuses=(1, 2),
                    whole=True, P1=12.0, P2=150.0, rx=4, ry=3):
    """Full-range fallback after net + band SGM over a ragged batch: the
    pairs whose share of pixels with a status in `statuses` (1 occluded, 2
    left-right mismatch, from refine_ragged) exceeds tau are solved again
    with full-range SGM (each pair's R from its maximum disparity,
    eval_real.py:143-146), one call per pair since their sizes differ.
    whole=True replaces the pair's disparity map, else only its flagged
    pixels.  out is updated in place; returns the indices of the pairs that
    fell back.  The reference has no counterpart: it fills inconsistent
    pixels from disp0 and its neighbours (stereonet.py:5372-5390)."""
    from . import refine as RF
    if status is None:
        raise ValueError("fallback_ragged needs the statuses of refine_ragged mode 1 or 2")
    if len(max_disps) != batch.B:
        raise ValueError("%d maximum disparities for %d pairs" % (len(max_disps), batch.B))
    flags = [torch.zeros_like(s_, dtype=torch.bool) for s_ in batch.unpack(status)]
    for f, s_ in zip(flags, batch.unpack(status)):
        for v in statuses:
            f |= s_ == v
    frac = torch.stack([f.float().mean() for f in flags]).cpu().numpy()
    picked = [n for n in range(batch.B) if frac[n] > tau]
    ls, rs, os_ = batch.unpack(left), batch.unpack(right), batch.unpack(out)
    for n in picked:
        w, h = batch.sizes[n]
        R = int(float(max_disps[n]) + 2) // 2 + 1
        fr = RF.full_range(ls[n], rs[n], w, h, R, P1=P1, P2=P2, rx=rx, ry=ry).reshape(h, w)
        if whole:
            os_[n].copy_(fr)
        else:
            os_[n][flags[n]] = fr[flags[n]]
    return picked


class CapturedPipeline:
    """fn(*static inputs) captured once in a HIP graph (torch.cuda.CUDAGraph)
    after `warmup` eager runs; replay() reruns the whole recorded launch
    sequence on new data written into the same input tensors."""

    def __init__(self, fn, *inputs, warmup=2):
        self.inputs = inputs
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            for _ in range(warmup):
                fn(*inputs)
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.outputs = fn(*inputs)
        torch.cuda.synchronize()

    def replay(self):
        self.graph.replay()
        return self.outputs
