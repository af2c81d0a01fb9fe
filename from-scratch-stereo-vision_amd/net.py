"""StereoNet on the GPU: forward, hand-written backward, losses, Adam,
checkpoints and the predict / predict_at_scale / predict_auto drivers.

Mirrors stereonet.py:999-1388 (network, losses, Adam), :1572-1689 (targets,
train_grads, flat parameter order), :1766-1874 (text checkpoints),
:3833-3960 (inference at arbitrary resolution) and :5891-5897 (predict).
Every layer runs in libstereo_mi.so (csrc/net.hip) in FP32; the float64
reference semantics are kept wherever they decide control flow (working
widths, checkpoint text).  Batches may be ragged: each pair can have its own
size (see csrc/net.hip).
"""

import ctypes
import math

import numpy as np
import torch

from . import _lib as L

_P = ctypes.c_void_p
_I = ctypes.c_int
_F = ctypes.c_float
_D = ctypes.c_double
_LL = ctypes.c_longlong
_SZ = ctypes.c_size_t

L.extra_signatures({
    "sv_conv_fwd": (_I, [_P, _P, _P, _LL, _I, _I, _I, _P, _P, _I, _I, _I, _I, _P, _P, _LL, _I, _I,
                         _I, _I, _F, _P, _LL, _P]),
    "sv_conv_bwd_data": (_I, [_P, _P, _LL, _P, _I, _I, _I, _I, _I, _I, _I, _P, _P, _LL, _I, _I, _I,
                              _I, _P, _P]),
    "sv_conv_bwd_weight_workspace_bytes": (_I, [_I, _I, _I, _I, _I, ctypes.POINTER(_SZ)]),
    "sv_conv_bwd_weight": (_I, [_P, _P, _P, _LL, _I, _I, _I, _P, _P, _LL, _I, _I, _I, _I, _I, _I,
                                _P, _P, _P, _SZ, _P]),
    "sv_act_bwd": (_I, [_P, _P, _LL, _F, _P]),
    "sv_costvol_bwd": (_I, [_P, _P, _P, _LL, _I, _I, _P, _LL, _P, _P, _I, _I, _I, _P]),
    "sv_softargmin_fwd": (_I, [_P, _P, _LL, _I, _F, _P, _LL, _P, _I, _I, _P]),
    "sv_softargmin_bwd": (_I, [_P, _P, _P, _P, _P, _LL, _I, _LL, _P, _I, _I, _P]),
    "sv_conv3_softargmin": (_I, [_P, _P, _LL, _I, _I, _P, _P, _F, _P, _LL, _P, _LL, _I, _I, _I,
                                 _P]),
    "sv_load_pair": (_I, [_P, _P, _I, _P, _LL, _P, _LL, _P, _P, _LL, _P, _P, _I, _I, _I, _P]),
    "sv_refine_in": (_I, [_P, _P, _LL, _P, _P, _LL, _P, _P, _I, _I, _F, _P, _LL, _I, _I, _I, _P]),
    "sv_bilinear_fwd": (_I, [_P, _I, _P, _LL, _I, _P, _P, _LL, _I, _F, _P, _I, _I, _I, _P]),
    "sv_bilinear_bwd": (_I, [_P, _P, _LL, _I, _I, _P, _P, _LL, _F, _I, _I, _I, _P]),
    "sv_avgpool2": (_I, [_P, _I, _P, _LL, _P, _P, _LL, _I, _F, _I, _I, _I, _P]),
    "sv_copy_channel": (_I, [_P, _I, _P, _LL, _P, _LL, _I, _I, _I, _P]),
    "sv_pseudo_huber": (_I, [_P, _P, _P, _P, _LL, _I, _F, _D, _D, _P, _P, _P, _P, _I, _P]),
    "sv_disparity_ce": (_I, [_P, _P, _P, _P, _LL, _LL, _I, _F, _I, _D, _D, _P, _P, _P, _P, _I, _P]),
    "sv_adam_step": (_I, [_P, _P, _P, _P, _LL, _F, _F, _F, _F, _F, _F, _F, _P]),
    "sv_select_kth_workspace_bytes": (_I, [_I, ctypes.POINTER(_SZ)]),
    "sv_select_kth": (_I, [_P, _P, _LL, _P, _P, _I, _LL, _P, _SZ, _P]),
    "sv_pool_valid": (_I, [_P, _P, _LL, _P, _P, _LL, _I, _I, _I, _P]),
    "sv_axpby": (_I, [_P, _P, _LL, _F, _F, _P]),
    "sv_widen": (_I, [_P, _P, _LL, _P, _LL, _P, _I, _I, _P]),
    "sv_auto_widths": (_I, [_P, _I, _P, _P, _D, _I, _I, _I, _P, _P, _P, _P, _P, _P]),
    "sv_conv_mfma_pack_bytes": (_I, [_I, _I, _I, ctypes.POINTER(_SZ)]),
    "sv_conv_fwd_mfma": (_I, [_P, _P, _P, _LL, _I, _I, _I, _P, _P, _I, _I, _I, _I, _P, _P, _LL, _I,
                              _LL, _I, _F, _P, _LL, _P]),
})


def _bf16_bits(a):
    """float -> bfloat16 bit patterns, round to nearest even."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def _pack_mfma(w, ic, oc, kd):
    """Weight [OC][IC][KD*9] -> the B operand of v_mfma_f32_16x16x16_bf16
    for k = tap*IC + ic: pack[ks][t][lane][i] = B[16ks + 4(lane/16) + i]
    [16t + lane%16] (csrc/mfma_conv.hip)."""
    K = ic * kd * 9
    ks = (K + 15) // 16
    nt = (oc + 15) // 16
    B = np.zeros((ks * 16, nt * 16), dtype=np.float32)
    wr = w.reshape(oc, ic, kd * 9)
    k = np.arange(K)
    B[:K, :oc] = wr[:, k % ic, k // ic].T
    arr = B.reshape(ks, 4, 4, nt, 16).transpose(0, 3, 1, 4, 2).reshape(ks, nt, 64, 4)
    return _bf16_bits(arr)

LEVEL_WEIGHTS = [0.5, 0.7, 1.0]   # stereonet.py:1615
CE_WEIGHT = 2.0                   # stereonet.py:1616
SLOPE = 0.2                       # LeakyReLU slope, stereonet.py:701
INVALID_PENALTY = 8.0             # CostVolume.invalid_mask, stereonet.py:827


def _co(n, k=3, s=1, p=1, d=1):
    """Conv output size with Python floor division (stereonet.py:483-487)."""
    return (n + 2 * p - (d * (k - 1) + 1)) // s + 1


def _call(name, *args):
    L.check(getattr(L.lib(), name)(*args))


def _p(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


class _Conv:
    """One convolution of the network: weight/bias offsets in the flat
    parameter vector and its geometry (2D: kd=1; 3D: kd=3)."""

    def __init__(self, ic, oc, kd, stride, dil, off):
        self.ic, self.oc, self.kd, self.s, self.dil = ic, oc, kd, stride, dil
        self.p = dil  # (dil * (k - 1)) // 2 with k = 3 (stereonet.py:471); 1 for 3D
        self.nw = oc * ic * kd * 9
        self.w_off = off
        self.b_off = off + self.nw
        self.end = self.b_off + oc


class StereoNet:
    """StereoNet (stereonet.py:1172-1268) with parameters in one flat FP32
    device vector ordered like all_params (:1195-1197): feature c1..c4,
    filter c1..c3, refinement 1 convs + c3, refinement 2 convs + c3, each as
    (w, b)."""

    def __init__(self, rng=None, feat_C=8, cost_C=4, D=8, max_disp=32,
                 refine_C=8, refine_dilations=(1, 2, 4), seed=None):
        self.feat_C, self.cost_C, self.D, self.max_disp = feat_C, cost_C, D, max_disp
        self.refine_C = refine_C
        self.refine_dilations = tuple(refine_dilations)
        self.target_peak = 20          # stereonet.py:1193
        off = 0
        self.feat = []
        for i, (ic, s) in enumerate(((1, 2), (feat_C, 2), (feat_C, 1), (feat_C, 1))):
            c = _Conv(ic, feat_C, 1, s, 1, off)
            self.feat.append(c)
            off = c.end
        self.filt = []
        for ic, oc in ((feat_C, cost_C), (cost_C, cost_C), (cost_C, 1)):
            c = _Conv(ic, oc, 3, 1, 1, off)
            self.filt.append(c)
            off = c.end
        self.ref = []
        for _ in range(2):
            convs, prev = [], 2
            for d in self.refine_dilations:
                c = _Conv(prev, refine_C, 1, 1, d, off)
                convs.append(c)
                off = c.end
                prev = refine_C
            c = _Conv(prev, 1, 1, 1, 1, off)
            convs.append(c)
            off = c.end
            self.ref.append(convs)
        self.nparams = off
        dev = L.device()
        self.params = torch.zeros(off, dtype=torch.float32, device=dev)
        self.grads = torch.zeros_like(self.params)
        self._init(seed if seed is not None else (rng if isinstance(rng, int) else 1))
        self._plans = {}
        self._cache = None
        self.bf16 = False
        self._packs = {}
        # per-layer kernel variants chosen by the autotuner: launch shape ->
        # ((knob, value), ...); and a recorder of conv launches it uses
        self._conv_tune = {}
        self._conv_hook = None
        # fused launches (identical results): the pass input with its
        # guides, the last cost-filter layer with the soft-argmin, each
        # refinement level's upsampling with its first convolution
        self.fused = True

    # -- precision -------------------------------------------------------
    def use_bf16(self, on=True):
        """Run every convolution with IC % 4 == 0 and OC >= 8 on the matrix
        cores (BF16 operands, FP32 accumulation); the rest, soft-argmin,
        losses and Adam stay FP32.  Weights are packed now; call again after
        changing the parameters."""
        self.bf16 = bool(on)
        self._packs = {}
        if not on:
            return
        host = self.params.detach().cpu().numpy()
        for c in self.convs():
            if c.ic % 4 == 0 and c.oc >= 8:
                pk = _pack_mfma(host[c.w_off:c.b_off], c.ic, c.oc, c.kd)
                self._packs[c.w_off] = torch.from_numpy(pk.view(np.int16).copy()).to(self.params.device)

    # -- parameters ------------------------------------------------------
    def convs(self):
        return self.feat + self.filt + self.ref[0] + self.ref[1]

    def _init(self, seed):
        """He initialisation (stereonet.py:466-478, :590-593) from a seeded
        numpy generator; the last filter layer is scaled by 0.25 (:1083)."""
        g = np.random.default_rng(seed)
        p = np.zeros(self.nparams, dtype=np.float64)
        for c in self.convs():
            std = math.sqrt(2.0 / (c.ic * c.kd * 9))
            p[c.w_off:c.b_off] = g.normal(size=c.nw) * std
        c3 = self.filt[2]
        p[c3.w_off:c3.b_off] *= 0.25
        self.params.copy_(torch.from_numpy(p).float())

    def all_params(self):
        """[(param_view, grad_view)] per block, in the reference's order."""
        out = []
        for c in self.convs():
            out.append((self.params[c.w_off:c.b_off], self.grads[c.w_off:c.b_off]))
            out.append((self.params[c.b_off:c.end], self.grads[c.b_off:c.end]))
        return out

    def modules(self):
        return self.convs()

    # -- plans -------------------------------------------------------------
    def _plan(self, dims):
        key = tuple(dims)
        pl = self._plans.get(key)
        if pl is None:
            pl = _Plan(self, dims)
            if len(self._plans) > 64:
                self._plans.clear()
            self._plans[key] = pl
        return pl

    # -- forward -------------------------------------------------------------
    def forward_batch(self, left, right, dims, guide_half=None, guide_full=None):
        """Forward over a (possibly ragged) batch.  left/right: sequences of
        per-pair tensors or one contiguous tensor [N][H][W] (uint8 or
        float32, on the GPU); dims: [(H, W)] per pair.  Returns the plan whose
        buffers hold d0, d1, d2 (views via plan.out(level, n))."""
        pl = self._plan(dims)
        pl.load_images(left, right)
        if guide_half is not None or guide_full is not None:
            pl.load_guides(guide_half, guide_full)
        pl.forward()
        self._cache = pl
        return pl

    def forward_resized(self, left, right, W, H, dims, src=None):
        """Forward over N pairs given at one source size (left/right:
        [N][H][W] uint8 or float32 on the GPU), each bilinearly resized
        (resize_gray, stereonet.py:3843-3845) straight into the plan's input
        buffer at its own working size dims[n] = (h, w).  src: a _Level of
        per-pair source sizes for inputs packed [N][src.slot] instead."""
        pl = self._plan(dims)
        pl.load_resized(left, right, W, H, src)
        pl.forward()
        self._cache = pl
        return pl

    def forward(self, left, right, H, W, guide_half, guide_full):
        """StereoNet.forward (stereonet.py:1203-1233): one pair; note the
        (H, W) argument order.  Returns (d0, d1, d2) in the input's kind."""
        kind = _kind(left)
        lt, rt = _img(left), _img(right)
        pl = self.forward_batch([lt.reshape(-1)], [rt.reshape(-1)], [(H, W)],
                                guide_half=_img(guide_half), guide_full=_img(guide_full))
        return tuple(_ret(pl.out(l, 0), kind) for l in range(3))

    def backward(self, g0, g1, g2, gcost_extra=None):
        """StereoNet.backward (stereonet.py:1235-1267) for the last forward
        of one pair: accumulates into self.grads."""
        pl = self._cache
        if pl is None or pl.N != 1:
            raise RuntimeError("backward() needs a preceding single-pair forward()")
        pl.seed_grads([_img(g0)], [_img(g1)], [_img(g2)],
                      None if gcost_extra is None else [_img(gcost_extra)])
        pl.backward()


def _kind(a):
    if isinstance(a, torch.Tensor):
        return "torch"
    if isinstance(a, np.ndarray):
        return "numpy"
    return "list"


def _img(a, dtype=None):
    if a is None:
        return None
    dev = L.device()
    if isinstance(a, torch.Tensor):
        t = a
    elif isinstance(a, np.ndarray):
        t = torch.from_numpy(np.ascontiguousarray(a))
    else:
        t = torch.tensor(a, dtype=torch.float64)
    if dtype is not None:
        t = t.to(dtype)
    elif t.dtype not in (torch.uint8, torch.float32):
        t = t.to(torch.float32)
    return t.to(dev).contiguous()


def _ret(t, kind):
    if kind == "torch":
        return t
    a = t.detach().cpu().numpy().astype(np.float64)
    return a if kind == "numpy" else a.ravel().tolist()


_TABLES = {}


def _dims_table(hw, dev):
    """Device copy of a dims table, cached (host-to-device copies cost a
    synchronising transfer each)."""
    key = (tuple(hw), str(dev))
    t = _TABLES.get(key)
    if t is None:
        if len(_TABLES) > 512:
            _TABLES.clear()
        t = torch.tensor([v for d in hw for v in d], dtype=torch.int32, device=dev)
        _TABLES[key] = t
    return t


class _Level:
    """Per-sample dims of one resolution level: a device table and the slot
    size (largest h*w).  With `table` given, the table is that device view
    (filled on the device) and `hw` only bounds each sample's dims."""

    def __init__(self, hw, dev, reps=1, table=None):
        hw = list(hw) * reps
        self.hw = hw
        self.table = _dims_table(hw, dev) if table is None else table
        self.maxh = max(h for h, _ in hw)
        self.maxw = max(w for _, w in hw)
        self.slot = max(h * w for h, w in hw)


class _Plan:
    """Buffers and dims tables of one forward/backward over a batch of pairs
    with given sizes."""

    def __init__(self, net, dims, tables=None):
        """dims: [(h, w)] per pair.  tables: optional device int32 block of
        18N entries ([full N][half N][f4 N][full2 2N][f1_2 2N][f4_2 2N]
        (h, w) pairs) that the device fills before each forward; dims then
        bound every pair's size and size the buffers and grids."""
        self.net = net
        self.dims = list(dims)
        self.N = N = len(dims)
        dev = L.device()
        self.dev = dev
        full = [(h, w) for h, w in dims]
        half = [(h // 2, w // 2) for h, w in dims]
        f1 = [(_co(h, s=2), _co(w, s=2)) for h, w in dims]
        f4 = [(_co(h, s=2), _co(w, s=2)) for h, w in f1]
        if any(h < 1 or w < 1 for h, w in half):
            raise ValueError("image too small for the network: %s" % (dims,))
        t = (lambda a, b: None) if tables is None else (lambda a, b: tables[2 * a * N:2 * b * N])
        self.full, self.half = _Level(full, dev, table=t(0, 1)), _Level(half, dev, table=t(1, 2))
        self.f4 = _Level(f4, dev, table=t(2, 3))
        self.full2 = _Level(full, dev, 2, table=t(3, 5))
        self.f1_2 = _Level(f1, dev, 2, table=t(5, 7))
        self.f4_2 = _Level(f4, dev, 2, table=t(7, 9))
        C, Cc, D, RC = net.feat_C, net.cost_C, net.D, net.refine_C
        z = lambda n: torch.empty(max(int(n), 1), dtype=torch.float32, device=dev)
        # activations (the forward tape)
        self.img = z(2 * N * self.full.slot)                # [2N][H][W]
        self.fa = [z(2 * N * C * self.f1_2.slot)] + [z(2 * N * C * self.f4_2.slot) for _ in range(3)]
        self.ca = [z(N * Cc * D * self.f4.slot), z(N * Cc * D * self.f4.slot), z(N * D * self.f4.slot)]
        self.probs = z(N * D * self.f4.slot)
        self.d0 = z(N * self.f4.slot)
        nd = len(net.refine_dilations)
        self.rin = [z(N * 2 * self.half.slot), z(N * 2 * self.full.slot)]
        self.ra = [[z(N * RC * lv.slot) for _ in range(nd)] for lv in (self.half, self.full)]
        self.d1 = z(N * self.half.slot)
        self.d2 = z(N * self.full.slot)
        self._g = None
        self._wt = None
        self.has_extra = False
        # launch-shape tag of the per-layer kernel variants: the passes of
        # predict_auto run one plan at different per-pair sizes (filled on
        # the device), so each pass keys its own choices
        self.tag = 0

    # slot strides
    def out(self, level, n):
        lv = (self.f4, self.half, self.full)[level]
        buf = (self.d0, self.d1, self.d2)[level]
        h, w = lv.hw[n]
        return buf[n * lv.slot: n * lv.slot + h * w]

    def load_images(self, left, right):
        """Copy the pairs into the float image buffer ([2N][H][W], left
        images first); uint8 is read as value / 255."""
        N = self.N
        for side, imgs in ((0, left), (1, right)):
            if isinstance(imgs, torch.Tensor) and imgs.dim() >= 2 and imgs.shape[0] == N and all(
                    d == self.dims[0] for d in self.dims):
                src = imgs.reshape(N, -1).contiguous()
                dt = L.SV_U8 if src.dtype == torch.uint8 else L.SV_F32
                if src.dtype not in (torch.uint8, torch.float32):
                    src = src.float()
                _call("sv_copy_channel", _p(src), dt, _p(self.full.table), src.shape[1],
                      _p(self.img[side * N * self.full.slot:]), self.full.slot, 0, N,
                      self.full.slot, L.stream())
            else:
                for n in range(N):
                    t = imgs[n]
                    if t.dtype not in (torch.uint8, torch.float32):
                        t = t.float()
                    t = t.reshape(-1).contiguous()
                    dt = L.SV_U8 if t.dtype == torch.uint8 else L.SV_F32
                    tab = self.full.table[2 * n:2 * n + 2]
                    dst = self.img[(side * N + n) * self.full.slot:]
                    _call("sv_copy_channel", _p(t), dt, _p(tab), t.numel(), _p(dst),
                          self.full.slot, 0, 1, t.numel(), L.stream())
        # guides: avg-pooled left image -> channel 1 of the level-1 input,
        # the left image itself -> channel 1 of the level-2 input
        self._guides()

    def load_resized(self, left, right, W, H, src=None):
        N = self.N
        if src is None:
            src = _Level([(H, W)] * N, self.dev)
        if getattr(self.net, "fused", False):
            l = left.reshape(N, -1)
            r = right.reshape(N, -1)
            if l.dtype not in (torch.uint8, torch.float32):
                l, r = l.float(), r.float()
            l, r = l.contiguous(), r.contiguous()
            _call("sv_load_pair", _p(l), _p(r), L.SV_U8 if l.dtype == torch.uint8 else L.SV_F32,
                  _p(src.table), src.slot, _p(self.img), self.full.slot, _p(self.full.table),
                  _p(self.rin[0]), self.half.slot, _p(self.half.table), _p(self.rin[1]), N,
                  self.full.maxh, self.full.maxw, L.stream())
            return
        for side, imgs in ((0, left), (1, right)):
            s = imgs.reshape(N, -1)
            if s.dtype not in (torch.uint8, torch.float32):
                s = s.float()
            s = s.contiguous()
            dt = L.SV_U8 if s.dtype == torch.uint8 else L.SV_F32
            _call("sv_bilinear_fwd", _p(s), dt, _p(src.table), src.slot, 1,
                  _p(self.img[side * N * self.full.slot:]), _p(self.full.table), self.full.slot, 0,
                  1.0, None, N, self.full.maxh, self.full.maxw, L.stream())
        self._guides()

    def _guides(self):
        N = self.N
        _call("sv_avgpool2", _p(self.img), L.SV_F32, _p(self.full.table), self.full.slot,
              _p(self.rin[0]), _p(self.half.table), 2 * self.half.slot, 1, 1.0, N,
              self.half.maxh, self.half.maxw, L.stream())
        _call("sv_copy_channel", _p(self.img), L.SV_F32, _p(self.full.table), self.full.slot,
              _p(self.rin[1]), 2 * self.full.slot, 1, N, self.full.slot, L.stream())

    def load_guides(self, gh, gf):
        """Explicit guide images (StereoNet.forward's guide_half/guide_full)."""
        N = self.N
        if gh is not None:
            _call("sv_copy_channel", _p(gh.float().contiguous()), L.SV_F32, _p(self.half.table),
                  self.half.slot, _p(self.rin[0]), 2 * self.half.slot, 1, N, self.half.slot,
                  L.stream())
        if gf is not None:
            _call("sv_copy_channel", _p(gf.float().contiguous()), L.SV_F32, _p(self.full.table),
                  self.full.slot, _p(self.rin[1]), 2 * self.full.slot, 1, N, self.full.slot,
                  L.stream())

    def _conv(self, c, x, xlv, y, ylv, N, mode, D=1, xr=None, resid=None, rstride=0):
        P = self.net.params
        pk = self.net._packs.get(c.w_off) if self.net.bf16 else None
        xst = c.ic * (1 if xr is not None else D) * xlv.slot
        if pk is not None:
            _call("sv_conv_fwd_mfma", _p(x), _p(xr), _p(xlv.table), xst, c.ic, D, c.kd, _p(pk),
                  _p(P[c.b_off:]), c.oc, c.s, c.p, c.dil, _p(y), _p(ylv.table),
                  c.oc * D * ylv.slot, N, D * ylv.slot, mode, SLOPE, _p(resid), rstride,
                  L.stream())
            return
        fn = "sv_conv_fwd"
        args = (_p(x), _p(xr), _p(xlv.table), xst, c.ic, D, c.kd, _p(P[c.w_off:]), _p(P[c.b_off:]),
                c.oc, c.s, c.p, c.dil, _p(y), _p(ylv.table), c.oc * D * ylv.slot, N, ylv.maxh,
                ylv.maxw, mode, SLOPE, _p(resid), rstride, L.stream())
        key = (c.w_off, D, N, ylv.maxh, ylv.maxw, self.tag)
        if self.net._conv_hook is not None:
            self.net._conv_hook(key, fn, args, y[:c.oc * D * ylv.slot * N], (xlv.table, ylv.table))
        variant = self.net._conv_tune.get(key)
        if variant is None:
            _call(fn, *args)
            return
        # this layer's own kernel variant (identical results), then back
        lib = L.lib()
        old = []
        for knob, v in variant:
            cur = ctypes.c_int(0)
            L.check(lib.sv_tune_get(knob, ctypes.byref(cur)))
            old.append((knob, cur.value))
            L.check(lib.sv_tune_set(knob, v))
        try:
            _call(fn, *args)
        finally:
            for knob, v in old:
                lib.sv_tune_set(knob, v)

    def forward(self):
        net, N = self.net, self.N
        C, Cc, D = net.feat_C, net.cost_C, net.D
        # siamese features on 2N images (stereonet.py:1036-1046)
        x, xlv = self.img, self.full2
        lvs = [self.f1_2, self.f4_2, self.f4_2, self.f4_2]
        for i, c in enumerate(net.feat):
            self._conv(c, x, xlv, self.fa[i], lvs[i], 2 * N, 1 if i < 3 else 0)
            x, xlv = self.fa[i], lvs[i]
        # cost volume fused into the first 3D conv, then two more (:1088-1094)
        fl = self.fa[3]
        fr = self.fa[3][N * C * self.f4.slot:]
        c1, c2, c3 = net.filt
        self._conv(c1, fl, self.f4, self.ca[0], self.f4, N, 1, D=D, xr=fr)
        self._conv(c2, self.ca[0], self.f4, self.ca[1], self.f4, N, 1, D=D)
        if getattr(net, "fused", False) and c3.oc == 1 and D in (8, 12):
            # last layer + soft-argmin in one launch (the filtered costs stay
            # in registers)
            _call("sv_conv3_softargmin", _p(self.ca[1]), _p(self.f4.table), c3.ic * D * self.f4.slot,
                  c3.ic, D, _p(net.params[c3.w_off:]), _p(net.params[c3.b_off:]), INVALID_PENALTY,
                  _p(self.d0), self.f4.slot, _p(self.probs), D * self.f4.slot, N, self.f4.maxh,
                  self.f4.maxw, L.stream())
        else:
            self._conv(c3, self.ca[1], self.f4, self.ca[2], self.f4, N, 0, D=D)
            # soft-argmin with the invalid-disparity penalty (:1216-1218)
            _call("sv_softargmin_fwd", _p(self.ca[2]), _p(self.f4.table), D * self.f4.slot, D,
                  INVALID_PENALTY, _p(self.d0), self.f4.slot, _p(self.probs), N, self.f4.slot,
                  L.stream())
        # two refinement levels (:1220-1230)
        src, slv = self.d0, self.f4
        for r, (lv, dout) in enumerate(((self.half, self.d1), (self.full, self.d2))):
            convs = net.ref[r]
            c1 = convs[0]
            first = 0
            if getattr(net, "fused", False) and c1.ic == 2 and c1.oc in (8, 32) and 1 <= c1.dil <= 4:
                # upsampling + first convolution in one launch
                _call("sv_refine_in", _p(src), _p(slv.table), slv.slot, _p(self.rin[r]),
                      _p(lv.table), lv.slot, _p(net.params[c1.w_off:]), _p(net.params[c1.b_off:]),
                      c1.oc, c1.dil, SLOPE, _p(self.ra[r][0]), c1.oc * lv.slot, N, lv.maxh,
                      lv.maxw, L.stream())
                first = 1
            else:
                _call("sv_bilinear_fwd", _p(src), L.SV_F32, _p(slv.table), slv.slot, 1,
                      _p(self.rin[r]), _p(lv.table), 2 * lv.slot, 0, 2.0, None, N, lv.maxh,
                      lv.maxw, L.stream())
            x = self.rin[r] if first == 0 else self.ra[r][0]
            for i, c in enumerate(convs[:-1]):
                if i < first:
                    continue
                self._conv(c, x, lv, self.ra[r][i], lv, N, 1)
                x = self.ra[r][i]
            self._conv(convs[-1], x, lv, dout, lv, N, 2, resid=self.rin[r], rstride=2 * lv.slot)
            src, slv = dout, lv

    # -- backward ------------------------------------------------------------
    def _grad_bufs(self):
        if self._g is None:
            net, N = self.net, self.N
            C, Cc, D, RC = net.feat_C, net.cost_C, net.D, net.refine_C
            z = lambda n: torch.empty(max(int(n), 1), dtype=torch.float32, device=self.dev)
            g = {}
            g["g0"], g["g1"], g["g2"] = z(N * self.f4.slot), z(N * self.half.slot), z(N * self.full.slot)
            g["gextra"] = z(N * D * self.f4.slot)
            g["gcost"] = z(N * D * self.f4.slot)
            g["ca"] = z(N * max(Cc, C) * D * self.f4.slot)
            g["cb"] = z(N * max(Cc, C) * D * self.f4.slot)
            fs = max(self.f4_2.slot, self.f1_2.slot)
            g["fa"], g["fb"] = z(2 * N * C * fs), z(2 * N * C * fs)
            g["rg"], g["rskip"], g["gd"] = z(N * self.full.slot), z(N * self.full.slot), z(N * self.full.slot)
            g["rx"], g["ry"] = z(N * RC * self.full.slot), z(N * RC * self.full.slot)
            g["up1"], g["up0"] = z(N * self.half.slot), z(N * self.f4.slot)
            self._g = g
        return self._g

    def seed_grads(self, g0, g1, g2, gextra=None):
        g = self._grad_bufs()
        for lvl, (lv, src) in enumerate(((self.f4, g0), (self.half, g1), (self.full, g2))):
            dst = g[("g0", "g1", "g2")[lvl]]
            for n, t in enumerate(src):
                h, w = lv.hw[n]
                dst[n * lv.slot:n * lv.slot + h * w].copy_(t.reshape(-1).float())
        self.has_extra = gextra is not None
        if gextra is not None:
            D = self.net.D
            for n, t in enumerate(gextra):
                h, w = self.f4.hw[n]
                g["gextra"][n * D * self.f4.slot:n * D * self.f4.slot + D * h * w].copy_(
                    t.reshape(-1).float())

    def _wgrad(self, c, x, xlv, gy, ylv, N, D=1, xr=None):
        net = self.net
        npos = max(h * w for h, w in ylv.hw) * D
        nchunk = int(max(1, min(64, npos // (64 * 4))))
        nb = ctypes.c_size_t(0)
        _call("sv_conv_bwd_weight_workspace_bytes", N, c.ic, c.oc, c.kd, nchunk, ctypes.byref(nb))
        ws = L.workspace.get(nb.value)
        # the fused cost-volume layer reads C-channel feature maps (no depth)
        xst = c.ic * (1 if xr is not None else D) * xlv.slot
        _call("sv_conv_bwd_weight", _p(x), _p(xr), _p(xlv.table), xst, c.ic, D,
              c.kd, _p(gy), _p(ylv.table), c.oc * D * ylv.slot, c.oc, c.s, c.p, c.dil, N, nchunk,
              _p(net.grads[c.w_off:]), _p(net.grads[c.b_off:]), _p(ws), nb.value, L.stream())

    def _dgrad(self, c, gy, ylv, gx, xlv, N, D=1, ic_count=0):
        # stride-1 layers run as transposed convolutions on the forward
        # engines (flipped weights in a small per-plan buffer)
        ws = None
        if c.s == 1:
            need = (ic_count or c.ic) * c.oc * c.kd * 9 + (ic_count or c.ic)
            if self._wt is None or self._wt.numel() < need:
                self._wt = torch.empty(max(need, 4096), dtype=torch.float32, device=self.dev)
            ws = self._wt
        _call("sv_conv_bwd_data", _p(gy), _p(ylv.table), c.oc * D * ylv.slot, _p(self.net.params[c.w_off:]),
              c.ic, c.oc, D, c.kd, c.s, c.p, c.dil, _p(gx), _p(xlv.table),
              (ic_count or c.ic) * D * xlv.slot, N, xlv.maxh, xlv.maxw, ic_count, _p(ws), L.stream())

    def _act(self, g, y, n, slope):
        _call("sv_act_bwd", _p(g), _p(y), n, slope, L.stream())

    def backward(self):
        """StereoNet.backward (stereonet.py:1235-1267) for the seeded output
        gradients; parameter gradients accumulate into net.grads."""
        net, N, g = self.net, self.N, self._grad_bufs()
        C, Cc, D = net.feat_C, net.cost_C, net.D
        nd = len(net.refine_dilations)
        gout = g["g2"]
        for r in (1, 0):
            lv, dout = ((self.half, self.d1), (self.full, self.d2))[r]
            n_el = N * lv.slot
            # output ReLU of disp + residual (:1154-1156), then the skip copy
            g["rg"][:n_el].copy_(gout[:n_el])
            self._act(g["rg"], dout, n_el, 0.0)
            g["rskip"][:n_el].copy_(g["rg"][:n_el])
            convs, acts = net.ref[r], self.ra[r]
            cur = g["rg"]
            for i in range(nd, -1, -1):
                c = convs[i]
                if i < nd:
                    self._act(cur, acts[i], N * c.oc * lv.slot, SLOPE)
                self._wgrad(c, acts[i - 1] if i > 0 else self.rin[r], lv, cur, lv, N)
                nxt = (g["rx"], g["ry"])[i % 2] if i > 0 else g["gd"]
                self._dgrad(c, cur, lv, nxt, lv, N, ic_count=(0 if i > 0 else 1))
                cur = nxt
            # skip connection into the disparity input (:1165-1167)
            _call("sv_axpby", _p(g["rskip"]), _p(g["gd"]), n_el, 1.0, 1.0, L.stream())
            # x2 upsample of the coarser level, plus that level's own gradient
            slv, sgrad, dst = ((self.f4, g["g0"], g["up0"]), (self.half, g["g1"], g["up1"]))[r]
            _call("sv_bilinear_bwd", _p(g["gd"]), _p(lv.table), lv.slot, 0, 1, _p(dst),
                  _p(slv.table), slv.slot, 2.0, N, slv.maxh, slv.maxw, L.stream())
            _call("sv_axpby", _p(sgrad), _p(dst), N * slv.slot, 1.0, 1.0, L.stream())
            gout = dst
        # soft-argmin (+ cross-entropy's direct cost gradient), :1253-1257
        _call("sv_softargmin_bwd", _p(self.probs), _p(self.d0), _p(g["up0"]),
              _p(g["gextra"]) if self.has_extra else None, _p(self.f4.table),
              D * self.f4.slot, D, self.f4.slot, _p(g["gcost"]), N, self.f4.slot, L.stream())
        c1, c2, c3 = net.filt
        self._wgrad(c3, self.ca[1], self.f4, g["gcost"], self.f4, N, D=D)
        self._dgrad(c3, g["gcost"], self.f4, g["ca"], self.f4, N, D=D)
        self._act(g["ca"], self.ca[1], N * Cc * D * self.f4.slot, SLOPE)
        self._wgrad(c2, self.ca[0], self.f4, g["ca"], self.f4, N, D=D)
        self._dgrad(c2, g["ca"], self.f4, g["cb"], self.f4, N, D=D)
        self._act(g["cb"], self.ca[0], N * Cc * D * self.f4.slot, SLOPE)
        fl = self.fa[3]
        fr = self.fa[3][N * C * self.f4.slot:]
        self._wgrad(c1, fl, self.f4, g["cb"], self.f4, N, D=D, xr=fr)
        self._dgrad(c1, g["cb"], self.f4, g["ca"], self.f4, N, D=D)   # d loss / d volume
        gf = g["fa"]
        _call("sv_costvol_bwd", _p(fl), _p(fr), _p(self.f4.table), C * self.f4.slot, C, D,
              _p(g["ca"]), C * D * self.f4.slot, _p(gf), _p(gf[N * C * self.f4.slot:]), N,
              self.f4.maxh, self.f4.maxw, L.stream())
        # siamese feature backward on all 2N images at once (:1259-1267)
        lvs_in = [self.full2, self.f1_2, self.f4_2, self.f4_2]
        lvs_out = [self.f1_2, self.f4_2, self.f4_2, self.f4_2]
        cur, other = g["fa"], g["fb"]
        for i in (3, 2, 1, 0):
            c = net.feat[i]
            if i < 3:
                self._act(cur, self.fa[i], 2 * N * C * lvs_out[i].slot, SLOPE)
            self._wgrad(c, self.fa[i - 1] if i > 0 else self.img, lvs_in[i], cur, lvs_out[i], 2 * N)
            if i > 0:
                self._dgrad(c, cur, lvs_out[i], other, lvs_in[i], 2 * N)
                cur, other = other, cur


# ---------------------------------------------------------------- losses

class _Targets:
    """build_targets (stereonet.py:1595-1612) on the GPU for a batch of
    equally sized samples."""

    def __init__(self, pl, disp, valid):
        N = pl.N
        dev = pl.dev
        z = lambda n, dt=torch.float32: torch.empty(max(int(n), 1), dtype=dt, device=dev)
        self.d = [z(N * pl.f4.slot), z(N * pl.half.slot), z(N * pl.full.slot)]
        self.v = [z(N * pl.f4.slot, torch.uint8), z(N * pl.half.slot, torch.uint8),
                  z(N * pl.full.slot, torch.uint8)]
        for n in range(N):
            h, w = pl.full.hw[n]
            self.d[2][n * pl.full.slot:n * pl.full.slot + h * w].copy_(disp[n].reshape(-1).float())
            self.v[2][n * pl.full.slot:n * pl.full.slot + h * w].copy_(valid[n].reshape(-1).to(torch.uint8))
        # pool_disp = avgpool2 * 0.5, pool_valid = all four (:1572-1592)
        q = _Level([(h // 2, w // 2) for h, w in pl.half.hw], dev)
        self.qlv = q
        _call("sv_avgpool2", _p(self.d[2]), L.SV_F32, _p(pl.full.table), pl.full.slot, _p(self.d[1]),
              _p(pl.half.table), pl.half.slot, 0, 0.5, N, pl.half.maxh, pl.half.maxw, L.stream())
        _call("sv_pool_valid", _p(self.v[2]), _p(pl.full.table), pl.full.slot, _p(self.v[1]),
              _p(pl.half.table), pl.half.slot, N, pl.half.maxh, pl.half.maxw, L.stream())
        _call("sv_avgpool2", _p(self.d[1]), L.SV_F32, _p(pl.half.table), pl.half.slot, _p(self.d[0]),
              _p(q.table), pl.f4.slot, 0, 0.5, N, q.maxh, q.maxw, L.stream())
        _call("sv_pool_valid", _p(self.v[1]), _p(pl.half.table), pl.half.slot, _p(self.v[0]),
              _p(q.table), pl.f4.slot, N, q.maxh, q.maxw, L.stream())


def train_grads_batch(net, lefts, rights, disps, valids, dims, sync=True):
    """train_grads over a batch (see _train_grads_batch); returns the
    per-sample losses and the plan."""
    loss, pl, _ = _train_grads_batch(net, lefts, rights, disps, valids, dims, sync)
    return loss, pl


def _train_grads_batch(net, lefts, rights, disps, valids, dims, sync=True):
    """train_grads (stereonet.py:1619-1653) over a batch: per-sample losses
    (three pseudo-Huber levels with LEVEL_WEIGHTS, cross-entropy with
    CE_WEIGHT), gradients averaged over the batch like train_parallel.py's
    host average (:146-153).  Leaves the mean gradient in net.grads; returns
    the per-sample losses (float64 numpy, or the device tensor when
    sync=False) and the plan."""
    for (h, w) in dims:
        if _co(_co(h, s=2), s=2) != (h // 2) // 2 or _co(_co(w, s=2), s=2) != (w // 2) // 2:
            raise ValueError("training needs H and W divisible by 4 (got %dx%d)" % (w, h))
    net.grads.zero_()
    pl = net.forward_batch(lefts, rights, dims)
    N = pl.N
    tg = _Targets(pl, disps, valids)
    g = pl._grad_bufs()
    loss = torch.zeros(N, dtype=torch.float64, device=pl.dev)
    nblk = 8
    part = torch.empty(2 * N * nblk, dtype=torch.float64, device=pl.dev)
    scale = torch.empty(N, dtype=torch.float32, device=pl.dev)
    outs = (pl.d0, pl.d1, pl.d2)
    lvs = (pl.f4, pl.half, pl.full)
    for lvl in range(3):
        _call("sv_pseudo_huber", _p(outs[lvl]), _p(tg.d[lvl]), _p(tg.v[lvl]), _p(lvs[lvl].table),
              lvs[lvl].slot, N, 1.0, LEVEL_WEIGHTS[lvl], 1.0 / N, _p(g[("g0", "g1", "g2")[lvl]]),
              _p(loss), _p(part), _p(scale), nblk, L.stream())
    _call("sv_disparity_ce", _p(pl.probs), _p(tg.d[0]), _p(tg.v[0]), _p(pl.f4.table),
          net.D * pl.f4.slot, pl.f4.slot, net.D, 0.8, N, CE_WEIGHT, 1.0 / N, _p(g["gextra"]),
          _p(loss), _p(part), _p(scale), nblk, L.stream())
    pl.has_extra = True
    pl.backward()
    return (loss.cpu().numpy() if sync else loss), pl, tg


def _targets_out(pl, tg, kind):
    """The targets dict build_targets returns (stereonet.py:1595-1612) for
    sample 0 of a plan: per-level disparities 'd' and validity 'v' (quarter,
    half, full), the guides and the (W, H) dims of each level."""
    q = tg.qlv.hw[0]
    hw = (q, pl.half.hw[0], pl.full.hw[0])
    d, v = [], []
    for lvl in range(3):
        n = hw[lvl][0] * hw[lvl][1]
        dd, vv = tg.d[lvl][:n], tg.v[lvl][:n]
        if kind == "torch":
            d.append(dd)
            v.append(vv.bool())
        else:
            da = dd.detach().cpu().numpy().astype(np.float64)
            va = vv.detach().cpu().numpy().astype(bool)
            d.append(da if kind == "numpy" else da.tolist())
            v.append(va if kind == "numpy" else va.tolist())
    h2, w2 = pl.half.hw[0]
    h, w = pl.full.hw[0]
    gh = pl.rin[0][h2 * w2:2 * h2 * w2]
    gf = pl.rin[1][h * w:2 * h * w]
    return {"d": d, "v": v, "guide_half": _ret(gh, kind), "guide_full": _ret(gf, kind),
            "dims": [(hw[0][1], hw[0][0]), (w2, h2), (w, h)]}


def train_grads(net, sample):
    """train_grads (stereonet.py:1619-1653) on one sample dict
    {left, right, disp, valid, W, H}: returns (loss, d2, targets) with the
    gradient left in net.grads.  targets is build_targets' dict (:1595-1612),
    in the sample's kind (lists, numpy arrays or device tensors)."""
    W, H = sample["W"], sample["H"]
    kind = _kind(sample["left"])
    l, r = _img(sample["left"]), _img(sample["right"])
    d = _img(sample["disp"], torch.float32)
    v = _img(np.asarray(sample["valid"], dtype=np.uint8)
             if not isinstance(sample["valid"], torch.Tensor) else sample["valid"], torch.uint8)
    losses, pl, tg = _train_grads_batch(net, [l], [r], [d], [v], [(H, W)])
    return float(losses[0]), _ret(pl.out(2, 0), kind), _targets_out(pl, tg, kind)


class Adam:
    """Adam (stereonet.py:1354-1387): elementwise clip, moments,
    running-product bias correction, on the device.

    params is what the reference passes, net.all_params(): a list of
    (param, grad) pairs, here float32 device views.  Blocks that lie back to
    back in one storage (all_params' views of the net's flat vectors) are
    merged, so the whole network is one launch.  A StereoNet may be passed
    instead of its all_params()."""

    def __init__(self, params, lr=1.0e-3, b1=0.9, b2=0.999, eps=1.0e-8):
        if isinstance(params, StereoNet):
            params = params.all_params()
        blocks = list(params)
        for i, blk in enumerate(blocks):
            if not (isinstance(blk, (tuple, list)) and len(blk) == 2):
                raise TypeError("Adam expects (param, grad) pairs as all_params() returns")
            p, g = blk
            if not (isinstance(p, torch.Tensor) and isinstance(g, torch.Tensor)):
                raise TypeError("Adam block %d: param and grad must be device tensors "
                                "(views from StereoNet.all_params())" % i)
            if p.dtype != torch.float32 or g.dtype != torch.float32 or not p.is_cuda or not g.is_cuda:
                raise TypeError("Adam block %d: float32 device tensors expected" % i)
            if not (p.is_contiguous() and g.is_contiguous()) or p.numel() != g.numel():
                raise ValueError("Adam block %d: contiguous param/grad views of equal length "
                                 "expected" % i)
        self.params = blocks
        self.lr, self.b1, self.b2, self.eps = lr, b1, b2, eps
        # runs of consecutive blocks adjacent in memory: (param ptr owner,
        # grad owner, element count); one launch each
        runs = []
        for p, g in blocks:
            if p.numel() == 0:
                continue
            if runs:
                rp, rg, n = runs[-1]
                if (p.data_ptr() == rp.data_ptr() + 4 * n and g.data_ptr() == rg.data_ptr() + 4 * n
                        and p.untyped_storage().data_ptr() == rp.untyped_storage().data_ptr()
                        and g.untyped_storage().data_ptr() == rg.untyped_storage().data_ptr()):
                    runs[-1] = (rp, rg, n + p.numel())
                    continue
            runs.append((p, g, p.numel()))
        self._runs = []
        self.m, self.v = [], []
        for p, g, n in runs:
            m = torch.zeros(n, dtype=torch.float32, device=p.device)
            v = torch.zeros(n, dtype=torch.float32, device=p.device)
            self._runs.append((p, g, n, m, v))
        # per-block moment views, as the reference's lists of lists
        k = 0
        for p, g in blocks:
            if p.numel() == 0:
                self.m.append(p.new_zeros(0))
                self.v.append(p.new_zeros(0))
                continue
            while True:
                rp, rg, n, m, v = self._runs[k]
                off = (p.data_ptr() - rp.data_ptr()) // 4
                if 0 <= off < n and p.untyped_storage().data_ptr() == rp.untyped_storage().data_ptr():
                    break
                k += 1
            self.m.append(m[off:off + p.numel()])
            self.v.append(v[off:off + p.numel()])
        self.b1t = 1.0
        self.b2t = 1.0

    def step(self, clip=5.0):
        self.b1t *= self.b1
        self.b2t *= self.b2
        c1 = 1.0 / (1.0 - self.b1t)
        c2 = 1.0 / (1.0 - self.b2t)
        for p, g, n, m, v in self._runs:
            _call("sv_adam_step", p.data_ptr(), g.data_ptr(), _p(m), _p(v), n,
                  self.lr, self.b1, self.b2, self.eps, c1, c2, clip, L.stream())


def train_step(net, opt, sample):
    total, pred, tgt = train_grads(net, sample)
    opt.step()
    return total, pred, tgt


def flat_params(net):
    return net.params.detach().cpu().double().numpy().tolist()


def set_flat_params(net, flat):
    net.params.copy_(torch.as_tensor(np.asarray(flat, dtype=np.float64)).float())


def flat_grads(net):
    return net.grads.detach().cpu().double().numpy().tolist()


def set_flat_grads(net, flat):
    net.grads.copy_(torch.as_tensor(np.asarray(flat, dtype=np.float64)).float())


# ----------------------------------------------------------- checkpoints

def save_weights(net, path):
    """save_weights (stereonet.py:1766-1787): header line, block count, one
    line per block with '%.9g' values."""
    params = net.all_params()
    with open(path, "w") as f:
        f.write("stereonet feat_C=%d cost_C=%d D=%d max_disp=%d refine_C=%d "
                "refine_dil=%s target_peak=%d\n"
                % (net.feat_C, net.cost_C, net.D, net.max_disp, net.refine_C,
                   ",".join(str(d) for d in net.refine_dilations),
                   int(getattr(net, "target_peak", 20))))
        f.write("%d\n" % len(params))
        host = net.params.detach().cpu().double().numpy()
        for c in net.convs():
            for a, b in ((c.w_off, c.b_off), (c.b_off, c.end)):
                f.write("%d " % (b - a))
                f.write(" ".join("%.9g" % v for v in host[a:b]))
                f.write("\n")


def read_checkpoint_config(path):
    """read_checkpoint_config (stereonet.py:1790-1807)."""
    with open(path) as f:
        first = f.readline().strip()
    if not first.startswith("stereonet "):
        return None
    cfg = {}
    for tok in first.split()[1:]:
        if "=" not in tok:
            continue
        k, v = tok.split("=", 1)
        cfg[k] = tuple(int(d) for d in v.split(",")) if k == "refine_dil" else int(v)
    return cfg


def load_weights(net, path):
    """load_weights (stereonet.py:1856-1874); ValueError on a mismatch."""
    with open(path) as f:
        lines = f.read().split("\n")
    off = 1 if lines[0].startswith("stereonet ") else 0
    count = int(lines[off])
    blocks = net.all_params()
    if count != len(blocks):
        raise ValueError("checkpoint does not match model")
    host = np.zeros(net.nparams)
    pos = 0
    for k in range(count):
        parts = lines[k + 1 + off].split()
        n = int(parts[0])
        if n != blocks[k][0].numel():
            raise ValueError("parameter %d size mismatch" % k)
        host[pos:pos + n] = [float(v) for v in parts[1:n + 1]]
        pos += n
    net.params.copy_(torch.from_numpy(host).float())


_LEGACY = [(8, 4, 4, (1, 2)), (8, 4, 8, (1, 2, 4)), (12, 6, 8, (1, 2, 4)), (16, 8, 8, (1, 2, 4))]


def load_model(path, rng=None):
    """load_model (stereonet.py:1810-1853): build the architecture the
    header names (or, for a headerless file, the shipped architecture whose
    block sizes match), then load it."""
    cfg = read_checkpoint_config(path)
    if cfg is not None:
        net = StereoNet(rng, feat_C=cfg.get("feat_C", 8), cost_C=cfg.get("cost_C", 4),
                        D=cfg.get("D", 8), max_disp=cfg.get("max_disp", 32),
                        refine_C=cfg.get("refine_C", 8),
                        refine_dilations=cfg.get("refine_dil", (1, 2, 4)))
        net.target_peak = cfg.get("target_peak", 20)
        load_weights(net, path)
        return net
    with open(path) as f:
        n_blocks = int(f.readline().strip())
        sizes = [int(f.readline().split(" ", 1)[0]) for _ in range(n_blocks)]
    for (fc, cc, rc, dil) in _LEGACY:
        trial = StereoNet(None, feat_C=fc, cost_C=cc, D=8, max_disp=32, refine_C=rc,
                          refine_dilations=dil)
        tp = trial.all_params()
        if len(tp) == n_blocks and all(tp[i][0].numel() == sizes[i] for i in range(n_blocks)):
            load_weights(trial, path)
            return trial
    raise ValueError("checkpoint %s does not match any known architecture; it may be from a "
                     "different build" % path)


# ------------------------------------------------------------- inference

_KTH = {}


def _select(buf, lv, N, ks):
    """sorted(values)[k] per sample (the k-th smallest), as float32 on the
    device (no host round trip once the k table is cached)."""
    key = (tuple(ks), str(buf.device))
    kt = _KTH.get(key)
    if kt is None:
        if len(_KTH) > 256:
            _KTH.clear()
        kt = _KTH[key] = torch.tensor(ks, dtype=torch.int64, device=buf.device)
    out = torch.empty(N, dtype=torch.float32, device=buf.device)
    nb = ctypes.c_size_t(0)
    _call("sv_select_kth_workspace_bytes", N, ctypes.byref(nb))
    ws = torch.empty(nb.value, dtype=torch.uint8, device=buf.device)
    _call("sv_select_kth", _p(buf), _p(lv.table), lv.slot, _p(kt), _p(out), N, lv.slot, _p(ws),
          nb.value, L.stream())
    return out


def _peak_index(n, q=0.95):
    return int(n * q)


def predict_batch(net, lefts, rights, dims):
    """predict (stereonet.py:5891-5897) for a ragged batch: returns the plan
    (full-resolution outputs via plan.out(2, n))."""
    return net.forward_batch(lefts, rights, dims)


def predict(net, left, right, W, H):
    """predict (stereonet.py:5891-5897): guide_half = avgpool2(left),
    forward, return the full-resolution disparity."""
    kind = _kind(left)
    pl = net.forward_batch([_img(left).reshape(-1)], [_img(right).reshape(-1)], [(H, W)])
    return _ret(pl.out(2, 0).clone(), kind)


def _width_for(pk, W, target_peak, max_width):
    """width_for inside predict_auto (stereonet.py:3905-3913)."""
    w = int(W * target_peak / (pk if pk > 1.0e-6 else 1.0e-6) + 0.5)
    if w > W:
        w = W
    if w > max_width:
        w = max_width
    if w < 48:
        w = 48
    return w


def _work_size(W, H, work_width):
    """predict_at_scale's working size (stereonet.py:3943-3947)."""
    if work_width >= W:
        return W, H, 1.0
    scale = work_width / float(W)
    Hw = int(H * scale + 0.5)
    if Hw < 8:
        Hw = 8
    return work_width, Hw, scale


class AutoResult:
    """Result of predict_auto_batch: full-resolution disparities [N][H][W]
    (float32, on the GPU) plus the decisions taken per pair.  The decisions
    stay on the GPU until `works` or `trace` is read."""

    def __init__(self, full, trace_dev, passes):
        self.full = full
        self._trace_dev = trace_dev
        self._passes = passes
        self._trace = None

    @property
    def trace(self):
        """Working width of every pass each pair ran ([[w0, w1, ...]])."""
        if self._trace is None:
            t = self._trace_dev.reshape(-1, self._passes).cpu().numpy()
            self._trace = [[int(w) for w in row if w > 0] for row in t]
        return self._trace

    @property
    def works(self):
        """Final working width per pair."""
        return [t[-1] for t in self.trace]


class _AutoState:
    """Device state of predict_auto_batch for one batch layout and setting:
    the plan sized for each pair's largest working size, its dims tables,
    and the per-pair widths, flags and 1/scale."""

    def __init__(self, net, sizes, max_width):
        dev = L.device()
        N = len(sizes)
        cap = []
        for W, H in sizes:
            Wc, Hc, _ = _work_size(W, H, max(48, min(W, max_width)))
            cap.append((Hc, Wc))
        self.tables = torch.zeros(18 * N, dtype=torch.int32, device=dev)
        self.plan = _Plan(net, cap, tables=self.tables)
        self.works = torch.zeros(N, dtype=torch.int32, device=dev)
        self.active = torch.zeros(N, dtype=torch.int32, device=dev)
        self.inv = torch.zeros(N, dtype=torch.float32, device=dev)


def predict_auto_batch(net, left, right, W, H, target_peak=None, probe_width=64,
                       max_width=320, passes=3, batch=None):
    """predict_auto (stereonet.py:3872-3929) over N pairs: probe at
    probe_width, choose each pair's working width from its 95th-percentile
    disparity, rerun the pairs whose width moved by more than 8%.
    left/right: [N][H][W] uint8 or float32 on the GPU, or, with batch (a
    RaggedBatch of pairs of their own sizes; W and H unused), packed
    [N][batch.slot].

    Nothing returns to the host: peaks are selected on the GPU, the width
    rule runs there in float64 (sv_auto_widths), and every pass is launched
    for all N pairs at each pair's largest possible working size, pairs that
    are done carrying zero dims so that each kernel skips them.  The
    sequence is therefore fixed for given sizes and settings and can be
    captured in a graph.  Returns AutoResult with full [N][H][W] (or
    [N][slot] for a ragged batch)."""
    N = left.shape[0]
    tp = float(getattr(net, "target_peak", 20)) if target_peak is None else float(target_peak)
    if batch is None:
        sizes = [(W, H)] * N
        src = _Level([(H, W)] * N, left.device)
    else:
        sizes = batch.sizes
        src = batch.level
    probe = []
    for w, h in sizes:
        pw = probe_width if probe_width < w else w
        ph = int(h * pw / float(w) + 0.5)
        probe.append((ph if ph >= 8 else 8, pw))
    key = ("auto", tuple(sizes), max_width)
    stt = net._plans.get(key)
    if stt is None:
        stt = net._plans[key] = _AutoState(net, sizes, max_width)
    trace = torch.empty(N * passes, dtype=torch.int32, device=left.device)
    full = torch.empty(N, src.slot, dtype=torch.float32, device=left.device)
    pl = net.forward_resized(left, right, W, H, probe, src=src)
    peaks = _select(pl.d2, pl.full, N, [_peak_index(h * w) for h, w in probe])
    plw = stt.plan
    for it in range(passes):
        _call("sv_auto_widths", _p(peaks), N, _p(src.table), _p(pl.full.table), tp, max_width, it,
              passes, _p(stt.works), _p(stt.active), _p(trace), _p(stt.tables), _p(stt.inv),
              L.stream())
        plw.load_resized(left, right, W, H, src)
        plw.tag = it + 1
        plw.forward()
        net._cache = plw
        # back to full resolution, values scaled by 1/scale (:3957-3959);
        # pairs that are done keep their previous result
        _call("sv_bilinear_fwd", _p(plw.d2), L.SV_F32, _p(plw.full.table), plw.full.slot, 1,
              _p(full), _p(src.table), src.slot, 0, 1.0, _p(stt.inv), N, src.maxh, src.maxw,
              L.stream())
        if it == passes - 1:
            break
        peaks = _select(full, src, N, [_peak_index(w * h) for w, h in sizes])
    if batch is None:
        full = full.reshape(N, H, W)
    return AutoResult(full, trace, passes)


def predict_at_scale(net, left, right, W, H, work_width=96, verbose=True):
    """predict_at_scale (stereonet.py:3932-3960): returns (disp_full,
    disp_work, W_work, H_work)."""
    kind = _kind(left)
    l, r = _img(left).reshape(1, H, W), _img(right).reshape(1, H, W)
    if work_width >= W:
        d = predict(net, l.reshape(-1), r.reshape(-1), W, H)
        d = _img(d, torch.float32) if kind != "torch" else d
        out = _ret(d, kind)
        return out, out, W, H
    Ww, Hw, scale = _work_size(W, H, work_width)
    if verbose:
        print("  running at %dx%d (input %dx%d, scale %.3f)" % (Ww, Hw, W, H, scale))
    dev = L.device()
    plw = net.forward_resized(l, r, W, H, [(Hw, Ww)])
    dw = plw.out(2, 0).clone()
    full = torch.empty(H * W, dtype=torch.float32, device=dev)
    lv = _Level([(H, W)], dev)
    _call("sv_bilinear_fwd", _p(plw.d2), L.SV_F32, _p(plw.full.table), plw.full.slot, 1, _p(full),
          _p(lv.table), H * W, 0, float(1.0 / scale), None, 1, H, W, L.stream())
    return _ret(full, kind), _ret(dw, kind), Ww, Hw


def predict_auto(net, left, right, W, H, target_peak=None, probe_width=64,
                 max_width=320, passes=3, verbose=True):
    """predict_auto (stereonet.py:3872-3929) for one pair: returns
    (disp_full, disp_work, W_work, H_work)."""
    kind = _kind(left)
    l, r = _img(left).reshape(1, H, W), _img(right).reshape(1, H, W)
    res = predict_auto_batch(net, l, r, W, H, target_peak, probe_width, max_width, passes)
    work = res.works[0]
    if verbose:
        print("  working widths %s" % res.trace[0])
    full = res.full[0].reshape(-1)
    Ww, Hw, _ = _work_size(W, H, work)
    # the working-resolution map of the final pass
    if work >= W:
        dw = full
    else:
        dw = net.forward_resized(l, r, W, H, [(Hw, Ww)]).out(2, 0).clone()
    return _ret(full, kind), _ret(dw, kind), Ww, Hw


def widen(old, feat_C, cost_C, refine_C, seed=77):
    """Net2Net widening as widen_net.py:18-41 does it: every convolution of
    the wider network receives the old weights in its top-left block,
    weights from new input channels into old output channels are zero (so
    the wider net computes exactly the old output), new output channels keep
    their He-initialised weights and get zero bias."""
    new = StereoNet(None, feat_C=feat_C, cost_C=cost_C, D=old.D, max_disp=old.max_disp,
                    refine_C=refine_C, refine_dilations=old.refine_dilations, seed=seed)
    new.target_peak = old.target_peak
    po = old.params.detach().cpu().double().numpy()
    pn = new.params.detach().cpu().double().numpy().copy()
    co, cn = old.convs(), new.convs()
    assert len(co) == len(cn)
    for a, b in zip(co, cn):
        kk = a.kd * 9
        wa = po[a.w_off:a.b_off].reshape(a.oc, a.ic, kk)
        wb = pn[b.w_off:b.b_off].reshape(b.oc, b.ic, kk)
        wb[:a.oc, :, :] = 0.0
        wb[:a.oc, :a.ic, :] = wa
        bb = np.zeros(b.oc)
        bb[:a.oc] = po[a.b_off:a.end]
        pn[b.w_off:b.b_off] = wb.ravel()
        pn[b.b_off:b.end] = bb
    new.params.copy_(torch.from_numpy(pn).float())
    return new
