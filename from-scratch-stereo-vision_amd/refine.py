"""Full-resolution census + band SGM + left-right refinement on the GPU.

Drop-in counterparts of stereonet.py:5115-5391, :5865-5884 and :2231-2258
with the reference's names, argument order and defaults.  Inputs may be
Python lists (returned as lists, like the reference), numpy arrays (returned
as numpy) or torch tensors (returned as CUDA tensors, no host round trip).
Images may be flat (W*H), (H, W) or a batch (B, H, W); uint8 images are
read as gray = value / 255 (the reference's PGM convention,
stereonet.py:284,300), float images are used as they are.

Every call runs hand-written HIP kernels from libstereo_mi.so; there is no
CPU fallback.
"""

import ctypes

import numpy as np
import torch

from . import _lib as L

_DTYPES = {torch.uint8: L.SV_U8, torch.float32: L.SV_F32, torch.float64: L.SV_F64}

_CI, _CD, _CP, _CSZ = ctypes.c_int, ctypes.c_double, ctypes.c_void_p, ctypes.c_size_t
L.extra_signatures({
    "sv_slic": (_CI, [_CP, _CI, _CI, _CI, _CI, _CI, _CD, _CI, _CP, ctypes.POINTER(_CI), _CP, _CP,
                      _CSZ, _CP]),
    "sv_plane_snap_workspace_bytes": (_CI, [_CI, _CI, _CI, _CI, ctypes.POINTER(_CSZ)]),
    "sv_plane_snap": (_CI, [_CP, _CP, _CI, _CP, _CI, _CI, _CI, _CI, _CI, _CD, _CD, _CI,
                            ctypes.c_uint, _CP, _CP, _CP, _CSZ, _CP]),
})


def _kind(a):
    if isinstance(a, torch.Tensor):
        return "torch"
    if isinstance(a, np.ndarray):
        return "numpy"
    return "list"


def _dev(a, dtype=None):
    dev = L.device()
    if isinstance(a, torch.Tensor):
        t = a
    elif isinstance(a, np.ndarray):
        t = torch.from_numpy(np.ascontiguousarray(a))
    else:
        t = torch.tensor(a, dtype=torch.float64 if dtype is None else dtype)
    if dtype is not None and t.dtype != dtype:
        t = t.to(dtype)
    return t.to(dev).contiguous()


def _image(a):
    t = _dev(a)
    if t.dtype not in _DTYPES:
        t = t.to(torch.float64)
    return t, _DTYPES[t.dtype]


def _batch(t, W, H):
    n = t.numel()
    if W <= 0 or H <= 0 or n % (W * H) != 0 or n == 0:
        raise ValueError("array of %d values does not hold W x H = %d x %d images" % (n, W, H))
    return n // (W * H)


def _ret(t, kind, shape=None):
    if kind == "torch":
        return t.reshape(shape) if shape is not None else t
    a = t.cpu().numpy()
    if shape is not None:
        a = a.reshape(shape)
    return a if kind == "numpy" else a.ravel().tolist()


def _shape_of(a):
    return tuple(a.shape) if isinstance(a, (torch.Tensor, np.ndarray)) else None


def census(img, W, H, rx=2, ry=2):
    """census, stereonet.py:5115-5142.  Returns the 64-bit words (as Python
    ints, uint64 numpy, or an int64 tensor holding the same bits)."""
    kind = _kind(img)
    t, dt = _image(img)
    B = _batch(t, W, H)
    out = torch.empty(B * W * H, dtype=torch.int64, device=t.device)
    L.check(L.lib().sv_census(L.ptr(t), dt, B, W, H, rx, ry, L.ptr(out), L.stream()))
    if kind == "torch":
        return out.reshape(_shape_of(img))
    a = out.cpu().numpy().view(np.uint64)
    if kind == "numpy":
        return a.reshape(_shape_of(img))
    return [int(v) for v in a]


def _words(c):
    if isinstance(c, torch.Tensor):
        return _dev(c, torch.int64)
    a = np.asarray(c if isinstance(c, np.ndarray) else [int(v) for v in c], dtype=np.uint64)
    return _dev(a.view(np.int64))


def band_costs(cl, cr, gl, gr, base, W, H, R, ad_weight, ad_trunc, invalid,
               prior=None, prior_w=0.0, prior_t=4.0):
    """band_costs, stereonet.py:5156-5189: costs for d = base[p] + k - R,
    pixel-major [p][k] (W*H*(2R+1) values)."""
    kind = _kind(cl)
    wl, wr = _words(cl), _words(cr)
    B = _batch(wl, W, H)
    g_l, dt = _image(gl)
    g_r, dt2 = _image(gr)
    if dt2 != dt:
        g_r = g_r.to(g_l.dtype)
    b = _dev(base, torch.int32)
    pr = _dev(prior, torch.float64) if (prior is not None and prior_w > 0.0) else None
    K = 2 * R + 1
    C = torch.empty(B * W * H * K, dtype=torch.float64, device=wl.device)
    L.check(L.lib().sv_band_costs(L.ptr(wl), L.ptr(wr), L.ptr(g_l), L.ptr(g_r), dt, L.ptr(b), B, W,
                                  H, R, float(ad_weight), float(ad_trunc), float(invalid), L.ptr(pr),
                                  float(prior_w), float(prior_t), L.ptr(C), L.stream()))
    return _ret(C, kind)


def _aggregate_path(C, base, W, H, K, P1, P2, dx, dy, S):
    """_aggregate_path, stereonet.py:5192-5241: adds one path's costs into S
    in place (S is a list, a float64 numpy array or a CUDA float64 tensor)."""
    c = _dev(C, torch.float64)
    b = _dev(base, torch.int32)
    B = _batch(b, W, H)
    if isinstance(S, torch.Tensor) and S.is_cuda and S.dtype == torch.float64 and S.is_contiguous():
        s = S
    else:
        s = _dev(S, torch.float64)
    nbytes = ctypes.c_size_t(0)
    L.check(L.lib().sv_aggregate_path_workspace_bytes(B, W, H, K, ctypes.byref(nbytes)))
    ws = L.workspace.get(nbytes.value)
    L.check(L.lib().sv_aggregate_path(L.ptr(c), L.ptr(b), B, W, H, K, float(P1), float(P2), dx, dy,
                                      L.ptr(s), L.ptr(ws), nbytes.value, L.stream()))
    if s is not S:
        res = s.cpu().numpy()
        if isinstance(S, np.ndarray):
            S.reshape(-1)[:] = res
        elif isinstance(S, torch.Tensor):
            S.copy_(torch.from_numpy(res).reshape(S.shape))
        else:
            S[:] = res.tolist()
    return None


def _ws_bytes(mode, B, W, H, R, P1, P2, ad_weight, ad_trunc, paths, rx, ry, prior_w, prior_t):
    n = ctypes.c_size_t(0)
    L.check(L.lib().sv_refine_workspace_bytes(mode, B, W, H, R, float(P1), float(P2),
                                              float(ad_weight), float(ad_trunc), int(paths), rx,
                                              ry, float(prior_w), float(prior_t), ctypes.byref(n)))
    return n.value


def _pair(left, right, disp0, W, H):
    l, dt = _image(left)
    r, dt2 = _image(right)
    if dt2 != dt:
        r = r.to(l.dtype)
    d0 = _dev(disp0, torch.float64)
    B = _batch(l, W, H)
    if r.numel() != l.numel() or d0.numel() != l.numel():
        raise ValueError("left, right and disp0 must have the same size")
    return l, r, dt, d0, B


def refine_fullres(left, right, disp0, W, H, R=3, P1=1.0, P2=12.0,
                   ad_weight=0.0, ad_trunc=20.0, paths=8, uniq=0.0,
                   rx=2, ry=2, prior_w=0.0, prior_t=4.0, return_conf=False):
    """refine_fullres, stereonet.py:5244-5305 (same defaults)."""
    kind, shape = _kind(disp0), _shape_of(disp0)
    l, r, dt, d0, B = _pair(left, right, disp0, W, H)
    out = torch.empty(B * W * H, dtype=torch.float64, device=l.device)
    conf = torch.empty_like(out) if return_conf else None
    nb = _ws_bytes(0, B, W, H, R, P1, P2, ad_weight, ad_trunc, paths, rx, ry, prior_w, prior_t)
    ws = L.workspace.get(nb)
    L.check(L.lib().sv_refine_fullres(L.ptr(l), L.ptr(r), dt, L.ptr(d0), B, W, H, R, float(P1),
                                      float(P2), float(ad_weight), float(ad_trunc), int(paths),
                                      float(uniq), rx, ry, float(prior_w), float(prior_t),
                                      L.ptr(out), L.ptr(conf), L.ptr(ws), nb, L.stream()))
    if return_conf:
        return _ret(out, kind, shape), _ret(conf, kind, shape)
    return _ret(out, kind, shape)


def full_range(left, right, W, H, R, P1=12.0, P2=150.0, rx=4, ry=3, paths=8, lr=False,
               tol=1.0, return_conf=False):
    """Full-range SGM, the control of eval_real.py:143-146: equals
    refine_fullres(left, right, [float(R)] * (W*H), W, H, R=R, P1=P1, P2=P2,
    rx=rx, ry=ry, paths=paths) (and with return_conf its conf), or with
    lr=True lr_refine(..., tol=tol) on that constant centre, returning
    (disparity, status).  R may be one int or one per pair of a batch
    (B, H, W): the pairs then run in one launch when the constant-centre
    kernels apply, else in one call per distinct R.  The centre map is never
    built on the host."""
    kind = _kind(left)
    l, dt = _image(left)
    r, dt2 = _image(right)
    if dt2 != dt:
        r = r.to(l.dtype)
    B = _batch(l, W, H)
    if r.numel() != l.numel():
        raise ValueError("left and right must have the same size")
    shape = _shape_of(left)
    per_pair = not isinstance(R, (int, np.integer))
    Rs = [int(v) for v in R] if per_pair else [int(R)] * B
    if len(Rs) != B or min(Rs) < 0:
        raise ValueError("R must be a non-negative int or one per pair (%d pairs)" % B)
    Rmax = max(Rs)
    mode = 1 if lr else 0
    out = torch.empty(B * W * H, dtype=torch.float64, device=l.device)
    conf = torch.empty_like(out) if (return_conf and not lr) else None
    status = torch.empty(B * W * H, dtype=torch.uint8, device=l.device) if lr else None
    fast = ctypes.c_int(0)
    L.check(L.lib().sv_full_range_fast(mode, W, H, Rmax, float(P1), float(P2), int(paths), rx, ry,
                                       ctypes.byref(fast)))
    uniform = len(set(Rs)) == 1
    if uniform or fast.value:
        rtab = None if uniform else _rtab(Rs, l.device)
        groups = [(list(range(B)), Rmax, rtab)]
    else:
        groups = [([i for i in range(B) if Rs[i] == v], v, None) for v in sorted(set(Rs))]
    hw = W * H
    l2, r2 = l.reshape(B, hw), r.reshape(B, hw)
    o2 = out.reshape(B, hw)
    for idx, Rg, rtab in groups:
        whole = len(idx) == B
        sl, sr = (l2, r2) if whole else (l2[idx].contiguous(), r2[idx].contiguous())
        go = o2 if whole else torch.empty(len(idx), hw, dtype=torch.float64, device=l.device)
        gc = None if conf is None else (conf.reshape(B, hw) if whole else torch.empty_like(go))
        gs = None if status is None else (status.reshape(B, hw) if whole else
                                          torch.empty(len(idx), hw, dtype=torch.uint8,
                                                      device=l.device))
        n = ctypes.c_size_t(0)
        L.check(L.lib().sv_full_range_workspace_bytes(mode, len(idx), W, H, Rg, float(P1),
                                                      float(P2), int(paths), rx, ry,
                                                      ctypes.byref(n)))
        ws = L.workspace.get(n.value)
        L.check(L.lib().sv_full_range(mode, L.ptr(sl), L.ptr(sr), dt, len(idx), W, H, Rg,
                                      L.ptr(rtab), float(P1), float(P2), int(paths), rx, ry,
                                      float(tol), L.ptr(go), L.ptr(gc), L.ptr(gs), L.ptr(ws),
                                      n.value, L.stream()))
        if not whole:
            ix = torch.as_tensor(idx, device=l.device)
            o2[ix] = go
            if gc is not None:
                conf.reshape(B, hw)[ix] = gc
            if gs is not None:
                status.reshape(B, hw)[ix] = gs
    if lr:
        st = status if kind == "torch" else status.to(torch.int32)
        return _ret(out, kind, shape), _ret(st, kind, shape)
    if return_conf:
        return _ret(out, kind, shape), _ret(conf, kind, shape)
    return _ret(out, kind, shape)


_RTABS = {}


def _rtab(Rs, dev):
    """Device int32 table of per-pair centres, cached (a host-to-device copy
    synchronises)."""
    key = (tuple(Rs), str(dev))
    t = _RTABS.get(key)
    if t is None:
        if len(_RTABS) > 256:
            _RTABS.clear()
        t = torch.tensor(Rs, dtype=torch.int32, device=dev)
        _RTABS[key] = t
    return t


def _mirror1(img, W, H):
    """_mirror1, stereonet.py:5308-5314 (horizontal flip)."""
    kind, shape = _kind(img), _shape_of(img)
    t = _dev(img)
    B = _batch(t, W, H)
    return _ret(t.reshape(B, H, W).flip(-1).contiguous().reshape(-1), kind, shape)


def left_to_right(dl, W, H):
    """left_to_right, stereonet.py:5317-5340."""
    kind, shape = _kind(dl), _shape_of(dl)
    d = _dev(dl, torch.float64)
    B = _batch(d, W, H)
    dr = torch.empty_like(d)
    L.check(L.lib().sv_left_to_right(L.ptr(d), B, W, H, L.ptr(dr), L.stream()))
    return _ret(dr, kind, shape)


_LR_KW = ("R", "P1", "P2", "ad_weight", "ad_trunc", "paths", "uniq", "rx", "ry", "prior_w",
          "prior_t")


def lr_refine(left, right, disp0, W, H, tol=1.0, **kw):
    """lr_refine, stereonet.py:5343-5391.  Returns (disparity, status) with
    status 0 ok, 1 occluded, 2 mismatch."""
    bad = set(kw) - set(_LR_KW)
    if bad:
        raise TypeError("refine_fullres() got unexpected keyword argument(s) %s" % sorted(bad))
    p = dict(R=3, P1=1.0, P2=12.0, ad_weight=0.0, ad_trunc=20.0, paths=8, uniq=0.0, rx=2, ry=2,
             prior_w=0.0, prior_t=4.0)
    p.update(kw)
    kind, shape = _kind(disp0), _shape_of(disp0)
    l, r, dt, d0, B = _pair(left, right, disp0, W, H)
    out = torch.empty(B * W * H, dtype=torch.float64, device=l.device)
    status = torch.empty(B * W * H, dtype=torch.uint8, device=l.device)
    nb = _ws_bytes(1, B, W, H, p["R"], p["P1"], p["P2"], p["ad_weight"], p["ad_trunc"],
                   p["paths"], p["rx"], p["ry"], p["prior_w"], p["prior_t"])
    ws = L.workspace.get(nb)
    L.check(L.lib().sv_lr_refine(L.ptr(l), L.ptr(r), dt, L.ptr(d0), B, W, H, float(tol), p["R"],
                                 float(p["P1"]), float(p["P2"]), float(p["ad_weight"]),
                                 float(p["ad_trunc"]), int(p["paths"]), float(p["uniq"]), p["rx"],
                                 p["ry"], float(p["prior_w"]), float(p["prior_t"]), L.ptr(out),
                                 L.ptr(status), L.ptr(ws), nb, L.stream()))
    st = status if kind == "torch" else status.to(torch.int32)
    return _ret(out, kind, shape), _ret(st, kind, shape)


def median_filter(disp, W, H, radius=1, passes=1):
    """median_filter, stereonet.py:2231-2258 (radius 0..2 on the GPU)."""
    kind, shape = _kind(disp), _shape_of(disp)
    d = _dev(disp, torch.float64)
    B = _batch(d, W, H)
    out = torch.empty_like(d)
    tmp = torch.empty_like(d) if passes > 1 else None
    L.check(L.lib().sv_median_filter(L.ptr(d), B, W, H, radius, passes, L.ptr(out), L.ptr(tmp),
                                     L.stream()))
    return _ret(out, kind, shape)


_PS = ctypes.c_size_t


def _rgb(rgb, W, H):
    """RGB as [B][H][W][3] on the device: uint8 (read as value * (1/255)) or
    float64 in [0, 1] (lists: the reference's interleaved floats)."""
    t = _dev(rgb)
    if t.dtype == torch.float32:
        t = t.to(torch.float64)
    n = t.numel()
    if n == 0 or n % (3 * W * H):
        raise ValueError("rgb of %d values does not hold W x H = %d x %d RGB images" % (n, W, H))
    return t, _DTYPES[t.dtype], n // (3 * W * H)


def _ps_ws(B, W, H, step):
    nb = _PS(0)
    L.check(L.lib().sv_plane_snap_workspace_bytes(B, W, H, step, ctypes.byref(nb)))
    return L.workspace.get(nb.value), nb.value


def _ps_check(flag):
    if int(flag.item()) != 0:
        raise RuntimeError("plane_snap: a segment had more than 256 pixels outside every SLIC "
                           "window (outside what the GPU kernels support)")


def slic(rgb, W, H, step=12, compactness=10.0, iters=5):
    """slic, stereonet.py:5394-5448: (labels, number of centres); labels -1
    where no centre ever reached."""
    kind = _kind(rgb)
    t, dt, B = _rgb(rgb, W, H)
    lab = torch.empty(B * W * H, dtype=torch.int32, device=t.device)
    flag = torch.zeros(1, dtype=torch.int32, device=t.device)
    k = ctypes.c_int(0)
    ws, nb = _ps_ws(B, W, H, step)
    L.check(L.lib().sv_slic(L.ptr(t), dt, B, W, H, step, float(compactness), iters, L.ptr(lab),
                            ctypes.byref(k), L.ptr(flag), L.ptr(ws), nb, L.stream()))
    _ps_check(flag)
    shape = None if kind == "list" else (tuple(rgb.shape[:-1]) if rgb.shape[-1] == 3 and rgb.ndim > 1
                                         else None)
    return _ret(lab, kind, shape), k.value


def _snap_dev(d, t, dt, B, W, H, valid, vkind, step, thr, min_inl, trials, seed):
    out = torch.empty_like(d)
    flag = torch.zeros(1, dtype=torch.int32, device=d.device)
    ws, nb = _ps_ws(B, W, H, step)
    L.check(L.lib().sv_plane_snap(L.ptr(d), L.ptr(t), dt, L.ptr(valid), vkind, B, W, H, step,
                                  float(thr), float(min_inl), int(trials), int(seed) & 0xFFFFFFFF,
                                  L.ptr(out), L.ptr(flag), L.ptr(ws), nb, L.stream()))
    _ps_check(flag)
    return out


def plane_snap(disp, rgb, W, H, valid=None, step=12, thr=1.0, min_inl=0.6, trials=40,
               seed=12345):
    """plane_snap, stereonet.py:5498-5563: a slanted plane per SLIC segment
    (RANSAC from the reference's xorshift32 stream, least-squares refit on
    the inliers); pixels on it, and the segment's invalid pixels, take the
    plane's value (floored at 0)."""
    kind, shape = _kind(disp), _shape_of(disp)
    d = _dev(disp, torch.float64)
    B = _batch(d, W, H)
    t, dt, Bc = _rgb(rgb, W, H)
    if Bc != B:
        raise ValueError("%d disparity maps and %d RGB images" % (B, Bc))
    v = None
    if valid is not None:
        v = valid if isinstance(valid, torch.Tensor) else torch.from_numpy(
            np.ascontiguousarray(np.asarray(valid, dtype=bool)))
        v = v.to(d.device).reshape(-1).to(torch.uint8).contiguous()
        if v.numel() != d.numel():
            raise ValueError("%d validity flags for %d disparities" % (v.numel(), d.numel()))
    out = _snap_dev(d, t, dt, B, W, H, v, 0, step, thr, min_inl, trials, seed)
    return _ret(out, kind, shape)


def refine_disparity(left, right, rgb, disp0, W, H, R=10, P1=12.0, P2=150.0,
                     rx=4, ry=3, tol=2.0, planes=True, plane_step=12,
                     plane_thr=1.0, plane_min_inl=0.4, verbose=False):
    """refine_disparity, stereonet.py:5865-5884: lr_refine (both views), then
    (planes=True with an RGB image) plane_snap on the pixels whose status is
    0, then one 3x3 median."""
    if planes and rgb is not None:
        kind, shape = _kind(disp0), _shape_of(disp0)
        d, st = lr_refine(left, right, _dev(disp0, torch.float64), W, H, tol=tol, R=R, P1=P1,
                          P2=P2, rx=rx, ry=ry)
        B = _batch(d, W, H)
        if verbose:
            n = float(W * H * B)
            print("  full-res refinement: %.1f%% occluded, %.1f%% mismatched"
                  % (100.0 * int((st == 1).sum()) / n, 100.0 * int((st == 2).sum()) / n))
        t, dt, Bc = _rgb(rgb, W, H)
        if Bc != B:
            raise ValueError("%d disparity maps and %d RGB images" % (B, Bc))
        sn = _snap_dev(d.reshape(-1).contiguous(), t, dt, B, W, H, st.reshape(-1).contiguous(), 1,
                       plane_step, plane_thr, plane_min_inl, 40, 12345)
        out = median_filter(sn, W, H, radius=1, passes=1)
        stt = st if kind == "torch" else st.to(torch.int32)
        return _ret(out, kind, shape), _ret(stt.reshape(-1), kind, shape)
    kind, shape = _kind(disp0), _shape_of(disp0)
    l, r, dt, d0, B = _pair(left, right, disp0, W, H)
    out = torch.empty(B * W * H, dtype=torch.float64, device=l.device)
    status = torch.empty(B * W * H, dtype=torch.uint8, device=l.device)
    nb = _ws_bytes(2, B, W, H, R, P1, P2, 0.0, 20.0, 8, rx, ry, 0.0, 4.0)
    ws = L.workspace.get(nb)
    L.check(L.lib().sv_refine_disparity(L.ptr(l), L.ptr(r), dt, L.ptr(d0), B, W, H, R, float(P1),
                                        float(P2), rx, ry, float(tol), L.ptr(out), L.ptr(status),
                                        L.ptr(ws), nb, L.stream()))
    if verbose:
        n = float(W * H * B)
        print("  full-res refinement: %.1f%% occluded, %.1f%% mismatched"
              % (100.0 * int((status == 1).sum()) / n, 100.0 * int((status == 2).sum()) / n))
    st = status if kind == "torch" else status.to(torch.int32)
    return _ret(out, kind, shape), _ret(st, kind, shape)
