"""CPU oracle for the dense-disparity path -- TEST INFRASTRUCTURE ONLY.

A float64 C restatement of the reference (stereonet.py) lives in
``stereo_oracle.c``; this module builds it with gcc and exposes numpy
wrappers.  Image decoding (read_image: Python's zlib for inflate), the
Middlebury scoring of eval_real.py and the disparity-to-3D functions
(Camera arithmetic, background_region, grid_mesh) are restated in numpy and
plain Python at the end of this module; each is pinned by golden vectors
the reference itself wrote (tests/test_images.py, test_eval_harness.py,
test_geometry.py).  Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
leg may import this package, and only as the checker or the timed CPU
baseline; the shipped library never calls it.

Parity status of each function (which check pins it) is listed in the C
file's header and in the tests (tests/test_oracle.py).
"""

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "stereo_oracle.c")
LIB = os.path.join(HERE, "libstereo_oracle.so")

_lib = None


def build(force=True):
    """Compile the oracle with gcc (binary64, no contraction)."""
    if not force and os.path.exists(LIB) and os.path.getmtime(LIB) >= os.path.getmtime(SRC):
        return LIB
    cmd = ["gcc", "-O2", "-ffp-contract=off", "-fno-fast-math", "-fPIC",
           "-shared", "-o", LIB + ".tmp", SRC, "-lm"]
    subprocess.check_call(cmd)
    os.replace(LIB + ".tmp", LIB)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB) or os.path.getmtime(LIB) < os.path.getmtime(SRC):
            build()
        _lib = ctypes.CDLL(LIB)
        _declare(_lib)
    return _lib


P = ctypes.c_void_p
I = ctypes.c_int
L_ = ctypes.c_long
D_ = ctypes.c_double


def _declare(l):
    sig = {
        "or_census": (I, [P, I, I, I, I, P]),
        "or_band_costs": (None, [P, P, P, P, P, I, I, I, D_, D_, D_, P, D_, D_, P]),
        "or_aggregate_path": (None, [P, P, I, I, I, D_, D_, I, I, P]),
        "or_refine_fullres": (I, [P, P, P, I, I, I, D_, D_, D_, D_, I, D_, I, I, D_, D_, P, P]),
        "or_left_to_right": (None, [P, I, I, P]),
        "or_lr_refine": (I, [P, P, P, I, I, D_, I, D_, D_, D_, D_, I, D_, I, I, D_, D_, P, P]),
        "or_median_filter": (None, [P, I, I, I, I, P]),
        "or_refine_disparity_plain": (I, [P, P, P, I, I, I, D_, D_, I, I, D_, P, P]),
        "or_conv2d_fwd": (None, [P, I, I, I, P, P, I, I, I, I, I, P]),
        "or_conv2d_bwd": (None, [P, I, I, I, P, I, I, I, I, I, P, P, P, P]),
        "or_conv3d_fwd": (None, [P, I, I, I, I, P, P, I, I, P]),
        "or_conv3d_bwd": (None, [P, I, I, I, I, P, I, I, P, P, P, P]),
        "or_costvol_fwd": (None, [P, P, I, I, I, I, P]),
        "or_costvol_bwd": (None, [P, P, I, I, I, I, P, P, P]),
        "or_softargmin_fwd": (None, [P, I, I, I, P, P, P]),
        "or_softargmin_bwd": (None, [P, P, I, I, I, P, P]),
        "or_bilinear_fwd": (None, [P, I, I, I, I, I, P]),
        "or_bilinear_bwd": (None, [P, I, I, I, I, I, P]),
        "or_avgpool2": (None, [P, I, I, I, P]),
        "or_pseudo_huber": (D_, [P, P, P, L_, D_, P, P]),
        "or_disparity_ce": (D_, [P, P, P, I, I, I, D_, P, P]),
        "or_adam_step": (None, [P, P, P, P, L_, D_, D_, D_, D_, D_, D_, D_]),
        "or_net_num_params": (L_, [I, I, I, I, I, P]),
        "or_net_forward": (None, [I, I, I, I, I, P, P, P, P, I, I, P, P, P, P, P, P, P]),
        "or_train_grads": (D_, [I, I, I, I, I, P, P, P, P, P, P, I, I, P, P]),
        "or_predict": (None, [I, I, I, I, I, P, P, P, P, I, I, P]),
        "or_peak": (D_, [P, L_, D_]),
        "or_predict_at_scale": (None, [I, I, I, I, I, P, P, P, P, I, I, I, P, P, P]),
        "or_predict_auto": (None, [I, I, I, I, I, P, P, D_, P, P, I, I, I, I, I, P, P, P, P]),
        "or_mirror_h": (None, [P, I, I, I, P]),
        "or_lr_consistency": (None, [P, P, I, I, D_, P]),
        "or_fill_invalid": (None, [P, P, I, I, P]),
        "or_slic": (I, [P, I, I, I, D_, I, P]),
        "or_plane_snap": (L_, [P, P, I, I, P, I, D_, D_, I, ctypes.c_uint32, P]),
    }
    for name, (res, args) in sig.items():
        f = getattr(l, name)
        f.restype = res
        f.argtypes = args


def _f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


# ----------------------------------------------------------------- refinement

def census(img, W, H, rx=2, ry=2):
    img = _f64(img).ravel()
    out = np.zeros(W * H, dtype=np.uint64)
    if lib().or_census(_p(img), W, H, rx, ry, _p(out)) != 0:
        raise ValueError("census window wider than 64 bits")
    return out


def band_costs(cl, cr, gl, gr, base, W, H, R, ad_weight, ad_trunc, invalid,
               prior=None, prior_w=0.0, prior_t=4.0):
    cl = np.ascontiguousarray(cl, dtype=np.uint64)
    cr = np.ascontiguousarray(cr, dtype=np.uint64)
    gl, gr = _f64(gl).ravel(), _f64(gr).ravel()
    base = np.ascontiguousarray(base, dtype=np.int32)
    prior = _f64(prior).ravel() if prior is not None else np.zeros(W * H)
    C = np.zeros(W * H * (2 * R + 1))
    lib().or_band_costs(_p(cl), _p(cr), _p(gl), _p(gr), _p(base), W, H, R,
                        ad_weight, ad_trunc, invalid, _p(prior), prior_w,
                        prior_t, _p(C))
    return C


def aggregate_path(C, base, W, H, K, P1, P2, dx, dy, S=None):
    C = _f64(C).ravel()
    base = np.ascontiguousarray(base, dtype=np.int32)
    S = np.zeros(W * H * K) if S is None else S
    lib().or_aggregate_path(_p(C), _p(base), W, H, K, P1, P2, dx, dy, _p(S))
    return S


def refine_fullres(left, right, disp0, W, H, R=3, P1=1.0, P2=12.0,
                   ad_weight=0.0, ad_trunc=20.0, paths=8, uniq=0.0, rx=2,
                   ry=2, prior_w=0.0, prior_t=4.0, return_conf=False):
    left, right, disp0 = _f64(left).ravel(), _f64(right).ravel(), _f64(disp0).ravel()
    out = np.zeros(W * H)
    conf = np.zeros(W * H)
    rc = lib().or_refine_fullres(_p(left), _p(right), _p(disp0), W, H, R, P1,
                                 P2, ad_weight, ad_trunc, paths, uniq, rx, ry,
                                 prior_w, prior_t, _p(out), _p(conf))
    if rc != 0:
        raise ValueError("census window wider than 64 bits")
    return (out, conf) if return_conf else out


def left_to_right(dl, W, H):
    dl = _f64(dl).ravel()
    dr = np.zeros(W * H)
    lib().or_left_to_right(_p(dl), W, H, _p(dr))
    return dr


def lr_refine(left, right, disp0, W, H, tol=1.0, R=3, P1=1.0, P2=12.0,
              ad_weight=0.0, ad_trunc=20.0, paths=8, uniq=0.0, rx=2, ry=2,
              prior_w=0.0, prior_t=4.0):
    left, right, disp0 = _f64(left).ravel(), _f64(right).ravel(), _f64(disp0).ravel()
    out = np.zeros(W * H)
    status = np.zeros(W * H, dtype=np.int32)
    rc = lib().or_lr_refine(_p(left), _p(right), _p(disp0), W, H, tol, R, P1,
                            P2, ad_weight, ad_trunc, paths, uniq, rx, ry,
                            prior_w, prior_t, _p(out), _p(status))
    if rc != 0:
        raise ValueError("census window wider than 64 bits")
    return out, status


def median_filter(disp, W, H, radius=1, passes=1):
    disp = _f64(disp).ravel()
    out = np.zeros(W * H)
    lib().or_median_filter(_p(disp), W, H, radius, passes, _p(out))
    return out


def refine_disparity_plain(left, right, disp0, W, H, R=10, P1=12.0, P2=150.0,
                           rx=4, ry=3, tol=2.0):
    left, right, disp0 = _f64(left).ravel(), _f64(right).ravel(), _f64(disp0).ravel()
    out = np.zeros(W * H)
    status = np.zeros(W * H, dtype=np.int32)
    rc = lib().or_refine_disparity_plain(_p(left), _p(right), _p(disp0), W, H,
                                         R, P1, P2, rx, ry, tol, _p(out),
                                         _p(status))
    if rc != 0:
        raise ValueError("census window wider than 64 bits")
    return out, status


# -------------------------------------------------------------------- layers

def conv_out(n, k, s, p, d):
    return (n + 2 * p - (d * (k - 1) + 1)) // s + 1


def conv2d_fwd(x, IC, H, W, w, b, OC, k, stride=1, pad=None, dil=1):
    pad = (dil * (k - 1)) // 2 if pad is None else pad
    x, w, b = _f64(x).ravel(), _f64(w).ravel(), _f64(b).ravel()
    OH, OW = conv_out(H, k, stride, pad, dil), conv_out(W, k, stride, pad, dil)
    y = np.zeros(OC * OH * OW)
    lib().or_conv2d_fwd(_p(x), IC, H, W, _p(w), _p(b), OC, k, stride, pad, dil, _p(y))
    return y, OH, OW


def conv2d_bwd(x, IC, H, W, w, OC, k, gy, stride=1, pad=None, dil=1):
    pad = (dil * (k - 1)) // 2 if pad is None else pad
    x, w, gy = _f64(x).ravel(), _f64(w).ravel(), _f64(gy).ravel()
    gx = np.zeros(IC * H * W)
    gw = np.zeros(w.size)
    gb = np.zeros(OC)
    lib().or_conv2d_bwd(_p(x), IC, H, W, _p(w), OC, k, stride, pad, dil,
                        _p(gy), _p(gx), _p(gw), _p(gb))
    return gx, gw, gb


def conv3d_fwd(x, IC, D, H, W, w, b, OC, k=3):
    x, w, b = _f64(x).ravel(), _f64(w).ravel(), _f64(b).ravel()
    y = np.zeros(OC * D * H * W)
    lib().or_conv3d_fwd(_p(x), IC, D, H, W, _p(w), _p(b), OC, k, _p(y))
    return y


def conv3d_bwd(x, IC, D, H, W, w, OC, gy, k=3):
    x, w, gy = _f64(x).ravel(), _f64(w).ravel(), _f64(gy).ravel()
    gx = np.zeros(IC * D * H * W)
    gw = np.zeros(w.size)
    gb = np.zeros(OC)
    lib().or_conv3d_bwd(_p(x), IC, D, H, W, _p(w), OC, k, _p(gy), _p(gx), _p(gw), _p(gb))
    return gx, gw, gb


def costvol_fwd(fl, fr, C, D, H, W):
    fl, fr = _f64(fl).ravel(), _f64(fr).ravel()
    vol = np.zeros(C * D * H * W)
    lib().or_costvol_fwd(_p(fl), _p(fr), C, D, H, W, _p(vol))
    return vol


def costvol_bwd(fl, fr, C, D, H, W, gvol):
    fl, fr, gvol = _f64(fl).ravel(), _f64(fr).ravel(), _f64(gvol).ravel()
    gfl, gfr = np.zeros(C * H * W), np.zeros(C * H * W)
    lib().or_costvol_bwd(_p(fl), _p(fr), C, D, H, W, _p(gvol), _p(gfl), _p(gfr))
    return gfl, gfr


def softargmin_fwd(cost, D, H, W, penalty=None):
    cost = _f64(cost).ravel()
    pen = None if penalty is None else _f64(penalty).ravel()
    disp, probs = np.zeros(H * W), np.zeros(D * H * W)
    lib().or_softargmin_fwd(_p(cost), D, H, W, _p(pen), _p(disp), _p(probs))
    return disp, probs


def softargmin_bwd(probs, disp, D, H, W, gdisp):
    probs, disp, gdisp = _f64(probs).ravel(), _f64(disp).ravel(), _f64(gdisp).ravel()
    g = np.zeros(D * H * W)
    lib().or_softargmin_bwd(_p(probs), _p(disp), D, H, W, _p(gdisp), _p(g))
    return g


def bilinear_fwd(x, C, H, W, H2, W2):
    x = _f64(x).ravel()
    y = np.zeros(C * H2 * W2)
    lib().or_bilinear_fwd(_p(x), C, H, W, H2, W2, _p(y))
    return y


def bilinear_bwd(gy, C, H, W, H2, W2):
    gy = _f64(gy).ravel()
    gx = np.zeros(C * H * W)
    lib().or_bilinear_bwd(_p(gy), C, H, W, H2, W2, _p(gx))
    return gx


def avgpool2(img, C, H, W):
    img = _f64(img).ravel()
    out = np.zeros(C * (H // 2) * (W // 2))
    lib().or_avgpool2(_p(img), C, H, W, _p(out))
    return out, H // 2, W // 2


def _u8(v):
    return None if v is None else np.ascontiguousarray(np.asarray(v).ravel().astype(np.uint8))


def pseudo_huber(pred, target, valid, alpha=1.0):
    pred, target = _f64(pred).ravel(), _f64(target).ravel()
    v = _u8(valid)
    grad = np.zeros(pred.size)
    cnt = ctypes.c_long(0)
    loss = lib().or_pseudo_huber(_p(pred), _p(target), _p(v), pred.size, alpha,
                                 _p(grad), ctypes.byref(cnt))
    return loss, grad, cnt.value


def disparity_ce(probs, target, valid, D, H, W, sigma=0.8):
    probs, target = _f64(probs).ravel(), _f64(target).ravel()
    v = _u8(valid)
    g = np.zeros(D * H * W)
    cnt = ctypes.c_long(0)
    loss = lib().or_disparity_ce(_p(probs), _p(target), _p(v), D, H, W, sigma,
                                 _p(g), ctypes.byref(cnt))
    return loss, g, cnt.value


class Adam:
    """Adam.step of stereonet.py:1368-1387 over one flat float64 vector."""

    def __init__(self, params, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.p = params
        self.m = np.zeros_like(params)
        self.v = np.zeros_like(params)
        self.lr, self.b1, self.b2, self.eps = lr, b1, b2, eps
        self.b1t = 1.0
        self.b2t = 1.0

    def step(self, grads, clip=5.0):
        self.b1t *= self.b1
        self.b2t *= self.b2
        g = _f64(grads)
        lib().or_adam_step(_p(self.p), _p(g), _p(self.m), _p(self.v), self.p.size,
                           self.lr, self.b1, self.b2, self.eps, self.b1t, self.b2t, clip)


# ------------------------------------------------------------------- network

class Arch:
    def __init__(self, feat_C=12, cost_C=6, D=12, refine_C=8, dil=(1, 2, 4),
                 target_peak=34.0):
        self.feat_C, self.cost_C, self.D, self.refine_C = feat_C, cost_C, D, refine_C
        self.dil = np.array(dil, dtype=np.int32)
        self.target_peak = target_peak

    def args(self):
        return (self.feat_C, self.cost_C, self.D, self.refine_C, len(self.dil),
                _p(self.dil))

    def num_params(self):
        return lib().or_net_num_params(*self.args())


def net_forward(arch, params, left, right, H, W, guide_half, guide_full):
    params = _f64(params)
    left, right = _f64(left).ravel(), _f64(right).ravel()
    gh, gf = _f64(guide_half).ravel(), _f64(guide_full).ravel()
    h = conv_out(conv_out(H, 3, 2, 1, 1), 3, 2, 1, 1)
    w = conv_out(conv_out(W, 3, 2, 1, 1), 3, 2, 1, 1)
    d0 = np.zeros(h * w)
    d1 = np.zeros((H // 2) * (W // 2))
    d2 = np.zeros(H * W)
    probs = np.zeros(arch.D * h * w)
    dims = np.zeros(2, dtype=np.int32)
    lib().or_net_forward(*arch.args(), _p(params), _p(left), _p(right), H, W,
                         _p(gh), _p(gf), _p(d0), _p(d1), _p(d2), _p(probs), _p(dims))
    return d0, d1, d2, probs, (h, w)


def train_grads(arch, params, left, right, disp, valid, W, H):
    params = _f64(params)
    left, right, disp = _f64(left).ravel(), _f64(right).ravel(), _f64(disp).ravel()
    v = _u8(valid)
    G = np.zeros(params.size)
    d2 = np.zeros(W * H)
    loss = lib().or_train_grads(*arch.args(), _p(params), _p(left), _p(right),
                                _p(disp), _p(v), W, H, _p(G), _p(d2))
    return loss, G, d2


def predict(arch, params, left, right, W, H):
    params = _f64(params)
    left, right = _f64(left).ravel(), _f64(right).ravel()
    out = np.zeros(W * H)
    lib().or_predict(*arch.args(), _p(params), _p(left), _p(right), W, H, _p(out))
    return out


def peak(disp, q=0.95):
    d = _f64(disp).ravel()
    return lib().or_peak(_p(d), d.size, q)


def predict_at_scale(arch, params, left, right, W, H, work_width=96):
    params = _f64(params)
    left, right = _f64(left).ravel(), _f64(right).ravel()
    full = np.zeros(W * H)
    work = np.zeros(W * H)
    dims = np.zeros(2, dtype=np.int32)
    lib().or_predict_at_scale(*arch.args(), _p(params), _p(left), _p(right), W, H,
                              work_width, _p(full), _p(work), _p(dims))
    Ww, Hw = int(dims[0]), int(dims[1])
    return full, work[:Ww * Hw].copy(), Ww, Hw


def predict_auto(arch, params, left, right, W, H, target_peak=None,
                 probe_width=64, max_width=320, passes=3):
    tp = arch.target_peak if target_peak is None else target_peak
    params = _f64(params)
    left, right = _f64(left).ravel(), _f64(right).ravel()
    full = np.zeros(W * H)
    work = np.zeros(W * H)
    dims = np.zeros(2, dtype=np.int32)
    trace = np.zeros(8, dtype=np.int32)
    lib().or_predict_auto(*arch.args(), _p(params), float(tp), _p(left), _p(right),
                          W, H, probe_width, max_width, passes, _p(full), _p(work),
                          _p(dims), _p(trace))
    Ww, Hw = int(dims[0]), int(dims[1])
    return full, work[:Ww * Hw].copy(), Ww, Hw, trace


def load_weights_txt(path):
    """Parse the text checkpoint format of stereonet.py:1766-1874 into
    (config dict or None, list of float64 blocks)."""
    lines = open(path).read().split("\n")
    cfg = None
    off = 0
    if lines[0].startswith("stereonet "):
        cfg = {}
        for tok in lines[0].split()[1:]:
            if "=" not in tok:
                continue
            k, v = tok.split("=", 1)
            cfg[k] = tuple(int(d) for d in v.split(",")) if k == "refine_dil" else int(v)
        off = 1
    n = int(lines[off])
    blocks = []
    for k in range(n):
        parts = lines[off + 1 + k].split()
        cnt = int(parts[0])
        blocks.append(np.array([float(v) for v in parts[1:cnt + 1]]))
    return cfg, blocks


# ------------------------------------------------ network LR consistency

def mirror_h(img, C, W, H):
    """mirror_h, stereonet.py:4974-4983."""
    img = _f64(img).ravel()
    out = np.zeros(C * W * H)
    lib().or_mirror_h(_p(img), C, W, H, _p(out))
    return out


def lr_consistency(disp_l, disp_r, W, H, tol=1.5):
    """lr_consistency, stereonet.py:4986-5008 (bool array)."""
    dl, dr = _f64(disp_l).ravel(), _f64(disp_r).ravel()
    v = np.zeros(W * H, dtype=np.uint8)
    lib().or_lr_consistency(_p(dl), _p(dr), W, H, float(tol), _p(v))
    return v.astype(bool)


def fill_invalid(disp, valid, W, H):
    """fill_invalid, stereonet.py:5011-5046."""
    d = _f64(disp).ravel()
    v = np.ascontiguousarray(np.asarray(valid, dtype=bool).ravel()).astype(np.uint8)
    out = np.zeros(W * H)
    lib().or_fill_invalid(_p(d), _p(v), W, H, _p(out))
    return out


def predict_consistent(arch, params, left, right, W, H, work_width=None, tol=1.5):
    """predict_consistent, stereonet.py:5049-5076: (disp, valid fraction),
    plus the validity mask for the tests."""
    if work_width is None:
        d_l, _, Ww, _, _ = predict_auto(arch, params, left, right, W, H)
    else:
        d_l, _, Ww, _ = predict_at_scale(arch, params, left, right, W, H, work_width)
    ml, mr = mirror_h(left, 1, W, H), mirror_h(right, 1, W, H)
    d_rm, _, _, _ = predict_at_scale(arch, params, mr, ml, W, H, Ww)
    d_r = mirror_h(d_rm, 1, W, H)
    valid = lr_consistency(d_l, d_r, W, H, tol)
    frac = int(valid.sum()) / float(W * H)
    return fill_invalid(d_l, valid, W, H), frac, valid


# ------------------------------------------------------------ planes

def slic(rgb, W, H, step=12, compactness=10.0, iters=5):
    """slic, stereonet.py:5394-5448: (labels int32, number of centres)."""
    rgb = _f64(rgb).ravel()
    lab = np.zeros(W * H, dtype=np.int32)
    k = lib().or_slic(_p(rgb), W, H, step, float(compactness), iters, _p(lab))
    return lab, int(k)


def plane_snap(disp, rgb, W, H, valid=None, step=12, thr=1.0, min_inl=0.6, trials=40,
               seed=12345):
    """plane_snap, stereonet.py:5498-5563."""
    d, rgb = _f64(disp).ravel(), _f64(rgb).ravel()
    v = None if valid is None else np.ascontiguousarray(np.asarray(valid, dtype=bool).ravel()).astype(np.uint8)
    out = np.zeros(W * H)
    lib().or_plane_snap(_p(d), _p(rgb), W, H, _p(v), step, float(thr), float(min_inl), trials,
                        seed, _p(out))
    return out


def refine_disparity_planes(left, right, rgb, disp0, W, H, R=10, P1=12.0, P2=150.0, rx=4, ry=3,
                            tol=2.0, plane_step=12, plane_thr=1.0, plane_min_inl=0.4):
    """refine_disparity with planes=True and an RGB image, stereonet.py:5865-5884:
    lr_refine, plane_snap on the status-0 pixels, one 3x3 median."""
    d, st = lr_refine(left, right, disp0, W, H, tol=tol, R=R, P1=P1, P2=P2, rx=rx, ry=ry)
    d = plane_snap(d, rgb, W, H, valid=np.asarray(st) == 0, step=plane_step, thr=plane_thr,
                   min_inl=plane_min_inl)
    return median_filter(d, W, H, 1, 1), np.asarray(st)


# ------------------------------------------------------------ image input
# Pinned against the reference's own read_image outputs on 73 generated
# PNG / Netpbm files (tests/golden/images.npz).  Inflate is Python's zlib
# (an independent DEFLATE implementation); the rest is restated here.

PNG_SIG = b"\x89PNG\r\n\x1a\n"


def _paeth(a, b, c):
    """stereonet.py:3436-3453."""
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    if pa <= pb and pa <= pc:
        return a
    return b if pb <= pc else c


def _unfilter(raw, H, bpp, stride):
    """stereonet.py:3456-3496, one row at a time; Sub as a per-lane cumulative
    sum mod 256, Up vectorised, Average and Paeth as loops."""
    out = np.zeros((H, stride), dtype=np.int64)
    raw = np.frombuffer(raw, dtype=np.uint8)
    for y in range(H):
        f = int(raw[y * (stride + 1)])
        row = raw[y * (stride + 1) + 1:(y + 1) * (stride + 1)].astype(np.int64)
        up = out[y - 1] if y > 0 else np.zeros(stride, dtype=np.int64)
        cur = out[y]
        if f == 0:
            cur[:] = row
        elif f == 1:
            for lane in range(min(bpp, stride)):
                cur[lane::bpp] = np.cumsum(row[lane::bpp]) & 0xFF
        elif f == 2:
            cur[:] = (row + up) & 0xFF
        elif f in (3, 4):
            r, u = row.tolist(), up.tolist()
            o = [0] * stride
            for i in range(stride):
                a = o[i - bpp] if i >= bpp else 0
                if f == 3:
                    o[i] = (r[i] + ((a + u[i]) >> 1)) & 0xFF
                else:
                    c = u[i - bpp] if i >= bpp else 0
                    o[i] = (r[i] + _paeth(a, u[i], c)) & 0xFF
            cur[:] = o
        else:
            raise ValueError("unknown PNG filter type %d on row %d" % (f, y))
    return out


def _planes(r, g, b):
    return (0.299 * r + 0.587 * g + 0.114 * b,
            np.stack([r, g, b], axis=-1).reshape(-1))


def read_png_bytes(data):
    """read_png, stereonet.py:3499-3601: (gray, rgb, W, H) as float64 numpy."""
    import zlib
    if data[0:8] != PNG_SIG:
        raise ValueError("not a PNG file")
    pos, W, H, depth, ctype, interlace, pal, idat = 8, 0, 0, 0, 0, 0, None, b""
    while pos < len(data):
        n = int.from_bytes(data[pos:pos + 4], "big")
        t, body = data[pos + 4:pos + 8], data[pos + 8:pos + 8 + n]
        pos += 12 + n
        if t == b"IHDR":
            W, H = int.from_bytes(body[0:4], "big"), int.from_bytes(body[4:8], "big")
            depth, ctype, interlace = body[8], body[9], body[12]
        elif t == b"PLTE":
            pal = np.frombuffer(body, dtype=np.uint8)
        elif t == b"IDAT":
            idat += body
        elif t == b"IEND":
            break
    if interlace:
        raise ValueError("Adam7 interlaced PNG is not supported")
    if depth not in (1, 2, 4, 8, 16):
        raise ValueError("unsupported PNG bit depth %d" % depth)
    ch = {0: 1, 2: 3, 3: 1, 4: 2, 6: 4}.get(ctype)
    if ch is None:
        raise ValueError("unsupported PNG colour type %d" % ctype)
    cmf, flg = idat[0], idat[1]                      # zlib_decompress :3421-3433
    if (cmf & 0x0F) != 8 or ((cmf << 8) | flg) % 31 or flg & 0x20:
        raise ValueError("bad zlib header")
    raw = zlib.decompressobj(-15).decompress(idat[2:])
    stride = (W * ch * depth + 7) // 8
    img = _unfilter(raw, H, max(1, ch * depth // 8), stride)
    if depth == 8:
        s = img[:, :W * ch].reshape(H, W, ch)
    elif depth == 16:
        s = (img[:, 0:2 * W * ch:2] << 8 | img[:, 1:2 * W * ch:2]).reshape(H, W, ch)
    else:
        bit = np.arange(W) * depth
        shift = 8 - depth - (bit & 7)
        s = ((img[:, bit >> 3] >> shift) & ((1 << depth) - 1)).reshape(H, W, 1)
    if ctype == 3:
        if pal is None:
            raise ValueError("indexed PNG with no PLTE chunk")
        p = 3 * s[..., 0].reshape(-1)
        r, g, b = pal[p] / 255.0, pal[p + 1] / 255.0, pal[p + 2] / 255.0
    else:
        inv = 1.0 / 65535.0 if depth == 16 else 1.0 / ((1 << depth) - 1)
        v = s.reshape(-1, ch) * inv
        r, g, b = (v[:, 0], v[:, 0], v[:, 0]) if ctype in (0, 4) else (v[:, 0], v[:, 1], v[:, 2])
    gray, rgb = _planes(r, g, b)
    return gray, rgb, W, H


def read_pnm_bytes(data):
    """read_pnm, stereonet.py:244-324 (tokens, comments, one whitespace byte
    after maxval, 8- or 16-bit binary samples, ASCII samples)."""
    magic = data[0:2]
    if magic not in (b"P2", b"P3", b"P5", b"P6"):
        raise ValueError("not a PGM/PPM file")
    toks, i = [], 2
    while len(toks) < 3 and i < len(data):
        c = data[i]
        if c == 35:
            while i < len(data) and data[i] not in (10, 13):
                i += 1
        elif c in (32, 9, 10, 13):
            i += 1
        else:
            j = i
            while j < len(data) and data[j] not in (32, 9, 10, 13):
                j += 1
            toks.append(int(data[i:j]))
            i = j
    W, H, maxval = toks
    ch = 3 if magic in (b"P3", b"P6") else 1
    count = W * H * ch
    if magic in (b"P2", b"P3"):
        v = np.array([int(t) for t in data[i:].split()[:count]], dtype=np.int64)
    elif maxval < 256:
        v = np.frombuffer(data, dtype=np.uint8, count=count, offset=i + 1).astype(np.int64)
    else:
        v = np.frombuffer(data, dtype=">u2", count=count, offset=i + 1).astype(np.int64)
    s = v * (1.0 / maxval)
    if ch == 3:
        s3 = s.reshape(-1, 3)
        return _planes(s3[:, 0], s3[:, 1], s3[:, 2])[0], s, W, H
    return s, np.repeat(s, 3), W, H


def read_image_bytes(data):
    """read_image, stereonet.py:3604-3613: PNG by its signature, else Netpbm."""
    if data[0:8] == PNG_SIG:
        return read_png_bytes(data)
    return read_pnm_bytes(data)


# ------------------------------------------------------ Middlebury scoring
# Pinned against the reference harness's own load_gt / masks / score /
# constant outputs (tests/golden/eval.npz).

def gt_from_gray(g, scale):
    """load_gt, eval_real.py:78-87, from the decoded gray plane."""
    lv = np.asarray(g, dtype=np.float64) * 255.0
    known = lv >= 0.5
    return np.where(known, lv / scale, 0.0), known


def nonocc(dl, kl, dr, kr, W, H):
    """masks, eval_real.py:90-102 (int() truncates toward zero)."""
    dl, dr = dl.reshape(H, W), dr.reshape(H, W)
    kl, kr = kl.reshape(H, W).astype(bool), kr.reshape(H, W).astype(bool)
    t = np.arange(W)[None, :] - dl + 0.5
    inside = kl & (t > -1.0) & (t < W)
    xr = np.where(inside, np.trunc(np.where(inside, t, 0.0)), 0).astype(np.int64)
    rows = np.arange(H)[:, None].repeat(W, 1)
    ok = inside & kr[rows, xr] & (np.abs(dl - dr[rows, xr]) <= 1.0)
    return ok.reshape(-1)


def score(pred, gt, mask):
    """score, eval_real.py:105-121, summed left to right like the reference."""
    d = np.abs(np.asarray(pred, dtype=np.float64) - np.asarray(gt, dtype=np.float64))
    d = d[np.asarray(mask, dtype=bool).ravel()]
    if d.size == 0:
        return None
    e = 0.0
    for v in d.tolist():
        e += v
    n = d.size
    return (e / n, 100.0 * int((d > 1.0).sum()) / n, 100.0 * int((d > 2.0).sum()) / n,
            100.0 * int((d > 4.0).sum()) / n)


def constant(gt, mask):
    """constant, eval_real.py:124-127, sequential sums."""
    v = np.asarray(gt, dtype=np.float64)[np.asarray(mask, dtype=bool).ravel()].tolist()
    s = 0.0
    for x in v:
        s += x
    m = s / len(v)
    a = 0.0
    for x in v:
        a += abs(x - m)
    return a / len(v)


# -------------------------------------------------------- disparity to 3D
# Pinned against the reference's own outputs (tests/golden/geometry.npz).

def _depth(cam, d):
    """Camera.depth_from_disparity, stereonet.py:1982-1986 (cam = fx, cx, cy,
    baseline, doffs)."""
    fx, cx, cy, bl, doffs = cam
    den = np.asarray(d, dtype=np.float64) + doffs
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.where(den < 1.0e-6, 0.0, (fx * bl) / den)


def disparity_to_points(disp, W, H, cam, min_disp=0.5, max_depth=1.0e6):
    """stereonet.py:2006-2021: (points [n][3] with NaN rows, valid)."""
    d = np.asarray(disp, dtype=np.float64).ravel()
    fx, cx, cy = cam[0], cam[1], cam[2]
    Z = _depth(cam, d)
    ok = ~(d < min_disp) & ~((Z <= 0.0) | (Z > max_depth))
    x = (np.arange(d.size) % W).astype(np.float64)
    y = ((np.arange(d.size) // W) % H).astype(np.float64)
    pts = np.stack([(x - cx) * Z / fx, (y - cy) * Z / fx, Z], -1)
    pts[~ok] = np.nan
    return pts, ok


def robust_min_disparity(disp, floor=0.4, factor=0.38):
    """stereonet.py:1989-2003."""
    v = np.sort(np.asarray(disp, dtype=np.float64).ravel())
    m = v[v.size // 2] * factor
    return floor if floor > m else m


def _win_first_min(v, r):
    """Per row of v [rows][W]: first strict minimum over [x - r, x + r]
    clipped, and its column."""
    rows, W = v.shape
    x = np.arange(W)
    a = np.maximum(x - r, 0)
    b = np.minimum(x + r, W - 1)
    bv = v[:, a].copy()
    bi = np.broadcast_to(a, v.shape).copy()
    for k in range(1, 2 * r + 1):
        xx = a + k
        inb = xx <= b
        cand = v[:, np.minimum(xx, W - 1)]
        upd = inb[None, :] & (cand < bv)
        bv = np.where(upd, cand, bv)
        bi = np.where(upd, np.minimum(xx, W - 1)[None, :], bi)
    return bv, bi


def min_filter_arg(v, W, H, r, rv=None):
    """min_filter_arg, stereonet.py:5566-5605 (rows, then columns)."""
    rv = r if rv is None else rv
    v2 = np.asarray(v, dtype=np.float64).reshape(H, W)
    rowv, rc = _win_first_min(v2, r)
    ri = np.arange(H)[:, None] * W + rc
    if rv == 0:
        return rowv.ravel(), ri.ravel()
    cv, cr = _win_first_min(rowv.T, rv)  # [W][H]: rows of the transpose are columns
    out = cv.T
    src_rows = cr.T  # the row the minimum came from, per pixel
    oi = ri[src_rows, np.arange(W)[None, :]]
    return out.ravel(), oi.ravel()


def _box_sum_means(v, r):
    """Per row: sum over [x - r, x + r] clipped, left to right, / count."""
    rows, W = v.shape
    x = np.arange(W)
    a = np.maximum(x - r, 0)
    b = np.minimum(x + r, W - 1)
    s = np.zeros_like(v)
    for k in range(0, 2 * r + 1):
        xx = a + k
        inb = (xx <= b)[None, :]
        s = np.where(inb, s + v[:, np.minimum(xx, W - 1)], s)
    return s / (b - a + 1).astype(np.float64)[None, :]


def smooth_background(m, disp, W, H, r):
    """smooth_background, stereonet.py:5608-5634."""
    t = _box_sum_means(np.asarray(m, dtype=np.float64).reshape(H, W), r)
    v = _box_sum_means(t.T, r).T.ravel()
    d = np.asarray(disp, dtype=np.float64).ravel()
    return np.where(v < d, v, d)


def background_region(disp, W, H, radius=None, jump_ratio=0.15, min_jump=2.0):
    """background_region, stereonet.py:5637-5671: (mask, m, src)."""
    if radius is None:
        radius = int(W / 16) if W >= 64 else 4
    d = np.asarray(disp, dtype=np.float64).ravel()
    m, src = min_filter_arg(d, W, H, radius, 0)
    m = smooth_background(m, d, W, H, radius // 2 if radius >= 2 else 1)
    thr = np.where(jump_ratio * d < min_jump, min_jump, jump_ratio * d)
    mask = (d - m > thr).reshape(H, W)
    for _ in range(2):
        g = mask.copy()
        g[:, 1:] |= mask[:, :-1]
        g[:, :-1] |= mask[:, 1:]
        g[1:, :] |= mask[:-1, :]
        g[:-1, :] |= mask[1:, :]
        mask = g
    return mask.ravel(), m, src


def grid_mesh(disp, W, H, cam, step=2, fill=True, jump=1.5, jump_rel=0.06, min_disp=None):
    """grid_mesh, stereonet.py:5770-5862, as plain loops; returns the dict
    with numpy arrays (pos [n][3], uv [n][2], fill, disp, idx [t][3])."""
    d = np.asarray(disp, dtype=np.float64).ravel()
    fx, cx, cy = cam[0], cam[1], cam[2]
    if min_disp is None:
        min_disp = robust_min_disparity(d)
    xs = list(range(0, W, step))
    if xs[-1] != W - 1:
        xs.append(W - 1)
    ys = list(range(0, H, step))
    if ys[-1] != H - 1:
        ys.append(H - 1)
    nx, ny = len(xs), len(ys)
    pos, uv, flag, dv, idx = [], [], [], [], []

    def build(val, is_fill, region, src):
        vid = [-1] * (nx * ny)
        for j, y in enumerate(ys):
            for i, x in enumerate(xs):
                p = y * W + x
                if region is not None and not region[p]:
                    continue
                v = float(val[p])
                if not v >= min_disp or (not is_fill and x - v < 0):
                    continue
                Z = float(_depth(cam, v))
                if Z <= 0.0:
                    continue
                vid[j * nx + i] = len(dv)
                pos.append(((x - cx) * Z / fx, (y - cy) * Z / fx, Z))
                if is_fill:
                    s = int(src[p])
                    uv.append(((s % W + 0.5) / W, (s // W + 0.5) / H))
                else:
                    uv.append(((x + 0.5) / W, (y + 0.5) / H))
                flag.append(1 if is_fill else 0)
                dv.append(v)

        def ok(a, b, c):
            if min(a, b, c) < 0:
                return False
            da, db, dc = dv[a], dv[b], dv[c]
            lo = da if da < (db if db < dc else dc) else (db if db < dc else dc)
            hi = da if da > (db if db > dc else dc) else (db if db > dc else dc)
            lim = jump if jump > jump_rel * hi else jump_rel * hi
            return hi - lo <= lim

        n = 0
        for j in range(ny - 1):
            for i in range(nx - 1):
                a, b = vid[j * nx + i], vid[j * nx + i + 1]
                c, e = vid[(j + 1) * nx + i], vid[(j + 1) * nx + i + 1]
                if ok(a, b, c):
                    idx.append((a, c, b))
                    n += 1
                if ok(b, c, e):
                    idx.append((b, c, e))
                    n += 1
        return n

    main = build(d, False, None, None)
    nfill = 0
    if fill:
        mask, m, src = background_region(d, W, H)
        nfill = build(m, True, mask, src)
    zs = sorted(pos[k][2] for k in range(len(flag)) if not flag[k])
    pivot = zs[len(zs) // 2] if zs else 1.0
    return {"pos": np.array(pos, dtype=np.float64).reshape(-1, 3),
            "uv": np.array(uv, dtype=np.float64).reshape(-1, 2),
            "fill": np.array(flag, dtype=np.int64), "disp": np.array(dv, dtype=np.float64),
            "idx": np.array(idx, dtype=np.int64).reshape(-1, 3), "main_tris": main,
            "fill_tris": nfill, "pivot": pivot}


# ------------------------------------------------------ Gaussian splatting
# SplatRenderer.render (stereonet.py:2706-2852) restated in numpy: the
# reference's scalar projection per Gaussian (its own fsqrt / fexp, stable
# sort by depth), then each Gaussian composited over its footprint box with
# elementwise array operations in the reference's order (exp by the same
# 15-term series; the 2^k scaling is exact here since k >= -7).  Pinned by
# the reference's own renders (tests/golden/splat.npz, tests/test_mesh.py).

_LN2 = 0.6931471805599453


def _rsqrt_ref(x):
    if x <= 0.0:
        return 0.0
    y = x
    if y > 4.0:
        s = 1.0
        while y > 4.0:
            y *= 0.25
            s *= 2.0
        y = s
    elif y < 0.25:
        s = 1.0
        while y < 0.25:
            y *= 4.0
            s *= 0.5
        y = s
    else:
        y = 1.0
    for _ in range(24):
        y = 0.5 * (y + x / y)
    return y


def _exp_series(x):
    """The reference's fexp on an array of values in [-700, 700]."""
    x = np.asarray(x, dtype=np.float64)
    k = np.trunc(x / _LN2 + np.where(x >= 0.0, 0.5, -0.5)).astype(np.int64)
    r = x - k * _LN2
    term = np.ones_like(x)
    total = np.ones_like(x)
    for n in range(1, 16):
        term = term * r / n
        total = total + term
    return np.ldexp(total, k)


def _rexp(x):
    if x > 700.0:
        return 1.0e308
    if x < -700.0:
        return 0.0
    k = int(x / _LN2 + (0.5 if x >= 0.0 else -0.5))
    r = x - k * _LN2
    term = 1.0
    total = 1.0
    for n in range(1, 16):
        term = term * r / n
        total += term
    if k >= 0:
        for _ in range(k):
            total *= 2.0
    else:
        for _ in range(-k):
            total *= 0.5
    return total


def splat_render(cloud, cam, W, H, R, t, blur=0.30, cutoff=3.0, background=(0.0, 0.0, 0.0)):
    """cloud [14][n] (mx my mz sxx sxy sxz syy syz szz cr cg cb raw_op
    raw_sc), cam = (fx, cx, cy); returns C [H*W*3]."""
    cl = np.asarray(cloud, dtype=np.float64)
    fx, cx, cy = cam[0], cam[1], cam[2]
    items = []
    for i in range(cl.shape[1]):
        X, Y, Z = cl[0, i], cl[1, i], cl[2, i]
        xc = R[0] * X + R[1] * Y + R[2] * Z + t[0]
        yc = R[3] * X + R[4] * Y + R[5] * Z + t[1]
        zc = R[6] * X + R[7] * Y + R[8] * Z + t[2]
        if zc <= 1.0e-4:
            continue
        u = fx * xc / zc + cx
        v = fx * yc / zc + cy
        f_z = fx / zc
        f_z2 = -fx / (zc * zc)
        jx = [f_z, 0.0, f_z2 * xc]
        jy = [0.0, f_z, f_z2 * yc]
        m0 = [jx[0] * R[0] + jx[1] * R[3] + jx[2] * R[6], jx[0] * R[1] + jx[1] * R[4] + jx[2] * R[7],
              jx[0] * R[2] + jx[1] * R[5] + jx[2] * R[8]]
        m1 = [jy[0] * R[0] + jy[1] * R[3] + jy[2] * R[6], jy[0] * R[1] + jy[1] * R[4] + jy[2] * R[7],
              jy[0] * R[2] + jy[1] * R[5] + jy[2] * R[8]]
        sxx, sxy, sxz, syy, syz, szz = (float(cl[k, i]) for k in (3, 4, 5, 6, 7, 8))

        def sm(m):
            return [sxx * m[0] + sxy * m[1] + sxz * m[2], sxy * m[0] + syy * m[1] + syz * m[2],
                    sxz * m[0] + syz * m[1] + szz * m[2]]
        sm0, sm1 = sm(m0), sm(m1)
        a00 = m0[0] * sm0[0] + m0[1] * sm0[1] + m0[2] * sm0[2]
        a01 = m0[0] * sm1[0] + m0[1] * sm1[1] + m0[2] * sm1[2]
        a11 = m1[0] * sm1[0] + m1[1] * sm1[1] + m1[2] * sm1[2]
        s = _rexp(float(cl[13, i]))
        c00 = s * a00 + blur
        c01 = s * a01
        c11 = s * a11 + blur
        det = c00 * c11 - c01 * c01
        if det < 1.0e-12:
            continue
        inv = 1.0 / det
        tr = 0.5 * (c00 + c11)
        disc = _rsqrt_ref(tr * tr - det if tr * tr - det > 0.0 else 0.0)
        lmax = tr + disc
        rad = cutoff * _rsqrt_ref(lmax if lmax > 1.0e-12 else 1.0e-12)
        if rad < 0.5:
            rad = 0.5
        if rad > 0.35 * (W + H):
            continue
        x0, x1 = int(u - rad), int(u + rad) + 1
        y0, y1 = int(v - rad), int(v + rad) + 1
        if x1 < 0 or y1 < 0 or x0 >= W or y0 >= H:
            continue
        x0, y0, x1, y1 = max(x0, 0), max(y0, 0), min(x1, W), min(y1, H)
        rop = float(cl[12, i])
        op = 1.0 / (1.0 + _rexp(-rop)) if rop >= 0.0 else _rexp(rop) / (1.0 + _rexp(rop))
        items.append((zc, i, u, v, c11 * inv, -c01 * inv, c00 * inv, x0, x1, y0, y1, op))
    items.sort(key=lambda e: e[0])
    T = np.ones((H, W))
    C = np.zeros((H, W, 3))
    for (zc, i, u, v, i00, i01, i11, x0, x1, y0, y1, op) in items:
        Tp = T[y0:y1, x0:x1]
        dy = (np.arange(y0, y1, dtype=np.float64) - v)[:, None]
        dx = (np.arange(x0, x1, dtype=np.float64) - u)[None, :]
        q = i00 * dx * dx + 2.0 * i01 * dx * dy + i11 * dy * dy
        live = (Tp >= 0.004) & (q <= 9.0)
        a = op * _exp_series(np.where(live, -0.5 * q, 0.0))
        live &= a >= 1.0e-4
        a = np.minimum(a, 0.99)
        w = Tp * a
        cb = C[y0:y1, x0:x1]
        for ch, col in enumerate((cl[9, i], cl[10, i], cl[11, i])):
            cb[..., ch] = np.where(live, cb[..., ch] + w * col, cb[..., ch])
        T[y0:y1, x0:x1] = np.where(live, Tp * (1.0 - a), Tp)
    for ch in range(3):
        C[..., ch] = C[..., ch] + T * background[ch]
    return C.reshape(-1)
