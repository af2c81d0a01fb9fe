"""Disparity to 3D on the GPU: drop-in counterparts of Camera, read_calib,
robust_min_disparity, disparity_to_points (stereonet.py:416-452,
:1951-2021), min_filter_arg, smooth_background, background_region and
grid_mesh (:5566-5671, :5770-5862).

The per-pixel and per-vertex work runs in HIP kernels (libstereo_mi.so:
sv_disparity_to_points, sv_kth_f64, sv_min_filter_arg,
sv_smooth_background, sv_background_region, sv_grid_mesh_build), in
float64 with the reference's operation order, so every value is identical
to the reference's.  Lists in give lists out (with None for unusable
points, as the reference); numpy arrays and torch tensors are accepted, and
the *_gpu forms return CUDA tensors.  grid_mesh reads its vertex and
triangle counts back once (they size the returned arrays).
"""

import ctypes

import numpy as np
import torch

from . import _lib as L

_P, _I, _D, _LL, _SZ = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_longlong, \
    ctypes.c_size_t
L.extra_signatures({
    "sv_disparity_to_points": (_I, [_P, _I, _I, _I, _P, _D, _D, _P, _P, _P]),
    "sv_kth_f64_workspace_bytes": (_I, [_I, ctypes.POINTER(_SZ)]),
    "sv_kth_f64": (_I, [_P, _LL, _LL, _LL, _P, _P, _SZ, _P]),
    "sv_kth_f64_batch": (_I, [_P, _LL, _LL, _I, _LL, _LL, _P, _I, _D, _P, _P, _SZ, _P]),
    "sv_min_filter_arg": (_I, [_P, _I, _I, _I, _I, _I, _P, _P, _P, _P, _P]),
    "sv_smooth_background": (_I, [_P, _P, _I, _I, _I, _I, _P, _P, _P]),
    "sv_background_region": (_I, [_P, _I, _I, _I, _I, _D, _D, _P, _P, _P, _P, _P, _P]),
    "sv_grid_mesh_workspace_bytes": (_I, [_I, _I, _I, _I, ctypes.POINTER(_SZ)]),
    "sv_grid_mesh_build": (_I, [_P, _P, _P, _I, _I, _I, _I, _P, _P, _D, _D, _P, _P, _P, _P, _P,
                                _LL, _LL, _P, _P, _SZ, _P]),
})


class Camera:
    """Camera, stereonet.py:1951-1992: pinhole intrinsics plus a stereo
    baseline; Z = fx * baseline / (d + doffs)."""

    def __init__(self, fx, cx, cy, baseline, doffs=0.0):
        self.fx = fx
        self.cx = cx
        self.cy = cy
        self.baseline = baseline
        self.doffs = doffs

    @staticmethod
    def synthetic(W, H, baseline=0.10):
        """60-degree horizontal field of view: fx = (W / 2) / tan(30 deg)."""
        return Camera((W * 0.5) / 0.5773502691896257, W * 0.5, H * 0.5, baseline, 0.0)

    @staticmethod
    def from_calib(path):
        c = read_calib(path)
        return Camera(c["fx"], c["cx"], c["cy"], c["baseline"], c["doffs"])

    def depth_from_disparity(self, d):
        den = d + self.doffs
        if den < 1.0e-6:
            return 0.0
        return self.fx * self.baseline / den

    def unproject(self, u, v, Z):
        return (u - self.cx) * Z / self.fx, (v - self.cy) * Z / self.fx, Z

    def _host(self):
        return (ctypes.c_double * 5)(self.fx, self.cx, self.cy, self.baseline, self.doffs)


def read_calib(path):
    """read_calib, stereonet.py:416-452: fx, cx, cy, baseline and doffs from
    a Middlebury calib.txt (cam0=[f 0 cx; 0 f cy; 0 0 1], baseline=,
    doffs=); missing entries stay None (doffs 0)."""
    out = {"fx": None, "cx": None, "cy": None, "baseline": None, "doffs": 0.0}
    with open(path, "r") as f:
        for line in f:
            line = line.strip()
            if "=" not in line:
                continue
            key, val = (t.strip() for t in line.split("=", 1))
            if key == "cam0":
                rows = val.replace("[", "").replace("]", "").split(";")
                first = rows[0].split()
                out["fx"], out["cx"] = float(first[0]), float(first[2])
                if len(rows) > 1:
                    out["cy"] = float(rows[1].split()[2])
            elif key == "baseline":
                out["baseline"] = float(val)
            elif key == "doffs":
                out["doffs"] = float(val)
    return out


def _kind(a):
    if isinstance(a, torch.Tensor):
        return "torch"
    return "numpy" if isinstance(a, np.ndarray) else "list"


def _dev(a, dtype=torch.float64):
    if isinstance(a, torch.Tensor):
        t = a
    else:
        t = torch.from_numpy(np.ascontiguousarray(np.asarray(a, dtype=np.float64)))
    return t.to(device=L.device(), dtype=dtype).reshape(-1).contiguous()


def _out(t, kind, like=None):
    if kind == "torch":
        return t
    a = t.cpu().numpy()
    if kind == "numpy":
        return a.reshape(np.shape(like)) if like is not None else a
    return a.ravel().tolist()


def _count(t, W, H):
    n = t.numel()
    if W <= 0 or H <= 0 or n == 0 or n % (W * H):
        raise ValueError("array of %d values does not hold W x H = %d x %d maps" % (n, W, H))
    return n // (W * H)


def kth_smallest(x, k, stride=1, n=None):
    """The k-th smallest (0-based) of n float64 values x[0], x[stride], ...
    (a CUDA tensor), as a 0-d device tensor."""
    n = x.numel() // stride if n is None else n
    if not 0 <= k < n:
        raise ValueError("k = %d outside [0, %d)" % (k, n))
    nb = ctypes.c_size_t(0)
    L.check(L.lib().sv_kth_f64_workspace_bytes(1, ctypes.byref(nb)))
    ws = L.workspace.get(nb.value)
    out = torch.empty((), dtype=torch.float64, device=x.device)
    L.check(L.lib().sv_kth_f64(L.ptr(x), stride, n, k, L.ptr(out), L.ptr(ws), nb.value,
                               L.stream()))
    return out


def robust_min_disparity(disp, floor=0.4, factor=0.38):
    """robust_min_disparity, stereonet.py:1989-2003: max(floor, median *
    factor), the median being sorted(disp)[len // 2]."""
    d = _dev(disp)
    if d.numel() == 0:
        raise IndexError("list index out of range")
    med = float(kth_smallest(d, d.numel() // 2))
    m = med * factor
    return floor if floor > m else m  # fmax(floor, m)


def disparity_to_points_gpu(disp, W, H, cam, min_disp=0.5, max_depth=1.0e6):
    """(points [n][3] float64 with NaN where unusable, valid [n] uint8)."""
    d = _dev(disp)
    B = _count(d, W, H)
    pts = torch.empty((d.numel(), 3), dtype=torch.float64, device=d.device)
    valid = torch.empty(d.numel(), dtype=torch.uint8, device=d.device)
    L.check(L.lib().sv_disparity_to_points(L.ptr(d), B, W, H, cam._host(), float(min_disp),
                                           float(max_depth), L.ptr(pts), L.ptr(valid), L.stream()))
    return pts, valid


def disparity_to_points(disp, W, H, cam, min_disp=0.5, max_depth=1.0e6):
    """disparity_to_points, stereonet.py:2006-2021: (X, Y, Z) per pixel, None
    where the disparity is below min_disp or the depth is not in
    (0, max_depth]."""
    kind = _kind(disp)
    pts, valid = disparity_to_points_gpu(disp, W, H, cam, min_disp, max_depth)
    if kind == "torch":
        return pts, valid
    p = pts.cpu().numpy()
    v = valid.cpu().numpy().astype(bool)
    if kind == "numpy":
        return p, v
    return [tuple(r) if ok else None for r, ok in zip(p.tolist(), v.tolist())]


def min_filter_arg(v, W, H, r, rv=None):
    """min_filter_arg, stereonet.py:5566-5605: (minimum over the clipped
    (2r+1) x (2rv+1) window, flat index of the first strict minimum), rows
    first then columns."""
    kind = _kind(v)
    t = _dev(v)
    B = _count(t, W, H)
    rv = r if rv is None else rv
    out = torch.empty_like(t)
    idx = torch.empty(t.numel(), dtype=torch.int32, device=t.device)
    tmp = torch.empty_like(t) if rv else None
    tmpi = torch.empty_like(idx) if rv else None
    L.check(L.lib().sv_min_filter_arg(L.ptr(t), B, W, H, int(r), int(rv), L.ptr(out), L.ptr(idx),
                                      L.ptr(tmp), L.ptr(tmpi), L.stream()))
    if kind == "torch":
        return out, idx
    return _out(out, kind), _out(idx.to(torch.int64), kind)


def smooth_background(m, disp, W, H, r):
    """smooth_background, stereonet.py:5608-5634: box means over rows then
    columns (radius r, clipped), clamped to at most disp."""
    kind = _kind(m)
    mt, d = _dev(m), _dev(disp)
    B = _count(mt, W, H)
    tmp, out = torch.empty_like(mt), torch.empty_like(mt)
    L.check(L.lib().sv_smooth_background(L.ptr(mt), L.ptr(d), B, W, H, int(r), L.ptr(tmp),
                                         L.ptr(out), L.stream()))
    return _out(out, kind)


def background_region_gpu(disp, W, H, radius=None, jump_ratio=0.15, min_jump=2.0):
    """(mask uint8, smoothed background disparity, source pixel int32) as
    CUDA tensors, for one or more maps of W x H."""
    d = _dev(disp)
    B = _count(d, W, H)
    if radius is None:
        radius = int(W / 16) if W >= 64 else 4
    mask = torch.empty(d.numel(), dtype=torch.uint8, device=d.device)
    m, tmp = torch.empty_like(d), torch.empty_like(d)
    src = torch.empty(d.numel(), dtype=torch.int32, device=d.device)
    tmask = torch.empty_like(mask)
    L.check(L.lib().sv_background_region(L.ptr(d), B, W, H, int(radius), float(jump_ratio),
                                         float(min_jump), L.ptr(mask), L.ptr(m), L.ptr(src),
                                         L.ptr(tmp), L.ptr(tmask), L.stream()))
    return mask, m, src


def background_region(disp, W, H, radius=None, jump_ratio=0.15, min_jump=2.0):
    """background_region, stereonet.py:5637-5671: (mask, background
    disparity, source pixel): pixels whose disparity exceeds the smoothed
    row-wise minimum within radius by max(jump_ratio * d, min_jump), grown by
    2 px."""
    kind = _kind(disp)
    mask, m, src = background_region_gpu(disp, W, H, radius, jump_ratio, min_jump)
    if kind == "torch":
        return mask, m, src
    mk = mask.cpu().numpy().astype(bool)
    if kind == "numpy":
        return mk, m.cpu().numpy(), src.cpu().numpy().astype(np.int64)
    return mk.tolist(), m.cpu().numpy().tolist(), src.cpu().numpy().tolist()


def _grid_dims(W, H, step):
    nx = (W + step - 1) // step + (1 if (W - 1) % step else 0)
    ny = (H + step - 1) // step + (1 if (H - 1) % step else 0)
    return nx, ny


def _kth_batch(x, mstride, stride, B, n, k=0, ns=None, nstride=0, empty=0.0):
    nb = ctypes.c_size_t(0)
    L.check(L.lib().sv_kth_f64_workspace_bytes(B, ctypes.byref(nb)))
    ws = L.workspace.get(nb.value)
    out = torch.empty(B, dtype=torch.float64, device=x.device)
    L.check(L.lib().sv_kth_f64_batch(L.ptr(x), mstride, stride, B, n, k, L.ptr(ns), nstride,
                                     float(empty), L.ptr(out), L.ptr(ws), nb.value, L.stream()))
    return out


def grid_mesh_batch(disp, W, H, cam, step=2, fill=True, jump=1.5, jump_rel=0.06,
                    min_disp=None):
    """grid_mesh for B maps of one size ([B][H][W] or B*W*H values) with one
    launch per stage for all maps.  Returns a list of B dicts as
    grid_mesh_gpu's (tensor views into shared buffers).  The counts and
    pivots are read back once for the whole batch."""
    d = _dev(disp)
    B = _count(d, W, H)
    if step < 1:
        raise ValueError("step must be >= 1")
    dev = d.device
    npix = W * H
    if min_disp is None:  # robust_min_disparity per map: max(0.4, median * 0.38)
        med = _kth_batch(d, npix, 1, B, npix, npix // 2)
        m = med * 0.38
        mind = torch.where(torch.full_like(m, 0.4) > m, torch.full_like(m, 0.4), m)
    else:
        mind = torch.full((B,), float(min_disp), dtype=torch.float64, device=dev)
    nx, ny = _grid_dims(W, H, step)
    builds = 2 if fill else 1
    vcap, tcap = builds * nx * ny, max(builds * 2 * (nx - 1) * (ny - 1), 1)
    pos = torch.empty((B, vcap, 3), dtype=torch.float64, device=dev)
    uv = torch.empty((B, vcap, 2), dtype=torch.float64, device=dev)
    flag = torch.empty((B, vcap), dtype=torch.uint8, device=dev)
    dv = torch.empty((B, vcap), dtype=torch.float64, device=dev)
    idx = torch.empty((B, tcap, 3), dtype=torch.int32, device=dev)
    totals = torch.zeros((B, 4), dtype=torch.int32, device=dev)
    nb = ctypes.c_size_t(0)
    L.check(L.lib().sv_grid_mesh_workspace_bytes(B, W, H, step, ctypes.byref(nb)))
    ws = L.workspace.get(nb.value)
    camh = cam._host()
    lib = L.lib()
    outs = (L.ptr(pos), L.ptr(uv), L.ptr(flag), L.ptr(dv), L.ptr(idx), vcap, tcap, L.ptr(totals))
    L.check(lib.sv_grid_mesh_build(L.ptr(d), None, None, B, W, H, step, camh, L.ptr(mind),
                                   float(jump), float(jump_rel), *outs, L.ptr(ws), nb.value,
                                   L.stream()))
    if fill:
        mask, m, src = background_region_gpu(d, W, H)
        L.check(lib.sv_grid_mesh_build(L.ptr(m), L.ptr(mask), L.ptr(src), B, W, H, step, camh,
                                       L.ptr(mind), float(jump), float(jump_rel), *outs, L.ptr(ws),
                                       nb.value, L.stream()))
    # pivot: the median depth of each map's main vertices (1.0 when none)
    piv = _kth_batch(pos.reshape(-1)[2:], 3 * vcap, 3, B, vcap, ns=totals, nstride=4, empty=1.0)
    tot = totals.cpu().tolist()
    pivots = piv.cpu().tolist()
    res = []
    for b in range(B):
        mv, mt, fv, ft = tot[b]
        n_v, n_t = mv + fv, mt + ft
        res.append({"pos": pos[b, :n_v], "uv": uv[b, :n_v], "fill": flag[b, :n_v],
                    "disp": dv[b, :n_v], "idx": idx[b, :n_t], "main_tris": mt, "fill_tris": ft,
                    "pivot": pivots[b], "fx": cam.fx, "cx": cam.cx, "cy": cam.cy, "W": W, "H": H})
    return res


def grid_mesh_gpu(disp, W, H, cam, step=2, fill=True, jump=1.5, jump_rel=0.06, min_disp=None):
    """grid_mesh with every array left on the GPU: a dict with pos [n][3],
    uv [n][2], fill [n] uint8, disp [n], idx [t][3] int32 tensors, the
    counts main_tris / fill_tris, pivot and the camera fields."""
    d = _dev(disp)
    if d.numel() != W * H:
        raise ValueError("grid_mesh takes one W x H = %d x %d map, not %d values"
                         % (W, H, d.numel()))
    return grid_mesh_batch(d, W, H, cam, step, fill, jump, jump_rel, min_disp)[0]


def grid_mesh(disp, W, H, cam, step=2, fill=True, jump=1.5, jump_rel=0.06, min_disp=None):
    """grid_mesh, stereonet.py:5770-5862: the viewer's dense grid mesh -- a
    vertex every `step` pixels (plus the last row and column) where the
    disparity is at least min_disp, x - d >= 0 and the depth positive; two
    triangles per cell unless the disparity jumps by more than max(jump,
    jump_rel * d); with fill=True a second grid over background_region's
    mask, textured from each value's source pixel.  Returns the reference's
    dict of flat lists."""
    g = grid_mesh_gpu(disp, W, H, cam, step, fill, jump, jump_rel, min_disp)
    out = dict(g)
    out["pos"] = g["pos"].cpu().numpy().ravel().tolist()
    out["uv"] = g["uv"].cpu().numpy().ravel().tolist()
    out["fill"] = g["fill"].cpu().numpy().astype(np.int64).tolist()
    out["disp"] = g["disp"].cpu().numpy().tolist()
    out["idx"] = g["idx"].cpu().numpy().astype(np.int64).ravel().tolist()
    return out
