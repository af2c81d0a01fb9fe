"""Mesh half of the downstream 3D output: drop-in counterparts of delaunay,
cull_triangles, cull_slanted, select_points (stereonet.py:2038-2482),
write_obj (:3042-3095), background_layer and merge_layer (:5674-5769).

Delaunay insertion, the culls and the vertex selection are serial walks
over a few hundred vertices, so they run as native host code in
libstereo_mi.so (sv_delaunay, sv_cull_triangles, sv_cull_slanted,
sv_select_points) with the reference's arithmetic and visiting order; the
per-pixel inputs (3D points, background region) come from the GPU kernels of
geometry.py.  Triangle and vertex lists are returned as lists of tuples, as
the reference does.

Per-pixel 3D points (pts3d) are accepted in either form: the reference's
list with None for unusable pixels, or the (points [n][3], valid [n]) pair
of tensors that geometry.disparity_to_points returns for tensor input.
"""

import ctypes

import numpy as np
import torch

from . import _lib as L
from . import geometry as G
from . import refmath as R

_P, _I, _D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
_PI = ctypes.POINTER(ctypes.c_int)
L.extra_signatures({
    "sv_delaunay": (_I, [_P, _I, _P, _I, _PI]),
    "sv_cull_triangles": (_I, [_P, _P, _I, _P, _I, _D, _D, _P, _PI]),
    "sv_cull_slanted": (_I, [_P, _P, _I, _P, _I, _D, _P, _PI]),
    "sv_select_points": (_I, [_P, _I, _I, _P, _I, _D, _P, _I, _PI]),
})


def _np_ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _tris_np(tris):
    t = np.ascontiguousarray(np.asarray(tris, dtype=np.int32).reshape(-1, 3))
    return t


def _tris_out(a, n):
    return [tuple(int(v) for v in r) for r in a[:n].tolist()]


def delaunay(points):
    """delaunay (stereonet.py:2052-2141): Bowyer-Watson over (x, y) points;
    index triples in the reference's order."""
    xy = np.ascontiguousarray(np.asarray(points, dtype=np.float64).reshape(-1, 2))
    n = xy.shape[0]
    if n < 3:
        return []
    cap = 2 * n + 8
    while True:
        out = np.empty((cap, 3), dtype=np.int32)
        cnt = ctypes.c_int(0)
        rc = L.lib().sv_delaunay(_np_ptr(xy), n, _np_ptr(out), cap, ctypes.byref(cnt))
        if rc == L.SV_OK:
            return _tris_out(out, cnt.value)
        if cnt.value <= cap:
            L.check(rc)
        cap = cnt.value


def _points_host(pts3d):
    """(points [n][3] float64, valid [n] uint8) numpy arrays of either form."""
    if isinstance(pts3d, tuple) and len(pts3d) == 2:
        p, v = pts3d
        p = p.detach().cpu().numpy() if isinstance(p, torch.Tensor) else np.asarray(p)
        v = v.detach().cpu().numpy() if isinstance(v, torch.Tensor) else np.asarray(v)
        return np.ascontiguousarray(p, dtype=np.float64).reshape(-1, 3), \
            np.ascontiguousarray(v.astype(np.uint8)).reshape(-1)
    n = len(pts3d)
    p = np.zeros((n, 3), dtype=np.float64)
    v = np.zeros(n, dtype=np.uint8)
    for i, q in enumerate(pts3d):
        if q is not None:
            p[i] = q
            v[i] = 1
    return p, v


def _vertex_points(points2d, pts3d, W):
    """The 3D point and valid flag of every mesh vertex (its pixel's)."""
    xy = np.asarray(points2d, dtype=np.int64).reshape(-1, 2)
    if isinstance(pts3d, tuple) and len(pts3d) == 2 and isinstance(pts3d[0], torch.Tensor):
        idx = torch.from_numpy(xy[:, 1] * W + xy[:, 0]).to(pts3d[0].device)
        vp = pts3d[0].reshape(-1, 3)[idx].cpu().numpy()
        vv = pts3d[1].reshape(-1)[idx].cpu().numpy().astype(np.uint8)
        return np.ascontiguousarray(vp), np.ascontiguousarray(vv)
    p, v = _points_host(pts3d)
    idx = xy[:, 1] * W + xy[:, 0]
    return np.ascontiguousarray(p[idx]), np.ascontiguousarray(v[idx])


def cull_triangles(points2d, tris, pts3d, W, max_z_ratio=1.4, max_edge_factor=4.0):
    """cull_triangles (stereonet.py:2144-2199): drops triangles with an
    unusable vertex, a depth ratio above max_z_ratio, or a longest edge
    beyond max_edge_factor x the median longest edge."""
    t = _tris_np(tris)
    if t.shape[0] == 0:
        return []
    vp, vv = _vertex_points(points2d, pts3d, W)
    out = np.empty_like(t)
    cnt = ctypes.c_int(0)
    L.check(L.lib().sv_cull_triangles(_np_ptr(vp), _np_ptr(vv), vp.shape[0], _np_ptr(t),
                                      t.shape[0], float(max_z_ratio), float(max_edge_factor),
                                      _np_ptr(out), ctypes.byref(cnt)))
    return _tris_out(out, cnt.value)


def cull_slanted(points2d, tris, pts3d, cam, W, max_slant_cos=0.07):
    """cull_slanted (stereonet.py:2324-2372): drops triangles whose normal
    is nearly perpendicular to the ray from the camera to their centroid."""
    t = _tris_np(tris)
    if t.shape[0] == 0:
        return []
    vp, vv = _vertex_points(points2d, pts3d, W)
    out = np.empty_like(t)
    cnt = ctypes.c_int(0)
    L.check(L.lib().sv_cull_slanted(_np_ptr(vp), _np_ptr(vv), vp.shape[0], _np_ptr(t), t.shape[0],
                                    float(max_slant_cos), _np_ptr(out), ctypes.byref(cnt)))
    return _tris_out(out, cnt.value)


def select_points(disp, W, H, valid, target=500, edge_fraction=0.45):
    """select_points (stereonet.py:2375-2482): corners, a uniform grid, then
    the strongest depth gradients at a minimum spacing; list of (x, y)."""
    d = disp.detach().double().cpu().numpy() if isinstance(disp, torch.Tensor) else \
        np.asarray(disp, dtype=np.float64)
    d = np.ascontiguousarray(d.reshape(-1))
    if d.size != W * H:
        raise ValueError("disparity of %d values is not W x H = %d x %d" % (d.size, W, H))
    v = None
    if valid is not None:
        v = valid.detach().cpu().numpy() if isinstance(valid, torch.Tensor) else np.asarray(valid)
        v = np.ascontiguousarray(v.reshape(-1).astype(np.uint8))
    cap = 4 + W * H
    out = np.empty((cap, 2), dtype=np.int32)
    cnt = ctypes.c_int(0)
    L.check(L.lib().sv_select_points(_np_ptr(d), W, H, _np_ptr(v), int(target),
                                     float(edge_fraction), _np_ptr(out), cap, ctypes.byref(cnt)))
    return [(int(x), int(y)) for x, y in out[:cnt.value].tolist()]


def _rgb_host(rgb):
    if isinstance(rgb, torch.Tensor):
        return rgb.detach().double().cpu().numpy().reshape(-1)
    return np.asarray(rgb, dtype=np.float64).reshape(-1)


def write_obj(path, points2d, tris, pts3d, rgb, W, H, max_z_ratio=1.6):
    """write_obj (stereonet.py:3042-3095): Wavefront OBJ with per-vertex
    colour (Y negated), faces without unusable vertices or a depth ratio
    above max_z_ratio; returns (vertices, faces) written."""
    vp, vv = _vertex_points(points2d, pts3d, W)
    c = _rgb_host(rgb)
    xy = np.asarray(points2d, dtype=np.int64).reshape(-1, 2)
    lines = ["# Reconstructed from a stereo pair\n",
             "# StereoNet disparity -> Delaunay triangulation -> mesh\n",
             "# %d vertices\n" % len(xy)]
    index_map = [0] * len(xy)
    nv = 0
    for idx in range(len(xy)):
        if not vv[idx]:
            index_map[idx] = -1
            continue
        x, y = int(xy[idx, 0]), int(xy[idx, 1])
        o = 3 * (y * W + x)
        nv += 1
        index_map[idx] = nv
        P = [float(q) for q in vp[idx]]
        lines.append("v %.6f %.6f %.6f %.4f %.4f %.4f\n" % (P[0], -P[1], P[2], float(c[o]),
                                                           float(c[o + 1]), float(c[o + 2])))
    nf = 0
    for (i, j, k) in _tris_np(tris).tolist():
        a, b, cc = index_map[i], index_map[j], index_map[k]
        if a < 0 or b < 0 or cc < 0:
            continue
        zs = (float(vp[i][2]), float(vp[j][2]), float(vp[k][2]))
        zmin = R.fmin(zs[0], R.fmin(zs[1], zs[2]))
        zmax = R.fmax(zs[0], R.fmax(zs[1], zs[2]))
        if zmin <= 0.0 or zmax > zmin * max_z_ratio:
            continue
        lines.append("f %d %d %d\n" % (a, b, cc))
        nf += 1
    with open(path, "w") as f:
        f.write("".join(lines))
    return nv, nf


def background_layer(disp, rgb, W, H, cam, n_points, radius=None, jump_ratio=0.15,
                     min_jump=2.0, min_disp=0.5):
    """background_layer (stereonet.py:5674-5740): a thin second surface
    behind foreground edges (background_region on the GPU, its colour copied
    from the source pixel, unprojected at the background disparity), a
    vertex grid over it plus its even-coordinate outline, Delaunay, triangles
    whose centroid pixel is masked, cull_triangles.  Returns (points2d,
    tris, pts3d, rgb) with pts3d as (points, valid) tensors and rgb a
    float64 tensor [3*W*H]."""
    d = G._dev(disp)
    mask, m, src = G.background_region_gpu(d, W, H, radius, jump_ratio, min_jump)
    N = W * H
    dev = d.device
    maskb = mask.reshape(-1).to(torch.bool)
    count = int(maskb.sum())
    if count < 12:
        return [], [], (torch.zeros((N, 3), dtype=torch.float64, device=dev),
                        torch.zeros(N, dtype=torch.uint8, device=dev)), \
            torch.zeros(3 * N, dtype=torch.float64, device=dev)
    c = G._dev(rgb).reshape(N, 3)
    srcl = src.reshape(-1).long().clamp(0, N - 1)
    lrgb = torch.where(maskb[:, None], c[srcl], torch.zeros_like(c)).reshape(-1).contiguous()
    # unproject the background disparity where masked (below min_disp or
    # outside the mask: no point), as the reference's per-pixel loop does
    md = torch.where(maskb, m.reshape(-1).double(), torch.full_like(d, -1.0))
    pts, valid = G.disparity_to_points_gpu(md, W, H, cam, min_disp, 1.0e308)
    mk = maskb.reshape(H, W).cpu().numpy()
    # vertices: a grid at the main mesh's density, then the outline
    s = int(R.fsqrt(N / float(n_points if n_points > 0 else 1)) + 0.5)
    if s < 2:
        s = 2
    ys, xs = np.mgrid[0:H:s, 0:W:s]
    sel = mk[ys, xs]
    pts2 = [(int(x), int(y)) for y, x in zip(ys[sel].tolist(), xs[sel].tolist())]
    pad = np.zeros((H + 2, W + 2), dtype=bool)
    pad[1:-1, 1:-1] = mk
    interior = pad[1:-1, :-2] & pad[1:-1, 2:] & pad[:-2, 1:-1] & pad[2:, 1:-1]
    yy, xx = np.mgrid[0:H, 0:W]
    border = (xx == 0) | (xx == W - 1) | (yy == 0) | (yy == H - 1)
    edge = mk & (border | ~interior) & (xx % 2 == 0) & (yy % 2 == 0)
    ey, ex = np.nonzero(edge)
    pts2 += [(int(x), int(y)) for y, x in zip(ey.tolist(), ex.tolist())]
    p3 = (pts, valid)
    if len(pts2) < 3:
        return [], [], p3, lrgb
    tris = delaunay(pts2)
    keep = []
    for (a, b, cc) in tris:
        cx = int((pts2[a][0] + pts2[b][0] + pts2[cc][0]) / 3.0)
        cy = int((pts2[a][1] + pts2[b][1] + pts2[cc][1]) / 3.0)
        if mk[cy, cx]:
            keep.append((a, b, cc))
    keep = cull_triangles(pts2, keep, p3, W)
    return pts2, keep, p3, lrgb


def merge_layer(points2d, tris, pts3d, rgb, layer, W, H):
    """merge_layer (stereonet.py:5743-5767): the main mesh and the layer
    side by side on a 2W-wide canvas (layer vertex (x, y) -> (x + W, y));
    returns (points2d, tris, pts3d, rgb, 2W) with pts3d as (points, valid)
    tensors and rgb a float64 tensor."""
    lp, lt, l3, lrgb = layer
    dev = L.device()
    p, v = pts3d if (isinstance(pts3d, tuple) and isinstance(pts3d[0], torch.Tensor)) else \
        tuple(torch.from_numpy(a).to(dev) for a in _points_host(pts3d))
    lpp, lpv = l3 if (isinstance(l3, tuple) and isinstance(l3[0], torch.Tensor)) else \
        tuple(torch.from_numpy(a).to(dev) for a in _points_host(l3))
    P = torch.cat([p.reshape(H, W, 3), lpp.reshape(H, W, 3)], dim=1).reshape(-1, 3).contiguous()
    V = torch.cat([v.reshape(H, W), lpv.reshape(H, W)], dim=1).reshape(-1).contiguous()
    c = G._dev(rgb).reshape(H, W, 3)
    lc = G._dev(lrgb).reshape(H, W, 3)
    C = torch.cat([c, lc], dim=1).reshape(-1).contiguous()
    n = len(points2d)
    pts = [tuple(q) for q in points2d] + [(x + W, y) for (x, y) in lp]
    tr = [tuple(t) for t in tris] + [(a + n, b + n, c2 + n) for (a, b, c2) in lt]
    return pts, tr, (P, V), C, 2 * W
