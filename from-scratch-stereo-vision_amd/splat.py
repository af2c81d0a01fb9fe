"""Gaussian half of the downstream 3D output: drop-in counterparts of
GaussianCloud, gaussians_from_mesh, SplatRenderer, CloudAdam,
refine_gaussians, look_at, cloud_centroid, orbit_views and reconstruct
(stereonet.py:2485-3187).

The cloud lives on the GPU as one float64 block [14][capacity] (mean,
covariance upper triangle, colour, raw opacity, raw scale) plus gradients
[5][capacity]; GaussianCloud exposes the reference's per-parameter names
as views of it.  Construction, projection, compositing, the backward sweep,
the photometric loss and Adam run in HIP kernels (csrc/splat.hip) with the
reference's float64 arithmetic (its own fsqrt / fexp / flog), so renders
and gradients equal the reference's.  Camera poses and per-run constants
are host arithmetic with the same primitives (refmath.py).

render returns a float64 CUDA tensor [H*W*3] (the reference returns a
list); backward accepts a tensor or a list.
"""

import ctypes
import os

import numpy as np
import torch

from . import _lib as L
from . import geometry as G
from . import mesh as M
from . import refmath as R

_P, _I, _D, _LL, _SZ = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_longlong, \
    ctypes.c_size_t
L.extra_signatures({
    "sv_gaussians_from_mesh": (_I, [_P, _I, _P, _I, _P, _P, _P, _I, _I, _D, _D, _D, _P, _I, _P,
                                    _P]),
    "sv_splat_workspace_bytes": (_I, [_I, ctypes.POINTER(_SZ)]),
    "sv_splat_prepare": (_I, [_P, _I, _I, _P, _P, _D, _D, _D, _I, _I, _D, _D, _P, _SZ, _P, _P]),
    "sv_splat_render": (_I, [_P, _I, _I, _I, _D, _D, _D, _P, _P, _P]),
    "sv_splat_backward": (_I, [_P, _I, _I, _I, _P, _P, _P, _LL, _P, _I, _P]),
    "sv_photometric_mse": (_I, [_P, _P, _LL, _P, _P, _P, _P]),
    "sv_cloud_adam": (_I, [_P, _P, _P, _P, _I, _I, _D, _D, _D, _D, _D, _P]),
})

ALPHA_MAX = 0.99
IDENTITY_R = [1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0]
IDENTITY_T = [0.0, 0.0, 0.0]

_PARAMS = ("mx", "my", "mz", "sxx", "sxy", "sxz", "syy", "syz", "szz", "cr", "cg", "cb",
           "raw_op", "raw_sc")
_GRADS = ("g_cr", "g_cg", "g_cb", "g_op", "g_sc")


class GaussianCloud:
    """GaussianCloud (stereonet.py:2485-2536) on the GPU: params [14][cap],
    grads [5][cap] float64; the reference's attribute names (mx, ..., raw_sc,
    g_cr, ..., g_sc) are views of the first count() columns."""

    def __init__(self, capacity=0, device=None):
        dev = device or L.device()
        cap = max(int(capacity), 1)
        self.params = torch.zeros((14, cap), dtype=torch.float64, device=dev)
        self.grads = torch.zeros((5, cap), dtype=torch.float64, device=dev)
        self.n = 0

    @property
    def cap(self):
        return self.params.shape[1]

    def count(self):
        return self.n

    def zero_grads(self):
        self.grads.zero_()

    def __getattr__(self, name):
        if name in _PARAMS:
            return self.params[_PARAMS.index(name), :self.n]
        if name in _GRADS:
            return self.grads[_GRADS.index(name), :self.n]
        raise AttributeError(name)

    @staticmethod
    def from_arrays(arr):
        """A cloud from a [14][n] array (the reference's lists in _PARAMS
        order)."""
        a = torch.as_tensor(np.asarray(arr, dtype=np.float64))
        gc = GaussianCloud(a.shape[1])
        gc.params[:, :a.shape[1]] = a.to(gc.params.device)
        gc.n = a.shape[1]
        return gc

    def arrays(self):
        """[14][n] float64 numpy copy of the parameters."""
        return self.params[:, :self.n].cpu().numpy()


def _points_dev(pts3d):
    """(points [n][3], valid [n] uint8) CUDA tensors of either pts3d form."""
    if isinstance(pts3d, tuple) and len(pts3d) == 2 and isinstance(pts3d[0], torch.Tensor):
        return pts3d[0].reshape(-1, 3).double().contiguous(), \
            pts3d[1].reshape(-1).to(torch.uint8).contiguous()
    p, v = M._points_host(pts3d)
    dev = L.device()
    return torch.from_numpy(p).to(dev), torch.from_numpy(v).to(dev)


def gaussians_from_mesh(points2d, tris, pts3d, rgb, W, H, thickness=0.35, init_opacity=0.92,
                        scale=1.0):
    """gaussians_from_mesh (stereonet.py:2546-2659): one Gaussian per
    triangle with usable vertices and a depth ratio within 1.6 -- the
    triangle's exact uniform-density covariance plus a normal thickness, the
    mean vertex colour, the initial opacity and log scale -- in triangle
    order."""
    dev = L.device()
    xy = torch.as_tensor(np.asarray(points2d, dtype=np.int32).reshape(-1, 2)).to(dev).contiguous()
    t = torch.as_tensor(M._tris_np(tris)).to(dev).contiguous()
    pts, valid = _points_dev(pts3d)
    c = G._dev(rgb)
    if pts.shape[0] != W * H or c.numel() != 3 * W * H:
        raise ValueError("pts3d / rgb do not hold W x H = %d x %d pixels" % (W, H))
    inv_log = 0.0 if scale <= 0.0 else R.flog(scale)
    p0 = R.fclamp(init_opacity, 1.0e-4, 1.0 - 1.0e-4)
    raw_op = R.flog(p0 / (1.0 - p0))
    gc = GaussianCloud(max(t.shape[0], 1))
    cnt = torch.zeros(1, dtype=torch.int32, device=dev)
    L.check(L.lib().sv_gaussians_from_mesh(L.ptr(xy), xy.shape[0], L.ptr(t), t.shape[0],
                                           L.ptr(pts), L.ptr(valid), L.ptr(c), W, H,
                                           float(thickness), raw_op, inv_log, L.ptr(gc.params),
                                           gc.cap, L.ptr(cnt), L.stream()))
    gc.n = int(cnt.item())
    gc.zero_grads()
    return gc


class SplatRenderer:
    """SplatRenderer (stereonet.py:2706-2962): EWA projection of every
    Gaussian, depth-sorted front-to-back compositing (alpha capped at
    ALPHA_MAX), and the back-to-front gradient sweep into the cloud's colour,
    opacity and scale."""

    def __init__(self, cam, W, H, blur=0.30, cutoff=3.0):
        self.cam = cam
        self.W = W
        self.H = H
        self.blur = blur
        self.cutoff = cutoff
        self.cache = None

    def _prepare(self, gc, R_, t):
        n = gc.count()
        nb = ctypes.c_size_t(0)
        L.check(L.lib().sv_splat_workspace_bytes(n, ctypes.byref(nb)))
        ws = torch.empty(max(nb.value, 1), dtype=torch.uint8, device=gc.params.device)
        info = torch.zeros(2, dtype=torch.int64, device=gc.params.device)
        Rh = (ctypes.c_double * 9)(*[float(v) for v in R_])
        th = (ctypes.c_double * 3)(*[float(v) for v in t])
        L.check(L.lib().sv_splat_prepare(L.ptr(gc.params), n, gc.cap, Rh, th, self.cam.fx,
                                         self.cam.cx, self.cam.cy, self.W, self.H, self.blur,
                                         self.cutoff, L.ptr(ws), nb.value, L.ptr(info),
                                         L.stream()))
        return ws, n, info

    def render(self, gc, R_, t, background=(0.0, 0.0, 0.0)):
        """Colour [H*W*3] float64 (CUDA tensor)."""
        ws, n, info = self._prepare(gc, R_, t)
        HW = self.W * self.H
        C = torch.empty(3 * HW, dtype=torch.float64, device=gc.params.device)
        T = torch.empty(HW, dtype=torch.float64, device=gc.params.device)
        L.check(L.lib().sv_splat_render(L.ptr(ws), n, self.W, self.H, float(background[0]),
                                        float(background[1]), float(background[2]), L.ptr(C),
                                        L.ptr(T), L.stream()))
        self.cache = (ws, n, info, T, gc)
        return C

    def backward(self, gC):
        """Adds d(sum gC * C)/d(colour, raw opacity, raw scale) of the last
        render to its cloud's gradients."""
        ws, n, info, T, gc = self.cache
        g = G._dev(gC)
        total = int(info[1].item())  # footprint area of the kept items
        contrib = torch.empty(max(5 * total, 1), dtype=torch.float64, device=ws.device)
        L.check(L.lib().sv_splat_backward(L.ptr(ws), n, self.W, self.H, L.ptr(g), L.ptr(T),
                                          L.ptr(contrib), total, L.ptr(gc.grads), gc.cap,
                                          L.stream()))


class CloudAdam:
    """CloudAdam (stereonet.py:2965-3007): Adam over colour (lr_colour),
    raw opacity (lr_op) and raw scale (lr_scale), gradients clipped to
    +-10, colours clamped to [0, 1]."""

    def __init__(self, gc, lr_colour=0.05, lr_op=0.05, lr_scale=0.02):
        self.gc = gc
        self.lr = (lr_colour, lr_op, lr_scale)
        self.m = torch.zeros((5, gc.cap), dtype=torch.float64, device=gc.params.device)
        self.v = torch.zeros_like(self.m)
        self.b1t = 1.0
        self.b2t = 1.0

    def step(self):
        b1, b2 = 0.9, 0.999
        self.b1t *= b1
        self.b2t *= b2
        c1 = 1.0 / (1.0 - self.b1t)
        c2 = 1.0 / (1.0 - self.b2t)
        gc = self.gc
        L.check(L.lib().sv_cloud_adam(L.ptr(gc.params), L.ptr(gc.grads), L.ptr(self.m),
                                      L.ptr(self.v), gc.count(), gc.cap, c1, c2, *self.lr,
                                      L.stream()))


def refine_gaussians(gc, renderer, target_rgb, iters=60, log_every=10,
                     background=(0.0, 0.0, 0.0), verbose=True):
    """refine_gaussians (stereonet.py:3010-3039): Adam on the mean squared
    photometric error of the render from the original viewpoint; returns
    the per-iteration losses."""
    opt = CloudAdam(gc)
    target = G._dev(target_rgb)
    n = target.numel()
    gC = torch.empty(n, dtype=torch.float64, device=target.device)
    part = torch.empty(256, dtype=torch.float64, device=target.device)
    losses = torch.empty(max(iters, 1), dtype=torch.float64, device=target.device)
    for it in range(1, iters + 1):
        gc.zero_grads()
        C = renderer.render(gc, IDENTITY_R, IDENTITY_T, background)
        L.check(L.lib().sv_photometric_mse(L.ptr(C), L.ptr(target), n, L.ptr(gC),
                                           L.ptr(losses[it - 1:]), L.ptr(part), L.stream()))
        renderer.backward(gC)
        opt.step()
        if verbose and (it % log_every == 0 or it == 1):
            loss = float(losses[it - 1])
            psnr = 10.0 * R.flog(1.0 / R.fmax(loss, 1.0e-12)) / 2.302585092994046
            print("    refine %3d   MSE %.5f   PSNR %.2f dB" % (it, loss, psnr))
    return losses[:iters].cpu().tolist()


def look_at(eye, target, up=(0.0, -1.0, 0.0)):
    """look_at (stereonet.py:2670-2701): world-to-camera (R, t)."""
    fx_ = target[0] - eye[0]
    fy_ = target[1] - eye[1]
    fz_ = target[2] - eye[2]
    fl = R.fsqrt(fx_ * fx_ + fy_ * fy_ + fz_ * fz_)
    if fl < 1.0e-12:
        fl = 1.0
    fx_ /= fl
    fy_ /= fl
    fz_ /= fl
    rx = fy_ * up[2] - fz_ * up[1]
    ry = fz_ * up[0] - fx_ * up[2]
    rz = fx_ * up[1] - fy_ * up[0]
    rl = R.fsqrt(rx * rx + ry * ry + rz * rz)
    if rl < 1.0e-12:
        rx, ry, rz, rl = 1.0, 0.0, 0.0, 1.0
    rx /= rl
    ry /= rl
    rz /= rl
    ux = ry * fz_ - rz * fy_
    uy = rz * fx_ - rx * fz_
    uz = rx * fy_ - ry * fx_
    Rm = [rx, ry, rz, ux, uy, uz, fx_, fy_, fz_]
    t = [-(Rm[0] * eye[0] + Rm[1] * eye[1] + Rm[2] * eye[2]),
         -(Rm[3] * eye[0] + Rm[4] * eye[1] + Rm[5] * eye[2]),
         -(Rm[6] * eye[0] + Rm[7] * eye[1] + Rm[8] * eye[2])]
    return Rm, t


def cloud_centroid(gc):
    """cloud_centroid (stereonet.py:3098-3107): mean of the Gaussian means,
    summed in index order."""
    n = gc.count()
    if n == 0:
        return (0.0, 0.0, 1.0)
    sx = sy = sz = 0.0
    m = gc.params[:3, :n].cpu().numpy()
    for i in range(n):
        sx += float(m[0, i])
        sy += float(m[1, i])
        sz += float(m[2, i])
    return (sx / n, sy / n, sz / n)


def orbit_views(gc, n_views=5, sweep=0.40, rise=0.12):
    """orbit_views (stereonet.py:3110-3136): camera poses on an arc about
    the cloud's centroid; [(angle, (R, t))]."""
    cx, cy, cz = cloud_centroid(gc)
    radius = R.fsqrt(cx * cx + cy * cy + cz * cz)
    if radius < 1.0e-6:
        radius = 1.0
    views = []
    for i in range(n_views):
        t = -1.0 + 2.0 * (i / float(R.fmax(1, n_views - 1)))
        ang = t * sweep
        cos_a = R.fcos(ang)
        sin_a = R.fsqrt(R.fmax(0.0, 1.0 - cos_a * cos_a))
        if ang < 0.0:
            sin_a = -sin_a
        eye = (cx - radius * sin_a, cy + t * rise * radius, cz - radius * cos_a)
        views.append((ang, look_at(eye, (cx, cy, cz), (0.0, -1.0, 0.0))))
    return views


def reconstruct(disp, rgb, W, H, cam, valid=None, n_points=520, refine_iters=60,
                out_prefix="scene", n_views=5, verbose=True):
    """reconstruct (stereonet.py:3139-3187): unproject, select vertices,
    Delaunay + cull, write the OBJ mesh, build and photometrically refine the
    Gaussian cloud, render the original and n_views novel viewpoints (PPM
    files).  Returns the reference's dict (render as a CUDA tensor, points3d
    as the (points, valid) pair)."""
    from .images import write_ppm
    if verbose:
        print("  unprojecting %d pixels" % (W * H))
    pts3d = G.disparity_to_points_gpu(disp, W, H, cam, min_disp=0.4)
    if verbose:
        print("  selecting mesh vertices")
    pts2d = M.select_points(disp, W, H, valid, target=n_points)
    if verbose:
        print("  triangulating %d vertices" % len(pts2d))
    tris = M.cull_triangles(pts2d, M.delaunay(pts2d), pts3d, W)
    nv, nf = M.write_obj(out_prefix + "_mesh.obj", pts2d, tris, pts3d, rgb, W, H)
    if verbose:
        print("  mesh: %d vertices, %d faces -> %s_mesh.obj" % (nv, nf, out_prefix))
    gc = gaussians_from_mesh(pts2d, tris, pts3d, rgb, W, H)
    if verbose:
        print("  built %d anisotropic Gaussians" % gc.count())
    renderer = SplatRenderer(cam, W, H)
    if refine_iters > 0:
        if verbose:
            print("  photometric refinement:")
        refine_gaussians(gc, renderer, rgb, iters=refine_iters, verbose=verbose)
    C0 = renderer.render(gc, IDENTITY_R, IDENTITY_T)
    write_ppm(out_prefix + "_render_original.ppm", C0.cpu().tolist(), W, H)
    views = orbit_views(gc, n_views=n_views)
    for vi in range(len(views)):
        ang, (Rm, t) = views[vi]
        Cv = renderer.render(gc, Rm, t)
        write_ppm("%s_novel_%d.ppm" % (out_prefix, vi), Cv.cpu().tolist(), W, H)
    if verbose:
        print("  rendered %d novel views" % len(views))
    return {"gaussians": gc, "mesh": (pts2d, tris, nv, nf), "render": C0, "points3d": pts3d,
            "renderer": renderer}
