"""Seeded synthetic rectified stereo pairs (host side, numpy).

The generator follows the scene model of the reference's own generators
(stereonet.py:1435-1565 random-dot stereograms, :3658-3816 textured scenes,
make_dataset.py:201-209 rendered scenes) without reusing their code:

* the left view is band-limited noise texture (uniform noise, 3x3 box blur
  twice, contrast stretch, rounded to uint8), or optionally a multiscale
  texture (value-noise octaves plus that fine noise);
* the disparity field is piecewise planar: 3-6 slanted planes over convex
  (Voronoi) regions, plus optional thin fronto-parallel foreground boxes;
* the right view is a z-buffered forward warp of a wider virtual left view
  (larger disparity wins, ``xr = int(x - d + 0.5)`` with truncation toward
  zero as at stereonet.py:3675), so every right column has real content;
  disocclusions are stretched from the background side, then Gaussian noise
  is added and the result is rounded to uint8;
* ``valid`` marks left pixels visible in the right view (the z-buffer test of
  stereonet.py:3808-3814);
* ``disp0`` is the coarse estimate the band refinement starts from: ground
  truth plus integer noise in [-n, n], stored as float32-representable values
  so that every consumer derives the same band centre.
"""

import numpy as np


def _box_blur3(a):
    """3x3 mean filter with edge replication."""
    p = np.pad(a, 1, mode="edge")
    out = np.zeros_like(a)
    H, W = a.shape
    for dy in range(3):
        for dx in range(3):
            out += p[dy:dy + H, dx:dx + W]
    return out / 9.0


def _texture(rng, H, W):
    t = rng.uniform(0.0, 1.0, size=(H, W))
    t = _box_blur3(_box_blur3(t))
    lo, hi = np.percentile(t, 1.0), np.percentile(t, 99.0)
    t = (t - lo) / max(hi - lo, 1e-9)
    return np.clip(t * 235.0 + 10.0, 0.0, 255.0)


def _multiscale_texture(rng, H, W):
    """Texture with structure at every scale: octaves of value noise from 4
    cells across up to cells of ~4 px (Catmull-Rom resampled, amplitude
    x0.7 per octave) plus the band-limited fine noise, stretched to uint8
    range.  Unlike the fine noise alone it survives the 10-20x downscaling
    the network's working resolution applies to large pairs."""
    from .scenes import _cubic_matrix
    acc = np.zeros((H, W))
    amp, cells = 1.0, 4
    while cells <= max(4, W // 4):
        cy = max(2, int(round(cells * H / float(W))))
        lat = rng.random((cy + 1, cells + 1))
        acc += amp * (_cubic_matrix(cy + 1, H) @ lat @ _cubic_matrix(cells + 1, W).T)
        amp *= 0.7
        cells *= 2
    fine = _box_blur3(_box_blur3(rng.uniform(0.0, 1.0, size=(H, W))))
    t = acc / max(acc.std(), 1e-9) + 0.6 * fine / max(fine.std(), 1e-9)
    lo, hi = np.percentile(t, 1.0), np.percentile(t, 99.0)
    t = (t - lo) / max(hi - lo, 1e-9)
    return np.clip(t * 235.0 + 10.0, 0.0, 255.0)


def _plane_field(rng, H, W, dmin, dmax, n_planes):
    """Voronoi regions (convex), each carrying a slanted plane d=ax+by+c."""
    sx = rng.uniform(0, W, size=n_planes)
    sy = rng.uniform(0, H, size=n_planes)
    yy, xx = np.mgrid[0:H, 0:W].astype(np.float64)
    dist = (xx[None] - sx[:, None, None]) ** 2 + (yy[None] - sy[:, None, None]) ** 2
    lab = np.argmin(dist, axis=0)
    span = float(dmax - dmin)
    d = np.zeros((H, W))
    for i in range(n_planes):
        c = rng.uniform(dmin, dmax)
        a = rng.uniform(-0.25, 0.25) * span / W
        b = rng.uniform(-0.25, 0.25) * span / H
        plane = c + a * (xx - sx[i]) + b * (yy - sy[i])
        m = lab == i
        d[m] = plane[m]
    return np.clip(d, dmin, dmax)


def _add_boxes(rng, d, n_boxes, dmax, wmin=3, wmax=8):
    H, W = d.shape
    for _ in range(n_boxes):
        bw = int(rng.integers(wmin, wmax + 1))
        bh = int(rng.integers(max(4, H // 4), max(5, (3 * H) // 4)))
        x0 = int(rng.integers(0, max(1, W - bw)))
        y0 = int(rng.integers(0, max(1, H - bh)))
        val = rng.uniform(0.85, 1.0) * dmax
        region = d[y0:y0 + bh, x0:x0 + bw]
        d[y0:y0 + bh, x0:x0 + bw] = np.maximum(region, val)
    return d


def _warp_right(left, disp):
    """Forward warp with z-buffer; returns (right, zbuf).  Holes take the
    nearest filled pixel on the background side (the smaller disparity of
    the two nearest filled neighbours, left on a tie), 0 if none."""
    H, W = left.shape
    right = np.full((H, W), -1.0)
    zbuf = np.full((H, W), -1.0)
    xs = np.arange(W, dtype=np.float64)
    for y in range(H):
        d = disp[y]
        xr = np.trunc(xs - d + 0.5).astype(np.int64)
        ok = (xr >= 0) & (xr < W)
        # increasing disparity, so the nearest surface is written last
        order = np.argsort(d[ok], kind="stable")
        src = np.nonzero(ok)[0][order]
        dst = xr[ok][order]
        right[y, dst] = left[y, src]
        zbuf[y, dst] = d[src]
    filled = zbuf >= 0.0
    cols = np.broadcast_to(np.arange(W), (H, W))
    li = np.maximum.accumulate(np.where(filled, cols, -1), axis=1)
    ri = np.minimum.accumulate(np.where(filled, cols, W)[:, ::-1], axis=1)[:, ::-1]
    rows = np.arange(H)[:, None]
    zl = np.where(li >= 0, zbuf[rows, np.clip(li, 0, W - 1)], np.inf)
    zr = np.where(ri < W, zbuf[rows, np.clip(ri, 0, W - 1)], np.inf)
    vl = right[rows, np.clip(li, 0, W - 1)]
    vr = right[rows, np.clip(ri, 0, W - 1)]
    fill = np.where(zl <= zr, vl, vr)
    fill = np.where(np.isinf(zl) & np.isinf(zr), 0.0, fill)
    right = np.where(filled, right, fill)
    return right, zbuf


def make_pair(W, H, seed, dmin=4.0, dmax=30.0, n_planes=None, n_boxes=0,
              noise_sigma=2.0, coarse_noise=6, texture="band"):
    """One synthetic pair. Returns a dict with uint8 ``left``/``right``
    (H, W), float64 ``disp`` (ground truth in the left view), bool ``valid``
    and float64 ``disp0`` (float32-representable coarse estimate).
    texture: "band" (band-limited fine noise) or "multiscale"."""
    rng = np.random.default_rng(seed)
    if n_planes is None:
        n_planes = int(rng.integers(3, 7))
    pad = int(np.ceil(dmax)) + 2
    Wv = W + pad
    tex = _texture(rng, H, Wv) if texture == "band" else _multiscale_texture(rng, H, Wv)
    disp_v = _plane_field(rng, H, Wv, dmin, dmax, n_planes)
    if n_boxes:
        disp_v = _add_boxes(rng, disp_v, n_boxes, dmax)
    right_v, zbuf_v = _warp_right(tex, disp_v)
    right_v = right_v + rng.normal(0.0, noise_sigma, size=right_v.shape)
    left = np.clip(np.rint(tex[:, :W]), 0, 255).astype(np.uint8)
    right = np.clip(np.rint(right_v[:, :W]), 0, 255).astype(np.uint8)
    disp = disp_v[:, :W].copy()
    xs = np.arange(W, dtype=np.float64)[None, :]
    xr = np.trunc(xs - disp + 0.5).astype(np.int64)
    inb = (xr >= 0) & (xr < W)
    zr = np.take_along_axis(zbuf_v, np.clip(xr, 0, Wv - 1), axis=1)
    valid = inb & (zr <= disp + 1e-6)
    if coarse_noise > 0:
        jitter = rng.integers(-coarse_noise, coarse_noise + 1, size=disp.shape)
    else:
        jitter = np.zeros(disp.shape, dtype=np.int64)
    disp0 = (disp + jitter).astype(np.float32).astype(np.float64)
    return {"left": left, "right": right, "disp": disp, "valid": valid,
            "disp0": disp0, "W": W, "H": H, "max_disp": float(dmax)}


def band_pair(seed, W=192, H=144):
    """The band-refinement-only workload: 192x144, disparity 4-30 px,
    3-6 planes, sigma 2 noise, coarse estimate = truth + U{-6..6}."""
    return make_pair(W, H, seed, dmin=4.0, dmax=30.0, n_boxes=0)


def pipeline_pair(seed, W=460, H=370):
    """The full-pipeline workload: 460x370, disparity 10-80 px, slanted
    planes plus 2-4 thin foreground boxes, sigma 2 noise."""
    rng = np.random.default_rng(10_000 + seed)
    nb = int(rng.integers(2, 5))
    return make_pair(W, H, seed, dmin=10.0, dmax=80.0, n_boxes=nb)


def gray(u8):
    """uint8 image -> float64 gray in [0, 1] the way the reference's PGM
    reader builds it (value times 1/maxval, stereonet.py:284,300)."""
    return u8.astype(np.float64) * (1.0 / 255.0)


def ragged_spec(n=64, seed=0):
    """Sizes and disparity ranges of an uneven batch: W ~ U{200..1390},
    H = round(u * W) with u ~ U[0.75, 0.80] clipped to [150, 1110], maximum
    disparity ~ U[16, 240] but at most W / 3 (a larger one would leave no
    overlap between the views of the narrow pairs).  Returns [(W, H, dmax)]."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        W = int(rng.integers(200, 1391))
        H = int(min(max(round(rng.uniform(0.75, 0.80) * W), 150), 1110))
        dmax = float(min(rng.uniform(16.0, 240.0), W / 3.0))
        out.append((W, H, dmax))
    return out


def ragged_pair(args):
    """One pair of an uneven batch: (W, H, dmax, seed) -> make_pair with
    slanted planes over [dmax / 4, dmax], two thin boxes and multiscale
    texture (large pairs are matched by the network 10-20x downscaled)."""
    W, H, dmax, seed = args
    return make_pair(W, H, seed, dmin=0.25 * dmax, dmax=dmax, n_boxes=2, texture="multiscale")
