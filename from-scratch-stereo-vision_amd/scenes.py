"""Ray-cast synthetic stereo scenes with exact ground truth.

The scene model follows the reference's rendered-dataset generator
(make_dataset.py:52-205, make :219-227): two pinhole cameras on a horizontal
baseline look at a textured ground plane, a textured back wall, 3-6 spheres
and 2-4 boxes, shaded with one Lambert light plus ambient.  Depth is
rendered first and the baseline is then chosen so that the nearest visible
point lands at the requested peak disparity (B = d_target * Zmin / f), and
both views are rendered independently, so occlusions and perspective are
real rather than warped.  Disparity = f * B / Z for every pixel of the left
view that hits geometry (NaN for rays that escape).

This is test and training data, not part of the matching path: it runs on
the host in numpy.  Textures are multi-octave value noise (random lattices
upsampled with Catmull-Rom cubics); images are returned as uint8 gray
(0.299 r + 0.587 g + 0.114 b).
"""

import numpy as np


def _cubic_matrix(n_in, n_out):
    """[n_out][n_in] Catmull-Rom interpolation weights resampling n_in
    lattice values onto n_out points spanning the same extent (edge
    indices clamped)."""
    pos = (np.arange(n_out) + 0.5) * (n_in / float(n_out)) - 0.5
    i0 = np.floor(pos).astype(int)
    t = pos - i0
    w = np.stack([(-t ** 3 + 2 * t ** 2 - t) / 2, (3 * t ** 3 - 5 * t ** 2 + 2) / 2,
                  (-3 * t ** 3 + 4 * t ** 2 + t) / 2, (t ** 3 - t ** 2) / 2], -1)
    m = np.zeros((n_out, n_in))
    rows = np.arange(n_out)
    for j in range(4):
        np.add.at(m, (rows, np.clip(i0 + j - 1, 0, n_in - 1)), w[:, j])
    return m


def value_noise(rng, size, octaves, persistence=0.55):
    """Sum of `octaves` random lattices (5, 9, 17, ... points) resampled to
    size x size with Catmull-Rom cubics, amplitudes falling by
    `persistence`, normalised by the amplitude total and clipped to [0, 1]."""
    acc = np.zeros((size, size))
    amp, total, cells = 1.0, 0.0, 4
    for _ in range(octaves):
        lattice = rng.random((cells + 1, cells + 1))
        m = _cubic_matrix(cells + 1, size)
        acc += amp * np.clip(m @ lattice @ m.T, 0.0, 1.0)
        total += amp
        amp *= persistence
        cells *= 2
    return acc / total


class Scene:
    """One random scene (make_dataset.py:52-99): camera focal length from a
    52-68 degree horizontal field of view, ground at y = -0.85, wall at
    z in [2.6, 4.0], spheres and boxes standing on the ground at 0.8-2.3 m."""

    def __init__(self, rng, W, H):
        self.W, self.H = W, H
        half_fov = np.deg2rad(rng.uniform(26.0, 34.0))
        self.f = (W * 0.5) / np.tan(half_fov)
        self.cx, self.cy = W / 2.0, H / 2.0
        self.ground_y = -0.85
        self.wall_z = rng.uniform(2.6, 4.0)
        self.spheres = []
        for _ in range(int(rng.integers(3, 7))):
            r = rng.uniform(0.20, 0.50)
            c = np.array([rng.uniform(-0.95, 0.95), self.ground_y + r, rng.uniform(0.85, 2.30)])
            self.spheres.append((c, r, rng.uniform(0.15, 0.97, 3)))
        self.boxes = []
        for _ in range(int(rng.integers(2, 5))):
            half = rng.uniform(0.14, 0.48, 3)
            c = np.array([rng.uniform(-1.0, 1.0), self.ground_y + half[1], rng.uniform(0.80, 2.30)])
            self.boxes.append((c - half, c + half, rng.uniform(0.15, 0.97, 3)))
        self.tex_ground = value_noise(rng, 512, int(rng.integers(3, 6)))
        self.tex_wall = value_noise(rng, 512, int(rng.integers(3, 6)))
        self.tex_sphere = value_noise(rng, 256, int(rng.integers(4, 7)))
        self.ground_scale = rng.uniform(0.8, 2.6)
        self.wall_scale = rng.uniform(0.25, 0.9)
        self.ground_contrast = rng.uniform(0.10, 0.45)
        self.wall_contrast = rng.uniform(0.08, 0.40)
        self.sphere_contrast = rng.uniform(0.15, 0.55)
        light = rng.normal(size=3) * np.array([1.0, 0.6, 0.7]) + np.array([0.0, 0.9, -0.5])
        self.light = light / np.linalg.norm(light)
        self.ambient = rng.uniform(0.22, 0.40)

    @staticmethod
    def _lookup(tex, u, v):
        """Nearest texel at wrapped (u, v); rays that miss (non-finite
        coordinates) read texel 0 and are discarded by the depth test."""
        n = tex.shape[0] - 1
        u = np.where(np.isfinite(u), u, 0.0)
        v = np.where(np.isfinite(v), v, 0.0)
        return tex[((v % 1.0) * n).astype(int), ((u % 1.0) * n).astype(int)]

    def render(self, eye_x):
        """Render from a camera at (eye_x, 0, 0) looking down +z: returns
        (rgb [H][W][3] in [0, 1], z depth [H][W] with NaN for misses).
        Rays that miss a surface carry inf/NaN hit points; the depth test
        discards them, so those floating-point warnings are silenced."""
        with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
            return self._render(eye_x)

    def _render(self, eye_x):
        H, W = self.H, self.W
        ys, xs = np.mgrid[0:H, 0:W].astype(np.float64)
        d = np.stack([(xs - self.cx) / self.f, -(ys - self.cy) / self.f, np.ones_like(xs)], -1)
        d /= np.linalg.norm(d, axis=-1, keepdims=True)
        o = np.array([eye_x, 0.0, 0.0])
        t_best = np.full((H, W), np.inf)
        albedo = np.zeros((H, W, 3))
        normal = np.zeros((H, W, 3))

        def take(t, alb, nrm):
            hit = np.isfinite(t) & (t > 0) & (t < t_best)
            t_best[hit] = t[hit]
            albedo[hit] = alb[hit] if alb.ndim == 3 else alb
            normal[hit] = nrm[hit] if nrm.ndim == 3 else nrm
            return hit

        # ground plane y = ground_y, in front of the wall
        t = np.where(d[..., 1] < -1e-6, (self.ground_y - o[1]) / d[..., 1], np.inf)
        p = o + t[..., None] * d
        t = np.where((p[..., 2] > 0) & (p[..., 2] < self.wall_z), t, np.inf)
        g = self._lookup(self.tex_ground, p[..., 0] * self.ground_scale, p[..., 2] * self.ground_scale)
        base = np.stack([0.40 + self.ground_contrast * g, 0.38 + self.ground_contrast * g,
                         0.35 + self.ground_contrast * g], -1)
        take(t, base, np.array([0.0, 1.0, 0.0]))
        # back wall z = wall_z, above the ground
        t = np.where(d[..., 2] > 1e-6, (self.wall_z - o[2]) / d[..., 2], np.inf)
        p = o + t[..., None] * d
        t = np.where(p[..., 1] > self.ground_y, t, np.inf)
        g = self._lookup(self.tex_wall, (p[..., 0] + 3.0) * self.wall_scale,
                         (p[..., 1] + 2.0) * self.wall_scale)
        base = np.stack([0.55 + self.wall_contrast * g, 0.54 + self.wall_contrast * g,
                         0.52 + self.wall_contrast * g], -1)
        take(t, base, np.array([0.0, 0.0, -1.0]))
        for c, r, alb in self.spheres:
            oc = c - o
            tca = d @ oc
            d2 = oc @ oc - tca * tca
            inside = d2 <= r * r
            t = np.where(inside, tca - np.sqrt(np.maximum(r * r - d2, 0.0)), np.inf)
            p = o + t[..., None] * d
            n = (p - c) / r
            u = np.arctan2(n[..., 2], n[..., 0]) / (2 * np.pi) + 0.5
            v = np.arccos(np.clip(n[..., 1], -1.0, 1.0)) / np.pi
            g = self._lookup(self.tex_sphere, u, v)
            k = 1.0 - self.sphere_contrast + 2.0 * self.sphere_contrast * g
            take(t, alb * k[..., None], n)
        for lo, hi, alb in self.boxes:
            inv = 1.0 / np.where(np.abs(d) < 1e-9, 1e-9, d)
            t0, t1 = (lo - o) * inv, (hi - o) * inv
            tmin = np.minimum(t0, t1).max(-1)
            tmax = np.maximum(t0, t1).min(-1)
            t = np.where(tmax >= np.maximum(tmin, 0.0), tmin, np.inf)
            p = o + t[..., None] * d
            n = np.zeros_like(p)
            for a in range(3):
                n[..., a] = np.where(np.abs(p[..., a] - lo[a]) < 1e-3, -1.0,
                                     np.where(np.abs(p[..., a] - hi[a]) < 1e-3, 1.0, 0.0))
            g = self._lookup(self.tex_wall, p[..., 0] * 2.0, p[..., 1] * 2.0)
            take(t, alb * (0.82 + 0.36 * g)[..., None], n)
        nl = np.linalg.norm(normal, axis=-1, keepdims=True)
        normal = normal / np.where(nl < 1e-9, 1.0, nl)
        lam = np.clip(normal @ self.light, 0.0, 1.0)
        rgb = np.clip(albedo * (self.ambient + (1.0 - self.ambient) * lam)[..., None], 0.0, 1.0)
        z = np.where(np.isfinite(t_best), t_best * d[..., 2], np.nan)
        return rgb, z


def _gray_u8(rgb):
    g = rgb[..., 0] * 0.299 + rgb[..., 1] * 0.587 + rgb[..., 2] * 0.114
    return np.clip(np.round(g * 255.0), 0, 255).astype(np.uint8)


def render_pair(seed, W, H, peak_lo=22.0, peak_hi=40.0):
    """A rendered pair (make_dataset.py:219-227 with the target peak drawn
    from U[peak_lo, peak_hi] by its own seeded generator): returns dict with
    left/right uint8 [H][W], disp float64 [H][W] (NaN where no geometry),
    valid bool [H][W]."""
    target = np.random.default_rng(seed).uniform(peak_lo, peak_hi)
    rng = np.random.default_rng(seed)
    sc = Scene(rng, W, H)
    rgb_l, z = sc.render(0.0)
    zmin = np.nanmin(z)
    baseline = target * zmin / sc.f
    rgb_r, _ = sc.render(baseline)
    disp = sc.f * baseline / z
    valid = np.isfinite(disp)
    return {"left": _gray_u8(rgb_l), "right": _gray_u8(rgb_r), "disp": disp, "valid": valid,
            "W": W, "H": H, "peak": float(target)}
