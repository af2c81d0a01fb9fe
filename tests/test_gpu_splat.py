"""The Gaussian half of the downstream 3D output on the GPU against the
reference's own outputs (tests/golden/splat.npz, written by make_splat.py):
gaussians_from_mesh, the splat renderer's forward (original and a novel
viewpoint with a background) and backward, CloudAdam, refine_gaussians,
orbit_views, background_layer / merge_layer and reconstruct.  Everything is
float64 with the reference's own series for sqrt / exp / log, so the
comparisons are exact except the reported loss (a sum in a different
order).  The backward is also checked against central differences of the
GPU forward on the reference test's own random cloud
(test_stereonet.py:219-250, relative error below 1e-5)."""

import os

import numpy as np
import pytest
import torch

import oracle
from conftest import GOLDEN, package

pytestmark = pytest.mark.gpu
G = np.load(os.path.join(GOLDEN, "splat.npz"))
SCENES = ["s48", "s64"]


@pytest.fixture(scope="module")
def sv():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return package()


def _setup(sv, name):
    W, H, npts = (int(v) for v in G[name + "_wh"])
    cam = sv.Camera.synthetic(W, H)
    disp = torch.from_numpy(G[name + "_disp"]).cuda()
    p3 = sv.geometry.disparity_to_points_gpu(disp, W, H, cam, min_disp=0.4)
    pts2 = [tuple(p) for p in G[name + "_sel"].tolist()]
    tris = [tuple(t) for t in G[name + "_tris"].tolist()]
    rgb = torch.from_numpy(G[name + "_rgb"]).cuda()
    return W, H, npts, cam, disp, p3, pts2, tris, rgb


def _eq(a, b):
    a = a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape, (a.shape, b.shape)
    if np.array_equal(a, b):
        return
    d = np.abs(a.astype(np.float64) - b.astype(np.float64))
    raise AssertionError("max |diff| %.3g at %d of %d" % (d.max(), (d > 0).sum(), d.size))


@pytest.mark.parametrize("name", SCENES)
def test_cloud_render_backward_adam_match_reference(sv, name):
    W, H, npts, cam, disp, p3, pts2, tris, rgb = _setup(sv, name)
    print("scene", name, W, H, len(tris), flush=True)
    gc = sv.gaussians_from_mesh(pts2, tris, p3, rgb, W, H)
    _eq(gc.arrays(), G[name + "_gc"])
    r = sv.SplatRenderer(cam, W, H)
    C = r.render(gc, sv.IDENTITY_R, sv.IDENTITY_T)
    _eq(C, G[name + "_render"])
    gc.zero_grads()
    r.render(gc, sv.IDENTITY_R, sv.IDENTITY_T)
    r.backward(torch.from_numpy(G[name + "_bw_w"]).cuda())
    _eq(gc.grads[:, :gc.count()], G[name + "_grads"])
    opt = sv.CloudAdam(gc)
    opt.step()
    _eq(gc.arrays(), G[name + "_adam1"])
    views = sv.orbit_views(gc, n_views=3)
    _eq(np.array([[a] + list(Rm) + list(t) for a, (Rm, t) in views]), G[name + "_views"])
    Rm, t = views[0][1]
    _eq(r.render(gc, Rm, t, background=(0.1, 0.2, 0.3)), G[name + "_render_novel"])


@pytest.mark.parametrize("name", SCENES)
def test_refine_gaussians_matches_reference(sv, name):
    W, H, npts, cam, disp, p3, pts2, tris, rgb = _setup(sv, name)
    gc = sv.gaussians_from_mesh(pts2, tris, p3, rgb, W, H, thickness=0.5, init_opacity=0.8,
                                scale=1.3)
    _eq(gc.arrays(), G[name + "_gc_kw"])
    hist = sv.refine_gaussians(gc, sv.SplatRenderer(cam, W, H), rgb, iters=3, verbose=False)
    assert np.allclose(hist, G[name + "_refine_hist"], rtol=1e-12, atol=0)
    _eq(gc.arrays(), G[name + "_refine_gc"])


@pytest.mark.parametrize("name", SCENES)
def test_background_layer_and_merge_match_reference(sv, name):
    W, H, npts, cam, disp, p3, pts2, tris, rgb = _setup(sv, name)
    lp, lt, l3, lrgb = sv.background_layer(disp, rgb, W, H, cam, npts)
    assert lp == [tuple(p) for p in G[name + "_bgl_pts"].tolist()]
    assert lt == [tuple(t) for t in G[name + "_bgl_tris"].tolist()]
    _eq(lrgb, G[name + "_bgl_rgb"])
    want = G[name + "_bgl_p3"]
    ok = ~np.isnan(want[:, 0])
    _eq(l3[1].cpu().numpy().astype(bool), ok)
    _eq(l3[0].cpu().numpy()[ok], want[ok])
    mp, mt, (P, V), C, W2 = sv.merge_layer(pts2, tris, p3, rgb, (lp, lt, l3, lrgb), W, H)
    assert W2 == 2 * W
    assert mp == [tuple(p) for p in G[name + "_merge_pts"].tolist()]
    assert mt == [tuple(t) for t in G[name + "_merge_tris"].tolist()]
    wm = G[name + "_merge_p3"]
    okm = ~np.isnan(wm[:, 0])
    _eq(V.cpu().numpy().astype(bool), okm)
    _eq(P.cpu().numpy()[okm], wm[okm])
    _eq(C, G[name + "_merge_rgb"])


@pytest.mark.parametrize("name", SCENES)
def test_reconstruct_matches_reference(sv, name, tmp_path):
    W, H, npts, cam, disp, p3, pts2, tris, rgb = _setup(sv, name)
    res = sv.reconstruct(disp, rgb, W, H, cam, n_points=npts, refine_iters=2,
                         out_prefix=str(tmp_path / "sc"), n_views=2, verbose=False)
    _eq(res["render"], G[name + "_rec_render"])
    _eq(res["gaussians"].arrays(), G[name + "_rec_gc"])
    assert tuple(res["mesh"][2:]) == tuple(G[name + "_rec_mesh"].tolist())
    for f in ("sc_mesh.obj", "sc_render_original.ppm", "sc_novel_0.ppm", "sc_novel_1.ppm"):
        assert (tmp_path / f).exists()


class _Rng:
    """xorshift64* of the reference's tests (stereonet.py:149-173)."""
    M = 0xFFFFFFFFFFFFFFFF

    def __init__(self, seed):
        self.s = seed & self.M

    def uniform(self):
        s = self.s
        s ^= (s << 13) & self.M
        s ^= s >> 7
        s ^= (s << 17) & self.M
        self.s = s
        return (((s * 2685821657736338717) & self.M) >> 11) * (1.0 / 9007199254740992.0)

    def range(self, lo, hi):
        return lo + (hi - lo) * self.uniform()


def test_backward_matches_central_differences(sv):
    """test_renderer_gradients (test_stereonet.py:219-250) on the GPU."""
    rng = _Rng(777)
    W, H = 20, 16
    cam = sv.Camera.synthetic(W, H, baseline=0.1)
    cols = []
    for _ in range(5):
        sc = rng.range(0.004, 0.02)
        mx, my, mz = rng.range(-0.3, 0.3), rng.range(-0.25, 0.25), rng.range(0.6, 1.8)
        sxx, syy, szz = sc * rng.range(0.6, 1.4), sc * rng.range(0.6, 1.4), sc * 0.2
        sxy, sxz, syz = sc * rng.range(-0.2, 0.2), sc * rng.range(-0.1, 0.1), sc * rng.range(-0.1, 0.1)
        cr, cg, cb = rng.range(0.1, 0.9), rng.range(0.1, 0.9), rng.range(0.1, 0.9)
        rop, rsc = rng.range(-0.5, 1.5), rng.range(-0.3, 0.3)
        cols.append([mx, my, mz, sxx, sxy, sxz, syy, syz, szz, cr, cg, cb, rop, rsc])
    gc = sv.GaussianCloud.from_arrays(np.array(cols).T)
    r = sv.SplatRenderer(cam, W, H)
    wts = torch.tensor([0.3 + 0.013 * i for i in range(3 * W * H)], dtype=torch.float64).cuda()

    def loss():
        return float((r.render(gc, sv.IDENTITY_R, sv.IDENTITY_T) * wts).sum())

    gc.zero_grads()
    loss()
    r.backward(wts)
    ana = gc.grads[:, :5].cpu().numpy()
    worst = 0.0
    for i in range(5):
        for row, g in ((9, 0), (12, 3), (13, 4)):  # cr, raw_op, raw_sc
            p0 = float(gc.params[row, i])
            gc.params[row, i] = p0 + 1e-6
            lp = loss()
            gc.params[row, i] = p0 - 1e-6
            lm = loss()
            gc.params[row, i] = p0
            num = (lp - lm) / 2e-6
            a = float(ana[g, i])
            worst = max(worst, abs(num - a) / max(abs(num), abs(a), 1e-12))
    print("renderer gradients: max rel err %.2e" % worst)
    assert worst < 1e-5
