"""GPU parity of plane snapping (stereonet.py:5394-5563, :5865-5884): SLIC
labels, plane_snap and refine_disparity(planes=True) bit-exact against the
reference's golden vectors and the oracle."""
import os

import numpy as np
import pytest
import torch

import oracle
from conftest import GOLDEN, package

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sv():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return package()


def _golden():
    return np.load(os.path.join(GOLDEN, "planes.npz"))


def test_slic_and_snap_golden(sv):
    """Reference outputs through lists, as its harnesses pass them."""
    z = _golden()
    for n in ("a", "b"):
        W, H = int(z[n + "_W"]), int(z[n + "_H"])
        rgb = list(map(float, z[n + "_rgb"]))
        for step in (12, 8):
            lab, k = sv.slic(rgb, W, H, step)
            assert k == int(z["%s_k%d" % (n, step)])
            assert np.array_equal(np.array(lab), z["%s_lab%d" % (n, step)]), (n, step)
        d = list(map(float, z[n + "_disp"]))
        s = sv.plane_snap(d, rgb, W, H)
        assert isinstance(s, list) and np.array_equal(np.array(s), z[n + "_snap"]), n
        s = sv.plane_snap(d, rgb, W, H, valid=list(map(bool, z[n + "_valid"])), min_inl=0.4)
        assert np.array_equal(np.array(s), z[n + "_snapv"]), n


def test_refine_disparity_planes_golden(sv):
    z = _golden()
    W, H = int(z["b_W"]), int(z["b_H"])
    d, st = sv.refine_disparity(list(map(float, z["b_left"])), list(map(float, z["b_right"])),
                                list(map(float, z["b_rgb"])), list(map(float, z["b_disp0"])), W, H,
                                planes=True)
    assert np.array_equal(np.array(d), z["b_refined"])
    assert np.array_equal(np.array(st), z["b_status"])


def _rgb_u8(rng, H, W, regions):
    """Colour blocks with noise: SLIC centres drift toward the blocks."""
    base = rng.integers(0, 256, size=(regions, 3))
    yy, xx = np.mgrid[0:H, 0:W]
    idx = ((xx * regions) // max(W, 1) + (yy * 3) // max(H, 1)) % regions
    img = base[idx] + rng.integers(-20, 21, size=(H, W, 3))
    return np.clip(img, 0, 255).astype(np.uint8)


@pytest.mark.parametrize("W,H,step", [(460, 370, 12), (101, 37, 12), (64, 48, 8), (7, 5, 12),
                                      (33, 90, 16), (50, 50, 3)])
def test_slic_matches_oracle(sv, W, H, step):
    """uint8 RGB (read as value * (1/255)), two images per launch, against the
    oracle: every label identical."""
    print("slic", W, H, step, flush=True)
    rng = np.random.default_rng(W + H + step)
    imgs = [_rgb_u8(rng, H, W, 5), rng.integers(0, 256, size=(H, W, 3)).astype(np.uint8)]
    t = torch.from_numpy(np.stack(imgs)).cuda()
    lab, k = sv.slic(t, W, H, step)
    lab = lab.reshape(2, -1).cpu().numpy()
    for b, img in enumerate(imgs):
        ref, rk = oracle.slic(img.astype(np.float64).ravel() * (1.0 / 255.0), W, H, step)
        assert rk == k
        assert np.array_equal(lab[b], ref), b


def _planar(rng, H, W):
    yy, xx = np.mgrid[0:H, 0:W].astype(np.float64)
    d = np.where(xx < W // 2, 8.0 + 0.03 * xx + 0.01 * yy, 20.0 - 0.02 * xx)
    d = d + rng.normal(0.0, 0.25, size=(H, W))
    d[rng.uniform(size=(H, W)) < 0.04] += 9.0
    return d


@pytest.mark.parametrize("W,H", [(460, 370), (97, 61), (30, 20)])
def test_plane_snap_matches_oracle(sv, W, H):
    """Planar disparity with noise and outliers, a colour image whose blocks
    follow the planes, with and without a validity mask, two maps per
    launch: bit-exact against the oracle."""
    print("plane_snap", W, H, flush=True)
    rng = np.random.default_rng(W * 3 + H)
    ds = [_planar(rng, H, W) for _ in range(2)]
    rgbs = [_rgb_u8(rng, H, W, 4) for _ in range(2)]
    valid = rng.uniform(size=(2, H, W)) > 0.1
    td = torch.from_numpy(np.stack(ds)).cuda()
    tr = torch.from_numpy(np.stack(rgbs)).cuda()
    for v in (None, valid):
        tv = None if v is None else torch.from_numpy(v).cuda()
        out = sv.plane_snap(td, tr, W, H, valid=tv).reshape(2, -1).cpu().numpy()
        for b in range(2):
            ref = oracle.plane_snap(ds[b], rgbs[b].astype(np.float64).ravel() * (1.0 / 255.0), W, H,
                                    valid=None if v is None else v[b])
            assert np.array_equal(out[b], ref), b


def test_refine_disparity_planes_matches_oracle(sv):
    """The 192x144 band workload with an RGB image: planes=True through the
    GPU equals the oracle's lr_refine + plane_snap + median."""
    p = sv.synth.band_pair(3)
    W, H = p["W"], p["H"]
    rng = np.random.default_rng(7)
    rgb = _rgb_u8(rng, H, W, 6)
    l8 = torch.from_numpy(p["left"]).cuda()
    r8 = torch.from_numpy(p["right"]).cuda()
    d0 = torch.from_numpy(np.asarray(p["disp0"], dtype=np.float64)).cuda()
    d, st = sv.refine_disparity(l8, r8, torch.from_numpy(rgb).cuda(), d0, W, H, planes=True)
    rd, rst = oracle.refine_disparity_planes(sv.synth.gray(p["left"]), sv.synth.gray(p["right"]),
                                             rgb.astype(np.float64).ravel() * (1.0 / 255.0),
                                             p["disp0"], W, H)
    assert np.array_equal(st.cpu().numpy().ravel().astype(np.int32), rst)
    assert np.array_equal(d.cpu().numpy().ravel(), rd)


def _stray_image(seed):
    """Images on which some labelled pixels fall outside every SLIC window in
    the later iterations (found by search): each keeps its earlier label,
    and its segment's member list must merge it in pixel order."""
    rng = np.random.default_rng(seed)
    step = int(rng.choice([4, 6, 8]))
    W, H = int(rng.integers(20, 60)), int(rng.integers(12, 40))
    if seed % 3 == 0:
        img = np.full((H, W, 3), 200.0)
        img[rng.uniform(size=(H, W)) < 0.08] = 20.0
    else:
        bs = int(rng.integers(2, 6))
        base = rng.integers(0, 256, size=((H + bs - 1) // bs, (W + bs - 1) // bs, 3))
        img = np.repeat(np.repeat(base, bs, 0), bs, 1)[:H, :W].astype(float)
    return img.astype(np.uint8), W, H, step


@pytest.mark.parametrize("seed", [9, 63, 88, 103, 169])
def test_slic_and_snap_with_strays_match_oracle(sv, seed):
    print("strays seed", seed, flush=True)
    img, W, H, step = _stray_image(seed)
    rgbf = img.astype(np.float64).ravel() * (1.0 / 255.0)
    lab, k = sv.slic(torch.from_numpy(img).cuda(), W, H, step)
    ref, rk = oracle.slic(rgbf, W, H, step)
    assert rk == k and np.array_equal(lab.reshape(-1).cpu().numpy(), ref)
    rng = np.random.default_rng(seed + 1)
    yy, xx = np.mgrid[0:H, 0:W].astype(np.float64)
    d = 5.0 + 0.1 * xx + 0.05 * yy + rng.normal(0.0, 0.2, size=(H, W))
    out = sv.plane_snap(torch.from_numpy(d).cuda(), torch.from_numpy(img).cuda(), W, H, step=step)
    assert np.array_equal(out.reshape(-1).cpu().numpy(), oracle.plane_snap(d, rgbf, W, H, step=step))
