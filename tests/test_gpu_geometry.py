"""GPU parity of disparity-to-3D (stereonet.py:1951-2021, :5566-5862):
robust_min_disparity, disparity_to_points, min_filter_arg,
smooth_background, background_region and grid_mesh bit-exact against the
reference's own outputs (tests/golden/geometry.npz) and, at frame size,
against the oracle."""
import os

import numpy as np
import pytest
import torch

import oracle
from conftest import GOLDEN, package

pytestmark = pytest.mark.gpu
GRIDS = (("g2f", dict(step=2, fill=True), "syn"),
         ("g1n", dict(step=1, fill=False, min_disp=0.5), "mid"),
         ("g3f", dict(step=3, fill=True, jump=0.8, jump_rel=0.1), "mid"))


@pytest.fixture(scope="module")
def sv():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return package()


def _golden():
    return np.load(os.path.join(GOLDEN, "geometry.npz"))


def _eq(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return a.shape == b.shape and np.array_equal(a, b, equal_nan=True)


def _cam(sv, c):
    return sv.Camera(*[float(v) for v in c])


def test_golden(sv):
    z = _golden()
    for c in z["cases"]:
        W, H = (int(v) for v in z[c + "_wh"])
        d = z[c + "_disp"]
        dl = d.tolist()
        assert sv.robust_min_disparity(dl) == float(z[c + "_robust"])
        for cn in ("syn", "mid"):
            cam = _cam(sv, z[c + "_cam_" + cn])
            pts = sv.disparity_to_points(dl, W, H, cam)
            ref = z[c + "_pts_" + cn]
            ok = z[c + "_ptsvalid_" + cn]
            assert [p is not None for p in pts] == ok.tolist()
            assert all(tuple(r) == p for r, p, k in zip(ref.tolist(), pts, ok) if k)
        m, i = sv.min_filter_arg(dl, W, H, 3, 2)
        assert _eq(m, z[c + "_mfa_m"]) and list(i) == z[c + "_mfa_i"].tolist()
        m0, i0 = sv.min_filter_arg(dl, W, H, 5, 0)
        assert _eq(m0, z[c + "_mfa0_m"]) and list(i0) == z[c + "_mfa0_i"].tolist()
        assert _eq(sv.smooth_background(m0, dl, W, H, 2), z[c + "_smooth"])
        mk, bm, bs = sv.background_region(dl, W, H)
        assert mk == z[c + "_bg_mask"].tolist() and _eq(bm, z[c + "_bg_m"])
        assert bs == z[c + "_bg_src"].tolist()
        for gname, kw, cn in GRIDS:
            g = sv.grid_mesh(dl, W, H, _cam(sv, z[c + "_cam_" + cn]), **kw)
            k = "%s_%s_" % (c, gname)
            for f in ("pos", "uv", "disp"):
                assert _eq(g[f], z[k + f].ravel()), (k, f)
            assert g["fill"] == z[k + "fill"].ravel().tolist()
            assert g["idx"] == z[k + "idx"].ravel().tolist(), k
            assert [g["main_tris"], g["fill_tris"]] == list(z[k + "counts"])
            assert g["pivot"] == float(z[k + "pivot"])


def test_batch_background_region(sv):
    """Three maps per launch equal three single calls and the oracle."""
    rng = np.random.default_rng(4)
    W, H = 97, 53
    maps = [np.where(rng.random((H, W)) < 0.1, 20.0, 5.0) + rng.normal(0, 0.3, (H, W))
            for _ in range(3)]
    t = torch.from_numpy(np.stack(maps)).cuda()
    mk, m, src = sv.geometry.background_region_gpu(t, W, H)
    for b in range(3):
        m1, s1, r1 = oracle.background_region(maps[b].ravel(), W, H)
        sl = slice(b * W * H, (b + 1) * W * H)
        assert np.array_equal(mk[sl].cpu().numpy().astype(bool), m1)
        assert np.array_equal(m[sl].cpu().numpy(), s1)
        assert np.array_equal(src[sl].cpu().numpy(), r1)


@pytest.mark.parametrize("step,fill", [(2, True), (1, False), (4, True)])
def test_frame_sized_vs_oracle(sv, step, fill):
    p = sv.synth.pipeline_pair(3)
    W, H = 460, 370
    d = np.asarray(p["disp"], np.float64).ravel() + np.random.default_rng(5).normal(0, 0.1, W * H)
    cam = sv.Camera.synthetic(W, H)
    g = sv.geometry.grid_mesh_gpu(torch.from_numpy(d).cuda(), W, H, cam, step=step, fill=fill)
    r = oracle.grid_mesh(d, W, H, (cam.fx, cam.cx, cam.cy, cam.baseline, cam.doffs), step=step,
                         fill=fill)
    for f in ("pos", "uv", "disp"):
        assert np.array_equal(g[f].cpu().numpy(), r[f]), f
    assert np.array_equal(g["fill"].cpu().numpy(), r["fill"])
    assert np.array_equal(g["idx"].cpu().numpy(), r["idx"])
    assert (g["main_tris"], g["fill_tris"], g["pivot"]) == (r["main_tris"], r["fill_tris"],
                                                            r["pivot"])
    assert sv.robust_min_disparity(torch.from_numpy(d).cuda()) == oracle.robust_min_disparity(d)


def test_kth_smallest(sv):
    rng = np.random.default_rng(8)
    x = np.concatenate([rng.normal(0, 1e3, 100001), [-0.0, 0.0, np.inf, -np.inf, 1e-310]])
    t = torch.from_numpy(x).cuda()
    s = np.sort(x)
    for k in (0, 1, 77, x.size // 2, x.size - 2, x.size - 1):
        assert float(sv.geometry.kth_smallest(t, k)) == s[k]
    # strided: every third value
    v = float(sv.geometry.kth_smallest(t, 1000, stride=3, n=x.size // 3))
    assert v == np.sort(x[0:3 * (x.size // 3):3])[1000]


def test_points_tensor_and_errors(sv):
    W, H = 31, 7
    d = torch.linspace(-1.0, 40.0, W * H, dtype=torch.float64).cuda()
    cam = sv.Camera(50.0, 15.0, 3.0, 0.2, 1.5)
    pts, valid = sv.disparity_to_points(d, W, H, cam, min_disp=0.5, max_depth=9.0)
    p0, v0 = oracle.disparity_to_points(d.cpu().numpy(), W, H, (50.0, 15.0, 3.0, 0.2, 1.5), 0.5,
                                        9.0)
    assert np.array_equal(valid.cpu().numpy().astype(bool), v0)
    assert np.array_equal(pts.cpu().numpy(), p0, equal_nan=True)
    with pytest.raises(ValueError):
        sv.grid_mesh(d, W + 1, H, cam)
    with pytest.raises(ValueError):
        sv.min_filter_arg(d, 5, 5, 2)


@pytest.mark.parametrize("W,H", [(4103, 3), (460, 5)])
def test_row_filters_both_paths(sv, W, H):
    """Rows wider than the LDS stage (4096) take the global-memory kernels;
    both give the oracle's values, with ties resolved to the first minimum."""
    rng = np.random.default_rng(W)
    v = np.round(rng.normal(10, 3, W * H))
    t = torch.from_numpy(v).cuda()
    for r, rv in ((7, 0), (7, 1), (28, 0)):
        m, i = sv.min_filter_arg(t, W, H, r, rv)
        m0, i0 = oracle.min_filter_arg(v, W, H, r, rv)
        assert np.array_equal(m.cpu().numpy(), m0) and np.array_equal(i.cpu().numpy(), i0)
    m0, _ = oracle.min_filter_arg(v, W, H, 9, 0)
    s = sv.smooth_background(torch.from_numpy(m0).cuda(), t, W, H, 4)
    assert np.array_equal(s.cpu().numpy(), oracle.smooth_background(m0, v, W, H, 4))


def test_batch_equals_single(sv):
    """grid_mesh_batch over five maps equals five grid_mesh_gpu calls: the
    maps differ in their vertex and triangle counts, one has no usable
    pixel (every count 0, pivot 1.0)."""
    W, H = 83, 61
    rng = np.random.default_rng(12)
    maps = []
    for b in range(5):
        p = sv.synth.band_pair(20 + b, W=W, H=H)
        maps.append(np.asarray(p["disp"], np.float64) + rng.normal(0, 0.1, (H, W)))
    maps[3][:] = 0.0  # nothing above the floor
    D = torch.from_numpy(np.stack(maps)).cuda()
    cam = sv.Camera.synthetic(W, H)
    res = sv.geometry.grid_mesh_batch(D, W, H, cam, step=2, fill=True)
    for b in range(5):
        one = sv.geometry.grid_mesh_gpu(D[b], W, H, cam, step=2, fill=True)
        ref = oracle.grid_mesh(maps[b].ravel(), W, H, (cam.fx, cam.cx, cam.cy, cam.baseline,
                                                       cam.doffs), step=2, fill=True)
        for f in ("pos", "uv", "fill", "disp", "idx"):
            assert torch.equal(res[b][f], one[f]), (b, f)
            assert np.array_equal(res[b][f].cpu().numpy(), ref[f]), (b, f)
        assert (res[b]["main_tris"], res[b]["fill_tris"], res[b]["pivot"]) == \
            (ref["main_tris"], ref["fill_tris"], ref["pivot"])
    assert res[3]["pos"].shape[0] == 0 and res[3]["pivot"] == 1.0


def test_row_min_with_nans(sv):
    """NaNs follow the scan's rules: a NaN where a window starts sticks, a
    NaN later in the window is passed over; the oracle's scan agrees."""
    W, H = 300, 6
    rng = np.random.default_rng(31)
    v = np.round(rng.normal(10, 3, (H, W)))
    v[1, 0] = np.nan  # the scan's start: sticks for the windows that begin there
    v[3, 150] = np.nan
    v[5, [7, 8, 299]] = np.nan
    t = torch.from_numpy(v.ravel()).cuda()
    for r in (1, 5, 28, 400):
        m, i = sv.min_filter_arg(t, W, H, r, 0)
        m0, i0 = oracle.min_filter_arg(v.ravel(), W, H, r, 0)
        assert np.array_equal(m.cpu().numpy(), m0, equal_nan=True)
        assert np.array_equal(i.cpu().numpy(), i0)


@pytest.mark.parametrize("r", [31, 32, 33, 40])
@pytest.mark.parametrize("W,H", [(300, 70), (97, 45)])
def test_smooth_background_both_column_paths(sv, r, W, H):
    """Column smoothing runs from an LDS tile for r <= 32 and from global
    memory beyond (background_region's radius reaches 33 at W >= 1056):
    both paths equal the oracle, on heights that are not multiples of 32."""
    rng = np.random.default_rng(r * 1000 + W)
    v = np.round(rng.normal(10, 3, W * H))
    t = torch.from_numpy(v).cuda()
    m0, _ = oracle.min_filter_arg(v, W, H, 9, 0)
    s = sv.smooth_background(torch.from_numpy(m0).cuda(), t, W, H, r)
    assert np.array_equal(s.cpu().numpy(), oracle.smooth_background(m0, v, W, H, r))
