"""Disparity to 3D on the CPU: the oracle's Camera arithmetic,
robust_min_disparity, disparity_to_points, min_filter_arg,
smooth_background, background_region and grid_mesh against the reference's
own outputs (tests/golden/geometry.npz, made by tests/golden/make_geometry.py),
closed forms from the reference's own tests, and planted mistakes."""
import os

import numpy as np

import oracle
from conftest import GOLDEN, package

GRIDS = (("g2f", dict(step=2, fill=True), "syn"),
         ("g1n", dict(step=1, fill=False, min_disp=0.5), "mid"),
         ("g3f", dict(step=3, fill=True, jump=0.8, jump_rel=0.1), "mid"))


def _golden():
    return np.load(os.path.join(GOLDEN, "geometry.npz"))


def _eq(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and np.array_equal(a, b, equal_nan=a.dtype.kind == "f")


def test_oracle_matches_reference():
    z = _golden()
    for c in z["cases"]:
        W, H = (int(v) for v in z[c + "_wh"])
        d = z[c + "_disp"]
        assert oracle.robust_min_disparity(d) == float(z[c + "_robust"])
        for cn in ("syn", "mid"):
            p, v = oracle.disparity_to_points(d, W, H, z[c + "_cam_" + cn])
            assert _eq(p, z[c + "_pts_" + cn]) and _eq(v, z[c + "_ptsvalid_" + cn])
        m, i = oracle.min_filter_arg(d, W, H, 3, 2)
        assert _eq(m, z[c + "_mfa_m"]) and _eq(i, z[c + "_mfa_i"])
        m0, i0 = oracle.min_filter_arg(d, W, H, 5, 0)
        assert _eq(m0, z[c + "_mfa0_m"]) and _eq(i0, z[c + "_mfa0_i"])
        assert _eq(oracle.smooth_background(m0, d, W, H, 2), z[c + "_smooth"])
        mk, bm, bs = oracle.background_region(d, W, H)
        assert _eq(mk, z[c + "_bg_mask"]) and _eq(bm, z[c + "_bg_m"]) and _eq(bs, z[c + "_bg_src"])
        for gname, kw, cn in GRIDS:
            g = oracle.grid_mesh(d, W, H, z[c + "_cam_" + cn], **kw)
            k = "%s_%s_" % (c, gname)
            for f in ("pos", "uv", "fill", "disp", "idx"):
                assert _eq(g[f].ravel(), z[k + f].ravel()), (k, f)
            assert [g["main_tris"], g["fill_tris"]] == list(z[k + "counts"])
            assert g["pivot"] == float(z[k + "pivot"])


def test_reference_closed_forms():
    """test_stereonet.py:535-565: a near square in front of a far wall --
    the background layer covers the square's edges with the wall's
    disparity and leaves the open wall alone; the grid never bridges the
    depth edge; a vertex's texture coordinate sits on its own pixel."""
    W, H = 100, 64
    step = np.full(W * H, 4.0)
    for y in range(20, 44):
        step[y * W + 40:y * W + 60] = 16.0
    mask, m, src = oracle.background_region(step, W, H)
    inside = [y * W + x for y in range(20, 44) for x in (40, 41, 58, 59)]
    assert mask[inside].all() and np.all(np.abs(m[inside] - 4.0) < 1e-9)
    assert not mask[5 * W + 5] and not mask[60 * W + 90]
    cam = (W * 0.5 / 0.5773502691896257, W * 0.5, H * 0.5, 0.1, 0.0)
    g = oracle.grid_mesh(step, W, H, cam, step=2, fill=False, min_disp=0.5)
    ds = g["disp"][g["idx"]]
    assert (ds.max(1) - ds.min(1) <= 5.0).all() and g["main_tris"] > 0
    k = len(g["disp"]) // 2
    x, y = int(g["uv"][k, 0] * W), int(g["uv"][k, 1] * H)
    Z = float(oracle._depth(cam, g["disp"][k]))
    assert abs((x - cam[1]) * Z / cam[0] - g["pos"][k, 0]) < 1e-9


def test_planted_mistakes_fail():
    """A min filter that keeps the last minimum instead of the first, or a
    grid that emits a cell's triangles as (a, b, c), no longer matches."""
    z = _golden()
    c = "w80"
    W, H = (int(v) for v in z[c + "_wh"])
    d = np.round(z[c + "_disp"])  # ties, so first and last minima differ
    v2 = d.reshape(H, W)
    last = np.array([[max(range(max(x - 5, 0), min(x + 5, W - 1) + 1),
                          key=lambda xx: (-v2[y, xx], xx)) for x in range(W)] for y in range(H)])
    first = oracle.min_filter_arg(d, W, H, 5, 0)[1].reshape(H, W) % W
    assert not np.array_equal(last, first)
    g = oracle.grid_mesh(z[c + "_disp"], W, H, z[c + "_cam_syn"], step=2, fill=True)
    ref = z[c + "_g2f_idx"].reshape(-1, 3)
    swapped = g["idx"].copy()
    swapped[0::2] = swapped[0::2][:, [0, 2, 1]]
    assert np.array_equal(g["idx"], ref) and not np.array_equal(swapped, ref)


def test_read_calib(tmp_path):
    geo = package().geometry
    p = tmp_path / "calib.txt"
    p.write_text("cam0=[3979.911 0 1244.772; 0 3979.911 1019.507; 0 0 1]\n"
                 "cam1=[3979.911 0 1369.115; 0 3979.911 1019.507; 0 0 1]\n"
                 "doffs=124.343\nbaseline=193.001\nwidth=2964\n")
    c = geo.read_calib(str(p))
    assert c == {"fx": 3979.911, "cx": 1244.772, "cy": 1019.507, "baseline": 193.001,
                 "doffs": 124.343}
    cam = geo.Camera.from_calib(str(p))
    assert cam.depth_from_disparity(100.0) == 3979.911 * 193.001 / (100.0 + 124.343)
    assert cam.depth_from_disparity(-124.343) == 0.0
