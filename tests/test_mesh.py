"""The mesh half of the downstream 3D output on the host (no GPU):
delaunay, select_points, cull_triangles, cull_slanted and write_obj against
the reference's own outputs (tests/golden/splat.npz, written by
make_splat.py), plus the Delaunay empty-circumcircle property that the
reference's test_geometry checks (test_stereonet.py:364-385)."""

import os

import numpy as np
import pytest

import oracle
from conftest import GOLDEN, package

G = np.load(os.path.join(GOLDEN, "splat.npz"))
SCENES = ["s48", "s64"]


@pytest.fixture(scope="module")
def sv():
    return package()


def _cam(W, H):
    return ((W * 0.5) / 0.5773502691896257, W * 0.5, H * 0.5, 0.10, 0.0)


def _pts3d(name):
    W, H, _ = G[name + "_wh"]
    p, ok = oracle.disparity_to_points(G[name + "_disp"], W, H, _cam(W, H), min_disp=0.4)
    return p, ok.astype(np.uint8)


@pytest.mark.parametrize("case", ["dl_rand", "dl_grid"])
def test_delaunay_matches_reference_order(sv, case):
    pts = [tuple(p) for p in G[case + "_pts"].tolist()]
    got = sv.delaunay(pts)
    want = [tuple(t) for t in G[case + "_tris"].tolist()]
    assert got == want


def test_delaunay_empty_circumcircle():
    sv = package()
    pts = G["dl_rand_pts"]
    tris = sv.delaunay([tuple(p) for p in pts.tolist()])
    viol = 0
    for i, j, k in tris:
        (ax, ay), (bx, by), (cx, cy) = pts[i], pts[j], pts[k]
        d = 2.0 * (ax * (by - cy) + bx * (cy - ay) + cx * (ay - by))
        if abs(d) < 1e-12:
            continue
        a2, b2, c2 = ax * ax + ay * ay, bx * bx + by * by, cx * cx + cy * cy
        ux = (a2 * (by - cy) + b2 * (cy - ay) + c2 * (ay - by)) / d
        uy = (a2 * (cx - bx) + b2 * (ax - cx) + c2 * (bx - ax)) / d
        r2 = (ax - ux) ** 2 + (ay - uy) ** 2
        dd = (pts[:, 0] - ux) ** 2 + (pts[:, 1] - uy) ** 2
        dd[[i, j, k]] = np.inf
        viol += int((dd < r2 - 1e-7).sum())
    assert len(tris) > 250 and viol == 0


def test_delaunay_small_inputs(sv):
    assert sv.delaunay([]) == [] and sv.delaunay([(0, 0), (1, 1)]) == []
    assert sv.delaunay([(0, 0), (4, 0), (0, 3)]) == [(0, 1, 2)] or \
        len(sv.delaunay([(0, 0), (4, 0), (0, 3)])) == 1


@pytest.mark.parametrize("name", SCENES)
def test_select_points_matches_reference(sv, name):
    W, H, npts = (int(v) for v in G[name + "_wh"])
    got = sv.select_points(G[name + "_disp"], W, H, None, target=npts)
    assert got == [tuple(p) for p in G[name + "_sel"].tolist()]
    gv = sv.select_points(G[name + "_disp"], W, H, G[name + "_valid"], target=npts)
    assert gv == [tuple(p) for p in G[name + "_selv"].tolist()]


@pytest.mark.parametrize("name", SCENES)
def test_culls_match_reference(sv, name):
    W, H, _ = (int(v) for v in G[name + "_wh"])
    pts2 = [tuple(p) for p in G[name + "_sel"].tolist()]
    raw = [tuple(t) for t in G[name + "_tris_raw"].tolist()]
    assert sv.delaunay(pts2) == raw
    p3 = _pts3d(name)
    tris = sv.cull_triangles(pts2, raw, p3, W)
    assert tris == [tuple(t) for t in G[name + "_tris"].tolist()]
    sl = sv.cull_slanted(pts2, tris, p3, None, W)
    assert sl == [tuple(t) for t in G[name + "_tris_slant"].tolist()]
    # a slant threshold that bites: fewer triangles, an ordered subset
    sl2 = sv.cull_slanted(pts2, tris, p3, None, W, max_slant_cos=0.9)
    assert len(sl2) < len(tris) and all(t in tris for t in sl2)


@pytest.mark.parametrize("name", SCENES)
def test_write_obj_text_matches_reference(sv, name, tmp_path):
    W, H, _ = (int(v) for v in G[name + "_wh"])
    pts2 = [tuple(p) for p in G[name + "_sel"].tolist()]
    tris = [tuple(t) for t in G[name + "_tris"].tolist()]
    path = str(tmp_path / "m.obj")
    nv, nf = sv.write_obj(path, pts2, tris, _pts3d(name), G[name + "_rgb"], W, H)
    assert (nv, nf) == tuple(G[name + "_obj_counts"].tolist())
    assert open(path, "rb").read() == G[name + "_obj"].tobytes()


@pytest.mark.parametrize("name", SCENES)
def test_oracle_splat_render_matches_reference(name):
    """The numpy restatement of SplatRenderer.render (the splat workload's
    CPU baseline) against the reference's renders, original and novel view
    with a background."""
    W, H, _ = (int(v) for v in G[name + "_wh"])
    cam = _cam(W, H)
    ident = [1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0]
    got = oracle.splat_render(G[name + "_gc"], cam, W, H, ident, [0.0, 0.0, 0.0])
    assert np.array_equal(got, G[name + "_render"])
    v = G[name + "_views"][0]
    got = oracle.splat_render(G[name + "_adam1"], cam, W, H, list(v[1:10]), list(v[10:13]),
                              background=(0.1, 0.2, 0.3))
    assert np.array_equal(got, G[name + "_render_novel"])
