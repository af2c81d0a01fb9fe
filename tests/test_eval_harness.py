"""Middlebury scoring on the CPU: the oracle's load_gt / masks / score /
constant against the reference harness's own outputs (tests/golden/eval.npz,
made by tests/golden/make_eval.py), planted mistakes, scenes() on the same
directory tree, and write_pgm / write_ppm's byte format."""
import os

import numpy as np
import pytest

import oracle
from conftest import GOLDEN, package


def _golden():
    return np.load(os.path.join(GOLDEN, "eval.npz"))


def _case(z, c):
    W, H, scale, ct = z[c + "_meta"]
    gl = oracle.read_image_bytes(z[c + "_fl"].tobytes())[0]
    gr = oracle.read_image_bytes(z[c + "_fr"].tobytes())[0]
    return int(W), int(H), float(scale), gl, gr


def test_oracle_matches_reference_harness():
    z = _golden()
    for c in z["cases"]:
        W, H, scale, gl, gr = _case(z, c)
        dl, kl = oracle.gt_from_gray(gl, scale)
        dr, kr = oracle.gt_from_gray(gr, scale)
        no = oracle.nonocc(dl, kl, dr, kr, W, H)
        assert np.array_equal(dl, z[c + "_gt"]) and np.array_equal(kl, z[c + "_known"])
        assert np.array_equal(no, z[c + "_nonocc"])
        assert 0 < no.sum() < kl.sum() < W * H
        p = z[c + "_pred"]
        assert np.array_equal(np.array(oracle.score(p, dl, no)), z[c + "_score_nonocc"])
        assert np.array_equal(np.array(oracle.score(p, dl, kl)), z[c + "_score_all"])
        assert oracle.constant(dl, no) == float(z[c + "_constant"])
    assert oracle.score([1.0], [1.0], [False]) is None


def test_planted_mistakes_fail():
    """Rounding x - d + 0.5 down instead of toward zero, or a strict
    tolerance in the left-right test, changes the non-occluded masks."""
    z = _golden()
    diffs = {"floor": 0, "strict": 0}
    for c in z["cases"]:
        W, H, scale, gl, gr = _case(z, c)
        dl, kl = oracle.gt_from_gray(gl, scale)
        dr, kr = oracle.gt_from_gray(gr, scale)
        d2, k2 = dl.reshape(H, W), kl.reshape(H, W)
        r2, q2 = dr.reshape(H, W), kr.reshape(H, W)
        fl = np.zeros((H, W), bool)
        st = np.zeros((H, W), bool)
        for y in range(H):
            for x in range(W):
                if not k2[y, x]:
                    continue
                t = x - d2[y, x] + 0.5
                for mode, xr in (("floor", int(np.floor(t))), ("strict", int(t))):
                    if 0 <= xr < W and q2[y, xr]:
                        dd = abs(d2[y, x] - r2[y, xr])
                        ok = dd <= 1.0 if mode == "floor" else dd < 1.0
                        (fl if mode == "floor" else st)[y, x] = ok
        diffs["floor"] += int(not np.array_equal(fl.ravel(), z[c + "_nonocc"]))
        diffs["strict"] += int(not np.array_equal(st.ravel(), z[c + "_nonocc"]))
    assert diffs["floor"] > 0 and diffs["strict"] > 0


def test_scenes_tree(tmp_path):
    ev = package().evaluation
    z = _golden()
    for i, yn in enumerate(z["tree"]):
        year, name = str(yn).split("/")
        d = tmp_path / year / name
        d.mkdir(parents=True)
        idx = [str(t) for t in z["tree"] if str(t).startswith(year + "/")].index(str(yn))
        if not (year == "2006" and idx == 3):
            (d / ev.LAYOUT[year][2]).write_bytes(b"x")
    (tmp_path / "2001").mkdir()
    got = ["%s/%s/%s" % (s["year"], s["name"], s["split"]) for s in ev.scenes(str(tmp_path))]
    assert got == [str(s) for s in z["scenes"]]


def test_write_pgm_ppm(tmp_path):
    sv = package()
    v = [0.25, -0.1, 1.2, 0.5, 0.75, 0.0]
    sv.write_pgm(str(tmp_path / "a.pgm"), v, 3, 2, 0.0, 1.0)
    b = (tmp_path / "a.pgm").read_bytes()
    assert b == b"P5\n3 2\n255\n" + bytes([64, 0, 255, 128, 191, 0])
    sv.write_pgm(str(tmp_path / "b.pgm"), [2.0, 2.0], 2, 1)
    assert (tmp_path / "b.pgm").read_bytes()[-2:] == bytes([0, 0])
    sv.write_ppm(str(tmp_path / "c.ppm"), [0.0, 0.5, 1.0], 1, 1)
    assert (tmp_path / "c.ppm").read_bytes() == b"P6\n1 1\n255\n" + bytes([0, 128, 255])
