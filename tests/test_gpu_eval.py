"""GPU parity of the Middlebury scoring harness (eval_real.py:78-172):
ground truth and masks bit-exact against the reference harness's own outputs
(tests/golden/eval.npz) and the oracle; the bad-pixel percentages exact and
EPE / constant within 1e-12 relative (float64 sums in a fixed parallel
order instead of left to right); evaluate() end to end."""
import os
import pickle

import numpy as np
import pytest
import torch

import oracle
from conftest import GOLDEN, package, V5

import sys  # noqa: E402
sys.path.insert(0, GOLDEN)
import make_eval  # noqa: E402  (its PNG writer)

pytestmark = pytest.mark.gpu
REL = 1e-12


@pytest.fixture(scope="module")
def sv():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return package()


def _golden():
    return np.load(os.path.join(GOLDEN, "eval.npz"))


def _write_case(z, c, d):
    pl, pr = d / (str(c) + "_l.png"), d / (str(c) + "_r.png")
    pl.write_bytes(z[c + "_fl"].tobytes())
    pr.write_bytes(z[c + "_fr"].tobytes())
    W, H, scale, _ = z[c + "_meta"]
    return {"gtl": str(pl), "gtr": str(pr), "scale": float(scale)}, int(W), int(H)


def _close(a, b):
    assert (a is None) == (b is None)
    if a is None:
        return
    assert abs(a[0] - b[0]) <= REL * abs(b[0])
    assert tuple(a[1:]) == tuple(b[1:])


def test_golden(sv, tmp_path):
    ev = sv.evaluation
    z = _golden()
    for c in z["cases"]:
        sc, W, H = _write_case(z, c, tmp_path)
        d, k, W1, H1 = ev.load_gt(sc["gtl"], sc["scale"])
        assert (W1, H1) == (W, H)
        assert np.array_equal(np.array(d), z[c + "_gt"])
        assert np.array_equal(np.array(k), z[c + "_known"])
        gt, known, no, _, _ = ev.masks(sc)
        assert np.array_equal(np.array(gt), z[c + "_gt"])
        assert np.array_equal(np.array(no), z[c + "_nonocc"])
        assert isinstance(gt, list) and isinstance(no, list)
        p = list(map(float, z[c + "_pred"]))
        _close(ev.score(p, gt, no), tuple(z[c + "_score_nonocc"]))
        _close(ev.score(p, gt, known), tuple(z[c + "_score_all"]))
        cg = ev.constant(gt, no)
        assert abs(cg - float(z[c + "_constant"])) <= REL * float(z[c + "_constant"])
    assert ev.score([1.0, 2.0], [1.0, 2.0], [False, False]) is None
    with pytest.raises(ZeroDivisionError):
        ev.constant([1.0], [False])


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_frame_sized_batch_vs_oracle(sv, dtype):
    ev = sv.evaluation
    rng = np.random.default_rng(3)
    B, W, H = 3, 460, 370
    lv = rng.integers(0, 256, size=(B, H, W)).astype(np.float64)
    lv[lv < 30] = 0
    gl = torch.from_numpy(lv / 255.0 * 1.0).cuda()
    rvv = np.clip(lv + rng.integers(-3, 4, size=lv.shape), 0, 255)
    gr = torch.from_numpy(rvv / 255.0).cuda()
    dl, kl = ev.gt_from_gray(gl, 3.0)
    dr, kr = ev.gt_from_gray(gr, 3.0)
    no = ev.nonocc_mask(dl, kl, dr, kr, W, H)
    pred = (dl + torch.from_numpy(rng.normal(0, 3, size=(B, H, W))).cuda()).to(dtype)
    stats, _, _ = ev.score_stats(pred, dl, no, B=B)
    absd = ev.abs_dev_sums(dl.reshape(-1), no.reshape(-1), stats, B=B)
    for b in range(B):
        d0, k0 = oracle.gt_from_gray(gl[b].cpu().numpy().ravel(), 3.0)
        d1, k1 = oracle.gt_from_gray(gr[b].cpu().numpy().ravel(), 3.0)
        assert np.array_equal(dl[b].cpu().numpy().ravel(), d0)
        assert np.array_equal(kl[b].cpu().numpy().ravel().astype(bool), k0)
        n0 = oracle.nonocc(d0, k0, d1, k1, W, H)
        assert np.array_equal(no[b].cpu().numpy().ravel().astype(bool), n0)
        pb = pred[b].double().cpu().numpy().ravel()
        ref = oracle.score(pb, d0, n0)
        s = stats[b].cpu().numpy()
        _close((s[1] / s[0], 100.0 * s[2] / s[0], 100.0 * s[3] / s[0], 100.0 * s[4] / s[0]), ref)
        cref = oracle.constant(d0, n0)
        assert abs(float(absd[b]) / s[0] - cref) <= REL * cref
    # a map with nothing in its mask scores n = 0
    stats, _, _ = ev.score_stats(pred, dl, torch.zeros_like(no), B=B)
    assert float(stats[:, 0].abs().sum()) == 0.0


def _scene(sv, tmp_path, W=96, H=64):
    """A synthetic Middlebury-layout scene: RGB views as PNG, ground truth as
    grey-level PNG (level = 3 * disparity) for both views."""
    p = sv.synth.band_pair(5, W=W, H=H)
    d = tmp_path / "2006" / "Synth"
    d.mkdir(parents=True)
    rng = np.random.default_rng(11)
    for name, img in (("view1.png", p["left"]), ("view5.png", p["right"])):
        (d / name).write_bytes(make_eval.gray_png(rng, np.asarray(img, np.uint8), 2))
    disp = np.asarray(p["disp"], np.float64).reshape(H, W)
    valid = np.asarray(p["valid"], bool).reshape(H, W)
    lv = np.where(valid, np.clip(np.round(disp * 3.0), 0, 255), 0).astype(np.uint8)
    dr = np.asarray(sv.left_to_right(disp.ravel().tolist(), W, H))
    rv = np.clip(np.round(dr.reshape(H, W) * 3.0), 0, 255).astype(np.uint8)
    (d / "disp1.png").write_bytes(make_eval.gray_png(rng, lv, 0))
    (d / "disp5.png").write_bytes(make_eval.gray_png(rng, rv, 0))
    return sv.evaluation.scenes(str(tmp_path))[0]


@pytest.mark.parametrize("mode,cons", [("net", False), ("net", True), ("refine", False),
                                       ("sgm", False)])
def test_evaluate(sv, tmp_path, mode, cons):
    ev = sv.evaluation
    net = sv.load_model(V5)
    sc = _scene(sv, tmp_path)
    r = ev.evaluate(net, sc, consistency=cons, mode=mode, dump=str(tmp_path / "dump"),
                    save_dir=str(tmp_path / "out"))
    with open(tmp_path / "dump" / "2006_Synth.pkl", "rb") as f:  # written by evaluate above
        dmp = pickle.load(f)
    gt, known, no, W, H = ev.masks(sc)
    assert dmp["gt"] == gt and dmp["nonocc"] == no and dmp["known"] == known
    pred = dmp["pred"]
    _close(r["nonocc"], oracle.score(pred, gt, no))
    _close(r["all"], oracle.score(pred, gt, known))
    assert abs(r["const"] - oracle.constant(gt, no)) <= REL * oracle.constant(gt, no)
    assert r["size"] == "%dx%d" % (W, H)
    assert (tmp_path / "out" / "2006_Synth.pgm").read_bytes().startswith(b"P5\n96 64\n255\n")
    if mode == "sgm":
        R = int(max(g for g, m in zip(gt, no) if m) + 2) // 2 + 1
        lg = sv.read_image(sc["left"])[0]
        rg = sv.read_image(sc["right"])[0]
        ref = sv.refine_fullres(lg, rg, [float(R)] * (W * H), W, H, R=R, P1=12.0, P2=150.0,
                                rx=4, ry=3)
        assert np.array_equal(np.array(pred), np.array(ref))
    assert r["nonocc"][2] < 60.0
    txt = ev.table([(dict(sc), r)], "t")
    assert "| 2006 Synth | train |" in txt and "**mean train (1)**" in txt
