"""Pins the CPU oracle (oracle/stereo_oracle.c) before anything is compared
against it.

* refinement: golden vectors written from the reference (tests/golden/*.npz)
  plus closed forms (P1 = P2 = 0 gives L = C; constant centre equals a
  brute-force minimum over disparity paths up to a per-pixel constant);
* network layers: torch float64 library routines and central differences
  (the tolerances of test_stereonet.py:99-215);
* predict_auto: the predictions embedded in the reference's demo page;
* train_grads: a golden vector written from the reference;
* Adam: torch.optim.Adam.
"""

import itertools
import os

import numpy as np
import pytest
import torch
import torch.nn.functional as F

import oracle
from conftest import GOLDEN, V5

REFINE_CASES = ["refine_band_r10", "refine_defaults", "refine_fullrange",
                "refine_float", "refine_paths4", "refine_jumps",
                "refine_tiny_7x5", "refine_tiny_9x1", "refine_tiny_1x6"]
DIRS8 = [(1, 0), (-1, 0), (0, 1), (0, -1), (1, 1), (-1, 1), (1, -1), (-1, -1)]


def _kw(z):
    kw = dict(R=int(z["R"]), P1=float(z["P1"]), P2=float(z["P2"]),
              rx=int(z["rx"]), ry=int(z["ry"]))
    for k in ("ad_weight", "ad_trunc", "uniq", "prior_w", "prior_t", "paths"):
        if k in z.files:
            kw[k] = z[k].item()
    return kw


@pytest.mark.parametrize("name", REFINE_CASES)
def test_refinement_golden(name):
    z = np.load(os.path.join(GOLDEN, name + ".npz"))
    W, H = int(z["W"]), int(z["H"])
    kw = _kw(z)
    R, K = kw["R"], 2 * kw["R"] + 1
    cl = oracle.census(z["left"], W, H, kw["rx"], kw["ry"])
    cr = oracle.census(z["right"], W, H, kw["rx"], kw["ry"])
    assert np.array_equal(cl, z["census_l"]) and np.array_equal(cr, z["census_r"])
    nb = (2 * kw["rx"] + 1) * (2 * kw["ry"] + 1) - 1
    aw, at = kw.get("ad_weight", 0.0), kw.get("ad_trunc", 20.0)
    pw, pt = kw.get("prior_w", 0.0), kw.get("prior_t", 4.0)
    inv = nb + aw * at + pw * pt + 4.0
    C = oracle.band_costs(cl, cr, z["left"], z["right"], z["base"], W, H, R,
                          aw, at, inv, z["disp0"], pw, pt)
    assert np.array_equal(C, z["C"].astype(np.float64))
    for i, (dx, dy) in enumerate(DIRS8):
        Ld = oracle.aggregate_path(C, z["base"], W, H, K, kw["P1"], kw["P2"], dx, dy)
        assert np.array_equal(Ld, z["L_dirs"][i].astype(np.float64)), (dx, dy)
    out, conf = oracle.refine_fullres(z["left"], z["right"], z["disp0"], W, H,
                                      return_conf=True, **kw)
    assert np.array_equal(out, z["out"]) and np.array_equal(conf, z["conf"])
    assert np.array_equal(oracle.left_to_right(z["disp0"], W, H), z["d0r"])
    lo, st = oracle.lr_refine(z["left"], z["right"], z["disp0"], W, H,
                              tol=float(z["tol"]), **kw)
    assert np.array_equal(lo, z["lr_out"])
    assert np.array_equal(st, z["lr_status"].astype(np.int32))
    assert np.array_equal(oracle.median_filter(lo, W, H), z["median"])


def test_refine_disparity_plain_golden():
    z = np.load(os.path.join(GOLDEN, "refine_disparity_plain.npz"))
    W, H = int(z["W"]), int(z["H"])
    out, st = oracle.refine_disparity_plain(z["left"], z["right"], z["disp0"], W, H)
    assert np.array_equal(out, z["out"])
    assert np.array_equal(st, z["status"].astype(np.int32))


def test_median_golden():
    z = np.load(os.path.join(GOLDEN, "median.npz"))
    W, H = int(z["W"]), int(z["H"])
    for r, p in ((1, 1), (1, 2), (2, 1)):
        got = oracle.median_filter(z["disp"], W, H, radius=r, passes=p)
        assert np.array_equal(got, z["r%d_p%d" % (r, p)])
    # interior pixels are the plain 3x3 median (scipy), an independent check
    import scipy.ndimage as ndi
    d = z["disp"].reshape(H, W)
    ref = ndi.median_filter(d, size=3, mode="nearest")
    got = oracle.median_filter(z["disp"], W, H).reshape(H, W)
    assert np.array_equal(got[1:-1, 1:-1], ref[1:-1, 1:-1])


def test_left_to_right_golden():
    z = np.load(os.path.join(GOLDEN, "left_to_right.npz"))
    W, H = int(z["W"]), int(z["H"])
    assert np.array_equal(oracle.left_to_right(z["dl"], W, H), z["dr"])
    # constant disparity d: every right pixel that some left pixel reaches
    # holds d; the d rightmost right-view columns are holes filled with d
    dl = np.full(3 * 20, 4.0)
    dr = oracle.left_to_right(dl, 20, 3)
    assert np.all(dr == 4.0)


def test_census_identical_images_zero_distance():
    # test_stereonet.py:493-495
    rng = np.random.default_rng(0)
    img = rng.uniform(size=96 * 64)
    c = oracle.census(img, 96, 64, 4, 3)
    assert all(bin(int(a) ^ int(a)).count("1") == 0 for a in c)
    # closed form on a tiny image: bit order MSB = first offset (dy outer)
    img = np.array([[0.1, 0.9, 0.5], [0.3, 0.5, 0.7], [0.2, 0.8, 0.6]])
    c = oracle.census(img.ravel(), 3, 3, 1, 1)
    # centre 0.5: neighbours (dy,dx) in order (-1,-1),(-1,0),(-1,1),(0,-1),(0,1),(1,-1),(1,0),(1,1)
    bits = [0.1 < 0.5, 0.9 < 0.5, 0.5 < 0.5, 0.3 < 0.5, 0.7 < 0.5, 0.2 < 0.5, 0.8 < 0.5, 0.6 < 0.5]
    want = int("".join("1" if b else "0" for b in bits), 2)
    assert int(c[4]) == want


def test_sgm_zero_penalties_is_identity():
    """P1 = P2 = 0: every candidate equals the previous minimum, so L = C."""
    z = np.load(os.path.join(GOLDEN, "refine_band_r10.npz"))
    W, H, R = int(z["W"]), int(z["H"]), int(z["R"])
    K = 2 * R + 1
    C = z["C"].astype(np.float64)
    for dx, dy in DIRS8:
        assert np.array_equal(oracle.aggregate_path(C, z["base"], W, H, K, 0.0, 0.0, dx, dy), C)


def test_sgm_constant_centre_matches_brute_force_paths():
    """With a constant band centre the recurrence is textbook SGM
    (Hirschmueller 2008): L(p, d) equals the minimum over all disparity
    sequences along the path of sum(C) + penalties, minus a constant per
    pixel.  Enumerate every sequence on short lines."""
    rng = np.random.default_rng(3)
    P1, P2 = 3.0, 7.0
    for W, R in ((5, 1), (4, 2)):
        K = 2 * R + 1
        C = rng.integers(0, 12, size=(W, K)).astype(np.float64)
        base = np.full(W, R, dtype=np.int32)
        L = oracle.aggregate_path(C.ravel(), base, W, 1, K, P1, P2, 1, 0).reshape(W, K)
        for p in range(W):
            brute = np.full(K, np.inf)
            for seq in itertools.product(range(K), repeat=p + 1):
                cost = sum(C[i, seq[i]] for i in range(p + 1))
                for i in range(1, p + 1):
                    jump = abs(seq[i] - seq[i - 1])
                    cost += 0.0 if jump == 0 else (P1 if jump == 1 else P2)
                brute[seq[-1]] = min(brute[seq[-1]], cost)
            diff = L[p] - brute
            assert np.allclose(diff, diff[0]), (W, R, p)


# --------------------------------------------------------------- network

def _t(a):
    return torch.from_numpy(np.asarray(a, dtype=np.float64))


@pytest.mark.parametrize("k,s,d", [(3, 2, 1), (3, 1, 2), (3, 1, 1), (3, 1, 4)])
def test_conv2d_matches_torch(k, s, d):
    rng = np.random.default_rng(1)
    IC, OC, H, W = 3, 4, 11, 9
    x = rng.normal(size=(IC, H, W))
    w = rng.normal(size=(OC, IC, k, k))
    b = rng.normal(size=OC)
    p = (d * (k - 1)) // 2
    y, OH, OW = oracle.conv2d_fwd(x, IC, H, W, w, b, OC, k, s, p, d)
    ref = F.conv2d(_t(x)[None], _t(w), _t(b), stride=s, padding=p, dilation=d)[0]
    assert ref.shape == (OC, OH, OW)
    assert np.allclose(y.reshape(OC, OH, OW), ref.numpy(), rtol=1e-13, atol=1e-13)
    # backward against autograd
    xt = _t(x).requires_grad_()
    wt = _t(w).requires_grad_()
    bt = _t(b).requires_grad_()
    gy = rng.normal(size=(OC, OH, OW))
    (F.conv2d(xt[None], wt, bt, stride=s, padding=p, dilation=d)[0] * _t(gy)).sum().backward()
    gx, gw, gb = oracle.conv2d_bwd(x, IC, H, W, w, OC, k, gy, s, p, d)
    assert np.allclose(gx, xt.grad.numpy().ravel(), rtol=1e-12, atol=1e-12)
    assert np.allclose(gw, wt.grad.numpy().ravel(), rtol=1e-12, atol=1e-12)
    assert np.allclose(gb, bt.grad.numpy(), rtol=1e-12, atol=1e-12)


def test_conv3d_matches_torch():
    rng = np.random.default_rng(2)
    IC, OC, D, H, W = 3, 2, 5, 4, 6
    x = rng.normal(size=(IC, D, H, W))
    w = rng.normal(size=(OC, IC, 3, 3, 3))
    b = rng.normal(size=OC)
    y = oracle.conv3d_fwd(x, IC, D, H, W, w, b, OC)
    xt = _t(x).requires_grad_()
    wt = _t(w).requires_grad_()
    bt = _t(b).requires_grad_()
    ref = F.conv3d(xt[None], wt, bt, padding=1)[0]
    assert np.allclose(y, ref.detach().numpy().ravel(), rtol=1e-13, atol=1e-13)
    gy = rng.normal(size=ref.shape)
    (ref * _t(gy)).sum().backward()
    gx, gw, gb = oracle.conv3d_bwd(x, IC, D, H, W, w, OC, gy)
    assert np.allclose(gx, xt.grad.numpy().ravel(), rtol=1e-12, atol=1e-12)
    assert np.allclose(gw, wt.grad.numpy().ravel(), rtol=1e-12, atol=1e-12)
    assert np.allclose(gb, bt.grad.numpy(), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("H,W,H2,W2", [(3, 3, 6, 6), (5, 7, 10, 14), (9, 13, 4, 5), (6, 4, 17, 3)])
def test_bilinear_matches_torch(H, W, H2, W2):
    rng = np.random.default_rng(4)
    x = rng.normal(size=(2, H, W))
    y = oracle.bilinear_fwd(x, 2, H, W, H2, W2)
    xt = _t(x).requires_grad_()
    ref = F.interpolate(xt[None], size=(H2, W2), mode="bilinear", align_corners=False)[0]
    assert np.allclose(y, ref.detach().numpy().ravel(), rtol=1e-13, atol=1e-13)
    gy = rng.normal(size=(2, H2, W2))
    (ref * _t(gy)).sum().backward()
    assert np.allclose(oracle.bilinear_bwd(gy, 2, H, W, H2, W2), xt.grad.numpy().ravel(),
                       rtol=1e-12, atol=1e-12)
    # test_stereonet.py:392-396: a constant field stays constant
    c = oracle.bilinear_fwd(np.full(16, 2.5), 1, 4, 4, 9, 9)
    assert c.max() - c.min() < 1e-12


def test_avgpool_and_softargmin_match_torch():
    rng = np.random.default_rng(5)
    x = rng.normal(size=(2, 7, 9))
    out, H2, W2 = oracle.avgpool2(x, 2, 7, 9)
    ref = F.avg_pool2d(_t(x)[None], 2)[0]
    assert np.allclose(out, ref.numpy().ravel(), rtol=1e-14, atol=1e-14)
    D, H, W = 6, 3, 4
    cost = rng.normal(size=(D, H, W)) * 3
    pen = np.zeros((D, H, W))
    pen[3:, :, :2] = 8.0
    disp, probs = oracle.softargmin_fwd(cost, D, H, W, pen)
    p = torch.softmax(-(_t(cost) + _t(pen)), dim=0)
    assert np.allclose(probs, p.numpy().ravel(), rtol=1e-13, atol=1e-15)
    dref = (p * torch.arange(D, dtype=torch.float64)[:, None, None]).sum(0)
    assert np.allclose(disp, dref.numpy().ravel(), rtol=1e-13, atol=1e-13)
    # test_stereonet.py:386-390: soft-argmin finds a V minimum at 3
    d, _ = oracle.softargmin_fwd([abs(k - 3) * 4.0 for k in range(5)], 5, 1, 1)
    assert abs(d[0] - 3.0) < 0.01


def _numgrad(f, arr, i, eps=1e-5):
    old = arr[i]
    arr[i] = old + eps
    lp = f()
    arr[i] = old - eps
    lm = f()
    arr[i] = old
    return (lp - lm) / (2 * eps)


def _relerr(a, b):
    return abs(a - b) / max(1e-9, abs(a) + abs(b))


def test_stereo_primitive_gradients_central_differences():
    """test_stereonet.py:149-215 tolerances: < 1e-6 (cost volume, soft-argmin,
    bilinear), < 1e-5 (cross-entropy)."""
    rng = np.random.default_rng(21)
    C, H, W, D = 2, 5, 6, 3
    fl = rng.normal(size=C * H * W)
    fr = rng.normal(size=C * H * W)
    wts = 0.5 + 0.07 * np.arange(C * D * H * W)
    lcv = lambda: float(np.dot(wts, oracle.costvol_fwd(fl, fr, C, D, H, W)))
    gfl, gfr = oracle.costvol_bwd(fl, fr, C, D, H, W, wts)
    worst = max(max(_relerr(_numgrad(lcv, fl, i), gfl[i]), _relerr(_numgrad(lcv, fr, i), gfr[i]))
                for i in range(0, C * H * W, 4))
    assert worst < 1e-6
    D2, H2, W2 = 5, 4, 4
    cost = rng.normal(size=D2 * H2 * W2)
    gw = 1.0 + 0.11 * np.arange(H2 * W2)
    lsa = lambda: float(np.dot(gw, oracle.softargmin_fwd(cost, D2, H2, W2)[0]))
    disp, probs = oracle.softargmin_fwd(cost, D2, H2, W2)
    gc = oracle.softargmin_bwd(probs, disp, D2, H2, W2, gw)
    assert max(_relerr(_numgrad(lsa, cost, i), gc[i]) for i in range(0, cost.size, 3)) < 1e-6
    D3, H3, W3 = 6, 3, 3
    cost = rng.normal(size=D3 * H3 * W3) * 2
    tgt = rng.uniform(0.5, 4.5, size=H3 * W3)
    val = np.ones(H3 * W3, dtype=bool)

    def lce():
        _, pr = oracle.softargmin_fwd(cost, D3, H3, W3)
        return oracle.disparity_ce(pr, tgt, val, D3, H3, W3)[0]

    _, pr = oracle.softargmin_fwd(cost, D3, H3, W3)
    _, gce, _ = oracle.disparity_ce(pr, tgt, val, D3, H3, W3)
    assert max(_relerr(_numgrad(lce, cost, i), gce[i]) for i in range(0, cost.size, 4)) < 1e-5
    # the CE value is the soft-target cross-entropy -sum q log p
    q = np.exp(-((np.arange(D3)[:, None] - tgt[None]) ** 2) / (2 * 0.8 ** 2))
    q /= q.sum(0)
    ce_ref = float(-(q * np.log(pr.reshape(D3, -1))).sum(0).mean())
    assert abs(oracle.disparity_ce(pr, tgt, val, D3, H3, W3)[0] - ce_ref) < 1e-12


def test_cost_volume_minimises_at_true_shift():
    # test_stereonet.py:405-415
    rng = np.random.default_rng(11)
    base = rng.uniform(size=40)
    shift = 3
    fL = base[8:32]
    fR = base[8 + shift:32 + shift]
    v = oracle.costvol_fwd(fL, fR, 1, 6, 1, 24).reshape(6, 24)
    assert int(np.argmin(v[:, 8:22].sum(1))) == shift


def test_pseudo_huber_closed_form():
    pred = np.array([0.0, 3.0, -4.0, 1.0])
    tgt = np.array([0.0, 0.0, 0.0, 1.0])
    val = np.array([1, 1, 1, 0], dtype=bool)
    loss, g, n = oracle.pseudo_huber(pred, tgt, val)
    assert n == 3
    want = (0.0 + (np.sqrt(10) - 1) + (np.sqrt(17) - 1)) / 3
    assert abs(loss - want) < 1e-15
    assert np.allclose(g, [0.0, 3 / np.sqrt(10) / 3, -4 / np.sqrt(17) / 3, 0.0])


def test_adam_matches_torch():
    rng = np.random.default_rng(6)
    p0 = rng.normal(size=50)
    opt = oracle.Adam(p0.copy(), lr=1e-2)
    pt = torch.tensor(p0.copy(), requires_grad=True)
    topt = torch.optim.Adam([pt], lr=1e-2, betas=(0.9, 0.999), eps=1e-8)
    for _ in range(5):
        g = rng.normal(size=50)
        opt.step(g)
        pt.grad = torch.tensor(g)
        topt.step()
    assert np.allclose(opt.p, pt.detach().numpy(), rtol=1e-12, atol=1e-14)
    # the pre-clip: a gradient of 100 acts as 5
    a = oracle.Adam(np.zeros(1), lr=0.1)
    b = oracle.Adam(np.zeros(1), lr=0.1)
    a.step(np.array([100.0]))
    b.step(np.array([5.0]))
    assert a.p[0] == b.p[0]


def test_train_grads_golden():
    z = np.load(os.path.join(GOLDEN, "train_grads_tiny.npz"))
    arch = oracle.Arch(int(z["feat_C"]), int(z["cost_C"]), int(z["D"]),
                       int(z["refine_C"]), tuple(z["refine_dil"]))
    loss, G, d2 = oracle.train_grads(arch, z["params"], z["left"], z["right"],
                                     z["disp"], z["valid"], int(z["W"]), int(z["H"]))
    assert abs(loss - float(z["loss"])) <= 1e-12 * abs(float(z["loss"]))
    assert np.abs(G - z["grads"]).max() <= 1e-12 * np.abs(z["grads"]).max()
    assert np.abs(d2 - z["d2"]).max() < 1e-12


def test_train_grads_central_differences():
    """The composite loss's gradient against central differences on a few
    parameters of every block (test_stereonet.py:99-145 style)."""
    z = np.load(os.path.join(GOLDEN, "train_grads_tiny.npz"))
    arch = oracle.Arch(int(z["feat_C"]), int(z["cost_C"]), int(z["D"]),
                       int(z["refine_C"]), tuple(z["refine_dil"]))
    P = z["params"].copy()
    args = (z["left"], z["right"], z["disp"], z["valid"], int(z["W"]), int(z["H"]))
    _, G, _ = oracle.train_grads(arch, P, *args)
    f = lambda: oracle.train_grads(arch, P, *args)[0]
    rng = np.random.default_rng(0)
    idx = rng.choice(P.size, size=24, replace=False)
    worst = 0.0
    for i in idx:
        ng = _numgrad(f, P, i, eps=1e-6)
        if abs(ng) + abs(G[i]) < 1e-7:
            continue
        worst = max(worst, _relerr(ng, G[i]))
    assert worst < 1e-4


def _demo_gray(rgb):
    a = rgb.astype(np.float64) / 255.0
    return 0.299 * a[..., 0] + 0.587 * a[..., 1] + 0.114 * a[..., 2]


@pytest.mark.parametrize("key", ["cones", "art"])
def test_predict_auto_matches_demo_page(key):
    """web/stereonet_demo.html:998 stores the reference's predict_auto output
    on real Middlebury pairs as round(d * 64); the oracle must reproduce every
    quantised value."""
    dp = np.load(os.path.join(GOLDEN, "demo_page.npz"))
    cfg, blocks = oracle.load_weights_txt(V5)
    P = np.concatenate(blocks)
    assert np.array_equal(P, dp["ckpt_flat"])
    arch = oracle.Arch(cfg["feat_C"], cfg["cost_C"], cfg["D"], cfg["refine_C"],
                       cfg["refine_dil"], cfg["target_peak"])
    gl, gr = _demo_gray(dp[key + "_left_rgb"]), _demo_gray(dp[key + "_right_rgb"])
    H, W = gl.shape
    full, _, _, _, _ = oracle.predict_auto(arch, P, gl, gr, W, H)
    q = np.clip(np.round(full * 64), 0, 65535).astype(np.uint16)
    assert np.array_equal(q, dp[key + "_raw"])


def _consistency_golden():
    return np.load(os.path.join(GOLDEN, "consistency.npz"))


def test_consistency_helpers_golden():
    """mirror_h, lr_consistency (two tolerances) and fill_invalid against
    the reference's own outputs on crafted maps: columns landing exactly on
    0 and W-1, just outside, a row entirely outside (left unchanged by the
    fill), quarter-pixel values whose difference equals the tolerance."""
    z = _consistency_golden()
    W, H = int(z["W"]), int(z["H"])
    assert np.array_equal(oracle.mirror_h(z["img"], 2, W, H), z["mirror"])
    for name, tol in (("t15", 1.5), ("t075", 0.75)):
        v = oracle.lr_consistency(z["dl"], z["dr"], W, H, tol)
        assert np.array_equal(v, z["valid_" + name]), name
        assert np.array_equal(oracle.fill_invalid(z["dl"], v, W, H), z["fill_" + name]), name


def test_consistency_closed_forms():
    """A double flip is the identity; two constant maps of the same value c
    agree wherever x - c lands in [0, W-1] and nowhere else; filling an
    all-valid map changes nothing; a run between valid values a < b takes a,
    a run at the row start takes its right neighbour."""
    rng = np.random.default_rng(5)
    W, H = 23, 4
    img = rng.uniform(size=3 * W * H)
    assert np.array_equal(oracle.mirror_h(oracle.mirror_h(img, 3, W, H), 3, W, H), img)
    c = 6.5
    v = oracle.lr_consistency(np.full(W * H, c), np.full(W * H, c), W, H, 0.0).reshape(H, W)
    x = np.arange(W)
    assert np.array_equal(v, np.broadcast_to((x - c >= 0) & (x - c <= W - 1), (H, W)))
    d = rng.uniform(0, 30, size=W * H)
    assert np.array_equal(oracle.fill_invalid(d, np.ones(W * H, bool), W, H), d)
    row = np.array([3.0, 9.0, 9.0, 2.0, 5.0, 1.0])
    valid = np.array([False, True, False, True, False, False])
    assert np.array_equal(oracle.fill_invalid(row, valid, 6, 1), [9.0, 9.0, 2.0, 2.0, 2.0, 2.0])


@pytest.mark.parametrize("name,ww", [("auto", None), ("w64", 64)])
def test_predict_consistent_golden(name, ww):
    """predict_consistent (v5 weights, 96x64 pair) against the reference's
    output: same valid fraction, disparities within 1e-12 px (the oracle's
    libm exp differs from the reference's hand-made one by ~1e-16)."""
    z = _consistency_golden()
    cfg, blocks = oracle.load_weights_txt(V5)
    arch = oracle.Arch(cfg["feat_C"], cfg["cost_C"], cfg["D"], cfg["refine_C"],
                       cfg["refine_dil"], cfg["target_peak"])
    Wp, Hp = int(z["Wp"]), int(z["Hp"])
    d, frac, _ = oracle.predict_consistent(arch, np.concatenate(blocks), z["left"], z["right"],
                                           Wp, Hp, ww)
    assert frac == float(z["pcfrac_" + name])
    assert np.abs(d - z["pc_" + name]).max() <= 1e-12



def test_planes_golden():
    """SLIC labels (step 12 and 8), plane_snap with and without a validity
    mask, and refine_disparity(planes=True) against the reference's own
    outputs (tests/golden/planes.npz), bit for bit."""
    z = np.load(os.path.join(GOLDEN, "planes.npz"))
    for n in ("a", "b"):
        W, H = int(z[n + "_W"]), int(z[n + "_H"])
        for step in (12, 8):
            lab, k = oracle.slic(z[n + "_rgb"], W, H, step)
            assert k == int(z["%s_k%d" % (n, step)])
            assert np.array_equal(lab, z["%s_lab%d" % (n, step)])
        assert np.array_equal(oracle.plane_snap(z[n + "_disp"], z[n + "_rgb"], W, H), z[n + "_snap"])
        assert np.array_equal(oracle.plane_snap(z[n + "_disp"], z[n + "_rgb"], W, H,
                                                valid=z[n + "_valid"], min_inl=0.4), z[n + "_snapv"])
    d, st = oracle.refine_disparity_planes(z["b_left"], z["b_right"], z["b_rgb"], z["b_disp0"], 64, 48)
    assert np.array_equal(d, z["b_refined"]) and np.array_equal(st, z["b_status"])


def test_plane_snap_closed_form():
    """An exactly planar map with a uniform colour: every segment fits the
    plane exactly, so plane_snap returns the plane's own values (floored at
    0) everywhere SLIC labelled."""
    W, H = 48, 36
    yy, xx = np.mgrid[0:H, 0:W].astype(np.float64)
    d = 0.25 * xx - 0.125 * yy + 3.0
    rgb = np.full((H, W, 3), 0.5)
    out = oracle.plane_snap(d.ravel(), rgb.ravel(), W, H)
    lab, _ = oracle.slic(rgb.ravel(), W, H, 12)
    m = lab >= 0
    assert np.abs(out[m] - np.maximum(d.ravel()[m], 0.0)).max() < 1e-9
