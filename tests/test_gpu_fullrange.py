"""Full-range SGM with the constant centre (the control of
eval_real.py:143-146) on the constant-centre kernels, against the CPU oracle
and against the general GPU path given the constant centre map.

Bar: disparities, confidences and statuses exactly equal (integer path;
the float64 epilogue is evaluated in the reference's order).  Band lengths
cover every bucket of the constant-centre kernels exactly and with soft
pads (K below the bucket), and lengths past the largest bucket (general
kernels).
"""

import numpy as np
import pytest
import torch

import oracle
from conftest import package

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sv():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return package()


def _cu(a, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(a)).to(dtype).cuda()


# K = 2R + 1: 41 (bucket 45, soft pads), 45, 65, 71 and 85 (bucket 85), 121
# and 125 (bucket 125), 141 (general kernels)
@pytest.mark.parametrize("R", [20, 22, 32, 35, 42, 60, 62, 70])
def test_full_range_matches_oracle(sv, R):
    W, H = 200, 60
    print("full range R=%d %dx%d" % (R, W, H), flush=True)
    p = sv.synth.make_pair(W, H, R, dmin=4.0, dmax=min(2.0 * R - 2, 120.0))
    gl, gr = sv.synth.gray(p["left"]), sv.synth.gray(p["right"])
    kw = dict(R=R, P1=12.0, P2=150.0, rx=4, ry=3)
    out, conf = sv.full_range(_cu(p["left"], torch.uint8), _cu(p["right"], torch.uint8), W, H,
                              return_conf=True, **kw)
    d0 = np.full(W * H, float(R))
    ref, rconf = oracle.refine_fullres(gl, gr, d0, W, H, return_conf=True, **kw)
    assert np.array_equal(out.cpu().numpy().ravel(), ref)
    assert np.array_equal(conf.cpu().numpy().ravel(), rconf)


@pytest.mark.parametrize("R,paths", [(42, 8), (42, 4), (30, 4), (100, 8)])
def test_full_range_lr_matches_oracle(sv, R, paths):
    W, H = 160, 48
    print("full range lr R=%d paths=%d" % (R, paths), flush=True)
    p = sv.synth.make_pair(W, H, 100 + R, dmin=4.0, dmax=min(2.0 * R - 2, 100.0))
    gl, gr = sv.synth.gray(p["left"]), sv.synth.gray(p["right"])
    kw = dict(R=R, P1=12.0, P2=150.0, rx=4, ry=3, paths=paths)
    out, st = sv.full_range(_cu(p["left"], torch.uint8), _cu(p["right"], torch.uint8), W, H,
                            lr=True, tol=1.0, **kw)
    d0 = np.full(W * H, float(R))
    rout, rst = oracle.lr_refine(gl, gr, d0, W, H, tol=1.0, **kw)
    assert np.array_equal(st.cpu().numpy().ravel().astype(np.int32), rst)
    assert np.array_equal(out.cpu().numpy().ravel(), rout)


def test_full_range_pipeline_size(sv):
    """460x370, R = 42 (K = 85, the largest full range of the 460x370
    workload): the oracle on one pair; a batch of three equals the general
    GPU path given the constant centre map, pair by pair."""
    W, H, R = 460, 370, 42
    pairs = [sv.synth.pipeline_pair(s) for s in range(3)]
    L8 = _cu(np.stack([q["left"] for q in pairs]), torch.uint8)
    R8 = _cu(np.stack([q["right"] for q in pairs]), torch.uint8)
    kw = dict(R=R, P1=12.0, P2=150.0, rx=4, ry=3)
    out = sv.full_range(L8, R8, W, H, **kw)
    gen = sv.refine_fullres(L8, R8, torch.full((3, H, W), float(R), dtype=torch.float64,
                                                device="cuda"), W, H, **kw)
    assert torch.equal(out, gen)
    one = sv.full_range(L8[1:2], R8[1:2], W, H, **kw)
    assert torch.equal(one.reshape(-1), out[1].reshape(-1))
    gl, gr = sv.synth.gray(pairs[0]["left"]), sv.synth.gray(pairs[0]["right"])
    ref = oracle.refine_fullres(gl, gr, np.full(W * H, float(R)), W, H, **kw)
    assert np.array_equal(out[0].cpu().numpy().ravel(), ref)


@pytest.mark.parametrize("W,H,R", [(7, 5, 22), (9, 1, 42), (1, 6, 20), (40, 30, 62), (30, 8, 42)])
def test_full_range_small_images(sv, W, H, R):
    """Images narrower than the band (every candidate of the far entries is
    invalid), single rows and columns; lr mode with W < R takes the general
    kernels (the right view's centre is then 0, not R)."""
    p = sv.synth.make_pair(W, H, 7, dmin=1.0, dmax=3.0)
    gl, gr = sv.synth.gray(p["left"]), sv.synth.gray(p["right"])
    kw = dict(R=R, P1=12.0, P2=150.0, rx=4, ry=3)
    d0 = np.full(W * H, float(R))
    out = sv.full_range(gl.ravel().tolist(), gr.ravel().tolist(), W, H, **kw)
    assert isinstance(out, list)
    assert np.array_equal(np.asarray(out), oracle.refine_fullres(gl, gr, d0, W, H, **kw))
    lo, st = sv.full_range(_cu(gl), _cu(gr), W, H, lr=True, tol=2.0, **kw)
    rlo, rst = oracle.lr_refine(gl, gr, d0, W, H, tol=2.0, **kw)
    assert np.array_equal(st.cpu().numpy().ravel().astype(np.int32), rst)
    assert np.array_equal(lo.cpu().numpy().ravel(), rlo)


def test_full_range_u16_path_falls_back(sv):
    """invalid + P2 > 255 leaves the u8 path: the general kernels, same
    results as refine_fullres with the constant map."""
    W, H, R = 120, 40, 30
    p = sv.synth.make_pair(W, H, 3, dmin=4.0, dmax=50.0)
    gl, gr = sv.synth.gray(p["left"]), sv.synth.gray(p["right"])
    kw = dict(R=R, P1=20.0, P2=400.0, rx=4, ry=3)
    out = sv.full_range(_cu(p["left"], torch.uint8), _cu(p["right"], torch.uint8), W, H, **kw)
    ref = oracle.refine_fullres(gl, gr, np.full(W * H, float(R)), W, H, **kw)
    assert np.array_equal(out.cpu().numpy().ravel(), ref)


@pytest.mark.parametrize("lr", [False, True])
@pytest.mark.parametrize("Rs", [[20, 35, 42, 30], [12, 70, 12, 40]])
def test_full_range_per_pair_centres(sv, lr, Rs):
    """One R per pair (eval_real.py's rule gives each pair its own): one
    launch with a centre table when every band fits the constant-centre
    kernels (first set), one call per distinct R otherwise (second set,
    K = 141); each pair equals its own single call."""
    W, H = 150, 40
    pairs = [sv.synth.make_pair(W, H, 30 + i, dmin=4.0, dmax=2.0 * r - 2) for i, r in enumerate(Rs)]
    L8 = _cu(np.stack([q["left"] for q in pairs]), torch.uint8)
    R8 = _cu(np.stack([q["right"] for q in pairs]), torch.uint8)
    kw = dict(P1=12.0, P2=150.0, rx=4, ry=3)
    res = sv.full_range(L8, R8, W, H, Rs, lr=lr, tol=1.0, **kw)
    for i, r in enumerate(Rs):
        one = sv.full_range(L8[i:i + 1], R8[i:i + 1], W, H, r, lr=lr, tol=1.0, **kw)
        if lr:
            assert torch.equal(res[0][i], one[0][0]) and torch.equal(res[1][i], one[1][0])
        else:
            assert torch.equal(res[i], one[0])
    gl, gr = sv.synth.gray(pairs[0]["left"]), sv.synth.gray(pairs[0]["right"])
    d0 = np.full(W * H, float(Rs[0]))
    if lr:
        rout, rst = oracle.lr_refine(gl, gr, d0, W, H, tol=1.0, R=Rs[0], **kw)
        assert np.array_equal(res[0][0].cpu().numpy().ravel(), rout)
        assert np.array_equal(res[1][0].cpu().numpy().ravel().astype(np.int32), rst)
    else:
        ref = oracle.refine_fullres(gl, gr, d0, W, H, R=Rs[0], **kw)
        assert np.array_equal(res[0].cpu().numpy().ravel(), ref)
