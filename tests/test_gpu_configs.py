"""The uneven-batch and training workloads at their own sizes.

* Uneven batch (64 pairs of 200x150 .. 1390x1110, disparity 16-240 px, net +
  band in one captured graph): a ragged batch holding the extremes -- a
  1390x1110 pair at disparity up to 240 and a 200x150 pair -- and sizes
  between, through predict_auto + refine_disparity(planes=False) captured
  in one HIP graph.  The replay equals the eager run, each pair equals the
  same pair run alone, and the 1390x1110 pair's refinement is bit-exact
  against the oracle given the GPU's own network disparity.
* Training (260x192, batch 16, pseudo-labels from full-range SGM): the batch
  gradient equals the mean of the 16 single-pair gradients, some dozens of
  Adam steps lower the loss, and the full-range fallback replaces exactly
  the flagged pixels at that size.
"""

import importlib

import numpy as np
import pytest
import torch

import oracle
from conftest import package, V5

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sv():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return package()


def _m(name):
    return importlib.import_module("from-scratch-stereo-vision_amd." + name)


UNEVEN = [(1390, 1110, 240.0), (200, 150, 16.0), (700, 540, 120.0), (460, 370, 80.0),
          (1012, 781, 200.0)]


def test_uneven_batch_extremes_graph(sv):
    R, N = _m("ragged"), _m("net")
    pairs = [sv.synth.ragged_pair((w, h, d, 700 + i)) for i, (w, h, d) in enumerate(UNEVEN)]
    sizes = [(w, h) for w, h, _ in UNEVEN]
    print("uneven batch", sizes, flush=True)
    b = R.RaggedBatch(sizes)
    l = b.pack([p["left"] for p in pairs])
    r = b.pack([p["right"] for p in pairs])
    net = N.load_model(V5)
    eout, est, res = R.net_band_ragged(net, b, l, r)
    eout, est = eout.clone(), est.clone()
    d0 = torch.empty(b.B, b.slot, dtype=torch.float64, device="cuda")
    N._call("sv_widen", N._p(res.full), N._p(b.table), b.slot, N._p(d0), b.slot, None, b.B, b.slot,
            N.L.stream())
    torch.cuda.synchronize()
    cap = R.CapturedPipeline(lambda a, c: R.net_band_ragged(net, b, a, c)[:2], l, r, warmup=1)
    for _ in range(2):
        out, st = cap.replay()
        torch.cuda.synchronize()
        assert all(torch.equal(u, v) for u, v in zip(b.unpack(out), b.unpack(eout)))
        assert all(torch.equal(u, v) for u, v in zip(b.unpack(st), b.unpack(est)))
    outs, sts = b.unpack(eout), b.unpack(est)
    for n, (p, (w, h, _)) in enumerate(zip(pairs, UNEVEN)):
        print("pair %d alone: %dx%d" % (n, w, h), flush=True)
        lt = torch.as_tensor(p["left"]).cuda()[None]
        rt = torch.as_tensor(p["right"]).cuda()[None]
        one = N.predict_auto_batch(net, lt, rt, w, h)
        assert one.trace[0] == res.trace[n], (n, one.trace, res.trace[n])
        od = one.full.reshape(1, h, w).double()
        o1, s1 = sv.refine_disparity(lt, rt, None, od, w, h, planes=False)
        assert torch.equal(o1.reshape(h, w), outs[n]), n
        assert torch.equal(s1.reshape(h, w).to(torch.uint8), sts[n]), n
    # the largest pair against the oracle, from the GPU's own disp0
    w, h, _ = UNEVEN[0]
    gl, gr = sv.synth.gray(pairs[0]["left"]), sv.synth.gray(pairs[0]["right"])
    disp0 = b.unpack(d0)[0].cpu().numpy().ravel()
    print("oracle refine_disparity on %dx%d" % (w, h), flush=True)
    ref, rst = oracle.refine_disparity_plain(gl, gr, disp0, w, h)
    assert np.array_equal(sts[0].cpu().numpy().ravel().astype(np.int32), rst)
    assert np.array_equal(outs[0].cpu().numpy().ravel(), ref)


@pytest.fixture(scope="module")
def c5_data(sv):
    T = _m("train")
    d = T.make_dataset([10000 + i for i in range(16)], 260, 192, workers=1)
    disp, valid = T.pseudo_labels(d["left"], d["right"], 260, 192, max_disp=40.0)
    return d, disp, valid


def test_training_batch16_gradient_is_mean_of_singles(sv, c5_data):
    N = _m("net")
    d, disp, valid = c5_data
    net = N.load_model(V5)
    l = torch.as_tensor(d["left"]).cuda().float() * (1.0 / 255.0)
    r = torch.as_tensor(d["right"]).cuda().float() * (1.0 / 255.0)
    dims = [(192, 260)] * 16
    losses, _ = N.train_grads_batch(net, l, r, disp, valid, dims)
    gb = net.grads.clone()
    acc = torch.zeros_like(gb)
    for i in range(16):
        li, _ = N.train_grads_batch(net, l[i:i + 1], r[i:i + 1], disp[i:i + 1], valid[i:i + 1],
                                    dims[:1])
        assert abs(li[0] - losses[i]) <= 1e-6 * abs(li[0])
        acc += net.grads
    acc /= 16
    rel = float((gb - acc).norm() / acc.norm())
    print("batch-16 gradient vs mean of singles: rel L2 %.2e" % rel)
    assert rel <= 1e-5


def test_training_batch16_loss_falls_and_fallback_exact(sv, c5_data):
    T, N = _m("train"), _m("net")
    d, disp, valid = c5_data
    net = N.load_model(V5)
    l = torch.as_tensor(d["left"]).cuda()
    r = torch.as_tensor(d["right"]).cuda()
    hist = T.train_net(net, l, r, disp, valid, steps=48, batch=16, lr=5e-4, jitter=True,
                       log_every=16)
    print("loss per 16 steps:", hist)
    assert len(hist) == 3 and hist[-1][1] < hist[0][1], hist
    md = [float(np.max(g[v])) for g, v in zip(d["disp"], d["valid"])]
    none = T.infer_with_fallback(net, l, r, 260, 192, md, tau=2.0)
    assert none.fallback_pairs == []
    every = T.infer_with_fallback(net, l, r, 260, 192, md, tau=-1.0, statuses=(1, 2))
    assert every.fallback_pairs == list(range(16))
    fr = T.full_range(l, r, 260, 192, md)
    flag = every.flag
    assert bool(flag.any())
    assert torch.equal(every.status, none.status)
    assert torch.equal(every.disp[flag], fr[flag])
    assert torch.equal(every.disp[~flag], none.disp[~flag])
    # the full-range control equals refine_fullres with each pair's constant centre
    for i in (0, 7):
        R_ = T.full_range_R(md[i])
        c = torch.full((1, 192, 260), float(R_), dtype=torch.float64, device="cuda")
        ref = sv.refine_fullres(l[i:i + 1], r[i:i + 1], c, 260, 192, R=R_, P1=12.0, P2=150.0,
                                rx=4, ry=3)
        assert torch.equal(fr[i], ref.reshape(192, 260))
