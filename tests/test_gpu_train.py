"""GPU training on pseudo-labels and the full-range fallback."""
import numpy as np
import pytest
import torch

from conftest import package

pytestmark = pytest.mark.gpu


def _mods():
    import importlib
    T = importlib.import_module("from-scratch-stereo-vision_amd.train")
    N = importlib.import_module("from-scratch-stereo-vision_amd.net")
    return T, N


def test_pseudo_labels_match_full_range_lr():
    T, _ = _mods()
    sv = package()
    d = T.make_dataset([10000, 10001, 10002, 10003], 96, 64, workers=1)
    disp, valid = T.pseudo_labels(d["left"], d["right"], 96, 64, max_disp=42.0)
    # eval_real.py:143-146: R = int(max + 2) // 2 + 1 (truncation)
    assert (T.full_range_R(42.0), T.full_range_R(40.0), T.full_range_R(39.9)) == (23, 22, 21)
    R = T.full_range_R(42.0)
    for i in range(4):
        ref, st = sv.lr_refine(d["left"][i], d["right"][i], np.full((64, 96), float(R)), 96, 64,
                               tol=1.0, R=R, P1=12.0, P2=150.0, rx=4, ry=3)
        np.testing.assert_array_equal(valid[i].cpu().numpy().ravel(), (np.asarray(st) == 0).ravel())
        got = disp[i].cpu().numpy().ravel()
        np.testing.assert_array_equal(got, np.asarray(ref, dtype=np.float32).ravel())


def test_training_lowers_the_loss():
    T, N = _mods()
    dev = torch.device("cuda")
    d = T.make_dataset([10000 + i for i in range(8)], 96, 64, workers=1)
    disp, valid = T.pseudo_labels(d["left"], d["right"], 96, 64, max_disp=42.0)
    net = N.StereoNet(seed=5, feat_C=12, cost_C=6, D=12, refine_C=8)
    l = torch.as_tensor(d["left"]).to(dev)
    r = torch.as_tensor(d["right"]).to(dev)
    hist = T.train_net(net, l, r, disp, valid, steps=60, batch=8, lr=2e-3, jitter=False,
                       log_every=20)
    assert len(hist) == 3
    assert hist[-1][1] < hist[0][1], hist


def test_fallback_takes_full_range_on_flagged_pixels():
    T, N = _mods()
    dev = torch.device("cuda")
    d = T.make_dataset([90000, 90001, 90002, 90003], 128, 96, workers=1)
    net = N.load_model(package().__path__[0] + "/weights/stereonet_c5.txt")
    l = torch.as_tensor(d["left"]).to(dev)
    r = torch.as_tensor(d["right"]).to(dev)
    md = [float(np.max(g[v])) for g, v in zip(d["disp"], d["valid"])]
    none = T.infer_with_fallback(net, l, r, 128, 96, md, tau=2.0)
    assert none.fallback_pairs == []
    every = T.infer_with_fallback(net, l, r, 128, 96, md, tau=-1.0, statuses=(1, 2))
    assert every.fallback_pairs == [0, 1, 2, 3]
    fr = T.full_range(l, r, 128, 96, md)
    flag = every.flag
    assert torch.equal(every.status, none.status)
    assert torch.equal(every.disp[flag], fr[flag])
    assert torch.equal(every.disp[~flag], none.disp[~flag])
