"""The allocation-free entry point sv_refine_disparity_noalloc (workspace in
the library's static device scratch) against sv_refine_disparity (workspace
from the caller): identical disparities and statuses on two calls in a row,
the second smaller than the first, so it runs over the first call's
leftovers in the scratch.  Sizes past the scratch are refused before any
GPU work.  Runs last (collected after every other test file)."""

import ctypes
import importlib

import numpy as np
import pytest
import torch

from conftest import package

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sv():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return package()


def _batch(sv, B, W, H, seed):
    pairs = [sv.synth.make_pair(W, H, seed + i, dmin=10.0, dmax=60.0, n_boxes=2) for i in range(B)]
    l = torch.from_numpy(np.stack([p["left"] for p in pairs])).cuda()
    r = torch.from_numpy(np.stack([p["right"] for p in pairs])).cuda()
    d0 = torch.from_numpy(np.stack([p["disp0"] for p in pairs])).cuda()
    return l, r, d0


def _noalloc(L, l, r, d0, B, W, H, R=10, tol=2.0):
    out = torch.empty(B * W * H, dtype=torch.float64, device="cuda")
    st = torch.empty(B * W * H, dtype=torch.uint8, device="cuda")
    rc = L.lib().sv_refine_disparity_noalloc(L.ptr(l), L.ptr(r), L.SV_U8, L.ptr(d0), B, W, H, R,
                                             12.0, 150.0, 4, 3, tol, L.ptr(out), L.ptr(st),
                                             L.stream())
    return rc, out, st


@pytest.mark.parametrize("first,second", [((1, 460, 370), (3, 192, 144)),
                                          ((3, 300, 200), (1, 97, 61))])
def test_noalloc_equals_refine_disparity(sv, first, second):
    L = importlib.import_module("from-scratch-stereo-vision_amd._lib")
    for k, (B, W, H) in enumerate((first, second)):
        print("noalloc call %d: B=%d %dx%d" % (k, B, W, H), flush=True)
        l, r, d0 = _batch(sv, B, W, H, 40 + 10 * k)
        rc, out, st = _noalloc(L, l, r, d0, B, W, H)
        L.check(rc)
        ref, rst = sv.refine_disparity(l, r, None, d0, W, H, planes=False)
        torch.cuda.synchronize()
        assert torch.equal(out, ref.reshape(-1))
        assert torch.equal(st, rst.reshape(-1).to(torch.uint8))


def test_noalloc_refuses_oversize(sv):
    """Two 460x370 pairs need about 166 MB of workspace: refused with
    SV_EWORKSPACE, nothing queued (the outputs keep their fill)."""
    L = importlib.import_module("from-scratch-stereo-vision_amd._lib")
    B, W, H = 2, 460, 370
    l = torch.zeros(B, H, W, dtype=torch.uint8, device="cuda")
    d0 = torch.zeros(B, H, W, dtype=torch.float64, device="cuda")
    out = torch.full((B * W * H,), 7.0, dtype=torch.float64, device="cuda")
    st = torch.full((B * W * H,), 9, dtype=torch.uint8, device="cuda")
    rc = L.lib().sv_refine_disparity_noalloc(L.ptr(l), L.ptr(l), L.SV_U8, L.ptr(d0), B, W, H, 10,
                                             12.0, 150.0, 4, 3, 2.0, L.ptr(out), L.ptr(st),
                                             L.stream())
    assert rc == L.SV_EWORKSPACE
    assert b"scratch" in L.lib().sv_last_error()
    torch.cuda.synchronize()
    assert bool((out == 7.0).all()) and bool((st == 9).all())
