"""Ragged batches (pairs of different sizes in one launch per stage) and
their capture in a HIP graph: every pair's result must equal the same pair
run alone through the uniform entry points."""
import importlib

import numpy as np
import pytest
import torch

from conftest import package

pytestmark = pytest.mark.gpu

SIZES = [(200, 150), (256, 180), (173, 201), (300, 160), (96, 64)]


def _mods():
    R = importlib.import_module("from-scratch-stereo-vision_amd.ragged")
    N = importlib.import_module("from-scratch-stereo-vision_amd.net")
    return R, N


def _pairs():
    sv = package()
    return [sv.synth.make_pair(w, h, 40 + i, dmin=4.0, dmax=36.0, n_boxes=1)
            for i, (w, h) in enumerate(SIZES)]


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_ragged_refinement_equals_per_pair(mode):
    R, _ = _mods()
    sv = package()
    ps = _pairs()
    b = R.RaggedBatch(SIZES)
    l = b.pack([p["left"] for p in ps])
    r = b.pack([p["right"] for p in ps])
    d0 = b.pack([p["disp0"].astype(np.float64) for p in ps], torch.float64)
    out, st = R.refine_ragged(b, l, r, d0, mode=mode, R=10)
    outs = b.unpack(out)
    sts = b.unpack(st) if st is not None else None
    for n, (p, (w, h)) in enumerate(zip(ps, SIZES)):
        lt = torch.as_tensor(p["left"]).cuda()
        rt = torch.as_tensor(p["right"]).cuda()
        dt = torch.as_tensor(p["disp0"].astype(np.float64)).cuda()
        if mode == 0:
            ref = sv.refine_fullres(lt, rt, dt, w, h, R=10, P1=12.0, P2=150.0, rx=4, ry=3)
            rst = None
        elif mode == 1:
            ref, rst = sv.lr_refine(lt, rt, dt, w, h, tol=2.0, R=10, P1=12.0, P2=150.0, rx=4, ry=3)
        else:
            ref, rst = sv.refine_disparity(lt, rt, None, dt, w, h, planes=False)
        assert torch.equal(outs[n], ref.reshape(h, w)), (mode, n)
        if rst is not None:
            assert torch.equal(sts[n], rst.reshape(h, w).to(torch.uint8)), (mode, n)


def test_ragged_predict_auto_equals_per_pair():
    R, N = _mods()
    ps = _pairs()
    net = N.load_model(package().__path__[0] + "/weights/stereonet_c5.txt")
    b = R.RaggedBatch(SIZES)
    l = b.pack([p["left"] for p in ps])
    r = b.pack([p["right"] for p in ps])
    res = N.predict_auto_batch(net, l, r, None, None, batch=b)
    fulls = b.unpack(res.full)
    for n, (p, (w, h)) in enumerate(zip(ps, SIZES)):
        one = N.predict_auto_batch(net, torch.as_tensor(p["left"]).cuda()[None],
                                   torch.as_tensor(p["right"]).cuda()[None], w, h)
        assert one.trace[0] == res.trace[n], (n, one.trace, res.trace)
        assert torch.equal(fulls[n], one.full[0]), n


def _same(b, x, y):
    # pixels past each pair's size are padding and carry no result
    return all(torch.equal(u, v) for u, v in zip(b.unpack(x), b.unpack(y)))


def test_captured_pipeline_replays_match_eager():
    R, N = _mods()
    ps = _pairs()
    net = N.load_model(package().__path__[0] + "/weights/stereonet_c5.txt")
    b = R.RaggedBatch(SIZES)
    l = b.pack([p["left"] for p in ps])
    r = b.pack([p["right"] for p in ps])
    eager_out, eager_st, _ = R.net_band_ragged(net, b, l, r)
    eager_out, eager_st = eager_out.clone(), eager_st.clone()
    cap = R.CapturedPipeline(lambda a, c: R.net_band_ragged(net, b, a, c)[:2], l, r)
    for _ in range(3):
        out, st = cap.replay()
        torch.cuda.synchronize()
        assert _same(b, out, eager_out)
        assert _same(b, st, eager_st)
        # eager work between replays
        R.refine_ragged(b, l, r, eager_out, mode=0)
        torch.cuda.synchronize()
    # new data through the same graph: swap the views' roles
    ps2 = [dict(p, left=np.ascontiguousarray(p["left"][::-1]),
                right=np.ascontiguousarray(p["right"][::-1])) for p in ps]
    l.copy_(b.pack([p["left"] for p in ps2]))
    r.copy_(b.pack([p["right"] for p in ps2]))
    out, st = cap.replay()
    ref_out, ref_st, _ = R.net_band_ragged(net, b, l.clone(), r.clone())
    torch.cuda.synchronize()
    assert _same(b, out, ref_out)
    assert _same(b, st, ref_st)


@pytest.mark.parametrize("overlap", [1, 2])
def test_parts_on_two_streams_captured_match_eager(overlap):
    """The batch cut into two parts, each with its own network object on its
    own stream (scratch buffers are per stream), both captured in one graph:
    every replay equals the single-stream eager run, pair by pair.  With
    the horizontal SGM paths on the library's side stream (overlap 2) the
    eager runs fork onto it and the captured ones keep to their own
    streams."""
    print("parts on two streams, sgm_overlap", overlap, flush=True)
    R, N = _mods()
    A = importlib.import_module("from-scratch-stereo-vision_amd.autotune")
    A.set_knob("sgm_overlap", overlap)
    try:
        _parts_on_two_streams(R, N)
    finally:
        A.reset()


def _parts_on_two_streams(R, N):
    ps = _pairs()
    path = package().__path__[0] + "/weights/stereonet_c5.txt"
    b = R.RaggedBatch(SIZES)
    l = b.pack([p["left"] for p in ps])
    r = b.pack([p["right"] for p in ps])
    eager_out, eager_st, _ = R.net_band_ragged(N.load_model(path), b, l, r)
    ref_out = [t.clone() for t in b.unpack(eager_out)]
    ref_st = [t.clone() for t in b.unpack(eager_st)]
    groups = [[0, 1, 2], [3, 4]]
    bs = [R.RaggedBatch([SIZES[i] for i in g]) for g in groups]
    ls = [bb.pack([ps[i]["left"] for i in g]) for bb, g in zip(bs, groups)]
    rs = [bb.pack([ps[i]["right"] for i in g]) for bb, g in zip(bs, groups)]
    nets = [N.load_model(path) for _ in groups]
    streams = [torch.cuda.Stream() for _ in groups]

    def both(*_):
        cur = torch.cuda.current_stream()
        outs = []
        for k, s in enumerate(streams):
            s.wait_stream(cur)
            with torch.cuda.stream(s):
                outs.append(R.net_band_ragged(nets[k], bs[k], ls[k], rs[k])[:2])
        for s in streams:
            cur.wait_stream(s)
        return outs

    cap = R.CapturedPipeline(both, l, r)
    for _ in range(3):
        outs = cap.replay()
        torch.cuda.synchronize()
        for (out, st), bb, g in zip(outs, bs, groups):
            for u, v, i in zip(bb.unpack(out), bb.unpack(st), g):
                assert torch.equal(u, ref_out[i]), i
                assert torch.equal(v, ref_st[i]), i
        # eager work between replays
        R.refine_ragged(b, l, r, eager_out, mode=0)
        torch.cuda.synchronize()


def test_autotuner_knobs_keep_results_identical():
    R, N = _mods()
    A = importlib.import_module("from-scratch-stereo-vision_amd.autotune")
    ps = _pairs()
    net = N.load_model(package().__path__[0] + "/weights/stereonet_c5.txt")
    b = R.RaggedBatch(SIZES)
    l = b.pack([p["left"] for p in ps])
    r = b.pack([p["right"] for p in ps])
    try:
        # tune() raises if any variant changes a single output value
        def run():
            out, st, _ = R.net_band_ragged(net, b, l, r)
            return b.unpack(out) + b.unpack(st)  # padding past each pair carries no result

        res = A.tune(run, reps=1)
        for k, (_, values) in A.KNOBS.items():
            assert res[k] in values
            assert A.get_knob(k) == res[k]
        assert len(res["_trials"]) >= 3
    finally:
        A.reset()
    assert all(A.get_knob(k) == 0 for k in A.KNOBS)


def test_per_layer_conv_variants_keep_results_identical():
    R, N = _mods()
    A = importlib.import_module("from-scratch-stereo-vision_amd.autotune")
    ps = _pairs()
    net = N.load_model(package().__path__[0] + "/weights/stereonet_c5.txt")
    b = R.RaggedBatch(SIZES)
    l = b.pack([p["left"] for p in ps])
    r = b.pack([p["right"] for p in ps])
    res0 = N.predict_auto_batch(net, l, r, None, None, batch=b)
    ref = [t.clone() for t in b.unpack(res0.full)]
    # tune_layers raises if any variant changes a layer's output
    chosen = A.tune_layers(net, lambda: N.predict_auto_batch(net, l, r, None, None, batch=b), reps=1)
    assert len(chosen) >= 15 and len(net._conv_tune) == len(chosen)
    res1 = N.predict_auto_batch(net, l, r, None, None, batch=b)
    assert all(torch.equal(u, v) for u, v in zip(b.unpack(res1.full), ref))
    assert all(A.get_knob(k) == 0 for k in A.KNOBS)


def test_side_streams_of_two_callers_match_sequential():
    """With the horizontal SGM paths on a side stream (sgm_overlap 2), two
    eager callers on two streams each get their own side stream: refining
    two batches side by side gives the outputs of refining them one after
    the other."""
    print("two eager callers with side streams", flush=True)
    R, N = _mods()
    A = importlib.import_module("from-scratch-stereo-vision_amd.autotune")
    ps = _pairs()
    b = R.RaggedBatch(SIZES)
    l = b.pack([p["left"] for p in ps])
    r = b.pack([p["right"] for p in ps])
    d0 = torch.zeros_like(l, dtype=torch.float64)
    for i, t in enumerate(b.unpack(d0)):
        t.fill_(3.0 + i)
    A.set_knob("sgm_overlap", 2)
    try:
        ref = [R.refine_ragged(b, l, r, d0, mode=m)[0].clone() for m in (0, 2)]
        torch.cuda.synchronize()
        streams = [torch.cuda.Stream(), torch.cuda.Stream()]
        cur = torch.cuda.current_stream()
        for _ in range(3):
            outs = []
            for m, s in zip((0, 2), streams):
                s.wait_stream(cur)
                with torch.cuda.stream(s):
                    outs.append(R.refine_ragged(b, l, r, d0, mode=m)[0])
            for s in streams:
                cur.wait_stream(s)
            torch.cuda.synchronize()
            for o, e in zip(outs, ref):  # the padding past each pair carries no result
                assert all(torch.equal(u, v) for u, v in zip(b.unpack(o), b.unpack(e)))
    finally:
        A.reset()


QSIZES = [(460, 370), (97, 64), (33, 17), (600, 90), (80, 500), (256, 180)]


def _qpairs():
    sv = package()
    return [sv.synth.make_pair(w, h, 70 + i, dmin=2.0, dmax=min(30.0, w / 3.0), n_boxes=1)
            for i, (w, h) in enumerate(QSIZES)]


@pytest.mark.parametrize("mode", [0, 2])
def test_sgm_work_queue_equals_grid(mode):
    """Persistent SGM workgroups over the per-XCD work queues (knob
    sgm_queue 2) give the grid launches' results bit for bit, under every
    prefetch depth, horizontal tile and overlap setting."""
    R, _ = _mods()
    A = importlib.import_module("from-scratch-stereo-vision_amd.autotune")
    ps = _qpairs()
    b = R.RaggedBatch(QSIZES)
    l = b.pack([p["left"] for p in ps])
    r = b.pack([p["right"] for p in ps])
    d0 = b.pack([p["disp0"].astype(np.float64) for p in ps], torch.float64)
    try:
        A.set_knob("sgm_split", 2)
        A.set_knob("sgm_queue", 1)
        ref, rst = R.refine_ragged(b, l, r, d0, mode=mode, R=10)
        ref = ref.clone()
        rst = rst.clone() if rst is not None else None
        for pf in (2, 8):
            for ht in (8, 16):
                for ov in (1, 2):
                    print("sgm queue mode=%d prefetch=%d htile=%d overlap=%d" % (mode, pf, ht, ov),
                          flush=True)
                    A.apply({"sgm_queue": 2, "sgm_prefetch": pf, "sgm_htile": ht,
                             "sgm_overlap": ov})
                    out, st = R.refine_ragged(b, l, r, d0, mode=mode, R=10)
                    torch.cuda.synchronize()
                    for n, (u, v) in enumerate(zip(b.unpack(out), b.unpack(ref))):
                        assert torch.equal(u, v), (pf, ht, ov, n)
                    if st is not None:
                        for n, (u, v) in enumerate(zip(b.unpack(st), b.unpack(rst))):
                            assert torch.equal(u, v), (pf, ht, ov, n)
    finally:
        A.reset()


def test_sgm_work_queue_graph_replays():
    """The queue build and the persistent kernels captured in a HIP graph:
    every replay, with eager work between them, equals the eager run (the
    queue heads are cleared by a kernel inside the graph on each replay)."""
    print("sgm queue graph replays", flush=True)
    R, _ = _mods()
    A = importlib.import_module("from-scratch-stereo-vision_amd.autotune")
    ps = _qpairs()
    b = R.RaggedBatch(QSIZES)
    l = b.pack([p["left"] for p in ps])
    r = b.pack([p["right"] for p in ps])
    d0 = b.pack([p["disp0"].astype(np.float64) for p in ps], torch.float64)
    try:
        A.apply({"sgm_split": 2, "sgm_queue": 1})
        eager, est = R.refine_ragged(b, l, r, d0, mode=2)
        eager, est = eager.clone(), est.clone()
        A.set_knob("sgm_queue", 2)
        cap = R.CapturedPipeline(lambda a, c: R.refine_ragged(b, a, c, d0, mode=2), l, r)
        for k in range(3):
            out, st = cap.replay()
            torch.cuda.synchronize()
            assert all(torch.equal(u, v) for u, v in zip(b.unpack(out), b.unpack(eager))), k
            assert all(torch.equal(u, v) for u, v in zip(b.unpack(st), b.unpack(est))), k
            R.refine_ragged(b, l, r, d0, mode=0)  # eager work between replays
            torch.cuda.synchronize()
    finally:
        A.reset()


def test_sgm_work_queue_uniform_batch():
    """A uniform batch through the queue kernels equals the grid launches."""
    print("sgm queue uniform batch", flush=True)
    sv = package()
    A = importlib.import_module("from-scratch-stereo-vision_amd.autotune")
    ps = [sv.synth.band_pair(s) for s in range(4)]
    W, H = ps[0]["W"], ps[0]["H"]
    l = torch.stack([torch.as_tensor(p["left"]) for p in ps]).cuda()
    r = torch.stack([torch.as_tensor(p["right"]) for p in ps]).cuda()
    d0 = torch.stack([torch.as_tensor(p["disp0"].astype(np.float64)) for p in ps]).cuda()
    try:
        A.apply({"sgm_split": 2, "sgm_queue": 1})
        ref, rst = sv.lr_refine(l, r, d0, W, H, tol=2.0, R=10, P1=12.0, P2=150.0, rx=4, ry=3)
        ref, rst = ref.clone(), rst.clone()
        A.set_knob("sgm_queue", 2)
        out, st = sv.lr_refine(l, r, d0, W, H, tol=2.0, R=10, P1=12.0, P2=150.0, rx=4, ry=3)
        torch.cuda.synchronize()
        assert torch.equal(out, ref)
        assert torch.equal(st, rst)
    finally:
        A.reset()


def test_fallback_ragged_whole_and_flagged_pixels():
    """fallback_ragged: tau above 1 changes nothing; tau below 0 replaces
    every pair by its full-range solution; flagged-pixel mode takes the
    full-range value exactly on the occluded/mismatched pixels."""
    print("fallback_ragged", flush=True)
    R, _ = _mods()
    sv = package()
    ps = _pairs()
    b = R.RaggedBatch(SIZES)
    l = b.pack([p["left"] for p in ps])
    r = b.pack([p["right"] for p in ps])
    d0 = b.pack([p["disp0"].astype(np.float64) for p in ps], torch.float64)
    md = [36.0] * len(SIZES)
    out, st = R.refine_ragged(b, l, r, d0, mode=2)
    base = [t.clone() for t in b.unpack(out)]
    assert R.fallback_ragged(b, l, r, out, st, md, tau=1.01) == []
    assert all(torch.equal(u, v) for u, v in zip(b.unpack(out), base))
    fulls = [sv.full_range(li, ri, w, h, 20).reshape(h, w)
             for li, ri, (w, h) in zip(b.unpack(l), b.unpack(r), SIZES)]
    part = out.clone()
    picked = R.fallback_ragged(b, l, r, part, st, md, tau=-1.0, whole=False)
    assert picked == list(range(len(SIZES)))
    for o, bs, s_, f in zip(b.unpack(part), base, b.unpack(st), fulls):
        flag = (s_ == 1) | (s_ == 2)
        assert torch.equal(o[flag], f[flag]) and torch.equal(o[~flag], bs[~flag])
    R.fallback_ragged(b, l, r, out, st, md, tau=-1.0)
    assert all(torch.equal(u, f) for u, f in zip(b.unpack(out), fulls))
