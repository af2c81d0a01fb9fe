"""GPU parity of the refinement path (census, band costs, SGM paths,
refine_fullres, left_to_right, lr_refine, median_filter, refine_disparity)
against the reference's golden vectors and the CPU oracle.

Bar: bit-exact for census words, costs, path costs, statuses; disparities
exactly equal as float64 (the required tolerance is 1e-4 px; the GPU path
evaluates the reference's float64 epilogue in the same order, so the test
demands equality).
"""

import os

import numpy as np
import pytest
import torch

import oracle
from conftest import GOLDEN, package

pytestmark = pytest.mark.gpu

REFINE_CASES = ["refine_band_r10", "refine_defaults", "refine_fullrange",
                "refine_float", "refine_paths4", "refine_jumps",
                "refine_tiny_7x5", "refine_tiny_9x1", "refine_tiny_1x6"]
DIRS8 = [(1, 0), (-1, 0), (0, 1), (0, -1), (1, 1), (-1, 1), (1, -1), (-1, -1)]


@pytest.fixture(scope="module")
def sv():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return package()


def _kw(z):
    kw = dict(R=int(z["R"]), P1=float(z["P1"]), P2=float(z["P2"]),
              rx=int(z["rx"]), ry=int(z["ry"]))
    for k in ("ad_weight", "ad_trunc", "uniq", "prior_w", "prior_t", "paths"):
        if k in z.files:
            kw[k] = z[k].item()
    return kw


def _cu(a, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(a)).to(dtype).cuda()


@pytest.mark.parametrize("name", REFINE_CASES)
def test_refinement_golden(sv, name):
    print("case", name, flush=True)
    z = np.load(os.path.join(GOLDEN, name + ".npz"))
    W, H = int(z["W"]), int(z["H"])
    kw = _kw(z)
    R, K = kw["R"], 2 * kw["R"] + 1
    left, right, d0 = _cu(z["left"]), _cu(z["right"]), _cu(z["disp0"])
    cl = sv.census(left, W, H, kw["rx"], kw["ry"])
    cr = sv.census(right, W, H, kw["rx"], kw["ry"])
    assert np.array_equal(cl.cpu().numpy().view(np.uint64), z["census_l"])
    assert np.array_equal(cr.cpu().numpy().view(np.uint64), z["census_r"])
    nb = (2 * kw["rx"] + 1) * (2 * kw["ry"] + 1) - 1
    aw, at = kw.get("ad_weight", 0.0), kw.get("ad_trunc", 20.0)
    pw, pt = kw.get("prior_w", 0.0), kw.get("prior_t", 4.0)
    inv = nb + aw * at + pw * pt + 4.0
    base = _cu(z["base"], torch.int32)
    C = sv.band_costs(cl, cr, left, right, base, W, H, R, aw, at, inv, d0, pw, pt)
    assert np.array_equal(C.cpu().numpy(), z["C"].astype(np.float64))
    for i, (dx, dy) in enumerate(DIRS8):
        S = torch.zeros(W * H * K, dtype=torch.float64, device="cuda")
        sv._aggregate_path(C, base, W, H, K, kw["P1"], kw["P2"], dx, dy, S)
        assert np.array_equal(S.cpu().numpy(), z["L_dirs"][i].astype(np.float64)), (dx, dy)
    out, conf = sv.refine_fullres(left, right, d0, W, H, return_conf=True, **kw)
    assert np.array_equal(out.cpu().numpy(), z["out"])
    assert np.array_equal(conf.cpu().numpy(), z["conf"])
    assert np.array_equal(sv.left_to_right(d0, W, H).cpu().numpy(), z["d0r"])
    lo, st = sv.lr_refine(left, right, d0, W, H, tol=float(z["tol"]), **kw)
    assert np.array_equal(st.cpu().numpy().astype(np.int8), z["lr_status"])
    assert np.array_equal(lo.cpu().numpy(), z["lr_out"])
    assert np.array_equal(sv.median_filter(lo, W, H).cpu().numpy(), z["median"])


def test_refine_disparity_plain_golden(sv):
    z = np.load(os.path.join(GOLDEN, "refine_disparity_plain.npz"))
    W, H = int(z["W"]), int(z["H"])
    out, st = sv.refine_disparity(z["left"], z["right"], None, z["disp0"], W, H, planes=False)
    assert np.array_equal(out, z["out"])
    assert np.array_equal(st.astype(np.int8), z["status"])


def test_median_golden(sv):
    z = np.load(os.path.join(GOLDEN, "median.npz"))
    W, H = int(z["W"]), int(z["H"])
    for r, p in ((1, 1), (1, 2), (2, 1)):
        got = sv.median_filter(z["disp"], W, H, radius=r, passes=p)
        assert np.array_equal(got, z["r%d_p%d" % (r, p)]), (r, p)


def test_left_to_right_golden(sv):
    z = np.load(os.path.join(GOLDEN, "left_to_right.npz"))
    W, H = int(z["W"]), int(z["H"])
    assert np.array_equal(sv.left_to_right(z["dl"], W, H), z["dr"])


def test_list_inputs_return_lists(sv):
    z = np.load(os.path.join(GOLDEN, "refine_defaults.npz"))
    W, H = int(z["W"]), int(z["H"])
    out = sv.refine_fullres(list(z["left"]), list(z["right"]), list(z["disp0"]), W, H)
    assert isinstance(out, list) and len(out) == W * H
    assert np.array_equal(np.array(out), z["out"])


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_band_pair_uint8_matches_oracle(sv, seed):
    """The band-refinement workload (192x144, R=10, census 9x7, P1=12,
    P2=150, tol=2) from uint8 images: identical to the oracle run on the
    float gray the reference would read."""
    print("band pair seed", seed, flush=True)
    p = sv.synth.band_pair(seed)
    W, H = p["W"], p["H"]
    kw = dict(R=10, P1=12.0, P2=150.0, rx=4, ry=3)
    gl, gr = sv.synth.gray(p["left"]), sv.synth.gray(p["right"])
    l8, r8 = _cu(p["left"], torch.uint8), _cu(p["right"], torch.uint8)
    d0 = _cu(p["disp0"])
    out, conf = sv.refine_fullres(l8, r8, d0, W, H, return_conf=True, **kw)
    ref, rconf = oracle.refine_fullres(gl, gr, p["disp0"], W, H, return_conf=True, **kw)
    assert np.array_equal(out.cpu().numpy().ravel(), ref)
    assert np.array_equal(conf.cpu().numpy().ravel(), rconf)
    lo, st = sv.refine_disparity(l8, r8, None, d0, W, H, planes=False)
    rlo, rst = oracle.refine_disparity_plain(gl, gr, p["disp0"], W, H)
    assert np.array_equal(st.cpu().numpy().ravel().astype(np.int32), rst)
    assert np.array_equal(lo.cpu().numpy().ravel(), rlo)


def test_batched_equals_single(sv):
    """Three pairs in one launch give the same result as three launches."""
    pairs = [sv.synth.band_pair(10 + i, W=96, H=72) for i in range(3)]
    W, H = 96, 72
    L8 = _cu(np.stack([p["left"] for p in pairs]), torch.uint8)
    R8 = _cu(np.stack([p["right"] for p in pairs]), torch.uint8)
    D0 = _cu(np.stack([p["disp0"] for p in pairs]))
    out, st = sv.lr_refine(L8, R8, D0, W, H, tol=2.0, R=10, P1=12.0, P2=150.0, rx=4, ry=3)
    for i in range(3):
        o1, s1 = sv.lr_refine(L8[i], R8[i], D0[i], W, H, tol=2.0, R=10, P1=12.0, P2=150.0, rx=4, ry=3)
        assert torch.equal(out[i], o1) and torch.equal(st[i], s1)


def test_pipeline_size_matches_oracle(sv):
    """Full 460x370 pair, disparity 10-80, both views + median."""
    p = sv.synth.pipeline_pair(0)
    W, H = p["W"], p["H"]
    gl, gr = sv.synth.gray(p["left"]), sv.synth.gray(p["right"])
    out, st = sv.refine_disparity(_cu(p["left"], torch.uint8), _cu(p["right"], torch.uint8), None,
                                  _cu(p["disp0"]), W, H, planes=False)
    rout, rst = oracle.refine_disparity_plain(gl, gr, p["disp0"], W, H)
    assert np.array_equal(st.cpu().numpy().ravel().astype(np.int32), rst)
    assert np.array_equal(out.cpu().numpy().ravel(), rout)


@pytest.mark.parametrize("R", [42, 122])
def test_full_range_matches_oracle(sv, R):
    """Full-range SGM (eval_real.py:143-146: constant centre R, K = 2R+1)
    with 4 and 8 lanes per scanline."""
    p = sv.synth.make_pair(200, 60, 5, dmin=10.0, dmax=min(2.0 * R - 4, 120.0))
    W, H = 200, 60
    gl, gr = sv.synth.gray(p["left"]), sv.synth.gray(p["right"])
    d0 = np.full(W * H, float(R))
    kw = dict(R=R, P1=12.0, P2=150.0, rx=4, ry=3)
    out, conf = sv.refine_fullres(_cu(p["left"], torch.uint8), _cu(p["right"], torch.uint8),
                                  _cu(d0), W, H, return_conf=True, **kw)
    ref, rconf = oracle.refine_fullres(gl, gr, d0, W, H, return_conf=True, **kw)
    assert np.array_equal(out.cpu().numpy().ravel(), ref)
    assert np.array_equal(conf.cpu().numpy().ravel(), rconf)


def test_float_path_wide_band_matches_oracle(sv):
    """General float64 path with several lanes per scanline (K = 41)."""
    p = sv.synth.make_pair(64, 40, 9, dmin=4.0, dmax=20.0)
    W, H = 64, 40
    gl, gr = sv.synth.gray(p["left"]), sv.synth.gray(p["right"])
    kw = dict(R=20, P1=2.5, P2=30.5, ad_weight=0.25, ad_trunc=15.0, uniq=0.02, rx=3, ry=3,
              prior_w=0.1, prior_t=5.0)
    lo, st = sv.lr_refine(_cu(gl), _cu(gr), _cu(p["disp0"]), W, H, tol=1.5, **kw)
    rlo, rst = oracle.lr_refine(gl, gr, p["disp0"], W, H, tol=1.5, **kw)
    assert np.array_equal(st.cpu().numpy().ravel().astype(np.int32), rst)
    assert np.array_equal(lo.cpu().numpy().ravel(), rlo)


def test_u16_path_large_p2(sv):
    """P2 large enough that path costs leave uint8 (invalid + P2 > 255)."""
    p = sv.synth.band_pair(4, W=80, H=48)
    W, H = 80, 48
    gl, gr = sv.synth.gray(p["left"]), sv.synth.gray(p["right"])
    kw = dict(R=10, P1=20.0, P2=400.0, rx=4, ry=3)
    out = sv.refine_fullres(_cu(p["left"], torch.uint8), _cu(p["right"], torch.uint8),
                            _cu(p["disp0"]), W, H, **kw)
    assert np.array_equal(out.cpu().numpy().ravel(), oracle.refine_fullres(gl, gr, p["disp0"], W, H, **kw))


@pytest.mark.parametrize("P2", [65469.0, 65470.0, 70000.0])
def test_u16_path_limit(sv, P2):
    """Path costs reach invalid + P2 (invalid = 66 for the 9x7 census): up to
    65535 they stay on the u16 integer path, beyond it the float64 path
    takes over; both exact against the oracle."""
    p = sv.synth.band_pair(6, W=48, H=24)
    W, H = 48, 24
    gl, gr = sv.synth.gray(p["left"]), sv.synth.gray(p["right"])
    kw = dict(R=5, P1=7.0, P2=P2, rx=4, ry=3)
    out = sv.refine_fullres(_cu(p["left"], torch.uint8), _cu(p["right"], torch.uint8),
                            _cu(p["disp0"]), W, H, **kw)
    assert np.array_equal(out.cpu().numpy().ravel(), oracle.refine_fullres(gl, gr, p["disp0"], W, H, **kw))


@pytest.mark.parametrize("W,H", [(460, 370), (203, 77), (64, 32), (65, 33), (7, 5), (9, 1),
                                 (1, 6), (4, 40), (130, 2)])
def test_census_9x7_uint8_matches_oracle(sv, W, H):
    """The 9x7 census of 8-bit images (four pixels per thread, packed
    compares) against the oracle on the gray image: partial pixel groups,
    tiles and rows, images narrower than the window, ties (few gray
    levels), three images per launch."""
    print("census 9x7", W, H, flush=True)
    rng = np.random.default_rng(W * 1000 + H)
    imgs = [rng.integers(0, 256, size=(H, W)), rng.integers(0, 4, size=(H, W)),
            np.full((H, W), 255)]
    batch = torch.as_tensor(np.stack(imgs).astype(np.uint8)).cuda()
    got = sv.census(batch, W, H, 4, 3).cpu().numpy().view(np.uint64).reshape(3, -1)
    for n, img in enumerate(imgs):
        ref = oracle.census(img.astype(np.float64).ravel() / 255.0, W, H, 4, 3)
        assert np.array_equal(got[n], np.asarray(ref, dtype=np.uint64).ravel()), n


def test_errors(sv):
    img = torch.zeros(20 * 10, dtype=torch.uint8, device="cuda")
    with pytest.raises(NotImplementedError):
        sv.census(img, 20, 10, rx=5, ry=5)
    with pytest.raises(ValueError):
        sv.refine_fullres(img, img, torch.zeros(7, dtype=torch.float64, device="cuda"), 20, 10)
    # planes=True with an RGB image of the wrong size
    with pytest.raises(ValueError):
        sv.refine_disparity(img, img, [0.5] * 599, torch.zeros(200, dtype=torch.float64,
                                                                   device="cuda"), 20, 10)
    # plane_snap's SLIC step outside what the GPU kernels support
    with pytest.raises(ValueError):
        sv.plane_snap(torch.zeros(200, dtype=torch.float64, device="cuda"), [0.5] * 600, 20, 10,
                      step=17)


@pytest.mark.parametrize("W,H", [(700, 9), (3600, 3)])
def test_left_to_right_random_rows(sv, W, H):
    """Row-parallel z-buffer warp and hole fill (and the one-thread-per-row
    kernel past 3500 columns) against the oracle: collisions, negative and
    huge disparities, long hole runs."""
    rng = np.random.default_rng(W)
    dl = rng.uniform(-3.0, 60.0, size=(H, W))
    dl[:, ::7] = -0.75
    dl[0, 100:300] = 1e9
    dl = np.round(dl * 8) / 8
    got = sv.left_to_right(_cu(dl), W, H).cpu().numpy().ravel()
    assert np.array_equal(got, oracle.left_to_right(dl, W, H))


def test_lr_check_random_rows(sv):
    """Status and fills of lr_refine at a width the row kernel handles,
    against the oracle on a synthetic pair (both views refined)."""
    p = sv.synth.make_pair(640, 24, 12, dmin=5.0, dmax=60.0)
    W, H = 640, 24
    gl, gr = sv.synth.gray(p["left"]), sv.synth.gray(p["right"])
    kw = dict(R=10, P1=12.0, P2=150.0, rx=4, ry=3)
    lo, st = sv.lr_refine(_cu(p["left"], torch.uint8), _cu(p["right"], torch.uint8),
                          _cu(p["disp0"]), W, H, tol=2.0, **kw)
    rlo, rst = oracle.lr_refine(gl, gr, p["disp0"], W, H, tol=2.0, **kw)
    assert np.array_equal(st.cpu().numpy().ravel().astype(np.int32), rst)
    assert np.array_equal(lo.cpu().numpy().ravel(), rlo)


@pytest.fixture
def sgm_split():
    """Packed SGM kernels with two lanes per scanline (tuning knob 8 = 2);
    the knobs are reset afterwards."""
    import importlib
    A = importlib.import_module("from-scratch-stereo-vision_amd.autotune")
    A.set_knob("sgm_split", 2)
    yield A
    A.reset()


@pytest.mark.parametrize("name", REFINE_CASES)
@pytest.mark.parametrize("variant", [(2, 8), (4, 16), (8, 16)], ids=lambda v: "pf%d_ht%d" % v)
def test_two_lanes_per_scanline_golden(sv, sgm_split, name, variant):
    """The split-band SGM kernels (each scanline's band over a lane pair)
    reproduce every golden vector exactly, under each prefetch depth and
    horizontal tile width, including images narrower and shorter than a
    wave's 32 scanlines."""
    sgm_split.set_knob("sgm_prefetch", variant[0])
    sgm_split.set_knob("sgm_htile", variant[1])
    test_refinement_golden(sv, name)


@pytest.mark.parametrize("seed", [0, 1])
def test_two_lanes_per_scanline_matches_oracle(sv, sgm_split, seed):
    """Split-band SGM on the band workload and the full 460x370 pipeline
    pair: identical to the oracle."""
    test_band_pair_uint8_matches_oracle(sv, seed)
    if seed == 0:
        test_pipeline_size_matches_oracle(sv)


@pytest.fixture
def wta_coalesced():
    """Band epilogue reading coalesced 64-pixel runs through LDS (tuning
    knob 9 = 2); the knobs are reset afterwards."""
    import importlib
    A = importlib.import_module("from-scratch-stereo-vision_amd.autotune")
    A.set_knob("wta_reads", 2)
    yield A
    A.reset()


@pytest.mark.parametrize("name", ["refine_band_r10", "refine_jumps", "refine_paths4",
                                  "refine_tiny_7x5", "refine_tiny_9x1", "refine_tiny_1x6"])
def test_coalesced_epilogue_golden(sv, wta_coalesced, name):
    """The coalesced epilogue reproduces the golden vectors (K = 21 cases
    take it, others the per-pixel kernels), including images of fewer than
    64 pixels and 4 paths."""
    test_refinement_golden(sv, name)


def test_coalesced_epilogue_matches_oracle(sv, wta_coalesced):
    """Pixel counts not a multiple of 64 (a partial last run), the band
    workload and the full 460x370 pipeline pair: identical to the oracle."""
    test_band_pair_uint8_matches_oracle(sv, 3)
    test_pipeline_size_matches_oracle(sv)
    p = sv.synth.band_pair(8, W=67, H=13)
    W, H = 67, 13
    gl, gr = sv.synth.gray(p["left"]), sv.synth.gray(p["right"])
    kw = dict(R=10, P1=12.0, P2=150.0, rx=4, ry=3)
    out, st = sv.lr_refine(_cu(p["left"], torch.uint8), _cu(p["right"], torch.uint8),
                           _cu(p["disp0"]), W, H, tol=2.0, **kw)
    rout, rst = oracle.lr_refine(gl, gr, p["disp0"], W, H, tol=2.0, **kw)
    assert np.array_equal(st.cpu().numpy().ravel().astype(np.int32), rst)
    assert np.array_equal(out.cpu().numpy().ravel(), rout)
