"""Write tests/golden/eval.npz: the reference Middlebury harness's own
outputs (eval_real.py:55-127 -- scenes, load_gt, masks, score, constant) on
ground-truth PNG files made here.

Run in the development container only (the reference checkout is not present
anywhere else):

    python tests/golden/make_eval.py /root/reference

eval_real.py is imported from the reference checkout (its module body loads
stereonet.py as source text); only the functions above are called.
Outputs are compressed .npz (no pickles).
"""

import importlib.util
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import make_images  # noqa: E402


def load_eval(ref_root):
    spec = importlib.util.spec_from_file_location("eval_real_reference",
                                                  os.path.join(ref_root, "eval_real.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def gt_pair(rng, W, H, scale):
    """Left/right ground-truth grey levels (0 = unknown): a slanted plane
    plus a box, the right view warped from the left with holes; some left
    pixels near x = 0 whose x - d + 0.5 falls in (-1, 0)."""
    y, x = np.mgrid[0:H, 0:W]
    d = 6.0 + 0.08 * x + 0.05 * y
    d[H // 4:H // 2, W // 3:W // 2] += 9.0
    lv = np.round(d * scale)
    lv[rng.random((H, W)) < 0.08] = 0
    lv[:, 0] = np.round((0.5 + rng.random(H) * 0.9) * scale)  # near the left edge
    rv = np.zeros_like(lv)
    for yy in range(H):
        for xx in range(W):
            if lv[yy, xx]:
                xr = int(xx - lv[yy, xx] / scale + 0.5)
                if 0 <= xr < W:
                    rv[yy, xr] = max(rv[yy, xr], lv[yy, xx] + rng.integers(-4, 5))
    rv[rng.random((H, W)) < 0.05] = 0
    return np.clip(lv, 0, 255).astype(np.uint8), np.clip(rv, 0, 255).astype(np.uint8)


def gray_png(rng, lv, ctype=0):
    """An 8-bit PNG of grey levels lv (ctype 0), or of an RGB image with all
    three channels equal to lv (ctype 2)."""
    H, W = lv.shape
    if ctype == 0:
        rows = [bytes(lv[y]) for y in range(H)]
        bpp = 1
    else:
        rows = [bytes(np.repeat(lv[y], 3)) for y in range(H)]
        bpp = 3
    raw = make_images._filter(rows, bpp, [int(t) for t in rng.integers(0, 5, size=H)])
    import struct
    import zlib
    ihdr = struct.pack(">IIBBBBB", W, H, 8, ctype, 0, 0, 0)
    return (b"\x89PNG\r\n\x1a\n" + make_images._chunk(b"IHDR", ihdr)
            + make_images._chunk(b"IDAT", zlib.compress(raw, 9)) + make_images._chunk(b"IEND", b""))


def main(ref_root):
    ev = load_eval(ref_root)
    rng = np.random.default_rng(77)
    data = {}
    cases = [("s3_gray", 53, 31, 3.0, 0), ("s4_gray", 47, 29, 4.0, 0), ("s3_rgb", 41, 23, 3.0, 2)]
    with tempfile.TemporaryDirectory() as td:
        for name, W, H, scale, ct in cases:
            lv, rv = gt_pair(rng, W, H, scale)
            fl, fr = gray_png(rng, lv, ct), gray_png(rng, rv, ct)
            pl, pr = os.path.join(td, name + "_l.png"), os.path.join(td, name + "_r.png")
            open(pl, "wb").write(fl)
            open(pr, "wb").write(fr)
            d, k, W1, H1 = ev.load_gt(pl, scale)
            sc = {"gtl": pl, "gtr": pr, "scale": scale}
            gt, known, no, _, _ = ev.masks(sc)
            pred = (np.asarray(gt) + rng.normal(0, 2.5, size=W * H)).tolist()
            data[name + "_fl"] = np.frombuffer(fl, np.uint8)
            data[name + "_fr"] = np.frombuffer(fr, np.uint8)
            data[name + "_meta"] = np.array([W, H, scale, ct], dtype=np.float64)
            data[name + "_gt"] = np.array(d)
            data[name + "_known"] = np.array(k, dtype=bool)
            data[name + "_nonocc"] = np.array(no, dtype=bool)
            data[name + "_pred"] = np.array(pred)
            data[name + "_score_nonocc"] = np.array(ev.score(pred, gt, no))
            data[name + "_score_all"] = np.array(ev.score(pred, gt, known))
            data[name + "_constant"] = np.array(ev.constant(gt, no))
        assert ev.score([1.0, 2.0], [1.0, 2.0], [False, False]) is None
        # a scene tree for scenes(): years, scene folders with and without GT
        tree = {"2006": ["Aloe", "Baby1", "Bowling1", "Cloth1", "Flowerpots", "Lampshade1"],
                "2005": ["Art", "Books"], "2003": ["cones", "teddy"]}
        for year, names in tree.items():
            for i, n in enumerate(names):
                d = os.path.join(td, "root", year, n)
                os.makedirs(d)
                if not (year == "2006" and i == 3):  # one scene without ground truth
                    open(os.path.join(d, ev.LAYOUT[year][2]), "wb").write(b"x")
        os.makedirs(os.path.join(td, "root", "2001"))
        got = ev.scenes(os.path.join(td, "root"))
        data["scenes"] = np.array(["%s/%s/%s" % (s["year"], s["name"], s["split"]) for s in got])
        data["tree"] = np.array(["%s/%s" % (y, n) for y, ns in tree.items() for n in ns])
    data["cases"] = np.array([c[0] for c in cases])
    np.savez_compressed(os.path.join(HERE, "eval.npz"), **data)
    print("wrote eval.npz")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
