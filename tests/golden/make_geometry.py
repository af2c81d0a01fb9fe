"""Write tests/golden/geometry.npz: the reference's own outputs of Camera,
robust_min_disparity, disparity_to_points, min_filter_arg,
smooth_background, background_region and grid_mesh (stereonet.py:1951-2021,
:5566-5862) on disparity maps made here.

Run in the development container only (the reference checkout is not present
anywhere else):

    python tests/golden/make_geometry.py /root/reference

The reference is loaded the way its own harnesses load it (exec of the
source text).  Outputs are compressed .npz (no pickles).
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def load_reference(ref_root):
    src = open(os.path.join(ref_root, "stereonet.py")).read().replace(
        "if __name__ == '__main__':\n    run_demo()\n", "")
    SN = {"__name__": "stereonet_reference"}
    exec(src, SN)
    return SN


def disp_map(rng, W, H):
    """A far slanted wall, two near boxes (depth edges), noise, a few
    pixels below every floor and a strip at x < d."""
    y, x = np.mgrid[0:H, 0:W]
    d = 4.0 + 0.03 * x + 0.02 * y
    d[H // 5:H // 2, W // 4:W // 2] = 16.0 + 0.05 * x[H // 5:H // 2, W // 4:W // 2]
    d[H // 2:H - 3, W // 2 + 3:W - 4] = 11.0
    d += rng.normal(0, 0.2, size=d.shape)
    d[rng.random(d.shape) < 0.02] = 0.1
    return d.ravel()


def main(ref_root):
    SN = load_reference(ref_root)
    rng = np.random.default_rng(99)
    out = {}
    cases = [("w80", 80, 40), ("w40", 40, 33), ("w67", 67, 21)]
    for name, W, H in cases:
        d = disp_map(rng, W, H)
        dl = d.tolist()
        out[name + "_disp"] = d
        out[name + "_wh"] = np.array([W, H])
        cams = {"syn": SN["Camera"].synthetic(W, H),
                "mid": SN["Camera"](W * 1.1, W * 0.47, H * 0.52, 0.193, 2.5)}
        out[name + "_robust"] = np.array(SN["robust_min_disparity"](dl))
        for cn, cam in cams.items():
            out[name + "_cam_" + cn] = np.array([cam.fx, cam.cx, cam.cy, cam.baseline, cam.doffs])
            pts = SN["disparity_to_points"](dl, W, H, cam)
            out[name + "_pts_" + cn] = np.array([p if p is not None else (np.nan,) * 3 for p in pts])
            out[name + "_ptsvalid_" + cn] = np.array([p is not None for p in pts])
        m, i = SN["min_filter_arg"](dl, W, H, 3, 2)
        out[name + "_mfa_m"], out[name + "_mfa_i"] = np.array(m), np.array(i)
        m0, i0 = SN["min_filter_arg"](dl, W, H, 5, 0)
        out[name + "_mfa0_m"], out[name + "_mfa0_i"] = np.array(m0), np.array(i0)
        out[name + "_smooth"] = np.array(SN["smooth_background"](m0, dl, W, H, 2))
        mask, bm, bs = SN["background_region"](dl, W, H)
        out[name + "_bg_mask"], out[name + "_bg_m"], out[name + "_bg_src"] = \
            np.array(mask), np.array(bm), np.array(bs)
        for gname, kw, cn in (("g2f", dict(step=2, fill=True), "syn"),
                              ("g1n", dict(step=1, fill=False, min_disp=0.5), "mid"),
                              ("g3f", dict(step=3, fill=True, jump=0.8, jump_rel=0.1), "mid")):
            g = SN["grid_mesh"](dl, W, H, cams[cn], **kw)
            k = "%s_%s_" % (name, gname)
            for f in ("pos", "uv", "fill", "disp", "idx"):
                out[k + f] = np.array(g[f])
            out[k + "counts"] = np.array([g["main_tris"], g["fill_tris"]])
            out[k + "pivot"] = np.array(g["pivot"])
    out["cases"] = np.array([c[0] for c in cases])
    np.savez_compressed(os.path.join(HERE, "geometry.npz"), **out)
    print("wrote geometry.npz")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
