"""Write tests/golden/splat.npz: the reference's own outputs of the mesh and
Gaussian-splatting half of the downstream 3D (stereonet.py:2038-2199,
:2324-2484, :2485-3041, :3042-3187, :5674-5769) on inputs made here.

Run in the development container only (the reference checkout is not present
anywhere else):

    python tests/golden/make_splat.py /root/reference

The reference is loaded the way its own harnesses load it (exec of the
source text).  It is used for what no closed form pins:
* delaunay's triangle choice and order on degenerate (integer-grid,
  cocircular) vertex sets, which downstream order depends on;
* select_points' sampling (grid walk, ranking ties, minimum spacing);
* the float64 arithmetic of cull_triangles / cull_slanted,
  gaussians_from_mesh, the splat renderer's forward and backward,
  CloudAdam and refine_gaussians, all of which use the reference's own
  fsqrt / fexp / flog series rather than library routines;
* background_layer / merge_layer and write_obj's text.
Outputs are compressed .npz (no pickles).
"""

import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def load_reference(ref_root):
    src = open(os.path.join(ref_root, "stereonet.py")).read().replace(
        "if __name__ == '__main__':\n    run_demo()\n", "")
    SN = {"__name__": "stereonet_reference"}
    exec(src, SN)
    return SN


def scene(rng, W, H):
    """Disparity: a slanted wall, a near box with a sharp edge, a far strip
    below min_disp; colour: smooth gradients plus a checker on the box."""
    y, x = np.mgrid[0:H, 0:W].astype(np.float64)
    d = 5.0 + 0.04 * x + 0.03 * y
    box = (y >= H // 4) & (y < H // 2 + 4) & (x >= W // 3) & (x < W // 3 + W // 4)
    d[box] = 14.0 + 0.02 * x[box]
    d[:, :3] = 0.2
    d += rng.normal(0, 0.05, size=d.shape)
    r = 0.2 + 0.6 * x / W
    g = 0.3 + 0.5 * y / H
    b = 0.5 + 0.3 * np.sin(x / 5.0) * np.cos(y / 7.0)
    chk = ((x.astype(int) // 3 + y.astype(int) // 3) % 2).astype(np.float64)
    r[box] = 0.1 + 0.8 * chk[box]
    rgb = np.stack([r, g, b], axis=-1)
    return d.ravel(), rgb.ravel()


def cloud_arrays(gc):
    keys = ("mx", "my", "mz", "sxx", "sxy", "sxz", "syy", "syz", "szz", "cr", "cg", "cb",
            "raw_op", "raw_sc")
    return np.array([getattr(gc, k) for k in keys], dtype=np.float64)


def grad_arrays(gc):
    return np.array([gc.g_cr, gc.g_cg, gc.g_cb, gc.g_op, gc.g_sc], dtype=np.float64)


def main(ref_root):
    SN = load_reference(ref_root)
    out = {}
    # Delaunay on random points (the reference test's setup) and on a
    # degenerate integer grid with duplicated rows of collinear points
    rng = SN["Rng"](4242)
    pts = [(rng.range(0, 100), rng.range(0, 100)) for _ in range(160)]
    out["dl_rand_pts"] = np.array(pts)
    out["dl_rand_tris"] = np.array(SN["delaunay"](pts), dtype=np.int32)
    grid = [(x, y) for y in range(0, 40, 5) for x in range(0, 50, 7)] + [(3, 3), (17, 22), (44, 1)]
    out["dl_grid_pts"] = np.array(grid, dtype=np.float64)
    out["dl_grid_tris"] = np.array(SN["delaunay"](grid), dtype=np.int32)

    nrng = np.random.default_rng(17)
    for name, W, H, npts in (("s48", 48, 36, 90), ("s64", 64, 40, 140)):
        disp, rgb = scene(nrng, W, H)
        dl, cl = disp.tolist(), rgb.tolist()
        cam = SN["Camera"].synthetic(W, H)
        out[name + "_disp"], out[name + "_rgb"] = disp, rgb
        out[name + "_wh"] = np.array([W, H, npts])
        pts3d = SN["disparity_to_points"](dl, W, H, cam, min_disp=0.4)
        valid = [not (i % 11 == 0 and (i // W) % 5 == 1) for i in range(W * H)]
        out[name + "_valid"] = np.array(valid)
        p2 = SN["select_points"](dl, W, H, None, target=npts)
        p2v = SN["select_points"](dl, W, H, valid, target=npts)
        out[name + "_sel"] = np.array(p2, dtype=np.int32)
        out[name + "_selv"] = np.array(p2v, dtype=np.int32)
        raw = SN["delaunay"](p2)
        out[name + "_tris_raw"] = np.array(raw, dtype=np.int32)
        tris = SN["cull_triangles"](p2, raw, pts3d, W)
        out[name + "_tris"] = np.array(tris, dtype=np.int32)
        out[name + "_tris_slant"] = np.array(SN["cull_slanted"](p2, tris, pts3d, cam, W),
                                             dtype=np.int32).reshape(-1, 3)
        gc = SN["gaussians_from_mesh"](p2, tris, pts3d, cl, W, H)
        out[name + "_gc"] = cloud_arrays(gc)
        r = SN["SplatRenderer"](cam, W, H)
        C0 = r.render(gc, SN["IDENTITY_R"], SN["IDENTITY_T"])
        out[name + "_render"] = np.array(C0)
        wts = [0.3 + 0.013 * (i % 97) - 0.002 * (i % 13) for i in range(3 * W * H)]
        out[name + "_bw_w"] = np.array(wts)
        gc.zero_grads()
        r.render(gc, SN["IDENTITY_R"], SN["IDENTITY_T"])
        r.backward(wts)
        out[name + "_grads"] = grad_arrays(gc)
        opt = SN["CloudAdam"](gc)
        opt.step()
        out[name + "_adam1"] = cloud_arrays(gc)
        views = SN["orbit_views"](gc, n_views=3)
        out[name + "_views"] = np.array([[a] + list(R) + list(t) for a, (R, t) in views])
        R, t = views[0][1]
        out[name + "_render_novel"] = np.array(r.render(gc, R, t, background=(0.1, 0.2, 0.3)))
        gc2 = SN["gaussians_from_mesh"](p2, tris, pts3d, cl, W, H, thickness=0.5,
                                        init_opacity=0.8, scale=1.3)
        out[name + "_gc_kw"] = cloud_arrays(gc2)
        hist = SN["refine_gaussians"](gc2, r, cl, iters=3, verbose=False)
        out[name + "_refine_hist"] = np.array(hist)
        out[name + "_refine_gc"] = cloud_arrays(gc2)
        with tempfile.TemporaryDirectory() as td:
            path = os.path.join(td, "m.obj")
            nv, nf = SN["write_obj"](path, p2, tris, pts3d, cl, W, H)
            out[name + "_obj"] = np.frombuffer(open(path, "rb").read(), dtype=np.uint8)
            out[name + "_obj_counts"] = np.array([nv, nf])
        lp, lt, l3, lrgb = SN["background_layer"](dl, cl, W, H, cam, npts)
        out[name + "_bgl_pts"] = np.array(lp, dtype=np.int32).reshape(-1, 2)
        out[name + "_bgl_tris"] = np.array(lt, dtype=np.int32).reshape(-1, 3)
        out[name + "_bgl_rgb"] = np.array(lrgb)
        out[name + "_bgl_p3"] = np.array([p if p is not None else (np.nan,) * 3 for p in l3])
        mp, mt, m3, mc, W2 = SN["merge_layer"](p2, tris, pts3d, cl, (lp, lt, l3, lrgb), W, H)
        out[name + "_merge_pts"] = np.array(mp, dtype=np.int32)
        out[name + "_merge_tris"] = np.array(mt, dtype=np.int32)
        out[name + "_merge_p3"] = np.array([p if p is not None else (np.nan,) * 3 for p in m3])
        out[name + "_merge_rgb"] = np.array(mc)
        with tempfile.TemporaryDirectory() as td:
            res = SN["reconstruct"](dl, cl, W, H, cam, n_points=npts, refine_iters=2,
                                    out_prefix=os.path.join(td, "sc"), n_views=2, verbose=False)
        out[name + "_rec_render"] = np.array(res["render"])
        out[name + "_rec_gc"] = cloud_arrays(res["gaussians"])
        out[name + "_rec_mesh"] = np.array(res["mesh"][2:])
    np.savez_compressed(os.path.join(HERE, "splat.npz"), **out)
    print("wrote splat.npz")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
