"""MI355X-native (gfx950) dense-disparity path of
Ray-5106/From-Scratch-Stereo-Vision.

The package re-exports the reference's stereonet.py names on the path with
the same signatures, so ``SN = vars(package)`` stands in for the namespace
the reference's harnesses build with ``exec`` (eval_real.py:41-44).  All
arithmetic runs in hand-written HIP kernels (libstereo_mi.so, built by
build.py); torch supplies device memory and streams only.
"""

from .refine import (census, band_costs, _aggregate_path, refine_fullres, full_range, _mirror1,
                     left_to_right, lr_refine, median_filter, refine_disparity, slic, plane_snap)
from .net import (StereoNet, Adam, LEVEL_WEIGHTS, CE_WEIGHT, train_grads, train_grads_batch,
                  train_step, flat_params, set_flat_params, flat_grads, set_flat_grads,
                  save_weights, read_checkpoint_config, load_weights, load_model, predict,
                  predict_batch, predict_at_scale, predict_auto, predict_auto_batch, widen)
from .ragged import RaggedBatch, refine_ragged, net_band_ragged, CapturedPipeline
from .consistency import (mirror_h, lr_consistency, fill_invalid, predict_consistent,
                          predict_consistent_batch)
from .images import (read_pnm, read_png, read_image, read_image_gpu, read_images_gpu, write_pgm,
                     write_ppm)
from .geometry import (Camera, read_calib, robust_min_disparity, disparity_to_points,
                       min_filter_arg, smooth_background, background_region, grid_mesh)
from .mesh import (delaunay, cull_triangles, cull_slanted, select_points, write_obj,
                   background_layer, merge_layer)
from .splat import (GaussianCloud, gaussians_from_mesh, SplatRenderer, CloudAdam,
                    refine_gaussians, look_at, cloud_centroid, orbit_views, reconstruct,
                    ALPHA_MAX, IDENTITY_R, IDENTITY_T)
from . import autotune, scenes, synth, evaluation, geometry, mesh, splat

__all__ = ["census", "band_costs", "_aggregate_path", "refine_fullres", "full_range", "_mirror1",
           "left_to_right", "lr_refine", "median_filter", "refine_disparity", "slic", "plane_snap",
           "StereoNet", "Adam",
           "LEVEL_WEIGHTS", "CE_WEIGHT", "train_grads", "train_grads_batch", "train_step",
           "flat_params", "set_flat_params", "flat_grads", "set_flat_grads", "save_weights",
           "read_checkpoint_config", "load_weights", "load_model", "predict", "predict_batch",
           "predict_at_scale", "predict_auto", "predict_auto_batch", "widen", "RaggedBatch",
           "refine_ragged", "net_band_ragged", "CapturedPipeline", "mirror_h", "lr_consistency",
           "fill_invalid", "predict_consistent", "predict_consistent_batch", "read_pnm", "read_png",
           "read_image", "read_image_gpu", "read_images_gpu", "write_pgm", "write_ppm",
           "Camera", "read_calib", "robust_min_disparity", "disparity_to_points", "min_filter_arg",
           "smooth_background", "background_region", "grid_mesh",
           "delaunay", "cull_triangles", "cull_slanted", "select_points", "write_obj",
           "background_layer", "merge_layer", "GaussianCloud", "gaussians_from_mesh",
           "SplatRenderer", "CloudAdam", "refine_gaussians", "look_at", "cloud_centroid",
           "orbit_views", "reconstruct", "ALPHA_MAX", "IDENTITY_R", "IDENTITY_T",
           "autotune", "scenes", "synth", "evaluation", "geometry", "mesh", "splat"]
