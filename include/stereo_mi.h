/*
 * stereo_mi.h -- C ABI of libstereo_mi.so, the gfx950 (MI355X) implementation
 * of the dense-disparity path of Ray-5106/From-Scratch-Stereo-Vision.
 *
 * The reference has no FFI: its boundary is the set of module-level Python
 * functions of stereonet.py that harnesses bind by name (eval_real.py:41-44,
 * train_parallel.py:25-26).  Each entry point below replaces one of those
 * functions (cited file:line, paths relative to the reference checkout) and
 * keeps its argument meaning; the Python package binds them with ctypes.
 *
 * Conventions
 *   - every pointer is a device pointer (hipMalloc / torch CUDA memory);
 *   - arrays are row-major; B independent pairs are stacked on the slowest
 *     axis ([B][H][W]); images may be u8, f32 or f64 (SV_U8/SV_F32/SV_F64);
 *     disparities are f64, statuses u8;
 *   - `stream` is a hipStream_t (NULL = the default stream); calls are
 *     asynchronous on it and never synchronise or allocate;
 *   - scratch memory comes from a caller workspace sized by the matching
 *     *_workspace_bytes query;
 *   - return 0 on success, a negative SV_E* code otherwise, with a message
 *     in sv_last_error() (thread-local).
 */
#ifndef STEREO_MI_H
#define STEREO_MI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SV_OK 0
#define SV_EINVAL -1
#define SV_EUNSUPPORTED -2
#define SV_EHIP -3
#define SV_EWORKSPACE -4

#define SV_U8 0
#define SV_F32 1
#define SV_F64 2

/* ------------------------------------------------------------ library */

/* Text of the last error on the calling thread. */
const char* sv_last_error(void);
/* ABI version (increments on incompatible changes). */
int sv_version(void);

/* ------------------------------------------- full-resolution refinement */

/* census(img, W, H, rx=2, ry=2), stereonet.py:5115-5142.  One 64-bit word
 * per pixel, first window offset in the most significant used bit.
 * SV_EUNSUPPORTED when (2rx+1)(2ry+1)-1 > 64. */
int sv_census(const void* img, int dtype, int n_images, int W, int H, int rx, int ry,
              uint64_t* out, void* stream);

/* band_costs(cl, cr, gl, gr, base, W, H, R, ad_weight, ad_trunc, invalid,
 * prior, prior_w, prior_t), stereonet.py:5156-5189.  C is [B][H][W][2R+1]
 * f64; prior may be NULL when prior_w <= 0. */
int sv_band_costs(const uint64_t* cl, const uint64_t* cr, const void* gl, const void* gr,
                  int dtype, const int* base, int B, int W, int H, int R, double ad_weight,
                  double ad_trunc, double invalid, const double* prior, double prior_w,
                  double prior_t, double* C, void* stream);

/* _aggregate_path(C, base, W, H, K, P1, P2, dx, dy, S), stereonet.py:5192-5241:
 * adds one path's costs into S ([B][H][W][K] f64, updated in place). */
int sv_aggregate_path_workspace_bytes(int B, int W, int H, int K, size_t* bytes);
int sv_aggregate_path(const double* C, const int* base, int B, int W, int H, int K, double P1,
                      double P2, int dx, int dy, double* S, void* ws, size_t ws_bytes,
                      void* stream);

/* Workspace for sv_refine_fullres (mode 0), sv_lr_refine (mode 1) and
 * sv_refine_disparity (mode 2) at the given sizes and parameters. */
int sv_refine_workspace_bytes(int mode, int B, int W, int H, int R, double P1, double P2,
                              double ad_weight, double ad_trunc, int paths, int rx, int ry,
                              double prior_w, double prior_t, size_t* bytes);

/* refine_fullres(left, right, disp0, W, H, R=3, P1=1.0, P2=12.0,
 * ad_weight=0.0, ad_trunc=20.0, paths=8, uniq=0.0, rx=2, ry=2, prior_w=0.0,
 * prior_t=4.0, return_conf=False), stereonet.py:5244-5305.  out (and conf
 * when not NULL) are [B][H][W] f64. */
int sv_refine_fullres(const void* left, const void* right, int dtype, const double* disp0, int B,
                      int W, int H, int R, double P1, double P2, double ad_weight,
                      double ad_trunc, int paths, double uniq, int rx, int ry, double prior_w,
                      double prior_t, double* out, double* conf, void* ws, size_t ws_bytes,
                      void* stream);

/* left_to_right(dl, W, H), stereonet.py:5317-5340. */
int sv_left_to_right(const double* dl, int B, int W, int H, double* dr, void* stream);

/* lr_refine(left, right, disp0, W, H, tol=1.0, **kw), stereonet.py:5343-5391.
 * status is [B][H][W] u8 (0 ok, 1 occluded, 2 mismatch). */
int sv_lr_refine(const void* left, const void* right, int dtype, const double* disp0, int B,
                 int W, int H, double tol, int R, double P1, double P2, double ad_weight,
                 double ad_trunc, int paths, double uniq, int rx, int ry, double prior_w,
                 double prior_t, double* out, uint8_t* status, void* ws, size_t ws_bytes,
                 void* stream);

/* median_filter(disp, W, H, radius=1, passes=1), stereonet.py:2231-2258.
 * Radius 0..2 on the GPU; tmp ([B][H][W] f64) is needed when passes > 1;
 * out must not alias disp. */
int sv_median_filter(const double* disp, int B, int W, int H, int radius, int passes,
                     double* out, double* tmp, void* stream);

/* refine_disparity(left, right, rgb, disp0, W, H, R=10, P1=12.0, P2=150.0,
 * rx=4, ry=3, tol=2.0, planes=False), stereonet.py:5865-5884: lr_refine then
 * one 3x3 median pass. */
int sv_refine_disparity(const void* left, const void* right, int dtype, const double* disp0,
                        int B, int W, int H, int R, double P1, double P2, int rx, int ry,
                        double tol, double* out, uint8_t* status, void* ws, size_t ws_bytes,
                        void* stream);

#ifdef __cplusplus
}
#endif

#endif /* STEREO_MI_H */
