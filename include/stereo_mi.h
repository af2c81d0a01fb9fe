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

/* Full-range SGM, the control of eval_real.py:143-146:
 * refine_fullres(left, right, [float(R)] * (W*H), W, H, R=R, P1, P2, rx, ry)
 * (mode 0; conf as return_conf=True when not NULL) or lr_refine with that
 * constant centre (mode 1; status [B][H][W] u8).  Results equal
 * sv_refine_fullres / sv_lr_refine given the constant map.  The integer u8
 * path with 2R + 1 <= 125 (and W >= R in mode 1) runs constant-centre
 * kernels, which never read a centre map (sv_full_range_fast says which);
 * anything else the general kernels.  rtab (device int32 [B], may be NULL)
 * gives pair n its own centre in [0, R] within one launch (constant-centre
 * kernels only; R sizes the band).  No counterpart function in the
 * reference: it replaces the harness's refine_fullres call per pair. */
int sv_full_range_workspace_bytes(int mode, int B, int W, int H, int R, double P1, double P2,
                                  int paths, int rx, int ry, size_t* bytes);
int sv_full_range_fast(int mode, int W, int H, int R, double P1, double P2, int paths, int rx,
                       int ry, int* fast);
int sv_full_range(int mode, const void* left, const void* right, int dtype, int B, int W, int H,
                  int R, const int* rtab, double P1, double P2, int paths, int rx, int ry,
                  double tol, double* out, double* conf, uint8_t* status, void* ws,
                  size_t ws_bytes, void* stream);

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

/* mirror_h(img, C, W, H), stereonet.py:4974-4983: each of `rows` rows of W
 * elements reversed (rows = images x channels x H of a planar buffer).
 * dtype SV_U8, SV_F32 or SV_F64; out must not alias in. */
int sv_mirror_h(const void* in, int dtype, long long rows, int W, void* out, void* stream);

/* lr_consistency(disp_l, disp_r, W, H, tol=1.5), stereonet.py:4986-5008,
 * for B maps of W x H (dtype SV_F32 or SV_F64, evaluated in float64).
 * valid: [B][H][W] u8, 1 where the maps agree.  counts (B ints, may be
 * NULL) receives each map's number of agreeing pixels.  A NaN disparity,
 * on which the reference raises, is marked invalid. */
int sv_lr_consistency(const void* dl, const void* dr, int dtype, int B, int W, int H, double tol,
                      uint8_t* valid, int* counts, void* stream);

/* fill_invalid(disp, valid, W, H), stereonet.py:5011-5046, for B maps
 * (dtype SV_F32 or SV_F64); out must not alias disp. */
int sv_fill_invalid(const void* disp, const uint8_t* valid, int dtype, int B, int W, int H,
                    void* out, void* stream);

/* slic(rgb, W, H, step=12, compactness=10.0, iters=5), stereonet.py:5394-5448,
 * for B images: rgb [B][H][W][3] (SV_U8 read as value * (1/255), as the
 * reference's image reader scales 8-bit samples, or SV_F64 in [0,1]);
 * labels [B][H][W] int32 (-1 where no centre reached); *nseg (host, may be
 * NULL) receives the number of centres per image.  step 2..16.  The
 * workspace is sv_plane_snap_workspace_bytes(B, W, H, step).  status_out
 * (device int, may be NULL) is set non-zero if a segment had more than 256
 * pixels outside every window (results then incomplete). */
int sv_slic(const void* rgb, int rgb_dtype, int B, int W, int H, int step, double compactness,
            int iters, int* labels, int* nseg, int* status_out, void* ws, size_t ws_bytes,
            void* stream);

/* plane_snap(disp, rgb, W, H, valid=None, step=12, thr=1.0, min_inl=0.6,
 * trials=40, seed=12345), stereonet.py:5498-5563, for B images (f64 maps):
 * SLIC segments, a RANSAC plane per segment from the reference's xorshift32
 * stream, the least-squares refit on the inliers and the snap.  valid may
 * be NULL (all valid); valid_kind 0: non-zero bytes are valid, 1: zero
 * bytes are valid (an lr_refine status array).  trials must be 40, step
 * 2..16; out must not alias disp; status_out as for sv_slic. */
int sv_plane_snap_workspace_bytes(int B, int W, int H, int step, size_t* bytes);
int sv_plane_snap(const double* disp, const void* rgb, int rgb_dtype, const uint8_t* valid,
                  int valid_kind, int B, int W, int H, int step, double thr, double min_inl,
                  int trials, unsigned int seed, double* out, int* status_out, void* ws,
                  size_t ws_bytes, void* stream);

/* refine_disparity(left, right, rgb, disp0, W, H, R=10, P1=12.0, P2=150.0,
 * rx=4, ry=3, tol=2.0, planes=False), stereonet.py:5865-5884: lr_refine then
 * one 3x3 median pass. */
int sv_refine_disparity(const void* left, const void* right, int dtype, const double* disp0,
                        int B, int W, int H, int R, double P1, double P2, int rx, int ry,
                        double tol, double* out, uint8_t* status, void* ws, size_t ws_bytes,
                        void* stream);

/* No counterpart in the reference: refine_disparity(planes=False) exactly
 * as sv_refine_disparity, for callers that cannot allocate (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * workspace is the library's static device scratch, 96 MiB per GPU, held
 * until the process ends; no allocation of any kind is made (no stage
 * events, no side stream).  Limit: the workspace of B*W*H pixels at band
 * half-width R must fit the scratch (about 488 bytes per pixel at R = 10:
 * one 460x370 pair, or seven 192x144 pairs); larger calls return
 * SV_EWORKSPACE before anything is queued.  All calls on one GPU share its
 * scratch: calls that overlap on one GPU race. */
int sv_refine_disparity_noalloc(const void* left, const void* right, int dtype,
                                const double* disp0, int B, int W, int H, int R, double P1,
                                double P2, int rx, int ry, double tol, double* out,
                                uint8_t* status, void* stream);

/* The three refinement drivers over a ragged batch (pairs of different
 * sizes in one launch per stage): refine_fullres (mode 0, paths=8,
 * uniq=0), lr_refine (mode 1) and refine_disparity planes=False (mode 2),
 * stereonet.py:5244-5305, :5343-5391, :5865-5884, integer pipeline
 * (ad_weight = prior_w = 0, integral P1 and P2).  Pair n is dims[n] = (h, w)
 * (device int32 pairs; h <= maxH, w <= maxW, h*w <= slot) and occupies
 * pixels [n*slot, n*slot + h*w) of every per-pixel input and output, rows
 * packed at its own width.  status may be null in mode 0. */
int sv_refine_ragged_workspace_bytes(int mode, int B, long long slot, int maxW, int maxH, int R,
                                     double P1, double P2, int rx, int ry, size_t* bytes);
int sv_refine_ragged(int mode, const void* left, const void* right, int dtype, const double* disp0,
                     int B, const int* dims, int maxW, int maxH, long long slot, int R, double P1,
                     double P2, int rx, int ry, double tol, double* out, uint8_t* status, void* ws,
                     size_t ws_bytes, void* stream);

/* Tuning knobs (no reference counterpart): kernel variants with identical
 * results and different speed, set by the autotuner.  knob 0: prefetch
 * depth of the vertical/diagonal SGM kernel (2, 4, 8); 1: columns per LDS
 * tile of the horizontal SGM kernel (8, 16); 2: output rows per thread of
 * the tiled convolution (1, 4); 3: output columns per tile of the tiled
 * convolution (16, 32, 64); 4: weight source of the tiled convolution (1
 * LDS, 2 scalar cache); 5: horizontal SGM paths after the others (1) or on a
 * second, library-owned stream beside them (2; not while the caller's
 * stream is being captured into a graph: then after the others); 6: convolution engine (1
 * vector ALUs, 2 matrix cores with 3D layers sliding over depth, 3 matrix
 * cores with one depth per tile, 4 vector ALUs staging one input channel at
 * a time, 5 matrix cores reading 4-channel blocks per LDS read with the next
 * tile prefetched, 2D layers of 2 to 12 output channels; 6 matrix cores
 * over a column of output depths per tile, 3D layers of up to 16 output
 * channels; 7 direct, one output position per thread without LDS staging,
 * any output channel count, the first 3D layer's cost volume formed on the
 * fly; other shapes fall back to 2/3); 7:
 * 64-pixel groups per wave of the
 * matrix-core convolution (1, 2, 4); 8: lanes per scanline of the packed
 * SGM kernels (1, 2); 9: path-cost reads of the band epilogue (K = 21, u8
 * path costs): one lane per pixel (1) or coalesced 64-pixel runs summed per
 * lane and handed out through LDS (2); 10: output depths per tile of the
 * depth-column 3D kernel (4, 6, 12); 11: workgroup order of the two-lane
 * vertical/diagonal SGM kernel: in launch order, spread over the XCDs (1),
 * or one contiguous run of (columns, direction, pair) blocks per XCD, so a
 * pair's paths share that XCD's L2 (2); 12: scheduling of the two-lane SGM
 * kernels: one workgroup per block of 32 scanlines (1), or persistent
 * workgroups, as many as fit the device at once, pulling (image,
 * direction, scanline block) items from eight queues, one per XCD (a
 * workgroup starts on the queue of the XCD it runs on), built
 * on the device before each launch with the longest scanlines first and the
 * images dealt round-robin, a workgroup that empties its own queue going on
 * to the others (2), or the same with one queue for the whole device (3);
 * batches of more than 4096 images (pairs x views) keep the grid.
 * Value 0 restores the built-in
 * heuristic (LDS weights; horizontal paths after the others; matrix cores
 * for 4 to 16 output channels, 3D layers over depth columns of 6, or 4 when
 * 6 leaves CUs idle; the per-channel vector kernel when the input channels
 * do not fit one staging pass).
 * Process-wide; set between launches, not while kernels are being queued
 * from other threads. */
int sv_tune_set(int knob, int value);
int sv_tune_get(int knob, int* value);

/* Opt-in stage timing of the refinement entry points: sv_profile_enable(1)
 * clears and starts recording HIP events around each stage on the caller's
 * stream; sv_profile_read sums a stage's ("census", "costs", "sgm", "wta",
 * "lr", "median") elapsed milliseconds.  Instrumentation, no reference
 * counterpart. */
int sv_profile_enable(int on);
int sv_profile_read(const char* stage, double* total_ms, int* launches);

/* ------------------------------------------------------ StereoNet (FP32)
 *
 * Ragged batches: a tensor of N samples is N slots of `stride` floats;
 * sample n is dense [C][D][h_n][w_n] at ptr + n*stride, with (h_n, w_n)
 * read from a device table `dims` of 2N ints.  max_h/max_w size the grid.
 * Weights follow the reference layout [OC][IC][KD][3][3] (KD = 1 for 2D).
 */

/* Conv2d.forward (stereonet.py:489-526) and Conv3d.forward (:596-641).
 * xr != NULL: the input is CostVolume.forward(x, xr) (:771-799), formed on
 * the fly.  mode 0: bias; 1: LeakyReLU(slope) (:705-710); 2: max(resid +
 * conv, 0), the Refinement output (:1150-1163). */
int sv_conv_fwd(const float* x, const float* xr, const int* idims, long long xstride, int IC,
                int D, int KD, const float* w, const float* b, int OC, int s, int p, int dil,
                float* y, const int* odims, long long ystride, int N, int max_ho, int max_wo,
                int mode, float slope, const float* resid, long long rstride, void* stream);

/* Conv2d/Conv3d.backward input gradient (:528-568, :643-692); ic_count > 0
 * computes only the first ic_count input channels.  wt_ws (may be NULL):
 * device workspace of ICn*OC*KD*9 + ICn floats (ICn = ic_count or IC); given
 * it, a stride-1 layer runs as the forward convolution of g with the
 * flipped, channel-exchanged kernel on the forward engines; otherwise (or
 * for stride 2) a gather kernel on the vector ALUs. */
int sv_conv_bwd_data(const float* g, const int* odims, long long gstride, const float* w, int IC,
                     int OC, int D, int KD, int s, int p, int dil, float* gx, const int* idims,
                     long long gxstride, int N, int max_hi, int max_wi, int ic_count, float* wt_ws,
                     void* stream);

/* Conv2d/Conv3d.backward weight and bias gradients (accumulated into gw,
 * gb, like the reference's +=); partials is a workspace of the size the
 * query returns.  Each (sample, position chunk) sums its share into a
 * partial, on the matrix cores when the layer's input tile fits (up to 12
 * output and 512 weight columns), else on the vector ALUs; the partials are
 * then added in a fixed order, so the result is the same on every run. */
int sv_conv_bwd_weight_workspace_bytes(int N, int IC, int OC, int KD, int nchunk, size_t* bytes);
int sv_conv_bwd_weight(const float* x, const float* xr, const int* idims, long long xstride,
                       int IC, int D, int KD, const float* g, const int* odims, long long gstride,
                       int OC, int s, int p, int dil, int N, int nchunk, float* gw, float* gb,
                       float* partials, size_t ws_bytes, void* stream);

/* LeakyReLU.backward (:713-720) in place, from the post-activation values. */
int sv_act_bwd(float* g, const float* y, long long n, float slope, void* stream);

/* CostVolume.backward (:801-825). */
int sv_costvol_bwd(const float* fl, const float* fr, const int* dims, long long fstride, int C,
                   int D, const float* gvol, long long gvstride, float* gfl, float* gfr, int N,
                   int max_h, int max_w, void* stream);

/* SoftArgmin.forward (:861-889) with CostVolume.invalid_mask(penalty)
 * (:827-840); SoftArgmin.backward (:891-901) plus an optional direct cost
 * gradient (the cross-entropy term of train_grads, :1641-1650). */
int sv_softargmin_fwd(const float* cost, const int* dims, long long cstride, int D,
                      float penalty, float* disp, long long dstride, float* probs, int N,
                      int max_hw, void* stream);
/* The last cost-filter layer (Conv3d IC -> 1, :1079-1094, no activation)
 * fused with sv_softargmin_fwd: disp and probs as that call gives them after
 * sv_conv_fwd of the layer, bit for bit, without the filtered cost volume
 * (D = 8 or 12; SV_EUNSUPPORTED otherwise).  No counterpart function in the
 * reference: two of its layers in one launch. */
int sv_conv3_softargmin(const float* x, const int* dims, long long xstride, int IC, int D,
                        const float* w, const float* b, float penalty, float* disp,
                        long long dstride, float* probs, long long pstride, int N, int max_h,
                        int max_w, void* stream);
/* The network input of one pass at the working size (predict_at_scale's
 * resize, :3950-3953, and predict's guides, :5891-5897) in one launch:
 * both views resized (resize_gray, :3843-3845) into img [2N][fslot] (left
 * first), the left again as channel 1 of rin_f [N][2][fslot] and its 2x2
 * average (avgpool2, :979-995) as channel 1 of rin_h [N][2][hslot]; equal
 * to sv_bilinear_fwd + sv_copy_channel + sv_avgpool2 bit for bit.  No
 * counterpart function in the reference: three of its steps in one launch. */
int sv_load_pair(const void* l, const void* r, int dtype, const int* sdims, long long sstride,
                 float* img, long long fslot, const int* fdims, float* rin_h, long long hslot,
                 const int* hdims, float* rin_f, int N, int max_h, int max_w, void* stream);
/* Entry of a refinement level (StereoNet.forward :1220-1232): the coarser
 * disparity upsampled x2 and doubled into channel 0 of rin [N][2][slot],
 * then the first refinement convolution (2 -> OC, 3x3, dilation dil,
 * LeakyReLU slope) over rin into y; equal to sv_bilinear_fwd + sv_conv_fwd
 * bit for bit (OC 8 or 32, dilation 1..4).  No counterpart function in the
 * reference: two of its steps in one launch. */
int sv_refine_in(const float* src, const int* sdims, long long sslot, float* rin, const int* dims,
                 long long slot, const float* w, const float* b, int OC, int dil, float slope,
                 float* y, long long ystride, int N, int max_h, int max_w, void* stream);
int sv_softargmin_bwd(const float* probs, const float* disp, const float* gdisp,
                      const float* gextra, const int* dims, long long cstride, int D,
                      long long dstride, float* gcost, int N, int max_hw, void* stream);

/* BilinearUpsample.forward/backward (:913-973) and resize_gray
 * (:3843-3845): y = mul * bilinear(x) (muln, when not NULL, gives one
 * multiplier per sample) into output channel och; uint8 input reads as
 * value/255. */
int sv_bilinear_fwd(const void* x, int dtype, const int* idims, long long xstride, int C,
                    float* y, const int* odims, long long ystride, int och, float mul,
                    const float* muln, int N, int max_ho, int max_wo, void* stream);
int sv_bilinear_bwd(const float* gy, const int* odims, long long gystride, int ich, int C,
                    float* gx, const int* idims, long long gxstride, float mul, int N, int max_hi,
                    int max_wi, void* stream);

/* avgpool2 (:979-995), scaled by mul (pool_disp uses 0.5, :1572-1576). */
int sv_avgpool2(const void* x, int dtype, const int* idims, long long xstride, float* y,
                const int* odims, long long ystride, int och, float mul, int N, int max_ho,
                int max_wo, void* stream);
/* pool_valid (:1579-1592). */
int sv_pool_valid(const uint8_t* v, const int* idims, long long vstride, uint8_t* out,
                  const int* odims, long long ostride, int N, int max_ho, int max_wo,
                  void* stream);
/* Channel copy into a multi-channel slot (the guide channel of
 * Refinement's input, :1140-1146); uint8 reads as value/255. */
int sv_copy_channel(const void* x, int dtype, const int* dims, long long xstride, float* y,
                    long long ystride, int och, int N, int max_hw, void* stream);

/* pseudo_huber (:1326-1351) and disparity_ce (:1274-1323) per sample, with
 * float64 reductions: gradients scaled by weight*gscale/count_n, and
 * loss_out[n] += weight * loss_n.  part: 2*N*nblk doubles, scale: N floats. */
int sv_pseudo_huber(const float* pred, const float* target, const uint8_t* valid,
                    const int* dims, long long stride, int N, float alpha, double weight,
                    double gscale, float* grad, double* loss_out, double* part, float* scale,
                    int nblk, void* stream);
int sv_disparity_ce(const float* probs, const float* target, const uint8_t* valid,
                    const int* dims, long long pstride, long long tstride, int D, float sigma,
                    int N, double weight, double gscale, float* gcost, double* loss_out,
                    double* part, float* scale, int nblk, void* stream);

/* Adam.step (:1368-1387) with c1 = 1/(1-b1^t), c2 = 1/(1-b2^t). */
int sv_adam_step(float* p, const float* g, float* m, float* v, long long n, float lr, float b1,
                 float b2, float eps, float c1, float c2, float clip, void* stream);

/* sorted(values)[k] per sample: the selection inside _peak (:3866-3869).
 * max_count is the largest sample's element count; ws holds
 * sv_select_kth_workspace_bytes(N) bytes. */
int sv_select_kth_workspace_bytes(int N, size_t* bytes);
int sv_select_kth(const float* x, const int* dims, long long stride, const long long* kth,
                  float* out, int N, long long max_count, void* ws, size_t ws_bytes,
                  void* stream);

/* predict_auto's scale decision (:3897-3926, widths by width_for, working
 * size by predict_at_scale :3943-3947) for N pairs without a host round
 * trip.  Pair n is src_dims[n] = (H, W) and was probed at probe_dims[n]
 * (device int32 pairs).  From the selected peaks (pass 0: the probe's,
 * scaled by W / probe width) writes works[n], active[n], trace[n*passes +
 * pass] (-1 once a pair is done), inv[n] = 1/scale, and the dims tables of
 * the pass: int block of 18N = [full N][half N][f4 N][full2 2N][f1_2 2N]
 * [f4_2 2N] (h, w) pairs, all zero for a finished pair. */
int sv_auto_widths(const float* peaks, int N, const int* src_dims, const int* probe_dims,
                   double target_peak, int max_width, int pass, int passes, int* works, int* active,
                   int* trace, int* tables, float* inv, void* stream);

/* Elementwise helpers of the backward pass and of the network-to-
 * refinement hand-off (no reference counterpart: list arithmetic there). */
int sv_axpby(const float* x, float* y, long long n, float a, float b, void* stream);
int sv_widen(const float* x, const int* dims, long long xstride, double* y, long long ystride,
             const float* mul, int N, int max_hw, void* stream);

/* Image input, replacing read_pnm (stereonet.py:262-324), read_png with its
 * zlib/DEFLATE inflater and scanline unfilter (:3204-3601) and read_image
 * (:3604-3613).  The file bytes are decoded on the host (DEFLATE is a
 * serial bit stream); the samples go to the GPU for the float64 planes.
 *
 * sv_image_size: width and height from the PNG IHDR or the Netpbm header
 * (PNG chosen by its 8-byte signature, as read_image does).
 * sv_image_decode: the whole file into samples (host uint16, capacity
 * values; 3*W*H are written, three channel values per pixel), the scale rule in *mode (0: value / maxval
 * as value * (1/maxval); 1: palette entry / 255; 2: as 0, one-channel
 * Netpbm whose gray is the sample itself) and *maxval.  Malformed files
 * return SV_EINVAL with the reference's message where it has one; so do
 * images of more than 2^28 pixels.
 * sv_image_expand: on the GPU, samples (device) -> gray[npix] and
 * rgb[3*npix] float64 (either may be NULL), bit-identical to the lists
 * read_pnm / read_png return. */
int sv_image_size(const uint8_t* data, size_t n, int* W, int* H);
int sv_image_decode(const uint8_t* data, size_t n, uint16_t* samples, size_t capacity, int* mode,
                    int* maxval);
int sv_image_expand(const uint16_t* samples, long long npix, int mode, int maxval, double* gray,
                    double* rgb, void* stream);

/* Middlebury scoring, replacing eval_real.py's load_gt (:78-87), masks
 * (:90-102), score (:105-121) and constant (:124-127); device pointers.
 *
 * sv_gt_from_gray: level = gray * 255; known[i] = level >= 0.5, disp[i] =
 * level / scale where known, else 0.
 * sv_gt_nonocc: for B maps of W x H, nonocc[i] = known left pixel whose
 * xr = (int)(x - dl + 0.5) (toward zero) is inside the row, known on the
 * right, with |dl - dr(xr)| <= 1.
 * sv_score: per map, out[6b..6b+5] = (n, sum |pred - gt|, count > 1,
 * count > 2, count > 4, sum gt) over mask != 0; pred f32 or f64 (SV_F32 /
 * SV_F64).  sv_abs_dev: out[b] = sum |gt - m| over the mask with m =
 * stats[6b+5] / stats[6b] (stats from sv_score).  Sums are float64 in a
 * fixed order (deterministic; equal to the reference's left-to-right sums
 * to rounding); both take a workspace of sv_score_workspace_bytes. */
int sv_gt_from_gray(const double* gray, long long n, double scale, double* disp, uint8_t* known,
                    void* stream);
int sv_gt_nonocc(const double* dl, const uint8_t* kl, const double* dr, const uint8_t* kr, int B,
                 int W, int H, uint8_t* nonocc, void* stream);
int sv_score_workspace_bytes(int B, long long npix, size_t* bytes);
int sv_score(const void* pred, int pred_dtype, const double* gt, const uint8_t* mask, int B,
             long long npix, double* out, void* ws, size_t ws_bytes, void* stream);
int sv_abs_dev(const double* gt, const uint8_t* mask, const double* stats, int B, long long npix,
               double* out, void* ws, size_t ws_bytes, void* stream);

/* Disparity to 3D, replacing Camera (stereonet.py:1951-1992),
 * robust_min_disparity (:1989-2003), disparity_to_points (:2006-2021),
 * min_filter_arg (:5566-5605), smooth_background (:5608-5634),
 * background_region (:5637-5671) and grid_mesh (:5770-5862).  Device
 * pointers except cam: a host array {fx, cx, cy, baseline, doffs}.  float64
 * throughout, bit-identical to the reference.
 *
 * sv_disparity_to_points: pts[3i..3i+2] = (X, Y, Z), valid[i] = 0 (pts
 * NaN) where d < min_disp or Z is not in (0, max_depth].
 * sv_kth_f64: *out (device) = the k-th smallest (0-based) of x[0],
 * x[stride], ..., x[(n-1)*stride]; robust_min_disparity's median and
 * grid_mesh's pivot.  sv_kth_f64_batch: the same for B sets at x + b *
 * mstride, out[b]; with ns given, set b holds ns[b * nstride] values (at
 * most n) and k is half of that (empty sets give empty_value).  Workspace:
 * sv_kth_f64_workspace_bytes(B).
 * sv_min_filter_arg: B maps; minimum over the clipped window (radius r
 * along rows, rv along columns), idx = flat in-image index of the first
 * strict minimum (rows scanned first); rv > 0 needs tmp (B*W*H doubles)
 * and tmp_idx (B*W*H ints).
 * sv_smooth_background: box means of radius r over rows then columns,
 * clamped to at most disp; tmp holds B*W*H doubles.
 * sv_background_region: mask (uint8), m (smoothed background disparity),
 * src (int32 source pixel) for B maps; tmp (doubles) and tmp_mask (bytes)
 * of B*W*H each.
 * sv_grid_mesh_build: one build() of grid_mesh for B maps (main: region =
 * src = NULL; fill: the background mask and source pixels, val = m), with
 * min_disp[b] per map (device).  Map b's vertices go to [b*vcap, ...) of
 * pos (3 per vertex), uv (2), flag, dv and its triangles to [b*tcap, ...)
 * of idx (3 per triangle), vertex ids counted within the map, in the
 * reference's order.  totals[B][4] (device) = main vertices, main
 * triangles, fill vertices, fill triangles: the main build writes the
 * first two, the fill build appends after them and writes the last two.
 * vcap >= nx*ny per build and tcap >= 2*(nx-1)*(ny-1) per build (nx, ny the
 * grid sizes); ws is sv_grid_mesh_workspace_bytes(B, W, H, step). */
int sv_disparity_to_points(const double* disp, int B, int W, int H, const double* cam,
                           double min_disp, double max_depth, double* pts, uint8_t* valid,
                           void* stream);
int sv_kth_f64_workspace_bytes(int B, size_t* bytes);
int sv_kth_f64(const double* x, long long stride, long long n, long long k, double* out, void* ws,
               size_t ws_bytes, void* stream);
int sv_kth_f64_batch(const double* x, long long mstride, long long stride, int B, long long n,
                     long long k, const int* ns, int nstride, double empty_value, double* out,
                     void* ws, size_t ws_bytes, void* stream);
int sv_min_filter_arg(const double* v, int B, int W, int H, int r, int rv, double* out, int* idx,
                      double* tmp, int* tmp_idx, void* stream);
int sv_smooth_background(const double* m, const double* disp, int B, int W, int H, int r,
                         double* tmp, double* out, void* stream);
int sv_background_region(const double* disp, int B, int W, int H, int radius, double jump_ratio,
                         double min_jump, uint8_t* mask, double* m, int* src, double* tmp,
                         uint8_t* tmp_mask, void* stream);
int sv_grid_mesh_workspace_bytes(int B, int W, int H, int step, size_t* bytes);
int sv_grid_mesh_build(const double* val, const uint8_t* region, const int* src, int B, int W,
                       int H, int step, const double* cam, const double* min_disp, double jump,
                       double jump_rel, double* pos, double* uv, uint8_t* flag, double* dv, int* idx,
                       long long vcap, long long tcap, int* totals, void* ws, size_t ws_bytes,
                       void* stream);

/* Mesh and Gaussian cloud of the downstream 3D output (csrc/mesh.hip host
 * code, csrc/splat.hip kernels); float64 with the reference's own fsqrt /
 * fexp / flog, results equal to the reference's.
 *
 * sv_delaunay: delaunay, stereonet.py:2052-2141 (Bowyer-Watson); xy = n
 *   (x, y) points (host); up to cap index triples to tris, *count = their
 *   number (an error beyond cap).  Host code, no GPU call.
 * sv_cull_triangles: cull_triangles, :2144-2199; vp = the triangles'
 *   vertices' 3D points [nv][3] with valid flags vv (host), kept triples to
 *   out in input order.  Host code.
 * sv_cull_slanted: cull_slanted, :2324-2372 (same arrays).  Host code.
 * sv_select_points: select_points, :2375-2482 on a host disparity map
 *   (valid may be null); (x, y) pairs to xy.  Host code. */
int sv_delaunay(const double* xy, int n, int* tris, int cap, int* count);
int sv_cull_triangles(const double* vp, const uint8_t* vv, int nv, const int* tris, int nt,
                      double max_z_ratio, double max_edge_factor, int* out, int* count);
int sv_cull_slanted(const double* vp, const uint8_t* vv, int nv, const int* tris, int nt,
                    double max_slant_cos, int* out, int* count);
int sv_select_points(const double* disp, int W, int H, const uint8_t* valid, int target,
                     double edge_fraction, int* xy, int cap, int* count);

/* sv_gaussians_from_mesh: gaussians_from_mesh, :2546-2659; device arrays:
 *   vertices xy [nv][2], triangles [nt][3], per-pixel points [W*H][3] with
 *   valid flags, colour [W*H][3]; the kept triangles' Gaussians go to cloud
 *   [14][cap] (mx my mz sxx sxy sxz syy syz szz cr cg cb raw_op raw_sc) in
 *   triangle order, their number to *count (device int); raw_op and raw_sc
 *   are the initial raw opacity and log scale.
 * sv_splat_workspace_bytes / sv_splat_prepare: SplatRenderer._prepare,
 *   :2717-2803: projection with the world-to-camera R (row-major 3x3) and t
 *   (host), camera fx, cx, cy, blur and cutoff; the kept items ranked by
 *   depth (ties in Gaussian order) into the workspace; info (device, 2
 *   long longs, may be null) = kept items, their total footprint area.
 * sv_splat_render: SplatRenderer.render, :2807-2852, after prepare: colour C
 *   [H*W][3] and final transmittance T [H*W] (device).
 * sv_splat_backward: SplatRenderer.backward, :2854-2962: adds the gradients
 *   of sum(gC * C) to grads [5][cap] (colour r, g, b, raw opacity, raw
 *   scale); contrib (5 * total doubles) is scratch for the footprint area
 *   total of the prepared workspace.
 * sv_photometric_mse: refine_gaussians' loss, :3017-3027: gC = 2 (C - target)
 *   / n, mean squared error to *loss (device), part = 256 doubles of scratch.
 * sv_cloud_adam: CloudAdam.step, :2982-3007, m and v [5][cap]; c1, c2 the
 *   bias corrections 1 / (1 - b^t) of the caller's running products. */
int sv_gaussians_from_mesh(const int* xy, int nv, const int* tris, int nt, const double* pts,
                           const uint8_t* pvalid, const double* rgb, int W, int H,
                           double thickness, double raw_op, double raw_sc, double* cloud, int cap,
                           int* count, void* stream);
int sv_splat_workspace_bytes(int n, size_t* bytes);
int sv_splat_prepare(const double* cloud, int n, int cap, const double* R, const double* t,
                     double fx, double cx, double cy, int W, int H, double blur, double cutoff,
                     void* ws, size_t ws_bytes, long long* info, void* stream);
int sv_splat_render(const void* ws, int n, int W, int H, double br, double bg, double bb,
                    double* C, double* T, void* stream);
int sv_splat_backward(const void* ws, int n, int W, int H, const double* gC, const double* T,
                      double* contrib, long long total, double* grads, int cap, void* stream);
int sv_photometric_mse(const double* C, const double* target, long long n, double* gC,
                       double* loss, double* part, void* stream);
int sv_cloud_adam(double* cloud, const double* grads, double* m, double* v, int n, int cap,
                  double c1, double c2, double lr_colour, double lr_op, double lr_scale,
                  void* stream);

#ifdef __cplusplus
}
#endif

#endif /* STEREO_MI_H */
