// StereoNet layers for gfx950 in FP32 (stereonet.py:453-1388,
// Ray-5106/From-Scratch-Stereo-Vision): convolutions (2D and 3D share one
// kernel family; a 2D conv is depth 1), the cost volume (formed on the fly
// inside the first 3D convolution, never stored), soft-argmin, bilinear
// resampling, average pooling, the two losses with float64 reductions, Adam,
// and a k-th-smallest selection for predict_auto's peak.
//
// Ragged batches: every kernel takes a per-sample dims table (int h, w per
// sample) and a per-sample slot stride; sample n of a tensor lives at
// ptr + n*stride, laid out densely as [C][D][h_n][w_n].  Grids are sized for
// the largest sample and blocks outside a sample's extent exit (the tiled
// forward convolutions instead walk a list of the samples' own tiles), so
// pairs run at different working resolutions share one launch per layer.
//
// Determinism: backward passes are gathers or fixed-order two-stage
// reductions; no floating-point atomics.
#include <algorithm>
#include <cmath>

#include "common.h"
#include "conv_tiles.h"

namespace {

// ---------------------------------------------------------------- conv fwd
// Conv2d.forward (stereonet.py:489-526) and Conv3d.forward (:596-641) as one
// direct convolution.  Weights [OC][IC][KD][3][3]; spatial stride s, zero
// padding p, dilation dil; depth stride 1, padding KD/2.  Each thread makes
// OCB output channels of one output voxel; the weight reads are uniform
// across the wave and go through the scalar cache.
// SRC == 1: the input is the cost volume |fl(x) - fr(x - d)|, 2.0 where
// x < d (CostVolume.forward, :771-799), computed from the two feature maps.
// mode 0: bias only; 1: LeakyReLU(slope); 2: max(resid + conv, 0) with the
// residual read from `resid` (Refinement output, :1150-1163).
template <int OCB, int SRC>
__global__ void __launch_bounds__(256) k_conv_fwd(
    const float* __restrict__ x, const float* __restrict__ xr, const int* __restrict__ idims,
    long long xstride, int IC, int D, int KD, const float* __restrict__ w,
    const float* __restrict__ b, int OC, int s, int p, int dil, float* __restrict__ y,
    const int* __restrict__ odims, long long ystride, int tilesY, int mode, float slope,
    const float* __restrict__ resid, long long rstride) {
    const int n = blockIdx.z;
    const int od = blockIdx.y / tilesY;
    const int oy = (blockIdx.y - od * tilesY) * 4 + threadIdx.y;
    const int ox = blockIdx.x * 64 + threadIdx.x;
    const int hi = dh(idims, n), wi = dw(idims, n);
    const int ho = dh(odims, n), wo = dw(odims, n);
    if (ox >= wo || oy >= ho) return;
    const int pd = KD / 2;
    const long long iplane = (long long)hi * wi;
    const long long ivol = iplane * D;
    const long long oplane = (long long)ho * wo;
    const float* xs = x + (long long)n * xstride;
    const float* xrs = SRC == 1 ? xr + (long long)n * xstride : nullptr;
    float* ys = y + (long long)n * ystride + (long long)od * oplane + (long long)oy * wo + ox;
    const int iy0 = oy * s - p, ix0 = ox * s - p;
    for (int oc0 = 0; oc0 < OC; oc0 += OCB) {
        float acc[OCB];
#pragma unroll
        for (int j = 0; j < OCB; ++j) acc[j] = (oc0 + j < OC) ? b[oc0 + j] : 0.f;
        for (int ic = 0; ic < IC; ++ic) {
            for (int kd = 0; kd < KD; ++kd) {
                const int dd = od + kd - pd;
                if (dd < 0 || dd >= D) continue;
#pragma unroll
                for (int ky = 0; ky < 3; ++ky) {
                    const int iy = iy0 + ky * dil;
                    if (iy < 0 || iy >= hi) continue;
#pragma unroll
                    for (int kx = 0; kx < 3; ++kx) {
                        const int ix = ix0 + kx * dil;
                        if (ix < 0 || ix >= wi) continue;
                        float v;
                        if (SRC == 1) {
                            // channel ic of the left/right feature maps, shift dd
                            if (ix < dd) {
                                v = 2.0f;
                            } else {
                                const long long o = (long long)ic * iplane + (long long)iy * wi;
                                v = fabsf(xs[o + ix] - xrs[o + ix - dd]);
                            }
                        } else {
                            v = xs[(long long)ic * ivol + (long long)dd * iplane + (long long)iy * wi + ix];
                        }
                        const int wt = ((ic * KD + kd) * 3 + ky) * 3 + kx;
#pragma unroll
                        for (int j = 0; j < OCB; ++j)
                            if (oc0 + j < OC) acc[j] += v * w[(long long)(oc0 + j) * IC * KD * 9 + wt];
                    }
                }
            }
        }
#pragma unroll
        for (int j = 0; j < OCB; ++j) {
            if (oc0 + j >= OC) break;
            float v = acc[j];
            if (mode == 1) {
                v = v > 0.f ? v : v * slope;
            } else if (mode == 2) {
                const float r = resid[(long long)n * rstride + (long long)(oc0 + j) * D * oplane +
                                      (long long)od * oplane + (long long)oy * wo + ox];
                v = r + v;
                v = v > 0.f ? v : 0.f;
            }
            ys[(long long)(oc0 + j) * D * oplane] = v;
        }
    }
}

// Per-channel form of the LDS-tiled convolution, for layers whose input
// channels do not fit one staging pass of k_conv_tile (the widened
// network).  A block of 256 threads (TW columns x 256/TW rows of threads)
// makes one tile of TW output columns x RP*256/TW rows at one output depth
// (grid over the largest sample; blocks outside a sample exit); for each
// input channel it stages the input halo tile of all KD depth taps in LDS
// (zero outside the image, the cost volume formed while staging when
// SRC == 1), double-buffered so the next channel loads while this one is
// consumed.  Each thread accumulates RP
// output rows x OCB channels.  WS selects where the weights come from: 1,
// staged in LDS per channel (read back as float4 broadcasts); 2, straight
// from memory through the scalar cache into SGPR operands (the weights are
// uniform across the block), which leaves the LDS return path to the
// activations.  TW (64, 32 or 16) only changes how the plane is cut into
// tiles: narrow maps waste less of a narrower tile.  Every output is summed
// in the same order, one fused multiply-add per term, whatever TW, RP and WS
// are, so the variants agree bit for bit.
template <int KD, int SRC, int OCB, int RP, int TW, int WS>
__global__ void __launch_bounds__(256) k_conv_chan(
    const float* __restrict__ x, const float* __restrict__ xr, const int* __restrict__ idims,
    long long xstride, int IC, int D, const float* __restrict__ w, const float* __restrict__ b,
    int OC, int s, int p, int dil, float* __restrict__ y, const int* __restrict__ odims,
    long long ystride, int tilesY, int mode, float slope, const float* __restrict__ resid,
    long long rstride, int IHt, int IWt) {
    // OCB padded to a multiple of 4 for float4 weight reads
    constexpr int OCP = (OCB + 3) / 4 * 4;
    constexpr int WSZ = KD * 9 * OCP;
    constexpr int TY = 256 / TW;   // thread rows
    constexpr int TT_H = TY * RP;  // rows per tile: TY threads high, RP rows each
    extern __shared__ __align__(16) float smem_f[];
    const int n = blockIdx.z;
    const int od = blockIdx.y / tilesY;
    const int ty0 = (blockIdx.y - od * tilesY) * TT_H;
    const int tx0 = blockIdx.x * TW;
    const int hi = dh(idims, n), wi = dw(idims, n);
    const int ho = dh(odims, n), wo = dw(odims, n);
    if (tx0 >= wo || ty0 >= ho) return;  // uniform per block
    const int tx = threadIdx.x, ty = threadIdx.y;  // TW x TY
    const int tid = ty * TW + tx;
    const int plane_t = IHt * IWt;
    const int tsz = KD * plane_t;
    float* wsm = smem_f;                   // [2][KD*9][OCP]
    float* tile = smem_f + 2 * WSZ;        // [2][KD][IHt][IWt]
    const int iy0 = ty0 * s - p, ix0 = tx0 * s - p;
    const long long iplane = (long long)hi * wi;
    const float* xs = x + (long long)n * xstride;
    const float* xrs = SRC == 1 ? xr + (long long)n * xstride : nullptr;
    constexpr int PD = KD / 2;
    float acc[RP][OCP];
    auto stage = [&](int ic, int oc0, int slotb) {
        float* buf = tile + slotb * tsz;
        for (int e = tid; e < tsz; e += 256) {
            const int kd = e / plane_t;
            const int r = e - kd * plane_t;
            const int yy = r / IWt, xx = r - yy * IWt;
            const int dd = od + kd - PD, iy = iy0 + yy, ix = ix0 + xx;
            float v = 0.f;
            if (dd >= 0 && dd < D && iy >= 0 && iy < hi && ix >= 0 && ix < wi) {
                if (SRC == 1) {
                    if (ix < dd) {
                        v = 2.0f;
                    } else {
                        const long long o = (long long)ic * iplane + (long long)iy * wi;
                        v = fabsf(xs[o + ix] - xrs[o + ix - dd]);
                    }
                } else {
                    v = xs[((long long)ic * D + dd) * iplane + (long long)iy * wi + ix];
                }
            }
            buf[e] = v;
        }
        if (WS == 1) {
            float* wb = wsm + slotb * WSZ;
            for (int e = tid; e < WSZ; e += 256) {
                const int t = e / OCP, j = e - t * OCP;
                wb[e] = (j < OCB && oc0 + j < OC) ? w[((long long)(oc0 + j) * IC + ic) * KD * 9 + t] : 0.f;
            }
        }
    };
    for (int oc0 = 0; oc0 < OC; oc0 += OCB) {
#pragma unroll
        for (int i = 0; i < RP; ++i)
#pragma unroll
            for (int j = 0; j < OCP; ++j) acc[i][j] = 0.f;
        stage(0, oc0, 0);
        __syncthreads();
        for (int ic = 0; ic < IC; ++ic) {
            const float* cur = tile + (ic & 1) * tsz;
            const float4* wc = reinterpret_cast<const float4*>(wsm + (ic & 1) * WSZ);
            if (ic + 1 < IC) stage(ic + 1, oc0, (ic + 1) & 1);
#pragma unroll
            for (int kd = 0; kd < KD; ++kd)
#pragma unroll
                for (int ky = 0; ky < 3; ++ky)
#pragma unroll
                    for (int kx = 0; kx < 3; ++kx) {
                        float v[RP];
#pragma unroll
                        for (int i = 0; i < RP; ++i)
                            v[i] = cur[kd * plane_t + ((ty * RP + i) * s + ky * dil) * IWt + tx * s + kx * dil];
                        const int wt = (kd * 3 + ky) * 3 + kx;
                        if (WS == 1) {
#pragma unroll
                            for (int q = 0; q < OCP / 4; ++q) {
                                const float4 wv = wc[wt * (OCP / 4) + q];
#pragma unroll
                                for (int i = 0; i < RP; ++i) {
                                    acc[i][4 * q + 0] = __builtin_fmaf(v[i], wv.x, acc[i][4 * q + 0]);
                                    acc[i][4 * q + 1] = __builtin_fmaf(v[i], wv.y, acc[i][4 * q + 1]);
                                    acc[i][4 * q + 2] = __builtin_fmaf(v[i], wv.z, acc[i][4 * q + 2]);
                                    acc[i][4 * q + 3] = __builtin_fmaf(v[i], wv.w, acc[i][4 * q + 3]);
                                }
                            }
                        } else {
                            // weights of (oc0 + j, ic, tap wt): uniform addresses
                            const float* wg = w + ((long long)oc0 * IC + ic) * KD * 9 + wt;
#pragma unroll
                            for (int j = 0; j < OCB; ++j) {
                                const float wj = oc0 + j < OC ? wg[(long long)j * IC * KD * 9] : 0.f;
#pragma unroll
                                for (int i = 0; i < RP; ++i) acc[i][j] = __builtin_fmaf(v[i], wj, acc[i][j]);
                            }
                        }
                    }
            __syncthreads();
        }
        const int ox = tx0 + tx;
        const long long oplane = (long long)ho * wo;
        float* ys = y + (long long)n * ystride + (long long)od * oplane;
#pragma unroll
        for (int i = 0; i < RP; ++i) {
            const int oy = ty0 + ty * RP + i;
            if (ox >= wo || oy >= ho) continue;
#pragma unroll
            for (int j = 0; j < OCB; ++j) {
                const int oc = oc0 + j;
                if (oc >= OC) break;
                float v = acc[i][j] + b[oc];
                if (mode == 1) {
                    v = v > 0.f ? v : v * slope;
                } else if (mode == 2) {
                    v = resid[(long long)n * rstride + (long long)oc * D * oplane +
                              (long long)od * oplane + (long long)oy * wo + ox] + v;
                    v = v > 0.f ? v : 0.f;
                }
                ys[(long long)oc * D * oplane + (long long)oy * wo + ox] = v;
            }
        }
    }
}

// LDS-tiled form of k_conv_fwd on the vector ALUs.  A block of 256 threads
// (TW columns x 256/TW rows of threads) makes output tiles of TW columns x
// RP*256/TW rows at one output depth, each thread RP rows x OCB channels,
// walking the launch's tile list; per tile it stages the input of ICG
// channels at a time (all of them when they fit, then once for every
// output-channel group).  Tile list and staging: conv_tiles.h.
//
// WS selects where the weights come from: 1, staged in LDS with the input
// (read back as float4 broadcasts); 2, straight from memory through the
// scalar cache into SGPR operands (uniform across the block).  TW (64, 32 or
// 16) only changes how the plane is cut into tiles.  Every output is summed
// in the same order (input channel, then depth, row, column tap; one fused
// multiply-add per term; bias last) whatever TW, RP, WS and ICG are, so the
// variants agree bit for bit.
template <int KD, int SRC, int OCB, int RP, int TW, int WS>
__global__ void __launch_bounds__(256) k_conv_tile(
    const float* __restrict__ x, const float* __restrict__ xr, const int* __restrict__ idims,
    long long xstride, int IC, int D, const float* __restrict__ w, const float* __restrict__ b,
    int OC, int s, int p, int dil, float* __restrict__ y, const int* __restrict__ odims,
    long long ystride, int N, int mode, float slope, const float* __restrict__ resid,
    long long rstride, sv::ConvTiles ct) {
    // OCB padded to a multiple of 4 for float4 weight reads
    constexpr int OCP = (OCB + 3) / 4 * 4;
    constexpr int TAPS = KD * 9;
    constexpr int TY = 256 / TW;   // thread rows
    constexpr int TT_H = TY * RP;  // rows per tile: TY threads high, RP rows each
    extern __shared__ __align__(16) float smem_f[];
    const int IWt = ct.IWt, ICG = ct.ICG;
    const int plane_t = ct.IHt * IWt;
    const int tsz = KD * plane_t;
    float* tile = smem_f;                                     // [ICG][KD][IHt][IWt]
    float* wsm = smem_f + ICG * tsz;                          // [ICG][TAPS][OCP] (WS == 1)
    int* pre = reinterpret_cast<int*>(wsm + (WS == 1 ? ICG * TAPS * OCP : 0));  // [N + 1]
    __shared__ int wsum[4];
    const int tx = threadIdx.x, ty = threadIdx.y;  // TW x TY
    const int tid = ty * TW + tx;
    const int total = conv_tile_prefix<TW, TT_H>(odims, N, D, pre, wsum, tid);
    float acc[RP][OCP];
    for (int t = conv_first_tile(); t < total; t += gridDim.x) {
        const TileAt at = conv_tile_at<TW, TT_H>(t, pre, N, D, odims);
        const int n = at.n, od = at.od, ty0 = at.ty0, tx0 = at.tx0;
        const int hin = dh(idims, n), win = dw(idims, n);
        const int ho = dh(odims, n), wo = dw(odims, n);
        const float* xs = x + (long long)n * xstride;
        const float* xrs = SRC == 1 ? xr + (long long)n * xstride : nullptr;
        for (int oc0 = 0; oc0 < OC; oc0 += OCB) {
#pragma unroll
            for (int i = 0; i < RP; ++i)
#pragma unroll
                for (int j = 0; j < OCP; ++j) acc[i][j] = 0.f;
            for (int ic0 = 0; ic0 < IC; ic0 += ICG) {
                const int icn = min(ICG, IC - ic0);
                if (oc0 == 0 || ICG < IC)
                    conv_stage_inputs<KD, SRC>(tile, xs, xrs, ic0, icn, D, od, ty0 * s - p,
                                               tx0 * s - p, hin, win, ct, tid);
                if (WS == 1) {
                    const int WE = icn * TAPS * OCP;
                    for (int e = tid; e < WE; e += 256) {
                        const int c = e / (TAPS * OCP);
                        const int r = e - c * (TAPS * OCP);
                        const int tp = r / OCP, j = r - tp * OCP;
                        wsm[e] = (j < OCB && oc0 + j < OC)
                                     ? w[((long long)(oc0 + j) * IC + ic0 + c) * TAPS + tp] : 0.f;
                    }
                }
                __syncthreads();
                for (int c = 0; c < icn; ++c) {
                    const float* cur = tile + c * tsz;
                    const float4* wc = reinterpret_cast<const float4*>(wsm + c * TAPS * OCP);
#pragma unroll
                    for (int kd = 0; kd < KD; ++kd)
#pragma unroll
                        for (int ky = 0; ky < 3; ++ky)
#pragma unroll
                            for (int kx = 0; kx < 3; ++kx) {
                                float v[RP];
#pragma unroll
                                for (int i = 0; i < RP; ++i)
                                    v[i] = cur[kd * plane_t + ((ty * RP + i) * s + ky * dil) * IWt +
                                               tx * s + kx * dil];
                                const int wt = (kd * 3 + ky) * 3 + kx;
                                if (WS == 1) {
#pragma unroll
                                    for (int q = 0; q < OCP / 4; ++q) {
                                        const float4 wv = wc[wt * (OCP / 4) + q];
#pragma unroll
                                        for (int i = 0; i < RP; ++i) {
                                            acc[i][4 * q + 0] = __builtin_fmaf(v[i], wv.x, acc[i][4 * q + 0]);
                                            acc[i][4 * q + 1] = __builtin_fmaf(v[i], wv.y, acc[i][4 * q + 1]);
                                            acc[i][4 * q + 2] = __builtin_fmaf(v[i], wv.z, acc[i][4 * q + 2]);
                                            acc[i][4 * q + 3] = __builtin_fmaf(v[i], wv.w, acc[i][4 * q + 3]);
                                        }
                                    }
                                } else {
                                    // weights of (oc0 + j, ic0 + c, tap wt): uniform
                                    // addresses, channels past OC clamped (their
                                    // sums are never stored)
                                    const float* wg = w + (long long)(ic0 + c) * TAPS + wt;
#pragma unroll
                                    for (int j = 0; j < OCB; ++j) {
                                        const float wj = wg[(long long)min(oc0 + j, OC - 1) * IC * TAPS];
#pragma unroll
                                        for (int i = 0; i < RP; ++i)
                                            acc[i][j] = __builtin_fmaf(v[i], wj, acc[i][j]);
                                    }
                                }
                            }
                }
                __syncthreads();
            }
            const int ox = tx0 + tx;
            const long long oplane = (long long)ho * wo;
            float* ys = y + (long long)n * ystride + (long long)od * oplane;
#pragma unroll
            for (int i = 0; i < RP; ++i) {
                const int oy = ty0 + ty * RP + i;
                if (ox >= wo || oy >= ho) continue;
#pragma unroll
                for (int j = 0; j < OCB; ++j) {
                    const int oc = oc0 + j;
                    if (oc >= OC) break;
                    const long long o = (long long)oc * D * oplane + (long long)oy * wo + ox;
                    ys[o] = conv_finish(acc[i][j], b[oc], mode, slope, resid,
                                        (long long)n * rstride + (long long)od * oplane + o);
                }
            }
        }
    }
}

// ----------------------------------------------------------- conv bwd data
// Conv2d/Conv3d.backward input gradient (:528-568, :643-692) as a gather:
// gx[ic][id][iy][ix] = sum over oc and taps whose output lands in range.
template <int ICB>
__global__ void __launch_bounds__(256) k_conv_bwd_data(
    const float* __restrict__ g, const int* __restrict__ odims, long long gstride,
    const float* __restrict__ w, int IC, int OC, int D, int KD, int s, int p, int dil,
    float* __restrict__ gx, const int* __restrict__ idims, long long gxstride, int tilesY, int ICn) {
    const int n = blockIdx.z;
    const int id = blockIdx.y / tilesY;
    const int iy = (blockIdx.y - id * tilesY) * 4 + threadIdx.y;
    const int ix = blockIdx.x * 64 + threadIdx.x;
    const int hi = dh(idims, n), wi = dw(idims, n);
    const int ho = dh(odims, n), wo = dw(odims, n);
    if (ix >= wi || iy >= hi) return;
    const int pd = KD / 2;
    const long long oplane = (long long)ho * wo;
    const long long iplane = (long long)hi * wi;
    const float* gs = g + (long long)n * gstride;
    float* gxs = gx + (long long)n * gxstride + (long long)id * iplane + (long long)iy * wi + ix;
    for (int ic0 = 0; ic0 < ICn; ic0 += ICB) {
        float acc[ICB];
#pragma unroll
        for (int j = 0; j < ICB; ++j) acc[j] = 0.f;
        for (int oc = 0; oc < OC; ++oc) {
            for (int kd = 0; kd < KD; ++kd) {
                const int od = id - kd + pd;
                if (od < 0 || od >= D) continue;
#pragma unroll
                for (int ky = 0; ky < 3; ++ky) {
                    const int ty = iy + p - ky * dil;
                    if (ty < 0) continue;
                    const int oy = ty / s;
                    if (oy * s != ty || oy >= ho) continue;
#pragma unroll
                    for (int kx = 0; kx < 3; ++kx) {
                        const int tx = ix + p - kx * dil;
                        if (tx < 0) continue;
                        const int ox = tx / s;
                        if (ox * s != tx || ox >= wo) continue;
                        const float gv = gs[(long long)oc * D * oplane + (long long)od * oplane +
                                            (long long)oy * wo + ox];
                        const int wt = (kd * 3 + ky) * 3 + kx;
#pragma unroll
                        for (int j = 0; j < ICB; ++j)
                            if (ic0 + j < ICn)
                                acc[j] += gv * w[((long long)oc * IC + ic0 + j) * KD * 9 + wt];
                    }
                }
            }
        }
#pragma unroll
        for (int j = 0; j < ICB; ++j)
            if (ic0 + j < ICn) gxs[(long long)(ic0 + j) * D * iplane] = acc[j];
    }
}

// Weights of the transposed convolution: the input gradient of a stride-1
// convolution with same padding is the forward convolution of the output
// gradient with every kernel flipped in all its taps and input and output
// channels exchanged, wt[ic][oc][t] = w[oc][ic][TAPS - 1 - t] (first ICn input
// channels), with a zero bias zb.
__global__ void k_flip_weights(const float* __restrict__ w, int OC, int IC, int TAPS, int ICn,
                               float* __restrict__ wt, float* __restrict__ zb) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < ICn) zb[i] = 0.f;
    if (i >= ICn * OC * TAPS) return;
    const int ic = i / (OC * TAPS);
    const int r = i - ic * OC * TAPS;
    const int o = r / TAPS, t = r - o * TAPS;
    wt[i] = w[((long long)o * IC + ic) * TAPS + (TAPS - 1 - t)];
}

// --------------------------------------------------------- conv bwd weight
// Weight and bias gradient as a split reduction.  Block (c, kt, n) owns the
// weight columns kt*128 + t (one (ic, tap) pair per thread) for all output
// channels, walks output positions c, c + nchunk*64, ... of sample n in
// tiles of 64 (the gradient tile staged in LDS), and writes its partial sums;
// k_reduce_partials adds the partials in a fixed order.
constexpr int WG_COLS = 128;
constexpr int POS_TILE = 64;

template <int SRC, int OCMAX>
__global__ void __launch_bounds__(WG_COLS) k_conv_bwd_weight(
    const float* __restrict__ x, const float* __restrict__ xr, const int* __restrict__ idims,
    long long xstride, int IC, int D, int KD, const float* __restrict__ g,
    const int* __restrict__ odims, long long gstride, int OC, int s, int p, int dil, int nchunk,
    float* __restrict__ partials) {
    __shared__ float gt[OCMAX][POS_TILE];
    const int n = blockIdx.z;
    const int c = blockIdx.x;
    const int t = threadIdx.x;
    const int ncol = IC * KD * 9;
    const int col = blockIdx.y * WG_COLS + t;
    const int hi = dh(idims, n), wi = dw(idims, n);
    const int ho = dh(odims, n), wo = dw(odims, n);
    const long long oplane = (long long)ho * wo;
    const long long npos = oplane * D;
    const long long iplane = (long long)hi * wi;
    const int pd = KD / 2;
    int ic = 0, kd = 0, ky = 0, kx = 0;
    if (col < ncol) {
        ic = col / (KD * 9);
        const int r = col - ic * KD * 9;
        kd = r / 9;
        ky = (r - kd * 9) / 3;
        kx = r - kd * 9 - ky * 3;
    }
    const float* xs = x + (long long)n * xstride;
    const float* xrs = SRC == 1 ? xr + (long long)n * xstride : nullptr;
    const float* gs = g + (long long)n * gstride;
    float acc[OCMAX];
#pragma unroll
    for (int o = 0; o < OCMAX; ++o) acc[o] = 0.f;
    float bacc = 0.f;
    for (long long base = (long long)c * POS_TILE; base < npos; base += (long long)nchunk * POS_TILE) {
        __syncthreads();
        for (int e = t; e < OC * POS_TILE; e += WG_COLS) {
            const int o = e / POS_TILE, q = e - o * POS_TILE;
            const long long pos = base + q;
            gt[o][q] = pos < npos ? gs[(long long)o * npos + pos] : 0.f;
        }
        __syncthreads();
        if (t < OC && blockIdx.y == 0) {
            for (int q = 0; q < POS_TILE; ++q) bacc += gt[t][q];
        }
        if (col < ncol) {
            const int lim = (int)min((long long)POS_TILE, npos - base);
            // (depth, row, column) of the tile's first position, then stepped
            int od = (int)(base / oplane);
            const int rem0 = (int)(base - (long long)od * oplane);
            int oy = rem0 / wo, ox = rem0 - oy * wo;
            for (int q = 0; q < lim; ++q) {
                if (q > 0 && ++ox == wo) {
                    ox = 0;
                    if (++oy == ho) {
                        oy = 0;
                        ++od;
                    }
                }
                const int dd = od + kd - pd;
                const int iy = oy * s - p + ky * dil;
                const int ix = ox * s - p + kx * dil;
                float v = 0.f;
                if (dd >= 0 && dd < D && iy >= 0 && iy < hi && ix >= 0 && ix < wi) {
                    if (SRC == 1) {
                        if (ix < dd) {
                            v = 2.0f;
                        } else {
                            const long long off = (long long)ic * iplane + (long long)iy * wi;
                            v = fabsf(xs[off + ix] - xrs[off + ix - dd]);
                        }
                    } else {
                        v = xs[(long long)ic * iplane * D + (long long)dd * iplane + (long long)iy * wi + ix];
                    }
                }
#pragma unroll
                for (int o = 0; o < OCMAX; ++o)
                    if (o < OC) acc[o] += gt[o][q] * v;
            }
        }
    }
    const int ncolp = ncol + 1;  // + one bias column
    float* out = partials + ((long long)n * nchunk + c) * (long long)OC * ncolp;
    if (col < ncol) {
#pragma unroll
        for (int o = 0; o < OCMAX; ++o)
            if (o < OC) out[(long long)o * ncolp + col] = acc[o];
    }
    if (t < OC && blockIdx.y == 0) out[(long long)t * ncolp + ncol] = bacc;
}

// gw[o][col] += sum over (n, chunk) of partials, in a fixed order; the
// bias column goes to gb.  A block takes 64 consecutive entries (lane l,
// entry 64*blockIdx.x + l); wave w adds parts w, w + RP_WAVES, ... in
// float64 (each part's 64 entries are one coalesced read, several parts in
// flight), then wave 0 adds the waves' sums in wave order.
constexpr int RP_WAVES = 16;

__global__ void __launch_bounds__(64 * RP_WAVES) k_reduce_partials(
    const float* __restrict__ partials, int nparts, int OC, int ncol, float* __restrict__ gw,
    float* __restrict__ gb) {
    __shared__ double ws[RP_WAVES][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int ncolp = ncol + 1;
    const long long E = (long long)OC * ncolp;
    const int i = blockIdx.x * 64 + lane;
    double s = 0.0;
    if (i < E) {
        const float* p = partials + i;
        int q = wave;
        for (; q + 3 * RP_WAVES < nparts; q += 4 * RP_WAVES) {
            const float a = p[(long long)q * E], b = p[(long long)(q + RP_WAVES) * E];
            const float c = p[(long long)(q + 2 * RP_WAVES) * E], d = p[(long long)(q + 3 * RP_WAVES) * E];
            s += (double)a;
            s += (double)b;
            s += (double)c;
            s += (double)d;
        }
        for (; q < nparts; q += RP_WAVES) s += (double)p[(long long)q * E];
    }
    ws[wave][lane] = s;
    __syncthreads();
    if (wave != 0 || i >= E) return;
    double t = ws[0][lane];
#pragma unroll
    for (int w = 1; w < RP_WAVES; ++w) t += ws[w][lane];
    const int o = i / ncolp, col = i - o * ncolp;
    if (col < ncol)
        gw[(long long)o * ncol + col] += (float)t;
    else
        gb[o] += (float)t;
}

// ------------------------------------------------------------ elementwise
// LeakyReLU.backward (:713-720) in place: the post-activation value has the
// sign of the pre-activation for any slope >= 0.
typedef float v4f32w __attribute__((ext_vector_type(4)));

// Weight and bias gradient on the matrix cores (v_mfma_f32_4x4x1_16b_f32),
// the same split reduction as k_conv_bwd_weight (partials per (sample,
// chunk), added by k_reduce_partials).  Block (c, n): tiles c, c + nchunk,
// ... of sample n's output (16 x 16 positions at one depth), each staged
// with its input halo tile in LDS (conv_tiles.h staging; the cost volume
// |fl - fr| formed there when SRC == 1).  Lanes are weight
// columns k = (ic, kd, ky, kx); per 16 positions of a tile row, one VGPR holds
// the output gradient of 4 output channels at all 16 positions (lane 4j + i:
// g[oc0 + i][pos j]) and the MFMA's broadcast index picks position j:
// acc[i] += g[oc0 + i][pos j] * x(column, pos j).  Work units (column group
// of 64, row phase) are dealt to the 4 waves; a unit's rows are
// r = phase, phase + PQ, ...; the phases of a column group are added in order
// at the end, so the result does not depend on timing.
constexpr int WG_T = 16;  // tile: 16 x 16 output positions
constexpr int WG_UMAX = 2;  // work units per wave

template <int KD, int SRC, int OCG>
__global__ void __launch_bounds__(256) k_wgrad_m4(
    const float* __restrict__ x, const float* __restrict__ xr, const int* __restrict__ idims,
    long long xstride, int IC, int D, const float* __restrict__ g, const int* __restrict__ odims,
    long long gstride, int OC, int s, int p, int dil, int nchunk, float* __restrict__ partials,
    sv::ConvTiles ct) {
    extern __shared__ __align__(16) float smem_f[];
    const int IWt = ct.IWt;
    const int plane = ct.IHt * IWt;
    const int tsz = KD * plane;
    float* tile = smem_f;                        // [IC][KD][IHt][IWt]
    float* gt = smem_f + IC * tsz;               // [4*OCG][16][16]
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int n = blockIdx.z, c = blockIdx.x;
    const int hin = dh(idims, n), win = dw(idims, n);
    const int ho = dh(odims, n), wo = dw(odims, n);
    const int ncol = IC * KD * 9;
    const int NCG = (ncol + 63) / 64;
    const int PQ = NCG >= 4 ? 1 : 4 / NCG;
    const int nunits = NCG * PQ;
    const int txn = (wo + WG_T - 1) / WG_T, tyn = (ho + WG_T - 1) / WG_T;
    const int ntiles = (ho > 0 && wo > 0) ? txn * tyn * D : 0;
    const long long oplane = (long long)ho * wo;
    const float* xs = x + (long long)n * xstride;
    const float* xrs = SRC == 1 ? xr + (long long)n * xstride : nullptr;
    const float* gs = g + (long long)n * gstride;
    // per unit: this lane's column and its offset in the staged tile
    int coff[WG_UMAX], rph[WG_UMAX];
    bool cok[WG_UMAX];
#pragma unroll
    for (int k = 0; k < WG_UMAX; ++k) {
        const int u = wave + 4 * k;
        const int cg = u % NCG;
        rph[k] = u / NCG;
        const int col = cg * 64 + lane;
        cok[k] = u < nunits && col < ncol;
        const int cc = cok[k] ? col : 0;
        const int ic = cc / (KD * 9);
        const int r = cc - ic * KD * 9;
        const int kd = r / 9, ky = (r - kd * 9) / 3, kx = r - kd * 9 - ky * 3;
        coff[k] = ic * tsz + kd * plane + ky * dil * IWt + kx * dil;
    }
    v4f32w acc[WG_UMAX][OCG];
#pragma unroll
    for (int k = 0; k < WG_UMAX; ++k)
#pragma unroll
        for (int q = 0; q < OCG; ++q) acc[k][q] = (v4f32w){0.f, 0.f, 0.f, 0.f};
    float bsum = 0.f;  // bias: thread tid < 16*OC sums row tid%16 of channel tid/16
    for (int t = c; t < ntiles; t += nchunk) {
        const int od = t % D;
        const int tt = t / D;
        const int ty0 = (tt / txn) * WG_T, tx0 = (tt % txn) * WG_T;
        __syncthreads();
        conv_stage_inputs<KD, SRC>(tile, xs, xrs, 0, IC, D, od, ty0 * s - p, tx0 * s - p, hin, win,
                                   ct, tid);
        for (int e = tid; e < 4 * OCG * 256; e += 256) {
            const int oc = e >> 8, r = (e >> 4) & 15, j = e & 15;
            const int oy = ty0 + r, ox = tx0 + j;
            gt[e] = (oc < OC && oy < ho && ox < wo)
                        ? gs[(long long)oc * D * oplane + (long long)od * oplane + (long long)oy * wo + ox]
                        : 0.f;
        }
        __syncthreads();
        if (tid < 16 * OC) {
            const float* gr = gt + (tid >> 4) * 256 + (tid & 15) * 16;
#pragma unroll
            for (int j = 0; j < 16; ++j) bsum += gr[j];
        }
#pragma unroll
        for (int k = 0; k < WG_UMAX; ++k) {
            if (wave + 4 * k >= nunits) break;
            for (int r = rph[k]; r < WG_T; r += PQ) {
                float av[OCG];
#pragma unroll
                for (int q = 0; q < OCG; ++q) av[q] = gt[(4 * q + (lane & 3)) * 256 + r * 16 + (lane >> 2)];
                const float* xb = tile + coff[k] + r * s * IWt;
                float bv[16];
#pragma unroll
                for (int j = 0; j < 16; ++j) bv[j] = xb[j * s];
                auto step = [&](auto jc) {
                    constexpr int j = decltype(jc)::value;
#pragma unroll
                    for (int q = 0; q < OCG; ++q)
                        acc[k][q] = __builtin_amdgcn_mfma_f32_4x4x1f32(av[q], bv[j], acc[k][q], 4, j, 0);
                };
                TapSteps<0, 16>::run(step);
            }
        }
    }
    // combine: the row phases of each column group in order, then write
    __syncthreads();
    float* red = smem_f;  // [unit][4*OCG][64]
#pragma unroll
    for (int k = 0; k < WG_UMAX; ++k) {
        const int u = wave + 4 * k;
        if (u >= nunits) break;
#pragma unroll
        for (int q = 0; q < OCG; ++q)
#pragma unroll
            for (int i = 0; i < 4; ++i) red[(u * 4 * OCG + 4 * q + i) * 64 + lane] = acc[k][q][i];
    }
    float* bred = smem_f + nunits * 4 * OCG * 64;  // [OC][16]
    if (tid < 16 * OC) bred[tid] = bsum;
    __syncthreads();
    const int ncolp = ncol + 1;
    float* out = partials + ((long long)n * nchunk + c) * (long long)OC * ncolp;
    for (int e = tid; e < NCG * 64 * OC; e += 256) {
        const int oc = e / (NCG * 64);
        const int col = e - oc * (NCG * 64);
        if (col >= ncol) continue;
        const int cg = col >> 6, l = col & 63;
        float v = 0.f;
        for (int ph = 0; ph < PQ; ++ph) v += red[((ph * NCG + cg) * 4 * OCG + oc) * 64 + l];
        out[(long long)oc * ncolp + col] = v;
    }
    if (tid < OC) {
        float v = 0.f;
        for (int r = 0; r < 16; ++r) v += bred[tid * 16 + r];
        out[(long long)tid * ncolp + ncol] = v;
    }
}

__global__ void k_act_bwd(float* __restrict__ g, const float* __restrict__ y, long long n,
                          float slope) {
    const long long i = blockIdx.x * 256LL + threadIdx.x;
    if (i >= n) return;
    g[i] = y[i] > 0.f ? g[i] : g[i] * slope;
}

// CostVolume.backward (:801-825) as a gather over disparities:
// gfl[c][y][x] += sum_d gvol*sign, gfr[c][y][x] -= sum_d gvol(x+d)*sign.
__global__ void k_costvol_bwd(const float* __restrict__ fl, const float* __restrict__ fr,
                              const int* __restrict__ dims, long long fstride, int C, int D,
                              const float* __restrict__ gvol, long long gvstride,
                              float* __restrict__ gfl, float* __restrict__ gfr, int tilesY) {
    const int n = blockIdx.z;
    const int c = blockIdx.y / tilesY;
    const int y = (blockIdx.y - c * tilesY) * 4 + threadIdx.y;
    const int x = blockIdx.x * 64 + threadIdx.x;
    const int h = dh(dims, n), w = dw(dims, n);
    if (x >= w || y >= h) return;
    const long long plane = (long long)h * w;
    const long long fo = (long long)n * fstride + (long long)c * plane + (long long)y * w;
    const float* gv = gvol + (long long)n * gvstride + (long long)c * D * plane + (long long)y * w;
    const float l = fl[fo + x];
    const float r = fr[fo + x];
    float al = 0.f, ar = 0.f;
    for (int d = 0; d < D; ++d) {
        if (x >= d) {
            const float diff = l - fr[fo + x - d];
            const float sg = diff < 0.f ? -1.f : 1.f;
            al += gv[(long long)d * plane + x] * sg;
        }
        if (x + d < w) {
            const float diff = fl[fo + x + d] - r;
            const float sg = diff < 0.f ? -1.f : 1.f;
            ar -= gv[(long long)d * plane + x + d] * sg;
        }
    }
    gfl[fo + x] = al;
    gfr[fo + x] = ar;
}

// SoftArgmin.forward (:861-889) with CostVolume.invalid_mask (:827-840):
// p = softmax_d(-(cost + penalty*[d >= 1 and x < d])), disp = sum d*p.
__global__ void k_softargmin_fwd(const float* __restrict__ cost, const int* __restrict__ dims,
                                 long long cstride, int D, float penalty,
                                 float* __restrict__ disp, long long dstride,
                                 float* __restrict__ probs) {
    const int n = blockIdx.y;
    const int h = dh(dims, n), w = dw(dims, n);
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= h * w) return;
    const int x = i % w;
    const long long plane = (long long)h * w;
    const float* cs = cost + (long long)n * cstride;
    float* ps = probs + (long long)n * cstride;
    float best = -INFINITY;
    for (int d = 0; d < D; ++d) {
        const float c = cs[(long long)d * plane + i] + ((d >= 1 && x < d) ? penalty : 0.f);
        best = fmaxf(best, -c);
    }
    float total = 0.f;
    for (int d = 0; d < D; ++d) {
        const float c = cs[(long long)d * plane + i] + ((d >= 1 && x < d) ? penalty : 0.f);
        const float e = __expf(-c - best);
        ps[(long long)d * plane + i] = e;
        total += e;
    }
    const float inv = 1.f / total;
    float acc = 0.f;
    for (int d = 0; d < D; ++d) {
        const float pv = ps[(long long)d * plane + i] * inv;
        ps[(long long)d * plane + i] = pv;
        acc = __builtin_fmaf((float)d, pv, acc);
    }
    disp[(long long)n * dstride + i] = acc;
}

// The last cost-filter layer (Conv3d IC -> 1, 3x3x3, no activation,
// stereonet.py:1079-1094) fused with SoftArgmin.forward and the invalid
// mask (:861-889, :827-840).  A block makes an 8 x 8 pixel tile of one
// sample: wave g computes output depths [g*DG, (g+1)*DG) of its lane's
// pixel, the four waves' costs meet in LDS, and wave 0 runs the soft-argmin
// of each pixel, so the filtered costs never reach memory.  The input tile
// of one channel (every depth, one-pixel halo, zero outside) is staged in
// LDS at a time; each staged value is read once and fed to the (up to
// three) outputs it touches.  The sums follow the convolution engines'
// order -- input channel, depth tap, row, column, one fmaf per term
// (zero-padding terms add nothing and are skipped), bias last -- and the
// soft-argmin is k_softargmin_fwd's, so disparities and probabilities equal
// the two unfused launches bit for bit.
template <int D>
__global__ void __launch_bounds__(256) k_conv3_softargmin(
    const float* __restrict__ x, const int* __restrict__ dims, long long xstride, int IC,
    const float* __restrict__ w, const float* __restrict__ b, float penalty,
    float* __restrict__ disp, long long dstride, float* __restrict__ probs, long long pstride) {
    constexpr int T = 8, TH = T + 2, PL = TH * TH;
    constexpr int DG = (D + 3) / 4;  // output depths per wave
    __shared__ float tile[D * PL];
    __shared__ float cost[64 * (D + 1)];
    const int n = blockIdx.z;
    const int h = dh(dims, n), wd = dw(dims, n);
    const int tx0 = blockIdx.x * T, ty0 = blockIdx.y * T;
    if (tx0 >= wd || ty0 >= h) return;  // the whole block lies outside this sample
    const int lane = threadIdx.x & 63;
    const int g = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int tx = lane & (T - 1), ty = lane / T;
    const long long plane = (long long)h * wd;
    const float* xs = x + (long long)n * xstride;
    const int d0 = g * DG;
    float acc[DG];
#pragma unroll
    for (int j = 0; j < DG; ++j) acc[j] = 0.f;
    // this thread's staging positions (fixed over the channels) and the next
    // channel's values, loaded while the current one is being used
    constexpr int NE = (D * PL + 255) / 256;
    long long off[NE];
    bool in[NE];
#pragma unroll
    for (int k = 0; k < NE; ++k) {
        const int e = threadIdx.x + 256 * k;
        const int dd = e / PL;
        const int r = e - dd * PL;
        const int yy = r / TH, xx = r - yy * TH;
        const int iy = ty0 - 1 + yy, ix = tx0 - 1 + xx;
        in[k] = e < D * PL && iy >= 0 && iy < h && ix >= 0 && ix < wd;
        off[k] = (long long)dd * plane + (long long)iy * wd + ix;
    }
    float nxt[NE];
#pragma unroll
    for (int k = 0; k < NE; ++k) nxt[k] = in[k] ? xs[off[k]] : 0.f;
    for (int c = 0; c < IC; ++c) {
#pragma unroll
        for (int k = 0; k < NE; ++k)
            if (threadIdx.x + 256 * k < D * PL) tile[threadIdx.x + 256 * k] = nxt[k];
        __syncthreads();
        if (c + 1 < IC) {
            const float* xn = xs + (long long)(c + 1) * D * plane;
#pragma unroll
            for (int k = 0; k < NE; ++k) nxt[k] = in[k] ? xn[off[k]] : 0.f;
        }
        const float* wc = w + c * 27;
        const float* tp = tile + ty * TH + tx;
        // input depths d0-1 .. d0+DG feed outputs d0 .. d0+DG-1 (od = dd - kd + 1)
#pragma unroll
        for (int q = -1; q <= DG; ++q) {
            const int dd = d0 + q;
            if (dd < 0 || dd >= D) continue;
#pragma unroll
            for (int ky = 0; ky < 3; ++ky)
#pragma unroll
                for (int kx = 0; kx < 3; ++kx) {
                    const float v = tp[dd * PL + ky * TH + kx];
#pragma unroll
                    for (int kd = 2; kd >= 0; --kd) {
                        const int j = q - kd + 1;  // local output index
                        if (j < 0 || j >= DG || d0 + j >= D) continue;
                        acc[j] = __builtin_fmaf(v, wc[(kd * 3 + ky) * 3 + kx], acc[j]);
                    }
                }
        }
        __syncthreads();
    }
    const float bias = b[0];
#pragma unroll
    for (int j = 0; j < DG; ++j)
        if (d0 + j < D) cost[lane * (D + 1) + d0 + j] = acc[j] + bias;
    __syncthreads();
    const int ox = tx0 + tx, oy = ty0 + ty;
    if (g != 0 || ox >= wd || oy >= h) return;
    float cst[D];
    float best = -INFINITY;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        cst[d] = cost[lane * (D + 1) + d] + ((d >= 1 && ox < d) ? penalty : 0.f);
        best = fmaxf(best, -cst[d]);
    }
    float e[D];
    float total = 0.f;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        e[d] = __expf(-cst[d] - best);
        total += e[d];
    }
    const float inv = 1.f / total;
    const long long i = (long long)oy * wd + ox;
    float* ps = probs + (long long)n * pstride + i;
    float dsum = 0.f;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const float pv = e[d] * inv;
        ps[(long long)d * plane] = pv;
        dsum = __builtin_fmaf((float)d, pv, dsum);
    }
    disp[(long long)n * dstride + i] = dsum;
}

// SoftArgmin.backward (:891-901) plus an extra cost gradient (the
// cross-entropy's q - p, train_grads :1641-1650).
__global__ void k_softargmin_bwd(const float* __restrict__ probs, const float* __restrict__ disp,
                                 const float* __restrict__ gdisp, const float* __restrict__ gextra,
                                 const int* __restrict__ dims, long long cstride, int D,
                                 long long dstride, float* __restrict__ gcost) {
    const int n = blockIdx.y;
    const int h = dh(dims, n), w = dw(dims, n);
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= h * w) return;
    const long long plane = (long long)h * w;
    const float g = gdisp[(long long)n * dstride + i];
    const float di = disp[(long long)n * dstride + i];
    const long long o = (long long)n * cstride + i;
    for (int d = 0; d < D; ++d) {
        float v = g * probs[o + (long long)d * plane] * (di - (float)d);
        if (gextra) v += gextra[o + (long long)d * plane];
        gcost[o + (long long)d * plane] = v;
    }
}

// BilinearUpsample taps (:913-950), in float64 so the integer taps match
// the reference's exactly.
__device__ __forceinline__ void tap(int o, int in, int out, int& i0, int& i1, float& wt) {
    const double sc = (double)in / (double)out;
    double f = (o + 0.5) * sc - 0.5;
    if (f < 0.0) f = 0.0;
    int a = (int)f;
    if (a > in - 1) a = in - 1;
    i0 = a;
    i1 = a + 1 < in ? a + 1 : a;
    wt = (float)(f - a);
}

// tap() with the ratio in / out already formed (the same float64 steps)
__device__ __forceinline__ void tap_sc(int o, int in, double sc, int& i0, int& i1, float& wt) {
    double f = (o + 0.5) * sc - 0.5;
    if (f < 0.0) f = 0.0;
    int a = (int)f;
    if (a > in - 1) a = in - 1;
    i0 = a;
    i1 = a + 1 < in ? a + 1 : a;
    wt = (float)(f - a);
}

// One bilinear sample of plane xs (row length wi) from its taps: the
// four-term sum of BilinearUpsample.forward (:913-960).  Every kernel that
// forms a bilinear value calls this, so fused and unfused launches compute
// it identically.  uint8 planes read as value / 255.
template <typename T>
__device__ __forceinline__ float bl_sample(const T* __restrict__ xs, int wi, int y0, int y1,
                                           float wy, int x0, int x1, float wx) {
    auto rd = [&](int yy, int xx) -> float {
        const T v = xs[(long long)yy * wi + xx];
        if constexpr (sizeof(T) == 1)
            return (float)v * (1.0f / 255.0f);
        else
            return (float)v;
    };
    const float w00 = (1.f - wy) * (1.f - wx), w01 = (1.f - wy) * wx;
    const float w10 = wy * (1.f - wx), w11 = wy * wx;
    return rd(y0, x0) * w00 + rd(y0, x1) * w01 + rd(y1, x0) * w10 + rd(y1, x1) * w11;
}

// BilinearUpsample.forward (:913-960) / resize_gray (:3843-3845): y =
// mul * bilinear(x) into output channels [och, och + C).  T is the input
// element type (uint8 images are read as value / 255).  A thread makes
// BL_X outputs of one row, 64 columns apart, so the row's tap and the two
// float64 ratios are formed once for them.
constexpr int BL_X = 4;

template <typename T>
__global__ void k_bilinear_fwd(const T* __restrict__ x, const int* __restrict__ idims,
                               long long xstride, int C, float* __restrict__ y,
                               const int* __restrict__ odims, long long ystride, int och,
                               float mul, const float* __restrict__ muln, int tilesY) {
    const int n = blockIdx.z;
    const int c = blockIdx.y / tilesY;
    const int oy = (blockIdx.y - c * tilesY) * 4 + threadIdx.y;
    const int ox0 = blockIdx.x * 64 * BL_X + threadIdx.x;
    const int hi = dh(idims, n), wi = dw(idims, n);
    const int ho = dh(odims, n), wo = dw(odims, n);
    if (ox0 >= wo || oy >= ho || hi <= 0 || wi <= 0) return;  // empty source: pair skipped
    if (muln) mul = muln[n];
    int y0, y1;
    float wy;
    tap_sc(oy, hi, (double)hi / (double)ho, y0, y1, wy);
    const double scx = (double)wi / (double)wo;
    const T* xs = x + (long long)n * xstride + (long long)c * hi * wi;
    float* yrow = y + (long long)n * ystride + (long long)(och + c) * ho * wo + (long long)oy * wo;
#pragma unroll
    for (int k = 0; k < BL_X; ++k) {
        const int ox = ox0 + 64 * k;
        if (ox >= wo) break;
        int x0, x1;
        float wx;
        tap_sc(ox, wi, scx, x0, x1, wx);
        yrow[ox] = bl_sample(xs, wi, y0, y1, wy, x0, x1, wx) * mul;
    }
}

// BilinearUpsample.backward (:962-973) as a gather: each input pixel sums
// the output pixels whose taps reference it (found by scanning the bounded
// window of rows/columns that map near it).
__global__ void k_bilinear_bwd(const float* __restrict__ gy, const int* __restrict__ odims,
                               long long gystride, int ich, int C, float* __restrict__ gx,
                               const int* __restrict__ idims, long long gxstride, float mul,
                               int tilesY) {
    const int n = blockIdx.z;
    const int c = blockIdx.y / tilesY;
    const int iy = (blockIdx.y - c * tilesY) * 4 + threadIdx.y;
    const int ix = blockIdx.x * 64 + threadIdx.x;
    const int hi = dh(idims, n), wi = dw(idims, n);
    const int ho = dh(odims, n), wo = dw(odims, n);
    if (ix >= wi || iy >= hi) return;
    const float* gs = gy + (long long)n * gystride + (long long)(ich + c) * ho * wo;
    // output rows whose source coordinate lies within one pixel of iy
    const double sy = (double)ho / hi, sx = (double)wo / wi;
    int oy_lo = (int)floor((iy - 1.0 + 0.5) * sy - 0.5) - 1;
    int oy_hi = (int)ceil((iy + 1.0 + 0.5) * sy - 0.5) + 1;
    int ox_lo = (int)floor((ix - 1.0 + 0.5) * sx - 0.5) - 1;
    int ox_hi = (int)ceil((ix + 1.0 + 0.5) * sx - 0.5) + 1;
    oy_lo = max(oy_lo, 0);
    ox_lo = max(ox_lo, 0);
    oy_hi = min(oy_hi, ho - 1);
    ox_hi = min(ox_hi, wo - 1);
    // the clamp at 0 maps every output with f < 0 onto row/col 0
    if (iy == 0) oy_lo = 0;
    if (ix == 0) ox_lo = 0;
    if (iy == hi - 1) oy_hi = ho - 1;
    if (ix == wi - 1) ox_hi = wo - 1;
    float acc = 0.f;
    for (int oy = oy_lo; oy <= oy_hi; ++oy) {
        int y0, y1;
        float wy;
        tap(oy, hi, ho, y0, y1, wy);
        float fy = 0.f;
        if (y0 == iy) fy += 1.f - wy;
        if (y1 == iy) fy += wy;
        if (fy == 0.f && y0 != iy && y1 != iy) continue;
        for (int ox = ox_lo; ox <= ox_hi; ++ox) {
            int x0, x1;
            float wx;
            tap(ox, wi, wo, x0, x1, wx);
            float fx = 0.f;
            if (x0 == ix) fx += 1.f - wx;
            if (x1 == ix) fx += wx;
            if (x0 != ix && x1 != ix) continue;
            acc += gs[(long long)oy * wo + ox] * (fy * fx);
        }
    }
    gx[(long long)n * gxstride + (long long)c * hi * wi + (long long)iy * wi + ix] = acc * mul;
}

// avgpool2 (:979-995) into output channel och; T as in k_bilinear_fwd.
template <typename T>
__global__ void k_avgpool2(const T* __restrict__ x, const int* __restrict__ idims,
                           long long xstride, float* __restrict__ y, const int* __restrict__ odims,
                           long long ystride, int och, float mul) {
    const int n = blockIdx.z;
    const int oy = blockIdx.y;
    const int ox = blockIdx.x * 64 + threadIdx.x;
    const int hi = dh(idims, n), wi = dw(idims, n);
    const int ho = dh(odims, n), wo = dw(odims, n);
    if (ox >= wo || oy >= ho) return;
    const T* r0 = x + (long long)n * xstride + (long long)(2 * oy) * wi + 2 * ox;
    const T* r1 = r0 + wi;
    float s = (float)r0[0] + (float)r0[1] + (float)r1[0] + (float)r1[1];
    if constexpr (sizeof(T) == 1) s *= (1.0f / 255.0f);
    y[(long long)n * ystride + (long long)och * ho * wo + (long long)oy * wo + ox] = 0.25f * s * mul;
}

// copy (with uint8 -> value/255) into output channel och
template <typename T>
__global__ void k_copy_chan(const T* __restrict__ x, const int* __restrict__ dims,
                            long long xstride, float* __restrict__ y, long long ystride, int och) {
    const int n = blockIdx.y;
    const int h = dh(dims, n), w = dw(dims, n);
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= h * w) return;
    float v = (float)x[(long long)n * xstride + i];
    if constexpr (sizeof(T) == 1) v *= (1.0f / 255.0f);
    y[(long long)n * ystride + (long long)och * h * w + i] = v;
}

// The network's input for one pass at the working size (the resize of
// predict_at_scale, :3950-3953, and the guides of predict, :5891-5897): both
// views resized (resize_gray, :3843-3845) into img ([2N][fslot], left views
// first), the left view again as channel 1 of the full-resolution
// refinement input rin_f ([N][2][fslot]) and its 2x2 average (avgpool2,
// :979-995) as channel 1 of the half-resolution one rin_h ([N][2][hslot]).
// A thread makes one working-size pixel; those at even (x, y) inside the
// half-size map also form the three neighbouring samples for the average.
// Every value equals sv_bilinear_fwd / sv_copy_channel / sv_avgpool2 of the
// same data (bl_sample; the sums in avgpool2's order).
template <typename T>
__global__ void k_load_pair(const T* __restrict__ l, const T* __restrict__ r,
                            const int* __restrict__ sdims, long long sstride,
                            float* __restrict__ img, long long fslot, const int* __restrict__ fdims,
                            float* __restrict__ rin_h, long long hslot,
                            const int* __restrict__ hdims, float* __restrict__ rin_f, int N) {
    const int n = blockIdx.z;
    const int oy = blockIdx.y * 4 + threadIdx.y;
    const int ox = blockIdx.x * 64 + threadIdx.x;
    const int hi = dh(sdims, n), wi = dw(sdims, n);
    const int ho = dh(fdims, n), wo = dw(fdims, n);
    if (ox >= wo || oy >= ho || hi <= 0 || wi <= 0) return;  // past the map, or a pair skipped
    const double scy = (double)hi / (double)ho, scx = (double)wi / (double)wo;
    const T* ls = l + (long long)n * sstride;
    const T* rs = r + (long long)n * sstride;
    int y0, y1, x0, x1;
    float wy, wx;
    tap_sc(oy, hi, scy, y0, y1, wy);
    tap_sc(ox, wi, scx, x0, x1, wx);
    const float vl = bl_sample(ls, wi, y0, y1, wy, x0, x1, wx);
    const float vr = bl_sample(rs, wi, y0, y1, wy, x0, x1, wx);
    const long long i = (long long)oy * wo + ox;
    img[(long long)n * fslot + i] = vl;
    img[(long long)(N + n) * fslot + i] = vr;
    rin_f[(long long)n * 2 * fslot + (long long)ho * wo + i] = vl;
    const int hh = dh(hdims, n), wh = dw(hdims, n);
    if ((oy & 1) || (ox & 1) || (oy >> 1) >= hh || (ox >> 1) >= wh) return;
    int ya, yb, xa, xb;
    float wya, wxa;
    tap_sc(oy + 1, hi, scy, ya, yb, wya);
    tap_sc(ox + 1, wi, scx, xa, xb, wxa);
    const float v01 = bl_sample(ls, wi, y0, y1, wy, xa, xb, wxa);
    const float v10 = bl_sample(ls, wi, ya, yb, wya, x0, x1, wx);
    const float v11 = bl_sample(ls, wi, ya, yb, wya, xa, xb, wxa);
    const float sum = vl + v01 + v10 + v11;
    rin_h[(long long)n * 2 * hslot + (long long)hh * wh + (long long)(oy >> 1) * wh + (ox >> 1)] =
        0.25f * sum;
}

// Entry of one refinement level (StereoNet.forward :1220-1232): the
// coarser disparity upsampled x2 with its values doubled
// (BilinearUpsample, :913-960) becomes channel 0 of the level's input rin
// ([N][2][slot]; kept there for the residual and the backward pass), and the
// first refinement convolution (2 -> OC, 3x3, dilation dil, LeakyReLU,
// :1125-1129) runs over [upsampled, guide], the guide being rin's channel 1.
// A block makes a 16 x 16 output tile: both channels with a dil-pixel halo
// (zero outside) go to LDS, the upsampled one formed there with
// bl_sample.  Sums in the convolution engines' order (input channel, row,
// column; one fmaf per term; bias last), so every value equals
// sv_bilinear_fwd + sv_conv_fwd bit for bit.
template <int OC>
__global__ void __launch_bounds__(256) k_refine_in(
    const float* __restrict__ src, const int* __restrict__ sdims, long long sslot,
    float* __restrict__ rin, const int* __restrict__ dims, long long slot,
    const float* __restrict__ w, const float* __restrict__ b, int dil, float slope,
    float* __restrict__ y, long long ystride) {
    constexpr int T = 16, MAXH = T + 8;
    __shared__ float tile[2][MAXH * MAXH];
    const int n = blockIdx.z;
    const int h = dh(dims, n), wd = dw(dims, n);
    const int tx0 = blockIdx.x * T, ty0 = blockIdx.y * T;
    if (tx0 >= wd || ty0 >= h) return;  // the whole block lies outside this sample
    const int hi = dh(sdims, n), wi = dw(sdims, n);
    const int TH = T + 2 * dil;
    const double scy = (double)hi / (double)h, scx = (double)wi / (double)wd;
    const float* ss = src + (long long)n * sslot;
    float* rn = rin + (long long)n * 2 * slot;
    const long long plane = (long long)h * wd;
    for (int e = threadIdx.x; e < TH * TH; e += 256) {
        const int yy = e / TH, xx = e - yy * TH;
        const int iy = ty0 - dil + yy, ix = tx0 - dil + xx;
        float up = 0.f, gd = 0.f;
        if (iy >= 0 && iy < h && ix >= 0 && ix < wd) {
            int y0, y1, x0, x1;
            float wy, wx;
            tap_sc(iy, hi, scy, y0, y1, wy);
            tap_sc(ix, wi, scx, x0, x1, wx);
            up = bl_sample(ss, wi, y0, y1, wy, x0, x1, wx) * 2.0f;
            gd = rn[plane + (long long)iy * wd + ix];
            if (yy >= dil && yy < dil + T && xx >= dil && xx < dil + T)
                rn[(long long)iy * wd + ix] = up;
        }
        tile[0][e] = up;
        tile[1][e] = gd;
    }
    __syncthreads();
    const int tx = threadIdx.x & (T - 1), ty = threadIdx.x / T;
    const int ox = tx0 + tx, oy = ty0 + ty;
    if (ox >= wd || oy >= h) return;
    float acc[OC];
#pragma unroll
    for (int o = 0; o < OC; ++o) acc[o] = 0.f;
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int ky = 0; ky < 3; ++ky)
#pragma unroll
            for (int kx = 0; kx < 3; ++kx) {
                const float v = tile[c][(ty + ky * dil) * TH + tx + kx * dil];
#pragma unroll
                for (int o = 0; o < OC; ++o)
                    acc[o] = __builtin_fmaf(v, w[(o * 2 + c) * 9 + ky * 3 + kx], acc[o]);
            }
    float* ys = y + (long long)n * ystride + (long long)oy * wd + ox;
#pragma unroll
    for (int o = 0; o < OC; ++o) ys[(long long)o * plane] = conv_finish(acc[o], b[o], 1, slope, nullptr, 0);
}

// ------------------------------------------------------------------ losses
// pseudo_huber (:1326-1351) per sample: writes x/r into grad and per-block
// float64 partial sums of (r - alpha) and the valid count.
__global__ void k_ph_partial(const float* __restrict__ pred, const float* __restrict__ target,
                             const uint8_t* __restrict__ valid, const int* __restrict__ dims,
                             long long stride, float alpha, float* __restrict__ grad,
                             double* __restrict__ part, int nblk) {
    __shared__ double sl[256];
    __shared__ double sc[256];
    const int n = blockIdx.y;
    const int hw = dh(dims, n) * dw(dims, n);
    double l = 0.0, cnt = 0.0;
    const long long o = (long long)n * stride;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < hw; i += nblk * 256) {
        float g = 0.f;
        if (valid[o + i]) {
            const float xv = pred[o + i] - target[o + i];
            const float r = sqrtf(xv * xv + alpha * alpha);
            l += (double)(r - alpha);
            g = xv / r;
            cnt += 1.0;
        }
        grad[o + i] = g;
    }
    sl[threadIdx.x] = l;
    sc[threadIdx.x] = cnt;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) {
            sl[threadIdx.x] += sl[threadIdx.x + st];
            sc[threadIdx.x] += sc[threadIdx.x + st];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[((long long)n * nblk + blockIdx.x) * 2] = sl[0];
        part[((long long)n * nblk + blockIdx.x) * 2 + 1] = sc[0];
    }
}

// per sample: loss = weight * sum / count, scale = weight * gscale / count
// (0 when no pixel is valid); loss accumulates into out[n].
__global__ void k_loss_finalize(const double* __restrict__ part, int nblk, int N, double weight,
                                double gscale, double* __restrict__ loss_out,
                                float* __restrict__ scale) {
    const int n = blockIdx.x * 64 + threadIdx.x;
    if (n >= N) return;
    double s = 0.0, c = 0.0;
    for (int b = 0; b < nblk; ++b) {
        s += part[((long long)n * nblk + b) * 2];
        c += part[((long long)n * nblk + b) * 2 + 1];
    }
    if (c > 0.0) {
        loss_out[n] += weight * (s / c);
        scale[n] = (float)(weight * gscale / c);
    } else {
        scale[n] = 0.f;
    }
}

__global__ void k_scale_samples(float* __restrict__ g, const int* __restrict__ dims,
                                long long stride, int chans, const float* __restrict__ scale) {
    const int n = blockIdx.y;
    const long long m = (long long)chans * dh(dims, n) * dw(dims, n);
    const float sc = scale[n];
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < m; i += (long long)gridDim.x * 256)
        g[(long long)n * stride + i] *= sc;
}

// disparity_ce (:1274-1323) per sample: gcost = q - p (unscaled), partial
// sums of -q log p (p > 1e-12) and the count of valid pixels with qsum > 0.
__global__ void k_ce_partial(const float* __restrict__ probs, const float* __restrict__ target,
                             const uint8_t* __restrict__ valid, const int* __restrict__ dims,
                             long long pstride, long long tstride, int D, float sigma,
                             float* __restrict__ gcost, double* __restrict__ part, int nblk) {
    __shared__ double sl[256];
    __shared__ double sc[256];
    const int n = blockIdx.y;
    const int hw = dh(dims, n) * dw(dims, n);
    const float inv2s2 = 1.f / (2.f * sigma * sigma);
    double l = 0.0, cnt = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < hw; i += nblk * 256) {
        const long long po = (long long)n * pstride + i;
        bool use = valid[(long long)n * tstride + i] != 0;
        float qsum = 0.f;
        const float dg = target[(long long)n * tstride + i];
        if (use) {
            for (int d = 0; d < D; ++d) {
                const float e = (float)d - dg;
                qsum += __expf(-e * e * inv2s2);
            }
            use = qsum > 0.f;
        }
        if (!use) {
            for (int d = 0; d < D; ++d) gcost[po + (long long)d * hw] = 0.f;
            continue;
        }
        const float qinv = 1.f / qsum;
        for (int d = 0; d < D; ++d) {
            const float e = (float)d - dg;
            const float q = __expf(-e * e * inv2s2) * qinv;
            const float pk = probs[po + (long long)d * hw];
            if (pk > 1.0e-12f) l -= (double)q * (double)__logf(pk);
            gcost[po + (long long)d * hw] = q - pk;
        }
        cnt += 1.0;
    }
    sl[threadIdx.x] = l;
    sc[threadIdx.x] = cnt;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) {
            sl[threadIdx.x] += sl[threadIdx.x + st];
            sc[threadIdx.x] += sc[threadIdx.x + st];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[((long long)n * nblk + blockIdx.x) * 2] = sl[0];
        part[((long long)n * nblk + blockIdx.x) * 2 + 1] = sc[0];
    }
}

// -------------------------------------------------------------------- Adam
// Adam.step (:1368-1387): clip to +-clip, moments, bias-corrected update
// (c1 = 1/(1-b1^t), c2 = 1/(1-b2^t) from the running products).
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, long long n, float lr, float b1, float b2, float eps,
                       float c1, float c2, float clip) {
    const long long i = blockIdx.x * 256LL + threadIdx.x;
    if (i >= n) return;
    float gi = g[i];
    gi = gi > clip ? clip : (gi < -clip ? -clip : gi);
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
}

// ------------------------------------------------------------- selection
// k-th smallest value per sample (sorted(values)[k], _peak :3866-3869) by a
// radix select over order-preserving keys: three passes of 11, 11 and 10
// bits, most significant first.  Pass q histograms the keys that match the
// prefix chosen so far; the next launch moves the prefix to the bucket that
// holds the k-th key (every block of a sample computes the same bucket from
// the same counts, and block 0 records it for the launch after).
__device__ __forceinline__ uint32_t fkey(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

constexpr int SEL_PASSES = 3;
constexpr int SEL_BINS = 2048;
__device__ __forceinline__ int sel_shift(int q) { return q == 0 ? 21 : (q == 1 ? 10 : 0); }
__device__ __forceinline__ int sel_bins(int q) { return q == 2 ? 1024 : 2048; }

struct SelState {
    uint32_t prefix, k;
};

// (prefix, k) moved to the bucket of pass q's histogram that holds the k-th
// key: the bucket with excl <= k < incl, or the last bucket when k is past
// the end (as a serial scan that stops at the last bucket would).  Called by
// all 256 threads; each gets the result.
__device__ SelState sel_pick(const uint32_t* __restrict__ hist, int q, SelState s,
                             uint32_t* wsum, SelState* res) {
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int nb = sel_bins(q), per = nb / 256;
    uint32_t c[SEL_BINS / 256];
    uint32_t sum = 0;
#pragma unroll
    for (int j = 0; j < SEL_BINS / 256; ++j) {
        c[j] = j < per ? hist[t * per + j] : 0u;
        sum += c[j];
    }
    uint32_t v = sum;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t u = __shfl_up(v, o, 64);
        if (lane >= o) v += u;
    }
    if (lane == 63) wsum[wave] = v;
    __syncthreads();
    uint32_t run = v - sum;
    for (int w = 0; w < wave; ++w) run += wsum[w];
#pragma unroll
    for (int j = 0; j < SEL_BINS / 256; ++j) {
        if (j < per) {
            const uint32_t excl = run, incl = run + c[j];
            const int b = t * per + j;
            if ((excl <= s.k && s.k < incl) || (b == nb - 1 && incl <= s.k)) {
                res->prefix = s.prefix | ((uint32_t)b << sel_shift(q));
                res->k = s.k - excl;
            }
            run = incl;
        }
    }
    __syncthreads();
    const SelState r = *res;
    __syncthreads();
    return r;
}

// ghist: [SEL_PASSES][N][SEL_BINS] counts, st: [SEL_PASSES][N] states
__global__ void k_sel_init(const long long* __restrict__ kth, SelState* __restrict__ st,
                           uint32_t* __restrict__ ghist, int N) {
    const long long i = blockIdx.x * 256LL + threadIdx.x;
    if (i < (long long)SEL_PASSES * N * SEL_BINS) ghist[i] = 0;
    if (i < N) {
        st[i].prefix = 0;
        st[i].k = (uint32_t)kth[i];
    }
}

__global__ void __launch_bounds__(256) k_sel_hist(const float* __restrict__ x,
                                                  const int* __restrict__ dims, long long stride,
                                                  SelState* __restrict__ st,
                                                  uint32_t* __restrict__ ghist, int q, int N) {
    __shared__ uint32_t h[SEL_BINS];
    __shared__ uint32_t wsum[4];
    __shared__ SelState res;
    const int n = blockIdx.y;
    const int t = threadIdx.x;
    for (int i = t; i < SEL_BINS; i += 256) h[i] = 0;
    const long long m = (long long)dh(dims, n) * dw(dims, n);
    const float* xs = x + (long long)n * stride;
    // this block's first chunk of 256 x U values is requested before the
    // bucket choice (which waits on the previous pass's counts)
    constexpr int U = 16;
    constexpr long long CH = 256LL * U;
    long long b0 = (long long)blockIdx.x * CH;
    float v[U];
    auto load = [&]() {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const long long i = b0 + t + 256LL * u;
            v[u] = i < m ? xs[i] : 0.f;
        }
    };
    load();
    SelState s = st[(q == 0 ? 0 : q - 1) * N + n];
    if (q > 0) {
        s = sel_pick(ghist + ((long long)(q - 1) * N + n) * SEL_BINS, q - 1, s, wsum, &res);
        if (blockIdx.x == 0 && t == 0) st[(long long)q * N + n] = s;
    }
    __syncthreads();
    const int shift = sel_shift(q);
    const uint32_t bmask = (uint32_t)sel_bins(q) - 1u;
    const uint32_t mask = q == 0 ? 0u : (0xFFFFFFFFu << (shift + (q == 2 ? 10 : 11)));
    const uint32_t want = s.prefix & mask;
    while (b0 < m) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint32_t k = fkey(v[u]);
            const uint32_t bin = (k >> shift) & bmask;
            const bool act = b0 + t + 256LL * u < m && (k & mask) == want;
            if (act) atomicAdd(&h[bin], 1u);
        }
        b0 += (long long)gridDim.x * CH;
        if (b0 < m) load();
    }
    __syncthreads();
    uint32_t* gh = ghist + ((long long)q * N + n) * SEL_BINS;
    for (int i = t; i <= (int)bmask; i += 256) {
        const uint32_t c = h[i];
        if (c) atomicAdd(&gh[i], c);
    }
}

// the last pass's pick; one block of 256 threads per sample
__global__ void __launch_bounds__(256) k_sel_final(const SelState* __restrict__ st,
                                                   const uint32_t* __restrict__ ghist,
                                                   float* __restrict__ out, int N) {
    __shared__ uint32_t wsum[4];
    __shared__ SelState res;
    const int n = blockIdx.x;
    constexpr int q = SEL_PASSES - 1;
    const SelState s = sel_pick(ghist + ((long long)q * N + n) * SEL_BINS, q,
                                st[(long long)q * N + n], wsum, &res);
    if (threadIdx.x == 0) {
        const uint32_t pre = s.prefix;
        const uint32_t u = (pre & 0x80000000u) ? (pre & 0x7FFFFFFFu) : ~pre;
        out[n] = __uint_as_float(u);
    }
}

// float -> double widening (network disparity -> refinement centre)
__global__ void k_widen(const float* __restrict__ x, const int* __restrict__ dims,
                        long long xstride, double* __restrict__ y, long long ystride,
                        const float* __restrict__ mul) {
    const int n = blockIdx.y;
    const int hw = dh(dims, n) * dw(dims, n);
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= hw) return;
    y[(long long)n * ystride + i] = (double)x[(long long)n * xstride + i] * (double)(mul ? mul[n] : 1.f);
}

// predict_auto's width rule on the device (stereonet.py:3897-3926 with
// _peak :3866-3869 and predict_at_scale's working size :3943-3947), in the
// reference's float64 arithmetic.  One thread per pair; pair n is
// src_dims[n] = (H, W) and was probed at probe_dims[n].  pass 0 takes the
// probe's peak (times W / probe width) and activates every pair; a later
// pass keeps a pair active only if its width moved by more than 8%.  Writes the width used in this pass to trace[n][pass] (-1 when
// the pair is done) and the per-level dims tables of the pass (zero dims
// for a finished pair, which every kernel of the pass skips): table block
// [full N][half N][f4 N][full2 2N][f1_2 2N][f4_2 2N] of (h, w) pairs.
__device__ __forceinline__ int co_s2(int n) { return (n - 1) / 2 + 1; }

__global__ void k_auto_widths(const float* __restrict__ peaks, int N,
                              const int* __restrict__ src_dims, const int* __restrict__ probe_dims,
                              double tp, int max_width, int pass, int passes,
                              int* __restrict__ works, int* __restrict__ active,
                              int* __restrict__ trace, int* __restrict__ tables,
                              float* __restrict__ inv) {
    const int n = blockIdx.x * 64 + threadIdx.x;
    if (n >= N) return;
    const int H = dh(src_dims, n), W = dw(src_dims, n);
    double v = (double)peaks[n];
    v = v > 0.5 ? v : 0.5;
    const double pk = pass == 0 ? v * ((double)W / (double)dw(probe_dims, n)) : v;
    const double den = pk > 1.0e-6 ? pk : 1.0e-6;
    int w = (int)((double)W * tp / den + 0.5);
    if (w > W) w = W;
    if (w > max_width) w = max_width;
    if (w < 48) w = 48;
    bool act;
    if (pass == 0) {
        act = true;
        works[n] = w;
    } else {
        const int cur = works[n];
        act = active[n] != 0 && !((double)abs(w - cur) <= 0.08 * (double)cur);
        if (act) works[n] = w;
    }
    active[n] = act ? 1 : 0;
    trace[(long long)n * passes + pass] = act ? works[n] : -1;
    int hw = 0, ww = 0;
    double scale = 1.0;
    if (act) {
        const int wk = works[n];
        if (wk >= W) {
            hw = H;
            ww = W;
        } else {
            scale = (double)wk / (double)W;
            ww = wk;
            hw = (int)((double)H * scale + 0.5);
            if (hw < 8) hw = 8;
        }
    }
    inv[n] = act ? (float)(1.0 / scale) : 0.f;
    const int f1h = act ? co_s2(hw) : 0, f1w = act ? co_s2(ww) : 0;
    const int f4h = act ? co_s2(f1h) : 0, f4w = act ? co_s2(f1w) : 0;
    int* full = tables;
    int* half = tables + 2 * N;
    int* f4 = tables + 4 * N;
    int* full2 = tables + 6 * N;
    int* f1_2 = tables + 10 * N;
    int* f4_2 = tables + 14 * N;
    full[2 * n] = hw;
    full[2 * n + 1] = ww;
    half[2 * n] = hw / 2;
    half[2 * n + 1] = ww / 2;
    f4[2 * n] = f4h;
    f4[2 * n + 1] = f4w;
    for (int r = 0; r < 2; ++r) {
        const int m = r * N + n;
        full2[2 * m] = hw;
        full2[2 * m + 1] = ww;
        f1_2[2 * m] = f1h;
        f1_2[2 * m + 1] = f1w;
        f4_2[2 * m] = f4h;
        f4_2[2 * m + 1] = f4w;
    }
}

// pool_valid (:1579-1592): a coarse pixel is valid only if all four of its
// children are
__global__ void k_pool_valid(const uint8_t* __restrict__ v, const int* __restrict__ idims,
                             long long vstride, uint8_t* __restrict__ out,
                             const int* __restrict__ odims, long long ostride) {
    const int n = blockIdx.z;
    const int oy = blockIdx.y;
    const int ox = blockIdx.x * 64 + threadIdx.x;
    const int wi = dw(idims, n);
    const int ho = dh(odims, n), wo = dw(odims, n);
    if (ox >= wo || oy >= ho) return;
    const uint8_t* r0 = v + (long long)n * vstride + (long long)(2 * oy) * wi + 2 * ox;
    const uint8_t* r1 = r0 + wi;
    out[(long long)n * ostride + (long long)oy * wo + ox] = (r0[0] && r0[1] && r1[0] && r1[1]) ? 1 : 0;
}

// y = a*x + b*y over a flat range (slot tails hold unused values)
__global__ void k_axpby(const float* __restrict__ x, float* __restrict__ y, long long n, float a,
                        float b) {
    const long long i = blockIdx.x * 256LL + threadIdx.x;
    if (i >= n) return;
    y[i] = a * x[i] + b * y[i];
}

inline int tiles4(int h) { return (h + 3) / 4; }

}  // namespace

// =============================================================== C ABI
extern "C" {

int sv_conv_fwd(const float* x, const float* xr, const int* idims, long long xstride, int IC,
                int D, int KD, const float* w, const float* b, int OC, int s, int p, int dil,
                float* y, const int* odims, long long ystride, int N, int max_ho, int max_wo,
                int mode, float slope, const float* resid, long long rstride, void* stream) {
    SV_REQUIRE(x && idims && w && b && y && odims && N > 0 && IC > 0 && OC > 0 && D > 0,
               "sv_conv_fwd: bad arguments");
    SV_REQUIRE(KD == 1 || KD == 3, "sv_conv_fwd: depth kernel must be 1 or 3");
    SV_REQUIRE(mode != 2 || resid, "sv_conv_fwd: residual mode needs resid");
    hipStream_t st = (hipStream_t)stream;
    const int src = xr ? 1 : 0;
    SV_REQUIRE(xstride < (1LL << 31), "sv_conv_fwd: a sample's input must hold fewer than 2^31 values");
    // engine: the matrix-core kernels when chosen (tuning knob) or, by
    // default, for 4 to 16 output channels (3D layers: one column of output
    // depths per tile); the vector kernel otherwise or when the shape does
    // not fit a matrix-core kernel.  All sum every output in the same order
    // with single-rounding multiply-adds, so they agree bit for bit.
    {
        const int eng = sv::tune(SV_TUNE_CONV_ENGINE);
        if (eng == 5 && KD == 1 && D == 1 && !src) {
            const int rc = sv::conv_fwd_c4(x, idims, xstride, IC, w, b, OC, s, p, dil, y, odims,
                                           ystride, N, max_ho, max_wo, mode, slope, resid, rstride,
                                           st);
            if (rc != SV_EUNSUPPORTED) return rc;
        }
        if (eng == 7) {
            const int rc = sv::conv_fwd_direct(x, xr, idims, xstride, IC, D, KD, w, b, OC, s, p, dil,
                                               y, odims, ystride, N, max_ho, max_wo, mode, slope,
                                               resid, rstride, st);
            if (rc != SV_EUNSUPPORTED) return rc;
        }
        if (KD == 3 && (eng == 6 || (eng == 0 && OC >= 4 && OC <= 16))) {
            const int rc = sv::conv_fwd_col(x, xr, idims, xstride, IC, D, KD, w, b, OC, s, p, dil, y,
                                            odims, ystride, N, max_ho, max_wo, mode, slope, resid,
                                            rstride, st);
            if (rc != SV_EUNSUPPORTED) return rc;
        }
        if (eng >= 2 || (eng == 0 && OC >= 4 && OC <= 16)) {
            const int rc = sv::conv_fwd_m4(x, xr, idims, xstride, IC, D, KD, w, b, OC, s, p, dil, y,
                                           odims, ystride, N, max_ho, max_wo, mode, slope, resid,
                                           rstride, st);
            if (rc != SV_EUNSUPPORTED) return rc;
        }
    }
    {
        // LDS-tiled vector path.  Tile width: the widest of 64, 32, 16
        // columns that wastes at most 1/8 of the row on padding (or the
        // tuning knob's); rows per thread: 4 unless that leaves the GPU with
        // fewer than ~2 tiles per CU, then 1 (or the knob's); weights from
        // LDS (or the knob's source).  Input channels per staging pass: as
        // many as fit a 40 KB LDS budget.
        int TWc = sv::tune(SV_TUNE_CONV_TW);
        if (!TWc) {
            TWc = 16;
            for (int c : {64, 32}) {
                const int waste = (max_wo + c - 1) / c * c - max_wo;
                if (8 * waste <= max_wo) {
                    TWc = c;
                    break;
                }
            }
        }
        const int TYc = 256 / TWc;
        const long long ub4 = (long long)((max_wo + TWc - 1) / TWc) * ((max_ho + 4 * TYc - 1) / (4 * TYc)) * D * N;
        const int tuned = sv::tune(SV_TUNE_CONV_ROWS);
        const int RP = tuned ? tuned : (ub4 >= 2LL * sv::cu_count() ? 4 : 1);
        const int WSc = sv::tune(SV_TUNE_CONV_WSRC) ? sv::tune(SV_TUNE_CONV_WSRC) : 1;
        const int TT_H = TYc * RP;
        const int ocp = OC == 1 ? 4 : (OC <= 4 ? 4 : (OC <= 8 ? 8 : (OC <= 12 ? 12 : 16)));
        sv::ConvTiles ct;
        const long long shm = sv::conv_tiles(KD, IC, TWc, TT_H, s, dil, N,
                                             WSc == 1 ? (long long)KD * 9 * ocp : 0, 40 * 1024, ct);
        const int eng = sv::tune(SV_TUNE_CONV_ENGINE);
        const int IHc = (TT_H - 1) * s + 2 * dil + 1, IWc = (TWc - 1) * s + 2 * dil + 1;
        const size_t shc = 2 * ((size_t)KD * IHc * IWc + (size_t)KD * 9 * ocp) * sizeof(float);
        if ((eng == 4 || (eng == 0 && (shm < 0 || ct.ICG < IC))) && shc <= 64 * 1024) {
            // per-channel double-buffered kernel: the channels do not fit
            // one staging pass (or the knob asks for it)
            const int tYt = (max_ho + TT_H - 1) / TT_H;
            dim3 gc((max_wo + TWc - 1) / TWc, tYt * D, N), bkc(TWc, TYc);
#define LAUNCH_C(KD_, SRC_, OCB_, RP_, TW_)                                                       \
    hipLaunchKernelGGL((k_conv_chan<KD_, SRC_, OCB_, RP_, TW_, 1>), gc, bkc, shc, st, x, xr,     \
                       idims, xstride, IC, D, w, b, OC, s, p, dil, y, odims, ystride, tYt, mode, \
                       slope, resid, rstride, IHc, IWc)
#define LAUNCH_CT(KD_, SRC_, OCB_)                                                               \
    do {                                                                                         \
        if (RP == 4) {                                                                           \
            if (TWc == 64) LAUNCH_C(KD_, SRC_, OCB_, 4, 64);                                     \
            else if (TWc == 32) LAUNCH_C(KD_, SRC_, OCB_, 4, 32);                                \
            else LAUNCH_C(KD_, SRC_, OCB_, 4, 16);                                               \
        } else {                                                                                 \
            if (TWc == 64) LAUNCH_C(KD_, SRC_, OCB_, 1, 64);                                     \
            else if (TWc == 32) LAUNCH_C(KD_, SRC_, OCB_, 1, 32);                                \
            else LAUNCH_C(KD_, SRC_, OCB_, 1, 16);                                               \
        }                                                                                        \
    } while (0)
#define LAUNCH_CT_OC(KD_, SRC_)                                                                  \
    do {                                                                                         \
        if (OC == 1) LAUNCH_CT(KD_, SRC_, 1);                                                    \
        else if (OC <= 4) LAUNCH_CT(KD_, SRC_, 4);                                               \
        else if (OC <= 8) LAUNCH_CT(KD_, SRC_, 8);                                               \
        else if (OC <= 12) LAUNCH_CT(KD_, SRC_, 12);                                             \
        else LAUNCH_CT(KD_, SRC_, 16);                                                           \
    } while (0)
            if (KD == 1 && !src) LAUNCH_CT_OC(1, 0);
            else if (KD == 3 && !src) LAUNCH_CT_OC(3, 0);
            else if (KD == 3) LAUNCH_CT_OC(3, 1);
            else LAUNCH_CT_OC(1, 1);
#undef LAUNCH_CT_OC
#undef LAUNCH_CT
#undef LAUNCH_C
            SV_CHECK_LAUNCH();
            return SV_OK;
        }
        if (shm > 0) {
            const long long ub = (long long)((max_wo + TWc - 1) / TWc) * ((max_ho + TT_H - 1) / TT_H) * D * N;
            dim3 g(sv::conv_grid(ub, shm)), bk(TWc, TYc);
#define LAUNCH_TWS(KD_, SRC_, OCB_, RP_, TW_, WS_)                                              \
    hipLaunchKernelGGL((k_conv_tile<KD_, SRC_, OCB_, RP_, TW_, WS_>), g, bk, (size_t)shm, st, x, \
                       xr, idims, xstride, IC, D, w, b, OC, s, p, dil, y, odims, ystride, N,    \
                       mode, slope, resid, rstride, ct)
#define LAUNCH_TW(KD_, SRC_, OCB_, RP_, TW_)                                                      \
    do {                                                                                         \
        if (WSc == 1) LAUNCH_TWS(KD_, SRC_, OCB_, RP_, TW_, 1);                                  \
        else LAUNCH_TWS(KD_, SRC_, OCB_, RP_, TW_, 2);                                           \
    } while (0)
#define LAUNCH_T(KD_, SRC_, OCB_)                                                                \
    do {                                                                                         \
        if (RP == 4) {                                                                           \
            if (TWc == 64) LAUNCH_TW(KD_, SRC_, OCB_, 4, 64);                                    \
            else if (TWc == 32) LAUNCH_TW(KD_, SRC_, OCB_, 4, 32);                               \
            else LAUNCH_TW(KD_, SRC_, OCB_, 4, 16);                                              \
        } else {                                                                                 \
            if (TWc == 64) LAUNCH_TW(KD_, SRC_, OCB_, 1, 64);                                    \
            else if (TWc == 32) LAUNCH_TW(KD_, SRC_, OCB_, 1, 32);                               \
            else LAUNCH_TW(KD_, SRC_, OCB_, 1, 16);                                              \
        }                                                                                        \
    } while (0)
#define LAUNCH_T_OC(KD_, SRC_)                                                                   \
    do {                                                                                         \
        if (OC == 1) LAUNCH_T(KD_, SRC_, 1);                                                     \
        else if (OC <= 4) LAUNCH_T(KD_, SRC_, 4);                                                \
        else if (OC <= 8) LAUNCH_T(KD_, SRC_, 8);                                                \
        else if (OC <= 12) LAUNCH_T(KD_, SRC_, 12);                                              \
        else LAUNCH_T(KD_, SRC_, 16);                                                            \
    } while (0)
            if (KD == 1 && !src) LAUNCH_T_OC(1, 0);
            else if (KD == 3 && !src) LAUNCH_T_OC(3, 0);
            else if (KD == 3) LAUNCH_T_OC(3, 1);
            else LAUNCH_T_OC(1, 1);
#undef LAUNCH_T_OC
#undef LAUNCH_T
#undef LAUNCH_TW
#undef LAUNCH_TWS
            SV_CHECK_LAUNCH();
            return SV_OK;
        }
    }
    const int tY = tiles4(max_ho);
    dim3 grid((max_wo + 63) / 64, tY * D, N), blk(64, 4);
#define LAUNCH_FWD(OCB_, SRC_)                                                                \
    hipLaunchKernelGGL((k_conv_fwd<OCB_, SRC_>), grid, blk, 0, st, x, xr, idims, xstride, IC, D, \
                       KD, w, b, OC, s, p, dil, y, odims, ystride, tY, mode, slope, resid, rstride)
    if (src) {
        if (OC <= 8) LAUNCH_FWD(8, 1); else LAUNCH_FWD(16, 1);
    } else {
        if (OC == 1) LAUNCH_FWD(1, 0);
        else if (OC <= 8) LAUNCH_FWD(8, 0);
        else if (OC <= 12) LAUNCH_FWD(12, 0);
        else LAUNCH_FWD(16, 0);
    }
#undef LAUNCH_FWD
    SV_CHECK_LAUNCH();
    return SV_OK;
}

int sv_conv_bwd_data(const float* g, const int* odims, long long gstride, const float* w, int IC,
                     int OC, int D, int KD, int s, int p, int dil, float* gx, const int* idims,
                     long long gxstride, int N, int max_hi, int max_wi, int ic_count, float* wt_ws,
                     void* stream) {
    SV_REQUIRE(g && odims && w && gx && idims && N > 0, "sv_conv_bwd_data: bad arguments");
    const int ICn = ic_count > 0 ? ic_count : IC;
    if (wt_ws && s == 1 && p == dil) {
        // stride 1, same padding: the transposed convolution on the forward
        // engines (matrix cores for 4..16 output channels)
        const int TAPS = KD * 9;
        float* wt = wt_ws;
        float* zb = wt_ws + (size_t)ICn * OC * TAPS;
        const int n = ICn * OC * TAPS;
        hipLaunchKernelGGL(k_flip_weights, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                           w, OC, IC, TAPS, ICn, wt, zb);
        SV_CHECK_LAUNCH();
        return sv_conv_fwd(g, nullptr, odims, gstride, OC, D, KD, wt, zb, ICn, 1, p, dil, gx, idims,
                           gxstride, N, max_hi, max_wi, 0, 0.f, nullptr, 0, stream);
    }
    const int tY = tiles4(max_hi);
    dim3 grid((max_wi + 63) / 64, tY * D, N), blk(64, 4);
    hipStream_t st = (hipStream_t)stream;
    if (ICn <= 1)
        hipLaunchKernelGGL(k_conv_bwd_data<1>, grid, blk, 0, st, g, odims, gstride, w, IC, OC, D,
                           KD, s, p, dil, gx, idims, gxstride, tY, ICn);
    else if (ICn <= 8)
        hipLaunchKernelGGL(k_conv_bwd_data<8>, grid, blk, 0, st, g, odims, gstride, w, IC, OC, D,
                           KD, s, p, dil, gx, idims, gxstride, tY, ICn);
    else
        hipLaunchKernelGGL(k_conv_bwd_data<12>, grid, blk, 0, st, g, odims, gstride, w, IC, OC, D,
                           KD, s, p, dil, gx, idims, gxstride, tY, ICn);
    SV_CHECK_LAUNCH();
    return SV_OK;
}

int sv_conv_bwd_weight_workspace_bytes(int N, int IC, int OC, int KD, int nchunk, size_t* bytes) {
    SV_REQUIRE(bytes && N > 0 && nchunk > 0, "bad arguments");
    *bytes = (size_t)N * nchunk * OC * (IC * KD * 9 + 1) * sizeof(float);
    return SV_OK;
}

int sv_conv_bwd_weight(const float* x, const float* xr, const int* idims, long long xstride,
                       int IC, int D, int KD, const float* g, const int* odims, long long gstride,
                       int OC, int s, int p, int dil, int N, int nchunk, float* gw, float* gb,
                       float* partials, size_t ws_bytes, void* stream) {
    SV_REQUIRE(x && idims && g && odims && gw && gb && partials && N > 0 && nchunk > 0,
               "sv_conv_bwd_weight: bad arguments");
    SV_REQUIRE(OC <= 48, "sv_conv_bwd_weight: at most 48 output channels");
    size_t need = 0;
    sv_conv_bwd_weight_workspace_bytes(N, IC, OC, KD, nchunk, &need);
    if (ws_bytes < need) return sv::set_error(SV_EWORKSPACE, "conv weight gradient needs %zu bytes", need);
    const int ncol = IC * KD * 9;
    hipStream_t st = (hipStream_t)stream;
    // matrix-core kernel when every input channel's halo tile, the gradient
    // tile and the combine buffer fit 64 KB and the columns fit its units
    {
        sv::ConvTiles ct;
        const long long shm = sv::conv_tiles(KD, IC, WG_T, WG_T, s, dil, N, 0, 64 * 1024, ct);
        const int OCG = (OC + 3) / 4;
        const int NCG = (ncol + 63) / 64, PQ = NCG >= 4 ? 1 : 4 / NCG;
        const long long tb = 4LL * IC * KD * ct.IHt * ct.IWt + 4LL * 4 * OCG * WG_T * WG_T;
        const long long rb = 4LL * (NCG * PQ * 4 * OCG * 64 + OC * 16);
        const long long shn = std::max(tb, rb);
        if ((KD == 3 || !xr) && (KD == 1 || KD == 3) && shm > 0 && ct.ICG >= IC && OC <= 12 &&
            NCG * PQ <= 4 * WG_UMAX && shn <= 64 * 1024) {
            dim3 gn(nchunk, 1, N);
#define LAUNCH_WG(KD_, SRC_, OCG_)                                                               \
    hipLaunchKernelGGL((k_wgrad_m4<KD_, SRC_, OCG_>), gn, dim3(256), (size_t)shn, st, x, xr, idims, \
                       xstride, IC, D, g, odims, gstride, OC, s, p, dil, nchunk, partials, ct)
#define LAUNCH_WG_OC(KD_, SRC_)                                                                  \
    do {                                                                                         \
        if (OCG == 1) LAUNCH_WG(KD_, SRC_, 1);                                                   \
        else if (OCG == 2) LAUNCH_WG(KD_, SRC_, 2);                                              \
        else LAUNCH_WG(KD_, SRC_, 3);                                                            \
    } while (0)
            if (KD == 1) LAUNCH_WG_OC(1, 0);
            else if (xr) LAUNCH_WG_OC(3, 1);
            else LAUNCH_WG_OC(3, 0);
#undef LAUNCH_WG_OC
#undef LAUNCH_WG
            SV_CHECK_LAUNCH();
            const int tot = OC * (ncol + 1);
            hipLaunchKernelGGL(k_reduce_partials, dim3((tot + 63) / 64), dim3(64 * RP_WAVES), 0, st,
                               partials, N * nchunk, OC, ncol, gw, gb);
            SV_CHECK_LAUNCH();
            return SV_OK;
        }
    }
    dim3 grid(nchunk, (ncol + WG_COLS - 1) / WG_COLS, N);
#define LAUNCH_BW(SRC_, OCM_)                                                                  \
    hipLaunchKernelGGL((k_conv_bwd_weight<SRC_, OCM_>), grid, dim3(WG_COLS), 0, st, x, xr, idims, \
                       xstride, IC, D, KD, g, odims, gstride, OC, s, p, dil, nchunk, partials)
    if (xr) {
        if (OC <= 8) LAUNCH_BW(1, 8); else if (OC <= 24) LAUNCH_BW(1, 24); else LAUNCH_BW(1, 48);
    } else {
        if (OC <= 1) LAUNCH_BW(0, 1);
        else if (OC <= 8) LAUNCH_BW(0, 8);
        else if (OC <= 12) LAUNCH_BW(0, 12);
        else if (OC <= 24) LAUNCH_BW(0, 24);
        else LAUNCH_BW(0, 48);
    }
#undef LAUNCH_BW
    SV_CHECK_LAUNCH();
    const int tot = OC * (ncol + 1);
    hipLaunchKernelGGL(k_reduce_partials, dim3((tot + 63) / 64), dim3(64 * RP_WAVES), 0, st,
                       partials, N * nchunk, OC, ncol, gw, gb);
    SV_CHECK_LAUNCH();
    return SV_OK;
}

int sv_act_bwd(float* g, const float* y, long long n, float slope, void* stream) {
    SV_REQUIRE(g && y && n >= 0, "sv_act_bwd: bad arguments");
    if (n == 0) return SV_OK;
    hipLaunchKernelGGL(k_act_bwd, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, g, y, n, slope);
    SV_CHECK_LAUNCH();
    return SV_OK;
}

int sv_costvol_bwd(const float* fl, const float* fr, const int* dims, long long fstride, int C,
                   int D, const float* gvol, long long gvstride, float* gfl, float* gfr, int N,
                   int max_h, int max_w, void* stream) {
    SV_REQUIRE(fl && fr && dims && gvol && gfl && gfr && N > 0, "sv_costvol_bwd: bad arguments");
    const int tY = tiles4(max_h);
    hipLaunchKernelGGL(k_costvol_bwd, dim3((max_w + 63) / 64, tY * C, N), dim3(64, 4), 0,
                       (hipStream_t)stream, fl, fr, dims, fstride, C, D, gvol, gvstride, gfl, gfr,
                       tY);
    SV_CHECK_LAUNCH();
    return SV_OK;
}

int sv_softargmin_fwd(const float* cost, const int* dims, long long cstride, int D,
                      float penalty, float* disp, long long dstride, float* probs, int N,
                      int max_hw, void* stream) {
    SV_REQUIRE(cost && dims && disp && probs && N > 0, "sv_softargmin_fwd: bad arguments");
    hipLaunchKernelGGL(k_softargmin_fwd, dim3((max_hw + 255) / 256, N), dim3(256), 0,
                       (hipStream_t)stream, cost, dims, cstride, D, penalty, disp, dstride, probs);
    SV_CHECK_LAUNCH();
    return SV_OK;
}

int sv_conv3_softargmin(const float* x, const int* dims, long long xstride, int IC, int D,
                        const float* w, const float* b, float penalty, float* disp,
                        long long dstride, float* probs, long long pstride, int N, int max_h,
                        int max_w, void* stream) {
    SV_REQUIRE(x && dims && w && b && disp && probs && N > 0 && IC > 0 && max_h > 0 && max_w > 0,
               "sv_conv3_softargmin: bad arguments");
    dim3 grid((max_w + 7) / 8, (max_h + 7) / 8, N);
    switch (D) {
        case 12:
            hipLaunchKernelGGL(k_conv3_softargmin<12>, grid, dim3(256), 0, (hipStream_t)stream, x,
                               dims, xstride, IC, w, b, penalty, disp, dstride, probs, pstride);
            break;
        case 8:
            hipLaunchKernelGGL(k_conv3_softargmin<8>, grid, dim3(256), 0, (hipStream_t)stream, x,
                               dims, xstride, IC, w, b, penalty, disp, dstride, probs, pstride);
            break;
        default:
            return sv::set_error(SV_EUNSUPPORTED, "fused soft-argmin for %d disparities", D);
    }
    SV_CHECK_LAUNCH();
    return SV_OK;
}

int sv_load_pair(const void* l, const void* r, int dtype, const int* sdims, long long sstride,
                 float* img, long long fslot, const int* fdims, float* rin_h, long long hslot,
                 const int* hdims, float* rin_f, int N, int max_h, int max_w, void* stream) {
    SV_REQUIRE(l && r && sdims && img && fdims && rin_h && hdims && rin_f && N > 0 && max_h > 0 &&
                   max_w > 0, "sv_load_pair: bad arguments");
    dim3 grid((max_w + 63) / 64, (max_h + 3) / 4, N), blk(64, 4);
    hipStream_t st = (hipStream_t)stream;
    if (dtype == SV_U8)
        hipLaunchKernelGGL(k_load_pair<uint8_t>, grid, blk, 0, st, (const uint8_t*)l,
                           (const uint8_t*)r, sdims, sstride, img, fslot, fdims, rin_h, hslot,
                           hdims, rin_f, N);
    else if (dtype == SV_F32)
        hipLaunchKernelGGL(k_load_pair<float>, grid, blk, 0, st, (const float*)l, (const float*)r,
                           sdims, sstride, img, fslot, fdims, rin_h, hslot, hdims, rin_f, N);
    else
        return sv::set_error(SV_EINVAL, "sv_load_pair: images must be uint8 or float32");
    SV_CHECK_LAUNCH();
    return SV_OK;
}

int sv_refine_in(const float* src, const int* sdims, long long sslot, float* rin, const int* dims,
                 long long slot, const float* w, const float* b, int OC, int dil, float slope,
                 float* y, long long ystride, int N, int max_h, int max_w, void* stream) {
    SV_REQUIRE(src && sdims && rin && dims && w && b && y && N > 0 && max_h > 0 && max_w > 0,
               "sv_refine_in: bad arguments");
    if (dil < 1 || dil > 4)
        return sv::set_error(SV_EUNSUPPORTED, "sv_refine_in: dilation %d (1..4 supported)", dil);
    dim3 grid((max_w + 15) / 16, (max_h + 15) / 16, N);
    hipStream_t st = (hipStream_t)stream;
    switch (OC) {
        case 8:
            hipLaunchKernelGGL(k_refine_in<8>, grid, dim3(256), 0, st, src, sdims, sslot, rin, dims,
                               slot, w, b, dil, slope, y, ystride);
            break;
        case 32:
            hipLaunchKernelGGL(k_refine_in<32>, grid, dim3(256), 0, st, src, sdims, sslot, rin,
                               dims, slot, w, b, dil, slope, y, ystride);
            break;
        default:
            return sv::set_error(SV_EUNSUPPORTED, "sv_refine_in: %d output channels", OC);
    }
    SV_CHECK_LAUNCH();
    return SV_OK;
}

int sv_softargmin_bwd(const float* probs, const float* disp, const float* gdisp,
                      const float* gextra, const int* dims, long long cstride, int D,
                      long long dstride, float* gcost, int N, int max_hw, void* stream) {
    SV_REQUIRE(probs && disp && gdisp && dims && gcost && N > 0, "sv_softargmin_bwd: bad arguments");
    hipLaunchKernelGGL(k_softargmin_bwd, dim3((max_hw + 255) / 256, N), dim3(256), 0,
                       (hipStream_t)stream, probs, disp, gdisp, gextra, dims, cstride, D, dstride,
                       gcost);
    SV_CHECK_LAUNCH();
    return SV_OK;
}

int sv_bilinear_fwd(const void* x, int dtype, const int* idims, long long xstride, int C,
                    float* y, const int* odims, long long ystride, int och, float mul,
                    const float* muln, int N, int max_ho, int max_wo, void* stream) {
    SV_REQUIRE(x && idims && y && odims && N > 0 && C > 0, "sv_bilinear_fwd: bad arguments");
    const int tY = tiles4(max_ho);
    dim3 grid((max_wo + 64 * BL_X - 1) / (64 * BL_X), tY * C, N), blk(64, 4);
    hipStream_t st = (hipStream_t)stream;
    if (dtype == SV_U8)
        hipLaunchKernelGGL(k_bilinear_fwd<uint8_t>, grid, blk, 0, st, (const uint8_t*)x, idims,
                           xstride, C, y, odims, ystride, och, mul, muln, tY);
    else if (dtype == SV_F32)
        hipLaunchKernelGGL(k_bilinear_fwd<float>, grid, blk, 0, st, (const float*)x, idims, xstride,
                           C, y, odims, ystride, och, mul, muln, tY);
    else
        return sv::set_error(SV_EINVAL, "sv_bilinear_fwd: dtype %d", dtype);
    SV_CHECK_LAUNCH();
    return SV_OK;
}

int sv_bilinear_bwd(const float* gy, const int* odims, long long gystride, int ich, int C,
                    float* gx, const int* idims, long long gxstride, float mul, int N, int max_hi,
                    int max_wi, void* stream) {
    SV_REQUIRE(gy && odims && gx && idims && N > 0, "sv_bilinear_bwd: bad arguments");
    const int tY = tiles4(max_hi);
    hipLaunchKernelGGL(k_bilinear_bwd, dim3((max_wi + 63) / 64, tY * C, N), dim3(64, 4), 0,
                       (hipStream_t)stream, gy, odims, gystride, ich, C, gx, idims, gxstride, mul,
                       tY);
    SV_CHECK_LAUNCH();
    return SV_OK;
}

int sv_avgpool2(const void* x, int dtype, const int* idims, long long xstride, float* y,
                const int* odims, long long ystride, int och, float mul, int N, int max_ho,
                int max_wo, void* stream) {
    SV_REQUIRE(x && idims && y && odims && N > 0, "sv_avgpool2: bad arguments");
    dim3 grid((max_wo + 63) / 64, max_ho, N);
    hipStream_t st = (hipStream_t)stream;
    if (dtype == SV_U8)
        hipLaunchKernelGGL(k_avgpool2<uint8_t>, grid, dim3(64), 0, st, (const uint8_t*)x, idims,
                           xstride, y, odims, ystride, och, mul);
    else if (dtype == SV_F32)
        hipLaunchKernelGGL(k_avgpool2<float>, grid, dim3(64), 0, st, (const float*)x, idims, xstride,
                           y, odims, ystride, och, mul);
    else
        return sv::set_error(SV_EINVAL, "sv_avgpool2: dtype %d", dtype);
    SV_CHECK_LAUNCH();
    return SV_OK;
}

int sv_copy_channel(const void* x, int dtype, const int* dims, long long xstride, float* y,
                    long long ystride, int och, int N, int max_hw, void* stream) {
    SV_REQUIRE(x && dims && y && N > 0, "sv_copy_channel: bad arguments");
    dim3 grid((max_hw + 255) / 256, N);
    hipStream_t st = (hipStream_t)stream;
    if (dtype == SV_U8)
        hipLaunchKernelGGL(k_copy_chan<uint8_t>, grid, dim3(256), 0, st, (const uint8_t*)x, dims,
                           xstride, y, ystride, och);
    else if (dtype == SV_F32)
        hipLaunchKernelGGL(k_copy_chan<float>, grid, dim3(256), 0, st, (const float*)x, dims,
                           xstride, y, ystride, och);
    else
        return sv::set_error(SV_EINVAL, "sv_copy_channel: dtype %d", dtype);
    SV_CHECK_LAUNCH();
    return SV_OK;
}

// pseudo_huber over N samples: grad (scaled by weight*gscale/count per
// sample) and loss_out[n] += weight * loss_n.  part: 2*N*nblk doubles,
// scale: N floats.
int sv_pseudo_huber(const float* pred, const float* target, const uint8_t* valid,
                    const int* dims, long long stride, int N, float alpha, double weight,
                    double gscale, float* grad, double* loss_out, double* part, float* scale,
                    int nblk, void* stream) {
    SV_REQUIRE(pred && target && valid && dims && grad && loss_out && part && scale && N > 0 &&
                   nblk > 0,
               "sv_pseudo_huber: bad arguments");
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_ph_partial, dim3(nblk, N), dim3(256), 0, st, pred, target, valid, dims,
                       stride, alpha, grad, part, nblk);
    SV_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_loss_finalize, dim3((N + 63) / 64), dim3(64), 0, st, part, nblk, N,
                       weight, gscale, loss_out, scale);
    SV_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_scale_samples, dim3(64, N), dim3(256), 0, st, grad, dims, stride, 1, scale);
    SV_CHECK_LAUNCH();
    return SV_OK;
}

// disparity_ce over N samples; gcost scaled by weight*gscale/count.
int sv_disparity_ce(const float* probs, const float* target, const uint8_t* valid,
                    const int* dims, long long pstride, long long tstride, int D, float sigma,
                    int N, double weight, double gscale, float* gcost, double* loss_out,
                    double* part, float* scale, int nblk, void* stream) {
    SV_REQUIRE(probs && target && valid && dims && gcost && loss_out && part && scale && N > 0,
               "sv_disparity_ce: bad arguments");
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_ce_partial, dim3(nblk, N), dim3(256), 0, st, probs, target, valid, dims,
                       pstride, tstride, D, sigma, gcost, part, nblk);
    SV_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_loss_finalize, dim3((N + 63) / 64), dim3(64), 0, st, part, nblk, N,
                       weight, gscale, loss_out, scale);
    SV_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_scale_samples, dim3(64, N), dim3(256), 0, st, gcost, dims, pstride, D, scale);
    SV_CHECK_LAUNCH();
    return SV_OK;
}

int sv_adam_step(float* p, const float* g, float* m, float* v, long long n, float lr, float b1,
                 float b2, float eps, float c1, float c2, float clip, void* stream) {
    SV_REQUIRE(p && g && m && v && n > 0, "sv_adam_step: bad arguments");
    hipLaunchKernelGGL(k_adam, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, p, g, m, v, n, lr, b1, b2, eps, c1, c2, clip);
    SV_CHECK_LAUNCH();
    return SV_OK;
}

int sv_select_kth_workspace_bytes(int N, size_t* bytes) {
    SV_REQUIRE(bytes && N > 0, "bad arguments");
    *bytes = (size_t)N * SEL_PASSES * (sizeof(SelState) + SEL_BINS * sizeof(uint32_t));
    return SV_OK;
}

// ws: sv_select_kth_workspace_bytes(N) bytes; max_count: the largest sample
// size, which sizes the number of blocks per sample.
int sv_select_kth(const float* x, const int* dims, long long stride, const long long* kth,
                  float* out, int N, long long max_count, void* ws, size_t ws_bytes,
                  void* stream) {
    SV_REQUIRE(x && dims && kth && out && N > 0, "sv_select_kth: bad arguments");
    size_t need = 0;
    sv_select_kth_workspace_bytes(N, &need);
    if (!ws || ws_bytes < need) return sv::set_error(SV_EWORKSPACE, "sv_select_kth needs %zu bytes", need);
    hipStream_t st = (hipStream_t)stream;
    uint32_t* ghist = (uint32_t*)ws;
    SelState* state = (SelState*)(ghist + (size_t)SEL_PASSES * N * SEL_BINS);
    const long long ninit = (long long)SEL_PASSES * N * SEL_BINS;
    hipLaunchKernelGGL(k_sel_init, dim3((unsigned)((ninit + 255) / 256)), dim3(256), 0, st, kth,
                       state, ghist, N);
    SV_CHECK_LAUNCH();
    // one chunk of 4096 values per block up to 64 blocks per sample
    const int nblk = (int)std::min<long long>(64, std::max<long long>(1, (max_count + 4095) / 4096));
    for (int q = 0; q < SEL_PASSES; ++q) {
        hipLaunchKernelGGL(k_sel_hist, dim3(nblk, N), dim3(256), 0, st, x, dims, stride, state,
                           ghist, q, N);
        SV_CHECK_LAUNCH();
    }
    hipLaunchKernelGGL(k_sel_final, dim3(N), dim3(256), 0, st, state, ghist, out, N);
    SV_CHECK_LAUNCH();
    return SV_OK;
}

int sv_pool_valid(const uint8_t* v, const int* idims, long long vstride, uint8_t* out,
                  const int* odims, long long ostride, int N, int max_ho, int max_wo,
                  void* stream) {
    SV_REQUIRE(v && idims && out && odims && N > 0, "sv_pool_valid: bad arguments");
    hipLaunchKernelGGL(k_pool_valid, dim3((max_wo + 63) / 64, max_ho, N), dim3(64), 0,
                       (hipStream_t)stream, v, idims, vstride, out, odims, ostride);
    SV_CHECK_LAUNCH();
    return SV_OK;
}

int sv_axpby(const float* x, float* y, long long n, float a, float b, void* stream) {
    SV_REQUIRE(x && y && n >= 0, "sv_axpby: bad arguments");
    if (n == 0) return SV_OK;
    hipLaunchKernelGGL(k_axpby, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, x, y, n, a, b);
    SV_CHECK_LAUNCH();
    return SV_OK;
}

int sv_widen(const float* x, const int* dims, long long xstride, double* y, long long ystride,
             const float* mul, int N, int max_hw, void* stream) {
    SV_REQUIRE(x && dims && y && N > 0, "sv_widen: bad arguments");
    hipLaunchKernelGGL(k_widen, dim3((max_hw + 255) / 256, N), dim3(256), 0, (hipStream_t)stream,
                       x, dims, xstride, y, ystride, mul);
    SV_CHECK_LAUNCH();
    return SV_OK;
}

// predict_auto's per-pair width decision and dims tables (k_auto_widths).
int sv_auto_widths(const float* peaks, int N, const int* src_dims, const int* probe_dims,
                   double target_peak, int max_width, int pass, int passes, int* works, int* active,
                   int* trace, int* tables, float* inv, void* stream) {
    SV_REQUIRE(peaks && src_dims && probe_dims && works && active && trace && tables && inv && N > 0,
               "sv_auto_widths: bad arguments");
    SV_REQUIRE(pass >= 0 && pass < passes, "sv_auto_widths: pass %d outside [0, %d)", pass, passes);
    hipLaunchKernelGGL(k_auto_widths, dim3((N + 63) / 64), dim3(64), 0, (hipStream_t)stream, peaks,
                       N, src_dims, probe_dims, target_peak, max_width, pass, passes, works, active,
                       trace, tables, inv);
    SV_CHECK_LAUNCH();
    return SV_OK;
}

}  // extern "C"
