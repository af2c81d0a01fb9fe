// BF16 MFMA implicit-GEMM convolution forward for gfx950 (the 4x-widened
// StereoNet: 48/24/32 channels).  Same semantics, ragged-batch layout and
// epilogues as k_conv_fwd in net.hip (Conv2d.forward stereonet.py:489-526,
// Conv3d.forward :596-641, cost volume :771-799 formed on the fly); the
// products run on the matrix cores in BF16 with FP32 accumulation.
//
// GEMM view: out[pos][oc] = sum_k A[pos][k] * B[k][oc] with
// k = tap * IC + ic (tap-major; IC is a multiple of 4 so the four K values a
// lane feeds share one tap).  A is gathered from the FP32 activations and
// rounded to BF16 in registers (im2col never stored); B is the weight tensor
// pre-packed on the host into the MFMA operand layout, one 8-byte load per
// lane per K step.  Each wave owns 16 output positions x NT*16 channels and
// issues v_mfma_f32_16x16x16_bf16 per (K step, channel tile).
//
// Operand layout of v_mfma_f32_16x16x16_bf16 (wave of 64): lane l holds
//   A[m = l % 16][k = 4*(l / 16) + i], i = 0..3
//   B[k = 4*(l / 16) + i][n = l % 16]
//   D[m = 4*(l / 16) + i][n = l % 16]
#include <cmath>
#include <type_traits>

#include "common.h"
#include "conv_tiles.h"

namespace {

typedef short v4i16 __attribute__((ext_vector_type(4)));
typedef float v4f32 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ short bf16_rne(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}


template <int NT, int SRC>
__global__ void __launch_bounds__(256) k_conv_mfma(
    const float* __restrict__ x, const float* __restrict__ xr, const int* __restrict__ idims,
    long long xstride, int IC, int D, int KD, const v4i16* __restrict__ wpack,
    const float* __restrict__ b, int OC, int s, int p, int dil, int ksteps,
    float* __restrict__ y, const int* __restrict__ odims, long long ystride, int mode,
    float slope, const float* __restrict__ resid, long long rstride) {
    const int n = blockIdx.y;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int hi = dh(idims, n), wi = dw(idims, n);
    const int ho = dh(odims, n), wo = dw(odims, n);
    const long long oplane = (long long)ho * wo;
    const long long npos = oplane * D;
    const long long tile0 = ((long long)blockIdx.x * 4 + wave) * 16;
    if (tile0 >= npos) return;
    const int pd = KD / 2;
    const long long iplane = (long long)hi * wi;
    const float* xs = x + (long long)n * xstride;
    const float* xrs = SRC == 1 ? xr + (long long)n * xstride : nullptr;
    // this lane's A row: output position m = lane % 16
    const long long pos = tile0 + (lane & 15);
    const bool pin = pos < npos;
    int od = 0, oy = 0, ox = 0;
    if (pin) {
        od = (int)(pos / oplane);
        const int rem = (int)(pos - (long long)od * oplane);
        oy = rem / wo;
        ox = rem - oy * wo;
    }
    const int kq = 4 * (lane >> 4);  // this lane's K offset within a step
    const int Ktot = IC * KD * 9;
    v4f32 acc[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = (v4f32){0.f, 0.f, 0.f, 0.f};
    for (int ks = 0; ks < ksteps; ++ks) {
        const int k0 = ks * 16 + kq;
        v4i16 a = {0, 0, 0, 0};
        if (pin && k0 < Ktot) {
            const int tap = k0 / IC;
            const int ic0 = k0 - tap * IC;
            const int kd = tap / 9;
            const int ky = (tap - kd * 9) / 3;
            const int kx = tap - kd * 9 - ky * 3;
            const int dd = od + kd - pd;
            const int iy = oy * s - p + ky * dil;
            const int ix = ox * s - p + kx * dil;
            if (dd >= 0 && dd < D && iy >= 0 && iy < hi && ix >= 0 && ix < wi) {
                float v[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int ic = ic0 + i;
                    if (SRC == 1) {
                        if (ix < dd) {
                            v[i] = 2.0f;
                        } else {
                            const long long o = (long long)ic * iplane + (long long)iy * wi;
                            v[i] = fabsf(xs[o + ix] - xrs[o + ix - dd]);
                        }
                    } else {
                        v[i] = xs[(long long)ic * iplane * D + (long long)dd * iplane +
                                  (long long)iy * wi + ix];
                    }
                }
                a = (v4i16){bf16_rne(v[0]), bf16_rne(v[1]), bf16_rne(v[2]), bf16_rne(v[3])};
            }
        }
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const v4i16 bb = wpack[((long long)ks * NT + t) * 64 + lane];
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bb, acc[t], 0, 0, 0);
        }
    }
    // epilogue: D[m = kq + i][oc = t*16 + lane%16]
    float* ys = y + (long long)n * ystride;
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const int oc = t * 16 + (lane & 15);
        if (oc >= OC) continue;
        const float bias = b[oc];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const long long q = tile0 + kq + i;
            if (q >= npos) continue;
            float v = acc[t][i] + bias;
            if (mode == 1) {
                v = v > 0.f ? v : v * slope;
            } else if (mode == 2) {
                v = resid[(long long)n * rstride + (long long)oc * npos + q] + v;
                v = v > 0.f ? v : 0.f;
            }
            ys[(long long)oc * npos + q] = v;
        }
    }
}


// FP32 convolution on the matrix cores with v_mfma_f32_4x4x1_16b_f32, the
// same sums as k_conv_tile (net.hip) bit for bit: each MFMA step is one
// single-rounding multiply-add per output, taken in the same order (input
// channel, then depth, row, column tap; bias last).
//
// Operand layout (wave of 64): the 16 blocks of the instruction are 16
// groups of 4 output pixels, lane l <-> pixel l of a 64-pixel group (TW
// columns x 64/TW rows).  B (1x4 per block) is the input value of lane l's
// pixel at the current tap; A (4x1) is the weights of 4 output channels at
// that tap, broadcast from one block to all 16 (CBSZ = 4, ABID = tap % 16),
// so a single VGPR holds the 4-channel weights of 16 taps: lane 4t + i holds
// w[oc0 + i][ic][16j + t].  D: VGPR r of lane l is output channel oc0 + r
// of pixel l.  The weights of every tap stay in registers for a whole input
// channel; each tap costs one LDS read per 64 pixels for OCG MFMAs.
//
// A block (4 waves) makes tiles of TW columns x 4*P*(64/TW) rows, P pixel
// groups per wave, over the launch's tile list (conv_tiles.h); OCG groups of
// 4 output channels (OC <= 4*OCG).
template <int KD, int SRC, int OCG, int TW, int P>
__global__ void __launch_bounds__(256) k_conv_m4(
    const float* __restrict__ x, const float* __restrict__ xr, const int* __restrict__ idims,
    long long xstride, int IC, int D, const float* __restrict__ w, const float* __restrict__ b,
    int OC, int s, int p, int dil, float* __restrict__ y, const int* __restrict__ odims,
    long long ystride, int N, int mode, float slope, const float* __restrict__ resid,
    long long rstride, sv::ConvTiles ct) {
    constexpr int TAPS = KD * 9;
    constexpr int NJ = (TAPS + 15) / 16;  // weight VGPRs per (channel, 4 output channels)
    constexpr int GR = 64 / TW;           // rows of one 64-pixel group
    constexpr int TT_H = 4 * P * GR;      // tile rows
    extern __shared__ __align__(16) float smem_f[];
    const int IWt = ct.IWt, ICG = ct.ICG;
    const int plane_t = ct.IHt * IWt;
    const int tsz = KD * plane_t;
    float* tile = smem_f;                             // [ICG][KD][IHt][IWt]
    float* wl = smem_f + ICG * tsz;                   // [ICG][OCG][NJ][64]
    int* pre = reinterpret_cast<int*>(wl + ICG * OCG * NJ * 64);  // [N + 1]
    __shared__ int wsum[4];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int lc = lane % TW, lr = lane / TW;
    const int total = conv_tile_prefix<TW, TT_H>(odims, N, D, pre, wsum, tid);
    v4f32 acc[P][OCG];
    for (int t = conv_first_tile(); t < total; t += gridDim.x) {
        const TileAt at = conv_tile_at<TW, TT_H>(t, pre, N, D, odims);
        const int n = at.n, od = at.od, ty0 = at.ty0, tx0 = at.tx0;
        const int hin = dh(idims, n), win = dw(idims, n);
        const int ho = dh(odims, n), wo = dw(odims, n);
        const float* xs = x + (long long)n * xstride;
        const float* xrs = SRC == 1 ? xr + (long long)n * xstride : nullptr;
#pragma unroll
        for (int q = 0; q < P; ++q)
#pragma unroll
            for (int g = 0; g < OCG; ++g) acc[q][g] = (v4f32){0.f, 0.f, 0.f, 0.f};
        for (int ic0 = 0; ic0 < IC; ic0 += ICG) {
            const int icn = min(ICG, IC - ic0);
            conv_stage_inputs<KD, SRC>(tile, xs, xrs, ic0, icn, D, od, ty0 * s - p, tx0 * s - p,
                                       hin, win, ct, tid);
            const int WE = icn * OCG * NJ * 64;
            for (int e = tid; e < WE; e += 256) {
                const int l = e & 63;
                int q = e >> 6;
                const int j = q % NJ;
                q /= NJ;
                const int g = q % OCG, c = q / OCG;
                const int tp = 16 * j + (l >> 2), oc = 4 * g + (l & 3);
                wl[e] = (tp < TAPS && oc < OC) ? w[((long long)oc * IC + ic0 + c) * TAPS + tp] : 0.f;
            }
            __syncthreads();
            for (int c = 0; c < icn; ++c) {
                float wv[OCG][NJ];
#pragma unroll
                for (int g = 0; g < OCG; ++g)
#pragma unroll
                    for (int j = 0; j < NJ; ++j) wv[g][j] = wl[((c * OCG + g) * NJ + j) * 64 + lane];
                const float* cur = tile + c * tsz + lr * s * IWt + lc * s;
                auto step = [&](auto tpc) {
                    constexpr int tp = decltype(tpc)::value;
                    constexpr int kd = tp / 9, ky = (tp / 3) % 3, kx = tp % 3;
                    float bv[P];
#pragma unroll
                    for (int q = 0; q < P; ++q)
                        bv[q] = cur[kd * plane_t + ((wave * P + q) * GR * s + ky * dil) * IWt + kx * dil];
#pragma unroll
                    for (int g = 0; g < OCG; ++g)
#pragma unroll
                        for (int q = 0; q < P; ++q)
                            acc[q][g] = __builtin_amdgcn_mfma_f32_4x4x1f32(wv[g][tp / 16], bv[q], acc[q][g],
                                                                           4, tp % 16, 0);
                };
                TapSteps<0, TAPS>::run(step);
            }
            __syncthreads();
        }
        const int ox = tx0 + lc;
        const long long oplane = (long long)ho * wo;
        float* ys = y + (long long)n * ystride + (long long)od * oplane;
#pragma unroll
        for (int q = 0; q < P; ++q) {
            const int oy = ty0 + (wave * P + q) * GR + lr;
            if (ox >= wo || oy >= ho) continue;
#pragma unroll
            for (int g = 0; g < OCG; ++g)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int oc = 4 * g + r;
                    if (oc >= OC) break;
                    const long long o = (long long)oc * D * oplane + (long long)oy * wo + ox;
                    ys[o] = conv_finish(acc[q][g][r], b[oc], mode, slope, resid,
                                        (long long)n * rstride + (long long)od * oplane + o);
                }
        }
    }
}

// Depth-sliding form of k_conv_m4 for 3D convolutions (KD = 3): a block
// owns a spatial tile of a sample for a chunk of DC consecutive output
// depths.  The input planes of all IC channels live in a ring of three
// depth slots in LDS; output depth od reads planes od-1, od, od+1, and
// moving to od+1 stages just plane od+2 (the cost volume |fl - fr| of that
// plane when SRC == 1) in place of plane od-1, so a chunk stages DC + 2
// planes instead of 3*DC.  Tiles are numbered (sample, spatial tile, chunk),
// chunk fastest; the chunk length trades staging against the number of
// tiles.  The weights of every channel are staged once per block.  Sums,
// operand layout and epilogue are those of k_conv_m4.
template <int SRC>
__device__ __forceinline__ void stage_plane(float* ring, const float* __restrict__ xs,
                                            const float* __restrict__ xrs, int IC, int D, int dd,
                                            int slot, int iy0, int ix0, int hin, int win,
                                            const sv::ConvTiles& ct, int tid) {
    constexpr int U = 8;
    const int IWt = ct.IWt;
    const int plane_t = ct.IHt * IWt;
    const int iplane = hin * win;
    const int cstride = SRC == 1 ? iplane : D * iplane;
    for (int pp = tid; pp < plane_t; pp += 256) {
        const int yy = div_magic(pp, ct.mW);
        const int xx = pp - yy * IWt;
        const int iy = iy0 + yy, ix = ix0 + xx;
        const bool dok = dd >= 0 && dd < D && iy >= 0 && iy < hin && ix >= 0 && ix < win;
        const bool fill2 = SRC == 1 && dok && ix < dd;
        const bool ld = dok && !fill2;
        const int g0 = SRC == 1 ? iy * win + ix : dd * iplane + iy * win + ix;
        float* dst = ring + slot * plane_t + pp;
#pragma unroll 1
        for (int c0 = 0; c0 < IC; c0 += U) {
            float v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int g = g0 + (c0 + u) * cstride;
                v[u] = fill2 ? 2.0f : 0.f;
                if (ld && c0 + u < IC) {
                    if (SRC == 1) v[u] = fabsf(xs[g] - xrs[g - dd]);
                    else v[u] = xs[g];
                }
            }
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (c0 + u < IC) dst[(c0 + u) * 3 * plane_t] = v[u];
        }
    }
}

template <int SRC, int OCG, int TW, int P>
__global__ void __launch_bounds__(256) k_conv_m4d(
    const float* __restrict__ x, const float* __restrict__ xr, const int* __restrict__ idims,
    long long xstride, int IC, int D, const float* __restrict__ w, const float* __restrict__ b,
    int OC, int s, int p, int dil, float* __restrict__ y, const int* __restrict__ odims,
    long long ystride, int N, int mode, float slope, const float* __restrict__ resid,
    long long rstride, sv::ConvTiles ct, int DC) {
    constexpr int TAPS = 27;
    constexpr int NJ = 2;
    constexpr int GR = 64 / TW;
    constexpr int TT_H = 4 * P * GR;
    extern __shared__ __align__(16) float smem_f[];
    const int IWt = ct.IWt;
    const int plane_t = ct.IHt * IWt;
    float* ring = smem_f;                              // [IC][3][IHt][IWt]
    float* wl = smem_f + IC * 3 * plane_t;             // [IC][OCG][NJ][64]
    int* pre = reinterpret_cast<int*>(wl + IC * OCG * NJ * 64);  // [N + 1]
    __shared__ int wsum[4];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int lc = lane % TW, lr = lane / TW;
    // weights of every channel, once per block
    for (int e = tid; e < IC * OCG * NJ * 64; e += 256) {
        const int l = e & 63;
        int q = e >> 6;
        const int j = q % NJ;
        q /= NJ;
        const int g = q % OCG, c = q / OCG;
        const int tp = 16 * j + (l >> 2), oc = 4 * g + (l & 3);
        wl[e] = (tp < TAPS && oc < OC) ? w[((long long)oc * IC + c) * TAPS + tp] : 0.f;
    }
    const int nch = (D + DC - 1) / DC;
    const int total = conv_tile_prefix<TW, TT_H>(odims, N, nch, pre, wsum, tid);
    v4f32 acc[P][OCG];
    for (int t = conv_first_tile(); t < total; t += gridDim.x) {
        const TileAt at = conv_tile_at<TW, TT_H>(t, pre, N, nch, odims);
        const int n = at.n, ty0 = at.ty0, tx0 = at.tx0;
        const int od0 = at.od * DC, od1 = min(D, od0 + DC);
        const int hin = dh(idims, n), win = dw(idims, n);
        const int ho = dh(odims, n), wo = dw(odims, n);
        const int iy0 = ty0 * s - p, ix0 = tx0 * s - p;
        const float* xs = x + (long long)n * xstride;
        const float* xrs = SRC == 1 ? xr + (long long)n * xstride : nullptr;
        const long long oplane = (long long)ho * wo;
        float* ys = y + (long long)n * ystride;
        // planes od0-1, od0, od0+1; plane dd lives in slot (dd + 3) % 3
        stage_plane<SRC>(ring, xs, xrs, IC, D, od0 - 1, (od0 + 2) % 3, iy0, ix0, hin, win, ct, tid);
        stage_plane<SRC>(ring, xs, xrs, IC, D, od0, od0 % 3, iy0, ix0, hin, win, ct, tid);
        stage_plane<SRC>(ring, xs, xrs, IC, D, od0 + 1, (od0 + 1) % 3, iy0, ix0, hin, win, ct, tid);
        __syncthreads();
        for (int od = od0; od < od1; ++od) {
            if (od > od0) {
                stage_plane<SRC>(ring, xs, xrs, IC, D, od + 1, (od + 1) % 3, iy0, ix0, hin, win,
                                 ct, tid);
                __syncthreads();
            }
#pragma unroll
            for (int q = 0; q < P; ++q)
#pragma unroll
                for (int g = 0; g < OCG; ++g) acc[q][g] = (v4f32){0.f, 0.f, 0.f, 0.f};
            // ring slot of input plane od + kd - 1
            const int so0 = ((od + 2) % 3) * plane_t, so1 = (od % 3) * plane_t,
                      so2 = ((od + 1) % 3) * plane_t;
            for (int c = 0; c < IC; ++c) {
                float wv[OCG][NJ];
#pragma unroll
                for (int g = 0; g < OCG; ++g)
#pragma unroll
                    for (int j = 0; j < NJ; ++j) wv[g][j] = wl[((c * OCG + g) * NJ + j) * 64 + lane];
                const float* cur = ring + c * 3 * plane_t + lr * s * IWt + lc * s;
                auto step = [&](auto tpc) {
                    constexpr int tp = decltype(tpc)::value;
                    constexpr int kd = tp / 9, ky = (tp / 3) % 3, kx = tp % 3;
                    const int so = kd == 0 ? so0 : (kd == 1 ? so1 : so2);
                    float bv[P];
#pragma unroll
                    for (int q = 0; q < P; ++q)
                        bv[q] = cur[so + ((wave * P + q) * GR * s + ky * dil) * IWt + kx * dil];
#pragma unroll
                    for (int g = 0; g < OCG; ++g)
#pragma unroll
                        for (int q = 0; q < P; ++q)
                            acc[q][g] = __builtin_amdgcn_mfma_f32_4x4x1f32(wv[g][tp / 16], bv[q], acc[q][g],
                                                                           4, tp % 16, 0);
                };
                TapSteps<0, TAPS>::run(step);
            }
            const int ox = tx0 + lc;
#pragma unroll
            for (int q = 0; q < P; ++q) {
                const int oy = ty0 + (wave * P + q) * GR + lr;
                if (ox >= wo || oy >= ho) continue;
#pragma unroll
                for (int g = 0; g < OCG; ++g)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int oc = 4 * g + r;
                        if (oc >= OC) break;
                        const long long o = (long long)oc * D * oplane + (long long)od * oplane +
                                            (long long)oy * wo + ox;
                        ys[o] = conv_finish(acc[q][g][r], b[oc], mode, slope, resid,
                                            (long long)n * rstride + o);
                    }
            }
            __syncthreads();
        }
    }
}

}  // namespace

extern "C" {

// Pack size (bytes) of a weight tensor for sv_conv_fwd_mfma.
int sv_conv_mfma_pack_bytes(int IC, int OC, int KD, size_t* bytes) {
    SV_REQUIRE(bytes && IC > 0 && OC > 0, "bad arguments");
    const int ksteps = (IC * KD * 9 + 15) / 16;
    const int NT = (OC + 15) / 16;
    *bytes = (size_t)ksteps * NT * 64 * 4 * sizeof(short);
    return SV_OK;
}

// Conv forward on the matrix cores; wpack is the host-packed BF16 weight
// operand (see the layout note above).  IC must be a multiple of 4 and OC at
// most 64.
int sv_conv_fwd_mfma(const float* x, const float* xr, const int* idims, long long xstride, int IC,
                     int D, int KD, const void* wpack, const float* b, int OC, int s, int p,
                     int dil, float* y, const int* odims, long long ystride, int N,
                     long long max_npos, int mode, float slope, const float* resid,
                     long long rstride, void* stream) {
    SV_REQUIRE(x && idims && wpack && b && y && odims && N > 0, "sv_conv_fwd_mfma: bad arguments");
    SV_REQUIRE(IC % 4 == 0, "sv_conv_fwd_mfma: IC must be a multiple of 4 (got %d)", IC);
    SV_REQUIRE(OC >= 1 && OC <= 64, "sv_conv_fwd_mfma: OC must be 1..64 (got %d)", OC);
    SV_REQUIRE(KD == 1 || KD == 3, "sv_conv_fwd_mfma: depth kernel must be 1 or 3");
    SV_REQUIRE(mode != 2 || resid, "sv_conv_fwd_mfma: residual mode needs resid");
    const int ksteps = (IC * KD * 9 + 15) / 16;
    const int NT = (OC + 15) / 16;
    dim3 grid((unsigned)((max_npos + 63) / 64), N);
    hipStream_t st = (hipStream_t)stream;
    const v4i16* wp = (const v4i16*)wpack;
#define LAUNCH_M(NT_, SRC_)                                                                   \
    hipLaunchKernelGGL((k_conv_mfma<NT_, SRC_>), grid, dim3(256), 0, st, x, xr, idims, xstride, \
                       IC, D, KD, wp, b, OC, s, p, dil, ksteps, y, odims, ystride, mode, slope,   \
                       resid, rstride)
    if (xr) {
        switch (NT) {
            case 1: LAUNCH_M(1, 1); break;
            case 2: LAUNCH_M(2, 1); break;
            case 3: LAUNCH_M(3, 1); break;
            default: LAUNCH_M(4, 1); break;
        }
    } else {
        switch (NT) {
            case 1: LAUNCH_M(1, 0); break;
            case 2: LAUNCH_M(2, 0); break;
            case 3: LAUNCH_M(3, 0); break;
            default: LAUNCH_M(4, 0); break;
        }
    }
#undef LAUNCH_M
    SV_CHECK_LAUNCH();
    return SV_OK;
}


}  // extern "C"

namespace sv {

int conv_fwd_m4(const float* x, const float* xr, const int* idims, long long xstride, int IC,
                int D, int KD, const float* w, const float* b, int OC, int s, int p, int dil,
                float* y, const int* odims, long long ystride, int N, int max_ho, int max_wo,
                int mode, float slope, const float* resid, long long rstride, hipStream_t st) {
    const int src = xr ? 1 : 0;
    if (OC > 16 || (KD == 1 && src)) return SV_EUNSUPPORTED;
    // 64-pixel group shape: the widest of 64, 32, 16 columns that wastes at
    // most 1/8 of a row (or the tuning knob's); groups per wave: 2 (or the
    // knob's)
    int TWc = tune(SV_TUNE_CONV_TW);
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
    const int P = tune(SV_TUNE_CONV_MGROUPS) ? tune(SV_TUNE_CONV_MGROUPS) : 2;
    const int OCG = (OC + 3) / 4;
    const int TT_H = 4 * P * (64 / TWc);
    const int NJ = (KD * 9 + 15) / 16;
    ConvTiles ct;
    const int eng = tune(SV_TUNE_CONV_ENGINE);
    const long long ub3 = (long long)((max_wo + TWc - 1) / TWc) * ((max_ho + TT_H - 1) / TT_H) * N;
    if (KD == 3 && (eng == 2 || eng == 0)) {
        // depth-sliding kernel when every channel's three-plane ring and
        // weights fit 64 KB
        const long long shm3 = conv_tiles(3, IC, TWc, TT_H, s, dil, N, (long long)OCG * NJ * 64,
                                          64 * 1024, ct);
        if (shm3 > 0 && ct.ICG == IC) {
            // chunk of output depths per tile: all of them when that leaves
            // at least two spatial tiles per CU, else short enough to reach
            // two tiles (or one chunk per depth pair) per CU
            const long long want = 2LL * cu_count();
            int nch = (int)std::min<long long>(D, std::max<long long>(1, (want + ub3 - 1) / ub3));
            int DC = (D + nch - 1) / nch;
            if (DC < 2 && D >= 2) DC = 2;
            nch = (D + DC - 1) / DC;
            dim3 g3(conv_grid(ub3 * nch, shm3)), bk3(256);
#define LAUNCH_M4D(SRC_, OCG_, TW_, P_)                                                         \
    hipLaunchKernelGGL((k_conv_m4d<SRC_, OCG_, TW_, P_>), g3, bk3, (size_t)shm3, st, x, xr,     \
                       idims, xstride, IC, D, w, b, OC, s, p, dil, y, odims, ystride, N, mode,  \
                       slope, resid, rstride, ct, DC)
#define LAUNCH_M4D_P(SRC_, OCG_, TW_)                                                           \
    do {                                                                                        \
        if (P == 1) LAUNCH_M4D(SRC_, OCG_, TW_, 1);                                             \
        else if (P == 2) LAUNCH_M4D(SRC_, OCG_, TW_, 2);                                        \
        else LAUNCH_M4D(SRC_, OCG_, TW_, 4);                                                    \
    } while (0)
#define LAUNCH_M4D_TW(SRC_, OCG_)                                                               \
    do {                                                                                        \
        if (TWc == 64) LAUNCH_M4D_P(SRC_, OCG_, 64);                                            \
        else if (TWc == 32) LAUNCH_M4D_P(SRC_, OCG_, 32);                                       \
        else LAUNCH_M4D_P(SRC_, OCG_, 16);                                                      \
    } while (0)
#define LAUNCH_M4D_OC(SRC_)                                                                     \
    do {                                                                                        \
        if (OCG == 1) LAUNCH_M4D_TW(SRC_, 1);                                                   \
        else if (OCG == 2) LAUNCH_M4D_TW(SRC_, 2);                                              \
        else if (OCG == 3) LAUNCH_M4D_TW(SRC_, 3);                                              \
        else LAUNCH_M4D_TW(SRC_, 4);                                                            \
    } while (0)
            if (src) LAUNCH_M4D_OC(1);
            else LAUNCH_M4D_OC(0);
#undef LAUNCH_M4D_OC
#undef LAUNCH_M4D_TW
#undef LAUNCH_M4D_P
#undef LAUNCH_M4D
            SV_CHECK_LAUNCH();
            return SV_OK;
        }
    }
    const long long shm = conv_tiles(KD, IC, TWc, TT_H, s, dil, N, (long long)OCG * NJ * 64,
                                     40 * 1024, ct);
    if (shm < 0) return SV_EUNSUPPORTED;
    const long long ub = (long long)((max_wo + TWc - 1) / TWc) * ((max_ho + TT_H - 1) / TT_H) * D * N;
    dim3 g(conv_grid(ub, shm)), bk(256);
#define LAUNCH_M4(KD_, SRC_, OCG_, TW_, P_)                                                     \
    hipLaunchKernelGGL((k_conv_m4<KD_, SRC_, OCG_, TW_, P_>), g, bk, (size_t)shm, st, x, xr,    \
                       idims, xstride, IC, D, w, b, OC, s, p, dil, y, odims, ystride, N, mode,  \
                       slope, resid, rstride, ct)
#define LAUNCH_M4_P(KD_, SRC_, OCG_, TW_)                                                       \
    do {                                                                                        \
        if (P == 1) LAUNCH_M4(KD_, SRC_, OCG_, TW_, 1);                                         \
        else if (P == 2) LAUNCH_M4(KD_, SRC_, OCG_, TW_, 2);                                    \
        else LAUNCH_M4(KD_, SRC_, OCG_, TW_, 4);                                                \
    } while (0)
#define LAUNCH_M4_TW(KD_, SRC_, OCG_)                                                           \
    do {                                                                                        \
        if (TWc == 64) LAUNCH_M4_P(KD_, SRC_, OCG_, 64);                                        \
        else if (TWc == 32) LAUNCH_M4_P(KD_, SRC_, OCG_, 32);                                   \
        else LAUNCH_M4_P(KD_, SRC_, OCG_, 16);                                                  \
    } while (0)
#define LAUNCH_M4_OC(KD_, SRC_)                                                                 \
    do {                                                                                        \
        if (OCG == 1) LAUNCH_M4_TW(KD_, SRC_, 1);                                               \
        else if (OCG == 2) LAUNCH_M4_TW(KD_, SRC_, 2);                                          \
        else if (OCG == 3) LAUNCH_M4_TW(KD_, SRC_, 3);                                          \
        else LAUNCH_M4_TW(KD_, SRC_, 4);                                                        \
    } while (0)
    if (KD == 1) LAUNCH_M4_OC(1, 0);
    else if (src) LAUNCH_M4_OC(3, 1);
    else LAUNCH_M4_OC(3, 0);
#undef LAUNCH_M4_OC
#undef LAUNCH_M4_TW
#undef LAUNCH_M4_P
#undef LAUNCH_M4
    SV_CHECK_LAUNCH();
    return SV_OK;
}

}  // namespace sv
