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

#include "common.h"

namespace {

typedef short v4i16 __attribute__((ext_vector_type(4)));
typedef float v4f32 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ short bf16_rne(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}

__device__ __forceinline__ int dh(const int* d, int n) { return d[2 * n]; }
__device__ __forceinline__ int dw(const int* d, int n) { return d[2 * n + 1]; }

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


// ------------------------------------------------------------------ FP32
// The same convolution on the matrix cores in full FP32
// (v_mfma_f32_16x16x4_f32: FP32 products, FP32 accumulation), for the FP32
// network.  A block makes 16*CT output columns x 4*(4/CT) rows at one
// output depth; wave w owns 4/CT consecutive rows, each as CT 16-position
// M-tiles (four M-tiles per wave).  The input channels are
// taken four at a time: their halo tiles (all KD depth taps, zero outside
// the image, the cost volume formed while staging when SRC == 1) are staged
// in LDS, and the 4 x KD*9 products of the chunk are split so that lane
// group q = lane / 16 (the MFMA's K index) walks the KD*9 taps of channel q
// while the four groups advance together, one tap per MFMA.  Weights for
// the chunk sit in registers, one per tap and 16-channel N-tile.
//   A[m = lane % 16][k = lane / 16]: input at position m, channel q, tap t
//   B[k = lane / 16][n = lane % 16]: w[oc = n + 16 nt][ic0 + q][t]
//   D[m = 4 (lane / 16) + i][n = lane % 16]
template <int KD, int SRC, int NT, int CT>
__global__ void __launch_bounds__(256) k_conv_mf32(
    const float* __restrict__ x, const float* __restrict__ xr, const int* __restrict__ idims,
    long long xstride, int IC, int D, const float* __restrict__ w, const float* __restrict__ b,
    int OC, int s, int p, int dil, float* __restrict__ y, const int* __restrict__ odims,
    long long ystride, int tilesY, int mode, float slope, const float* __restrict__ resid,
    long long rstride, int IHt, int IWt) {
    constexpr int T = KD * 9;
    constexpr int PD = KD / 2;
    constexpr int RW = 4 / CT;  // rows per wave
    extern __shared__ __align__(16) float smem_m[];
    const int n = blockIdx.z;
    const int od = blockIdx.y / tilesY;
    const int ty0 = (blockIdx.y - od * tilesY) * 4 * RW;
    const int tx0 = blockIdx.x * 16 * CT;
    const int hi = dh(idims, n), wi = dw(idims, n);
    const int ho = dh(odims, n), wo = dw(odims, n);
    if (tx0 >= wo || ty0 >= ho) return;  // uniform per block
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int q = lane >> 4, mcol = lane & 15;
    const int plane_t = IHt * IWt;
    const int chan_t = KD * plane_t;
    const int iy0 = ty0 * s - p, ix0 = tx0 * s - p;
    const long long iplane = (long long)hi * wi;
    const float* xs = x + (long long)n * xstride;
    const float* xrs = SRC == 1 ? xr + (long long)n * xstride : nullptr;
    v4f32 acc[4][NT];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[mt][t] = (v4f32){0.f, 0.f, 0.f, 0.f};
    // this lane's A position within the tile for M-tile 0: row RW * wave,
    // column mcol; M-tile mt adds row mt / CT and 16 (mt % CT) columns
    const int apos = (wave * RW * s) * IWt + mcol * s;
    for (int ic0 = 0; ic0 < IC; ic0 += 4) {
        const int icc = min(4, IC - ic0);
        // stage the chunk's halo tiles
        for (int e = tid; e < icc * chan_t; e += 256) {
            const int c = e / chan_t;
            int r = e - c * chan_t;
            const int kd = r / plane_t;
            r -= kd * plane_t;
            const int yy = r / IWt, xx = r - yy * IWt;
            const int dd = od + kd - PD, iy = iy0 + yy, ix = ix0 + xx;
            const int ic = ic0 + c;
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
            smem_m[e] = v;
        }
        // this lane's weights for the chunk: channel ic0 + q, every tap
        float bw[NT][T];
        const bool qok = q < icc;
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const int oc = t * 16 + mcol;
            const bool ok = qok && oc < OC;
            const float* wp = w + ((long long)oc * IC + ic0 + q) * T;
#pragma unroll
            for (int k = 0; k < T; ++k) bw[t][k] = ok ? wp[k] : 0.f;
        }
        __syncthreads();
        // channel of this lane's group (a group past the chunk reads a loaded
        // channel; its weights are zero)
        const float* ab = smem_m + (qok ? q : icc - 1) * chan_t + apos;
#pragma unroll
        for (int kd = 0; kd < KD; ++kd)
#pragma unroll
            for (int ky = 0; ky < 3; ++ky)
#pragma unroll
                for (int kx = 0; kx < 3; ++kx) {
                    const int k = (kd * 3 + ky) * 3 + kx;
                    const float* ak = ab + kd * plane_t + ky * dil * IWt + kx * dil;
#pragma unroll
                    for (int mt = 0; mt < 4; ++mt) {
                        const float a = ak[(mt / CT) * s * IWt + (mt % CT) * 16 * s];
#pragma unroll
                        for (int t = 0; t < NT; ++t)
                            acc[mt][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, bw[t][k], acc[mt][t],
                                                                              0, 0, 0);
                    }
                }
        __syncthreads();
    }
    // epilogue: D[m = 4 q + i][oc = 16 t + mcol], position (row ty0 + RW
    // wave + mt / CT, column tx0 + 16 (mt % CT) + m)
    const long long oplane = (long long)ho * wo;
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const int oc = t * 16 + mcol;
        if (oc >= OC) continue;
        const float bias = b[oc];
#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int oy = ty0 + RW * wave + mt / CT;
                const int ox = tx0 + (mt % CT) * 16 + 4 * q + i;
                if (ox >= wo || oy >= ho) continue;
                float* ys = y + (long long)n * ystride + (long long)od * oplane + (long long)oy * wo;
                float v = acc[mt][t][i] + bias;
                if (mode == 1) {
                    v = v > 0.f ? v : v * slope;
                } else if (mode == 2) {
                    v = resid[(long long)n * rstride + (long long)oc * D * oplane +
                              (long long)od * oplane + (long long)oy * wo + ox] + v;
                    v = v > 0.f ? v : 0.f;
                }
                ys[(long long)oc * D * oplane + ox] = v;
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


// FP32 conv forward on the matrix cores (k_conv_mf32): same arguments and
// results as sv_conv_fwd up to FP32 rounding order; OC at most 48.
int sv_conv_fwd_mf32(const float* x, const float* xr, const int* idims, long long xstride, int IC,
                     int D, int KD, const float* w, const float* b, int OC, int s, int p, int dil,
                     float* y, const int* odims, long long ystride, int N, int max_ho, int max_wo,
                     int mode, float slope, const float* resid, long long rstride, void* stream) {
    SV_REQUIRE(x && idims && w && b && y && odims && N > 0 && IC > 0 && D > 0,
               "sv_conv_fwd_mf32: bad arguments");
    SV_REQUIRE(OC >= 1 && OC <= 48, "sv_conv_fwd_mf32: OC must be 1..48 (got %d)", OC);
    SV_REQUIRE(KD == 1 || KD == 3, "sv_conv_fwd_mf32: depth kernel must be 1 or 3");
    SV_REQUIRE(mode != 2 || resid, "sv_conv_fwd_mf32: residual mode needs resid");
    SV_REQUIRE(s >= 1 && dil >= 1, "sv_conv_fwd_mf32: stride and dilation must be positive");
    // column tiles per block: the widest of 4, 2, 1 (x16 columns) that
    // wastes at most 1/8 of a row (or the tiled convolution's tuning knob)
    int CT = sv::tune(SV_TUNE_CONV_TW) / 16;
    if (!CT) {
        CT = 1;
        for (int c : {4, 2}) {
            const int waste = (max_wo + 16 * c - 1) / (16 * c) * 16 * c - max_wo;
            if (8 * waste <= max_wo) {
                CT = c;
                break;
            }
        }
    }
    const int RW = 4 / CT;
    const int IHt = (4 * RW - 1) * s + 2 * dil + 1, IWt = (16 * CT - 1) * s + 2 * dil + 1;
    const size_t shm = (size_t)4 * KD * IHt * IWt * sizeof(float);
    SV_REQUIRE(shm <= 64 * 1024, "sv_conv_fwd_mf32: input tile of %zu bytes is too large", shm);
    const int NT = (OC + 15) / 16;
    const int tY = (max_ho + 4 * RW - 1) / (4 * RW);
    dim3 grid((max_wo + 16 * CT - 1) / (16 * CT), tY * D, N);
    hipStream_t st = (hipStream_t)stream;
#define LAUNCH_MF_CT(KD_, SRC_, NT_, CT_)                                                       \
    hipLaunchKernelGGL((k_conv_mf32<KD_, SRC_, NT_, CT_>), grid, dim3(256), shm, st, x, xr,      \
                       idims, xstride, IC, D, w, b, OC, s, p, dil, y, odims, ystride, tY, mode,  \
                       slope, resid, rstride, IHt, IWt)
#define LAUNCH_MF(KD_, SRC_, NT_)                                                               \
    do {                                                                                        \
        if (CT == 4) LAUNCH_MF_CT(KD_, SRC_, NT_, 4);                                           \
        else if (CT == 2) LAUNCH_MF_CT(KD_, SRC_, NT_, 2);                                      \
        else LAUNCH_MF_CT(KD_, SRC_, NT_, 1);                                                   \
    } while (0)
#define LAUNCH_MF_NT(KD_, SRC_)                                                                 \
    do {                                                                                        \
        if (NT == 1) LAUNCH_MF(KD_, SRC_, 1);                                                   \
        else if (NT == 2) LAUNCH_MF(KD_, SRC_, 2);                                              \
        else LAUNCH_MF(KD_, SRC_, 3);                                                           \
    } while (0)
    if (xr) {
        SV_REQUIRE(KD == 3, "sv_conv_fwd_mf32: the cost-volume input needs a 3-D kernel");
        LAUNCH_MF_NT(3, 1);
    } else if (KD == 3) {
        LAUNCH_MF_NT(3, 0);
    } else {
        LAUNCH_MF_NT(1, 0);
    }
#undef LAUNCH_MF_NT
#undef LAUNCH_MF
#undef LAUNCH_MF_CT
    SV_CHECK_LAUNCH();
    return SV_OK;
}

}  // extern "C"
