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
