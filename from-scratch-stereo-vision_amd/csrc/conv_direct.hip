// Direct convolution for small layers (Conv2d/Conv3d forward,
// stereonet.py:489-526, :596-641; the first 3D layer with the cost volume
// of CostVolume, :771-840, formed on the fly): one output position per
// thread, every input value read straight from memory (neighbouring threads
// read neighbouring addresses, so the loads of a wave coalesce and the 3x3
// window's overlap is served by the L1/L2 caches), no LDS staging and no
// barriers.  It suits layers with little arithmetic per launch -- the
// refinement's 8 -> 1 layers, the probe pass's maps of a few hundred pixels
// -- whose time the staged kernels spend mostly on their staging round trip
// and barriers.
//
// Sums: the same single-rounding multiply-adds in the same order as the
// other engines (input channel, depth, row, column tap; bias last; taps in
// the zero padding contribute fmaf(0, w, acc), which leaves an accumulator
// that started at +0 unchanged, like the staged kernels' zero-filled tiles
// or skipped planes), so the results agree bit for bit.
#include "common.h"
#include "conv_tiles.h"

namespace {

template <int OCB, int KD, int SRC>
__global__ void __launch_bounds__(256) k_conv_direct(
    const float* __restrict__ x, const float* __restrict__ xr, const int* __restrict__ idims,
    long long xstride, int IC, int D, const float* __restrict__ w, const float* __restrict__ b,
    int OC, int s, int p, int dil, float* __restrict__ y, const int* __restrict__ odims,
    long long ystride, int tilesY, int mode, float slope, const float* __restrict__ resid,
    long long rstride) {
    constexpr int TAPS = KD * 9;
    const int n = blockIdx.z;
    const int od = blockIdx.y / tilesY;
    const int oy = (blockIdx.y - od * tilesY) * 4 + threadIdx.y;
    const int ox = blockIdx.x * 64 + threadIdx.x;
    const int hin = dh(idims, n), win = dw(idims, n);
    const int ho = dh(odims, n), wo = dw(odims, n);
    if (ox >= wo || oy >= ho) return;
    const long long iplane = (long long)hin * win;
    const float* xs = x + (long long)n * xstride;
    const float* xrs = SRC ? xr + (long long)n * xstride : nullptr;
    const int iy0 = oy * s - p, ix0 = ox * s - p;
    // row / column offsets and validity of the 3 taps (clamped addresses,
    // zero where the tap lies outside the image)
    int ro[3], co[3], cx[3];
    bool rv[3], cv[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int iy = iy0 + k * dil, ix = ix0 + k * dil;
        rv[k] = iy >= 0 && iy < hin;
        cv[k] = ix >= 0 && ix < win;
        ro[k] = min(max(iy, 0), hin - 1) * win;
        co[k] = min(max(ix, 0), win - 1);
        cx[k] = ix;
    }
    for (int oc0 = 0; oc0 < OC; oc0 += OCB) {
        float acc[OCB];
#pragma unroll
        for (int j = 0; j < OCB; ++j) acc[j] = 0.f;
        for (int ic = 0; ic < IC; ++ic) {
#pragma unroll
            for (int kd = 0; kd < KD; ++kd) {
                const int dd = od + kd - KD / 2;
                const bool dv = dd >= 0 && dd < D;
                float v[9];
                if constexpr (SRC == 1) {
                    // cost of disparity dd: |left - right at x - dd|, the
                    // invalid cost 2.0 where x - dd leaves the image
                    const float* fl = xs + (long long)ic * iplane;
                    const float* fr = xrs + (long long)ic * iplane;
                    const int dc = min(max(dd, 0), D - 1);
#pragma unroll
                    for (int t = 0; t < 9; ++t) {
                        const int r = ro[t / 3], c = co[t % 3];
                        float cvl = fabsf(fl[r + c] - fr[r + max(c - dc, 0)]);
                        cvl = cx[t % 3] < dd ? 2.0f : cvl;
                        v[t] = (dv && rv[t / 3] && cv[t % 3]) ? cvl : 0.f;
                    }
                } else {
                    const float* xp = xs + ((long long)ic * D + min(max(dd, 0), D - 1)) * iplane;
#pragma unroll
                    for (int t = 0; t < 9; ++t) {
                        const float raw = xp[ro[t / 3] + co[t % 3]];
                        v[t] = (dv && rv[t / 3] && cv[t % 3]) ? raw : 0.f;
                    }
                }
#pragma unroll
                for (int t = 0; t < 9; ++t)
#pragma unroll
                    for (int j = 0; j < OCB; ++j) {
                        const int oc = min(oc0 + j, OC - 1);
                        acc[j] = __builtin_fmaf(v[t], w[((long long)oc * IC + ic) * TAPS + kd * 9 + t],
                                                acc[j]);
                    }
            }
        }
        const long long oplane = (long long)ho * wo;
        float* ys = y + (long long)n * ystride + (long long)od * oplane + (long long)oy * wo + ox;
#pragma unroll
        for (int j = 0; j < OCB; ++j) {
            const int oc = oc0 + j;
            if (oc >= OC) break;
            const long long o = (long long)oc * D * oplane;
            ys[o] = conv_finish(acc[j], b[oc], mode, slope, resid,
                                (long long)n * rstride + o + (long long)od * oplane +
                                    (long long)oy * wo + ox);
        }
    }
}

}  // namespace

namespace sv {

int conv_fwd_direct(const float* x, const float* xr, const int* idims, long long xstride, int IC,
                    int D, int KD, const float* w, const float* b, int OC, int s, int p, int dil,
                    float* y, const int* odims, long long ystride, int N, int max_ho, int max_wo,
                    int mode, float slope, const float* resid, long long rstride, hipStream_t st) {
    if (KD != 1 && KD != 3) return SV_EUNSUPPORTED;
    if (xr && (KD != 3 || s != 1 || p != 1 || dil != 1)) return SV_EUNSUPPORTED;
    const int tY = (max_ho + 3) / 4;
    dim3 grid((max_wo + 63) / 64, tY * D, N), blk(64, 4);
#define LAUNCH_D(OCB_, KD_, SRC_)                                                                \
    hipLaunchKernelGGL((k_conv_direct<OCB_, KD_, SRC_>), grid, blk, 0, st, x, xr, idims, xstride, \
                       IC, D, w, b, OC, s, p, dil, y, odims, ystride, tY, mode, slope, resid,      \
                       rstride)
#define LAUNCH_OC(KD_, SRC_)                   \
    if (OC == 1) LAUNCH_D(1, KD_, SRC_);        \
    else if (OC == 2) LAUNCH_D(2, KD_, SRC_);   \
    else if (OC <= 4) LAUNCH_D(4, KD_, SRC_);   \
    else if (OC <= 6) LAUNCH_D(6, KD_, SRC_);   \
    else if (OC <= 8) LAUNCH_D(8, KD_, SRC_);   \
    else if (OC <= 12) LAUNCH_D(12, KD_, SRC_); \
    else LAUNCH_D(16, KD_, SRC_)
    if (KD == 1) {
        LAUNCH_OC(1, 0);
    } else if (!xr) {
        LAUNCH_OC(3, 0);
    } else {
        LAUNCH_OC(3, 1);
    }
#undef LAUNCH_OC
#undef LAUNCH_D
    SV_CHECK_LAUNCH();
    return SV_OK;
}

}  // namespace sv
