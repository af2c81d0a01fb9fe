// 2D FP32 convolution on the matrix cores over 4-channel blocks: the same
// sums as k_conv_m4 (mfma_conv.hip) and k_conv_tile (net.hip) bit for bit --
// one v_mfma_f32_4x4x1_16b_f32 step per (input channel, tap), a single
// rounding each, input channel outer, taps in row-major order, bias last
// (Conv2d.forward, stereonet.py:489-526).
//
// What differs is how the operands reach the instruction:
// - the input tile sits in LDS as float4 per position holding 4 consecutive
//   input channels ([channel block][row][column][4]), so one ds_read_b128
//   hands a lane its pixel's value at one tap for 4 channels, i.e. 4 x OCG
//   MFMAs per LDS read instead of OCG (the read-instruction rate, not the
//   matrix cores, bounded the one-channel layout);
// - every weight lives in registers for the whole launch (lane 4t + i of
//   register (c, g) holds w[4g + i][c][t], the broadcast A operand of tap t);
// - the block walks its tiles with two LDS buffers: the next tile's inputs
//   are loaded into registers before this tile's MFMAs and stored after
//   them, so their latency is hidden behind the arithmetic.
// Operand layout otherwise as k_conv_m4: lane l <-> pixel l of a 64-pixel
// group (TW columns x 64/TW rows), D register r = output channel 4g + r.
#include <algorithm>

#include "common.h"
#include "conv_tiles.h"

namespace {

typedef float v4f32 __attribute__((ext_vector_type(4)));

constexpr int kC4MaxPerThread = 8;  // tile positions per thread (staging registers: 4 each)

struct C4Tiles {
    int IHt, IWt, E;   // halo tile rows, columns; positions (channel blocks x IHt x IWt)
    unsigned mW, mP;   // ceil(2^32 / IWt), ceil(2^32 / (IHt * IWt))
};

// The 4 channels of tile position e (channel block e / plane, row, column):
// raw values loaded from a position clamped into the image (every lane
// loads), and a 4-bit mask of the real ones (inside the image, below IC);
// the rest are zero.  Selecting at the LDS store keeps the compiler from
// waiting for the loads where they are issued.
__device__ __forceinline__ unsigned c4_load(float v[4], const float* __restrict__ xs, int e,
                                            int IC, int iy0, int ix0, int hin, int win,
                                            const C4Tiles& ct) {
    const int plane = ct.IHt * ct.IWt;
    const int cb = div_magic(e, ct.mP);
    const int pp = e - cb * plane;
    const int yy = div_magic(pp, ct.mW);
    const int xx = pp - yy * ct.IWt;
    const int iy = iy0 + yy, ix = ix0 + xx;
    const bool inb = e < ct.E && iy >= 0 && iy < hin && ix >= 0 && ix < win;
    const int iplane = hin * win;
    const float* src = xs + min(max(iy, 0), hin - 1) * win + min(max(ix, 0), win - 1);
    unsigned m = 0;
#pragma unroll
    for (int ci = 0; ci < 4; ++ci) {
        v[ci] = src[(long long)min(4 * cb + ci, IC - 1) * iplane];
        m |= (inb && 4 * cb + ci < IC) ? (1u << ci) : 0u;
    }
    return m;
}

__device__ __forceinline__ v4f32 c4_masked(const float v[4], unsigned m) {
    return (v4f32){(m & 1) ? v[0] : 0.f, (m & 2) ? v[1] : 0.f, (m & 4) ? v[2] : 0.f,
                   (m & 8) ? v[3] : 0.f};
}

// A block (4 waves) makes tiles of TW columns x 4*P*(64/TW) rows, P pixel
// groups per wave, over the launch's tile list (conv_tiles.h); OCG groups of
// 4 output channels (OC <= 4*OCG), CBM 4-channel blocks (IC <= 4*CBM).
template <int OCG, int TW, int P, int CBM>
__global__ void __launch_bounds__(256) k_conv_c4(
    const float* __restrict__ x, const int* __restrict__ idims, long long xstride, int IC,
    const float* __restrict__ w, const float* __restrict__ b, int OC, int s, int p, int dil,
    float* __restrict__ y, const int* __restrict__ odims, long long ystride, int N, int mode,
    float slope, const float* __restrict__ resid, long long rstride, C4Tiles ct) {
    constexpr int GR = 64 / TW;
    constexpr int TT_H = 4 * P * GR;
    constexpr int MAXE = kC4MaxPerThread;
    extern __shared__ __align__(16) float smem_f[];
    v4f32* buf = reinterpret_cast<v4f32*>(smem_f);      // [2][E]
    int* pre = reinterpret_cast<int*>(smem_f + 8 * ct.E);  // [N + 1]
    __shared__ int wsum[4];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int lc = lane % TW, lr = lane / TW;
    const int IWt = ct.IWt, plane = ct.IHt * ct.IWt;
    float wr[4 * CBM][OCG];
    {
        const int t = lane >> 2, i = lane & 3;
#pragma unroll
        for (int c = 0; c < 4 * CBM; ++c)
#pragma unroll
            for (int g = 0; g < OCG; ++g) {
                const int oc = 4 * g + i;
                wr[c][g] = (c < IC && t < 9 && oc < OC) ? w[((long long)oc * IC + c) * 9 + t] : 0.f;
            }
        // the weights complete here: otherwise the tile loop, which issues
        // the next tile's loads before its MFMAs, would wait for all of them
        __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
    }
    const int total = conv_tile_prefix<TW, TT_H>(odims, N, 1, pre, wsum, tid);
    int t = conv_first_tile();
    int cur = 0;
    TileAt at = {0, 0, 0, 0};
    if (t < total) {
        at = conv_tile_at<TW, TT_H>(t, pre, N, 1, odims);
        const int n = __builtin_amdgcn_readfirstlane(at.n);
        const int hin = dh(idims, n), win = dw(idims, n);
        const float* xs = x + (long long)n * xstride;
        for (int e = tid; e < ct.E; e += 256) {
            float v[4];
            const unsigned m = c4_load(v, xs, e, IC, at.ty0 * s - p, at.tx0 * s - p, hin, win, ct);
            buf[e] = c4_masked(v, m);
        }
    }
    __syncthreads();
    while (t < total) {
        const int n = __builtin_amdgcn_readfirstlane(at.n);
        const int ty0 = __builtin_amdgcn_readfirstlane(at.ty0);
        const int tx0 = __builtin_amdgcn_readfirstlane(at.tx0);
        const int tn = t + gridDim.x;
        TileAt an = at;
        float pf[MAXE][4];
        unsigned pm = 0;
        if (tn < total) {
            an = conv_tile_at<TW, TT_H>(tn, pre, N, 1, odims);
            const int nn = __builtin_amdgcn_readfirstlane(an.n);
            const int hin = dh(idims, nn), win = dw(idims, nn);
            const float* xs = x + (long long)nn * xstride;
            const int iy0 = __builtin_amdgcn_readfirstlane(an.ty0) * s - p;
            const int ix0 = __builtin_amdgcn_readfirstlane(an.tx0) * s - p;
#pragma unroll
            for (int k = 0; k < MAXE; ++k)
                pm |= c4_load(pf[k], xs, tid + 256 * k, IC, iy0, ix0, hin, win, ct) << (4 * k);
        }
        const v4f32* T = buf + cur * ct.E;
        v4f32 acc[P][OCG];
#pragma unroll
        for (int q = 0; q < P; ++q)
#pragma unroll
            for (int g = 0; g < OCG; ++g) acc[q][g] = (v4f32){0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int cb = 0; cb < CBM; ++cb) {
            if (4 * cb >= IC) break;
#pragma unroll
            for (int q = 0; q < P; ++q) {
                const int row = (wave * P + q) * GR + lr;
                const v4f32* base = T + cb * plane + row * s * IWt + lc * s;
                v4f32 bv[9];
#pragma unroll
                for (int tp = 0; tp < 9; ++tp) bv[tp] = base[(tp / 3) * dil * IWt + (tp % 3) * dil];
#pragma unroll
                for (int ci = 0; ci < 4; ++ci) {
                    if (4 * cb + ci >= IC) break;
                    const int c = 4 * cb + ci;
                    auto step = [&](auto tpc) {
                        constexpr int tp = decltype(tpc)::value;
#pragma unroll
                        for (int g = 0; g < OCG; ++g)
                            acc[q][g] = __builtin_amdgcn_mfma_f32_4x4x1f32(wr[c][g], bv[tp][ci],
                                                                           acc[q][g], 4, tp, 0);
                    };
                    TapSteps<0, 9>::run(step);
                }
            }
        }
        {
            const int ho = dh(odims, n), wo = dw(odims, n);
            const int ox = tx0 + lc;
            const long long oplane = (long long)ho * wo;
            float* ys = y + (long long)n * ystride;
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
                        const long long o = (long long)oc * oplane + (long long)oy * wo + ox;
                        ys[o] = conv_finish(acc[q][g][r], b[oc], mode, slope, resid,
                                            (long long)n * rstride + o);
                    }
            }
        }
        if (tn < total) {
            v4f32* D = buf + (cur ^ 1) * ct.E;
#pragma unroll
            for (int k = 0; k < MAXE; ++k) {
                const int e = tid + 256 * k;
                if (e < ct.E) D[e] = c4_masked(pf[k], pm >> (4 * k));
            }
        }
        __syncthreads();
        cur ^= 1;
        t = tn;
        at = an;
    }
}

}  // namespace

namespace sv {

int conv_fwd_c4(const float* x, const int* idims, long long xstride, int IC, const float* w,
                const float* b, int OC, int s, int p, int dil, float* y, const int* odims,
                long long ystride, int N, int max_ho, int max_wo, int mode, float slope,
                const float* resid, long long rstride, hipStream_t st) {
    if (IC > 16 || OC < 2 || OC > 12) return SV_EUNSUPPORTED;
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
    const int P = tune(SV_TUNE_CONV_MGROUPS) == 1 ? 1 : 2;
    const int TT_H = 4 * P * (64 / TWc);
    C4Tiles ct;
    ct.IHt = (TT_H - 1) * s + 2 * dil + 1;
    ct.IWt = (TWc - 1) * s + 2 * dil + 1;
    const int CB = (IC + 3) / 4;
    const long long plane = (long long)ct.IHt * ct.IWt;
    const long long E = CB * plane;
    const long long shm = 32 * E + 4LL * (N + 1);
    // the staging registers hold every position; the index splits by
    // multiply-high are exact below 2^15
    if (E > 256LL * kC4MaxPerThread || shm > 64 * 1024) return SV_EUNSUPPORTED;
    ct.E = (int)E;
    ct.mW = (unsigned)((0xFFFFFFFFull + (unsigned long long)ct.IWt) / (unsigned long long)ct.IWt);
    ct.mP = (unsigned)((0xFFFFFFFFull + (unsigned long long)plane) / (unsigned long long)plane);
    const long long ub = (long long)((max_wo + TWc - 1) / TWc) * ((max_ho + TT_H - 1) / TT_H) * N;
    dim3 g(conv_grid(ub, shm)), bk(256);
    const int OCG = (OC + 3) / 4;
#define LAUNCH_C4(OCG_, TW_, P_, CBM_)                                                           \
    hipLaunchKernelGGL((k_conv_c4<OCG_, TW_, P_, CBM_>), g, bk, (size_t)shm, st, x, idims,      \
                       xstride, IC, w, b, OC, s, p, dil, y, odims, ystride, N, mode, slope,      \
                       resid, rstride, ct)
#define LAUNCH_C4_CB(OCG_, TW_, P_)                                                              \
    do {                                                                                         \
        if (CB == 1) LAUNCH_C4(OCG_, TW_, P_, 1);                                                \
        else if (CB == 2) LAUNCH_C4(OCG_, TW_, P_, 2);                                           \
        else if (CB == 3) LAUNCH_C4(OCG_, TW_, P_, 3);                                           \
        else LAUNCH_C4(OCG_, TW_, P_, 4);                                                        \
    } while (0)
#define LAUNCH_C4_P(OCG_, TW_)                                                                   \
    do {                                                                                         \
        if (P == 1) LAUNCH_C4_CB(OCG_, TW_, 1);                                                  \
        else LAUNCH_C4_CB(OCG_, TW_, 2);                                                         \
    } while (0)
#define LAUNCH_C4_TW(OCG_)                                                                       \
    do {                                                                                         \
        if (TWc == 64) LAUNCH_C4_P(OCG_, 64);                                                    \
        else if (TWc == 32) LAUNCH_C4_P(OCG_, 32);                                               \
        else LAUNCH_C4_P(OCG_, 16);                                                              \
    } while (0)
    if (OCG == 1) LAUNCH_C4_TW(1);
    else if (OCG == 2) LAUNCH_C4_TW(2);
    else LAUNCH_C4_TW(3);
#undef LAUNCH_C4_TW
#undef LAUNCH_C4_P
#undef LAUNCH_C4_CB
#undef LAUNCH_C4
    SV_CHECK_LAUNCH();
    return SV_OK;
}

}  // namespace sv
