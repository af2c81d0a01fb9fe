// 3D convolution of the cost filter (Conv3d k = 3, stride 1, zero padding 1,
// stereonet.py:596-641; the first layer forms the cost volume |fl - fr(x - d)|
// of :771-799 on the fly) on the matrix cores, one column of output depths
// per tile.
//
// Sums: the same single-rounding multiply-adds in the same order as
// k_conv_tile (net.hip) and k_conv_m4 (mfma_conv.hip): per output, input
// channel, then depth tap, row tap, column tap; bias last.  Taps that read the
// zero padding above and below the depth range are skipped: an accumulator
// that starts at +0 never holds -0, so adding a +-0 product leaves it
// unchanged, and the results agree bit for bit with the kernels above.
//
// Why a depth column: a tile of the other kernels covers one output depth and
// reads 27 input taps per channel from LDS, each input plane staged for (and
// read by) three output depths.  Here a block owns a spatial tile for DC
// consecutive output depths and walks the input planes d' = od0-1 .. od0+DC
// once per input channel; each value read from LDS feeds the (up to) three
// output depths od = d'+1, d', d'-1 whose depth tap it is, so one LDS read
// serves 3 x OCG matrix-core steps, and a plane is staged once per tile
// instead of three times.  Visiting planes in increasing d' keeps each
// output's depth taps in increasing order, so the summation order above is
// preserved.
//
// MFMA operand layout (v_mfma_f32_4x4x1_16b_f32, wave of 64, as k_conv_m4):
// lane l <-> output pixel l of the wave's 64-pixel group (TW columns x 64/TW
// rows); B = the input value of lane l's pixel at the current tap; A = the
// weights of 4 output channels at that tap, broadcast from block tap % 16
// (CBSZ = 4), one VGPR holding 16 taps x 4 channels (lane 4t + i holds
// w[oc0 + i][ic][16j + t]); D: VGPR r of lane l is output channel oc0 + r of
// pixel l.  Accumulators: DC depths x OCG channel groups x 4 VGPRs.
//
// Staging, double-buffered over input channels: the global loads of channel
// ic + 1 are issued before channel ic's matrix-core work and written to the
// other LDS buffer after it, one barrier per channel.  For the cost-volume
// layer (SRC == 1) the buffers hold the left features' halo tile and a
// wider right-feature tile covering every shift of the chunk; the cost value
// is formed per tap in registers (the right value's LDS read, |fl - fr|, the
// 2.0 fill left of the shift, zero outside the image).
//
// Work list: tiles (sample, spatial tile, depth chunk), chunk fastest, dealt
// over a grid sized to fill the GPU once (conv_tiles.h).
#include <type_traits>

#include "common.h"
#include "conv_tiles.h"

namespace {

typedef float v4f32 __attribute__((ext_vector_type(4)));

// LDS row pitches: with 16-column tiles a wave's 64 lanes read 4 rows of 16,
// and a half-wave's two rows hit the same LDS banks unless their pitch is 16
// (mod 32) words apart; the cost-volume layer (SRC == 1), whose staging is
// small, pads its rows to 48 words for that.  Other shapes keep packed rows
// (a half-wave reads within one row).
template <int SRC, int TW, int DC>
struct ColGeom {
    static constexpr int GR = 64 / TW;                // rows of a 64-pixel group
    static constexpr int TT_H = 4 * GR;               // tile rows (one group per wave)
    static constexpr int IHt = TT_H + 2, IWt = TW + 2;  // input halo tile of one plane
    static constexpr int PLANE = IHt * IWt;           // staged values of one plane
    static constexpr int FW = IWt + DC + 1;           // right-feature tile columns (SRC == 1)
    static constexpr bool PAD = SRC == 1 && TW == 16;
    static constexpr int IP = PAD ? 48 : IWt;         // LDS row pitch of a plane
    static constexpr int FP = PAD ? 48 : FW;          // LDS row pitch of the right features
    static constexpr int PLANEP = IHt * IP;           // LDS words of a plane
    static constexpr int NE = SRC == 1 ? PLANE + IHt * FW : (DC + 2) * PLANE;  // staged values
    static constexpr int NL = SRC == 1 ? PLANEP + IHt * FP : (DC + 2) * PLANEP;  // LDS words
    static constexpr int KE = (NE + 255) / 256;       // per thread
    static constexpr int BUF = (NL + 3) / 4 * 4;      // one LDS buffer (floats)
    static_assert(!PAD || (IWt <= IP && FW <= FP), "padded pitches hold the rows");
    // LDS word of staged value idx
    __device__ __forceinline__ static int at(int idx) {
        if constexpr (!PAD) {
            return idx;
        } else if (SRC == 1) {
            if (idx < PLANE) {
                const int yy = idx / IWt;
                return yy * IP + (idx - yy * IWt);
            }
            const int j = idx - PLANE, yy = j / FW;
            return PLANEP + yy * FP + (j - yy * FW);
        } else {
            const int e = idx / PLANE, pp = idx - e * PLANE, yy = pp / IWt;
            return e * PLANEP + yy * IP + (pp - yy * IWt);
        }
    }
};

// Global loads of input channel ic for this tile into registers (zero
// outside the image and outside the depth range).
template <int SRC, int TW, int DC>
__device__ __forceinline__ void col_load(float* v, const float* __restrict__ xs,
                                         const float* __restrict__ xrs, int ic, int D, int od0,
                                         int iy0, int ix0, int hin, int win, int tid) {
    using G = ColGeom<SRC, TW, DC>;
    const int iplane = hin * win;
#pragma unroll
    for (int k = 0; k < G::KE; ++k) {
        const int idx = tid + 256 * k;
        float val = 0.f;
        if (idx < G::NE) {
            if (SRC == 1) {
                if (idx < G::PLANE) {
                    const int yy = idx / G::IWt, xx = idx - yy * G::IWt;
                    const int iy = iy0 + yy, ix = ix0 + xx;
                    if (iy >= 0 && iy < hin && ix >= 0 && ix < win)
                        val = xs[(long long)ic * iplane + iy * win + ix];
                } else {
                    const int j = idx - G::PLANE;
                    const int yy = j / G::FW, xf = j - yy * G::FW;
                    const int iy = iy0 + yy, ix = ix0 - od0 - DC + xf;
                    if (iy >= 0 && iy < hin && ix >= 0 && ix < win)
                        val = xrs[(long long)ic * iplane + iy * win + ix];
                }
            } else {
                const int e = idx / G::PLANE, pp = idx - e * G::PLANE;
                const int yy = pp / G::IWt, xx = pp - yy * G::IWt;
                const int dd = od0 - 1 + e, iy = iy0 + yy, ix = ix0 + xx;
                if (dd >= 0 && dd < D && iy >= 0 && iy < hin && ix >= 0 && ix < win)
                    val = xs[((long long)ic * D + dd) * iplane + iy * win + ix];
            }
        }
        v[k] = val;
    }
}

template <int SRC, int TW, int DC>
__device__ __forceinline__ void col_store(float* buf, const float* v, int tid) {
    using G = ColGeom<SRC, TW, DC>;
#pragma unroll
    for (int k = 0; k < G::KE; ++k) {
        const int idx = tid + 256 * k;
        if (idx < G::NE) buf[G::at(idx)] = v[k];
    }
}

template <int E, int N>
struct Unroll {
    template <class F>
    __device__ __forceinline__ static void run(F& f) {
        f(std::integral_constant<int, E>{});
        Unroll<E + 1, N>::run(f);
    }
};
template <int N>
struct Unroll<N, N> {
    template <class F>
    __device__ __forceinline__ static void run(F&) {}
};

template <int SRC, int OCG, int TW, int DC>
__global__ void __launch_bounds__(256) k_conv3_col(
    const float* __restrict__ x, const float* __restrict__ xr, const int* __restrict__ idims,
    long long xstride, int IC, int D, const float* __restrict__ w, const float* __restrict__ b,
    int OC, float* __restrict__ y, const int* __restrict__ odims, long long ystride, int N,
    int mode, float slope, const float* __restrict__ resid, long long rstride) {
    using G = ColGeom<SRC, TW, DC>;
    constexpr int NJ = 2;  // weight VGPRs per (channel, 4 output channels): taps 0-15, 16-26
    extern __shared__ __align__(16) float smem_f[];
    float* bufs = smem_f;                              // [2][BUF]
    float* wl = smem_f + 2 * G::BUF;                   // [IC][OCG][NJ][64]
    int* pre = reinterpret_cast<int*>(wl + IC * OCG * NJ * 64);  // [N + 1]
    __shared__ int wsum[4];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int lc = lane % TW, lr = lane / TW;
    const int yy0 = wave * G::GR + lr;  // this lane's row in the tile
    for (int e = tid; e < IC * OCG * NJ * 64; e += 256) {
        const int l = e & 63;
        int q = e >> 6;
        const int j = q % NJ;
        q /= NJ;
        const int g = q % OCG, c = q / OCG;
        const int tp = 16 * j + (l >> 2), oc = 4 * g + (l & 3);
        wl[e] = (tp < 27 && oc < OC) ? w[((long long)oc * IC + c) * 27 + tp] : 0.f;
    }
    const int nch = (D + DC - 1) / DC;
    const int total = conv_tile_prefix<TW, G::TT_H>(odims, N, nch, pre, wsum, tid);
    for (int t = conv_first_tile(); t < total; t += gridDim.x) {
        const TileAt at = conv_tile_at<TW, G::TT_H>(t, pre, N, nch, odims);
        const int n = at.n, ty0 = at.ty0, tx0 = at.tx0;
        const int od0 = at.od * DC;
        const int hin = dh(idims, n), win = dw(idims, n);
        const int ho = dh(odims, n), wo = dw(odims, n);
        const int iy0 = ty0 - 1, ix0 = tx0 - 1;
        const float* xs = x + (long long)n * xstride;
        const float* xrs = SRC == 1 ? xr + (long long)n * xstride : nullptr;
        const int oy = ty0 + yy0, ox = tx0 + lc;
        // per column/row tap: inside the image, and (cost volume) the column
        // relative to the chunk, x' - (od0 - 1): the 2.0 fill applies when
        // it is below the plane index e
        bool inb[9];
        int xq[9];
#pragma unroll
        for (int tp = 0; tp < 9; ++tp) {
            const int yi = oy + tp / 3 - 1, xi = ox + tp % 3 - 1;
            inb[tp] = yi >= 0 && yi < hin && xi >= 0 && xi < win;
            xq[tp] = xi - (od0 - 1);
        }
        v4f32 acc[DC][OCG];
#pragma unroll
        for (int e = 0; e < DC; ++e)
#pragma unroll
            for (int g = 0; g < OCG; ++g) acc[e][g] = (v4f32){0.f, 0.f, 0.f, 0.f};
        float pf[G::KE];
        col_load<SRC, TW, DC>(pf, xs, xrs, 0, D, od0, iy0, ix0, hin, win, tid);
        col_store<SRC, TW, DC>(bufs, pf, tid);
        __syncthreads();
        for (int ic = 0; ic < IC; ++ic) {
            const float* cb = bufs + (ic & 1) * G::BUF;
            if (ic + 1 < IC) col_load<SRC, TW, DC>(pf, xs, xrs, ic + 1, D, od0, iy0, ix0, hin, win, tid);
            float wv[OCG][NJ];
#pragma unroll
            for (int g = 0; g < OCG; ++g)
#pragma unroll
                for (int j = 0; j < NJ; ++j) wv[g][j] = wl[((ic * OCG + g) * NJ + j) * 64 + lane];
            float flv[9];
            if (SRC == 1) {
#pragma unroll
                for (int tp = 0; tp < 9; ++tp) flv[tp] = cb[(yy0 + tp / 3) * G::IP + lc + tp % 3];
            }
            // input planes d' = od0 - 1 + e, e = 0 .. DC + 1
            auto plane = [&](auto ec) {
                constexpr int e = decltype(ec)::value;
                const int dd = od0 - 1 + e;
                if (dd < 0 || dd >= D) return;  // zero padding of the depth range
                auto tap = [&](auto tc) {
                    constexpr int t9 = decltype(tc)::value;
                    constexpr int ky = t9 / 3, kx = t9 % 3;
                    float v;
                    if (SRC == 1) {
                        const float fr = cb[G::PLANEP + (yy0 + ky) * G::FP + lc + kx + DC + 1 - e];
                        v = fabsf(flv[t9] - fr);
                        v = xq[t9] < e ? 2.0f : v;
                        v = inb[t9] ? v : 0.f;
                    } else {
                        v = cb[e * G::PLANEP + (yy0 + ky) * G::IP + lc + kx];
                    }
                    // output depths od0 + o with depth tap kd = e - o
                    auto out = [&](auto oc_) {
                        constexpr int o = decltype(oc_)::value;
                        constexpr int kd = e - o;
                        if constexpr (o >= 0 && o < DC && kd >= 0 && kd <= 2) {
                            constexpr int tp = kd * 9 + t9;
#pragma unroll
                            for (int g = 0; g < OCG; ++g)
                                acc[o][g] = __builtin_amdgcn_mfma_f32_4x4x1f32(wv[g][tp / 16], v, acc[o][g],
                                                                               4, tp % 16, 0);
                        }
                    };
                    // increasing o = decreasing kd; each output gets one tap
                    // per plane, so the order across outputs does not matter
                    out(std::integral_constant<int, e - 2>{});
                    out(std::integral_constant<int, e - 1>{});
                    out(std::integral_constant<int, e>{});
                };
                Unroll<0, 9>::run(tap);
            };
            Unroll<0, DC + 2>::run(plane);
            if (ic + 1 < IC) col_store<SRC, TW, DC>(bufs + ((ic + 1) & 1) * G::BUF, pf, tid);
            __syncthreads();
        }
        if (ox < wo && oy < ho) {
            const long long oplane = (long long)ho * wo;
            float* ys = y + (long long)n * ystride;
#pragma unroll
            for (int o = 0; o < DC; ++o) {
                const int od = od0 + o;
                if (od >= D) break;
#pragma unroll
                for (int g = 0; g < OCG; ++g)
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int oc = 4 * g + r;
                        if (oc >= OC) break;
                        const long long oi = (long long)oc * D * oplane + (long long)od * oplane +
                                             (long long)oy * wo + ox;
                        ys[oi] = conv_finish(acc[o][g][r], b[oc], mode, slope, resid,
                                             (long long)n * rstride + oi);
                    }
            }
        }
    }
}

template <int SRC, int OCG, int TW, int DC>
int launch_col(const float* x, const float* xr, const int* idims, long long xstride, int IC, int D,
               const float* w, const float* b, int OC, float* y, const int* odims,
               long long ystride, int N, int max_ho, int max_wo, int mode, float slope,
               const float* resid, long long rstride, hipStream_t st) {
    using G = ColGeom<SRC, TW, DC>;
    const long long shm = 4LL * (2 * G::BUF + (long long)IC * OCG * 2 * 64) + 4LL * (N + 1);
    if (shm > 64 * 1024) return SV_EUNSUPPORTED;
    const int nch = (D + DC - 1) / DC;
    const long long tiles = (long long)((max_wo + TW - 1) / TW) * ((max_ho + G::TT_H - 1) / G::TT_H) *
                            nch * N;
    // resident blocks per CU: LDS, and about 3 waves per SIMD of registers
    const long long occ = std::max<long long>(1, std::min<long long>(3, 160 * 1024 / shm));
    long long grid = std::min<long long>(tiles, occ * sv::cu_count());
    if (grid >= 8) grid -= grid % 8;
    grid = std::max<long long>(grid, 1);
    hipLaunchKernelGGL((k_conv3_col<SRC, OCG, TW, DC>), dim3((unsigned)grid), dim3(256), (size_t)shm,
                       st, x, xr, idims, xstride, IC, D, w, b, OC, y, odims, ystride, N, mode, slope,
                       resid, rstride);
    SV_CHECK_LAUNCH();
    return SV_OK;
}

template <int SRC, int OCG, int TW>
int launch_col_dc(int DC, const float* x, const float* xr, const int* idims, long long xstride,
                  int IC, int D, const float* w, const float* b, int OC, float* y,
                  const int* odims, long long ystride, int N, int max_ho, int max_wo, int mode,
                  float slope, const float* resid, long long rstride, hipStream_t st) {
    // accumulators DC x OCG x 4 VGPRs: at most 96
    if (DC == 12 && OCG <= 2)
        return launch_col<SRC, OCG, TW, 12>(x, xr, idims, xstride, IC, D, w, b, OC, y, odims,
                                            ystride, N, max_ho, max_wo, mode, slope, resid,
                                            rstride, st);
    if (DC >= 6 && OCG <= 4)
        return launch_col<SRC, OCG, TW, 6>(x, xr, idims, xstride, IC, D, w, b, OC, y, odims,
                                           ystride, N, max_ho, max_wo, mode, slope, resid, rstride,
                                           st);
    return launch_col<SRC, OCG, TW, 4>(x, xr, idims, xstride, IC, D, w, b, OC, y, odims, ystride,
                                       N, max_ho, max_wo, mode, slope, resid, rstride, st);
}

}  // namespace

namespace sv {

int conv_fwd_col(const float* x, const float* xr, const int* idims, long long xstride, int IC,
                 int D, int KD, const float* w, const float* b, int OC, int s, int p, int dil,
                 float* y, const int* odims, long long ystride, int N, int max_ho, int max_wo,
                 int mode, float slope, const float* resid, long long rstride, hipStream_t st) {
    if (KD != 3 || s != 1 || p != 1 || dil != 1 || OC > 16 || D < 1) return SV_EUNSUPPORTED;
    const int src = xr ? 1 : 0;
    const int OCG = (OC + 3) / 4;
    // output columns per tile: the widest of 64, 32, 16 that wastes at most
    // 1/8 of a row (or the knob's)
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
    // depth chunk: the knob's, else 6 when that leaves about two tiles per
    // CU, 4 otherwise (measured on the 12-depth cost filter: 6 beats 12,
    // whose 96 accumulator registers halve the resident waves, and 4 only
    // wins where 6 leaves CUs idle)
    int DC = tune(SV_TUNE_CONV_DCHUNK);
    if (!DC) {
        const int TT_H = 4 * (64 / TWc);
        const long long sp = (long long)((max_wo + TWc - 1) / TWc) * ((max_ho + TT_H - 1) / TT_H) * N;
        DC = sp * ((D + 5) / 6) >= 2LL * cu_count() ? 6 : 4;
    }
#define COL_ARGS x, xr, idims, xstride, IC, D, w, b, OC, y, odims, ystride, N, max_ho, max_wo, mode, \
                 slope, resid, rstride, st
#define COL_TW(SRC_, OCG_)                                                                      \
    do {                                                                                        \
        if (TWc == 64) return launch_col_dc<SRC_, OCG_, 64>(DC, COL_ARGS);                      \
        if (TWc == 32) return launch_col_dc<SRC_, OCG_, 32>(DC, COL_ARGS);                      \
        return launch_col_dc<SRC_, OCG_, 16>(DC, COL_ARGS);                                     \
    } while (0)
#define COL_OC(SRC_)                                                                            \
    do {                                                                                        \
        if (OCG == 1) COL_TW(SRC_, 1);                                                          \
        if (OCG == 2) COL_TW(SRC_, 2);                                                          \
        if (OCG == 3) COL_TW(SRC_, 3);                                                          \
        COL_TW(SRC_, 4);                                                                        \
    } while (0)
    if (src) COL_OC(1);
    COL_OC(0);
#undef COL_OC
#undef COL_TW
#undef COL_ARGS
}

}  // namespace sv
