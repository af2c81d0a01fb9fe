// Tile lists and input staging shared by the two tiled forward convolutions
// (the vector kernel in net.hip and the matrix-core kernel in
// mfma_conv.hip).
//
// Work list: the output of sample n (its own h x w, every output depth) is
// cut into tiles of TW columns x TT_H rows; tiles are numbered sample after
// sample, depth fastest (neighbouring tiles then read overlapping input).
// Each block scans the per-sample tile counts into LDS once and walks tile
// numbers bslot, bslot + gridDim.x, ...; the grid is sized to fill the GPU
// once, so samples much smaller than the largest cost nothing.  Tile numbers
// are dealt so that each XCD (blocks go round-robin over the 8 XCDs) works
// through a contiguous range and shares its L2.
//
// Staging: the input halo tiles of `icn` channels (every depth tap, zero
// outside the image; the cost volume |fl(x) - fr(x - d)|, 2.0 where x < d,
// formed on the fly when SRC == 1) go to LDS in one pass with several
// independent loads per thread in flight.  Position -> (row, column) splits
// use multiply-high by a precomputed reciprocal, exact for numbers and
// divisors below 2^16.
#pragma once

#include <algorithm>
#include <type_traits>

#include "common.h"

namespace sv {

struct ConvTiles {
    int IHt, IWt;     // input tile (rows, columns) of one depth tap
    int ICG;          // input channels staged per pass
    unsigned mP, mW;  // ceil(2^32 / (IHt*IWt)), ceil(2^32 / IWt)
};

// Tile geometry for an output tile of TW x TT_H at stride s, dilation dil:
// fills ct (ICG the most channels whose input tiles plus per_ch_extra floats
// each fit `budget` bytes next to the (N+1)-entry prefix table) and returns
// the dynamic LDS bytes, or -1 when one channel does not fit 64 KB or the
// reciprocal split would not be exact.
inline long long conv_tiles(int KD, int IC, int TW, int TT_H, int s, int dil, int N,
                            long long per_ch_extra, long long budget, ConvTiles& ct) {
    const int IHt = (TT_H - 1) * s + 2 * dil + 1, IWt = (TW - 1) * s + 2 * dil + 1;
    const long long plane = (long long)IHt * IWt;
    const long long tsz = KD * plane;
    const long long per_ch = tsz + per_ch_extra;
    const long long pre_bytes = 4LL * (N + 1);
    long long icg = (budget - pre_bytes) / (4 * per_ch);
    icg = std::min<long long>(std::max<long long>(icg, 1), IC);
    icg = std::min<long long>(icg, 65536 / std::max<long long>(tsz, 1));
    const long long shm = 4 * icg * per_ch + pre_bytes;
    if (icg < 1 || plane >= 65536 || shm > 64 * 1024) return -1;
    ct.IHt = IHt;
    ct.IWt = IWt;
    ct.ICG = (int)icg;
    ct.mP = (unsigned)((0xFFFFFFFFull + (unsigned long long)plane) / (unsigned long long)plane);
    ct.mW = (unsigned)((0xFFFFFFFFull + (unsigned long long)IWt) / (unsigned long long)IWt);
    return shm;
}

// Blocks of a tile-list launch: enough to fill every CU once at the
// occupancy the LDS footprint allows, at most one per tile, a multiple of 8.
inline unsigned conv_grid(long long tiles_bound, long long shm) {
    const long long occ = std::max<long long>(1, std::min<long long>(8, 160 * 1024 / std::max<long long>(shm, 1)));
    long long G = std::min<long long>(tiles_bound, occ * cu_count());
    if (G >= 8) G -= G % 8;
    return (unsigned)std::max<long long>(G, 1);
}

// 2D matrix-core convolution over 4-channel blocks (conv_c4.hip); returns
// SV_EUNSUPPORTED when the shape does not fit it
int conv_fwd_c4(const float* x, const int* idims, long long xstride, int IC, const float* w,
                const float* b, int OC, int s, int p, int dil, float* y, const int* odims,
                long long ystride, int N, int max_ho, int max_wo, int mode, float slope,
                const float* resid, long long rstride, hipStream_t st);

// matrix-core variant of the tiled forward convolution (mfma_conv.hip);
// returns SV_EUNSUPPORTED when the shape does not fit it
int conv_fwd_m4(const float* x, const float* xr, const int* idims, long long xstride, int IC,
                int D, int KD, const float* w, const float* b, int OC, int s, int p, int dil,
                float* y, const int* odims, long long ystride, int N, int max_ho, int max_wo,
                int mode, float slope, const float* resid, long long rstride, hipStream_t st);

// matrix-core 3D convolution over a column of output depths per tile
// (conv3d_col.hip); returns SV_EUNSUPPORTED when the shape does not fit it
int conv_fwd_col(const float* x, const float* xr, const int* idims, long long xstride, int IC,
                 int D, int KD, const float* w, const float* b, int OC, int s, int p, int dil,
                 float* y, const int* odims, long long ystride, int N, int max_ho, int max_wo,
                 int mode, float slope, const float* resid, long long rstride, hipStream_t st);

// direct convolution, one output position per thread (conv_direct.hip;
// xr: the right features of a first 3D layer that forms the cost volume on
// the fly, else null); returns SV_EUNSUPPORTED when the shape does not fit
int conv_fwd_direct(const float* x, const float* xr, const int* idims, long long xstride, int IC,
                    int D, int KD, const float* w, const float* b, int OC, int s, int p, int dil,
                    float* y, const int* odims, long long ystride, int N, int max_ho, int max_wo,
                    int mode, float slope, const float* resid, long long rstride, hipStream_t st);

}  // namespace sv

namespace {

__device__ __forceinline__ int dh(const int* d, int n) { return d[2 * n]; }
__device__ __forceinline__ int dw(const int* d, int n) { return d[2 * n + 1]; }

__device__ __forceinline__ int div_magic(int e, unsigned m) { return (int)__umulhi((unsigned)e, m); }

// pre[0..N] = prefix sums of the per-sample tile counts; returns pre[N].
// Called by all 256 threads of the block.
template <int TW, int TT_H>
__device__ __forceinline__ int conv_tile_prefix(const int* __restrict__ odims, int N, int D,
                                                int* pre, int* wsum, int tid) {
    const int lane = tid & 63, wave = tid >> 6;
    if (N <= 64) {
        // one wave scans, one barrier
        if (wave == 0) {
            int c = 0;
            if (lane < N) {
                const int ho = dh(odims, lane), wo = dw(odims, lane);
                if (ho > 0 && wo > 0) c = (wo + TW - 1) / TW * ((ho + TT_H - 1) / TT_H) * D;
            }
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const int t = __shfl_up(c, o, 64);
                if (lane >= o) c += t;
            }
            if (lane < N) pre[lane + 1] = c;
            if (lane == 0) pre[0] = 0;
        }
        __syncthreads();
        return pre[N];
    }
    int carry = 0;
    for (int base = 0; base < N; base += 256) {
        const int i = base + tid;
        int c = 0;
        if (i < N) {
            const int ho = dh(odims, i), wo = dw(odims, i);
            if (ho > 0 && wo > 0) c = (wo + TW - 1) / TW * ((ho + TT_H - 1) / TT_H) * D;
        }
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int t = __shfl_up(c, o, 64);
            if (lane >= o) c += t;
        }
        if (lane == 63) wsum[wave] = c;
        __syncthreads();
        int add = carry;
        for (int q = 0; q < wave; ++q) add += wsum[q];
        if (i < N) pre[i + 1] = c + add;
        carry += wsum[0] + wsum[1] + wsum[2] + wsum[3];
        __syncthreads();
    }
    if (tid == 0) pre[0] = 0;
    __syncthreads();
    return pre[N];
}

// this block's first tile number (XCD-contiguous dealing)
__device__ __forceinline__ int conv_first_tile() {
    const int G = gridDim.x;
    return (G & 7) == 0 ? (blockIdx.x & 7) * (G >> 3) + (blockIdx.x >> 3) : blockIdx.x;
}

struct TileAt {
    int n, od, ty0, tx0;
};

// tile t -> sample n (the largest n with pre[n] <= t), depth and corner
template <int TW, int TT_H>
__device__ __forceinline__ TileAt conv_tile_at(int t, const int* pre, int N, int D,
                                               const int* __restrict__ odims) {
    int lo = 0, hi = N;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (pre[mid] <= t) lo = mid;
        else hi = mid;
    }
    TileAt a;
    a.n = lo;
    const int twn = (dw(odims, lo) + TW - 1) / TW;
    int local = t - pre[lo];
    a.od = local % D;
    local /= D;
    a.ty0 = (local / twn) * TT_H;
    a.tx0 = (local % twn) * TW;
    return a;
}

// Stage channels ic0 .. ic0+icn-1 of the input halo tile at (od, iy0, ix0)
// into tile[c][kd][yy][xx].  Each thread owns positions pp = tid, tid + 256,
// ... of the IHt x IWt plane (one row/column split per position) and walks
// the channels there, U channels x KD depth taps of loads in flight before
// their LDS stores; an element costs an address increment.  All 256 threads
// call it.
template <int KD, int SRC>
__device__ __forceinline__ void conv_stage_inputs(float* tile, const float* __restrict__ xs,
                                                  const float* __restrict__ xrs, int ic0, int icn,
                                                  int D, int od, int iy0, int ix0, int hin,
                                                  int win, const sv::ConvTiles& ct, int tid) {
    constexpr int U = KD == 1 ? 8 : 4;
    constexpr int PD = KD / 2;
    const int IWt = ct.IWt;
    const int plane_t = ct.IHt * IWt;
    const int iplane = hin * win;
    const int cstride = SRC == 1 ? iplane : D * iplane;  // next channel, same depth
    for (int pp = tid; pp < plane_t; pp += 256) {
        const int yy = div_magic(pp, ct.mW);
        const int xx = pp - yy * IWt;
        const int iy = iy0 + yy, ix = ix0 + xx;
        const bool inb = iy >= 0 && iy < hin && ix >= 0 && ix < win;
        const int roff = iy * win + ix;
        bool fill2[KD], ld[KD];
        int g0[KD];
#pragma unroll
        for (int kd = 0; kd < KD; ++kd) {
            const int dd = od + kd - PD;
            const bool dok = inb && dd >= 0 && dd < D;
            // cost volume: 2.0 left of the shift, else |fl - fr(x - dd)|
            fill2[kd] = SRC == 1 && dok && ix < dd;
            ld[kd] = dok && !fill2[kd];
            g0[kd] = SRC == 1 ? ic0 * iplane + roff : (ic0 * D + dd) * iplane + roff;
        }
        float* dst = tile + pp;
#pragma unroll 1
        for (int c0 = 0; c0 < icn; c0 += U) {
            float v[U][KD];
#pragma unroll
            for (int u = 0; u < U; ++u)
#pragma unroll
                for (int kd = 0; kd < KD; ++kd) {
                    const int g = g0[kd] + (c0 + u) * cstride;
                    v[u][kd] = fill2[kd] ? 2.0f : 0.f;
                    if (ld[kd] && c0 + u < icn) {
                        if (SRC == 1) v[u][kd] = fabsf(xs[g] - xrs[g - (od + kd - PD)]);
                        else v[u][kd] = xs[g];
                    }
                }
#pragma unroll
            for (int u = 0; u < U; ++u)
#pragma unroll
                for (int kd = 0; kd < KD; ++kd)
                    if (c0 + u < icn) dst[((c0 + u) * KD + kd) * plane_t] = v[u][kd];
        }
    }
}

// Compile-time loop over taps TP .. TAPS-1: f(std::integral_constant<int, tp>)
// (MFMA operand-select immediates need the tap as a constant).
template <int TP, int TAPS>
struct TapSteps {
    template <class F>
    __device__ __forceinline__ static void run(F& f) {
        f(std::integral_constant<int, TP>{});
        TapSteps<TP + 1, TAPS>::run(f);
    }
};
template <int TAPS>
struct TapSteps<TAPS, TAPS> {
    template <class F>
    __device__ __forceinline__ static void run(F&) {}
};

// conv epilogue: bias, then LeakyReLU (mode 1) or max(resid + v, 0) (mode 2)
__device__ __forceinline__ float conv_finish(float acc, float bias, int mode, float slope,
                                             const float* __restrict__ resid, long long ridx) {
    float v = acc + bias;
    if (mode == 1) {
        v = v > 0.f ? v : v * slope;
    } else if (mode == 2) {
        v = resid[ridx] + v;
        v = v > 0.f ? v : 0.f;
    }
    return v;
}

}  // namespace
