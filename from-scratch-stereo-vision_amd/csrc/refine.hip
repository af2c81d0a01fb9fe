// Full-resolution census + band semi-global matching + left-right check for
// gfx950 (MI355X).  Restates stereonet.py:5115-5391 and :2231-2258
// (Ray-5106/From-Scratch-Stereo-Vision) as batched GPU kernels.
//
// Data layout in HBM (B pairs of W x H, V views, K = 2R+1 band entries,
// Kp = G*KC padded band length):
//   images      [B][H][W]            u8 / f32 / f64 (row-major, as read)
//   census      [2][B][H][W]         u64, MSB = first window offset
//   base        [V][B][H][W]         i32 band centre max(0, int(d0 + 0.5))
//   costs       [V][B][H][W][Kp]     u8 (integer path) or f64 (general path)
//   path costs  [dir][V][B][H][W][Kp] u8/u16 (integer) or f64 (general)
//   disparity   [V][B][H][W]         f64 (parity with the float64 reference)
//   status      [B][H][W]            u8 in {0, 1, 2}
//
// The right view of lr_refine is matched in the left image's coordinates:
// census words of a mirrored image are a fixed bit permutation of the
// unmirrored words, and Hamming distance is invariant under that
// permutation, so costs, path sums and winners equal the reference's
// mirrored computation without building mirrored copies.
//
// Two arithmetic paths: the integer path (ad_weight = prior_w = 0 and
// integral, non-negative P1/P2: every cost is a popcount, every path cost an
// integer, so results are bit-exact), and the float64 general path for
// everything else, which evaluates the reference's expressions in the same
// order so results are bit-exact as well.  This file is compiled with
// -ffp-contract=off: no multiply-add is ever fused.
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>

#include "common.h"

namespace {

__constant__ int c_dx[8] = {1, -1, 0, 0, 1, -1, 1, -1};
__constant__ int c_dy[8] = {0, 0, 1, -1, 1, 1, -1, -1};
// summation order of the eight paths: the reference order for the left view;
// for the right view the order the mirrored computation visits them in
// (dx negated: 0<->1, 4<->5, 6<->7)
__constant__ int c_order[2][8] = {{0, 1, 2, 3, 4, 5, 6, 7}, {1, 0, 2, 3, 5, 4, 7, 6}};

// Geometry of a batch: image i (views of pair i % B) is h x w with rows
// packed at its own width, and starts i * slot pixels into every per-pixel
// buffer.  dims == nullptr: every image is W x H (slot = W * H); otherwise
// dims holds (h, w) per pair on the device and W, H bound every pair.
struct Geom {
    const int* dims;
    int W, H, B;
    long long slot;
    __device__ __forceinline__ int w(int i) const { return dims ? dims[2 * (i % B) + 1] : W; }
    __device__ __forceinline__ int h(int i) const { return dims ? dims[2 * (i % B)] : H; }
};

template <typename T>
__device__ __forceinline__ double to_gray(T v);
template <>
__device__ __forceinline__ double to_gray<uint8_t>(uint8_t v) {
    return (double)v * (1.0 / 255.0);
}
template <>
__device__ __forceinline__ double to_gray<float>(float v) {
    return (double)v;
}
template <>
__device__ __forceinline__ double to_gray<double>(double v) {
    return v;
}

template <typename A>
__device__ __forceinline__ A amin(A a, A b) {
    return b < a ? b : a;
}

template <typename A>
__device__ __forceinline__ A big();
template <>
__device__ __forceinline__ int big<int>() {
    return 0x3FFFFFFF;
}
template <>
__device__ __forceinline__ double big<double>() {
    return HUGE_VAL;
}

// two u16 lanes in one register (packed integer ops: v_pk_*_u16)
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
// LDS slots written as u32 pairs and read as single u16 entries: both
// pointer types must be allowed to alias
typedef uint32_t __attribute__((may_alias)) u32_alias;
typedef uint16_t __attribute__((may_alias)) u16_alias;

__device__ __forceinline__ u16x2 as_u16x2(uint32_t v) { return __builtin_bit_cast(u16x2, v); }
__device__ __forceinline__ uint32_t as_u32(u16x2 v) { return __builtin_bit_cast(uint32_t, v); }

// ------------------------------------------------------------------ census
// census, stereonet.py:5115-5142.  A block makes 64 x CR words (CR = 16 rows;
// each of its 64 x 4 threads makes the rows threadIdx.y + 4r): it stages the
// window-padded tile of its pixels in LDS with border-clamped coordinates
// (the reference clamps every read the same way), then each thread compares
// its windows from LDS.  Bit order: MSB = first offset, dy outer, dx inner.
// RX / RY >= 0 fix the window at compile time; -1 takes the runtime radii.
constexpr int CENSUS_ROWS = 16;

template <typename T, int RX, int RY>
__global__ void __launch_bounds__(256) k_census(const T* __restrict__ img, Geom g, int rx_,
                                                int ry_, uint64_t* __restrict__ out) {
    extern __shared__ __align__(16) unsigned char smem_c[];
    T* tile = reinterpret_cast<T*>(smem_c);
    const int rx = RX >= 0 ? RX : rx_, ry = RY >= 0 ? RY : ry_;
    const int TW = 64 + 2 * rx, TH = CENSUS_ROWS + 2 * ry;
    const int x0 = blockIdx.x * 64, y0 = blockIdx.y * CENSUS_ROWS;
    const int W = g.w(blockIdx.z), H = g.h(blockIdx.z);
    if (x0 >= W || y0 >= H) return;  // whole block outside this image
    const long long plane = g.slot;
    const T* im = img + (long long)blockIdx.z * plane;
    const int tid = threadIdx.y * 64 + threadIdx.x;
    if constexpr (RX >= 0 && RY >= 0) {
        // fixed tile: every load of the thread in flight before its stores
        constexpr int TWc = 64 + 2 * RX, NE = TWc * (CENSUS_ROWS + 2 * RY);
        constexpr int PER = (NE + 255) / 256;
        T v[PER];
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int e = min(tid + 256 * u, NE - 1);
            const int ty = e / TWc, tx = e - ty * TWc;
            const int yy = min(max(y0 - RY + ty, 0), H - 1);
            const int xx = min(max(x0 - RX + tx, 0), W - 1);
            v[u] = im[(long long)yy * W + xx];
        }
#pragma unroll
        for (int u = 0; u < PER; ++u)
            if (tid + 256 * u < NE) tile[tid + 256 * u] = v[u];
    } else {
        for (int e = tid; e < TW * TH; e += 256) {
            const int ty = e / TW, tx = e - ty * TW;
            const int yy = min(max(y0 - ry + ty, 0), H - 1);
            const int xx = min(max(x0 - rx + tx, 0), W - 1);
            tile[e] = im[(long long)yy * W + xx];
        }
    }
    __syncthreads();
    const int x = x0 + threadIdx.x;
    if (x >= W) return;
#pragma unroll 1
    for (int r = 0; r < CENSUS_ROWS / 4; ++r) {
        const int ly = threadIdx.y + 4 * r;
        const int y = y0 + ly;
        if (y >= H) break;
        const T* t0 = tile + ly * TW + threadIdx.x;  // window corner (-rx, -ry)
        const T c = t0[ry * TW + rx];
        uint64_t s;
        if constexpr (RX >= 0 && RY >= 0) {
            constexpr int NBITS = (2 * RX + 1) * (2 * RY + 1) - 1;
            uint32_t hi = 0, lo = 0;
            int i = 0;
#pragma unroll
            for (int dy = 0; dy <= 2 * RY; ++dy)
#pragma unroll
                for (int dx = 0; dx <= 2 * RX; ++dx) {
                    if (dy == RY && dx == RX) continue;
                    const int pos = NBITS - 1 - i;  // bit of this offset
                    const uint32_t bit = t0[dy * (64 + 2 * RX) + dx] < c ? 1u : 0u;
                    if (pos >= 32)
                        hi |= bit << (pos - 32);
                    else
                        lo |= bit << pos;
                    ++i;
                }
            s = ((uint64_t)hi << 32) | lo;
        } else {
            s = 0;
            for (int dy = 0; dy <= 2 * ry; ++dy)
                for (int dx = 0; dx <= 2 * rx; ++dx) {
                    if (dy == ry && dx == rx) continue;
                    s = (s << 1) | (uint64_t)(t0[dy * TW + dx] < c);
                }
        }
        out[(long long)blockIdx.z * plane + (long long)y * W + x] = s;
    }
}

// census with the 9 x 7 window on 8-bit images (the refinement default).
// Each thread makes 4 neighbouring pixels of C8_RPT rows.  The LDS tile
// holds the pixels widened to u16, so a thread's 12-pixel run of a window
// row is three aligned 8-byte words, and one packed u16 subtract serves two
// pixels: (n - c) mod 2^16 has bit 15 set exactly when n < c (both in
// 0..255).  That bit is shifted into a per-pixel u16 accumulator, 16
// offsets per accumulator; the four accumulators of a pixel hold its bits
// 61..46, 45..30, 29..14 and 13..0 (first offset most significant, dy
// outer, dx inner, as k_census).  Reads are border-clamped like k_census.
constexpr int C8_RPT = 2;
constexpr int C8_ROWS = 16 * C8_RPT;

__global__ void __launch_bounds__(256) k_census8(const uint8_t* __restrict__ img, Geom g,
                                                 uint64_t* __restrict__ out) {
    constexpr int RX = 4, RY = 3;
    constexpr int TW = 64 + 2 * RX, TH = C8_ROWS + 2 * RY, NE = TW * TH;
    constexpr int WR = C8_RPT + 2 * RY;  // window rows of a thread
    __shared__ __align__(16) uint16_t tile[NE];
    const int x0 = blockIdx.x * 64, y0 = blockIdx.y * C8_ROWS;
    const int W = g.w(blockIdx.z), H = g.h(blockIdx.z);
    if (x0 >= W || y0 >= H) return;  // whole block outside this image
    const long long plane = g.slot;
    const uint8_t* im = img + (long long)blockIdx.z * plane;
    const int tid = threadIdx.x;
    {
        constexpr int PER = (NE + 255) / 256;
        uint32_t v[PER];
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int e = min(tid + 256 * u, NE - 1);
            const int ty = e / TW, tx = e - ty * TW;
            const int yy = min(max(y0 - RY + ty, 0), H - 1);
            const int xx = min(max(x0 - RX + tx, 0), W - 1);
            v[u] = im[(long long)yy * W + xx];
        }
#pragma unroll
        for (int u = 0; u < PER; ++u)
            if (tid + 256 * u < NE) tile[tid + 256 * u] = (uint16_t)v[u];
    }
    __syncthreads();
    const int tx = tid & 15, ty = tid >> 4;
    const int xg = x0 + 4 * tx;  // this thread's first pixel
    if (xg >= W) return;
    const int ly0 = ty * C8_RPT;
    // w[r]: tile columns 4tx .. 4tx+11 of window row r as u16 pairs
    uint32_t w[WR][6];
#pragma unroll
    for (int r = 0; r < WR; ++r) {
        const uint2* src = reinterpret_cast<const uint2*>(tile + (ly0 + r) * TW + 4 * tx);
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const uint2 t = src[q];
            w[r][2 * q] = t.x;
            w[r][2 * q + 1] = t.y;
        }
    }
    // columns k, k+1 of window row r as one u16 pair
    auto pair_at = [&](int r, int k) -> uint32_t {
        return (k & 1) ? __builtin_amdgcn_alignbit(w[r][(k + 1) >> 1], w[r][(k - 1) >> 1], 16)
                       : w[r][k >> 1];
    };
#pragma unroll
    for (int r = 0; r < C8_RPT; ++r) {
        const int y = y0 + ly0 + r;
        if (y >= H) break;
        const u16x2 cA = as_u16x2(w[r + RY][2]), cB = as_u16x2(w[r + RY][3]);
        // offsets last to first: each bit enters at the top of its half
        // and moves down one place per later offset of its group, so the
        // group's first offset ends most significant (the 14-offset last
        // group ends two places up)
        uint32_t accA[4] = {0u, 0u, 0u, 0u}, accB[4] = {0u, 0u, 0u, 0u};
        int i = (2 * RX + 1) * (2 * RY + 1) - 2;
#pragma unroll
        for (int dy = 2 * RY; dy >= 0; --dy)
#pragma unroll
            for (int dx = 2 * RX; dx >= 0; --dx) {
                if (dy == RY && dx == RX) continue;
                const uint32_t dA = as_u32(as_u16x2(pair_at(r + dy, dx)) - cA);
                const uint32_t dB = as_u32(as_u16x2(pair_at(r + dy, dx + 2)) - cB);
                accA[i >> 4] = (accA[i >> 4] >> 1) | (dA & 0x80008000u);
                accB[i >> 4] = (accB[i >> 4] >> 1) | (dB & 0x80008000u);
                // keep one shift and one and-or per step (unfenced, the
                // compiler re-associates the chain into more operations)
                asm volatile("" : "+v"(accA[i >> 4]), "+v"(accB[i >> 4]));
                --i;
            }
        uint64_t* o = out + (long long)blockIdx.z * plane + (long long)y * W + xg;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (xg + j >= W) break;
            const uint32_t* acc = j < 2 ? accA : accB;
            const int sh = 16 * (j & 1);
            const uint32_t g0 = (acc[0] >> sh) & 0xFFFFu, g1 = (acc[1] >> sh) & 0xFFFFu;
            const uint32_t g2 = (acc[2] >> sh) & 0xFFFFu, g3 = ((acc[3] >> sh) & 0xFFFFu) >> 2;
            const uint32_t hi = (g0 << 14) | (g1 >> 2);
            const uint32_t lo = (g1 << 30) | (g2 << 14) | g3;
            o[j] = ((uint64_t)hi << 32) | lo;
        }
    }
}

// band centre, stereonet.py:5254-5257: max(0, int(disp0 + 0.5)), int()
// truncating toward zero.  Values beyond 2^30 (and NaN) put the whole band
// outside the image, which is what the reference's huge centre does.
__device__ __forceinline__ int band_centre(double d0) {
    const double t = d0 + 0.5;
    if (!(t < 1073741824.0)) return 1 << 30;
    return t < 1.0 ? 0 : (int)t;
}

__global__ void k_base(const double* __restrict__ d0, long long n, int* __restrict__ base) {
    const long long i = blockIdx.x * 256LL + threadIdx.x;
    if (i >= n) return;
    base[i] = band_centre(d0[i]);
}

// ------------------------------------------------------------- band costs
// band_costs, stereonet.py:5156-5189, integer path: C = popcount(a ^ b) or
// `invalid` where the candidate leaves the image.  RIGHT selects the right
// view matched in left coordinates: partner column x + d instead of x - d.
// KT > 0 fixes the band length (Kp = KT rounded up to 8): the partner words
// are read from clamped columns, all KT loads in flight, and the candidates
// outside the image take `invalid` afterwards.
// d0 != nullptr: the band centres are formed here from d0 (k_base's rule)
// and written to base for the later stages, instead of being read.
template <bool RIGHT, int KT>
__global__ void __launch_bounds__(256) k_costs_int(const uint64_t* __restrict__ cA,
                                                   const uint64_t* __restrict__ cB,
                                                   int* __restrict__ base,
                                                   const double* __restrict__ d0, Geom g, int R,
                                                   int K, int Kp, int invalid,
                                                   uint8_t* __restrict__ C) {
    // each wave: 64 pixels of one row (blockDim.y rows per block) x Kp bytes,
    // staged in the wave's own LDS region and written back as contiguous
    // 8-byte units (a lane's own Kp bytes are Kp apart from its neighbour's).
    // A wave reads back only what it wrote, in issue order, so no block
    // barrier is needed and a wave past the image leaves at once.
    extern __shared__ __align__(16) uint32_t cstage_all[];
    const int lane = threadIdx.x;
    uint32_t* cstage = cstage_all + threadIdx.y * 64 * (Kp / 4);
    const int x0 = blockIdx.x * 64;
    const int x = x0 + lane;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    const int W = g.w(blockIdx.z);
    if (x0 >= W || y >= g.h(blockIdx.z)) return;  // whole wave outside this image
    const long long plane = g.slot;
    const long long row0 = (long long)blockIdx.z * plane + (long long)y * W;
    if constexpr (KT > 0) {
        // The wave's partner words (every lane's KT columns, clamped to the
        // row) usually span a short run of the row: it is read once with
        // coalesced loads into the wave's LDS window and each lane takes its
        // KT words from there; a longer run is read per lane from memory.
        // Every lane takes part (a lane past the row stands in for the last
        // pixel and writes nothing).
        constexpr int KPT = (KT + 7) / 8 * 8;
        constexpr int WIN = 256;  // words per wave window
        uint64_t* win = reinterpret_cast<uint64_t*>(cstage_all + blockDim.y * 64 * (KPT / 4)) +
                        threadIdx.y * WIN;
        const bool live = x < W;
        const int xc = live ? x : W - 1;
        const long long pix = row0 + xc;
        const uint64_t a = cA[pix];
        int bc;
        if (d0) {
            bc = band_centre(d0[pix]);
            if (live) base[pix] = bc;
        } else {
            bc = base[pix];
        }
        const int b0 = bc - R;
        const uint64_t* rowB = cB + row0;
        const int e0 = RIGHT ? xc + b0 : xc - b0 - (KT - 1);
        int wlo = min(max(e0, 0), W - 1), whi = min(max(e0 + KT - 1, 0), W - 1);
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            wlo = min(wlo, __shfl_xor(wlo, o, 64));
            whi = max(whi, __shfl_xor(whi, o, 64));
        }
        wlo = __builtin_amdgcn_readfirstlane(wlo);
        whi = __builtin_amdgcn_readfirstlane(whi);
        uint64_t w[KT];
        if (whi - wlo < WIN) {
            for (int i = 0; wlo + 64 * i <= whi; ++i) {
                const int e = wlo + 64 * i + lane;
                if (e <= whi) win[64 * i + lane] = rowB[e];
            }
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int k = 0; k < KT; ++k) {
                const int d = b0 + k;
                const int xr = RIGHT ? xc + d : xc - d;
                w[k] = win[min(max(xr, 0), W - 1) - wlo];
            }
        } else {
#pragma unroll
            for (int k = 0; k < KT; ++k) {
                const int d = b0 + k;
                const int xr = RIGHT ? xc + d : xc - d;
                w[k] = rowB[min(max(xr, 0), W - 1)];
            }
        }
        if (live) {
#pragma unroll
            for (int k0 = 0; k0 < KPT; k0 += 4) {
                uint32_t packed = 0;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const int k = k0 + j;
                    if (k < KT) {
                        const int d = b0 + k;
                        const int xr = RIGHT ? x + d : x - d;
                        const bool ok = d >= 0 && xr >= 0 && xr < W;
                        const uint32_t v = ok ? (uint32_t)__popcll(a ^ w[k]) : (uint32_t)invalid;
                        packed |= v << (8 * j);
                    }
                }
                cstage[lane * (KPT / 4) + k0 / 4] = packed;
            }
        }
    } else if (x < W) {
        const long long pix = row0 + x;
        const uint64_t a = cA[pix];
        int bc;
        if (d0) {
            bc = band_centre(d0[pix]);
            base[pix] = bc;
        } else {
            bc = base[pix];
        }
        const int b0 = bc - R;
        const uint64_t* rowB = cB + row0;
        for (int k0 = 0; k0 < Kp; k0 += 4) {
            uint32_t packed = 0;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int k = k0 + j;
                uint32_t v = 0;
                if (k < K) {
                    const int d = b0 + k;
                    const int xr = RIGHT ? x + d : x - d;
                    const bool ok = d >= 0 && xr >= 0 && xr < W;
                    v = ok ? (uint32_t)__popcll(a ^ rowB[ok ? xr : 0]) : (uint32_t)invalid;
                }
                packed |= v << (8 * j);
            }
            cstage[lane * (Kp / 4) + k0 / 4] = packed;
        }
    }
    __builtin_amdgcn_wave_barrier();
    const int units = (min(64, W - x0) * Kp) / 8;
    const uint2* src = reinterpret_cast<const uint2*>(cstage);
    uint2* dst = reinterpret_cast<uint2*>(C + (row0 + x0) * Kp);
    for (int u = lane; u < units; u += 64) dst[u] = src[u];
}

// band_costs, general float64 path (absolute-difference and prior terms),
// evaluated in the reference's operation order.
template <bool RIGHT, typename T>
__global__ void __launch_bounds__(64) k_costs_f64(const uint64_t* __restrict__ cA,
                                                  const uint64_t* __restrict__ cB,
                                                  const T* __restrict__ gA, const T* __restrict__ gB,
                                                  const int* __restrict__ base,
                                                  const double* __restrict__ prior, Geom g, int R,
                                                  int K, int Kp, double ad_weight,
                                                  double ad_trunc, double invalid, double prior_w,
                                                  double prior_t, double* __restrict__ C) {
    const int x = blockIdx.x * 64 + threadIdx.x;
    const int y = blockIdx.y;
    const int W = g.w(blockIdx.z);
    if (x >= W || y >= g.h(blockIdx.z)) return;
    const long long plane = g.slot;
    const long long row0 = (long long)blockIdx.z * plane + (long long)y * W;
    const long long pix = row0 + x;
    const uint64_t a = cA[pix];
    const double il = to_gray<T>(gA[pix]);
    const int b0 = base[pix] - R;
    const double pr = prior_w > 0.0 ? prior[pix] : 0.0;
    double* o = C + pix * Kp;
    for (int k = 0; k < Kp; ++k) {
        double c = invalid;
        if (k >= K) {
            c = 0.0;
        } else {
            const int d = b0 + k;
            const int xr = RIGHT ? x + d : x - d;
            if (d >= 0 && xr >= 0 && xr < W) {
                double ad = il - to_gray<T>(gB[row0 + xr]);
                if (ad < 0) ad = -ad;
                ad = ad * 255.0;
                if (ad > ad_trunc) ad = ad_trunc;
                c = (double)__popcll(a ^ cB[row0 + xr]) + ad_weight * ad;
                if (prior_w > 0.0) {
                    double e = (double)(b0 + k) - pr;
                    if (e < 0) e = -e;
                    if (e > prior_t) e = prior_t;
                    c += prior_w * e;
                }
            }
        }
        o[k] = c;
    }
}

// ------------------------------------------------------------- SGM paths
template <typename CT, typename AT, int KC>
__device__ __forceinline__ void load_chunk(const CT* __restrict__ p, AT (&c)[KC]);

template <typename AT, int KC>
__device__ __forceinline__ void load_chunk_u8(const uint8_t* __restrict__ p, AT (&c)[KC]) {
#pragma unroll
    for (int q = 0; q < KC / 8; ++q) {
        const uint2 v = *reinterpret_cast<const uint2*>(p + 8 * q);
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            c[8 * q + t] = (AT)((v.x >> (8 * t)) & 0xFFu);
            c[8 * q + 4 + t] = (AT)((v.y >> (8 * t)) & 0xFFu);
        }
    }
}

template <typename AT, int KC>
__device__ __forceinline__ void load_chunk_f64(const double* __restrict__ p, AT (&c)[KC]) {
#pragma unroll
    for (int q = 0; q < KC; ++q) c[q] = (AT)p[q];
}

template <typename LT, typename AT, int KC>
__device__ __forceinline__ void store_chunk(LT* __restrict__ p, const AT (&L)[KC], int nvalid);

template <typename AT, int KC>
__device__ __forceinline__ void store_u8(uint8_t* __restrict__ p, const AT (&L)[KC], int nvalid) {
#pragma unroll
    for (int q = 0; q < KC / 8; ++q) {
        uint2 v = {0u, 0u};
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const uint32_t a = (8 * q + t) < nvalid ? ((uint32_t)L[8 * q + t] & 0xFFu) : 0u;
            const uint32_t b = (8 * q + 4 + t) < nvalid ? ((uint32_t)L[8 * q + 4 + t] & 0xFFu) : 0u;
            v.x |= a << (8 * t);
            v.y |= b << (8 * t);
        }
        *reinterpret_cast<uint2*>(p + 8 * q) = v;
    }
}

template <typename AT, int KC>
__device__ __forceinline__ void store_u16(uint16_t* __restrict__ p, const AT (&L)[KC], int nvalid) {
#pragma unroll
    for (int q = 0; q < KC / 4; ++q) {
        uint2 v = {0u, 0u};
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            const uint32_t a = (4 * q + t) < nvalid ? ((uint32_t)L[4 * q + t] & 0xFFFFu) : 0u;
            const uint32_t b = (4 * q + 2 + t) < nvalid ? ((uint32_t)L[4 * q + 2 + t] & 0xFFFFu) : 0u;
            v.x |= a << (16 * t);
            v.y |= b << (16 * t);
        }
        *reinterpret_cast<uint2*>(p + 4 * q) = v;
    }
}

template <typename AT, int KC>
__device__ __forceinline__ void store_f64(double* __restrict__ p, const AT (&L)[KC], int nvalid) {
#pragma unroll
    for (int q = 0; q < KC; ++q) p[q] = q < nvalid ? (double)L[q] : 0.0;
}

template <typename CT, typename AT, int KC>
__device__ __forceinline__ void load_any(const CT* p, AT (&c)[KC]) {
    if constexpr (sizeof(CT) == 1)
        load_chunk_u8<AT, KC>(reinterpret_cast<const uint8_t*>(p), c);
    else
        load_chunk_f64<AT, KC>(reinterpret_cast<const double*>(p), c);
}

template <typename LT, typename AT, int KC>
__device__ __forceinline__ void store_any(LT* p, const AT (&L)[KC], int nvalid) {
    if constexpr (sizeof(LT) == 1)
        store_u8<AT, KC>(reinterpret_cast<uint8_t*>(p), L, nvalid);
    else if constexpr (sizeof(LT) == 2)
        store_u16<AT, KC>(reinterpret_cast<uint16_t*>(p), L, nvalid);
    else
        store_f64<AT, KC>(reinterpret_cast<double*>(p), L, nvalid);
}

// _aggregate_path, stereonet.py:5192-5241, for every direction of a launch at
// once (blockIdx.y = direction).  G lanes own one scanline; lane j holds band
// entries [j*KC, (j+1)*KC) of the current pixel in registers.  The previous
// pixel's neighbour-minimum M[m] = min(L[m], L[m-1]+P1, L[m+1]+P1) (m in
// [-1, K]) is staged in LDS so that the re-centring shift
// s = base[p] - base[q] becomes an indexed read M[k+s], clamped into a
// sentinel slot outside [-1, K]; then
//   L[k] = C[k] + min(M[k+s], min_k' L_q[k'] + P2) - min_k' L_q[k'],
// which is the reference's candidate set in one read.  Scanlines:
// horizontal directions run one line per row; vertical and diagonal ones run
// one line per column, stepping rows, and a diagonal that leaves the image
// wraps to the one entering at the opposite edge (a border pixel, where the
// recurrence restarts with L = C), so every line has exactly H steps.
template <typename CT, typename LT, typename AT, int KC>
__global__ void __launch_bounds__(64) k_sgm(const CT* __restrict__ C, const int* __restrict__ base,
                                            LT* __restrict__ Lout, Geom g, int K, int G,
                                            AT P1, AT P2, int dir0, long long dirStride) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    AT* lds = reinterpret_cast<AT*>(smem_raw);
    const int W = g.w(blockIdx.z), H = g.h(blockIdx.z);
    const int Kp = G * KC;
    const int dir = dir0 + (int)blockIdx.y;
    const int dx = c_dx[dir], dy = c_dy[dir];
    const int lane = threadIdx.x;
    const int lpw = 64 / G;
    const int li = lane / G;
    const int j = lane - li * G;
    const bool horiz = dy == 0;
    const int nLines = horiz ? H : W;
    const int len = horiz ? W : H;
    if ((int)blockIdx.x * lpw >= nLines) return;
    const int line = blockIdx.x * lpw + li;
    const bool active = line < nLines;
    const int stride = Kp + 5;
    AT* slot = lds + li * stride;
    const AT INF = big<AT>();
    if (j == 0) {
        slot[0] = INF;
        slot[K + 3] = INF;
    }
    const long long plane = g.slot;
    const CT* Cb = C + (long long)blockIdx.z * plane * Kp;
    const int* Bb = base + (long long)blockIdx.z * plane;
    LT* Lb = Lout + (long long)blockIdx.y * dirStride + (long long)blockIdx.z * plane * Kp;
    const int kbase = j * KC;
    AT L[KC];
    AT mprev = 0;
    int bprev = 0;
    for (int i = 0; i < len; ++i) {
        int x, y;
        bool border;
        if (horiz) {
            y = line;
            x = dx > 0 ? i : W - 1 - i;
            border = i == 0;
        } else {
            y = dy > 0 ? i : H - 1 - i;
            if (dx == 0) {
                x = line;
                border = i == 0;
            } else {
                int t = (dx > 0 ? line + i : line - i) % W;
                if (t < 0) t += W;
                x = t;
                border = i == 0 || (dx > 0 ? x == 0 : x == W - 1);
            }
        }
        if (!active) {
            x = 0;
            y = 0;
        }
        const long long pix = (long long)y * W + x;
        AT c[KC];
        load_any<CT, AT, KC>(Cb + pix * Kp + kbase, c);
        const int b = Bb[pix];
        if (border) {
#pragma unroll
            for (int kk = 0; kk < KC; ++kk) L[kk] = c[kk];
        } else {
            const int s = b - bprev;
            const AT cap = mprev + P2;
#pragma unroll
            for (int kk = 0; kk < KC; ++kk) {
                const int idx = min(max(kbase + kk + s + 2, 0), K + 3);
                const AT best = amin(slot[idx], cap);
                L[kk] = c[kk] + best - mprev;
            }
        }
#pragma unroll
        for (int kk = 0; kk < KC; ++kk)
            if (kbase + kk >= K) L[kk] = INF;
        AT ml = L[0];
#pragma unroll
        for (int kk = 1; kk < KC; ++kk) ml = amin(ml, L[kk]);
        for (int off = G >> 1; off > 0; off >>= 1) ml = amin(ml, __shfl_xor(ml, off, G));
        AT leftN = __shfl_up(L[KC - 1], 1, G);
        AT rightN = __shfl_down(L[0], 1, G);
        if (j == 0) leftN = INF;
        if (j == G - 1) rightN = INF;
        if (G > 1) __syncthreads();
        if (active) {
#pragma unroll
            for (int kk = 0; kk < KC; ++kk) {
                const int k = kbase + kk;
                const AT lm = kk > 0 ? L[kk - 1] : leftN;
                const AT lp = kk < KC - 1 ? L[kk + 1] : rightN;
                const AT m = amin(L[kk], amin(lm, lp) + P1);
                if (k < K) slot[k + 2] = m;
                if (k == 0) slot[1] = L[kk] + P1;
                if (k == K - 1) slot[K + 2] = L[kk] + P1;
            }
            store_any<LT, AT, KC>(Lb + pix * Kp + kbase, L, K - kbase);
        }
        if (G > 1) __syncthreads();
        mprev = ml;
        bprev = b;
    }
}

// Specialisation of k_sgm for one lane per scanline (G = 1) with the band
// length K known at compile time (the pipeline's K = 21, the signature
// default K = 7), integer path.  Same recurrence and LDS staging as k_sgm;
// with K static every padding test and slot index resolves at compile time.
// A lane owns its scanline alone, so no barrier is needed: its LDS reads of
// the previous pixel's M precede its writes of the current one in program
// order.
template <int K, typename LT>
__global__ void __launch_bounds__(64) k_sgm_k(const uint8_t* __restrict__ C,
                                              const int* __restrict__ base,
                                              LT* __restrict__ Lout, Geom g, int Kp, int P1,
                                              int P2, int dir0, long long dirStride) {
    constexpr int SL = ((K + 4 + 3) / 4) * 4 + 1;  // odd stride spreads lanes over banks
    __shared__ int lds[64 * SL];
    const int W = g.w(blockIdx.z), H = g.h(blockIdx.z);
    const int dir = dir0 + (int)blockIdx.y;
    const int dx = c_dx[dir], dy = c_dy[dir];
    const int lane = threadIdx.x;
    const bool horiz = dy == 0;
    const int nLines = horiz ? H : W;
    const int len = horiz ? W : H;
    const int line = blockIdx.x * 64 + lane;
    if (line >= nLines) return;
    int* slot = lds + lane * SL;
    constexpr int INF = 0x3FFFFFFF;
    slot[0] = INF;
    slot[K + 3] = INF;
    const long long plane = g.slot;
    const uint8_t* Cb = C + (long long)blockIdx.z * plane * Kp;
    const int* Bb = base + (long long)blockIdx.z * plane;
    LT* Lb = Lout + (long long)blockIdx.y * dirStride + (long long)blockIdx.z * plane * Kp;
    int L[K];
    int mprev = 0, bprev = 0;
    for (int i = 0; i < len; ++i) {
        int x, y;
        bool border;
        if (horiz) {
            y = line;
            x = dx > 0 ? i : W - 1 - i;
            border = i == 0;
        } else {
            y = dy > 0 ? i : H - 1 - i;
            if (dx == 0) {
                x = line;
                border = i == 0;
            } else {
                int t = (dx > 0 ? line + i : line - i) % W;
                if (t < 0) t += W;
                x = t;
                border = i == 0 || (dx > 0 ? x == 0 : x == W - 1);
            }
        }
        const long long pix = (long long)y * W + x;
        int c[(K + 7) / 8 * 8];
        load_chunk_u8<int, (K + 7) / 8 * 8>(Cb + pix * Kp, c);
        const int b = Bb[pix];
        if (border) {
#pragma unroll
            for (int k = 0; k < K; ++k) L[k] = c[k];
        } else {
            const int s = b - bprev;
            const int cap = mprev + P2;
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const int idx = min(max(k + s + 2, 0), K + 3);
                L[k] = c[k] + min(slot[idx], cap) - mprev;
            }
        }
        int m = L[0];
#pragma unroll
        for (int k = 1; k < K; ++k) m = min(m, L[k]);
        slot[1] = L[0] + P1;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int lm = k > 0 ? L[k - 1] : INF;
            const int lp = k < K - 1 ? L[k + 1] : INF;
            slot[k + 2] = min(L[k], min(lm, lp) + P1);
        }
        slot[K + 2] = L[K - 1] + P1;
        int Ls[(K + 7) / 8 * 8];
#pragma unroll
        for (int k = 0; k < (K + 7) / 8 * 8; ++k) Ls[k] = k < K ? L[k] : 0;
        store_any<LT, int, (K + 7) / 8 * 8>(Lb + pix * Kp, Ls, K);
        mprev = m;
        bprev = b;
    }
}

// Packed form of k_sgm_k for the uint8 path-cost case (odd K >= 3, every
// value below 0xFFFF): band entries travel as u16 pairs, so each add / sub
// / min is one v_pk_*_u16 for two entries.  The LDS slot holds u16 M values
// with INF (0xFFFF) guard zones of K+2 entries on both sides, so the
// shifted read M[k+s] needs no per-entry clamp once s is clamped to
// [-(K+2), K+2] (a larger shift reads only INF, leaving the P2 candidate,
// as in the reference).  Neighbour pairs (L[2j-1], L[2j]) and
// (L[2j+1], L[2j+2]) come from v_alignbit.  The pad entry (index K) is
// kept at INF, which makes the computed M[K] equal L[K-1] + P1.

// Position of step i of a scanline (see k_sgm for the line geometry).
struct LinePos {
    int W, H, dx, dy, line;
    bool horiz;
    __device__ __forceinline__ long long at(int i, bool& border) const {
        int x, y;
        if (horiz) {
            y = line;
            x = dx > 0 ? i : W - 1 - i;
            border = i == 0;
        } else {
            y = dy > 0 ? i : H - 1 - i;
            if (dx == 0) {
                x = line;
                border = i == 0;
            } else {
                int t = (dx > 0 ? line + i : line - i) % W;
                if (t < 0) t += W;
                x = t;
                border = i == 0 || (dx > 0 ? x == 0 : x == W - 1);
            }
        }
        return (long long)y * W + x;
    }
};

// LinePos::at for consecutive steps without its modulo: the pixel index
// and column are carried from one step to the next, a diagonal wrapping
// from one image edge to the other (dx is -1, 0 or 1).
struct LineCursor {
    int x, p, dx, dyW, W;
    __device__ __forceinline__ void start(const LinePos& lp, int i) {
        bool bd;
        p = (int)lp.at(i, bd);
        dx = lp.dx;
        dyW = lp.dy * lp.W;
        W = lp.W;
        x = lp.horiz ? (lp.dx > 0 ? i : lp.W - 1 - i) : p - (lp.dy > 0 ? i : lp.H - 1 - i) * lp.W;
    }
    __device__ __forceinline__ void next() {
        const int xn = x + dx;  // in [-1, W]
        int adj = xn >= W ? W : 0;
        adj = xn < 0 ? -W : adj;
        x = xn - adj;
        p += dyW + dx - adj;
    }
    // border pixel (the recurrence restarts): the first step, or where a
    // diagonal re-enters at the opposite edge
    __device__ __forceinline__ bool border(int i) const {
        return i == 0 || (dx > 0 && x == 0) || (dx < 0 && x == W - 1);
    }
};

// Band geometry of the packed kernels for band length K.
template <int K>
struct PackedBand {
    static_assert(K % 2 == 1 && K >= 3, "odd band length");
    static constexpr int NP = (K + 1) / 2;       // u16 pairs, the last one padded
    static constexpr int NB = (2 * NP + 7) / 8;  // uint2 (8-byte) cost/path chunks
    static constexpr int OFF = K + 3;            // u16 index of M[0]; M[-1] at OFF-1
    static constexpr int SLOT16 = 3 * K + 6;
    static constexpr int SD = ((SLOT16 + 1) / 2) | 1;  // slot dwords per lane, odd
    // the dword reads of the re-centred minima (u16 index OFF + shift,
    // |shift| <= K + 2) start at or after dword 0 and end inside the slot
    static_assert(OFF - (K + 2) >= 0 && (OFF + K + 2) / 2 + NP < SD, "minima reads in the slot");
};

// One step of the packed recurrence for one scanline: costs cw (u8, 2*NB
// words) and band centre b of the current pixel in, path costs out (u8,
// entries >= K written as 0).  L, mprev and bprev carry the line's state;
// sl32 / s16 are the line's LDS slot (INF-initialised).
template <int K>
__device__ __forceinline__ void packed_step(u16x2 (&L)[PackedBand<K>::NP], uint32_t& mprev,
                                            int& bprev, u32_alias* sl32, u16_alias* s16,
                                            const uint32_t (&cw)[2 * PackedBand<K>::NB], int b,
                                            bool border, int P1, int P2,
                                            uint2 (&out)[PackedBand<K>::NB]) {
    constexpr int NP = PackedBand<K>::NP, NB = PackedBand<K>::NB, OFF = PackedBand<K>::OFF;
    const u16x2 p1v = {(unsigned short)P1, (unsigned short)P1};
    u16x2 c[NP];
#pragma unroll
    for (int j = 0; j < NP; ++j)
        c[j] = as_u16x2(__builtin_amdgcn_perm(0u, cw[j >> 1], (j & 1) ? 0x0c030c02u : 0x0c010c00u));
    if (border) {
#pragma unroll
        for (int j = 0; j < NP; ++j) L[j] = c[j];
    } else {
        const int sh = min(max(b - bprev, -(K + 2)), K + 2);
        const uint32_t cap = mprev + (uint32_t)P2;
        const u16x2 capv = {(unsigned short)cap, (unsigned short)cap};
        const u16x2 mpv = {(unsigned short)mprev, (unsigned short)mprev};
        // the 2*NP neighbour minima from u16 index q0 on: NP + 1 aligned
        // dwords, each pair funnel-shifted by 16 bits when q0 is odd
        const int q0 = OFF + sh;
        const u32_alias* gw = sl32 + (q0 >> 1);
        const uint32_t fs = (uint32_t)(q0 & 1) << 4;
        uint32_t wv[NP + 1];
#pragma unroll
        for (int j = 0; j <= NP; ++j) wv[j] = gw[j];
#pragma unroll
        for (int j = 0; j < NP; ++j) {
            const u16x2 m2 = as_u16x2(__builtin_amdgcn_alignbit(wv[j + 1], wv[j], fs));
            const u16x2 best = __builtin_elementwise_min(m2, capv);
            L[j] = c[j] + best - mpv;
        }
    }
    L[NP - 1].y = 0xFFFF;  // pad entry K
    u16x2 mm = L[0];
#pragma unroll
    for (int j = 1; j < NP; ++j) mm = __builtin_elementwise_min(mm, L[j]);
    const uint32_t m = min((uint32_t)mm.x, (uint32_t)mm.y);
    s16[OFF - 1] = (uint16_t)(L[0].x + P1);
#pragma unroll
    for (int j = 0; j < NP; ++j) {
        const uint32_t lo = j > 0 ? as_u32(L[j - 1]) : 0xFFFFFFFFu;
        const uint32_t hi = j + 1 < NP ? as_u32(L[j + 1]) : 0xFFFFFFFFu;
        const u16x2 left = as_u16x2(__builtin_amdgcn_alignbit(as_u32(L[j]), lo, 16));
        const u16x2 right = as_u16x2(__builtin_amdgcn_alignbit(hi, as_u32(L[j]), 16));
        const u16x2 nb = __builtin_elementwise_min(left, right) + p1v;
        sl32[OFF / 2 + j] = as_u32(__builtin_elementwise_min(L[j], nb));
    }
    // path costs: low bytes of the pairs
    u16x2 Ls[4 * NB];
#pragma unroll
    for (int j = 0; j < 4 * NB; ++j) Ls[j] = j < NP ? L[j] : (u16x2){0, 0};
    Ls[NP - 1].y = 0;
#pragma unroll
    for (int q = 0; q < NB; ++q) {
        out[q].x = __builtin_amdgcn_perm(as_u32(Ls[4 * q + 1]), as_u32(Ls[4 * q]), 0x06040200u);
        out[q].y = __builtin_amdgcn_perm(as_u32(Ls[4 * q + 3]), as_u32(Ls[4 * q + 2]), 0x06040200u);
    }
    mprev = m;
    // a value of its own (empty asm): otherwise bprev shares the register
    // of the prefetch ring and keeps it alive past its refill, which costs a
    // full memory wait per group of steps
    int bc = b;
    asm volatile("" : "+v"(bc));
    bprev = bc;
}

// Vertical and diagonal directions (and horizontal ones as a fallback):
// one lane per scanline, 64 lines per wave.  Neighbouring lanes hold
// neighbouring columns of one row, so cost reads and path-cost writes are
// coalesced; the costs of the next PF steps are prefetched into registers.
template <int K, int PF>
__global__ void __launch_bounds__(64) k_sgm_p(const uint8_t* __restrict__ C,
                                              const int* __restrict__ base,
                                              uint8_t* __restrict__ Lout, Geom g, int Kp,
                                              int P1, int P2, int dir0, long long dirStride) {
    using PB = PackedBand<K>;
    constexpr int NP = PB::NP, NB = PB::NB, SD = PB::SD;
    __shared__ uint32_t lds[64 * SD];
    const int W = g.w(blockIdx.z), H = g.h(blockIdx.z);
    const int dir = dir0 + (int)blockIdx.y;
    LinePos lp;
    lp.dx = c_dx[dir];
    lp.dy = c_dy[dir];
    lp.W = W;
    lp.H = H;
    lp.horiz = lp.dy == 0;
    const int lane = threadIdx.x;
    const int nLines = lp.horiz ? H : W;
    const int len = lp.horiz ? W : H;
    lp.line = blockIdx.x * 64 + lane;
    if (lp.line >= nLines) return;
    u32_alias* sl32 = reinterpret_cast<u32_alias*>(lds + lane * SD);
    u16_alias* s16 = reinterpret_cast<u16_alias*>(sl32);
    for (int i = 0; i < SD; ++i) sl32[i] = 0xFFFFFFFFu;
    const long long plane = g.slot;
    const uint8_t* Cb = C + (long long)blockIdx.z * plane * Kp;
    const int* Bb = base + (long long)blockIdx.z * plane;
    uint8_t* Lb = Lout + (long long)blockIdx.y * dirStride + (long long)blockIdx.z * plane * Kp;
    // prefetch ring: cost chunks and band centres of the next PF steps
    uint2 ring[PF][NB];
    int rb[PF];
#pragma unroll
    for (int u = 0; u < PF; ++u) {
        bool bd;
        const long long pq = lp.at(u < len ? u : 0, bd);
#pragma unroll
        for (int q = 0; q < NB; ++q) ring[u][q] = *reinterpret_cast<const uint2*>(Cb + pq * Kp + 8 * q);
        rb[u] = Bb[pq];
    }
    u16x2 L[NP];
    uint32_t mprev = 0;
    int bprev = 0;
    auto step = [&](int i, const uint2 (&cr)[NB], int b) {
        bool border;
        const long long pix = lp.at(i, border);
        uint32_t cw[2 * NB];
#pragma unroll
        for (int q = 0; q < NB; ++q) {
            cw[2 * q] = cr[q].x;
            cw[2 * q + 1] = cr[q].y;
        }
        uint2 o[NB];
        packed_step<K>(L, mprev, bprev, sl32, s16, cw, b, border, P1, P2, o);
#pragma unroll
        for (int q = 0; q < NB; ++q) *reinterpret_cast<uint2*>(Lb + pix * Kp + 8 * q) = o[q];
    };
    auto refill = [&](int u, int inext) {
        bool bd;
        const long long pq = lp.at(inext, bd);
#pragma unroll
        for (int q = 0; q < NB; ++q) ring[u][q] = *reinterpret_cast<const uint2*>(Cb + pq * Kp + 8 * q);
        rb[u] = Bb[pq];
    };
    // full groups of PF steps: no exit inside a group, so the ring stays in
    // place and the loads of step i + PF overlap steps i+1 .. i+PF-1
    int i0 = 0;
    for (; i0 + PF <= len; i0 += PF) {
#pragma unroll
        for (int u = 0; u < PF; ++u) {
            step(i0 + u, ring[u], rb[u]);
            refill(u, min(i0 + u + PF, len - 1));
        }
    }
#pragma unroll
    for (int u = 0; u < PF; ++u)
        if (i0 + u < len) step(i0 + u, ring[u], rb[u]);
}

// Horizontal directions (dy = 0): one lane per row, 64 rows per wave.  A
// lane's consecutive pixels are consecutive in memory but neighbouring lanes
// are a row apart, so direct per-lane access would touch 64 cache lines per
// instruction.  Instead the wave stages T columns of its 64 rows through LDS:
// each row's T*Kp cost bytes are one contiguous run, loaded and later stored
// by the whole wave; every lane then steps through its row in LDS, writing
// the path costs over the costs it has consumed.  TC columns per tile.
template <int K, int TC>
struct HorizTile {
    static constexpr int NB = PackedBand<K>::NB;
    static constexpr int T = TC;
    static constexpr int ROWD = T * NB * 2;                 // dwords of one row's run
    static constexpr int RS = ROWD + ((34 - ROWD % 32) % 32);  // row stride, = 2 mod 32
    static constexpr int BS = T + 1;                        // band-centre row stride
};

template <int K, int TC>
__global__ void __launch_bounds__(64) k_sgm_ph(const uint8_t* __restrict__ C,
                                               const int* __restrict__ base,
                                               uint8_t* __restrict__ Lout, Geom g, int Kp,
                                               int P1, int P2, int dir0, long long dirStride) {
    using PB = PackedBand<K>;
    using HT = HorizTile<K, TC>;
    const int W = g.w(blockIdx.z), H = g.h(blockIdx.z);
    if ((int)blockIdx.x * 64 >= H) return;  // whole block below this image
    constexpr int NP = PB::NP, NB = PB::NB, SD = PB::SD;
    constexpr int T = HT::T, RS = HT::RS, BS = HT::BS;
    static_assert(RS % 2 == 0, "8-byte aligned rows");
    __shared__ __align__(16) uint32_t tile[64 * RS];
    __shared__ int btile[64 * BS];
    __shared__ uint32_t lds[64 * SD];
    const int dx = c_dx[dir0 + (int)blockIdx.y];
    const int lane = threadIdx.x;
    const int y0 = blockIdx.x * 64;
    const int rows = min(64, H - y0);
    u32_alias* sl32 = reinterpret_cast<u32_alias*>(lds + lane * SD);
    u16_alias* s16 = reinterpret_cast<u16_alias*>(sl32);
    for (int i = 0; i < SD; ++i) sl32[i] = 0xFFFFFFFFu;
    const long long plane = g.slot;
    const uint8_t* Cb = C + (long long)blockIdx.z * plane * Kp + (long long)y0 * W * Kp;
    const int* Bb = base + (long long)blockIdx.z * plane + (long long)y0 * W;
    uint8_t* Lb = Lout + (long long)blockIdx.y * dirStride + (long long)blockIdx.z * plane * Kp +
                  (long long)y0 * W * Kp;
    const long long rowBytes = (long long)W * Kp;
    u16x2 L[NP];
    uint32_t mprev = 0;
    int bprev = 0;
    uint2* mine = reinterpret_cast<uint2*>(tile + lane * RS);
    const int* bmine = btile + lane * BS;
    // staging walks a (row, unit) grid of the tile, 64 units per
    // instruction; units past the chunk or rows past the image are clamped
    // onto the last valid one (a duplicate load and an identical LDS write)
    constexpr int CU = T * NB;  // 8-byte cost units of a full row run
    static_assert(CU % 8 == 0, "staging runs in groups of 8 instructions");
    for (int c0 = 0; c0 < W; c0 += T) {
        const int n = min(T, W - c0);
        const int xlo = dx > 0 ? c0 : W - c0 - n;
        const int units = n * NB;
        const uint8_t* Cx = Cb + (long long)xlo * Kp;
        const int* Bx = Bb + xlo;
#pragma unroll 1
        for (int it = 0; it < CU; it += 8) {
            uint2 v[8];
            int off[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int idx = (it + u) * 64 + lane;
                const int r = min(idx / CU, rows - 1);
                const int q = min(idx % CU, units - 1);
                off[u] = r * RS + 2 * q;
                v[u] = *reinterpret_cast<const uint2*>(Cx + r * rowBytes + 8 * q);
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) *reinterpret_cast<uint2*>(tile + off[u]) = v[u];
        }
#pragma unroll
        for (int it = 0; it < T; it += 8) {
            int v[8], off[8];
#pragma unroll
            for (int u = 0; u < 8 && it + u < T; ++u) {
                const int idx = (it + u) * 64 + lane;
                const int r = min(idx / T, rows - 1);
                const int c = min(idx % T, n - 1);
                off[u] = r * BS + c;
                v[u] = Bx[(long long)r * W + c];
            }
#pragma unroll
            for (int u = 0; u < 8 && it + u < T; ++u) btile[off[u]] = v[u];
        }
        __syncthreads();
        for (int j = 0; j < n; ++j) {
            const int col = dx > 0 ? j : n - 1 - j;
            uint32_t cw[2 * NB];
#pragma unroll
            for (int q = 0; q < NB; ++q) {
                const uint2 v = mine[col * NB + q];
                cw[2 * q] = v.x;
                cw[2 * q + 1] = v.y;
            }
            uint2 o[NB];
            packed_step<K>(L, mprev, bprev, sl32, s16, cw, bmine[col], c0 + j == 0, P1, P2, o);
#pragma unroll
            for (int q = 0; q < NB; ++q) mine[col * NB + q] = o[q];
        }
        __syncthreads();
        uint8_t* Lx = Lb + (long long)xlo * Kp;
#pragma unroll 1
        for (int it = 0; it < CU; it += 8) {
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int idx = (it + u) * 64 + lane;
                const int r = min(idx / CU, rows - 1);
                const int q = min(idx % CU, units - 1);
                *reinterpret_cast<uint2*>(Lx + r * rowBytes + 8 * q) =
                    *reinterpret_cast<const uint2*>(tile + r * RS + 2 * q);
            }
        }
        __syncthreads();
    }
}

// ------------------------------------------------ two lanes per scanline
// The same recurrence with each scanline's band split over a lane pair
// (lanes 2l, 2l+1): lane h holds u16 pairs [h*NPH, (h+1)*NPH) of the band
// (entries >= K are pads held at the sentinel), reads and writes its NW
// words of each pixel's costs and path costs, and the pair shares one LDS
// slot for the neighbour minima M.  Per step the pair swaps the pairs at
// the split (one DPP lane swap) for the M of the entries beside it and the
// two half-band minima (another swap); the re-centring read M[k+s] spans
// both halves of the shared slot, written by the pair in the previous step
// (one wave's LDS accesses complete in issue order).  Pads have both
// neighbours at the sentinel, so their M uses a saturating +P1 and stays at
// the sentinel; every other value is far below it, so the results equal
// packed_step's.  A wave covers 32 scanlines, so a launch has twice the
// waves of the one-lane kernels and each lane half the work per step.
template <int K>
struct PackedBand2 {
    static constexpr int NP = (K + 1) / 2;
    static constexpr int NPH = (((NP + 1) / 2) + 1) & ~1;  // pairs per lane, even
    static constexpr int NW = NPH / 2;                      // cost words per lane
    static constexpr int OFF = K + 3;                       // u16 index of M[0]
    static constexpr int SLOT16 = OFF + (K + 2) + 4 * NPH + 2;
    static constexpr int SD = ((SLOT16 + 1) / 2) | 1;       // slot dwords per pair, odd
    // sentinel bits of lane 1's pair j (entries 2*NPH + 2j, +1 at or past K)
    static constexpr uint32_t pad1(int j) {
        return (2 * NPH + 2 * j >= K ? 0x0000FFFFu : 0u) | (2 * NPH + 2 * j + 1 >= K ? 0xFFFF0000u : 0u);
    }
    static_assert(2 * NPH <= K, "lane 0 holds no pads");
    static_assert(OFF - (K + 2) >= 0 && (OFF + K + 2 + 2 * NPH) / 2 + NPH < SD,
                  "minima reads in the slot");
};

__device__ __forceinline__ uint32_t swap_pair(uint32_t v) {
    // quad_perm [1, 0, 3, 2]: exchange with the other lane of the pair.  It
    // reads the partner lane, so both lanes must execute it: callers apply it
    // to a lane-selected operand, never inside a branch on the lane's half
    // (a DPP read from a lane the branch disabled returns garbage)
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, false);
}

template <int K>
__device__ __forceinline__ void packed_step2(u16x2 (&L)[PackedBand2<K>::NPH], uint32_t& mprev,
                                             int& bprev, u32_alias* sl32, u16_alias* s16,
                                             const uint32_t (&cw)[PackedBand2<K>::NW], int b,
                                             bool border, int P1, int P2, int h,
                                             uint32_t (&out)[PackedBand2<K>::NW]) {
    using PB = PackedBand2<K>;
    constexpr int NPH = PB::NPH, NW = PB::NW, OFF = PB::OFF;
    const u16x2 p1v = {(unsigned short)P1, (unsigned short)P1};
    u16x2 c[NPH];
#pragma unroll
    for (int j = 0; j < NPH; ++j)
        c[j] = as_u16x2(__builtin_amdgcn_perm(0u, cw[j >> 1], (j & 1) ? 0x0c030c02u : 0x0c010c00u));
    if (border) {
#pragma unroll
        for (int j = 0; j < NPH; ++j) L[j] = c[j];
    } else {
        const int sh = min(max(b - bprev, -(K + 2)), K + 2);
        const uint32_t cap = mprev + (uint32_t)P2;
        const u16x2 capv = {(unsigned short)cap, (unsigned short)cap};
        const u16x2 mpv = {(unsigned short)mprev, (unsigned short)mprev};
        // the 2*NPH neighbour minima from u16 index q0 on: NPH + 1 aligned
        // dwords, each pair funnel-shifted by 16 bits when q0 is odd
        // (q0 >= OFF - (K + 2) > 0, and the last dword stays inside the slot)
        const int q0 = OFF + sh + 2 * NPH * h;
        const u32_alias* gw = sl32 + (q0 >> 1);
        const uint32_t fs = (uint32_t)(q0 & 1) << 4;
        uint32_t wv[NPH + 1];
#pragma unroll
        for (int j = 0; j <= NPH; ++j) wv[j] = gw[j];
#pragma unroll
        for (int j = 0; j < NPH; ++j) {
            const u16x2 m2 = as_u16x2(__builtin_amdgcn_alignbit(wv[j + 1], wv[j], fs));
            const u16x2 best = __builtin_elementwise_min(m2, capv);
            L[j] = c[j] + best - mpv;
        }
    }
#pragma unroll
    for (int j = 0; j < NPH; ++j)
        if (PB::pad1(j)) L[j] = as_u16x2(as_u32(L[j]) | (h ? PB::pad1(j) : 0u));
    u16x2 mm = L[0];
#pragma unroll
    for (int j = 1; j < NPH; ++j) mm = __builtin_elementwise_min(mm, L[j]);
    uint32_t m = min((uint32_t)mm.x, (uint32_t)mm.y);
    m = min(m, swap_pair(m));
    if (h == 0) s16[OFF - 1] = (uint16_t)(L[0].x + P1);
    const uint32_t recv = swap_pair(as_u32(h ? L[0] : L[NPH - 1]));
#pragma unroll
    for (int j = 0; j < NPH; ++j) {
        const uint32_t lo = j > 0 ? as_u32(L[j - 1]) : (h ? recv : 0xFFFFFFFFu);
        const uint32_t hi = j + 1 < NPH ? as_u32(L[j + 1]) : (h ? 0xFFFFFFFFu : recv);
        const u16x2 left = as_u16x2(__builtin_amdgcn_alignbit(as_u32(L[j]), lo, 16));
        const u16x2 right = as_u16x2(__builtin_amdgcn_alignbit(hi, as_u32(L[j]), 16));
        const u16x2 nb = __builtin_elementwise_add_sat(__builtin_elementwise_min(left, right), p1v);
        sl32[OFF / 2 + NPH * h + j] = as_u32(__builtin_elementwise_min(L[j], nb));
    }
    // path costs: low bytes of the pairs, pads 0
#pragma unroll
    for (int q = 0; q < NW; ++q) {
        uint32_t lo = as_u32(L[2 * q]), hi = as_u32(L[2 * q + 1]);
        if (PB::pad1(2 * q)) lo &= h ? ~PB::pad1(2 * q) : 0xFFFFFFFFu;
        if (PB::pad1(2 * q + 1)) hi &= h ? ~PB::pad1(2 * q + 1) : 0xFFFFFFFFu;
        out[q] = __builtin_amdgcn_perm(hi, lo, 0x06040200u);
    }
    mprev = m;
    int bc = b;
    asm volatile("" : "+v"(bc));
    bprev = bc;
}

// packed_step2 for a constant band centre: full-range matching (every
// pixel's band is disparities 0..K-1 around the centre R) has re-centring
// shift 0 at every step, so the neighbour minima M stay in the lane's
// registers and no LDS slot is used.  A launch may serve a band shorter than
// K: its entries from the real length up to K hold cost 255 in the cost
// array (soft pads).  On the u8 path every real path cost is at most
// invalid + P2 <= 255 and a soft pad's is at least its cost, 255, so soft
// pads never lower a real entry's minimum, its M or the line minimum, and
// the real entries equal those of a launch of the exact length.
template <int K>
__device__ __forceinline__ void packed_step2c(u16x2 (&L)[PackedBand2<K>::NPH],
                                              u16x2 (&M)[PackedBand2<K>::NPH], uint32_t& mprev,
                                              const uint32_t (&cw)[PackedBand2<K>::NW],
                                              bool border, int P1, int P2, int h,
                                              uint32_t (&out)[PackedBand2<K>::NW]) {
    using PB = PackedBand2<K>;
    constexpr int NPH = PB::NPH, NW = PB::NW;
    const u16x2 p1v = {(unsigned short)P1, (unsigned short)P1};
    u16x2 c[NPH];
#pragma unroll
    for (int j = 0; j < NPH; ++j)
        c[j] = as_u16x2(__builtin_amdgcn_perm(0u, cw[j >> 1], (j & 1) ? 0x0c030c02u : 0x0c010c00u));
    if (border) {
#pragma unroll
        for (int j = 0; j < NPH; ++j) L[j] = c[j];
    } else {
        const uint32_t cap = mprev + (uint32_t)P2;
        const u16x2 capv = {(unsigned short)cap, (unsigned short)cap};
        const u16x2 mpv = {(unsigned short)mprev, (unsigned short)mprev};
#pragma unroll
        for (int j = 0; j < NPH; ++j) L[j] = c[j] + __builtin_elementwise_min(M[j], capv) - mpv;
    }
#pragma unroll
    for (int j = 0; j < NPH; ++j)
        if (PB::pad1(j)) L[j] = as_u16x2(as_u32(L[j]) | (h ? PB::pad1(j) : 0u));
    u16x2 mm = L[0];
#pragma unroll
    for (int j = 1; j < NPH; ++j) mm = __builtin_elementwise_min(mm, L[j]);
    uint32_t m = min((uint32_t)mm.x, (uint32_t)mm.y);
    m = min(m, swap_pair(m));
    const uint32_t recv = swap_pair(as_u32(h ? L[0] : L[NPH - 1]));
#pragma unroll
    for (int j = 0; j < NPH; ++j) {
        const uint32_t lo = j > 0 ? as_u32(L[j - 1]) : (h ? recv : 0xFFFFFFFFu);
        const uint32_t hi = j + 1 < NPH ? as_u32(L[j + 1]) : (h ? 0xFFFFFFFFu : recv);
        const u16x2 left = as_u16x2(__builtin_amdgcn_alignbit(as_u32(L[j]), lo, 16));
        const u16x2 right = as_u16x2(__builtin_amdgcn_alignbit(hi, as_u32(L[j]), 16));
        const u16x2 nb = __builtin_elementwise_add_sat(__builtin_elementwise_min(left, right), p1v);
        M[j] = __builtin_elementwise_min(L[j], nb);
    }
#pragma unroll
    for (int q = 0; q < NW; ++q) {
        uint32_t lo = as_u32(L[2 * q]), hi = as_u32(L[2 * q + 1]);
        if (PB::pad1(2 * q)) lo &= h ? ~PB::pad1(2 * q) : 0xFFFFFFFFu;
        if (PB::pad1(2 * q + 1)) hi &= h ? ~PB::pad1(2 * q + 1) : 0xFFFFFFFFu;
        out[q] = __builtin_amdgcn_perm(hi, lo, 0x06040200u);
    }
    mprev = m;
}

// The block a k_sgm_p2 workgroup works on.  Workgroups go to the eight
// XCDs in turn (workgroup b to XCD b mod 8), each XCD with its own L2.  With
// xcd set, XCD k takes one contiguous run of the (column block, direction,
// pair) blocks instead: the blocks of one pair's directions then share an
// L2, and its three downward (and three upward) paths, which read the same
// cost row at the same step, find it there.
__device__ __forceinline__ uint3 sgm_block(int xcd) {
    const unsigned gx = gridDim.x, gy = gridDim.y;
    unsigned b = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
    if (xcd) {
        const unsigned nb = gx * gy * gridDim.z, q = nb / 8, r = nb % 8, k = b % 8, i = b / 8;
        b = (k < r ? k * (q + 1) : r * (q + 1) + (k - r) * q) + i;
    }
    return make_uint3(b % gx, (b / gx) % gy, b / (gx * gy));
}

// k_sgm_p with two lanes per scanline (32 lines per wave): one block of 32
// scanlines (bk.x) of direction dir0 + bk.y of image bk.z.  CB: constant
// band centre (packed_step2c; the band centres are not read).  lds: the
// wave's 32 * SD words.
template <int K, int PF, bool CB>
__device__ __forceinline__ void sgm_p2_block(const uint8_t* __restrict__ C,
                                             const int* __restrict__ base,
                                             uint8_t* __restrict__ Lout, const Geom& g, int Kp,
                                             int P1, int P2, int dir0, long long dirStride,
                                             uint3 bk, uint32_t* lds) {
    using PB = PackedBand2<K>;
    constexpr int NPH = PB::NPH, NW = PB::NW, SD = PB::SD;
    const int W = g.w(bk.z), H = g.h(bk.z);
    const int dir = dir0 + (int)bk.y;
    LinePos lp;
    lp.dx = c_dx[dir];
    lp.dy = c_dy[dir];
    lp.W = W;
    lp.H = H;
    lp.horiz = lp.dy == 0;
    const int lane = threadIdx.x, h = lane & 1, ln = lane >> 1;
    const int nLines = lp.horiz ? H : W;
    const int len = lp.horiz ? W : H;
    lp.line = bk.x * 32 + ln;
    if (lp.line >= nLines) return;  // both lanes of a pair leave together
    u32_alias* sl32 = reinterpret_cast<u32_alias*>(lds + ln * SD);
    u16_alias* s16 = reinterpret_cast<u16_alias*>(sl32);
    if constexpr (!CB)
        for (int i = h; i < SD; i += 2) sl32[i] = 0xFFFFFFFFu;
    const long long plane = g.slot;
    const uint8_t* Cb = C + (long long)bk.z * plane * Kp + 4 * NW * h;
    const int* Bb = base + (long long)bk.z * plane;
    uint8_t* Lb = Lout + (long long)bk.y * dirStride + (long long)bk.z * plane * Kp + 4 * NW * h;
    uint32_t ring[PF][NW];
    int rb[PF];
#pragma unroll
    for (int u = 0; u < PF; ++u) {
        bool bd;
        const long long pq = lp.at(u < len ? u : 0, bd);
#pragma unroll
        for (int q = 0; q < NW; ++q) ring[u][q] = *reinterpret_cast<const uint32_t*>(Cb + pq * Kp + 4 * q);
        rb[u] = CB ? 0 : Bb[pq];
    }
    // cur: the pixel of the step being made; pre: the next one to prefetch
    // (steps PF, PF + 1, ..., then the last step again)
    LineCursor cur, pre;
    cur.start(lp, 0);
    int ipre = min(PF, len - 1);
    pre.start(lp, ipre);
    u16x2 L[NPH];
    u16x2 M[CB ? NPH : 1];
    uint32_t mprev = 0;
    int bprev = 0;
    auto step = [&](int i, const uint32_t (&cr)[NW], int b) {
        const bool border = cur.border(i);
        const long long pix = cur.p;
        cur.next();
        uint32_t cw[NW];
#pragma unroll
        for (int q = 0; q < NW; ++q) cw[q] = cr[q];
        uint32_t o[NW];
        if constexpr (CB)
            packed_step2c<K>(L, M, mprev, cw, border, P1, P2, h, o);
        else
            packed_step2<K>(L, mprev, bprev, sl32, s16, cw, b, border, P1, P2, h, o);
#pragma unroll
        for (int q = 0; q < NW; ++q) *reinterpret_cast<uint32_t*>(Lb + pix * Kp + 4 * q) = o[q];
    };
    auto refill = [&](int u) {
        const long long pq = pre.p;
        if (ipre < len - 1) {
            pre.next();
            ++ipre;
        }
#pragma unroll
        for (int q = 0; q < NW; ++q) ring[u][q] = *reinterpret_cast<const uint32_t*>(Cb + pq * Kp + 4 * q);
        rb[u] = CB ? 0 : Bb[pq];
    };
    int i0 = 0;
    for (; i0 + PF <= len; i0 += PF) {
#pragma unroll
        for (int u = 0; u < PF; ++u) {
            step(i0 + u, ring[u], rb[u]);
            refill(u);
        }
    }
#pragma unroll
    for (int u = 0; u < PF; ++u)
        if (i0 + u < len) step(i0 + u, ring[u], rb[u]);
}

template <int K, int PF, bool CB = false>
__global__ void __launch_bounds__(64) k_sgm_p2(const uint8_t* __restrict__ C,
                                               const int* __restrict__ base,
                                               uint8_t* __restrict__ Lout, Geom g, int Kp,
                                               int P1, int P2, int dir0, long long dirStride,
                                               int xcd) {
    __shared__ uint32_t lds[32 * PackedBand2<K>::SD];
    sgm_p2_block<K, PF, CB>(C, base, Lout, g, Kp, P1, P2, dir0, dirStride, sgm_block(xcd), lds);
}

// ------------------------------------------------ persistent work queue
// A ragged batch's scanline blocks differ in length by up to the ratio of
// its largest to smallest image, and a grid in launch order starts the
// longest ones wherever they happen to fall.  k_sgm_queue orders the
// (image, direction, scanline block) items of one SGM kernel longest
// first, deals the images in snake order to nq queues (nq = 8: one per XCD,
// so that an image's directions, which read the same cost rows, share an
// L2; nq = 1: one queue for the device) and clears the queue heads; the
// persistent kernels below then pull items until every queue is empty, a
// workgroup starting on its own XCD's queue and going on to the others.
// Queue layout (ints): [0, 8) heads, [8, 16) item counts,
// [16, 24) first item, [32, 32 + BV) image order, [32 + BV, 32 + 2 BV)
// image offsets within their queue, then the items (image << 16 | j with
// j = block * ndir + direction).
constexpr int SGM_Q_HDR = 32;

__device__ __forceinline__ int sgm_q_len(const Geom& g, int z, int horiz) {
    return horiz ? g.w(z) : g.h(z);
}
__device__ __forceinline__ int sgm_q_items(const Geom& g, int z, int horiz, int ndir) {
    return ((horiz ? g.h(z) : g.w(z)) + 31) / 32 * ndir;
}

// Queue of the image at rank s: dealt in snake order (0..7, 7..0, ...), so
// that the longest images of each round of eight do not all go to queue 0.
__device__ __forceinline__ int sgm_q_of(int s, int nq) {
    const int i = s % nq;
    return (s / nq) & 1 ? nq - 1 - i : i;
}

// blockIdx.x: which queue of `qs` (0: vertical/diagonal kernel, 1:
// horizontal kernel), each of qints ints.
__global__ void __launch_bounds__(256) k_sgm_queue(Geom g, int BV, int ndirV, int ndirH, int nq,
                                                   int* __restrict__ qs, long long qints) {
    const int horiz = blockIdx.x;
    const int ndir = horiz ? ndirH : ndirV;
    int* q = qs + horiz * qints;
    int* order = q + SGM_Q_HDR;
    int* off = order + BV;
    int* items = off + BV;
    // rank by scanline length, longest first, ties by image index
    for (int z = threadIdx.x; z < BV; z += blockDim.x) {
        const int lz = sgm_q_len(g, z, horiz);
        int r = 0;
        for (int u = 0; u < BV; ++u) {
            const int lu = sgm_q_len(g, u, horiz);
            r += lu > lz || (lu == lz && u < z);
        }
        order[r] = z;
    }
    __syncthreads();
    __shared__ int cnt[8];
    if (threadIdx.x < 8) {
        int acc = 0;
        for (int s = 0; s < BV && (int)threadIdx.x < nq; ++s) {
            if (sgm_q_of(s, nq) != (int)threadIdx.x) continue;
            const int z = order[s];
            off[z] = acc;
            acc += sgm_q_items(g, z, horiz, ndir);
        }
        cnt[threadIdx.x] = acc;
    }
    __syncthreads();
    if (threadIdx.x < 8) {
        int first = 0;
        for (int k = 0; k < (int)threadIdx.x; ++k) first += cnt[k];
        q[threadIdx.x] = 0;
        q[8 + threadIdx.x] = cnt[threadIdx.x];
        q[16 + threadIdx.x] = first;
    }
    __syncthreads();
    for (int s = threadIdx.x; s < BV; s += blockDim.x) {
        const int z = order[s];
        const int n = sgm_q_items(g, z, horiz, ndir);
        int* dst = items + q[16 + sgm_q_of(s, nq)] + off[z];
        for (int j = 0; j < n; ++j) dst[j] = z << 16 | j;
    }
}

// The XCD this workgroup runs on (hardware register XCC_ID, bits 0-3).
__device__ __forceinline__ int xcd_id() {
    return __builtin_amdgcn_s_getreg((3 << 11) | 20) & 7;
}

// The next item of queue k for this workgroup (one wave), -1 when empty.
__device__ __forceinline__ int sgm_q_pull(int* __restrict__ q, int k) {
    int i = 0;
    if (threadIdx.x == 0) i = atomicAdd(q + k, 1);
    i = __builtin_amdgcn_readfirstlane(__shfl(i, 0));
    if (i >= q[8 + k]) return -1;
    return q[SGM_Q_HDR + 0] * 0 + i;
}

template <int K, int PF, bool CB = false>
__global__ void __launch_bounds__(64) k_sgm_p2q(const uint8_t* __restrict__ C,
                                                const int* __restrict__ base,
                                                uint8_t* __restrict__ Lout, Geom g, int Kp,
                                                int P1, int P2, int dir0, long long dirStride,
                                                int* __restrict__ q, int BV, int ndir, int nq) {
    __shared__ uint32_t lds[32 * PackedBand2<K>::SD];
    const int* items = q + SGM_Q_HDR + 2 * BV;
    const int home = nq == 8 ? xcd_id() : 0;
    for (int k = 0; k < nq; ++k) {
        const int qk = (home + k) % nq;
        for (;;) {
            const int i = sgm_q_pull(q, qk);
            if (i < 0) break;
            const int it = __builtin_amdgcn_readfirstlane(items[q[16 + qk] + i]);
            const int j = it & 0xFFFF;
            sgm_p2_block<K, PF, CB>(C, base, Lout, g, Kp, P1, P2, dir0, dirStride,
                                    make_uint3(j / ndir, j % ndir, it >> 16), lds);
        }
    }
}

// k_sgm_ph with two lanes per row: 32 rows per wave, the row tiles staged
// through LDS as in k_sgm_ph.
// LDS of one k_sgm_ph2 workgroup
template <int K, int TC, bool CB>
struct Ph2Lds {
    using HT = HorizTile<K, TC>;
    __align__(16) uint32_t tile[32 * HT::RS];
    int btile[CB ? 1 : 32 * HT::BS];
    uint32_t lds[CB ? 1 : 32 * PackedBand2<K>::SD];
};

// One block of 32 rows (bk.x) of direction dir0 + bk.y of image bk.z.
template <int K, int TC, bool CB>
__device__ __forceinline__ void sgm_ph2_block(const uint8_t* __restrict__ C,
                                              const int* __restrict__ base,
                                              uint8_t* __restrict__ Lout, const Geom& g, int Kp,
                                              int P1, int P2, int dir0, long long dirStride,
                                              uint3 bk, Ph2Lds<K, TC, CB>& sh) {
    using PB = PackedBand2<K>;
    using HT = HorizTile<K, TC>;
    constexpr int ROWS = 32;
    const int W = g.w(bk.z), H = g.h(bk.z);
    if ((int)bk.x * ROWS >= H) return;  // whole block below this image
    constexpr int NPH = PB::NPH, NW = PB::NW, SD = PB::SD;
    constexpr int NB = PackedBand<K>::NB;
    constexpr int T = HT::T, RS = HT::RS, BS = HT::BS;
    static_assert(RS % 2 == 0, "8-byte aligned rows");
    static_assert(4 * NW * 2 <= 8 * NB, "a lane pair's words fit the pixel's chunks");
    uint32_t* tile = sh.tile;
    int* btile = sh.btile;
    uint32_t* lds = sh.lds;
    const int dx = c_dx[dir0 + (int)bk.y];
    const int lane = threadIdx.x, h = lane & 1, ln = lane >> 1;
    const int y0 = bk.x * ROWS;
    const int rows = min(ROWS, H - y0);
    u32_alias* sl32 = reinterpret_cast<u32_alias*>(lds + (CB ? 0 : ln * SD));
    u16_alias* s16 = reinterpret_cast<u16_alias*>(sl32);
    if constexpr (!CB)
        for (int i = h; i < SD; i += 2) sl32[i] = 0xFFFFFFFFu;
    const long long plane = g.slot;
    const uint8_t* Cb = C + (long long)bk.z * plane * Kp + (long long)y0 * W * Kp;
    const int* Bb = base + (long long)bk.z * plane + (long long)y0 * W;
    uint8_t* Lb = Lout + (long long)bk.y * dirStride + (long long)bk.z * plane * Kp +
                  (long long)y0 * W * Kp;
    const long long rowBytes = (long long)W * Kp;
    u16x2 L[NPH];
    u16x2 M[CB ? NPH : 1];
    uint32_t mprev = 0;
    int bprev = 0;
    // this lane's words of a pixel: u32 offset 2*NB*col + NW*h in its row
    const uint32_t* mine = tile + ln * RS + NW * h;
    uint32_t* mine_w = tile + ln * RS + NW * h;
    const int* bmine = btile + (CB ? 0 : ln * BS);
    constexpr int CU = T * NB;          // 8-byte cost units of a full row run
    constexpr int TOT = ROWS * CU;      // units of a full tile
    constexpr int TB = ROWS * T;        // band centres of a full tile
    constexpr int NIT = (TOT + 511) / 512, NITB = CB ? 1 : (TB + 255) / 256;
    // the next tile's costs and band centres, loaded into registers while
    // this tile is swept and stored into LDS after it is written out.  The
    // staging covers the (row, unit) grid, clamped onto the last valid
    // element (a duplicate load and an identical LDS write).
    uint2 pv[NIT][8];
    int pb[NITB][4];
    auto fetch = [&](int c0n) {
        const int n = min(T, W - c0n);
        const int xlo = dx > 0 ? c0n : W - c0n - n;
        const int units = n * NB;
        const uint8_t* Cx = Cb + (long long)xlo * Kp;
        const int* Bx = Bb + xlo;
#pragma unroll
        for (int it = 0; it < NIT; ++it)
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int idx = it * 512 + u * 64 + lane;
                const int r = min(idx / CU, rows - 1);
                const int q = min(idx % CU, units - 1);
                pv[it][u] = *reinterpret_cast<const uint2*>(Cx + r * rowBytes + 8 * q);
            }
        if constexpr (!CB) {
#pragma unroll
            for (int it = 0; it < NITB; ++it)
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int idx = it * 256 + u * 64 + lane;
                    const int r = min(idx / T, rows - 1);
                    const int c = min(idx % T, n - 1);
                    pb[it][u] = Bx[(long long)r * W + c];
                }
        }
    };
    fetch(0);
    for (int c0 = 0; c0 < W; c0 += T) {
        const int n = min(T, W - c0);
        const int xlo = dx > 0 ? c0 : W - c0 - n;
        const int units = n * NB;
#pragma unroll
        for (int it = 0; it < NIT; ++it)
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                // the fetch's (row, unit) of this element, recomputed
                const int idx = it * 512 + u * 64 + lane;
                const int r = min(idx / CU, rows - 1);
                const int q = min(idx % CU, units - 1);
                *reinterpret_cast<uint2*>(tile + r * RS + 2 * q) = pv[it][u];
            }
        if constexpr (!CB) {
#pragma unroll
            for (int it = 0; it < NITB; ++it)
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int idx = it * 256 + u * 64 + lane;
                    const int r = min(idx / T, rows - 1);
                    const int c = min(idx % T, n - 1);
                    btile[r * BS + c] = pb[it][u];
                }
        }
        __syncthreads();
        if (c0 + T < W) fetch(c0 + T);
        for (int j = 0; j < n; ++j) {
            const int col = dx > 0 ? j : n - 1 - j;
            uint32_t cw[NW];
#pragma unroll
            for (int q = 0; q < NW; ++q) cw[q] = mine[col * 2 * NB + q];
            uint32_t o[NW];
            if constexpr (CB)
                packed_step2c<K>(L, M, mprev, cw, c0 + j == 0, P1, P2, h, o);
            else
                packed_step2<K>(L, mprev, bprev, sl32, s16, cw, bmine[col], c0 + j == 0, P1, P2,
                                h, o);
#pragma unroll
            for (int q = 0; q < NW; ++q) mine_w[col * 2 * NB + q] = o[q];
        }
        __syncthreads();
        uint8_t* Lx = Lb + (long long)xlo * Kp;
#pragma unroll 1
        for (int it = 0; it < TOT; it += 64 * 8) {
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int idx = it + u * 64 + lane;
                const int r = min(idx / CU, rows - 1);
                const int q = min(idx % CU, units - 1);
                *reinterpret_cast<uint2*>(Lx + r * rowBytes + 8 * q) =
                    *reinterpret_cast<const uint2*>(tile + r * RS + 2 * q);
            }
        }
        __syncthreads();
    }
}

template <int K, int TC, bool CB = false>
__global__ void __launch_bounds__(64) k_sgm_ph2(const uint8_t* __restrict__ C,
                                                const int* __restrict__ base,
                                                uint8_t* __restrict__ Lout, Geom g, int Kp,
                                                int P1, int P2, int dir0, long long dirStride) {
    __shared__ Ph2Lds<K, TC, CB> sh;
    sgm_ph2_block<K, TC, CB>(C, base, Lout, g, Kp, P1, P2, dir0, dirStride,
                             make_uint3(blockIdx.x, blockIdx.y, blockIdx.z), sh);
}

// k_sgm_ph2 as a persistent workgroup over the horizontal kernel's queue
template <int K, int TC, bool CB = false>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) k_sgm_ph2q(const uint8_t* __restrict__ C,
                                                 const int* __restrict__ base,
                                                 uint8_t* __restrict__ Lout, Geom g, int Kp,
                                                 int P1, int P2, int dir0, long long dirStride,
                                                 int* __restrict__ q, int BV, int ndir, int nq) {
    __shared__ Ph2Lds<K, TC, CB> sh;
    const int* items = q + SGM_Q_HDR + 2 * BV;
    const int home = nq == 8 ? xcd_id() : 0;
    for (int k = 0; k < nq; ++k) {
        const int qk = (home + k) % nq;
        for (;;) {
            const int i = sgm_q_pull(q, qk);
            if (i < 0) break;
            const int it = __builtin_amdgcn_readfirstlane(items[q[16 + qk] + i]);
            const int j = it & 0xFFFF;
            sgm_ph2_block<K, TC, CB>(C, base, Lout, g, Kp, P1, P2, dir0, dirStride,
                                     make_uint3(j / ndir, j % ndir, it >> 16), sh);
            __syncthreads();  // the next item restages the tiles
        }
    }
}

// ------------------------------------------------------------------- WTA
// refine_fullres epilogue, stereonet.py:5271-5302: path sums S, first
// minimum, band-edge passthrough, second best outside bk+-1, confidence,
// uniqueness, parabola.  G lanes per pixel, KC entries per lane.  All
// float64 expressions follow the reference's order.
template <typename LT, typename AT, int KC>
__global__ void __launch_bounds__(64) k_wta(const LT* __restrict__ Lv, long long dirStride,
                                            int ndir, int mirrorFrom, const int* __restrict__ base,
                                            const double* __restrict__ d0a,
                                            const double* __restrict__ d0b, Geom g, int K,
                                            int G, int R, double uniq, double* __restrict__ out,
                                            double* __restrict__ conf) {
    const int lane = threadIdx.x;
    const int lpw = 64 / G;
    const int li = lane / G;
    const int j = lane - li * G;
    const int Kp = G * KC;
    const int bv = blockIdx.y;
    const int B = g.B;
    const int HW = g.w(bv) * g.h(bv);
    if ((int)blockIdx.x * lpw >= HW) return;  // whole block past this image
    int p = blockIdx.x * lpw + li;
    const bool active = p < HW;
    if (!active) p = HW - 1;
    const long long pix = (long long)bv * g.slot + p;
    const int kbase = j * KC;
    const int ord = bv >= mirrorFrom ? 1 : 0;
    AT S[KC];
#pragma unroll
    for (int kk = 0; kk < KC; ++kk) S[kk] = 0;
    for (int t = 0; t < ndir; ++t) {
        const int dd = c_order[ord][t];
        AT l[KC];
        const LT* src = Lv + (long long)dd * dirStride + pix * Kp + kbase;
        if constexpr (sizeof(LT) == 1) {
            load_chunk_u8<AT, KC>(reinterpret_cast<const uint8_t*>(src), l);
        } else if constexpr (sizeof(LT) == 2) {
#pragma unroll
            for (int kk = 0; kk < KC; ++kk) l[kk] = (AT)reinterpret_cast<const uint16_t*>(src)[kk];
        } else {
            load_chunk_f64<AT, KC>(reinterpret_cast<const double*>(src), l);
        }
#pragma unroll
        for (int kk = 0; kk < KC; ++kk) S[kk] = S[kk] + l[kk];
    }
    const AT INF = big<AT>();
    AT bv_ = INF;
    int bk = INT_MAX;
#pragma unroll
    for (int kk = 0; kk < KC; ++kk)
        if (kbase + kk < K && S[kk] < bv_) {
            bv_ = S[kk];
            bk = kbase + kk;
        }
    for (int off = G >> 1; off > 0; off >>= 1) {
        const AT ov = __shfl_xor(bv_, off, G);
        const int ok = __shfl_xor(bk, off, G);
        if (ov < bv_ || (ov == bv_ && ok < bk)) {
            bv_ = ov;
            bk = ok;
        }
    }
    AT sv = INF;
    int found = 0;
    AT a = INF, b = INF;
#pragma unroll
    for (int kk = 0; kk < KC; ++kk) {
        const int k = kbase + kk;
        if (k < K && (k < bk - 1 || k > bk + 1)) {
            found = 1;
            if (S[kk] < sv) sv = S[kk];
        }
        if (k == bk - 1) a = S[kk];
        if (k == bk + 1) b = S[kk];
    }
    for (int off = G >> 1; off > 0; off >>= 1) {
        sv = amin(sv, __shfl_xor(sv, off, G));
        a = amin(a, __shfl_xor(a, off, G));
        b = amin(b, __shfl_xor(b, off, G));
        found |= __shfl_xor(found, off, G);
    }
    if (!active || j != 0) return;
    const int v = bv >= B ? 1 : 0;
    const double d0 = (v ? d0b : d0a)[(long long)(bv - v * B) * g.slot + p];
    double o, cf;
    if (bk == 0 || bk == K - 1) {
        o = d0;
        cf = 0.0;
    } else {
        const double bvd = (double)bv_;
        cf = (found && (double)sv < 1.0e29) ? ((double)sv - bvd) / (bvd + 1.0e-6) : 1.0;
        if (cf < uniq) {
            o = d0;
        } else {
            const double ad = (double)a, bd = (double)b;
            const double den = ad - 2.0 * bvd + bd;
            double off = 0.0;
            if (den > 1.0e-9) off = 0.5 * (ad - bd) / den;
            o = (double)(base[pix] - R + bk) + off;
        }
    }
    out[pix] = o;
    if (conf) conf[pix] = cf;
}

// k_wta with G = 1 and K and the path count known at compile time (integer
// path).  The NDIR path-cost chunks of a pixel are all requested before the
// first is summed.
template <int K, typename LT, int NDIR>
__global__ void __launch_bounds__(64) k_wta_k(const LT* __restrict__ Lv, long long dirStride,
                                              int mirrorFrom, const int* __restrict__ base,
                                              const double* __restrict__ d0a,
                                              const double* __restrict__ d0b, Geom g, int Kp,
                                              int R, double uniq, double* __restrict__ out,
                                              double* __restrict__ conf) {
    constexpr int K8 = (K + 7) / 8 * 8;
    constexpr int NU = K8 * (int)sizeof(LT) / 8;  // 8-byte units per chunk
    const int bv = blockIdx.y;
    const int B = g.B;
    const int p = blockIdx.x * 64 + threadIdx.x;
    if (p >= g.w(bv) * g.h(bv)) return;
    const long long pix = (long long)bv * g.slot + p;
    uint2 raw[NDIR][NU];
#pragma unroll
    for (int t = 0; t < NDIR; ++t) {
        const uint2* src = reinterpret_cast<const uint2*>(Lv + (long long)t * dirStride + pix * Kp);
#pragma unroll
        for (int q = 0; q < NU; ++q) raw[t][q] = src[q];
    }
    int S[K8];
#pragma unroll
    for (int k = 0; k < K8; ++k) S[k] = 0;
#pragma unroll
    for (int t = 0; t < NDIR; ++t) {
#pragma unroll
        for (int q = 0; q < NU; ++q) {
            const uint32_t w2[2] = {raw[t][q].x, raw[t][q].y};
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                if constexpr (sizeof(LT) == 1) {
#pragma unroll
                    for (int b = 0; b < 4; ++b) S[8 * q + 4 * h + b] += (w2[h] >> (8 * b)) & 0xFFu;
                } else {
#pragma unroll
                    for (int b = 0; b < 2; ++b) S[4 * q + 2 * h + b] += (w2[h] >> (16 * b)) & 0xFFFFu;
                }
            }
        }
    }
    int bvv = S[0], bk = 0;
#pragma unroll
    for (int k = 1; k < K; ++k)
        if (S[k] < bvv) {
            bvv = S[k];
            bk = k;
        }
    const int v = bv >= B ? 1 : 0;
    const double d0 = (v ? d0b : d0a)[(long long)(bv - v * B) * g.slot + p];
    double o, cf;
    if (bk == 0 || bk == K - 1) {
        o = d0;
        cf = 0.0;
    } else {
        int sv = 0x7FFFFFFF, a = 0, bb = 0;
        bool found = false;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            if (k < bk - 1 || k > bk + 1) {
                found = true;
                sv = min(sv, S[k]);
            }
            if (k == bk - 1) a = S[k];
            if (k == bk + 1) bb = S[k];
        }
        const double bvd = (double)bvv;
        cf = found ? ((double)sv - bvd) / (bvd + 1.0e-6) : 1.0;
        if (cf < uniq) {
            o = d0;
        } else {
            const double ad = (double)a, bd = (double)bb;
            const double den = ad - 2.0 * bvd + bd;
            double off = 0.0;
            if (den > 1.0e-9) off = 0.5 * (ad - bd) / den;
            o = (double)(base[pix] - R + bk) + off;
        }
    }
    out[pix] = o;
    if (conf) conf[pix] = cf;
}


// ------------------------------------------- constant band centre (full range)
// band_costs with the centre R at every pixel (full-range matching,
// eval_real.py:143-146): candidate k is disparity k.  A wave makes the costs
// of 64 pixels of one row: the partner row's census words its candidates
// reach (64 + Kp - 1, clamped into the row) are staged in the wave's LDS
// window, each lane takes its Kp words from there, entries k in [K, Kp) are
// set to 255 (the soft pads of packed_step2c) and the row run is written
// back as contiguous 8-byte units through the wave's LDS stage.
template <bool RIGHT>
__global__ void __launch_bounds__(256) k_costs_cb(const uint64_t* __restrict__ cA,
                                                  const uint64_t* __restrict__ cB, Geom g, int K,
                                                  const int* __restrict__ rtab, int Kp,
                                                  int invalid, uint8_t* __restrict__ C) {
    extern __shared__ __align__(16) uint32_t cb_lds[];
    const int lane = threadIdx.x;
    const int nwin = 64 + Kp - 1;
    const int wave_u32 = 64 * (Kp / 4) + 2 * ((nwin + 1) & ~1);
    uint32_t* cstage = cb_lds + threadIdx.y * wave_u32;
    uint64_t* win = reinterpret_cast<uint64_t*>(cstage + 64 * (Kp / 4));
    const int x0 = blockIdx.x * 64;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    const int W = g.w(blockIdx.z);
    if (x0 >= W || y >= g.h(blockIdx.z)) return;  // whole wave outside this image
    if (rtab) K = 2 * rtab[blockIdx.z] + 1;        // this pair's own band length
    const long long row0 = (long long)blockIdx.z * g.slot + (long long)y * W;
    const uint64_t* rowB = cB + row0;
    // window element i is partner column w0 + i
    const int w0 = RIGHT ? x0 : x0 - (Kp - 1);
    for (int i = lane; i < nwin; i += 64) win[i] = rowB[min(max(w0 + i, 0), W - 1)];
    __builtin_amdgcn_wave_barrier();
    const int x = x0 + lane;
    if (x < W) {
        const uint64_t a = cA[row0 + x];
        for (int k0 = 0; k0 < Kp; k0 += 4) {
            uint32_t packed = 0;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int k = k0 + j;
                const int xr = RIGHT ? x + k : x - k;
                uint32_t v = 255u;
                if (k < K) {
                    const uint64_t w = win[RIGHT ? lane + k : lane + Kp - 1 - k];
                    v = (xr >= 0 && xr < W) ? (uint32_t)__popcll(a ^ w) : (uint32_t)invalid;
                }
                packed |= v << (8 * j);
            }
            cstage[lane * (Kp / 4) + k0 / 4] = packed;
        }
    }
    __builtin_amdgcn_wave_barrier();
    const int units = (min(64, W - x0) * Kp) / 8;
    const uint2* src = reinterpret_cast<const uint2*>(cstage);
    uint2* dst = reinterpret_cast<uint2*>(C + (row0 + x0) * Kp);
    for (int u = lane; u < units; u += 64) dst[u] = src[u];
}

// The refine_fullres epilogue (stereonet.py:5271-5302) on u8 path costs of
// KP bytes per pixel with coalesced reads.  A wave takes 64 consecutive
// pixels of one image; for each path their 64 * KP bytes are one contiguous
// run, which lane l reads as the 4-byte words l, l + 64, ...  Each lane sums
// its words over the paths as packed u16 (at most 8 * 255); the sums go
// through LDS so that lane l then holds pixel l's KP sums, and the
// selection follows k_wta_k: first minimum over the K real entries,
// band-edge passthrough (disp0; the centre R itself for constant centres),
// second best outside bk +- 1, confidence, uniqueness, parabola, output
// base - R + bk + off (CB: base = R, so bk + off).  Integer sums do not
// depend on the order of the paths, so the mirrored view needs no
// reordering.
template <int KP, int NDIR, bool CB>
__global__ void __launch_bounds__(64) k_wta_lds(const uint8_t* __restrict__ Lv, long long dirStride,
                                                const int* __restrict__ base,
                                                const double* __restrict__ d0a,
                                                const double* __restrict__ d0b, Geom g, int K,
                                                const int* __restrict__ rtab, int R, double uniq,
                                                double* __restrict__ out,
                                                double* __restrict__ conf) {
    static_assert(KP % 4 == 0, "whole words per pixel");
    constexpr int NWD = KP / 4;          // words per pixel and path
    constexpr int NPR = KP / 2;          // u16 pairs per pixel
    constexpr int SP = NPR | 1;          // LDS row stride in pairs (odd: no bank conflicts)
    __shared__ uint32_t sums[64 * SP];
    const int bv = blockIdx.y;
    const int HW = g.w(bv) * g.h(bv);
    const int p0 = blockIdx.x * 64;
    if (p0 >= HW) return;
    if (CB && rtab) {  // this pair's own centre and band length
        R = rtab[bv % g.B];
        K = 2 * R + 1;
    }
    const int np = min(64, HW - p0);
    const int nwords = np * NWD;
    const int lane = threadIdx.x;
    const long long pix0 = (long long)bv * g.slot + p0;
    u16x2 S[2 * NWD];
#pragma unroll
    for (int j = 0; j < 2 * NWD; ++j) S[j] = (u16x2){0, 0};
#pragma unroll
    for (int t0 = 0; t0 < NDIR; t0 += 4) {
        uint32_t raw[4][NWD];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const uint32_t* src = reinterpret_cast<const uint32_t*>(Lv + (long long)(t0 + t) * dirStride + pix0 * KP);
#pragma unroll
            for (int i = 0; i < NWD; ++i) {
                const int w = lane + 64 * i;
                raw[t][i] = w < nwords ? src[w] : 0u;
            }
        }
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int i = 0; i < NWD; ++i) {
                S[2 * i] += as_u16x2(__builtin_amdgcn_perm(0u, raw[t][i], 0x0c010c00u));
                S[2 * i + 1] += as_u16x2(__builtin_amdgcn_perm(0u, raw[t][i], 0x0c030c02u));
            }
    }
#pragma unroll
    for (int i = 0; i < NWD; ++i) {
        const int w = lane + 64 * i;
        if (w < nwords) {
            const int pw = w / NWD;
            const int j = 2 * (w - pw * NWD);
            sums[pw * SP + j] = as_u32(S[2 * i]);
            sums[pw * SP + j + 1] = as_u32(S[2 * i + 1]);
        }
    }
    __syncthreads();
    if (lane >= np) return;
    u16x2 T[NPR];
#pragma unroll
    for (int j = 0; j < NPR; ++j) T[j] = as_u16x2(sums[lane * SP + j]);
#pragma unroll
    for (int j = 0; j < NPR; ++j) {
        if (2 * j >= K) T[j].x = 0xFFFF;
        if (2 * j + 1 >= K) T[j].y = 0xFFFF;
    }
    u16x2 mm = T[0];
#pragma unroll
    for (int j = 1; j < NPR; ++j) mm = __builtin_elementwise_min(mm, T[j]);
    const int bvv = min((int)mm.x, (int)mm.y);
    int bk = KP;
#pragma unroll
    for (int j = NPR - 1; j >= 0; --j) {
        if ((int)T[j].y == bvv) bk = 2 * j + 1;
        if ((int)T[j].x == bvv) bk = 2 * j;
    }
    const long long pix = pix0 + lane;
    double o, cf;
    const int v = bv >= g.B ? 1 : 0;
    auto disp0 = [&]() {
        return CB ? (double)R : (v ? d0b : d0a)[(long long)(bv - v * g.B) * g.slot + p0 + lane];
    };
    if (bk == 0 || bk == K - 1) {
        o = disp0();
        cf = 0.0;
    } else {
        int sv = 0x7FFFFFFF, a = 0, bb = 0;
        bool found = false;
#pragma unroll
        for (int j = 0; j < NPR; ++j) {
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                const int k = 2 * j + e;
                const int val = e ? (int)T[j].y : (int)T[j].x;
                if (k < K && (k < bk - 1 || k > bk + 1)) {
                    found = true;
                    sv = min(sv, val);
                }
                if (k == bk - 1) a = val;
                if (k == bk + 1) bb = val;
            }
        }
        const double bvd = (double)bvv;
        cf = found ? ((double)sv - bvd) / (bvd + 1.0e-6) : 1.0;
        if (cf < uniq) {
            o = disp0();
        } else {
            const double ad = (double)a, bd = (double)bb;
            const double den = ad - 2.0 * bvd + bd;
            double off = 0.0;
            if (den > 1.0e-9) off = 0.5 * (ad - bd) / den;
            o = (double)((CB ? R : base[pix]) - R + bk) + off;
        }
    }
    out[pix] = o;
    if (conf) conf[pix] = cf;
}

// a[i] = v, or the pair's own centre rtab[i / slot] when rtab is given
__global__ void k_fill_f64(double* __restrict__ a, long long n, double v,
                           const int* __restrict__ rtab, long long slot) {
    const long long i = blockIdx.x * 256LL + threadIdx.x;
    if (i < n) a[i] = rtab ? (double)rtab[i / slot] : v;
}

// ------------------------------------------------------------ LR check
// int(x - d + 0.5) in [0, W)  <=>  -1 < x - d + 0.5 < W (truncation).
__device__ __forceinline__ int warp_col(int x, double d, int W) {
    const double t = (double)x - d + 0.5;
    if (!(t > -1.0) || !(t < (double)W)) return -1;
    return (int)t;
}

// left_to_right, stereonet.py:5317-5340: one thread per row.  A z-buffer
// forward warp (strict >, init -1), then each run of holes (value < 0) takes
// min(nearest non-hole on the left, on the right), 0.0 if neither is below
// 1e8 -- the closed form of the reference's in-place cascade.
__global__ void k_left_to_right(const double* __restrict__ dl, Geom g, double* __restrict__ dr) {
    const int y = blockIdx.x * 64 + threadIdx.x;
    const int b = blockIdx.y;
    const int W = g.w(b), H = g.h(b);
    if (y >= H) return;
    const long long r = (long long)b * g.slot + (long long)y * W;
    for (int x = 0; x < W; ++x) dr[r + x] = -1.0;
    for (int x = 0; x < W; ++x) {
        const double d = dl[r + x];
        const int xr = warp_col(x, d, W);
        if (xr >= 0 && d > dr[r + xr]) dr[r + xr] = d;
    }
    int x = 0;
    while (x < W) {
        if (!(dr[r + x] < 0.0)) {
            ++x;
            continue;
        }
        int e = x;
        while (e < W && dr[r + e] < 0.0) ++e;
        const double va = x > 0 ? dr[r + x - 1] : 1.0e9;
        const double vb = e < W ? dr[r + e] : 1.0e9;
        double v = va < vb ? va : vb;
        v = v < 1.0e8 ? v : 0.0;
        for (int t = x; t < e; ++t) dr[r + t] = v;
        x = e;
    }
}

// lr_refine check and fills, stereonet.py:5358-5390: one thread per row.
__global__ void k_lr_check_fill(const double* __restrict__ dl, const double* __restrict__ dr,
                                const double* __restrict__ disp0, Geom g, double tol,
                                double* __restrict__ out, uint8_t* __restrict__ status) {
    const int y = blockIdx.x * 64 + threadIdx.x;
    const int b = blockIdx.y;
    const int W = g.w(b), H = g.h(b);
    if (y >= H) return;
    const long long r = (long long)b * g.slot + (long long)y * W;
    for (int x = 0; x < W; ++x) {
        const double d = dl[r + x];
        const int xr = warp_col(x, d, W);
        uint8_t st = 0;
        if (xr < 0) {
            st = 1;
        } else {
            const double e = dr[r + xr];
            if (!(fabs(d - e) <= tol)) st = e > d ? 1 : 2;
        }
        status[r + x] = st;
        out[r + x] = st == 2 ? disp0[r + x] : d;
    }
    int x = 0;
    while (x < W) {
        if (status[r + x] != 1) {
            ++x;
            continue;
        }
        int e = x;
        while (e < W && status[r + e] == 1) ++e;
        const double va = x > 0 ? out[r + x - 1] : 1.0e9;
        const double vb = e < W ? out[r + e] : 1.0e9;
        const double v = va < vb ? va : vb;
        for (int t = x; t < e; ++t) out[r + t] = v < 1.0e8 ? v : disp0[r + t];
        x = e;
    }
}

// ---------------------------------------------------- row-parallel LR
// One 256-thread block per image row.  nearest_valid() gives, for every
// column x, the nearest column at or left of x whose flag is set (-1 if
// none) and the nearest at or right of x (n if none), by per-thread chunk
// scans plus a block-wide prefix max / suffix min.
constexpr int ROW_THREADS = 256;

__device__ void nearest_valid(const uint8_t* __restrict__ flag, int n, int* __restrict__ lidx,
                              int* __restrict__ ridx, int* __restrict__ sc) {
    const int t = threadIdx.x;
    const int chunk = (n + ROW_THREADS - 1) / ROW_THREADS;
    const int a = min(t * chunk, n), b = min(a + chunk, n);
    int last = -1, first = n;
    for (int x = a; x < b; ++x)
        if (flag[x]) {
            if (first == n) first = x;
            last = x;
        }
    int* sl = sc;
    int* sr = sc + ROW_THREADS;
    sl[t] = last;
    sr[t] = first;
    __syncthreads();
    // inclusive prefix max of sl, suffix min of sr (Hillis-Steele)
    for (int off = 1; off < ROW_THREADS; off <<= 1) {
        const int vl = t >= off ? sl[t - off] : -1;
        const int vr = t + off < ROW_THREADS ? sr[t + off] : n;
        __syncthreads();
        sl[t] = max(sl[t], vl);
        sr[t] = min(sr[t], vr);
        __syncthreads();
    }
    int run = t > 0 ? sl[t - 1] : -1;
    for (int x = a; x < b; ++x) {
        if (flag[x]) run = x;
        lidx[x] = run;
    }
    run = t + 1 < ROW_THREADS ? sr[t + 1] : n;
    for (int x = b - 1; x >= a; --x) {
        if (flag[x]) run = x;
        ridx[x] = run;
    }
    __syncthreads();
}

__device__ __forceinline__ unsigned long long dkey(double d) {
    const unsigned long long u = (unsigned long long)__double_as_longlong(d);
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}
__device__ __forceinline__ double dval(unsigned long long k) {
    const unsigned long long u = (k >> 63) ? (k & 0x7FFFFFFFFFFFFFFFull) : ~k;
    return __longlong_as_double((long long)u);
}

// left_to_right, stereonet.py:5317-5340.  The z-buffer keeps the largest
// disparity landing on a column (and -1 where none does), which a max over
// order-preserving keys computes exactly in any order; holes (< 0) then
// take min(nearest non-hole left, right), 0.0 if neither is below 1e8.
__global__ void __launch_bounds__(ROW_THREADS) k_left_to_right_row(const double* __restrict__ dl,
                                                                   Geom g,
                                                                   double* __restrict__ dr) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    unsigned long long* key = reinterpret_cast<unsigned long long*>(smem_raw);
    int* lidx = reinterpret_cast<int*>(key + g.W);
    int* ridx = lidx + g.W;
    int* sc = ridx + g.W;
    uint8_t* flag = reinterpret_cast<uint8_t*>(sc + 2 * ROW_THREADS);
    const int y = blockIdx.x, b = blockIdx.y;
    const int W = g.w(b);
    if (y >= g.h(b)) return;  // whole block below this image
    const long long r = (long long)b * g.slot + (long long)y * W;
    const unsigned long long kinit = dkey(-1.0);
    for (int x = threadIdx.x; x < W; x += ROW_THREADS) key[x] = kinit;
    __syncthreads();
    for (int x = threadIdx.x; x < W; x += ROW_THREADS) {
        const double d = dl[r + x];
        const int xr = warp_col(x, d, W);
        if (xr >= 0 && d > -1.0) atomicMax(&key[xr], dkey(d));
    }
    __syncthreads();
    for (int x = threadIdx.x; x < W; x += ROW_THREADS) flag[x] = dval(key[x]) < 0.0 ? 0 : 1;
    __syncthreads();
    nearest_valid(flag, W, lidx, ridx, sc);
    for (int x = threadIdx.x; x < W; x += ROW_THREADS) {
        double v = dval(key[x]);
        if (!flag[x]) {
            const double va = lidx[x] >= 0 ? dval(key[lidx[x]]) : 1.0e9;
            const double vb = ridx[x] < W ? dval(key[ridx[x]]) : 1.0e9;
            v = va < vb ? va : vb;
            v = v < 1.0e8 ? v : 0.0;
        }
        dr[r + x] = v;
    }
}

// lr_refine's check and fills, stereonet.py:5358-5390, one block per row.
__global__ void __launch_bounds__(ROW_THREADS) k_lr_check_fill_row(
    const double* __restrict__ dl, const double* __restrict__ dr, const double* __restrict__ disp0,
    Geom g, double tol, double* __restrict__ out, uint8_t* __restrict__ status) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    double* o = reinterpret_cast<double*>(smem_raw);
    int* lidx = reinterpret_cast<int*>(o + g.W);
    int* ridx = lidx + g.W;
    int* sc = ridx + g.W;
    uint8_t* flag = reinterpret_cast<uint8_t*>(sc + 2 * ROW_THREADS);
    const int y = blockIdx.x, b = blockIdx.y;
    const int W = g.w(b);
    if (y >= g.h(b)) return;  // whole block below this image
    const long long r = (long long)b * g.slot + (long long)y * W;
    for (int x = threadIdx.x; x < W; x += ROW_THREADS) {
        const double d = dl[r + x];
        const int xr = warp_col(x, d, W);
        uint8_t st = 0;
        if (xr < 0) {
            st = 1;
        } else {
            const double e = dr[r + xr];
            if (!(fabs(d - e) <= tol)) st = e > d ? 1 : 2;
        }
        status[r + x] = st;
        o[x] = st == 2 ? disp0[r + x] : d;
        flag[x] = st != 1;
    }
    __syncthreads();
    nearest_valid(flag, W, lidx, ridx, sc);
    for (int x = threadIdx.x; x < W; x += ROW_THREADS) {
        double v = o[x];
        if (!flag[x]) {
            const double va = lidx[x] >= 0 ? o[lidx[x]] : 1.0e9;
            const double vb = ridx[x] < W ? o[ridx[x]] : 1.0e9;
            const double m = va < vb ? va : vb;
            v = m < 1.0e8 ? m : disp0[r + x];
        }
        out[r + x] = v;
    }
}

inline size_t row_smem(int W) { return (size_t)W * (8 + 4 + 4 + 1) + 2 * ROW_THREADS * 4 + 16; }
// rows wider than this use the one-thread-per-row kernels above
constexpr int ROW_PAR_MAX_W = 3500;

void launch_l2r(const double* dl, const Geom& g, double* dr, hipStream_t st) {
    if (g.W <= ROW_PAR_MAX_W)
        hipLaunchKernelGGL(k_left_to_right_row, dim3(g.H, g.B), dim3(ROW_THREADS), row_smem(g.W), st,
                           dl, g, dr);
    else
        hipLaunchKernelGGL(k_left_to_right, dim3((g.H + 63) / 64, g.B), dim3(64), 0, st, dl, g, dr);
}

void launch_lr_check(const double* dl, const double* dr, const double* d0, const Geom& g,
                     double tol, double* out, uint8_t* status, hipStream_t st) {
    if (g.W <= ROW_PAR_MAX_W)
        hipLaunchKernelGGL(k_lr_check_fill_row, dim3(g.H, g.B), dim3(ROW_THREADS), row_smem(g.W), st,
                           dl, dr, d0, g, tol, out, status);
    else
        hipLaunchKernelGGL(k_lr_check_fill, dim3((g.H + 63) / 64, g.B), dim3(64), 0, st, dl, dr, d0,
                           g, tol, out, status);
}

// ------------------------------------------------------------- median
// median_filter, stereonet.py:2231-2258: clamped (2R+1)^2 window, ascending
// order, element n//2.  The window is padded with +inf to a fixed size and
// sorted by an odd-even transposition network in registers.
template <int RAD>
__global__ void __launch_bounds__(64) k_median(const double* __restrict__ in, Geom g,
                                               double* __restrict__ out) {
    constexpr int N = (2 * RAD + 1) * (2 * RAD + 1);
    const int x = blockIdx.x * 64 + threadIdx.x;
    const int y = blockIdx.y;
    const int W = g.w(blockIdx.z), H = g.h(blockIdx.z);
    if (x >= W || y >= H) return;
    const long long plane = g.slot;
    const double* im = in + (long long)blockIdx.z * plane;
    const int y0 = max(y - RAD, 0), y1 = min(y + RAD, H - 1);
    const int x0 = max(x - RAD, 0), x1 = min(x + RAD, W - 1);
    double v[N];
    int n = 0;
#pragma unroll
    for (int dy = 0; dy < 2 * RAD + 1; ++dy)
#pragma unroll
        for (int dx = 0; dx < 2 * RAD + 1; ++dx) {
            const int yy = y0 + dy, xx = x0 + dx;
            const bool in_ = yy <= y1 && xx <= x1;
            v[dy * (2 * RAD + 1) + dx] = in_ ? im[(long long)yy * W + xx] : HUGE_VAL;
            n += in_ ? 1 : 0;
        }
#pragma unroll
    for (int pass = 0; pass < N; ++pass)
#pragma unroll
        for (int i = pass & 1; i + 1 < N; i += 2) {
            const double a = v[i], b = v[i + 1];
            const bool sw = b < a;
            v[i] = sw ? b : a;
            v[i + 1] = sw ? a : b;
        }
    double m = v[0];
#pragma unroll
    for (int i = 1; i < N; ++i)
        if (i == n / 2) m = v[i];
    out[(long long)blockIdx.z * plane + (long long)y * W + x] = m;
}

__global__ void k_add_into(const double* __restrict__ L, int Kp, int K, long long n,
                           double* __restrict__ S) {
    const long long i = blockIdx.x * 256LL + threadIdx.x;
    if (i >= n) return;
    const long long p = i / K;
    const int k = (int)(i - p * K);
    S[i] = S[i] + L[p * Kp + k];
}

__global__ void k_pad_costs(const double* __restrict__ C, int Kp, int K, long long n,
                            double* __restrict__ Cp) {
    const long long i = blockIdx.x * 256LL + threadIdx.x;
    if (i >= n) return;
    const long long p = i / Kp;
    const int k = (int)(i - p * Kp);
    Cp[i] = k < K ? C[p * K + k] : 0.0;
}

// ============================================================ host side

// Opt-in stage timing: when enabled, each refinement stage records a pair
// of HIP events around its launches on the caller's stream; sv_profile_read
// sums the elapsed times once the work has completed.  Disabled by default
// (then nothing is recorded).
struct ProfPair {
    int stage;
    hipEvent_t a, b;
};
enum { ST_CENSUS, ST_COSTS, ST_SGM, ST_WTA, ST_LR, ST_MEDIAN, ST_COUNT };
static const char* kStageNames[ST_COUNT] = {"census", "costs", "sgm", "wta", "lr", "median"};
static bool g_prof_on = false;
// set for the duration of an allocation-free call on this thread: no stage
// events are created and the SGM paths never fork onto the side stream
static thread_local bool t_noalloc = false;
static ProfPair* g_prof = nullptr;
static int g_prof_n = 0, g_prof_cap = 0;

struct StageTimer {
    hipEvent_t b = nullptr;
    int idx = -1;
    hipStream_t st;
    StageTimer(int stage, hipStream_t s) : st(s) {
        if (!g_prof_on || t_noalloc) return;
        if (g_prof_n == g_prof_cap) {
            const int cap = g_prof_cap ? 2 * g_prof_cap : 256;
            ProfPair* np = (ProfPair*)realloc(g_prof, sizeof(ProfPair) * cap);
            if (!np) return;
            g_prof = np;
            g_prof_cap = cap;
        }
        ProfPair& pp = g_prof[g_prof_n];
        if (hipEventCreate(&pp.a) != hipSuccess) return;
        if (hipEventCreate(&pp.b) != hipSuccess) {
            (void)hipEventDestroy(pp.a);
            return;
        }
        pp.stage = stage;
        idx = g_prof_n++;
        (void)hipEventRecord(pp.a, st);
    }
    void stop() {
        if (idx >= 0) (void)hipEventRecord(g_prof[idx].b, st);
        idx = -1;
    }
    ~StageTimer() { stop(); }
};

struct Plan {
    int K, G, KC, Kp;
    int nbits;
    bool integer;  // integer path
    int lt;        // path-cost element bytes: 1, 2 or 8
    double invalid;
    int ndir;
};

int choose_layout(int K, bool f64, int* G, int* KC) {
    static const int kc_int[] = {8, 16, 24, 32};
    static const int kc_f64[] = {8, 16};
    const int* kcs = f64 ? kc_f64 : kc_int;
    const int nk = f64 ? 2 : 4;
    int best = INT_MAX;
    for (int g = 1; g <= 64; g <<= 1)
        for (int t = 0; t < nk; ++t) {
            const int kp = g * kcs[t];
            if (kp >= K && kp < best) {
                best = kp;
                *G = g;
                *KC = kcs[t];
            }
        }
    if (best == INT_MAX) return sv::set_error(SV_EUNSUPPORTED, "band of %d entries is wider than the GPU path supports", K);
    return SV_OK;
}

int make_plan(int W, int H, int R, double P1, double P2, double ad_weight, double prior_w,
              double ad_trunc, double prior_t, int paths, int rx, int ry, Plan* pl) {
    SV_REQUIRE(W > 0 && H > 0, "image size must be positive (W=%d, H=%d)", W, H);
    SV_REQUIRE(R >= 0, "band half-width R must be >= 0 (R=%d)", R);
    SV_REQUIRE(rx >= 0 && ry >= 0, "census radii must be >= 0");
    pl->nbits = (2 * rx + 1) * (2 * ry + 1) - 1;
    if (pl->nbits > 64)
        return sv::set_error(SV_EUNSUPPORTED, "census window of %d bits exceeds 64", pl->nbits);
    pl->K = 2 * R + 1;
    pl->ndir = paths == 8 ? 8 : 4;
    pl->invalid = (double)pl->nbits + ad_weight * ad_trunc + prior_w * prior_t + 4.0;
    // integer path: every path cost L <= invalid + P2 must fit u16 storage
    const bool integral = P1 == std::floor(P1) && P2 == std::floor(P2) && P1 >= 0 && P2 >= 0 &&
                          P1 <= 65535 && pl->invalid + P2 <= 65535.0;
    pl->integer = ad_weight == 0.0 && prior_w == 0.0 && integral;
    if (pl->integer) {
        const double lmax = pl->invalid + P2;  // L <= C + P2
        pl->lt = lmax <= 255.0 ? 1 : 2;
    } else {
        pl->lt = 8;
    }
    return choose_layout(pl->K, !pl->integer, &pl->G, &pl->KC);
}

long long sgm_queue_ints(int BV, int W, int H);

struct RefineBufs {
    uint64_t* cen;   // [2][B][HW]
    int* base;       // [V][B][HW]
    void* C;         // [V][B][HW][Kp]
    void* L;         // [ndir][V][B][HW][Kp]
    double* d0r;     // [B][HW]            (lr only)
    double* disp;    // [V][B][HW]         (lr only: dl, dr)
    double* tmp;     // [B][HW]            (median)
    int* queue;      // 2 persistent SGM queues (sgm_queue_ints each)
    size_t bytes;
};

// W, H: bounds of every image (the SGM queues are sized by them)
RefineBufs carve(void* ws, const Plan& pl, int V, int B, long long slot, bool lr, bool median,
                 int W, int H) {
    sv::Carver cv(ws);
    const size_t hw = (size_t)slot;
    RefineBufs r;
    r.cen = cv.take<uint64_t>(2 * (size_t)B * hw);
    r.base = cv.take<int>((size_t)V * B * hw);
    const size_t csz = pl.integer ? 1 : 8;
    r.C = cv.take<unsigned char>((size_t)V * B * hw * pl.Kp * csz);
    r.L = cv.take<unsigned char>((size_t)pl.ndir * V * B * hw * pl.Kp * pl.lt);
    r.d0r = lr ? cv.take<double>((size_t)B * hw) : nullptr;
    r.disp = lr ? cv.take<double>((size_t)V * B * hw) : nullptr;
    r.tmp = median ? cv.take<double>((size_t)B * hw) : nullptr;
    r.queue = cv.take<int>(2 * (size_t)sgm_queue_ints(V * B, W, H));
    r.bytes = cv.used;
    return r;
}

template <typename T>
int launch_census(const T* img, int n, const Geom& g, int rx, int ry, uint64_t* out,
                  hipStream_t st) {
    dim3 grid((g.W + 63) / 64, (g.H + CENSUS_ROWS - 1) / CENSUS_ROWS, n);
    const size_t shm = (size_t)(64 + 2 * rx) * (CENSUS_ROWS + 2 * ry) * sizeof(T);
    if (std::is_same<T, uint8_t>::value && rx == 4 && ry == 3) {
        dim3 g8((g.W + 63) / 64, (g.H + C8_ROWS - 1) / C8_ROWS, n);
        hipLaunchKernelGGL(k_census8, g8, dim3(256), 0, st, (const uint8_t*)img, g, out);
    } else if (rx == 4 && ry == 3)
        hipLaunchKernelGGL((k_census<T, 4, 3>), grid, dim3(64, 4), shm, st, img, g, rx, ry, out);
    else
        hipLaunchKernelGGL((k_census<T, -1, -1>), grid, dim3(64, 4), shm, st, img, g, rx, ry, out);
    SV_CHECK_LAUNCH();
    return SV_OK;
}

int census_any(const void* img, int dtype, int n, const Geom& g, int rx, int ry, uint64_t* out,
               hipStream_t st) {
    switch (dtype) {
        case SV_U8: return launch_census((const uint8_t*)img, n, g, rx, ry, out, st);
        case SV_F32: return launch_census((const float*)img, n, g, rx, ry, out, st);
        case SV_F64: return launch_census((const double*)img, n, g, rx, ry, out, st);
    }
    return sv::set_error(SV_EINVAL, "unknown image dtype %d", dtype);
}

int launch_base(const double* d0, long long n, int* base, hipStream_t st) {
    hipLaunchKernelGGL(k_base, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, d0, n, base);
    SV_CHECK_LAUNCH();
    return SV_OK;
}

template <typename T>
int costs_f64_t(bool right, const uint64_t* cA, const uint64_t* cB, const void* gA,
                const void* gB, const int* base, const double* prior, int B, const Geom& g, int R,
                int K, int Kp, double aw, double at, double inv, double pw, double pt, double* C,
                hipStream_t st) {
    dim3 grid((g.W + 63) / 64, g.H, B);
    if (right)
        hipLaunchKernelGGL((k_costs_f64<true, T>), grid, dim3(64), 0, st, cA, cB, (const T*)gA,
                           (const T*)gB, base, prior, g, R, K, Kp, aw, at, inv, pw, pt, C);
    else
        hipLaunchKernelGGL((k_costs_f64<false, T>), grid, dim3(64), 0, st, cA, cB, (const T*)gA,
                           (const T*)gB, base, prior, g, R, K, Kp, aw, at, inv, pw, pt, C);
    SV_CHECK_LAUNCH();
    return SV_OK;
}

// Integer path: the band centres are formed from `prior` (the centre
// estimate) inside the costs kernel and written to base; the float path
// reads base, which launch_base filled.
int launch_costs(const Plan& pl, bool right, const uint64_t* cA, const uint64_t* cB,
                 const void* gA, const void* gB, int dtype, int* base, const double* prior,
                 int B, const Geom& g, int R, double aw, double at, double pw, double pt, void* C,
                 hipStream_t st) {
    if (pl.integer) {
        // four rows per block while their staging fits 64 KB of LDS
        const int rb = 4 * 64 * pl.Kp <= 64 * 1024 ? 4 : 1;
        dim3 grid((g.W + 63) / 64, (g.H + rb - 1) / rb, B);
        const bool k21 = pl.K == 21 && pl.Kp == 24;
        // the fixed-length kernel adds a 256-word partner window per wave
        const size_t shm = (size_t)rb * 64 * pl.Kp + (k21 ? (size_t)rb * 256 * 8 : 0);
        if (right && k21)
            hipLaunchKernelGGL((k_costs_int<true, 21>), grid, dim3(64, rb), shm, st, cA, cB, base,
                               prior, g, R, pl.K, pl.Kp, (int)pl.invalid, (uint8_t*)C);
        else if (k21)
            hipLaunchKernelGGL((k_costs_int<false, 21>), grid, dim3(64, rb), shm, st, cA, cB, base,
                               prior, g, R, pl.K, pl.Kp, (int)pl.invalid, (uint8_t*)C);
        else if (right)
            hipLaunchKernelGGL((k_costs_int<true, 0>), grid, dim3(64, rb), shm, st, cA, cB, base,
                               prior, g, R, pl.K, pl.Kp, (int)pl.invalid, (uint8_t*)C);
        else
            hipLaunchKernelGGL((k_costs_int<false, 0>), grid, dim3(64, rb), shm, st, cA, cB, base,
                               prior, g, R, pl.K, pl.Kp, (int)pl.invalid, (uint8_t*)C);
        SV_CHECK_LAUNCH();
        return SV_OK;
    }
    switch (dtype) {
        case SV_U8:
            return costs_f64_t<uint8_t>(right, cA, cB, gA, gB, base, prior, B, g, R, pl.K,
                                        pl.Kp, aw, at, pl.invalid, pw, pt, (double*)C, st);
        case SV_F32:
            return costs_f64_t<float>(right, cA, cB, gA, gB, base, prior, B, g, R, pl.K, pl.Kp,
                                      aw, at, pl.invalid, pw, pt, (double*)C, st);
        case SV_F64:
            return costs_f64_t<double>(right, cA, cB, gA, gB, base, prior, B, g, R, pl.K,
                                       pl.Kp, aw, at, pl.invalid, pw, pt, (double*)C, st);
    }
    return sv::set_error(SV_EINVAL, "unknown image dtype %d", dtype);
}

// A second stream for work that can overlap the caller's: the horizontal
// SGM paths run there while the vertical and diagonal ones run on the
// caller's stream (fork and join through events, so the caller's stream
// order is kept).  One per (device, caller stream), so callers on different
// streams never wait on each other's work; created on first use and kept
// until the process ends.  Past kSidePerDevice caller streams on one device
// the work stays on the caller's stream.
struct SideStream {
    hipStream_t caller = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
};

constexpr int kSidePerDevice = 16;
std::mutex g_side_mu;
SideStream g_side[64][kSidePerDevice];

// *out = the side stream of caller (null when the table is full)
int side_stream(hipStream_t caller, SideStream** out) {
    *out = nullptr;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess || dev < 0 || dev >= 64)
        return sv::set_error(SV_EHIP, "side stream: no current device");
    for (int i = 0; i < kSidePerDevice; ++i) {
        SideStream& ss = g_side[dev][i];
        if (ss.s && ss.caller == caller) {
            *out = &ss;
            return SV_OK;
        }
        if (!ss.s) {
            if ((e = hipStreamCreateWithFlags(&ss.s, hipStreamNonBlocking)) != hipSuccess ||
                (e = hipEventCreateWithFlags(&ss.fork, hipEventDisableTiming)) != hipSuccess ||
                (e = hipEventCreateWithFlags(&ss.join, hipEventDisableTiming)) != hipSuccess)
                return sv::set_error(SV_EHIP, "side stream: %s", hipGetErrorString(e));
            ss.caller = caller;
            *out = &ss;
            return SV_OK;
        }
    }
    return SV_OK;
}

template <typename CT, typename LT, typename AT, int KC>
int sgm_t(const Plan& pl, const void* C, const int* base, void* L, int BV, const Geom& g,
          double P1, double P2, int dir0, int ndir, hipStream_t st) {
    const int lpw = 64 / pl.G;
    const int nl = g.W > g.H ? g.W : g.H;
    dim3 grid((nl + lpw - 1) / lpw, ndir, BV);
    const size_t shm = (size_t)lpw * (pl.Kp + 5) * sizeof(AT);
    const long long dirStride = (long long)BV * g.slot * pl.Kp;
    hipLaunchKernelGGL((k_sgm<CT, LT, AT, KC>), grid, dim3(64), shm, st, (const CT*)C, base,
                       (LT*)L, g, pl.K, pl.G, (AT)P1, (AT)P2, dir0, dirStride);
    SV_CHECK_LAUNCH();
    return SV_OK;
}

template <typename LT, typename AT, int KC>
int wta_t(const Plan& pl, const void* L, int BV, int mirrorFrom, const int* base,
          const double* d0a, const double* d0b, const Geom& g, int R, double uniq,
          double* out, double* conf, hipStream_t st) {
    const int lpw = 64 / pl.G;
    dim3 grid((unsigned)((g.slot + lpw - 1) / lpw), BV);
    const long long dirStride = (long long)BV * g.slot * pl.Kp;
    hipLaunchKernelGGL((k_wta<LT, AT, KC>), grid, dim3(64), 0, st, (const LT*)L, dirStride,
                       pl.ndir, mirrorFrom, base, d0a, d0b, g, pl.K, pl.G, R, uniq, out, conf);
    SV_CHECK_LAUNCH();
    return SV_OK;
}

#define SV_DISPATCH_KC(KCV, CALL)                                              \
    switch (KCV) {                                                             \
        case 8: { constexpr int KC_ = 8; return CALL; }                        \
        case 16: { constexpr int KC_ = 16; return CALL; }                      \
        case 24: { constexpr int KC_ = 24; return CALL; }                      \
        case 32: { constexpr int KC_ = 32; return CALL; }                      \
    }                                                                          \
    return sv::set_error(SV_EUNSUPPORTED, "band chunk %d", KCV);

// Ints of one persistent-queue area for BV images of at most W x H.
long long sgm_queue_ints(int BV, int W, int H) {
    return SGM_Q_HDR + 2LL * BV + (long long)BV * (((W > H ? W : H) + 31) / 32) * 8;
}

// Workgroups of a persistent kernel that fit the device at once.
template <typename F>
int resident_blocks(F kernel, int* cache) {
    if (*cache > 0) return *cache;
    int dev = 0, cus = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, 64, 0) != hipSuccess)
        return 0;
    *cache = cus * (per > 0 ? per : 1);
    return *cache;
}

template <int K, int PF>
int launch_p2q(const void* C, const int* base, void* L, const Geom& g, int Kp, int P1, int P2,
               int dir0, long long dirStride, int* q, int BV, int ndir, int nq, long long items,
               hipStream_t st) {
    static int cache = 0;
    const int rb = resident_blocks(k_sgm_p2q<K, PF>, &cache);
    if (rb <= 0) return sv::set_error(SV_EHIP, "sgm queue: occupancy query failed");
    const unsigned nb = (unsigned)(items < rb ? items : rb);
    hipLaunchKernelGGL((k_sgm_p2q<K, PF>), dim3(nb), dim3(64), 0, st, (const uint8_t*)C, base,
                       (uint8_t*)L, g, Kp, P1, P2, dir0, dirStride, q, BV, ndir, nq);
    return SV_OK;
}

template <int K, int TC>
int launch_ph2q(const void* C, const int* base, void* L, const Geom& g, int Kp, int P1, int P2,
                int dir0, long long dirStride, int* q, int BV, int ndir, int nq, long long items,
                hipStream_t st) {
    static int cache = 0;
    const int rb = resident_blocks(k_sgm_ph2q<K, TC>, &cache);
    if (rb <= 0) return sv::set_error(SV_EHIP, "sgm queue: occupancy query failed");
    const unsigned nb = (unsigned)(items < rb ? items : rb);
    hipLaunchKernelGGL((k_sgm_ph2q<K, TC>), dim3(nb), dim3(64), 0, st, (const uint8_t*)C, base,
                       (uint8_t*)L, g, Kp, P1, P2, dir0, dirStride, q, BV, ndir, nq);
    return SV_OK;
}

template <int K, typename LT>
int sgm_k_t(const Plan& pl, const void* C, const int* base, void* L, int BV, const Geom& g,
            double P1, double P2, int dir0, int ndir, int* queue, hipStream_t st) {
    const int nl = g.W > g.H ? g.W : g.H;
    dim3 grid((nl + 63) / 64, ndir, BV);
    const long long dirStride = (long long)BV * g.slot * pl.Kp;
    // packed u16 kernels when every value stays below the u16 sentinel:
    // the horizontal directions (0, 1) through the LDS-staged kernel, the
    // others through the per-lane one
    if (sizeof(LT) == 1 && pl.invalid + P2 + P1 < 65000.0 && pl.Kp >= 8 * ((K + 1 + 7) / 8)) {
        const int hEnd = min(dir0 + ndir, 2);
        const int vBeg = max(dir0, 2), vEnd = dir0 + ndir;
        // both kinds of path in this call: horizontal ones on the caller
        // stream's side stream, except while the caller's stream is being
        // captured into a graph (captured work keeps to the capturing
        // stream and runs both kinds in order)
        // persistent workgroups over work queues (tuning knob 12 = 2): the
        // queues are built on the caller's stream before anything forks
        const bool split0 = sv::tune(SV_TUNE_SGM_SPLIT) == 2;
        const int qmode = sv::tune(SV_TUNE_SGM_QUEUE);
        const int nq = qmode == 3 ? 1 : 8;
        // (the queue build ranks the images in O(BV^2) on one block: up to
        // 4096 images, a few hundred microseconds at most)
        const bool useq = queue && split0 && qmode >= 2 && BV <= 4096 &&
                          ((g.W > g.H ? g.W : g.H) + 31) / 32 * 8 <= 0xFFFF;
        const long long qints = sgm_queue_ints(BV, g.W, g.H);
        if (useq) {
            hipLaunchKernelGGL(k_sgm_queue, dim3(2), dim3(256), 0, st, g, BV, max(vEnd - vBeg, 1),
                               max(hEnd - dir0, 1), nq, queue, qints);
            SV_CHECK_LAUNCH();
        }
        bool fork = hEnd > dir0 && vEnd > vBeg && sv::tune(SV_TUNE_SGM_OVERLAP) == 2 && !t_noalloc;
        if (fork) {
            hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
            if (hipStreamIsCapturing(st, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone)
                fork = false;
        }
        std::unique_lock<std::mutex> lock(g_side_mu, std::defer_lock);
        SideStream* ss = nullptr;
        hipStream_t hst = st;
        if (fork) {
            lock.lock();
            int rc = side_stream(st, &ss);
            if (rc) return rc;
            fork = ss != nullptr;
        }
        if (fork) {
            hipError_t e = hipEventRecord(ss->fork, st);
            if (e == hipSuccess) e = hipStreamWaitEvent(ss->s, ss->fork, 0);
            if (e != hipSuccess) return sv::set_error(SV_EHIP, "sgm fork: %s", hipGetErrorString(e));
            hst = ss->s;
        }
        const bool split = sv::tune(SV_TUNE_SGM_SPLIT) == 2;
        int qrc = SV_OK;
        if (hEnd > dir0 && useq) {
            const long long items = (long long)BV * ((g.H + 31) / 32) * (hEnd - dir0);
            int* qh = queue + qints;
            if (sv::tune(SV_TUNE_SGM_HTILE) == 8)
                qrc = launch_ph2q<K, 8>(C, base, L, g, pl.Kp, (int)P1, (int)P2, dir0, dirStride, qh,
                                        BV, hEnd - dir0, nq, items, hst);
            else
                qrc = launch_ph2q<K, 16>(C, base, L, g, pl.Kp, (int)P1, (int)P2, dir0, dirStride, qh,
                                         BV, hEnd - dir0, nq, items, hst);
        } else if (hEnd > dir0 && split) {
            dim3 hgrid((g.H + 31) / 32, hEnd - dir0, BV);
            if (sv::tune(SV_TUNE_SGM_HTILE) == 8)
                hipLaunchKernelGGL((k_sgm_ph2<K, 8>), hgrid, dim3(64), 0, hst, (const uint8_t*)C,
                                   base, (uint8_t*)L, g, pl.Kp, (int)P1, (int)P2, dir0, dirStride);
            else
                hipLaunchKernelGGL((k_sgm_ph2<K, 16>), hgrid, dim3(64), 0, hst, (const uint8_t*)C,
                                   base, (uint8_t*)L, g, pl.Kp, (int)P1, (int)P2, dir0, dirStride);
        } else if (hEnd > dir0) {
            dim3 hgrid((g.H + 63) / 64, hEnd - dir0, BV);
            if (sv::tune(SV_TUNE_SGM_HTILE) == 8)
                hipLaunchKernelGGL((k_sgm_ph<K, 8>), hgrid, dim3(64), 0, hst, (const uint8_t*)C,
                                   base, (uint8_t*)L, g, pl.Kp, (int)P1, (int)P2, dir0, dirStride);
            else
                hipLaunchKernelGGL((k_sgm_ph<K, 16>), hgrid, dim3(64), 0, hst, (const uint8_t*)C,
                                   base, (uint8_t*)L, g, pl.Kp, (int)P1, (int)P2, dir0, dirStride);
        }
        hipError_t lerr = hipGetLastError();
        const int xcd = sv::tune(SV_TUNE_SGM_ORDER) == 2;
        if (vEnd > vBeg && lerr == hipSuccess && qrc == SV_OK && useq) {
            const long long items = (long long)BV * ((g.W + 31) / 32) * (vEnd - vBeg);
            uint8_t* Lv = (uint8_t*)L + (long long)(vBeg - dir0) * dirStride;
            switch (sv::tune(SV_TUNE_SGM_PREFETCH)) {
                case 2:
                    qrc = launch_p2q<K, 2>(C, base, Lv, g, pl.Kp, (int)P1, (int)P2, vBeg, dirStride,
                                           queue, BV, vEnd - vBeg, nq, items, st);
                    break;
                case 8:
                    qrc = launch_p2q<K, 8>(C, base, Lv, g, pl.Kp, (int)P1, (int)P2, vBeg, dirStride,
                                           queue, BV, vEnd - vBeg, nq, items, st);
                    break;
                default:
                    qrc = launch_p2q<K, 4>(C, base, Lv, g, pl.Kp, (int)P1, (int)P2, vBeg, dirStride,
                                           queue, BV, vEnd - vBeg, nq, items, st);
            }
            lerr = hipGetLastError();
        } else if (vEnd > vBeg && lerr == hipSuccess && qrc == SV_OK && split) {
            dim3 vgrid((g.W + 31) / 32, vEnd - vBeg, BV);
            uint8_t* Lv = (uint8_t*)L + (long long)(vBeg - dir0) * dirStride;
            switch (sv::tune(SV_TUNE_SGM_PREFETCH)) {
                case 2:
                    hipLaunchKernelGGL((k_sgm_p2<K, 2>), vgrid, dim3(64), 0, st, (const uint8_t*)C,
                                       base, Lv, g, pl.Kp, (int)P1, (int)P2, vBeg, dirStride, xcd);
                    break;
                case 8:
                    hipLaunchKernelGGL((k_sgm_p2<K, 8>), vgrid, dim3(64), 0, st, (const uint8_t*)C,
                                       base, Lv, g, pl.Kp, (int)P1, (int)P2, vBeg, dirStride, xcd);
                    break;
                default:
                    hipLaunchKernelGGL((k_sgm_p2<K, 4>), vgrid, dim3(64), 0, st, (const uint8_t*)C,
                                       base, Lv, g, pl.Kp, (int)P1, (int)P2, vBeg, dirStride, xcd);
            }
            lerr = hipGetLastError();
        } else if (vEnd > vBeg && lerr == hipSuccess) {
            dim3 vgrid((g.W + 63) / 64, vEnd - vBeg, BV);
            uint8_t* Lv = (uint8_t*)L + (long long)(vBeg - dir0) * dirStride;
            switch (sv::tune(SV_TUNE_SGM_PREFETCH)) {
                case 2:
                    hipLaunchKernelGGL((k_sgm_p<K, 2>), vgrid, dim3(64), 0, st, (const uint8_t*)C,
                                       base, Lv, g, pl.Kp, (int)P1, (int)P2, vBeg, dirStride);
                    break;
                case 8:
                    hipLaunchKernelGGL((k_sgm_p<K, 8>), vgrid, dim3(64), 0, st, (const uint8_t*)C,
                                       base, Lv, g, pl.Kp, (int)P1, (int)P2, vBeg, dirStride);
                    break;
                default:
                    hipLaunchKernelGGL((k_sgm_p<K, 4>), vgrid, dim3(64), 0, st, (const uint8_t*)C,
                                       base, Lv, g, pl.Kp, (int)P1, (int)P2, vBeg, dirStride);
            }
            lerr = hipGetLastError();
        }
        if (fork) {  // join before reporting anything, so the caller's stream never runs ahead
            hipError_t e = hipEventRecord(ss->join, ss->s);
            if (e == hipSuccess) e = hipStreamWaitEvent(st, ss->join, 0);
            if (e != hipSuccess) return sv::set_error(SV_EHIP, "sgm join: %s", hipGetErrorString(e));
        }
        if (lerr != hipSuccess) return sv::set_error(SV_EHIP, "sgm launch: %s", hipGetErrorString(lerr));
        return qrc;
    }
    hipLaunchKernelGGL((k_sgm_k<K, LT>), grid, dim3(64), 0, st, (const uint8_t*)C, base, (LT*)L, g,
                       pl.Kp, (int)P1, (int)P2, dir0, dirStride);
    SV_CHECK_LAUNCH();
    return SV_OK;
}

// band lengths with a compile-time specialisation (G = 1, integer path)
inline bool has_k_specialisation(const Plan& pl) {
    return pl.integer && pl.G == 1 && (pl.K == 21 || pl.K == 7) && pl.Kp >= ((pl.K + 7) / 8) * 8;
}

// queue: sgm_queue_ints(BV, g.W, g.H) * 2 ints for the persistent kernels,
// or nullptr (grid launches only)
int launch_sgm(const Plan& pl, const void* C, const int* base, void* L, int BV, const Geom& g,
               double P1, double P2, int dir0, int ndir, hipStream_t st, int* queue = nullptr) {
    if (has_k_specialisation(pl)) {
        if (pl.K == 21)
            return pl.lt == 1
                       ? sgm_k_t<21, uint8_t>(pl, C, base, L, BV, g, P1, P2, dir0, ndir, queue, st)
                       : sgm_k_t<21, uint16_t>(pl, C, base, L, BV, g, P1, P2, dir0, ndir, queue, st);
        return pl.lt == 1
                   ? sgm_k_t<7, uint8_t>(pl, C, base, L, BV, g, P1, P2, dir0, ndir, queue, st)
                   : sgm_k_t<7, uint16_t>(pl, C, base, L, BV, g, P1, P2, dir0, ndir, queue, st);
    }
    if (!pl.integer) {
        switch (pl.KC) {
            case 8: return sgm_t<double, double, double, 8>(pl, C, base, L, BV, g, P1, P2, dir0, ndir, st);
            case 16: return sgm_t<double, double, double, 16>(pl, C, base, L, BV, g, P1, P2, dir0, ndir, st);
        }
        return sv::set_error(SV_EUNSUPPORTED, "band chunk %d", pl.KC);
    }
    if (pl.lt == 1) {
        SV_DISPATCH_KC(pl.KC, (sgm_t<uint8_t, uint8_t, int, KC_>(pl, C, base, L, BV, g, P1, P2, dir0, ndir, st)));
    }
    SV_DISPATCH_KC(pl.KC, (sgm_t<uint8_t, uint16_t, int, KC_>(pl, C, base, L, BV, g, P1, P2, dir0, ndir, st)));
}

template <int K, typename LT>
int wta_k_t(const Plan& pl, const void* L, int BV, int mirrorFrom, const int* base,
            const double* d0a, const double* d0b, const Geom& g, int R, double uniq,
            double* out, double* conf, hipStream_t st) {
    dim3 grid((unsigned)((g.slot + 63) / 64), BV);
    const long long dirStride = (long long)BV * g.slot * pl.Kp;
    if (pl.ndir == 8)
        hipLaunchKernelGGL((k_wta_k<K, LT, 8>), grid, dim3(64), 0, st, (const LT*)L, dirStride,
                           mirrorFrom, base, d0a, d0b, g, pl.Kp, R, uniq, out, conf);
    else
        hipLaunchKernelGGL((k_wta_k<K, LT, 4>), grid, dim3(64), 0, st, (const LT*)L, dirStride,
                           mirrorFrom, base, d0a, d0b, g, pl.Kp, R, uniq, out, conf);
    SV_CHECK_LAUNCH();
    return SV_OK;
}

int launch_wta(const Plan& pl, const void* L, int BV, int mirrorFrom, const int* base,
               const double* d0a, const double* d0b, const Geom& g, int R, double uniq,
               double* out, double* conf, hipStream_t st) {
    if (has_k_specialisation(pl) && pl.K == 21 && pl.Kp == 24 && pl.lt == 1 &&
        sv::tune(SV_TUNE_WTA_READS) == 2) {
        const long long dirStride = (long long)BV * g.slot * pl.Kp;
        dim3 grid((unsigned)((g.slot + 63) / 64), BV);
        if (pl.ndir == 8)
            hipLaunchKernelGGL((k_wta_lds<24, 8, false>), grid, dim3(64), 0, st, (const uint8_t*)L,
                               dirStride, base, d0a, d0b, g, pl.K, nullptr, R, uniq, out, conf);
        else
            hipLaunchKernelGGL((k_wta_lds<24, 4, false>), grid, dim3(64), 0, st, (const uint8_t*)L,
                               dirStride, base, d0a, d0b, g, pl.K, nullptr, R, uniq, out, conf);
        SV_CHECK_LAUNCH();
        return SV_OK;
    }
    if (has_k_specialisation(pl)) {
        if (pl.K == 21)
            return pl.lt == 1 ? wta_k_t<21, uint8_t>(pl, L, BV, mirrorFrom, base, d0a, d0b, g, R, uniq, out, conf, st)
                              : wta_k_t<21, uint16_t>(pl, L, BV, mirrorFrom, base, d0a, d0b, g, R, uniq, out, conf, st);
        return pl.lt == 1 ? wta_k_t<7, uint8_t>(pl, L, BV, mirrorFrom, base, d0a, d0b, g, R, uniq, out, conf, st)
                          : wta_k_t<7, uint16_t>(pl, L, BV, mirrorFrom, base, d0a, d0b, g, R, uniq, out, conf, st);
    }
    if (!pl.integer) {
        switch (pl.KC) {
            case 8: return wta_t<double, double, 8>(pl, L, BV, mirrorFrom, base, d0a, d0b, g, R, uniq, out, conf, st);
            case 16: return wta_t<double, double, 16>(pl, L, BV, mirrorFrom, base, d0a, d0b, g, R, uniq, out, conf, st);
        }
        return sv::set_error(SV_EUNSUPPORTED, "band chunk %d", pl.KC);
    }
    if (pl.lt == 1) {
        SV_DISPATCH_KC(pl.KC, (wta_t<uint8_t, int, KC_>(pl, L, BV, mirrorFrom, base, d0a, d0b, g, R, uniq, out, conf, st)));
    }
    SV_DISPATCH_KC(pl.KC, (wta_t<uint16_t, int, KC_>(pl, L, BV, mirrorFrom, base, d0a, d0b, g, R, uniq, out, conf, st)));
}

int launch_median(const double* in, const Geom& g, int radius, int passes, double* out,
                  double* tmp, hipStream_t st) {
    SV_REQUIRE(passes >= 0, "passes must be >= 0");
    if (passes == 0) {
        hipError_t e = hipMemcpyAsync(out, in, sizeof(double) * g.B * g.slot, hipMemcpyDeviceToDevice, st);
        if (e != hipSuccess) return sv::set_error(SV_EHIP, "median copy: %s", hipGetErrorString(e));
        return SV_OK;
    }
    if (radius < 0 || radius > 2)
        return sv::set_error(SV_EUNSUPPORTED, "median radius %d (the GPU path supports 0..2)", radius);
    SV_REQUIRE(passes == 1 || tmp, "multi-pass median needs a workspace");
    dim3 grid((g.W + 63) / 64, g.H, g.B);
    const double* src = in;
    for (int it = 0; it < passes; ++it) {
        // the last pass lands in out; earlier passes alternate so that the
        // final one does
        double* dst = ((passes - 1 - it) % 2 == 0) ? out : tmp;
        switch (radius) {
            case 0: hipLaunchKernelGGL(k_median<0>, grid, dim3(64), 0, st, src, g, dst); break;
            case 1: hipLaunchKernelGGL(k_median<1>, grid, dim3(64), 0, st, src, g, dst); break;
            case 2: hipLaunchKernelGGL(k_median<2>, grid, dim3(64), 0, st, src, g, dst); break;
        }
        SV_CHECK_LAUNCH();
        src = dst;
    }
    return SV_OK;
}

Geom uniform_geom(int B, int W, int H) {
    Geom g;
    g.dims = nullptr;
    g.W = W;
    g.H = H;
    g.B = B;
    g.slot = (long long)W * H;
    return g;
}

int check_ws(size_t need, void* ws, size_t have) {
    if (!ws || have < need)
        return sv::set_error(SV_EWORKSPACE, "workspace of %zu bytes needed, %zu given", need, have);
    return SV_OK;
}

// One band refinement of V views over B pairs.  View 0 matches left against
// right with centre disp0; view 1 (lr only) matches right against left with
// centre left_to_right(disp0).  Results land in outs[v*B*HW ...].
int refine_views(const Plan& pl, const RefineBufs& bf, const void* left, const void* right,
                 int dtype, const double* disp0, const double* d0r, int V, const Geom& g,
                 int R, double P1, double P2, double aw, double at, double uniq, int rx, int ry,
                 double pw, double pt, double* outs, double* conf, hipStream_t st) {
    const int B = g.B;
    const size_t hw = (size_t)g.slot;
    int rc;
    uint64_t* cL = bf.cen;
    uint64_t* cR = bf.cen + (size_t)B * hw;
    {
        StageTimer t(ST_CENSUS, st);
        if ((rc = census_any(left, dtype, B, g, rx, ry, cL, st))) return rc;
        if ((rc = census_any(right, dtype, B, g, rx, ry, cR, st))) return rc;
    }
    const size_t csz = pl.integer ? 1 : 8;
    StageTimer tc(ST_COSTS, st);
    for (int v = 0; v < V; ++v) {
        const double* d0 = v == 0 ? disp0 : d0r;
        int* base = bf.base + (size_t)v * B * hw;
        if (!pl.integer && (rc = launch_base(d0, (long long)B * hw, base, st))) return rc;
        void* C = (char*)bf.C + (size_t)v * B * hw * pl.Kp * csz;
        if (v == 0)
            rc = launch_costs(pl, false, cL, cR, left, right, dtype, base, d0, B, g, R, aw, at, pw, pt, C, st);
        else
            rc = launch_costs(pl, true, cR, cL, right, left, dtype, base, d0, B, g, R, aw, at, pw, pt, C, st);
        if (rc) return rc;
    }
    tc.stop();
    {
        StageTimer t(ST_SGM, st);
        if ((rc = launch_sgm(pl, bf.C, bf.base, bf.L, V * B, g, P1, P2, 0, pl.ndir, st, bf.queue)))
            return rc;
    }
    StageTimer t(ST_WTA, st);
    return launch_wta(pl, bf.L, V * B, V == 2 ? B : V * B, bf.base, disp0, d0r, g, R, uniq,
                      outs, conf, st);
}


// ------------------------------------------- constant band centre (full range)
// Band lengths with constant-centre kernels: a band of K entries runs on the
// smallest bucket KB >= K (soft pads in between, see packed_step2c).
constexpr int kCbBuckets[] = {45, 65, 85, 125};

int cb_bucket(int K) {
    for (int kb : kCbBuckets)
        if (K <= kb) return kb;
    return 0;
}

int cb_kp(int kb) {
    switch (kb) {
        case 45: return 8 * PackedBand<45>::NB;
        case 65: return 8 * PackedBand<65>::NB;
        case 85: return 8 * PackedBand<85>::NB;
        case 125: return 8 * PackedBand<125>::NB;
    }
    return 0;
}

struct CbBufs {
    uint64_t* cen;  // [2][B][HW]
    uint8_t* C;     // [V][B][HW][Kp]
    uint8_t* L;     // [ndir][V][B][HW][Kp]
    double* d0;     // [B][HW]   the centre R (lr only: fill value and check input)
    double* d0r;    // [B][HW]   left_to_right(d0) (lr only)
    double* disp;   // [V][B][HW] (lr only)
    size_t bytes;
};

CbBufs carve_cb(void* ws, int Kp, int ndir, int V, int B, long long slot, bool lr) {
    sv::Carver cv(ws);
    const size_t hw = (size_t)slot;
    CbBufs r;
    r.cen = cv.take<uint64_t>(2 * (size_t)B * hw);
    r.C = cv.take<uint8_t>((size_t)V * B * hw * Kp);
    r.L = cv.take<uint8_t>((size_t)ndir * V * B * hw * Kp);
    r.d0 = lr ? cv.take<double>((size_t)B * hw) : nullptr;
    r.d0r = lr ? cv.take<double>((size_t)B * hw) : nullptr;
    r.disp = lr ? cv.take<double>((size_t)V * B * hw) : nullptr;
    r.bytes = cv.used;
    return r;
}

int launch_costs_cb(bool right, const uint64_t* cA, const uint64_t* cB, int B, const Geom& g,
                    int K, const int* rtab, int Kp, int invalid, uint8_t* C, hipStream_t st) {
    const int rb = 4;
    const int nwin = 64 + Kp - 1;
    const size_t wave_u32 = 64 * (size_t)(Kp / 4) + 2 * (size_t)((nwin + 1) & ~1);
    dim3 grid((g.W + 63) / 64, (g.H + rb - 1) / rb, B);
    const size_t shm = rb * wave_u32 * 4;
    if (right)
        hipLaunchKernelGGL((k_costs_cb<true>), grid, dim3(64, rb), shm, st, cA, cB, g, K, rtab, Kp,
                           invalid, C);
    else
        hipLaunchKernelGGL((k_costs_cb<false>), grid, dim3(64, rb), shm, st, cA, cB, g, K, rtab, Kp,
                           invalid, C);
    SV_CHECK_LAUNCH();
    return SV_OK;
}

template <int KB>
int sgm_cb_t(const uint8_t* C, uint8_t* L, int BV, const Geom& g, int Kp, int P1, int P2,
             int ndir, hipStream_t st) {
    const long long dirStride = (long long)BV * g.slot * Kp;
    dim3 hgrid((g.H + 31) / 32, 2, BV);
    hipLaunchKernelGGL((k_sgm_ph2<KB, 8, true>), hgrid, dim3(64), 0, st, C, nullptr, L, g, Kp, P1,
                       P2, 0, dirStride);
    dim3 vgrid((g.W + 31) / 32, ndir - 2, BV);
    hipLaunchKernelGGL((k_sgm_p2<KB, 4, true>), vgrid, dim3(64), 0, st, C, nullptr,
                       L + 2 * dirStride, g, Kp, P1, P2, 2, dirStride,
                       (int)(sv::tune(SV_TUNE_SGM_ORDER) == 2));
    SV_CHECK_LAUNCH();
    return SV_OK;
}

template <int KB>
int wta_cb_t(const uint8_t* L, int BV, int ndir, const Geom& g, int K, const int* rtab, int Kp,
             int R, double* out, double* conf, hipStream_t st) {
    constexpr int KP = 8 * PackedBand<KB>::NB;
    if (Kp != KP) return sv::set_error(SV_EINVAL, "constant-centre bucket %d needs %d-byte rows", KB, KP);
    const long long dirStride = (long long)BV * g.slot * Kp;
    dim3 grid((unsigned)((g.slot + 63) / 64), BV);
    if (ndir == 8)
        hipLaunchKernelGGL((k_wta_lds<KP, 8, true>), grid, dim3(64), 0, st, L, dirStride, nullptr,
                           nullptr, nullptr, g, K, rtab, R, 0.0, out, conf);
    else
        hipLaunchKernelGGL((k_wta_lds<KP, 4, true>), grid, dim3(64), 0, st, L, dirStride, nullptr,
                           nullptr, nullptr, g, K, rtab, R, 0.0, out, conf);
    SV_CHECK_LAUNCH();
    return SV_OK;
}

// V views of B pairs with the constant centre R (or pair n's own centre
// rtab[n], device int32, all at most R): census, costs, the packed
// constant-centre SGM and its epilogue.  View 1 matches right against left.
int refine_views_cb(const CbBufs& bf, const void* left, const void* right, int dtype, int V,
                    const Geom& g, int R, const int* rtab, int kb, int Kp, int ndir, int P1, int P2,
                    int rx, int ry, double* outs, double* conf, hipStream_t st) {
    const int B = g.B;
    const int K = 2 * R + 1;
    const int invalid = (2 * rx + 1) * (2 * ry + 1) - 1 + 4;
    const size_t hw = (size_t)g.slot;
    uint64_t* cL = bf.cen;
    uint64_t* cR = bf.cen + (size_t)B * hw;
    int rc;
    {
        StageTimer t(ST_CENSUS, st);
        if ((rc = census_any(left, dtype, B, g, rx, ry, cL, st))) return rc;
        if ((rc = census_any(right, dtype, B, g, rx, ry, cR, st))) return rc;
    }
    {
        StageTimer t(ST_COSTS, st);
        for (int v = 0; v < V; ++v) {
            uint8_t* C = bf.C + (size_t)v * B * hw * Kp;
            rc = v == 0 ? launch_costs_cb(false, cL, cR, B, g, K, rtab, Kp, invalid, C, st)
                        : launch_costs_cb(true, cR, cL, B, g, K, rtab, Kp, invalid, C, st);
            if (rc) return rc;
        }
    }
    {
        StageTimer t(ST_SGM, st);
        switch (kb) {
            case 45: rc = sgm_cb_t<45>(bf.C, bf.L, V * B, g, Kp, P1, P2, ndir, st); break;
            case 65: rc = sgm_cb_t<65>(bf.C, bf.L, V * B, g, Kp, P1, P2, ndir, st); break;
            case 85: rc = sgm_cb_t<85>(bf.C, bf.L, V * B, g, Kp, P1, P2, ndir, st); break;
            case 125: rc = sgm_cb_t<125>(bf.C, bf.L, V * B, g, Kp, P1, P2, ndir, st); break;
            default: return sv::set_error(SV_EUNSUPPORTED, "no constant-centre bucket %d", kb);
        }
        if (rc) return rc;
    }
    StageTimer t(ST_WTA, st);
    switch (kb) {
        case 45: return wta_cb_t<45>(bf.L, V * B, ndir, g, K, rtab, Kp, R, outs, conf, st);
        case 65: return wta_cb_t<65>(bf.L, V * B, ndir, g, K, rtab, Kp, R, outs, conf, st);
        case 85: return wta_cb_t<85>(bf.L, V * B, ndir, g, K, rtab, Kp, R, outs, conf, st);
        case 125: return wta_cb_t<125>(bf.L, V * B, ndir, g, K, rtab, Kp, R, outs, conf, st);
    }
    return sv::set_error(SV_EUNSUPPORTED, "no constant-centre bucket %d", kb);
}

// Whether the constant-centre kernels serve this full-range call: integer u8
// path (census costs, integral P1/P2 with invalid + P2 <= 255) and a bucket
// for K.  The right view's centre is left_to_right of the constant map,
// again the constant R when W >= R (some pixel lands inside the view).
bool cb_applies(const Plan& pl, int W, int R, bool lr) {
    return pl.integer && pl.lt == 1 && cb_bucket(pl.K) != 0 && (!lr || W >= R);
}

// --------------------------------------------------- allocation-free calls
// The workspace of sv_refine_disparity_noalloc: a file-scope device array
// (each GPU the library runs on holds its own copy until the process ends).
// Host code never uses its name as an address: the device address comes
// from hipGetSymbolAddress, once per device, with that device current.
constexpr size_t kScratchBytes = (size_t)96 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mu;
void* g_scratch_ptr[64];

int scratch_ptr(void** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess || dev < 0 || dev >= 64)
        return sv::set_error(SV_EHIP, "scratch: no current device");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return sv::set_error(SV_EHIP, "scratch: hipGetSymbolAddress: %s", hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return SV_OK;
}

struct NoAllocScope {
    NoAllocScope() { t_noalloc = true; }
    ~NoAllocScope() { t_noalloc = false; }
};

}  // namespace

// =============================================================== C ABI
// -------------------------------------- network left-right consistency
// mirror_h, stereonet.py:4974-4983: `rows` rows of W elements, each reversed
// (every channel of every image of a dense planar buffer).  A block takes
// MIR_ROWS rows, every load of a thread in flight before its stores.
constexpr int MIR_ROWS = 8;

template <typename T>
__global__ void __launch_bounds__(256) k_mirror_h(const T* __restrict__ in, long long rows, int W,
                                                  T* __restrict__ out) {
    for (long long r0 = (long long)blockIdx.x * MIR_ROWS; r0 < rows;
         r0 += (long long)gridDim.x * MIR_ROWS) {
        for (int x0 = 0; x0 < W; x0 += 256) {
            const int x = x0 + threadIdx.x;
            T v[MIR_ROWS];
#pragma unroll
            for (int k = 0; k < MIR_ROWS; ++k)
                if (x < W && r0 + k < rows) v[k] = in[(r0 + k) * W + (W - 1 - x)];
#pragma unroll
            for (int k = 0; k < MIR_ROWS; ++k)
                if (x < W && r0 + k < rows) out[(r0 + k) * W + x] = v[k];
        }
    }
}

// lr_consistency, stereonet.py:4986-5008, in float64 from T maps: valid = 0
// where x - d leaves [0, W-1] (or is NaN, where the reference's int()
// raises) or |d - dr(x - d)| > tol, dr interpolated linearly between
// columns int(x - d) and the next (clamped to the row).
template <typename T>
__global__ void k_lr_consistency(const T* __restrict__ dl, const T* __restrict__ dr, int W, int H,
                                 double tol, uint8_t* __restrict__ valid) {
    const int b = blockIdx.z, y = blockIdx.y;
    const int x = blockIdx.x * 256 + threadIdx.x;
    if (x >= W) return;
    const long long row = ((long long)b * H + y) * W;
    const double d = (double)dl[row + x];
    const double xr = (double)x - d;
    bool ok = false;
    if (!(xr < 0.0 || xr > (double)(W - 1) || xr != xr)) {
        const int x0 = (int)xr;
        const int x1 = x0 + 1 < W ? x0 + 1 : x0;
        const double f = xr - (double)x0;
        const double v = (double)dr[row + x0] * (1.0 - f) + (double)dr[row + x1] * f;
        double diff = d - v;
        if (diff < 0.0) diff = -diff;
        ok = !(diff > tol);
    }
    valid[row + x] = ok ? 1 : 0;
}

// counts[b] = the number of 0/1 flags set in map b: one block per map,
// 16 bytes per load (one bit per flag byte), no atomics, nothing to clear.
__global__ void __launch_bounds__(1024) k_count_flags(const uint8_t* __restrict__ flags,
                                                      long long n, int* __restrict__ counts) {
    __shared__ int ws[16];
    const int b = blockIdx.x, tid = threadIdx.x;
    const uint8_t* f = flags + (long long)b * n;
    // [0, a) bytes, [a, e) 16-byte words, [e, n) bytes
    long long a = (long long)((16 - ((uintptr_t)f & 15)) & 15);
    if (a > n) a = n;
    const long long e = a + (n - a) / 16 * 16;
    int acc = 0;
    for (long long i = tid; i < a; i += 1024) acc += f[i];
    for (long long q = a + 16LL * tid; q < e; q += 16LL * 1024) {
        const uint4 v = *reinterpret_cast<const uint4*>(f + q);
        acc += __popc(v.x) + __popc(v.y) + __popc(v.z) + __popc(v.w);
    }
    for (long long i = e + tid; i < n; i += 1024) acc += f[i];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if ((tid & 63) == 0) ws[tid >> 6] = acc;
    __syncthreads();
    if (tid == 0) {
        int t = 0;
        for (int w = 0; w < 16; ++w) t += ws[w];
        counts[b] = t;
    }
}

// fill_invalid, stereonet.py:5011-5046: one wave per row, 64 columns per
// step.  A max-scan of (valid ? x : -1) gives each pixel's nearest valid
// column at or left of it (left to right, each step's carry kept in LDS),
// a min-scan of (valid ? x : W) from the right the nearest at or right of
// it; an invalid pixel takes the smaller of the two neighbours' values (the
// one there is when only one side has one), and a row without a valid pixel
// is copied.
constexpr int FILL_STEPS = 128;  // rows up to 8192 columns keep every carry

template <typename T>
__global__ void __launch_bounds__(256) k_fill_invalid(const T* __restrict__ disp,
                                                      const uint8_t* __restrict__ valid, int W,
                                                      int rows, T* __restrict__ out) {
    __shared__ int lcarry[4][FILL_STEPS];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int r = blockIdx.x * 4 + wave;
    if (r >= rows) return;  // whole wave
    const long long row = (long long)r * W;
    // left to right: copy every pixel, record the carry before each step
    int carry = -1;
    for (int c0 = 0; c0 < W; c0 += 64) {
        const int x = c0 + lane;
        if (lane == 0 && c0 / 64 < FILL_STEPS) lcarry[wave][c0 / 64] = carry;
        int v = (x < W && valid[row + x]) ? x : -1;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int t = __shfl_up(v, o, 64);
            if (lane >= o) v = max(v, t);
        }
        carry = max(__shfl(v, 63, 64), carry);
        if (x < W) out[row + x] = disp[row + x];
    }
    __builtin_amdgcn_wave_barrier();
    // right to left: nearest valid on each side, then the fill
    carry = W;
    for (int c0 = (W - 1) / 64 * 64; c0 >= 0; c0 -= 64) {
        const int x = c0 + lane;
        const bool in = x < W;
        const bool vx = in && valid[row + x];
        int rt = vx ? x : W, lt = vx ? x : -1;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int a = __shfl_down(rt, o, 64);
            if (lane + o < 64) rt = min(rt, a);
            const int bq = __shfl_up(lt, o, 64);
            if (lane >= o) lt = max(lt, bq);
        }
        rt = min(rt, carry);
        carry = __shfl(rt, 0, 64);
        if (in && !vx) {
            if (lt < 0) {
                if (c0 / 64 < FILL_STEPS) {
                    lt = lcarry[wave][c0 / 64];
                } else {
                    for (int xx = c0 - 1; xx >= 0; --xx)
                        if (valid[row + xx]) {
                            lt = xx;
                            break;
                        }
                }
            }
            const bool hl = lt >= 0, hr = rt < W;
            if (hl || hr) {
                const T a = hl ? disp[row + lt] : T(0), bb = hr ? disp[row + rt] : T(0);
                out[row + x] = !hl ? bb : (!hr ? a : (a < bb ? a : bb));
            }
        }
    }
}

// fill_invalid for rows of at most 64*NC columns: the row's flags and
// values are loaded into registers at once (NC columns per lane), the two
// scans run there, and a rejected pixel reads its two neighbours' values.
template <typename T, int NC>
__global__ void __launch_bounds__(256) k_fill_rows(const T* __restrict__ disp,
                                                   const uint8_t* __restrict__ valid, int W,
                                                   int rows, T* __restrict__ out) {
    const int lane = threadIdx.x & 63;
    const int r = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= rows) return;  // whole wave
    const long long row = (long long)r * W;
    bool v[NC];
    T d[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        const int x = 64 * c + lane;
        v[c] = x < W && valid[row + x];
        d[c] = x < W ? disp[row + x] : T(0);
    }
    int lt[NC], rt[NC];
    int carry = -1;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        int a = v[c] ? 64 * c + lane : -1;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int t = __shfl_up(a, o, 64);
            if (lane >= o) a = max(a, t);
        }
        lt[c] = max(a, carry);
        carry = __shfl(lt[c], 63, 64);
    }
    carry = W;
#pragma unroll
    for (int c = NC - 1; c >= 0; --c) {
        int a = v[c] ? 64 * c + lane : W;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int t = __shfl_down(a, o, 64);
            if (lane + o < 64) a = min(a, t);
        }
        rt[c] = min(a, carry);
        carry = __shfl(rt[c], 0, 64);
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        const int x = 64 * c + lane;
        if (x >= W) continue;
        T o = d[c];
        if (!v[c]) {
            const bool hl = lt[c] >= 0, hr = rt[c] < W;
            if (hl || hr) {
                const T a = hl ? disp[row + lt[c]] : T(0), bb = hr ? disp[row + rt[c]] : T(0);
                o = !hl ? bb : (!hr ? a : (a < bb ? a : bb));
            }
        }
        out[row + x] = o;
    }
}

template <typename T>
static void fill_t(const void* disp, const uint8_t* valid, int W, int rows, void* out,
                   hipStream_t st) {
    dim3 grid((unsigned)((rows + 3) / 4));
    if (W <= 512)
        hipLaunchKernelGGL((k_fill_rows<T, 8>), grid, dim3(256), 0, st, (const T*)disp, valid, W,
                           rows, (T*)out);
    else if (W <= 1024)
        hipLaunchKernelGGL((k_fill_rows<T, 16>), grid, dim3(256), 0, st, (const T*)disp, valid, W,
                           rows, (T*)out);
    else
        hipLaunchKernelGGL((k_fill_invalid<T>), grid, dim3(256), 0, st, (const T*)disp, valid, W,
                           rows, (T*)out);
}

template <typename T>
static int consistency_t(const void* dl, const void* dr, int B, int W, int H, double tol,
                         uint8_t* valid, int* counts, hipStream_t st) {
    hipLaunchKernelGGL((k_lr_consistency<T>), dim3((W + 255) / 256, H, B), dim3(256), 0, st,
                       (const T*)dl, (const T*)dr, W, H, tol, valid);
    SV_CHECK_LAUNCH();
    if (counts) {
        hipLaunchKernelGGL(k_count_flags, dim3(B), dim3(1024), 0, st, valid, (long long)W * H,
                           counts);
        SV_CHECK_LAUNCH();
    }
    return SV_OK;
}

extern "C" {

int sv_census(const void* img, int dtype, int n_images, int W, int H, int rx, int ry,
              uint64_t* out, void* stream) {
    SV_REQUIRE(img && out && n_images > 0 && W > 0 && H > 0, "sv_census: bad arguments");
    const int nbits = (2 * rx + 1) * (2 * ry + 1) - 1;
    if (rx < 0 || ry < 0 || nbits > 64)
        return sv::set_error(SV_EUNSUPPORTED, "census window of %d bits exceeds 64", nbits);
    return census_any(img, dtype, n_images, uniform_geom(n_images, W, H), rx, ry, out,
                      (hipStream_t)stream);
}

int sv_band_costs(const uint64_t* cl, const uint64_t* cr, const void* gl, const void* gr,
                  int dtype, const int* base, int B, int W, int H, int R, double ad_weight,
                  double ad_trunc, double invalid, const double* prior, double prior_w,
                  double prior_t, double* C, void* stream) {
    SV_REQUIRE(cl && cr && gl && gr && base && C && B > 0 && W > 0 && H > 0 && R >= 0,
               "sv_band_costs: bad arguments");
    SV_REQUIRE(prior_w <= 0.0 || prior, "sv_band_costs: prior needed when prior_w > 0");
    const int K = 2 * R + 1;
    hipStream_t st = (hipStream_t)stream;
    // costs are written unpadded ([p][K])
    const Geom ug = uniform_geom(B, W, H);
    const void* ga = gl;
    const void* gb = gr;
    switch (dtype) {
        case SV_U8:
            return costs_f64_t<uint8_t>(false, cl, cr, ga, gb, base, prior, B, ug, R, K, K,
                                        ad_weight, ad_trunc, invalid, prior_w, prior_t, C, st);
        case SV_F32:
            return costs_f64_t<float>(false, cl, cr, ga, gb, base, prior, B, ug, R, K, K,
                                      ad_weight, ad_trunc, invalid, prior_w, prior_t, C, st);
        case SV_F64:
            return costs_f64_t<double>(false, cl, cr, ga, gb, base, prior, B, ug, R, K, K,
                                       ad_weight, ad_trunc, invalid, prior_w, prior_t, C, st);
    }
    return sv::set_error(SV_EINVAL, "unknown image dtype %d", dtype);
}

int sv_aggregate_path_workspace_bytes(int B, int W, int H, int K, size_t* bytes) {
    SV_REQUIRE(bytes && B > 0 && W > 0 && H > 0 && K > 0, "bad arguments");
    int G = 1, KC = 8;
    int rc = choose_layout(K, true, &G, &KC);
    if (rc) return rc;
    const size_t n = (size_t)B * W * H * G * KC;
    *bytes = sv::align_up(n * 8) * 2;
    return SV_OK;
}

int sv_aggregate_path(const double* C, const int* base, int B, int W, int H, int K, double P1,
                      double P2, int dx, int dy, double* S, void* ws, size_t ws_bytes,
                      void* stream) {
    SV_REQUIRE(C && base && S && B > 0 && W > 0 && H > 0 && K > 0, "sv_aggregate_path: bad arguments");
    int dir = -1;
    for (int d = 0; d < 8; ++d) {
        static const int DX[8] = {1, -1, 0, 0, 1, -1, 1, -1};
        static const int DY[8] = {0, 0, 1, -1, 1, 1, -1, -1};
        if (DX[d] == dx && DY[d] == dy) dir = d;
    }
    SV_REQUIRE(dir >= 0, "direction (%d, %d) is not one of the eight SGM paths", dx, dy);
    Plan pl{};
    pl.K = K;
    pl.integer = false;
    pl.lt = 8;
    pl.ndir = 1;
    int rc = choose_layout(K, true, &pl.G, &pl.KC);
    if (rc) return rc;
    pl.Kp = pl.G * pl.KC;
    size_t need = 0;
    sv_aggregate_path_workspace_bytes(B, W, H, K, &need);
    if ((rc = check_ws(need, ws, ws_bytes))) return rc;
    hipStream_t st = (hipStream_t)stream;
    const long long n = (long long)B * W * H * pl.Kp;
    double* Cp = (double*)ws;
    double* L = (double*)((char*)ws + need / 2);
    hipLaunchKernelGGL(k_pad_costs, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, C, pl.Kp, K, n, Cp);
    SV_CHECK_LAUNCH();
    if ((rc = launch_sgm(pl, Cp, base, L, B, uniform_geom(B, W, H), P1, P2, dir, 1, st))) return rc;
    const long long m = (long long)B * W * H * K;
    hipLaunchKernelGGL(k_add_into, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, st, L, pl.Kp, K, m, S);
    SV_CHECK_LAUNCH();
    return SV_OK;
}

// mode: 0 refine_fullres, 1 lr_refine, 2 refine_disparity (planes=False)
int sv_refine_workspace_bytes(int mode, int B, int W, int H, int R, double P1, double P2,
                              double ad_weight, double ad_trunc, int paths, int rx, int ry,
                              double prior_w, double prior_t, size_t* bytes) {
    SV_REQUIRE(bytes && B > 0, "bad arguments");
    Plan pl{};
    int rc = make_plan(W, H, R, P1, P2, ad_weight, prior_w, ad_trunc, prior_t, paths, rx, ry, &pl);
    if (rc) return rc;
    pl.Kp = pl.G * pl.KC;
    const int V = mode == 0 ? 1 : 2;
    RefineBufs r = carve(nullptr, pl, V, B, (long long)W * H, mode != 0, mode == 2, W, H);
    *bytes = r.bytes;
    return SV_OK;
}

int sv_refine_fullres(const void* left, const void* right, int dtype, const double* disp0, int B,
                      int W, int H, int R, double P1, double P2, double ad_weight,
                      double ad_trunc, int paths, double uniq, int rx, int ry, double prior_w,
                      double prior_t, double* out, double* conf, void* ws, size_t ws_bytes,
                      void* stream) {
    SV_REQUIRE(left && right && disp0 && out && B > 0, "sv_refine_fullres: bad arguments");
    Plan pl{};
    int rc = make_plan(W, H, R, P1, P2, ad_weight, prior_w, ad_trunc, prior_t, paths, rx, ry, &pl);
    if (rc) return rc;
    pl.Kp = pl.G * pl.KC;
    RefineBufs bf = carve(ws, pl, 1, B, (long long)W * H, false, false, W, H);
    if ((rc = check_ws(bf.bytes, ws, ws_bytes))) return rc;
    return refine_views(pl, bf, left, right, dtype, disp0, nullptr, 1, uniform_geom(B, W, H), R,
                        P1, P2, ad_weight, ad_trunc, uniq, rx, ry, prior_w, prior_t, out, conf,
                        (hipStream_t)stream);
}

int sv_left_to_right(const double* dl, int B, int W, int H, double* dr, void* stream) {
    SV_REQUIRE(dl && dr && B > 0 && W > 0 && H > 0, "sv_left_to_right: bad arguments");
    launch_l2r(dl, uniform_geom(B, W, H), dr, (hipStream_t)stream);
    SV_CHECK_LAUNCH();
    return SV_OK;
}

static int lr_core(const void* left, const void* right, int dtype, const double* disp0,
                   const Geom& g, double tol, int R, double P1, double P2, double ad_weight,
                   double ad_trunc, int paths, double uniq, int rx, int ry, double prior_w,
                   double prior_t, double* out, uint8_t* status, void* ws, size_t ws_bytes,
                   bool median, hipStream_t st) {
    SV_REQUIRE(left && right && disp0 && out && status && g.B > 0, "lr_refine: bad arguments");
    Plan pl{};
    int rc = make_plan(g.W, g.H, R, P1, P2, ad_weight, prior_w, ad_trunc, prior_t, paths, rx, ry, &pl);
    if (rc) return rc;
    pl.Kp = pl.G * pl.KC;
    const int B = g.B;
    RefineBufs bf = carve(ws, pl, 2, B, g.slot, true, median, g.W, g.H);
    if ((rc = check_ws(bf.bytes, ws, ws_bytes))) return rc;
    {
        StageTimer t(ST_LR, st);
        launch_l2r(disp0, g, bf.d0r, st);
        SV_CHECK_LAUNCH();
    }
    rc = refine_views(pl, bf, left, right, dtype, disp0, bf.d0r, 2, g, R, P1, P2, ad_weight,
                      ad_trunc, uniq, rx, ry, prior_w, prior_t, bf.disp, nullptr, st);
    if (rc) return rc;
    const size_t hw = (size_t)g.slot;
    double* lrout = median ? bf.tmp : out;
    {
        StageTimer t(ST_LR, st);
        launch_lr_check(bf.disp, bf.disp + (size_t)B * hw, disp0, g, tol, lrout, status, st);
        SV_CHECK_LAUNCH();
    }
    if (median) {
        StageTimer t(ST_MEDIAN, st);
        return launch_median(lrout, g, 1, 1, out, nullptr, st);
    }
    return SV_OK;
}

int sv_lr_refine(const void* left, const void* right, int dtype, const double* disp0, int B,
                 int W, int H, double tol, int R, double P1, double P2, double ad_weight,
                 double ad_trunc, int paths, double uniq, int rx, int ry, double prior_w,
                 double prior_t, double* out, uint8_t* status, void* ws, size_t ws_bytes,
                 void* stream) {
    SV_REQUIRE(B > 0 && W > 0 && H > 0, "sv_lr_refine: bad sizes");
    return lr_core(left, right, dtype, disp0, uniform_geom(B, W, H), tol, R, P1, P2, ad_weight,
                   ad_trunc, paths, uniq, rx, ry, prior_w, prior_t, out, status, ws, ws_bytes,
                   false, (hipStream_t)stream);
}

int sv_refine_disparity(const void* left, const void* right, int dtype, const double* disp0,
                        int B, int W, int H, int R, double P1, double P2, int rx, int ry,
                        double tol, double* out, uint8_t* status, void* ws, size_t ws_bytes,
                        void* stream) {
    SV_REQUIRE(B > 0 && W > 0 && H > 0, "sv_refine_disparity: bad sizes");
    return lr_core(left, right, dtype, disp0, uniform_geom(B, W, H), tol, R, P1, P2, 0.0, 20.0, 8,
                   0.0, rx, ry, 0.0, 4.0, out, status, ws, ws_bytes, true, (hipStream_t)stream);
}

// Full-range SGM: refine_fullres (mode 0) or lr_refine (mode 1) with disp0 =
// the constant R at every pixel, K = 2R + 1 -- the control of eval_real.py:
// 143-146.  Results equal sv_refine_fullres / sv_lr_refine given that
// constant map (ad_weight = prior_w = 0, uniq = 0); conf (mode 0, may be
// null) as refine_fullres' return_conf.  rtab (device int32 [B], may be
// null): pair n's own centre, each in [0, R], all pairs in one launch (the
// constant-centre kernels only); R then sizes the band.
int sv_full_range_workspace_bytes(int mode, int B, int W, int H, int R, double P1, double P2,
                                  int paths, int rx, int ry, size_t* bytes) {
    SV_REQUIRE(bytes && B > 0 && (mode == 0 || mode == 1), "sv_full_range_workspace_bytes: bad arguments");
    Plan pl{};
    int rc = make_plan(W, H, R, P1, P2, 0.0, 0.0, 20.0, 4.0, paths, rx, ry, &pl);
    if (rc) return rc;
    const bool lr = mode == 1;
    const int V = lr ? 2 : 1;
    if (cb_applies(pl, W, R, lr)) {
        *bytes = carve_cb(nullptr, cb_kp(cb_bucket(pl.K)), pl.ndir, V, B, (long long)W * H, lr).bytes;
        return SV_OK;
    }
    pl.Kp = pl.G * pl.KC;
    // the general path plus the constant centre map
    *bytes = carve(nullptr, pl, V, B, (long long)W * H, lr, false, W, H).bytes +
             sv::align_up((size_t)B * W * H * sizeof(double));
    return SV_OK;
}

// Whether sv_full_range runs the constant-centre kernels for these
// arguments (1) or the general ones (0); per-pair centres need 1.
int sv_full_range_fast(int mode, int W, int H, int R, double P1, double P2, int paths, int rx,
                       int ry, int* fast) {
    SV_REQUIRE(fast && (mode == 0 || mode == 1), "sv_full_range_fast: bad arguments");
    Plan pl{};
    int rc = make_plan(W, H, R, P1, P2, 0.0, 0.0, 20.0, 4.0, paths, rx, ry, &pl);
    if (rc) return rc;
    *fast = cb_applies(pl, W, R, mode == 1) ? 1 : 0;
    return SV_OK;
}

int sv_full_range(int mode, const void* left, const void* right, int dtype, int B, int W, int H,
                  int R, const int* rtab, double P1, double P2, int paths, int rx, int ry,
                  double tol, double* out, double* conf, uint8_t* status, void* ws,
                  size_t ws_bytes, void* stream) {
    SV_REQUIRE(left && right && out && B > 0 && W > 0 && H > 0 && (mode == 0 || mode == 1),
               "sv_full_range: bad arguments");
    SV_REQUIRE(mode == 0 || status, "sv_full_range: lr mode needs a status buffer");
    Plan pl{};
    int rc = make_plan(W, H, R, P1, P2, 0.0, 0.0, 20.0, 4.0, paths, rx, ry, &pl);
    if (rc) return rc;
    const bool lr = mode == 1;
    const int V = lr ? 2 : 1;
    const Geom g = uniform_geom(B, W, H);
    hipStream_t st = (hipStream_t)stream;
    const long long n = (long long)B * W * H;
    const dim3 fgrid((unsigned)((n + 255) / 256));
    if (!cb_applies(pl, W, R, lr)) {
        if (rtab)
            return sv::set_error(SV_EUNSUPPORTED, "per-pair centres need the constant-centre "
                                 "kernels (integer u8 path, 2R + 1 <= 125, W >= R in lr mode)");
        pl.Kp = pl.G * pl.KC;
        RefineBufs bf = carve(ws, pl, V, B, (long long)W * H, lr, false, W, H);
        const size_t need = bf.bytes + sv::align_up((size_t)n * sizeof(double));
        if ((rc = check_ws(need, ws, ws_bytes))) return rc;
        double* d0 = reinterpret_cast<double*>((char*)ws + bf.bytes);
        hipLaunchKernelGGL(k_fill_f64, fgrid, dim3(256), 0, st, d0, n, (double)R, nullptr, g.slot);
        SV_CHECK_LAUNCH();
        if (!lr)
            return refine_views(pl, bf, left, right, dtype, d0, nullptr, 1, g, R, P1, P2, 0.0, 20.0,
                                0.0, rx, ry, 0.0, 4.0, out, conf, st);
        return lr_core(left, right, dtype, d0, g, tol, R, P1, P2, 0.0, 20.0, paths, 0.0, rx, ry, 0.0,
                       4.0, out, status, bf.cen, bf.bytes, false, st);
    }
    const int kb = cb_bucket(pl.K), Kp = cb_kp(kb);
    CbBufs bf = carve_cb(ws, Kp, pl.ndir, V, B, (long long)W * H, lr);
    if ((rc = check_ws(bf.bytes, ws, ws_bytes))) return rc;
    if (!lr)
        return refine_views_cb(bf, left, right, dtype, 1, g, R, rtab, kb, Kp, pl.ndir, (int)P1,
                               (int)P2, rx, ry, out, conf, st);
    {
        StageTimer t(ST_LR, st);
        hipLaunchKernelGGL(k_fill_f64, fgrid, dim3(256), 0, st, bf.d0, n, (double)R, rtab, g.slot);
        launch_l2r(bf.d0, g, bf.d0r, st);
        SV_CHECK_LAUNCH();
    }
    rc = refine_views_cb(bf, left, right, dtype, 2, g, R, rtab, kb, Kp, pl.ndir, (int)P1, (int)P2,
                         rx, ry, bf.disp, nullptr, st);
    if (rc) return rc;
    StageTimer t(ST_LR, st);
    launch_lr_check(bf.disp, bf.disp + (size_t)n, bf.d0, g, tol, out, status, st);
    SV_CHECK_LAUNCH();
    return SV_OK;
}

// refine_disparity(planes=False) as sv_refine_disparity, with the workspace
// in the library's static device scratch (kScratchBytes per GPU): no
// allocation of any kind (no stage events, no side stream).  Sizes whose
// workspace exceeds the scratch are rejected before anything is queued.
// Every call on one GPU shares that GPU's scratch: calls that overlap on one
// GPU (other streams, other threads) race.  No kernel reads scratch memory it
// has not written in the same call, so nothing is cleared between calls.
int sv_refine_disparity_noalloc(const void* left, const void* right, int dtype,
                                const double* disp0, int B, int W, int H, int R, double P1,
                                double P2, int rx, int ry, double tol, double* out,
                                uint8_t* status, void* stream) {
    SV_REQUIRE(left && right && disp0 && out && status && B > 0 && W > 0 && H > 0,
               "sv_refine_disparity_noalloc: bad arguments");
    Plan pl{};
    int rc = make_plan(W, H, R, P1, P2, 0.0, 0.0, 20.0, 4.0, 8, rx, ry, &pl);
    if (rc) return rc;
    pl.Kp = pl.G * pl.KC;
    const size_t need = carve(nullptr, pl, 2, B, (long long)W * H, true, true, W, H).bytes;
    if (need > kScratchBytes)
        return sv::set_error(SV_EWORKSPACE,
                             "sv_refine_disparity_noalloc: %zu bytes of workspace needed, the "
                             "scratch holds %zu (B*W*H too large for R=%d)",
                             need, kScratchBytes, R);
    void* ws = nullptr;
    if ((rc = scratch_ptr(&ws))) return rc;
    NoAllocScope scope;
    return lr_core(left, right, dtype, disp0, uniform_geom(B, W, H), tol, R, P1, P2, 0.0, 20.0, 8,
                   0.0, rx, ry, 0.0, 4.0, out, status, ws, kScratchBytes, true,
                   (hipStream_t)stream);
}

// Ragged batches: pair n is dims[n] = (h, w) (device int32 pairs, every h <=
// maxH, w <= maxW, h * w <= slot) and sits at n * slot pixels of every
// per-pixel input and output, rows packed at its own width.  mode 0:
// refine_fullres (8 paths, uniq 0, status unused), 1: lr_refine, 2:
// refine_disparity with planes=False.  Integer pipeline only (census
// costs, integral P1/P2).
int sv_refine_ragged_workspace_bytes(int mode, int B, long long slot, int maxW, int maxH, int R,
                                     double P1, double P2, int rx, int ry, size_t* bytes) {
    SV_REQUIRE(bytes && B > 0 && slot > 0 && maxW > 0 && maxH > 0 && mode >= 0 && mode <= 2,
               "sv_refine_ragged_workspace_bytes: bad arguments");
    Plan pl{};
    int rc = make_plan(maxW, maxH, R, P1, P2, 0.0, 0.0, 20.0, 4.0, 8, rx, ry, &pl);
    if (rc) return rc;
    SV_REQUIRE(pl.integer, "ragged batches need integral P1, P2");
    pl.Kp = pl.G * pl.KC;
    *bytes = carve(nullptr, pl, mode == 0 ? 1 : 2, B, slot, mode != 0, mode == 2, maxW, maxH).bytes;
    return SV_OK;
}

int sv_refine_ragged(int mode, const void* left, const void* right, int dtype, const double* disp0,
                     int B, const int* dims, int maxW, int maxH, long long slot, int R, double P1,
                     double P2, int rx, int ry, double tol, double* out, uint8_t* status, void* ws,
                     size_t ws_bytes, void* stream) {
    SV_REQUIRE(left && right && disp0 && out && dims && B > 0 && slot > 0 && maxW > 0 && maxH > 0,
               "sv_refine_ragged: bad arguments");
    SV_REQUIRE(mode >= 0 && mode <= 2, "sv_refine_ragged: mode %d is not 0, 1 or 2", mode);
    Geom g;
    g.dims = dims;
    g.W = maxW;
    g.H = maxH;
    g.B = B;
    g.slot = slot;
    hipStream_t st = (hipStream_t)stream;
    if (mode != 0) {
        SV_REQUIRE(status, "sv_refine_ragged: lr modes need a status buffer");
        return lr_core(left, right, dtype, disp0, g, tol, R, P1, P2, 0.0, 20.0, 8, 0.0, rx, ry, 0.0,
                       4.0, out, status, ws, ws_bytes, mode == 2, st);
    }
    Plan pl{};
    int rc = make_plan(maxW, maxH, R, P1, P2, 0.0, 0.0, 20.0, 4.0, 8, rx, ry, &pl);
    if (rc) return rc;
    SV_REQUIRE(pl.integer, "ragged batches need integral P1, P2");
    pl.Kp = pl.G * pl.KC;
    RefineBufs bf = carve(ws, pl, 1, B, slot, false, false, maxW, maxH);
    if ((rc = check_ws(bf.bytes, ws, ws_bytes))) return rc;
    return refine_views(pl, bf, left, right, dtype, disp0, nullptr, 1, g, R, P1, P2, 0.0, 20.0, 0.0,
                        rx, ry, 0.0, 4.0, out, nullptr, st);
}

// Stage timing (see StageTimer): on != 0 clears and starts collecting.
int sv_profile_enable(int on) {
    for (int i = 0; i < g_prof_n; ++i) {
        (void)hipEventDestroy(g_prof[i].a);
        (void)hipEventDestroy(g_prof[i].b);
    }
    g_prof_n = 0;
    g_prof_on = on != 0;
    return SV_OK;
}

// Total milliseconds and number of recorded launches of one stage ("census",
// "costs", "sgm", "wta", "lr", "median"); waits for the recorded events.
int sv_profile_read(const char* stage, double* total_ms, int* launches) {
    SV_REQUIRE(stage && total_ms && launches, "sv_profile_read: bad arguments");
    int id = -1;
    for (int i = 0; i < ST_COUNT; ++i)
        if (strcmp(stage, kStageNames[i]) == 0) id = i;
    SV_REQUIRE(id >= 0, "unknown stage '%s'", stage);
    double tot = 0.0;
    int cnt = 0;
    for (int i = 0; i < g_prof_n; ++i) {
        if (g_prof[i].stage != id) continue;
        hipError_t e = hipEventSynchronize(g_prof[i].b);
        if (e != hipSuccess) return sv::set_error(SV_EHIP, "profile: %s", hipGetErrorString(e));
        float ms = 0.f;
        e = hipEventElapsedTime(&ms, g_prof[i].a, g_prof[i].b);
        if (e != hipSuccess) return sv::set_error(SV_EHIP, "profile: %s", hipGetErrorString(e));
        tot += ms;
        ++cnt;
    }
    *total_ms = tot;
    *launches = cnt;
    return SV_OK;
}

int sv_median_filter(const double* disp, int B, int W, int H, int radius, int passes,
                     double* out, double* tmp, void* stream) {
    SV_REQUIRE(disp && out && B > 0 && W > 0 && H > 0, "sv_median_filter: bad arguments");
    return launch_median(disp, uniform_geom(B, W, H), radius, passes, out, tmp, (hipStream_t)stream);
}

int sv_mirror_h(const void* in, int dtype, long long rows, int W, void* out, void* stream) {
    SV_REQUIRE(in && out && rows > 0 && W > 0, "sv_mirror_h: bad arguments");
    SV_REQUIRE(in != out, "sv_mirror_h: out must not alias in");
    hipStream_t st = (hipStream_t)stream;
    dim3 grid((unsigned)std::min<long long>((rows + MIR_ROWS - 1) / MIR_ROWS, 1LL << 20));
    switch (dtype) {
        case SV_U8:
            hipLaunchKernelGGL((k_mirror_h<uint8_t>), grid, dim3(256), 0, st, (const uint8_t*)in,
                               rows, W, (uint8_t*)out);
            break;
        case SV_F32:
            hipLaunchKernelGGL((k_mirror_h<float>), grid, dim3(256), 0, st, (const float*)in, rows,
                               W, (float*)out);
            break;
        case SV_F64:
            hipLaunchKernelGGL((k_mirror_h<double>), grid, dim3(256), 0, st, (const double*)in,
                               rows, W, (double*)out);
            break;
        default:
            return sv::set_error(SV_EINVAL, "sv_mirror_h: unknown dtype %d", dtype);
    }
    SV_CHECK_LAUNCH();
    return SV_OK;
}

int sv_lr_consistency(const void* dl, const void* dr, int dtype, int B, int W, int H, double tol,
                      uint8_t* valid, int* counts, void* stream) {
    SV_REQUIRE(dl && dr && valid && B > 0 && W > 0 && H > 0, "sv_lr_consistency: bad arguments");
    SV_REQUIRE(H <= 65535 && B <= 65535, "sv_lr_consistency: %d images of height %d", B, H);
    hipStream_t st = (hipStream_t)stream;
    if (dtype == SV_F32) return consistency_t<float>(dl, dr, B, W, H, tol, valid, counts, st);
    if (dtype == SV_F64) return consistency_t<double>(dl, dr, B, W, H, tol, valid, counts, st);
    return sv::set_error(SV_EINVAL, "sv_lr_consistency: maps must be f32 or f64 (dtype %d)", dtype);
}

int sv_fill_invalid(const void* disp, const uint8_t* valid, int dtype, int B, int W, int H,
                    void* out, void* stream) {
    SV_REQUIRE(disp && valid && out && B > 0 && W > 0 && H > 0, "sv_fill_invalid: bad arguments");
    SV_REQUIRE(disp != out, "sv_fill_invalid: out must not alias disp");
    const long long rows = (long long)B * H;
    SV_REQUIRE(rows <= (long long)INT_MAX, "sv_fill_invalid: too many rows");
    hipStream_t st = (hipStream_t)stream;
    if (dtype == SV_F32)
        fill_t<float>(disp, valid, W, (int)rows, out, st);
    else if (dtype == SV_F64)
        fill_t<double>(disp, valid, W, (int)rows, out, st);
    else
        return sv::set_error(SV_EINVAL, "sv_fill_invalid: maps must be f32 or f64 (dtype %d)", dtype);
    SV_CHECK_LAUNCH();
    return SV_OK;
}

}  // extern "C"
