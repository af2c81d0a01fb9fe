// Middlebury-style evaluation on the GPU (eval_real.py:78-133): ground truth
// from a decoded disparity PNG, the non-occluded mask from the two views'
// ground truth, and the per-map error sums behind EPE / bad1 / bad2 / bad4
// and the constant-prediction baseline.
//
// Per-pixel rules are the reference's float64 arithmetic, bit for bit.  The
// sums are deterministic (a fixed partition of each map into chunks, a fixed
// reduction tree inside a block, the chunk partials added in chunk order) but
// not in the reference's left-to-right order: float64 sums agree with it to
// rounding, counts exactly.
#include <cstdint>

#include "common.h"

namespace {

constexpr int kThreads = 256;
constexpr int kStats = 6;  // n, sum |d|, >1, >2, >4, sum gt (over the mask)

// load_gt, eval_real.py:78-87: level = gray * 255; known where level >= 0.5,
// disparity = level / scale there, 0 elsewhere
__global__ void k_gt_from_gray(const double* __restrict__ g, long long n, double scale,
                               double* __restrict__ disp, uint8_t* __restrict__ known) {
    const long long i = blockIdx.x * (long long)kThreads + threadIdx.x;
    if (i >= n) return;
    const double lv = g[i] * 255.0;
    const bool k = lv >= 0.5;
    disp[i] = k ? lv / scale : 0.0;
    known[i] = k;
}

// masks, eval_real.py:90-102: a known left pixel is non-occluded when
// xr = int(x - dL + 0.5) (truncation toward zero) lands inside the row on a
// known right pixel whose disparity is within 1 of dL
__global__ void k_gt_nonocc(const double* __restrict__ dl, const uint8_t* __restrict__ kl,
                            const double* __restrict__ dr, const uint8_t* __restrict__ kr, int W,
                            long long total, uint8_t* __restrict__ nonocc) {
    const long long i = blockIdx.x * (long long)kThreads + threadIdx.x;
    if (i >= total) return;
    uint8_t v = 0;
    if (kl[i]) {
        const long long x = i % W, row = i - x;
        const double t = (double)x - dl[i] + 0.5;
        if (t > -1.0 && t < (double)W) {
            const long long xr = (long long)t;  // toward zero, as int()
            v = kr[row + xr] && fabs(dl[i] - dr[row + xr]) <= 1.0;
        }
    }
    nonocc[i] = v;
}

int chunks_for(long long npix) {
    long long p = (npix + kThreads * 16 - 1) / (kThreads * 16);
    return (int)(p < 1 ? 1 : p > 128 ? 128 : p);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// fixed-shape block sum: waves by shuffles, then wave 0 over the four waves
__device__ double block_sum(double v, double* lds) {
    v = wave_sum(v);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    __syncthreads();
    if (l == 0) lds[w] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int k = 0; k < kThreads / 64; ++k) s += lds[k];
    return s;
}

// four consecutive predictions as doubles (16-byte aligned loads)
__device__ __forceinline__ void load4(const float* p, double* v) {
    const float4 f = *reinterpret_cast<const float4*>(p);
    v[0] = f.x;
    v[1] = f.y;
    v[2] = f.z;
    v[3] = f.w;
}
__device__ __forceinline__ void load4(const double* p, double* v) {
    const double2 a = reinterpret_cast<const double2*>(p)[0];
    const double2 b = reinterpret_cast<const double2*>(p)[1];
    v[0] = a.x;
    v[1] = a.y;
    v[2] = b.x;
    v[3] = b.y;
}

// pixel range [lo, hi) of chunk c; V4 chunks are whole groups of four
template <bool V4>
__device__ __forceinline__ void chunk_range(long long npix, int P_, int c, long long& lo,
                                            long long& hi) {
    const long long chunk = V4 ? (npix / 4 + P_ - 1) / P_ * 4 : (npix + P_ - 1) / P_;
    lo = c * chunk;
    hi = lo + chunk < npix ? lo + chunk : npix;
}

// score, eval_real.py:105-121 (and the sum of constant's mean, :124-127):
// chunk (blockIdx.x) of map (blockIdx.y) -> partial[map][chunk][6].  V4:
// four pixels per thread and step (npix % 4 == 0, aligned maps)
template <typename P, bool V4>
__global__ void __launch_bounds__(kThreads) k_score(const P* __restrict__ pred,
                                                    const double* __restrict__ gt,
                                                    const uint8_t* __restrict__ mask,
                                                    long long npix, int P_, double* __restrict__ part) {
    __shared__ double lds[kThreads / 64];
    const int b = blockIdx.y, c = blockIdx.x;
    long long lo, hi;
    chunk_range<V4>(npix, P_, c, lo, hi);
    const long long base = (long long)b * npix;
    double n = 0, e = 0, b1 = 0, b2 = 0, b4 = 0, gs = 0;
    auto add = [&](double pv, double g) {
        const double d = fabs(pv - g);
        n += 1.0;
        e += d;
        b1 += d > 1.0;
        b2 += d > 2.0;
        b4 += d > 4.0;
        gs += g;
    };
    if constexpr (V4) {
        // two groups of four pixels per step, all their loads issued first
        for (long long i = lo + 4 * threadIdx.x; i < hi; i += 8 * kThreads) {
            const long long i2 = i + 4 * kThreads;
            const bool two = i2 < hi;
            const uint32_t mk = *reinterpret_cast<const uint32_t*>(mask + base + i);
            const uint32_t mk2 = two ? *reinterpret_cast<const uint32_t*>(mask + base + i2) : 0u;
            double pv[4], g[4], pv2[4], g2[4];
            load4(pred + base + i, pv);
            load4(gt + base + i, g);
            if (two) {
                load4(pred + base + i2, pv2);
                load4(gt + base + i2, g2);
            }
#pragma unroll
            for (int j = 0; j < 4; ++j)
                if ((mk >> (8 * j)) & 0xFFu) add(pv[j], g[j]);
            if (two) {
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    if ((mk2 >> (8 * j)) & 0xFFu) add(pv2[j], g2[j]);
            }
        }
    } else {
        for (long long i = lo + threadIdx.x; i < hi; i += kThreads)
            if (mask[base + i]) add((double)pred[base + i], gt[base + i]);
    }
    double* o = part + ((size_t)b * P_ + c) * kStats;
    const double r0 = block_sum(n, lds);
    const double r1 = block_sum(e, lds);
    const double r2 = block_sum(b1, lds);
    const double r3 = block_sum(b2, lds);
    const double r4 = block_sum(b4, lds);
    const double r5 = block_sum(gs, lds);
    if (threadIdx.x == 0) {
        o[0] = r0;
        o[1] = r1;
        o[2] = r2;
        o[3] = r3;
        o[4] = r4;
        o[5] = r5;
    }
}

// constant, eval_real.py:124-127: sum |gt - mean| over the mask, the mean
// taken from the score sums (sum gt / n)
template <bool V4>
__global__ void __launch_bounds__(kThreads) k_absdev(const double* __restrict__ gt,
                                                     const uint8_t* __restrict__ mask,
                                                     const double* __restrict__ stats, long long npix,
                                                     int P_, double* __restrict__ part) {
    __shared__ double lds[kThreads / 64];
    const int b = blockIdx.y, c = blockIdx.x;
    long long lo, hi;
    chunk_range<V4>(npix, P_, c, lo, hi);
    const long long base = (long long)b * npix;
    const double m = stats[b * kStats + 5] / stats[b * kStats + 0];
    double s = 0;
    if constexpr (V4) {
        for (long long i = lo + 4 * threadIdx.x; i < hi; i += 4 * kThreads) {
            const uint32_t mk = *reinterpret_cast<const uint32_t*>(mask + base + i);
            if (!mk) continue;
            double g[4];
            load4(gt + base + i, g);
#pragma unroll
            for (int j = 0; j < 4; ++j)
                if ((mk >> (8 * j)) & 0xFFu) s += fabs(g[j] - m);
        }
    } else {
        for (long long i = lo + threadIdx.x; i < hi; i += kThreads)
            if (mask[base + i]) s += fabs(gt[base + i] - m);
    }
    const double r = block_sum(s, lds);
    if (threadIdx.x == 0) part[(size_t)b * P_ + c] = r;
}

// chunk partials of each (map, statistic): one wave per map, lane l holding
// chunks l and l + 64, then the fixed shuffle tree
__global__ void __launch_bounds__(64) k_sum_parts(const double* __restrict__ part, int P_, int K,
                                                  double* __restrict__ out) {
    const int b = blockIdx.x, l = threadIdx.x;
    for (int k = 0; k < K; ++k) {
        double v = 0.0;
        if (l < P_) v = part[((size_t)b * P_ + l) * K + k];
        if (l + 64 < P_) v += part[((size_t)b * P_ + l + 64) * K + k];
        v = wave_sum(v);
        if (l == 0) out[b * K + k] = v;
    }
}

}  // namespace

extern "C" {

int sv_gt_from_gray(const double* gray, long long n, double scale, double* disp, uint8_t* known,
                    void* stream) {
    SV_REQUIRE(gray && disp && known && n > 0 && scale > 0.0, "sv_gt_from_gray: bad arguments");
    hipLaunchKernelGGL(k_gt_from_gray, dim3((unsigned)((n + kThreads - 1) / kThreads)),
                       dim3(kThreads), 0, (hipStream_t)stream, gray, n, scale, disp, known);
    SV_CHECK_LAUNCH();
    return SV_OK;
}

int sv_gt_nonocc(const double* dl, const uint8_t* kl, const double* dr, const uint8_t* kr, int B,
                 int W, int H, uint8_t* nonocc, void* stream) {
    SV_REQUIRE(dl && kl && dr && kr && nonocc && B > 0 && W > 0 && H > 0,
               "sv_gt_nonocc: bad arguments");
    const long long total = (long long)B * W * H;
    hipLaunchKernelGGL(k_gt_nonocc, dim3((unsigned)((total + kThreads - 1) / kThreads)),
                       dim3(kThreads), 0, (hipStream_t)stream, dl, kl, dr, kr, W, total, nonocc);
    SV_CHECK_LAUNCH();
    return SV_OK;
}

int sv_score_workspace_bytes(int B, long long npix, size_t* bytes) {
    SV_REQUIRE(bytes && B > 0 && npix > 0, "sv_score_workspace_bytes: bad arguments");
    *bytes = (size_t)B * chunks_for(npix) * kStats * sizeof(double);
    return SV_OK;
}

int sv_score(const void* pred, int pred_dtype, const double* gt, const uint8_t* mask, int B,
             long long npix, double* out, void* ws, size_t ws_bytes, void* stream) {
    SV_REQUIRE(pred && gt && mask && out && ws && B > 0 && B <= 65535 && npix > 0,
               "sv_score: bad arguments");
    SV_REQUIRE(pred_dtype == SV_F32 || pred_dtype == SV_F64, "sv_score: pred must be f32 or f64");
    const int P_ = chunks_for(npix);
    SV_REQUIRE(ws_bytes >= (size_t)B * P_ * kStats * sizeof(double), "sv_score: workspace too small");
    double* part = (double*)ws;
    hipStream_t s = (hipStream_t)stream;
    const size_t psz = pred_dtype == SV_F32 ? 4 : 8;
    const bool v4 = npix % 4 == 0 && !((uintptr_t)pred % (4 * psz)) && !((uintptr_t)gt % 16) &&
                    !((uintptr_t)mask % 4);
    const dim3 grid(P_, B);
    if (pred_dtype == SV_F32) {
        if (v4)
            hipLaunchKernelGGL((k_score<float, true>), grid, dim3(kThreads), 0, s, (const float*)pred,
                               gt, mask, npix, P_, part);
        else
            hipLaunchKernelGGL((k_score<float, false>), grid, dim3(kThreads), 0, s,
                               (const float*)pred, gt, mask, npix, P_, part);
    } else {
        if (v4)
            hipLaunchKernelGGL((k_score<double, true>), grid, dim3(kThreads), 0, s,
                               (const double*)pred, gt, mask, npix, P_, part);
        else
            hipLaunchKernelGGL((k_score<double, false>), grid, dim3(kThreads), 0, s,
                               (const double*)pred, gt, mask, npix, P_, part);
    }
    SV_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_sum_parts, dim3(B), dim3(64), 0, s, part, P_, kStats, out);
    SV_CHECK_LAUNCH();
    return SV_OK;
}

int sv_abs_dev(const double* gt, const uint8_t* mask, const double* stats, int B, long long npix,
               double* out, void* ws, size_t ws_bytes, void* stream) {
    SV_REQUIRE(gt && mask && stats && out && ws && B > 0 && B <= 65535 && npix > 0,
               "sv_abs_dev: bad arguments");
    const int P_ = chunks_for(npix);
    SV_REQUIRE(ws_bytes >= (size_t)B * P_ * sizeof(double), "sv_abs_dev: workspace too small");
    double* part = (double*)ws;
    hipStream_t s = (hipStream_t)stream;
    const bool v4 = npix % 4 == 0 && !((uintptr_t)gt % 16) && !((uintptr_t)mask % 4);
    if (v4)
        hipLaunchKernelGGL((k_absdev<true>), dim3(P_, B), dim3(kThreads), 0, s, gt, mask, stats, npix,
                           P_, part);
    else
        hipLaunchKernelGGL((k_absdev<false>), dim3(P_, B), dim3(kThreads), 0, s, gt, mask, stats,
                           npix, P_, part);
    SV_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_sum_parts, dim3(B), dim3(64), 0, s, part, P_, 1, out);
    SV_CHECK_LAUNCH();
    return SV_OK;
}

}  // extern "C"
