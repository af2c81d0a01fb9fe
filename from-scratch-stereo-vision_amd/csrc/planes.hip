// Plane snapping: SLIC superpixels, a RANSAC plane per segment, the
// least-squares refit and the snap rule of plane_snap (stereonet.py:
// 5394-5563), for B images of W x H.  float64 throughout, compiled without
// multiply-add contraction, and every order-dependent float64 sum (centre
// means, the plane refit) taken sequentially in the reference's pixel order,
// so labels and snapped values equal the reference's bit for bit.
//
// Parallel structure:
//  * SLIC assignment: a block per centre walks its window twice; the first
//    pass takes the atomic minimum of the distance bits per pixel
//    (non-negative doubles order like their bits), the second the smallest
//    centre index at that distance, which is the reference's "first centre
//    with the strictly smallest distance".
//  * A pixel no window reached keeps its previous label (the reference
//    never clears labels).  Such "strays" lie outside every current window;
//    each image keeps a list of them, and a centre's member list is its
//    window's members (row-major, so ascending pixel index) merged with its
//    strays, in ascending order.
//  * Centre update and plane refit: one wave per centre/segment stages
//    64 members at a time in LDS, and lane 0 adds them in order.
//  * The reference draws its RANSAC triples from one xorshift32 stream over
//    all segments in order, 3 draws per trial for each segment with at least
//    12 valid pixels.  A segment's starting state is the seed advanced by
//    3 * trials * (eligible segments before it) steps, computed from
//    precomputed powers of the (linear) xorshift map.
#include <climits>
#include <cstring>

#include "common.h"

namespace {

// ------------------------------------------------ xorshift32 jump tables
// The xorshift32 step as a 32x32 matrix over GF(2): col[k] = step(1 << k).
// JUMP[j] holds the columns of step^(120 * 2^j) (120 draws: 40 trials of 3).
struct Mat32 {
    uint32_t c[32];
};

constexpr uint32_t xs_step(uint32_t s) {
    s ^= s << 13;
    s ^= s >> 17;
    s ^= s << 5;
    return s;
}

constexpr uint32_t mat_apply(const Mat32& m, uint32_t v) {
    uint32_t r = 0;
    for (int k = 0; k < 32; ++k)
        if ((v >> k) & 1u) r ^= m.c[k];
    return r;
}

constexpr Mat32 mat_mul(const Mat32& a, const Mat32& b) {  // a after b
    Mat32 r{};
    for (int k = 0; k < 32; ++k) r.c[k] = mat_apply(a, b.c[k]);
    return r;
}

constexpr int JUMP_BITS = 16;

struct JumpTable {
    Mat32 j[JUMP_BITS];
};

constexpr JumpTable make_jumps(int draws) {
    Mat32 step{};
    for (int k = 0; k < 32; ++k) step.c[k] = xs_step(1u << k);
    Mat32 p{};
    for (int k = 0; k < 32; ++k) p.c[k] = 1u << k;  // identity
    for (int d = 0; d < draws; ++d) p = mat_mul(step, p);
    JumpTable t{};
    t.j[0] = p;
    for (int b = 1; b < JUMP_BITS; ++b) t.j[b] = mat_mul(t.j[b - 1], t.j[b - 1]);
    return t;
}

constexpr int PS_TRIALS = 40;
__constant__ JumpTable c_jump = make_jumps(3 * PS_TRIALS);

__device__ uint32_t jump_apply(const Mat32& m, uint32_t v) {
    uint32_t r = 0;
#pragma unroll
    for (int k = 0; k < 32; ++k)
        if ((v >> k) & 1u) r ^= m.c[k];
    return r;
}

// --------------------------------------------------------------- helpers
template <typename RT>
__device__ __forceinline__ double rgb_at(const RT* __restrict__ rgb, long long e) {
    if constexpr (sizeof(RT) == 1)
        return (double)rgb[e] * (1.0 / 255.0);
    else
        return (double)rgb[e];
}

struct SlicGeom {
    int W, H, step, KX, K;
    double m2;
};

// window of centre (cx, cy), clamped as slic does
__device__ __forceinline__ void window(double cx, double cy, const SlicGeom& g, int& x0, int& x1,
                                       int& y0, int& y1) {
    x0 = (int)cx - g.step;
    x1 = (int)cx + g.step;
    y0 = (int)cy - g.step;
    y1 = (int)cy + g.step;
    if (x0 < 0) x0 = 0;
    if (y0 < 0) y0 = 0;
    if (x1 >= g.W) x1 = g.W - 1;
    if (y1 >= g.H) y1 = g.H - 1;
}

template <typename RT>
__device__ __forceinline__ double slic_dd(const RT* __restrict__ rgbb, const double* c, int xx,
                                          int yy, const SlicGeom& g) {
    const long long i = (long long)yy * g.W + xx;
    const double a = rgb_at(rgbb, 3 * i) * 255.0 - c[0];
    const double b = rgb_at(rgbb, 3 * i + 1) * 255.0 - c[1];
    const double cc = rgb_at(rgbb, 3 * i + 2) * 255.0 - c[2];
    const double dx = xx - c[3], dy = yy - c[4];
    return a * a + b * b + cc * cc + g.m2 * (dx * dx + dy * dy);
}

// ------------------------------------------------------------------ SLIC
template <typename RT>
__global__ void k_slic_init(const RT* __restrict__ rgb, SlicGeom g, int B, double* __restrict__ cs) {
    const int k = blockIdx.x * 256 + threadIdx.x, b = blockIdx.y;
    if (k >= g.K) return;
    const int x = g.step / 2 + (k % g.KX) * g.step, y = g.step / 2 + (k / g.KX) * g.step;
    const long long i = (long long)b * g.W * g.H + (long long)y * g.W + x;
    double* c = cs + ((long long)b * g.K + k) * 5;
    c[0] = rgb_at(rgb, 3 * i) * 255.0;
    c[1] = rgb_at(rgb, 3 * i + 1) * 255.0;
    c[2] = rgb_at(rgb, 3 * i + 2) * 255.0;
    c[3] = (double)x;
    c[4] = (double)y;
}

__global__ void k_fill_i32(int* __restrict__ p, long long n, int v) {
    const long long i = blockIdx.x * 256LL + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_slic_clear(unsigned long long* __restrict__ dist, int* __restrict__ lab2,
                             int* __restrict__ nstray, long long n, int B) {
    const long long i = blockIdx.x * 256LL + threadIdx.x;
    if (i < n) {
        dist[i] = (unsigned long long)__double_as_longlong(1.0e30);
        lab2[i] = INT_MAX;
    }
    if (i < B) nstray[i] = 0;
}

// pass 1 (PASS == 0): atomic minimum of the distance bits; pass 2: the
// smallest centre index at the minimum
template <typename RT, int PASS>
__global__ void __launch_bounds__(256) k_slic_assign(const RT* __restrict__ rgb, SlicGeom g,
                                                     const double* __restrict__ cs,
                                                     unsigned long long* __restrict__ dist,
                                                     int* __restrict__ lab2) {
    const int ci = blockIdx.x, b = blockIdx.y;
    const long long N = (long long)g.W * g.H;
    const double* c = cs + ((long long)b * g.K + ci) * 5;
    const double cv[5] = {c[0], c[1], c[2], c[3], c[4]};
    int x0, x1, y0, y1;
    window(cv[3], cv[4], g, x0, x1, y0, y1);
    const int ww = x1 - x0 + 1, n = ww * (y1 - y0 + 1);
    const RT* rgbb = rgb + 3 * (long long)b * N;
    for (int e = threadIdx.x; e < n; e += 256) {
        const int yy = y0 + e / ww, xx = x0 + e % ww;
        const long long i = (long long)b * N + (long long)yy * g.W + xx;
        const unsigned long long d =
            (unsigned long long)__double_as_longlong(slic_dd(rgbb, cv, xx, yy, g));
        if (PASS == 0) {
            if (d < dist[i]) atomicMin(&dist[i], d);
        } else if (d == dist[i]) {
            atomicMin(&lab2[i], ci);
        }
    }
}

// labels of the pass; a labelled pixel that no window reached is a stray
__global__ void k_slic_commit(const int* __restrict__ lab2, int* __restrict__ lab,
                              int* __restrict__ strays, int* __restrict__ nstray, long long N,
                              int B) {
    const long long t = blockIdx.x * 256LL + threadIdx.x;
    if (t >= N * B) return;
    const int l2 = lab2[t];
    if (l2 != INT_MAX) {
        lab[t] = l2;
    } else if (lab[t] >= 0) {
        const int b = (int)(t / N);
        const int k = atomicAdd(&nstray[b], 1);
        strays[(long long)b * N + k] = (int)(t - (long long)b * N);
    }
}

// Ordered member walk of segment ci of image b: the window's pixels in
// row-major order (those with label ci) merged with the image's strays of
// label ci in ascending order.  f(i) is called for every member, in order,
// by lane 0 only, after the wave staged a chunk in LDS: stage(e, i, match)
// runs on every lane for its window element.  Strays of the segment are
// collected (sorted) into sl[] (capacity SCAP); returns false when that
// overflows.
constexpr int SCAP = 256;

struct StrayList {
    int n;
    int idx[SCAP];
};

__device__ bool collect_strays(const int* __restrict__ strays, int ns, const int* __restrict__ labb,
                               int ci, StrayList* sl) {
    const int lane = threadIdx.x;
    if (lane == 0) sl->n = 0;
    __builtin_amdgcn_wave_barrier();
    bool ok = true;
    for (int s0 = 0; s0 < ns; s0 += 64) {
        const int s = s0 + lane;
        const int i = s < ns ? strays[s] : -1;
        const bool m = s < ns && labb[i] == ci;
        const uint64_t bal = __ballot(m);
        const int base = sl->n;
        const int pos = base + __popcll(bal & ((1ull << lane) - 1ull));
        __builtin_amdgcn_wave_barrier();
        if (m) {
            if (pos < SCAP) sl->idx[pos] = i;
            else ok = false;
        }
        __builtin_amdgcn_wave_barrier();
        if (lane == 0) sl->n = min(SCAP, base + __popcll(bal));
        __builtin_amdgcn_wave_barrier();
    }
    ok = !__any(!ok);
    // insertion sort (the list is short), lane 0
    if (lane == 0) {
        for (int a = 1; a < sl->n; ++a) {
            const int v = sl->idx[a];
            int j = a - 1;
            while (j >= 0 && sl->idx[j] > v) {
                sl->idx[j + 1] = sl->idx[j];
                --j;
            }
            sl->idx[j + 1] = v;
        }
    }
    __builtin_amdgcn_wave_barrier();
    return ok;
}

// centre update: means of the members' colour (as read, [0,1]) and
// coordinates, summed in pixel order; one wave (64 threads) per centre
template <typename RT>
__global__ void __launch_bounds__(64) k_slic_update(const RT* __restrict__ rgb, SlicGeom g,
                                                    const double* __restrict__ cs_in,
                                                    double* __restrict__ cs_out,
                                                    const int* __restrict__ lab,
                                                    const int* __restrict__ strays,
                                                    const int* __restrict__ nstray,
                                                    int* __restrict__ err) {
    __shared__ StrayList sl;
    __shared__ int mi[64];
    __shared__ double mc[3][64];  // the chunk's colours, staged by every lane
    const int ci = blockIdx.x, b = blockIdx.y, lane = threadIdx.x;
    const long long N = (long long)g.W * g.H;
    const double* c = cs_in + ((long long)b * g.K + ci) * 5;
    int x0, x1, y0, y1;
    window(c[3], c[4], g, x0, x1, y0, y1);
    const int ww = x1 - x0 + 1, n = ww * (y1 - y0 + 1);
    const int* labb = lab + (long long)b * N;
    const RT* rgbb = rgb + 3 * (long long)b * N;
    if (!collect_strays(strays + (long long)b * N, nstray[b], labb, ci, &sl)) {
        if (lane == 0) atomicOr(err, 1);
        return;
    }
    // colour sums in member order (lane 0, over each chunk's members
    // compacted by ballot); coordinate sums and the count are integers,
    // exact in any order, so every lane adds its own
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    long long sx = 0, sy = 0, cnt = 0;
    int sp = 0;
    auto add_stray = [&](int i) {
        a0 += rgb_at(rgbb, 3LL * i);
        a1 += rgb_at(rgbb, 3LL * i + 1);
        a2 += rgb_at(rgbb, 3LL * i + 2);
        sx += i % g.W;
        sy += i / g.W;
        cnt += 1;
    };
    const uint64_t below = (1ull << lane) - 1ull;
    for (int e0 = 0; e0 < n; e0 += 64) {
        const int e = e0 + lane;
        bool m = false;
        int p = 0;
        if (e < n) {
            const int yy = y0 + e / ww, xx = x0 + e % ww;
            p = yy * g.W + xx;
            m = labb[p] == ci;
        }
        const uint64_t bal = __ballot(m);
        if (m) {
            const int pos = __popcll(bal & below);
            mi[pos] = p;
            mc[0][pos] = rgb_at(rgbb, 3LL * p);
            mc[1][pos] = rgb_at(rgbb, 3LL * p + 1);
            mc[2][pos] = rgb_at(rgbb, 3LL * p + 2);
            sx += p % g.W;
            sy += p / g.W;
            cnt += 1;
        }
        __builtin_amdgcn_wave_barrier();
        if (lane == 0) {
            const int nmem = __popcll(bal);
            for (int k = 0; k < nmem; ++k) {
                if (sp < sl.n)
                    while (sp < sl.n && sl.idx[sp] < mi[k]) add_stray(sl.idx[sp++]);
                a0 += mc[0][k];
                a1 += mc[1][k];
                a2 += mc[2][k];
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
    if (lane == 0)
        while (sp < sl.n) add_stray(sl.idx[sp++]);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        sx += __shfl_xor(sx, o, 64);
        sy += __shfl_xor(sy, o, 64);
        cnt += __shfl_xor(cnt, o, 64);
    }
    if (lane == 0) {
        double* o = cs_out + ((long long)b * g.K + ci) * 5;
        if (cnt != 0) {
            const double a5 = (double)cnt;
            o[0] = a0 / a5 * 255.0;
            o[1] = a1 / a5 * 255.0;
            o[2] = a2 / a5 * 255.0;
            o[3] = (double)sx / a5;
            o[4] = (double)sy / a5;
        } else {
            for (int q = 0; q < 5; ++q) o[q] = c[q];
        }
    }
}

// ------------------------------------------------------------ plane_snap
__device__ __forceinline__ bool is_valid(const uint8_t* __restrict__ vb, int kind, int i) {
    if (!vb) return true;
    return kind == 0 ? vb[i] != 0 : vb[i] == 0;
}

// Per segment: the number of valid members (eligibility for the draws).
__global__ void __launch_bounds__(64) k_ps_count(SlicGeom g, const double* __restrict__ cs,
                                                 const int* __restrict__ lab,
                                                 const int* __restrict__ strays,
                                                 const int* __restrict__ nstray,
                                                 const uint8_t* __restrict__ valid, int vkind,
                                                 int* __restrict__ npts) {
    const int ci = blockIdx.x, b = blockIdx.y, lane = threadIdx.x;
    const long long N = (long long)g.W * g.H;
    const double* c = cs + ((long long)b * g.K + ci) * 5;
    int x0, x1, y0, y1;
    window(c[3], c[4], g, x0, x1, y0, y1);
    const int ww = x1 - x0 + 1, n = ww * (y1 - y0 + 1);
    const int* labb = lab + (long long)b * N;
    const uint8_t* vb = valid ? valid + (long long)b * N : nullptr;
    int cnt = 0;
    for (int e = lane; e < n; e += 64) {
        const int p = (y0 + e / ww) * g.W + x0 + e % ww;
        if (labb[p] == ci && is_valid(vb, vkind, p)) ++cnt;
    }
    const int* sb = strays + (long long)b * N;
    for (int s = lane; s < nstray[b]; s += 64) {
        const int p = sb[s];
        if (labb[p] == ci && is_valid(vb, vkind, p)) ++cnt;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o, 64);
    if (lane == 0) npts[(long long)b * g.K + ci] = cnt;
}

// Starting xorshift state of every segment: the seed advanced by 120 draws
// per eligible segment before it.  One block of 1024 threads per image.
__global__ void __launch_bounds__(1024) k_ps_states(const int* __restrict__ npts, int K,
                                                    uint32_t seed, uint32_t* __restrict__ state) {
    __shared__ int wsum[16];
    __shared__ int carry;
    const int b = blockIdx.x, t = threadIdx.x, lane = t & 63, wave = t >> 6;
    if (t == 0) carry = 0;
    __syncthreads();
    for (int k0 = 0; k0 < K; k0 += 1024) {
        const int k = k0 + t;
        const int e = (k < K && npts[(long long)b * K + k] >= 12) ? 1 : 0;
        int v = e;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int u = __shfl_up(v, o, 64);
            if (lane >= o) v += u;
        }
        if (lane == 63) wsum[wave] = v;
        __syncthreads();
        int before = carry;
        for (int w = 0; w < wave; ++w) before += wsum[w];
        const int excl = before + v - e;  // eligible segments before k
        if (k < K) {
            uint32_t s = seed;
            for (int j = 0; j < JUMP_BITS; ++j)
                if ((excl >> j) & 1) s = jump_apply(c_jump.j[j], s);
            state[(long long)b * K + k] = s;
        }
        __syncthreads();
        if (t == 0) {
            int tot = 0;
            for (int w = 0; w < 16; ++w) tot += wsum[w];
            carry += tot;
        }
        __syncthreads();
    }
}

// _solve3, stereonet.py:5451-5463
__device__ bool solve3(const double A[9], const double bv[3], double out[3]) {
    const double a = A[0], bb = A[1], c = A[2], d = A[3], e = A[4], f = A[5], g = A[6],
                 h = A[7], k = A[8];
    const double det = a * (e * k - f * h) - bb * (d * k - f * g) + c * (d * h - e * g);
    if (fabs(det) < 1.0e-9) return false;
    out[0] = (bv[0] * (e * k - f * h) - bb * (bv[1] * k - f * bv[2]) + c * (bv[1] * h - e * bv[2])) / det;
    out[1] = (a * (bv[1] * k - f * bv[2]) - bv[0] * (d * k - f * g) + c * (d * bv[2] - bv[1] * g)) / det;
    out[2] = (a * (e * bv[2] - bv[1] * h) - bb * (d * bv[2] - bv[1] * g) + bv[0] * (d * h - e * g)) / det;
    return true;
}

// One wave per segment: member lists in LDS (window members merged with the
// segment's strays, ascending), RANSAC (lane 0 draws and solves, the wave
// counts), the inliers' least-squares plane (lane 0 sums in order), the snap.

__global__ void __launch_bounds__(64) k_ps_segment(const double* __restrict__ disp, SlicGeom g,
                                                   const double* __restrict__ cs,
                                                   const int* __restrict__ lab,
                                                   const int* __restrict__ strays,
                                                   const int* __restrict__ nstray,
                                                   const uint8_t* __restrict__ valid, int vkind,
                                                   const uint32_t* __restrict__ state, double thr,
                                                   double min_inl, double* __restrict__ out,
                                                   int* __restrict__ err) {
    __shared__ StrayList sl;
    // dynamic LDS, sized for the step: the valid members (ascending) then
    // the inliers, and their disparities.  The full member list is not kept:
    // the final snap re-scans the window (its writes are per pixel, in any
    // order), which leaves room for more segments per CU
    extern __shared__ __align__(16) unsigned char ps_dyn[];
    const int maxm = (2 * g.step + 1) * (2 * g.step + 1) + SCAP;
    double* pd = reinterpret_cast<double*>(ps_dyn);
    int* pts = reinterpret_cast<int*>(pd + maxm);
    __shared__ int mi[64];
    __shared__ int mv[64];
    __shared__ int counts[2];
    __shared__ double plane[4];
    const int ci = blockIdx.x, b = blockIdx.y, lane = threadIdx.x;
    const long long N = (long long)g.W * g.H;
    const double* c = cs + ((long long)b * g.K + ci) * 5;
    int x0, x1, y0, y1;
    window(c[3], c[4], g, x0, x1, y0, y1);
    const int ww = x1 - x0 + 1, n = ww * (y1 - y0 + 1);
    const int* labb = lab + (long long)b * N;
    const double* d = disp + (long long)b * N;
    const uint8_t* vb = valid ? valid + (long long)b * N : nullptr;
    if (!collect_strays(strays + (long long)b * N, nstray[b], labb, ci, &sl)) {
        if (lane == 0) atomicOr(err, 1);
        return;
    }
    // ordered member lists: without strays, each chunk's members are
    // appended by ballot compaction; with strays, lane 0 merges them in
    if (lane == 0) counts[0] = counts[1] = 0;
    __builtin_amdgcn_wave_barrier();
    const uint64_t below = (1ull << lane) - 1ull;
    int sp = 0;
    for (int e0 = 0; e0 < n; e0 += 64) {
        const int e = e0 + lane;
        bool m = false, v = false;
        int p = 0;
        if (e < n) {
            p = (y0 + e / ww) * g.W + x0 + e % ww;
            m = labb[p] == ci;
            v = m && is_valid(vb, vkind, p);
        }
        const uint64_t bm = __ballot(m), bv = __ballot(v);
        const int nm0 = counts[0], np0 = counts[1];
        if (sl.n == 0) {
            if (v) pts[np0 + __popcll(bv & below)] = p;
            __builtin_amdgcn_wave_barrier();
            if (lane == 0) {
                counts[0] = nm0 + __popcll(bm);
                counts[1] = np0 + __popcll(bv);
            }
        } else {
            if (m) {
                mi[__popcll(bm & below)] = p;
                mv[__popcll(bm & below)] = v ? 1 : 0;
            }
            __builtin_amdgcn_wave_barrier();
            if (lane == 0) {
                int nm = nm0, np = np0;
                const int c = __popcll(bm);
                for (int k = 0; k < c; ++k) {
                    const int q0 = mi[k];
                    while (sp < sl.n && sl.idx[sp] < q0) {
                        const int q = sl.idx[sp++];
                        ++nm;
                        if (is_valid(vb, vkind, q)) pts[np++] = q;
                    }
                    ++nm;
                    if (mv[k]) pts[np++] = q0;
                }
                counts[0] = nm;
                counts[1] = np;
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
    if (lane == 0) {
        int nm = counts[0], np = counts[1];
        while (sp < sl.n) {
            const int q = sl.idx[sp++];
            ++nm;
            if (is_valid(vb, vkind, q)) pts[np++] = q;
        }
        counts[0] = nm;
        counts[1] = np;
    }
    __builtin_amdgcn_wave_barrier();
    const int np = counts[1];
    if (np < 12) return;
    for (int k = lane; k < np; k += 64) pd[k] = d[pts[k]];
    __builtin_amdgcn_wave_barrier();
    // the plane test of stereonet.py:5535/5547 on staged points
    auto on_pt = [&](double a, double bb, double cc, int k) {
        const int i = pts[k];
        return fabs(a * (double)(i % g.W) + bb * (double)(i / g.W) + cc - pd[k]) <= thr;
    };
    // RANSAC, the trials side by side: lane t runs trial t, whose three
    // draws are steps 3t+1..3t+3 of the segment's xorshift stream, and counts
    // its inliers over all points.  The best trial is the highest count, the
    // earliest among equals (a later trial replaces the best only when
    // strictly better); trials whose 3x3 system is singular take no part
    int cnt = -1;
    double pl[3] = {0.0, 0.0, 0.0};
    if (lane < PS_TRIALS) {
        uint32_t s = state[(long long)b * g.K + ci];
        for (int q = 0; q < 3 * lane; ++q) s = xs_step(s);
        double A[9], bv[3];
        for (int q = 0; q < 3; ++q) {
            s = xs_step(s);
            const int i = pts[s % (uint32_t)np];
            A[3 * q] = (double)(i % g.W);
            A[3 * q + 1] = (double)(i / g.W);
            A[3 * q + 2] = 1.0;
            bv[q] = pd[s % (uint32_t)np];
        }
        if (solve3(A, bv, pl)) {
            cnt = 0;
            for (int k = 0; k < np; ++k) cnt += on_pt(pl[0], pl[1], pl[2], k) ? 1 : 0;
        }
    }
    int bestc = cnt, bt = lane;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const int oc = __shfl_xor(bestc, o, 64), ot = __shfl_xor(bt, o, 64);
        if (oc > bestc || (oc == bestc && ot < bt)) {
            bestc = oc;
            bt = ot;
        }
    }
    const bool have = bestc >= 0;
    double best[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) best[q] = __shfl(pl[q], bt, 64);
    if (!have || (double)bestc < min_inl * (double)np) return;
    // inliers, in order, compacted over pts by ballot
    {
        int base = 0;
        for (int k0 = 0; k0 < np; k0 += 64) {
            const int k = k0 + lane;
            const bool in = k < np && on_pt(best[0], best[1], best[2], k);
            const int q = in ? pts[k] : 0;
            const double dq = in ? pd[k] : 0.0;
            const uint64_t bi = __ballot(in);
            __builtin_amdgcn_wave_barrier();  // every read of this chunk before the writes
            if (in) {
                pts[base + __popcll(bi & below)] = q;
                pd[base + __popcll(bi & below)] = dq;
            }
            base += __popcll(bi);
            __builtin_amdgcn_wave_barrier();
        }
        if (lane == 0) counts[1] = base;
    }
    __builtin_amdgcn_wave_barrier();
    const int ni = counts[1];
    // _lsq_plane, stereonet.py:5466-5495: the coordinate sums are integers
    // (exact in any order); the disparity and the centred sums go in order
    long long isx = 0, isy = 0;
    for (int k = lane; k < ni; k += 64) {
        isx += pts[k] % g.W;
        isy += pts[k] / g.W;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        isx += __shfl_xor(isx, o, 64);
        isy += __shfl_xor(isy, o, 64);
    }
    if (lane == 0) {
        const double nn = (double)ni;
        const double sx = (double)isx, sy = (double)isy;
        double sd = 0.0;
        for (int k = 0; k < ni; ++k) sd += pd[k];
        const double mx = sx / nn, my = sy / nn, md = sd / nn;
        double sxx = 0.0, sxy = 0.0, syy = 0.0, sxd = 0.0, syd = 0.0;
        for (int k = 0; k < ni; ++k) {
            const double u = (double)(pts[k] % g.W) - mx;
            const double v = (double)(pts[k] / g.W) - my;
            const double w = pd[k] - md;
            sxx += u * u;
            sxy += u * v;
            syy += v * v;
            sxd += u * w;
            syd += v * w;
        }
        const double det = sxx * syy - sxy * sxy;
        if (fabs(det) < 1.0e-9) {
            plane[0] = 0.0;
            plane[1] = 0.0;
            plane[2] = md;
        } else {
            const double a = (sxd * syy - syd * sxy) / det;
            const double bb = (syd * sxx - sxd * sxy) / det;
            plane[0] = a;
            plane[1] = bb;
            plane[2] = md - a * mx - bb * my;
        }
    }
    __builtin_amdgcn_wave_barrier();
    const double fa = plane[0], fb = plane[1], fc = plane[2];
    double* ob = out + (long long)b * N;
    auto snap = [&](int i) {
        const double v = fa * (double)(i % g.W) + fb * (double)(i / g.W) + fc;
        if (fabs(v - d[i]) <= thr || (vb && !is_valid(vb, vkind, i))) ob[i] = v > 0.0 ? v : 0.0;
    };
    for (int e = lane; e < n; e += 64) {  // the window's members
        const int i = (y0 + e / ww) * g.W + x0 + e % ww;
        if (labb[i] == ci) snap(i);
    }
    for (int k = lane; k < sl.n; k += 64) snap(sl.idx[k]);  // and the strays
}

__global__ void k_copy_i32(const int* __restrict__ a, int* __restrict__ o) {
    if (threadIdx.x == 0) o[0] = a[0];
}

__global__ void k_copy_i32n(const int* __restrict__ a, int* __restrict__ o, long long n) {
    const long long i = blockIdx.x * 256LL + threadIdx.x;
    if (i < n) o[i] = a[i];
}

__global__ void k_copy_f64(const double* __restrict__ a, double* __restrict__ o, long long n) {
    const long long i = blockIdx.x * 256LL + threadIdx.x;
    if (i < n) o[i] = a[i];
}

struct PsWs {
    unsigned long long* dist;
    int *lab2, *lab, *strays, *nstray, *npts, *err;
    double *cs0, *cs1;
    uint32_t* state;
    size_t bytes;
};

PsWs carve_ps(void* ws, int B, long long N, int K) {
    PsWs w{};
    size_t off = 0;
    auto take = [&](size_t n) -> void* {
        off = (off + 255) & ~(size_t)255;
        void* p = ws ? (char*)ws + off : nullptr;
        off += n;
        return p;
    };
    w.dist = (unsigned long long*)take(sizeof(unsigned long long) * B * N);
    w.lab2 = (int*)take(sizeof(int) * B * N);
    w.lab = (int*)take(sizeof(int) * B * N);
    w.strays = (int*)take(sizeof(int) * B * N);
    w.nstray = (int*)take(sizeof(int) * B);
    w.npts = (int*)take(sizeof(int) * (size_t)B * K);
    w.err = (int*)take(sizeof(int));
    w.cs0 = (double*)take(sizeof(double) * 5 * (size_t)B * K);
    w.cs1 = (double*)take(sizeof(double) * 5 * (size_t)B * K);
    w.state = (uint32_t*)take(sizeof(uint32_t) * (size_t)B * K);
    w.bytes = off;
    return w;
}

SlicGeom slic_geom(int W, int H, int step, double compactness = 10.0) {
    SlicGeom g{};
    g.W = W;
    g.H = H;
    g.step = step;
    g.KX = (W - step / 2 + step - 1) / step;
    const int KY = (H - step / 2 + step - 1) / step;
    g.KX = W > step / 2 ? g.KX : 0;
    g.K = (H > step / 2 ? KY : 0) * g.KX;
    const double q = compactness / step;
    g.m2 = q * q;
    return g;
}

// SLIC; leaves labels in w.lab, the strays of the last pass in w.strays,
// and the centres of the last assignment in *last.
template <typename RT>
int run_slic(const RT* rgb, int B, const SlicGeom& g, int iters, const PsWs& w, double** last,
             hipStream_t st) {
    const long long N = (long long)g.W * g.H, BN = N * B;
    const unsigned nb = (unsigned)((BN + 255) / 256);
    hipLaunchKernelGGL(k_fill_i32, dim3(nb), dim3(256), 0, st, w.lab, BN, -1);
    hipLaunchKernelGGL((k_slic_init<RT>), dim3((g.K + 255) / 256, B), dim3(256), 0, st, rgb, g, B,
                       w.cs0);
    SV_CHECK_LAUNCH();
    double* cin = w.cs0;
    double* cout = w.cs1;
    for (int it = 0; it < iters; ++it) {
        hipLaunchKernelGGL(k_slic_clear, dim3(nb), dim3(256), 0, st, w.dist, w.lab2, w.nstray, BN, B);
        hipLaunchKernelGGL((k_slic_assign<RT, 0>), dim3(g.K, B), dim3(256), 0, st, rgb, g, cin,
                           w.dist, w.lab2);
        hipLaunchKernelGGL((k_slic_assign<RT, 1>), dim3(g.K, B), dim3(256), 0, st, rgb, g, cin,
                           w.dist, w.lab2);
        hipLaunchKernelGGL(k_slic_commit, dim3(nb), dim3(256), 0, st, w.lab2, w.lab, w.strays,
                           w.nstray, N, B);
        hipLaunchKernelGGL((k_slic_update<RT>), dim3(g.K, B), dim3(64), 0, st, rgb, g, cin, cout,
                           w.lab, w.strays, w.nstray, w.err);
        SV_CHECK_LAUNCH();
        if (it < iters - 1) std::swap(cin, cout);
    }
    *last = cin;  // the centres of the last assignment
    return SV_OK;
}

template <typename RT>
int plane_snap_t(const double* disp, const RT* rgb, const uint8_t* valid, int vkind, int B, int W,
                 int H, int step, double thr, double min_inl, uint32_t seed, double* out, void* ws,
                 size_t ws_bytes, hipStream_t st) {
    const SlicGeom g = slic_geom(W, H, step);
    const long long N = (long long)W * H;
    if (g.K >= (1 << JUMP_BITS))
        return sv::set_error(SV_EUNSUPPORTED, "sv_plane_snap: %d segments (at most %d)", g.K,
                             (1 << JUMP_BITS) - 1);
    PsWs w = carve_ps(ws, B, N, g.K > 0 ? g.K : 1);
    if (!ws || ws_bytes < w.bytes)
        return sv::set_error(SV_EWORKSPACE, "sv_plane_snap needs %zu bytes", w.bytes);
    hipLaunchKernelGGL(k_copy_f64, dim3((unsigned)((N * B + 255) / 256)), dim3(256), 0, st, disp,
                       out, N * B);
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(256), 0, st, w.err, 1LL, 0);
    SV_CHECK_LAUNCH();
    if (g.K == 0) return SV_OK;
    double* last = nullptr;
    int rc = run_slic(rgb, B, g, 5, w, &last, st);
    if (rc) return rc;
    hipLaunchKernelGGL(k_ps_count, dim3(g.K, B), dim3(64), 0, st, g, last, w.lab, w.strays,
                       w.nstray, valid, vkind, w.npts);
    hipLaunchKernelGGL(k_ps_states, dim3(B), dim3(1024), 0, st, w.npts, g.K, seed, w.state);
    const size_t shm = (size_t)((2 * step + 1) * (2 * step + 1) + SCAP) * (8 + 4);
    hipLaunchKernelGGL(k_ps_segment, dim3(g.K, B), dim3(64), shm, st, disp, g, last, w.lab,
                       w.strays, w.nstray, valid, vkind, w.state, thr, min_inl, out, w.err);
    SV_CHECK_LAUNCH();
    return SV_OK;
}

}  // namespace

extern "C" {

int sv_plane_snap_workspace_bytes(int B, int W, int H, int step, size_t* bytes) {
    SV_REQUIRE(bytes && B > 0 && W > 0 && H > 0 && step > 0, "sv_plane_snap_workspace_bytes: bad arguments");
    const SlicGeom g = slic_geom(W, H, step);
    *bytes = carve_ps(nullptr, B, (long long)W * H, g.K > 0 ? g.K : 1).bytes;
    return SV_OK;
}

int sv_slic(const void* rgb, int rgb_dtype, int B, int W, int H, int step, double compactness,
            int iters, int* labels, int* nseg, int* status_out, void* ws, size_t ws_bytes,
            void* stream) {
    SV_REQUIRE(rgb && labels && B > 0 && W > 0 && H > 0 && iters >= 0, "sv_slic: bad arguments");
    SV_REQUIRE(step >= 2 && step <= 16, "sv_slic: step %d (2..16 supported)", step);
    SV_REQUIRE((long long)W * H < INT_MAX, "sv_slic: image too large");
    hipStream_t st = (hipStream_t)stream;
    const SlicGeom g = slic_geom(W, H, step, compactness);
    const long long N = (long long)W * H;
    PsWs w = carve_ps(ws, B, N, g.K > 0 ? g.K : 1);
    if (!ws || ws_bytes < w.bytes) return sv::set_error(SV_EWORKSPACE, "sv_slic needs %zu bytes", w.bytes);
    if (nseg) *nseg = g.K;
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(256), 0, st, w.err, 1LL, 0);
    if (g.K == 0 || iters == 0) {
        hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)((N * B + 255) / 256)), dim3(256), 0, st,
                           labels, N * B, -1);
    } else {
        double* last = nullptr;
        int rc = rgb_dtype == SV_U8 ? run_slic((const uint8_t*)rgb, B, g, iters, w, &last, st)
               : rgb_dtype == SV_F64 ? run_slic((const double*)rgb, B, g, iters, w, &last, st)
               : sv::set_error(SV_EINVAL, "sv_slic: rgb must be u8 or f64 (dtype %d)", rgb_dtype);
        if (rc) return rc;
        hipLaunchKernelGGL(k_copy_i32n, dim3((unsigned)((N * B + 255) / 256)), dim3(256), 0, st,
                           w.lab, labels, N * B);
    }
    if (status_out) hipLaunchKernelGGL(k_copy_i32, dim3(1), dim3(64), 0, st, w.err, status_out);
    SV_CHECK_LAUNCH();
    return SV_OK;
}

int sv_plane_snap(const double* disp, const void* rgb, int rgb_dtype, const uint8_t* valid,
                  int valid_kind, int B, int W, int H, int step, double thr, double min_inl,
                  int trials, unsigned int seed, double* out, int* status_out, void* ws,
                  size_t ws_bytes, void* stream) {
    SV_REQUIRE(disp && rgb && out && B > 0 && W > 0 && H > 0, "sv_plane_snap: bad arguments");
    SV_REQUIRE(disp != out, "sv_plane_snap: out must not alias disp");
    SV_REQUIRE(step >= 2 && step <= 16, "sv_plane_snap: step %d (2..16 supported)", step);
    SV_REQUIRE(trials == PS_TRIALS, "sv_plane_snap: %d trials (40 supported)", trials);
    SV_REQUIRE(valid_kind == 0 || valid_kind == 1, "sv_plane_snap: valid_kind %d", valid_kind);
    SV_REQUIRE((long long)W * H < INT_MAX, "sv_plane_snap: image too large");
    hipStream_t st = (hipStream_t)stream;
    int rc;
    if (rgb_dtype == SV_U8)
        rc = plane_snap_t(disp, (const uint8_t*)rgb, valid, valid_kind, B, W, H, step, thr,
                          min_inl, (uint32_t)seed, out, ws, ws_bytes, st);
    else if (rgb_dtype == SV_F64)
        rc = plane_snap_t(disp, (const double*)rgb, valid, valid_kind, B, W, H, step, thr,
                          min_inl, (uint32_t)seed, out, ws, ws_bytes, st);
    else
        return sv::set_error(SV_EINVAL, "sv_plane_snap: rgb must be u8 or f64 (dtype %d)", rgb_dtype);
    if (rc) return rc;
    // the overflow flag of the per-segment stray lists, when asked for
    if (status_out) {
        const SlicGeom g = slic_geom(W, H, step);
        PsWs w = carve_ps(ws, B, (long long)W * H, g.K > 0 ? g.K : 1);
        hipLaunchKernelGGL(k_copy_i32, dim3(1), dim3(64), 0, st, w.err, status_out);
        SV_CHECK_LAUNCH();
    }
    return SV_OK;
}

}  // extern "C"
