// Disparity to 3D on the GPU (stereonet.py:1951-2035, :5566-5862): the
// pinhole camera's depth and unprojection, disparity_to_points, the
// robust disparity floor (a k-th smallest value), the background layer's
// min filter with source index, its smoothing and mask, and the viewer's
// dense grid mesh (vertices, texture coordinates, triangles).
//
// Everything is float64 with the reference's operation order and no
// contraction, so values are bit-identical.  The grid mesh keeps the
// reference's vertex and triangle order (row by row, left to right, the two
// triangles of a cell in turn): each grid row is one wave that counts its
// survivors by ballot, one block scans the row counts, and the same waves
// then write at their offsets.
#include <cstdint>
#include <cstring>

#include "common.h"

namespace {

struct Cam {
    double fx, cx, cy, baseline, doffs;
};

// Camera.depth_from_disparity, :1982-1986
__device__ __forceinline__ double depth_of(const Cam& c, double d) {
    const double den = d + c.doffs;
    if (den < 1.0e-6) return 0.0;
    return c.fx * c.baseline / den;
}

// disparity_to_points, :2006-2021 (None -> valid 0, point NaN)
__global__ void k_points(const double* __restrict__ disp, long long total, int W, int H, Cam cam,
                         double min_disp, double max_depth, double* __restrict__ pts,
                         uint8_t* __restrict__ valid) {
    const long long i = blockIdx.x * 256LL + threadIdx.x;
    if (i >= total) return;
    const double d = disp[i];
    const int x = (int)(i % W), y = (int)((i / W) % H);
    double Z = 0.0;
    bool ok = !(d < min_disp);
    if (ok) {
        Z = depth_of(cam, d);
        ok = !(Z <= 0.0 || Z > max_depth);
    }
    valid[i] = ok;
    const double nan = __longlong_as_double(0x7FF8000000000000LL);
    pts[3 * i] = ok ? ((double)x - cam.cx) * Z / cam.fx : nan;
    pts[3 * i + 1] = ok ? ((double)y - cam.cy) * Z / cam.fx : nan;
    pts[3 * i + 2] = ok ? Z : nan;
}

// ------------------------------------------------- k-th smallest float64
// Radix select over order-preserving 64-bit keys in six digit passes
// (five of 11 bits, one of 9).  Each pass is two launches: blocks count the
// keys that share the digits chosen so far in an LDS histogram and add it
// to the set's global one; then one block per set picks the digit holding
// the k-th key and clears the histogram for the next pass.
constexpr int kSelBits = 11;
constexpr int kSelBins = 1 << kSelBits;
constexpr int kSelPasses = 6;
constexpr int kSelPer = kSelBins / 256;  // bins per thread in the pick

struct Sel64 {
    unsigned long long prefix;
    long long k, n;
};

__device__ __forceinline__ unsigned long long key_of(double v) {
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    return (u >> 63) ? ~u : (u | 0x8000000000000000ULL);
}

__device__ __forceinline__ double value_of(unsigned long long k) {
    const unsigned long long u = (k >> 63) ? (k & 0x7FFFFFFFFFFFFFFFULL) : ~k;
    return __longlong_as_double((long long)u);
}

__device__ __forceinline__ void sel_digit(int pass, int& shift, int& width) {
    width = pass < kSelPasses - 1 ? kSelBits : 64 - kSelBits * (kSelPasses - 1);
    shift = pass < kSelPasses - 1 ? 64 - kSelBits * (pass + 1) : 0;
}

// map blockIdx.y: n values (ns[map * nstride] when ns is given, k = n / 2
// then; an empty map yields empty_value)
__global__ void k_sel64_init(unsigned* __restrict__ hist, Sel64* __restrict__ st, long long n,
                             long long k, const int* __restrict__ ns, int nstride,
                             double empty_value, double* __restrict__ out) {
    const int b = blockIdx.y, t = blockIdx.x * 256 + threadIdx.x;
    if (t < kSelBins) hist[(size_t)b * kSelBins + t] = 0;
    if (t == 0) {
        const long long nb = ns ? (long long)ns[(size_t)b * nstride] : n;
        st[b].prefix = 0;
        st[b].n = nb;
        st[b].k = ns ? nb / 2 : k;
        if (nb == 0) out[b] = empty_value;
    }
}

__global__ void __launch_bounds__(256) k_sel64_pass(const double* __restrict__ x, long long mstride,
                                                    long long stride, int pass,
                                                    Sel64* __restrict__ sts,
                                                    unsigned* __restrict__ hists,
                                                    double* __restrict__ out) {
    __shared__ unsigned h[kSelBins];
    const int t = threadIdx.x, mb = blockIdx.y;
    Sel64* st = sts + mb;
    unsigned* hist = hists + (size_t)mb * kSelBins;
    const long long n = st->n;
    if (n == 0) return;
    x += mb * mstride;
    int shift, width;
    sel_digit(pass, shift, width);
    const unsigned long long hi = pass ? ~0ULL << (shift + width) : 0ULL;
    const unsigned long long want = st->prefix & hi;
    const unsigned dmask = (1u << width) - 1u;
    for (int b = t; b < kSelBins; b += 256) h[b] = 0;
    __syncthreads();
    // wave-uniform loop; lanes sharing a bin add once (up to four bins per
    // wave and step: the early digits put most keys in a handful of bins),
    // the rest add one by one
    const int lane = t & 63;
    for (long long i0 = blockIdx.x * 256LL; i0 < n; i0 += (long long)gridDim.x * 256) {
        const long long i = i0 + t;
        unsigned bin = 0;
        bool pend = false;
        if (i < n) {
            const unsigned long long key = key_of(x[i * stride]);
            pend = (key & hi) == want;
            bin = (unsigned)(key >> shift) & dmask;
        }
#pragma unroll
        for (int it = 0; it < 4; ++it) {
            const unsigned long long act = __ballot(pend);
            if (!act) break;
            const int leader = __ffsll((long long)act) - 1;
            const unsigned lb = __shfl(bin, leader, 64);
            const bool eq = pend && bin == lb;
            const unsigned long long same = __ballot(eq);
            if (lane == leader) atomicAdd(&h[lb], (unsigned)__popcll(same));
            if (eq) pend = false;
        }
        if (pend) atomicAdd(&h[bin], 1u);
    }
    __syncthreads();
    for (int b = t; b < kSelBins; b += 256)
        if (h[b]) atomicAdd(&hist[b], h[b]);
}

// one block per set: the digit holding the k-th key, then the histogram
// cleared for the next pass (a kernel of its own: the boundary orders it
// after every block's histogram, with no device-wide fence in the counting)
__global__ void __launch_bounds__(256) k_sel64_pick(int pass, Sel64* __restrict__ sts,
                                                    unsigned* __restrict__ hists,
                                                    double* __restrict__ out) {
    __shared__ long long s[256];
    const int t = threadIdx.x, mb = blockIdx.x;
    Sel64* st = sts + mb;
    unsigned* hist = hists + (size_t)mb * kSelBins;
    if (st->n == 0) return;
    int shift, width;
    sel_digit(pass, shift, width);
    // thread t owns bins [kSelPer t, kSelPer (t + 1))
    unsigned c[kSelPer];
    long long sum = 0;
#pragma unroll
    for (int b = 0; b < kSelPer; ++b) c[b] = hist[t * kSelPer + b];
#pragma unroll
    for (int b = 0; b < kSelPer; ++b) sum += c[b];
    s[t] = sum;
    __syncthreads();
    for (int o = 1; o < 256; o <<= 1) {
        const long long v = t >= o ? s[t - o] : 0;
        __syncthreads();
        s[t] += v;
        __syncthreads();
    }
    const long long k = st->k, before = t ? s[t - 1] : 0;
    __syncthreads();  // every thread has read st->k
    if (k >= before && k < s[t]) {
        long long run = before;
#pragma unroll
        for (int b = 0; b < kSelPer; ++b) {
            if (k < run + c[b]) {
                const unsigned long long pre =
                    st->prefix | ((unsigned long long)(t * kSelPer + b) << shift);
                st->prefix = pre;
                st->k = k - run;
                if (pass == kSelPasses - 1) out[mb] = value_of(pre);
                break;
            }
            run += c[b];
        }
    }
#pragma unroll
    for (int b = 0; b < kSelPer; ++b) hist[t * kSelPer + b] = 0;
}

// ------------------------------------------------------ background layer
// min_filter_arg, :5566-5605, rows: the first strict minimum over the
// clipped window [x - r, x + r] and its flat index
__global__ void k_minrow(const double* __restrict__ v, long long rows, int W, int H, int r,
                         double* __restrict__ out, int* __restrict__ idx) {
    const long long i = blockIdx.x * 256LL + threadIdx.x;
    if (i >= rows * W) return;
    const int x = (int)(i % W);
    const long long row = i - x;
    const int a = x - r > 0 ? x - r : 0, b = x + r < W - 1 ? x + r : W - 1;
    long long bi = row + a;
    double bv = v[bi];
    for (int xx = a + 1; xx <= b; ++xx)
        if (v[row + xx] < bv) {
            bv = v[row + xx];
            bi = row + xx;
        }
    out[i] = bv;
    idx[i] = (int)(bi % ((long long)W * H));  // flat index within its own image
}

// the same with the row staged in LDS: one block per image row (W <= 4096)
constexpr int kRowMax = 4096;

__global__ void __launch_bounds__(256) k_minrow_lds(const double* __restrict__ v, int W, int H,
                                                    int r, double* __restrict__ out,
                                                    int* __restrict__ idx) {
    __shared__ double row[kRowMax];
    const long long base = (long long)blockIdx.x * W;
    for (int x = threadIdx.x; x < W; x += 256) row[x] = v[base + x];
    __syncthreads();
    const int rowoff = (int)(base % ((long long)W * H));  // the row's first in-image index
    for (int x = threadIdx.x; x < W; x += 256) {
        const int a = x - r > 0 ? x - r : 0, b = x + r < W - 1 ? x + r : W - 1;
        int bi = a;
        double bv = row[a];
        for (int xx = a + 1; xx <= b; ++xx)
            if (row[xx] < bv) {
                bv = row[xx];
                bi = xx;
            }
        out[base + x] = bv;
        idx[base + x] = rowoff + bi;
    }
}

// columns: the same over [y - rv, y + rv] of the row result, index carried
__global__ void k_mincol(const double* __restrict__ rv_, const int* __restrict__ ri, int B, int W,
                         int H, int rv, double* __restrict__ out, int* __restrict__ idx) {
    const long long i = blockIdx.x * 256LL + threadIdx.x;
    if (i >= (long long)B * W * H) return;
    const long long img = i / ((long long)W * H), p = i % ((long long)W * H);
    const int x = (int)(p % W), y = (int)(p / W);
    const double* v = rv_ + img * W * H;
    const int a = y - rv > 0 ? y - rv : 0, b = y + rv < H - 1 ? y + rv : H - 1;
    long long bi = (long long)a * W + x;
    double bv = v[bi];
    for (int yy = a + 1; yy <= b; ++yy) {
        const long long j = (long long)yy * W + x;
        if (v[j] < bv) {
            bv = v[j];
            bi = j;
        }
    }
    out[i] = bv;
    idx[i] = ri[img * W * H + bi];
}

// smooth_background, :5608-5634: row box means (sums left to right), then
// column box means (top to bottom), never in front of the pixel's surface
__global__ void k_smooth_row(const double* __restrict__ m, long long rows, int W, int r,
                             double* __restrict__ t) {
    const long long i = blockIdx.x * 256LL + threadIdx.x;
    if (i >= rows * W) return;
    const int x = (int)(i % W);
    const long long row = i - x;
    const int a = x - r > 0 ? x - r : 0, b = x + r < W - 1 ? x + r : W - 1;
    double s = 0.0;
    for (int xx = a; xx <= b; ++xx) s += m[row + xx];
    t[i] = s / (double)(b - a + 1);
}

__global__ void __launch_bounds__(256) k_smooth_row_lds(const double* __restrict__ m, int W, int r,
                                                        double* __restrict__ t) {
    __shared__ double row[kRowMax];
    const long long base = (long long)blockIdx.x * W;
    for (int x = threadIdx.x; x < W; x += 256) row[x] = m[base + x];
    __syncthreads();
    for (int x = threadIdx.x; x < W; x += 256) {
        const int a = x - r > 0 ? x - r : 0, b = x + r < W - 1 ? x + r : W - 1;
        double s = 0.0;
        for (int xx = a; xx <= b; ++xx) s += row[xx];
        t[base + x] = s / (double)(b - a + 1);
    }
}

__global__ void k_smooth_col(const double* __restrict__ t, const double* __restrict__ disp, int B,
                             int W, int H, int r, double* __restrict__ out) {
    const long long i = blockIdx.x * 256LL + threadIdx.x;
    if (i >= (long long)B * W * H) return;
    const long long img = (i / ((long long)W * H)) * W * H, p = i % ((long long)W * H);
    const int x = (int)(p % W), y = (int)(p / W);
    const int a = y - r > 0 ? y - r : 0, b = y + r < H - 1 ? y + r : H - 1;
    double s = 0.0;
    for (int yy = a; yy <= b; ++yy) s += t[img + (long long)yy * W + x];
    const double v = s / (double)(b - a + 1);
    out[i] = v < disp[i] ? v : disp[i];
}

// the same with a tile of 64 columns x 32 rows (plus r rows above and below)
// staged in LDS: each input is read from HBM once per tile instead of once
// per output row that covers it
constexpr int kColTile = 32, kColRMax = 32;

__global__ void __launch_bounds__(256) k_smooth_col_lds(const double* __restrict__ t,
                                                        const double* __restrict__ disp, int W,
                                                        int H, int r, double* __restrict__ out) {
    extern __shared__ double tile[];  // (kColTile + 2r) rows x 64 columns
    const int cx = threadIdx.x & 63, ry = threadIdx.x >> 6;
    const int x = blockIdx.x * 64 + cx, y0 = blockIdx.y * kColTile;
    const long long img = (long long)blockIdx.z * W * H;
    const int ya = y0 - r > 0 ? y0 - r : 0;
    const int ylast = y0 + kColTile - 1 < H - 1 ? y0 + kColTile - 1 : H - 1;
    const int yb = ylast + r < H - 1 ? ylast + r : H - 1;
    for (int rr = ry; rr <= yb - ya; rr += 4)
        tile[rr * 64 + cx] = x < W ? t[img + (long long)(ya + rr) * W + x] : 0.0;
    __syncthreads();
    if (x >= W) return;
    for (int y = y0 + ry; y <= ylast; y += 4) {
        const int a = y - r > 0 ? y - r : 0, b = y + r < H - 1 ? y + r : H - 1;
        double s = 0.0;
        for (int yy = a; yy <= b; ++yy) s += tile[(yy - ya) * 64 + cx];
        const double v = s / (double)(b - a + 1);
        const long long i = img + (long long)y * W + x;
        out[i] = v < disp[i] ? v : disp[i];
    }
}

// background_region, :5637-5671: the foreground side of a depth edge
__global__ void k_bg_thresh(const double* __restrict__ disp, const double* __restrict__ m,
                            long long total, double jump_ratio, double min_jump,
                            uint8_t* __restrict__ mask) {
    const long long i = blockIdx.x * 256LL + threadIdx.x;
    if (i >= total) return;
    const double d = disp[i];
    double thr = jump_ratio * d;
    if (thr < min_jump) thr = min_jump;
    mask[i] = d - m[i] > thr;
}

// one 4-neighbour growth step
__global__ void k_bg_grow(const uint8_t* __restrict__ in, int B, int W, int H,
                          uint8_t* __restrict__ out) {
    const long long i = blockIdx.x * 256LL + threadIdx.x;
    if (i >= (long long)B * W * H) return;
    const long long p = i % ((long long)W * H);
    const int x = (int)(p % W), y = (int)(p / W);
    out[i] = in[i] || (x > 0 && in[i - 1]) || (x < W - 1 && in[i + 1]) || (y > 0 && in[i - W]) ||
             (y < H - 1 && in[i + W]);
}

// ------------------------------------------------------------ grid mesh
struct Grid {
    const double* val;      // disparity (main) or background disparity (fill), map-major
    const uint8_t* region;  // fill: the background mask; main: null
    const int* src;         // fill: source pixel of each background value
    const double* min_disp; // per map
    int W, H, step, nx, ny, nx0, ny0, is_fill;
    Cam cam;
};

// where map b's vertices and triangles go: its own region of vcap vertices
// and tcap triangles, after base[b * 4 + {0, 1}] (the main build's counts)
// when building the fill
struct Out {
    double* pos;
    double* uv;
    uint8_t* flag;
    double* dv;
    int* idx;
    long long vcap, tcap;
};

__device__ __forceinline__ bool grid_vertex(const Grid& g, int b, int i, int j, double& d,
                                            double& Z, int& x, int& y, int& p) {
    x = i < g.nx0 ? i * g.step : g.W - 1;
    y = j < g.ny0 ? j * g.step : g.H - 1;
    p = y * g.W + x;
    const long long q = (long long)b * g.W * g.H + p;
    if (g.region && !g.region[q]) return false;
    d = g.val[q];
    if (!(d >= g.min_disp[b])) return false;
    if (!g.is_fill && (double)x - d < 0.0) return false;
    Z = depth_of(g.cam, d);
    return !(Z <= 0.0);
}

// grid_mesh's build(), :5806-5829: one wave per (grid row, map).  Count pass
// (WRITE = false) -> cnt[map][row]; write pass -> the map's vid table and
// vertex arrays at offs[map][row] (the scan of the counts, plus the
// vertices built before)
template <bool WRITE>
__global__ void __launch_bounds__(64) k_grid_vert(Grid g, Out o, int* __restrict__ cnt,
                                                  const int* __restrict__ offs,
                                                  int* __restrict__ vid) {
    const int j = blockIdx.x, b = blockIdx.y, l = threadIdx.x;
    const size_t row = (size_t)b * g.ny + j;
    int run = WRITE ? offs[row] : 0;
    int* vt = vid + (size_t)b * g.nx * g.ny;
    const size_t vb = (size_t)b * o.vcap;
    for (int i0 = 0; i0 < g.nx; i0 += 64) {
        const int i = i0 + l;
        double d = 0.0, Z = 0.0;
        int x = 0, y = 0, p = 0;
        const bool v = i < g.nx && grid_vertex(g, b, i, j, d, Z, x, y, p);
        const unsigned long long bal = __ballot(v);
        if (WRITE && i < g.nx) {
            if (v) {
                const int id = run + __popcll(bal & ((1ULL << l) - 1ULL));
                vt[j * g.nx + i] = id;
                const size_t a = vb + id;
                o.pos[3 * a] = ((double)x - g.cam.cx) * Z / g.cam.fx;
                o.pos[3 * a + 1] = ((double)y - g.cam.cy) * Z / g.cam.fx;
                o.pos[3 * a + 2] = Z;
                int ux = x, uy = y;
                if (g.is_fill) {
                    const int sp = g.src[(size_t)b * g.W * g.H + p];
                    ux = sp % g.W;
                    uy = sp / g.W;
                }
                o.uv[2 * a] = ((double)ux + 0.5) / (double)g.W;
                o.uv[2 * a + 1] = ((double)uy + 0.5) / (double)g.H;
                o.flag[a] = (uint8_t)g.is_fill;
                o.dv[a] = d;
            } else {
                vt[j * g.nx + i] = -1;
            }
        }
        run += __popcll(bal);
    }
    if (!WRITE && l == 0) cnt[row] = run;
}

__device__ __forceinline__ double fmax_(double a, double b) { return a > b ? a : b; }
__device__ __forceinline__ double fmin_(double a, double b) { return a < b ? a : b; }

// ok(a, b, c), :5831-5839
__device__ __forceinline__ bool tri_ok(int a, int b, int c, const double* dv, double jump,
                                       double jump_rel) {
    if (a < 0 || b < 0 || c < 0) return false;
    const double da = dv[a], db = dv[b], dc = dv[c];
    const double lo = fmin_(da, fmin_(db, dc)), hi = fmax_(da, fmax_(db, dc));
    return hi - lo <= fmax_(jump, jump_rel * hi);
}

// the cells of grid row j of map b (one wave): (a, c, b) then (b, c, e)
// when they pass
template <bool WRITE>
__global__ void __launch_bounds__(64) k_grid_tri(const int* __restrict__ vid, int nx, int ny,
                                                 Out o, double jump, double jump_rel,
                                                 int* __restrict__ cnt, const int* __restrict__ offs) {
    const int j = blockIdx.x, mb = blockIdx.y, l = threadIdx.x;
    const size_t row = (size_t)mb * (ny - 1) + j;
    int run = WRITE ? offs[row] : 0;
    const int* vt = vid + (size_t)mb * nx * ny;
    const double* dv = o.dv + (size_t)mb * o.vcap;
    int* idx = o.idx + 3 * (size_t)mb * o.tcap;
    const unsigned long long lt = (1ULL << l) - 1ULL;
    for (int i0 = 0; i0 < nx - 1; i0 += 64) {
        const int i = i0 + l;
        int a = -1, b = -1, c = -1, e = -1;
        if (i < nx - 1) {
            a = vt[j * nx + i];
            b = vt[j * nx + i + 1];
            c = vt[(j + 1) * nx + i];
            e = vt[(j + 1) * nx + i + 1];
        }
        const bool o1 = tri_ok(a, b, c, dv, jump, jump_rel);
        const bool o2 = tri_ok(b, c, e, dv, jump, jump_rel);
        const unsigned long long b1 = __ballot(o1), b2 = __ballot(o2);
        if (WRITE) {
            int t = run + __popcll(b1 & lt) + __popcll(b2 & lt);
            if (o1) {
                idx[3 * (size_t)t] = a;
                idx[3 * (size_t)t + 1] = c;
                idx[3 * (size_t)t + 2] = b;
                ++t;
            }
            if (o2) {
                idx[3 * (size_t)t] = b;
                idx[3 * (size_t)t + 1] = c;
                idx[3 * (size_t)t + 2] = e;
            }
        }
        run += __popcll(b1) + __popcll(b2);
    }
    if (!WRITE && l == 0) cnt[row] = run;
}

// exclusive scan of map blockIdx.x's n row counts (one block per map),
// offset by base[map * 4] when given; total[map * 4] = the sum
__global__ void __launch_bounds__(1024) k_scan_rows(const int* __restrict__ cnts, int n,
                                                    const int* __restrict__ base,
                                                    int* __restrict__ offss, int* __restrict__ total) {
    __shared__ int s[1024];
    const int t = threadIdx.x, mb = blockIdx.x, seg = (n + 1023) / 1024;
    const int* cnt = cnts + (size_t)mb * n;
    int* offs = offss + (size_t)mb * n;
    const int lo = t * seg < n ? t * seg : n, hi = lo + seg < n ? lo + seg : n;
    int sum = 0;
    for (int i = lo; i < hi; ++i) sum += cnt[i];
    s[t] = sum;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const int v = t >= o ? s[t - o] : 0;
        __syncthreads();
        s[t] += v;
        __syncthreads();
    }
    int run = (t ? s[t - 1] : 0) + (base ? base[(size_t)mb * 4] : 0);
    for (int i = lo; i < hi; ++i) {
        offs[i] = run;
        run += cnt[i];
    }
    if (t == 1023) total[(size_t)mb * 4] = s[1023];
}

unsigned blocks(long long n) { return (unsigned)((n + 255) / 256); }

// row minimum / row mean over B*H rows: LDS-staged rows when they fit
void launch_minrow(hipStream_t s, const double* v, int B, int W, int H, int r, double* out,
                   int* idx) {
    const long long rows = (long long)B * H;
    if (W <= kRowMax && rows <= 0x7FFFFFFFLL)
        hipLaunchKernelGGL(k_minrow_lds, dim3((unsigned)rows), dim3(256), 0, s, v, W, H, r, out, idx);
    else
        hipLaunchKernelGGL(k_minrow, dim3(blocks(rows * W)), dim3(256), 0, s, v, rows, W, H, r, out,
                           idx);
}

void launch_smooth_col(hipStream_t s, const double* t, const double* disp, int B, int W, int H,
                       int r, double* out) {
    if (r <= kColRMax && B <= 65535) {
        const dim3 grid((unsigned)((W + 63) / 64), (unsigned)((H + kColTile - 1) / kColTile), B);
        const size_t shm = (size_t)(kColTile + 2 * r) * 64 * sizeof(double);
        hipLaunchKernelGGL(k_smooth_col_lds, grid, dim3(256), shm, s, t, disp, W, H, r, out);
    } else {
        const long long total = (long long)B * W * H;
        hipLaunchKernelGGL(k_smooth_col, dim3(blocks(total)), dim3(256), 0, s, t, disp, B, W, H, r,
                           out);
    }
}

void launch_smooth_row(hipStream_t s, const double* m, int B, int W, int H, int r, double* t) {
    const long long rows = (long long)B * H;
    if (W <= kRowMax && rows <= 0x7FFFFFFFLL)
        hipLaunchKernelGGL(k_smooth_row_lds, dim3((unsigned)rows), dim3(256), 0, s, m, W, r, t);
    else
        hipLaunchKernelGGL(k_smooth_row, dim3(blocks(rows * W)), dim3(256), 0, s, m, rows, W, r, t);
}

}  // namespace

extern "C" {

int sv_disparity_to_points(const double* disp, int B, int W, int H, const double* cam,
                           double min_disp, double max_depth, double* pts, uint8_t* valid,
                           void* stream) {
    SV_REQUIRE(disp && cam && pts && valid && B > 0 && W > 0 && H > 0,
               "sv_disparity_to_points: bad arguments");
    const Cam c{cam[0], cam[1], cam[2], cam[3], cam[4]};
    const long long total = (long long)B * W * H;
    hipLaunchKernelGGL(k_points, dim3(blocks(total)), dim3(256), 0, (hipStream_t)stream, disp, total,
                       W, H, c, min_disp, max_depth, pts, valid);
    SV_CHECK_LAUNCH();
    return SV_OK;
}

int sv_kth_f64_workspace_bytes(int B, size_t* bytes) {
    SV_REQUIRE(bytes && B > 0, "sv_kth_f64_workspace_bytes: bad arguments");
    *bytes = (size_t)B * (kSelBins * sizeof(unsigned) + 64);
    return SV_OK;
}

int sv_kth_f64_batch(const double* x, long long mstride, long long stride, int B, long long n,
                     long long k, const int* ns, int nstride, double empty_value, double* out,
                     void* ws, size_t ws_bytes, void* stream) {
    SV_REQUIRE(x && out && ws && stride > 0 && B > 0 && B <= 65535 && n > 0,
               "sv_kth_f64_batch: bad arguments");
    SV_REQUIRE(ns || (k >= 0 && k < n), "sv_kth_f64_batch: k must be in [0, n)");
    static_assert(sizeof(Sel64) <= 64, "selection state");
    SV_REQUIRE(ws_bytes >= (size_t)B * (kSelBins * sizeof(unsigned) + 64),
               "sv_kth_f64_batch: workspace too small");
    unsigned* hist = (unsigned*)ws;
    Sel64* st = (Sel64*)((char*)ws + (size_t)B * kSelBins * sizeof(unsigned));
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_sel64_init, dim3(kSelBins / 256, B), dim3(256), 0, s, hist, st, n, k, ns,
                       nstride, empty_value, out);
    SV_CHECK_LAUNCH();
    // about 32 keys per thread: each block pays a fixed cost (histogram
    // clear and flush, the completion counter) that dominates small blocks
    const long long nb = (n + 8191) / 8192;
    const unsigned g = (unsigned)(nb < 256 ? nb : 256);
    for (int pass = 0; pass < kSelPasses; ++pass) {
        hipLaunchKernelGGL(k_sel64_pass, dim3(g, B), dim3(256), 0, s, x, mstride, stride, pass, st,
                           hist, out);
        SV_CHECK_LAUNCH();
        hipLaunchKernelGGL(k_sel64_pick, dim3(B), dim3(256), 0, s, pass, st, hist, out);
        SV_CHECK_LAUNCH();
    }
    return SV_OK;
}

int sv_kth_f64(const double* x, long long stride, long long n, long long k, double* out, void* ws,
               size_t ws_bytes, void* stream) {
    return sv_kth_f64_batch(x, 0, stride, 1, n, k, nullptr, 0, 0.0, out, ws, ws_bytes, stream);
}

int sv_min_filter_arg(const double* v, int B, int W, int H, int r, int rv, double* out, int* idx,
                      double* tmp, int* tmp_idx, void* stream) {
    SV_REQUIRE(v && out && idx && B > 0 && W > 0 && H > 0 && r >= 0 && rv >= 0,
               "sv_min_filter_arg: bad arguments");
    SV_REQUIRE(rv == 0 || (tmp && tmp_idx), "sv_min_filter_arg: rv > 0 needs tmp buffers");
    SV_REQUIRE((long long)W * H < 0x7FFFFFFFLL, "sv_min_filter_arg: image too large");
    hipStream_t s = (hipStream_t)stream;
    const long long total = (long long)B * W * H;
    if (rv == 0) {
        launch_minrow(s, v, B, W, H, r, out, idx);
        SV_CHECK_LAUNCH();
        return SV_OK;
    }
    launch_minrow(s, v, B, W, H, r, tmp, tmp_idx);
    SV_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_mincol, dim3(blocks(total)), dim3(256), 0, s, tmp, tmp_idx, B, W, H, rv, out,
                       idx);
    SV_CHECK_LAUNCH();
    return SV_OK;
}

int sv_smooth_background(const double* m, const double* disp, int B, int W, int H, int r,
                         double* tmp, double* out, void* stream) {
    SV_REQUIRE(m && disp && tmp && out && B > 0 && W > 0 && H > 0 && r >= 0,
               "sv_smooth_background: bad arguments");
    hipStream_t s = (hipStream_t)stream;
    const long long total = (long long)B * W * H;
    launch_smooth_row(s, m, B, W, H, r, tmp);
    SV_CHECK_LAUNCH();
    launch_smooth_col(s, tmp, disp, B, W, H, r, out);
    SV_CHECK_LAUNCH();
    return SV_OK;
}

int sv_background_region(const double* disp, int B, int W, int H, int radius, double jump_ratio,
                         double min_jump, uint8_t* mask, double* m, int* src, double* tmp,
                         uint8_t* tmp_mask, void* stream) {
    SV_REQUIRE(disp && mask && m && src && tmp && tmp_mask && B > 0 && W > 0 && H > 0 &&
                   radius >= 0,
               "sv_background_region: bad arguments");
    SV_REQUIRE((long long)W * H < 0x7FFFFFFFLL, "sv_background_region: image too large");
    hipStream_t s = (hipStream_t)stream;
    const long long total = (long long)B * W * H;
    // min filter (rows only) -> m (as scratch), src; row means -> tmp;
    // column means, clamped to the surface -> m
    launch_minrow(s, disp, B, W, H, radius, m, src);
    SV_CHECK_LAUNCH();
    const int rs = radius >= 2 ? radius / 2 : 1;
    launch_smooth_row(s, m, B, W, H, rs, tmp);
    SV_CHECK_LAUNCH();
    launch_smooth_col(s, tmp, disp, B, W, H, rs, m);
    SV_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_bg_thresh, dim3(blocks(total)), dim3(256), 0, s, disp, m, total,
                       jump_ratio, min_jump, mask);
    SV_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_bg_grow, dim3(blocks(total)), dim3(256), 0, s, mask, B, W, H, tmp_mask);
    SV_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_bg_grow, dim3(blocks(total)), dim3(256), 0, s, tmp_mask, B, W, H, mask);
    SV_CHECK_LAUNCH();
    return SV_OK;
}

int sv_grid_mesh_workspace_bytes(int B, int W, int H, int step, size_t* bytes) {
    SV_REQUIRE(bytes && B > 0 && W > 0 && H > 0 && step > 0,
               "sv_grid_mesh_workspace_bytes: bad arguments");
    const long long nx = (W + step - 1) / step + ((W - 1) % step ? 1 : 0);
    const long long ny = (H + step - 1) / step + ((H - 1) % step ? 1 : 0);
    *bytes = (size_t)B * (nx * ny + 2 * ny + 8) * sizeof(int);
    return SV_OK;
}

// One build of grid_mesh for B maps (main: val = disparity, region/src
// null; fill: val = background disparity, region = mask, src = source
// pixels).  Map b's vertices go to [b * vcap, ...) of pos/uv/flag/dv and its
// triangles to [b * tcap, ...) of idx, vertex ids counted within the map.
// totals is [B][4] on the device: (main vertices, main triangles, fill
// vertices, fill triangles); the main build writes the first two, the fill
// build appends after them and writes the last two.
int sv_grid_mesh_build(const double* val, const uint8_t* region, const int* src, int B, int W,
                       int H, int step, const double* cam, const double* min_disp, double jump,
                       double jump_rel, double* pos, double* uv, uint8_t* flag, double* dv, int* idx,
                       long long vcap, long long tcap, int* totals, void* ws, size_t ws_bytes,
                       void* stream) {
    SV_REQUIRE(val && cam && min_disp && pos && uv && flag && dv && idx && totals && ws && B > 0 &&
                   B <= 65535 && W > 0 && H > 0 && step > 0,
               "sv_grid_mesh_build: bad arguments");
    SV_REQUIRE((region == nullptr) == (src == nullptr), "sv_grid_mesh_build: region and src go together");
    const int nx0 = (W + step - 1) / step, ny0 = (H + step - 1) / step;
    const int nx = nx0 + ((W - 1) % step ? 1 : 0), ny = ny0 + ((H - 1) % step ? 1 : 0);
    const bool fill = region != nullptr;
    SV_REQUIRE(vcap >= (long long)nx * ny * (fill ? 2 : 1) &&
                   tcap >= 2LL * (nx - 1) * (ny - 1) * (fill ? 2 : 1),
               "sv_grid_mesh_build: output capacity below the worst case");
    size_t need = 0;
    sv_grid_mesh_workspace_bytes(B, W, H, step, &need);
    SV_REQUIRE(ws_bytes >= need, "sv_grid_mesh_build: workspace too small");
    SV_REQUIRE(ny <= 1 << 20, "sv_grid_mesh_build: too many grid rows");
    int* vid = (int*)ws;
    int* cnt = vid + (size_t)B * nx * ny;
    int* offs = cnt + (size_t)B * ny;
    Grid g{val, region, src, min_disp, W, H, step, nx, ny, nx0, ny0, fill ? 1 : 0,
           Cam{cam[0], cam[1], cam[2], cam[3], cam[4]}};
    Out o{pos, uv, flag, dv, idx, vcap, tcap};
    const int* vbase = fill ? totals : nullptr;
    const int* tbase = fill ? totals + 1 : nullptr;
    int* vtot = totals + (fill ? 2 : 0);
    int* ttot = totals + (fill ? 3 : 1);
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_grid_vert<false>, dim3(ny, B), dim3(64), 0, s, g, o, cnt, offs, vid);
    SV_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_scan_rows, dim3(B), dim3(1024), 0, s, cnt, ny, vbase, offs, vtot);
    SV_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_grid_vert<true>, dim3(ny, B), dim3(64), 0, s, g, o, cnt, offs, vid);
    SV_CHECK_LAUNCH();
    if (ny < 2 || nx < 2) {  // no cells: zero triangles
        hipLaunchKernelGGL(k_scan_rows, dim3(B), dim3(1024), 0, s, cnt, 0, tbase, offs, ttot);
        SV_CHECK_LAUNCH();
        return SV_OK;
    }
    hipLaunchKernelGGL(k_grid_tri<false>, dim3(ny - 1, B), dim3(64), 0, s, vid, nx, ny, o, jump,
                       jump_rel, cnt, offs);
    SV_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_scan_rows, dim3(B), dim3(1024), 0, s, cnt, ny - 1, tbase, offs, ttot);
    SV_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_grid_tri<true>, dim3(ny - 1, B), dim3(64), 0, s, vid, nx, ny, o, jump,
                       jump_rel, cnt, offs);
    SV_CHECK_LAUNCH();
    return SV_OK;
}

}  // extern "C"
