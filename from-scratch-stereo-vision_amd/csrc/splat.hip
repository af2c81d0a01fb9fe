// Gaussian cloud of the downstream 3D output on the GPU: one anisotropic
// Gaussian per mesh triangle (gaussians_from_mesh, stereonet.py:2546-2659),
// the EWA splat renderer's projection, front-to-back compositing and its
// backward sweep (SplatRenderer, :2706-2962), the photometric loss of
// refine_gaussians (:3010-3039) and CloudAdam (:2965-3007).
//
// All float64, evaluated in the reference's operation order with its own
// fsqrt / fexp / flog (refmath.h; this file is built without contraction):
// - projection: one thread per Gaussian; the kept ones are ranked by depth
//   with ties in index order (the reference's stable sort) by counting, so
//   every pixel sees the same sequence;
// - compositing: one thread per pixel, 16x16-pixel blocks; a block walks the
//   depth-ordered list in chunks, keeps (ballot-compacted, in order) the
//   Gaussians whose footprint box meets its tile, and each pixel composites
//   them in order exactly as the reference's inner loop does;
// - backward: the same walk back to front per pixel; each pixel writes its
//   contribution to every Gaussian it touches into that Gaussian's slot of
//   a per-footprint buffer (zero where the pixel adds nothing), and one
//   thread per Gaussian then sums its footprint in the reference's
//   row-major pixel order, so the gradients agree bit for bit (no atomics,
//   no reassociation; an added +0.0 changes no sum, since a sum that starts
//   at +0.0 is never -0.0).
#include <algorithm>
#include <cstdint>

#include "common.h"
#include "refmath.h"

namespace {

constexpr int kNP = 14;  // cloud parameters: mx my mz sxx sxy sxz syy syz szz cr cg cb raw_op raw_sc
enum { MX, MY, MZ, SXX, SXY, SXZ, SYY, SYZ, SZZ, CR, CG, CB, ROP, RSC };
// per-item doubles, then ints
enum { IZ, IU, IV, I00, I01, I11, IA00, IA01, IA11, IS, IOP, ICR, ICG, ICB, NID };
enum { JG, JX0, JX1, JY0, JY1, NJI };
constexpr int kTile = 16;
constexpr double kAlphaMax = 0.99;

// ---------------------------------------------------------- cloud from mesh
// One block of 1024 threads walks the triangles in chunks; a kept triangle's
// Gaussian lands at the next free slot (block scan), so the cloud keeps the
// triangle order.
__global__ void __launch_bounds__(1024) k_gauss_from_mesh(
    const int* __restrict__ xy, int nv, const int* __restrict__ tris, int nt,
    const double* __restrict__ pts, const uint8_t* __restrict__ pvalid,
    const double* __restrict__ rgb, int W, double thickness, double raw_op, double raw_sc,
    double* __restrict__ cloud, int cap, int* __restrict__ count) {
    __shared__ int wsum[16];
    __shared__ int base;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid == 0) base = 0;
    __syncthreads();
    for (int t0 = 0; t0 < nt; t0 += 1024) {
        const int t = t0 + tid;
        bool keep = false;
        double g[kNP];
        if (t < nt) {
            const double* P[3];
            int pix[3];
            keep = true;
            for (int q = 0; q < 3; ++q) {
                const int id = tris[3 * t + q];
                if (id < 0 || id >= nv) {
                    keep = false;
                    break;
                }
                pix[q] = xy[2 * id + 1] * W + xy[2 * id];
                if (!pvalid[pix[q]]) {
                    keep = false;
                    break;
                }
                P[q] = pts + 3LL * pix[q];
            }
            if (keep) {
                const double zmin = refm::fmin_(P[0][2], refm::fmin_(P[1][2], P[2][2]));
                const double zmax = refm::fmax_(P[0][2], refm::fmax_(P[1][2], P[2][2]));
                if (zmin <= 0.0 || zmax > zmin * 1.6) keep = false;
            }
            if (keep) {
                const double cxm = (P[0][0] + P[1][0] + P[2][0]) / 3.0;
                const double cym = (P[0][1] + P[1][1] + P[2][1]) / 3.0;
                const double czm = (P[0][2] + P[1][2] + P[2][2]) / 3.0;
                const double ux = P[1][0] - P[0][0], uy = P[1][1] - P[0][1], uz = P[1][2] - P[0][2];
                const double vx = P[2][0] - P[0][0], vy = P[2][1] - P[0][1], vz = P[2][2] - P[0][2];
                const double A = 1.0 / 18.0, B = -1.0 / 36.0;
                double sxx = A * (ux * ux + vx * vx) + B * 2.0 * ux * vx;
                double syy = A * (uy * uy + vy * vy) + B * 2.0 * uy * vy;
                double szz = A * (uz * uz + vz * vz) + B * 2.0 * uz * vz;
                double sxy = A * (ux * uy + vx * vy) + B * (ux * vy + vx * uy);
                double sxz = A * (ux * uz + vx * vz) + B * (ux * vz + vx * uz);
                double syz = A * (uy * uz + vy * vz) + B * (uy * vz + vy * uz);
                double nx = uy * vz - uz * vy, ny = uz * vx - ux * vz, nz = ux * vy - uy * vx;
                const double nlen = refm::fsqrt(nx * nx + ny * ny + nz * nz);
                if (nlen < 1.0e-12) {
                    keep = false;
                } else {
                    nx /= nlen;
                    ny /= nlen;
                    nz /= nlen;
                    const double tscale = thickness * refm::fsqrt(refm::fmax_(sxx + syy + szz, 1.0e-12));
                    const double t2 = tscale * tscale;
                    sxx += t2 * nx * nx;
                    syy += t2 * ny * ny;
                    szz += t2 * nz * nz;
                    sxy += t2 * nx * ny;
                    sxz += t2 * nx * nz;
                    syz += t2 * ny * nz;
                    double r = 0.0, gg = 0.0, b = 0.0;
                    for (int q = 0; q < 3; ++q) {
                        const long long o = 3LL * pix[q];
                        r += rgb[o];
                        gg += rgb[o + 1];
                        b += rgb[o + 2];
                    }
                    r /= 3.0;
                    gg /= 3.0;
                    b /= 3.0;
                    g[MX] = cxm;
                    g[MY] = cym;
                    g[MZ] = czm;
                    g[SXX] = sxx;
                    g[SXY] = sxy;
                    g[SXZ] = sxz;
                    g[SYY] = syy;
                    g[SYZ] = syz;
                    g[SZZ] = szz;
                    g[CR] = r;
                    g[CG] = gg;
                    g[CB] = b;
                    g[ROP] = raw_op;
                    g[RSC] = raw_sc;
                }
            }
        }
        // ordered slot: block-wide exclusive scan of the keep flags
        const unsigned long long bal = __ballot(keep);
        const int before = __popcll(bal & ((1ULL << lane) - 1ULL));
        if (lane == 0) wsum[wave] = __popcll(bal);
        __syncthreads();
        int off = base;
        for (int q = 0; q < wave; ++q) off += wsum[q];
        if (keep) {
            const int slot = off + before;
            if (slot < cap)
                for (int p = 0; p < kNP; ++p) cloud[(long long)p * cap + slot] = g[p];
        }
        __syncthreads();
        if (tid == 0)
            for (int q = 0; q < 16; ++q) base += wsum[q];
        __syncthreads();
    }
    if (tid == 0) *count = base;
}

// ---------------------------------------------------------------- projection
struct Proj {
    double R[9], t[3];
    double fx, cx, cy, blur, cutoff;
    int W, H;
};

// SplatRenderer._prepare for one Gaussian: the item fields, or false when
// the reference drops it (behind the camera, degenerate, too large, off
// screen).
__device__ bool project(const double* __restrict__ cloud, int cap, int i, const Proj& pj,
                        double* f, int* b) {
    const double X = cloud[MX * cap + i], Y = cloud[MY * cap + i], Z = cloud[MZ * cap + i];
    const double* R = pj.R;
    const double xc = R[0] * X + R[1] * Y + R[2] * Z + pj.t[0];
    const double yc = R[3] * X + R[4] * Y + R[5] * Z + pj.t[1];
    const double zc = R[6] * X + R[7] * Y + R[8] * Z + pj.t[2];
    if (zc <= 1.0e-4) return false;
    const double fx = pj.fx;
    const double u = fx * xc / zc + pj.cx;
    const double v = fx * yc / zc + pj.cy;
    const double f_z = fx / zc;
    const double f_z2 = -fx / (zc * zc);
    const double jx0 = f_z, jx1 = 0.0, jx2 = f_z2 * xc;
    const double jy0 = 0.0, jy1 = f_z, jy2 = f_z2 * yc;
    const double m0[3] = {jx0 * R[0] + jx1 * R[3] + jx2 * R[6], jx0 * R[1] + jx1 * R[4] + jx2 * R[7],
                          jx0 * R[2] + jx1 * R[5] + jx2 * R[8]};
    const double m1[3] = {jy0 * R[0] + jy1 * R[3] + jy2 * R[6], jy0 * R[1] + jy1 * R[4] + jy2 * R[7],
                          jy0 * R[2] + jy1 * R[5] + jy2 * R[8]};
    const double sxx = cloud[SXX * cap + i], sxy = cloud[SXY * cap + i], sxz = cloud[SXZ * cap + i];
    const double syy = cloud[SYY * cap + i], syz = cloud[SYZ * cap + i], szz = cloud[SZZ * cap + i];
    const double sm0[3] = {sxx * m0[0] + sxy * m0[1] + sxz * m0[2], sxy * m0[0] + syy * m0[1] + syz * m0[2],
                           sxz * m0[0] + syz * m0[1] + szz * m0[2]};
    const double sm1[3] = {sxx * m1[0] + sxy * m1[1] + sxz * m1[2], sxy * m1[0] + syy * m1[1] + syz * m1[2],
                           sxz * m1[0] + syz * m1[1] + szz * m1[2]};
    const double a00 = m0[0] * sm0[0] + m0[1] * sm0[1] + m0[2] * sm0[2];
    const double a01 = m0[0] * sm1[0] + m0[1] * sm1[1] + m0[2] * sm1[2];
    const double a11 = m1[0] * sm1[0] + m1[1] * sm1[1] + m1[2] * sm1[2];
    const double s = refm::fexp(cloud[RSC * cap + i]);
    const double c00 = s * a00 + pj.blur;
    const double c01 = s * a01;
    const double c11 = s * a11 + pj.blur;
    const double det = c00 * c11 - c01 * c01;
    if (det < 1.0e-12) return false;
    const double inv = 1.0 / det;
    const double tr = 0.5 * (c00 + c11);
    const double disc = refm::fsqrt(refm::fmax_(tr * tr - det, 0.0));
    const double lmax = tr + disc;
    double rad = pj.cutoff * refm::fsqrt(refm::fmax_(lmax, 1.0e-12));
    if (rad < 0.5) rad = 0.5;
    if (rad > 0.35 * (double)(pj.W + pj.H)) return false;
    // int() truncates toward zero; the operands stay far inside 2^62
    long long x0 = (long long)(u - rad), x1 = (long long)(u + rad) + 1;
    long long y0 = (long long)(v - rad), y1 = (long long)(v + rad) + 1;
    if (x1 < 0 || y1 < 0 || x0 >= pj.W || y0 >= pj.H) return false;
    if (x0 < 0) x0 = 0;
    if (y0 < 0) y0 = 0;
    if (x1 > pj.W) x1 = pj.W;
    if (y1 > pj.H) y1 = pj.H;
    f[IZ] = zc;
    f[IU] = u;
    f[IV] = v;
    f[I00] = c11 * inv;
    f[I01] = -c01 * inv;
    f[I11] = c00 * inv;
    f[IA00] = a00;
    f[IA01] = a01;
    f[IA11] = a11;
    f[IS] = s;
    f[IOP] = refm::sigmoid(cloud[ROP * cap + i]);
    f[ICR] = cloud[CR * cap + i];
    f[ICG] = cloud[CG * cap + i];
    f[ICB] = cloud[CB * cap + i];
    b[JX0] = (int)x0;
    b[JX1] = (int)x1;
    b[JY0] = (int)y0;
    b[JY1] = (int)y1;
    return true;
}

// items in Gaussian order (kept flag in keep[i])
__global__ void k_splat_project(const double* __restrict__ cloud, int n, int cap, Proj pj,
                                double* __restrict__ raw, int* __restrict__ rawi,
                                uint8_t* __restrict__ keep) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double f[NID];
    int b[NJI];
    const bool k = project(cloud, cap, i, pj, f, b);
    keep[i] = k ? 1 : 0;
    if (!k) return;
    for (int q = 0; q < NID; ++q) raw[(long long)q * n + i] = f[q];
    rawi[(long long)JG * n + i] = i;
    for (int q = JX0; q < NJI; ++q) rawi[(long long)q * n + i] = b[q];
}

// rank of each kept item: kept items in front (smaller depth), or equally
// deep and earlier; the items land in depth order, ties in Gaussian order.
// One wavefront per item, its lanes counting over strided candidates.
__global__ void __launch_bounds__(256) k_splat_rank(const double* __restrict__ raw,
                                                    const int* __restrict__ rawi,
                                                    const uint8_t* __restrict__ keep, int n,
                                                    double* __restrict__ items,
                                                    int* __restrict__ itemi) {
    const int lane = threadIdx.x & 63;
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= n || !keep[i]) return;
    const double zi = raw[(long long)IZ * n + i];
    int cnt = 0;
    for (int j = lane; j < n; j += 64) {
        const double zj = raw[(long long)IZ * n + j];
        cnt += (keep[j] && (zj < zi || (zj == zi && j < i))) ? 1 : 0;
    }
    for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o, 64);
    if (lane < NID) items[(long long)lane * n + cnt] = raw[(long long)lane * n + i];
    else if (lane - NID < NJI)
        itemi[(long long)(lane - NID) * n + cnt] = rawi[(long long)(lane - NID) * n + i];
}

// footprint offsets: area_off[r] = sum of the box areas of items 0 .. r-1
// (one block, items in chunks)
__global__ void __launch_bounds__(1024) k_splat_areas(const int* __restrict__ itemi, int n,
                                                      const int* __restrict__ nkept,
                                                      long long* __restrict__ area_off,
                                                      long long* __restrict__ info) {
    __shared__ long long wsum[16];
    __shared__ long long base;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m = *nkept;
    if (tid == 0) base = 0;
    __syncthreads();
    for (int r0 = 0; r0 < m; r0 += 1024) {
        const int r = r0 + tid;
        long long a = 0;
        if (r < m)
            a = (long long)(itemi[(long long)JX1 * n + r] - itemi[(long long)JX0 * n + r]) *
                (itemi[(long long)JY1 * n + r] - itemi[(long long)JY0 * n + r]);
        long long inc = a;
        for (int o = 1; o < 64; o <<= 1) {
            const long long t = __shfl_up(inc, o, 64);
            if (lane >= o) inc += t;
        }
        if (lane == 63) wsum[wave] = inc;
        __syncthreads();
        long long off = base;
        for (int q = 0; q < wave; ++q) off += wsum[q];
        if (r < m) area_off[r] = off + inc - a;
        __syncthreads();
        if (tid == 0)
            for (int q = 0; q < 16; ++q) base += wsum[q];
        __syncthreads();
    }
    if (tid == 0) {
        area_off[m] = base;
        if (info) {
            info[0] = m;
            info[1] = base;
        }
    }
}

__global__ void k_count_kept(const uint8_t* __restrict__ keep, int n, int* __restrict__ out) {
    __shared__ int s[256];
    int c = 0;
    for (int i = threadIdx.x; i < n; i += 256) c += keep[i];
    s[threadIdx.x] = c;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = s[0];
}

// the items of chunk [c0, c0 + 256) whose box meets this block's tile, in
// order, into lst; returns their number
__device__ int tile_chunk(const int* __restrict__ itemi, int n, int m, int c0, int tx0, int ty0,
                          int* lst, int* wc) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int r = c0 + tid;
    bool hit = false;
    if (r < m) {
        const int x0 = itemi[(long long)JX0 * n + r], x1 = itemi[(long long)JX1 * n + r];
        const int y0 = itemi[(long long)JY0 * n + r], y1 = itemi[(long long)JY1 * n + r];
        hit = x0 < tx0 + kTile && x1 > tx0 && y0 < ty0 + kTile && y1 > ty0;
    }
    const unsigned long long bal = __ballot(hit);
    if (lane == 0) wc[wave] = __popcll(bal);
    __syncthreads();
    int off = 0;
    for (int q = 0; q < wave; ++q) off += wc[q];
    if (hit) lst[off + __popcll(bal & ((1ULL << lane) - 1ULL))] = r;
    const int tot = wc[0] + wc[1] + wc[2] + wc[3];
    __syncthreads();
    return tot;
}

__global__ void __launch_bounds__(256) k_splat_render(const double* __restrict__ items,
                                                      const int* __restrict__ itemi, int n,
                                                      const int* __restrict__ nkept, int W, int H,
                                                      double br, double bgc, double bb,
                                                      double* __restrict__ C,
                                                      double* __restrict__ Tout) {
    __shared__ int lst[256];
    __shared__ int wc[4];
    const int m = *nkept;
    const int tx0 = blockIdx.x * kTile, ty0 = blockIdx.y * kTile;
    const int x = tx0 + (threadIdx.x & (kTile - 1)), y = ty0 + threadIdx.x / kTile;
    const bool in = x < W && y < H;
    double T = 1.0, c0 = 0.0, c1 = 0.0, c2 = 0.0;
    for (int r0 = 0; r0 < m; r0 += 256) {
        const int cnt = tile_chunk(itemi, n, m, r0, tx0, ty0, lst, wc);
        for (int q = 0; q < cnt && in; ++q) {
            const int r = lst[q];
            if (x < itemi[(long long)JX0 * n + r] || x >= itemi[(long long)JX1 * n + r] ||
                y < itemi[(long long)JY0 * n + r] || y >= itemi[(long long)JY1 * n + r])
                continue;
            const double Tp = T;
            if (Tp < 0.004) continue;
            const double dy = (double)y - items[(long long)IV * n + r];
            const double dx = (double)x - items[(long long)IU * n + r];
            const double i00 = items[(long long)I00 * n + r], i01 = items[(long long)I01 * n + r];
            const double i11 = items[(long long)I11 * n + r];
            const double qv = i00 * dx * dx + 2.0 * i01 * dx * dy + i11 * dy * dy;
            if (qv > 9.0) continue;
            double a = items[(long long)IOP * n + r] * refm::fexp(-0.5 * qv);
            if (a < 1.0e-4) continue;
            if (a > kAlphaMax) a = kAlphaMax;
            const double w = Tp * a;
            c0 += w * items[(long long)ICR * n + r];
            c1 += w * items[(long long)ICG * n + r];
            c2 += w * items[(long long)ICB * n + r];
            T = Tp * (1.0 - a);
        }
    }
    if (!in) return;
    const long long p = (long long)y * W + x;
    C[3 * p] = c0 + T * br;
    C[3 * p + 1] = c1 + T * bgc;
    C[3 * p + 2] = c2 + T * bb;
    Tout[p] = T;
}

// backward sweep: per pixel, back to front; each touched item's five
// contributions go to its footprint slot (the opacity and scale ones 0.0
// where alpha was clipped, as the reference adds nothing there)
__global__ void __launch_bounds__(256) k_splat_backward(
    const double* __restrict__ items, const int* __restrict__ itemi, int n,
    const int* __restrict__ nkept, int W, int H, const double* __restrict__ gC,
    const double* __restrict__ Tfinal, const long long* __restrict__ area_off,
    double* __restrict__ contrib) {
    __shared__ int lst[256];
    __shared__ int wc[4];
    const int m = *nkept;
    const int tx0 = blockIdx.x * kTile, ty0 = blockIdx.y * kTile;
    const int x = tx0 + (threadIdx.x & (kTile - 1)), y = ty0 + threadIdx.x / kTile;
    const bool in = x < W && y < H;
    const long long p = (long long)y * W + x;
    double T = in ? Tfinal[p] : 1.0;
    double S0 = 0.0, S1 = 0.0, S2 = 0.0;
    const double gr = in ? gC[3 * p] : 0.0, gg = in ? gC[3 * p + 1] : 0.0, gb = in ? gC[3 * p + 2] : 0.0;
    const long long total = area_off[m];
    const int nch = (m + 255) / 256;
    for (int ch = nch - 1; ch >= 0; --ch) {
        const int cnt = tile_chunk(itemi, n, m, ch * 256, tx0, ty0, lst, wc);
        for (int q = cnt - 1; q >= 0 && in; --q) {
            const int r = lst[q];
            const int x0 = itemi[(long long)JX0 * n + r], x1 = itemi[(long long)JX1 * n + r];
            const int y0 = itemi[(long long)JY0 * n + r], y1 = itemi[(long long)JY1 * n + r];
            if (x < x0 || x >= x1 || y < y0 || y >= y1) continue;
            const double dy = (double)y - items[(long long)IV * n + r];
            const double dx = (double)x - items[(long long)IU * n + r];
            const double i00 = items[(long long)I00 * n + r], i01 = items[(long long)I01 * n + r];
            const double i11 = items[(long long)I11 * n + r];
            const double qv = i00 * dx * dx + 2.0 * i01 * dx * dy + i11 * dy * dy;
            if (qv > 9.0) continue;
            const double ex = refm::fexp(-0.5 * qv);
            const double op = items[(long long)IOP * n + r];
            double a = op * ex;
            if (a < 1.0e-4) continue;
            bool clipped = false;
            if (a > kAlphaMax) {
                a = kAlphaMax;
                clipped = true;
            }
            const double Ti = T / (1.0 - a);
            T = Ti;
            const double cr = items[(long long)ICR * n + r], cg = items[(long long)ICG * n + r];
            const double cb = items[(long long)ICB * n + r];
            const double w = Ti * a;
            const double inv1a = 1.0 / (1.0 - a);
            const double dC_da = gr * (Ti * cr - S0 * inv1a) + gg * (Ti * cg - S1 * inv1a) +
                                 gb * (Ti * cb - S2 * inv1a);
            double dop = 0.0, dsc = 0.0;
            if (!clipped) {
                const double dop_draw = op * (1.0 - op);
                dop = dC_da * ex * dop_draw;
                const double px = i00 * dx + i01 * dy;
                const double py = i01 * dx + i11 * dy;
                const double a00 = items[(long long)IA00 * n + r], a01 = items[(long long)IA01 * n + r];
                const double a11 = items[(long long)IA11 * n + r];
                const double Ap_x = a00 * px + a01 * py;
                const double Ap_y = a01 * px + a11 * py;
                const double dq_ds = -(px * Ap_x + py * Ap_y);
                dsc = dC_da * a * (-0.5 * dq_ds) * items[(long long)IS * n + r];
            }
            const long long slot = area_off[r] + (long long)(y - y0) * (x1 - x0) + (x - x0);
            contrib[slot] = gr * w;
            contrib[total + slot] = gg * w;
            contrib[2 * total + slot] = gb * w;
            contrib[3 * total + slot] = dop;
            contrib[4 * total + slot] = dsc;
            S0 += w * cr;
            S1 += w * cg;
            S2 += w * cb;
        }
    }
}

// one block of NW wavefronts per item.  Its lanes load CH = NW * 64 * EPL
// consecutive footprint entries of each of the five contributions (the next
// chunk in flight while this one is summed), keep the entries that hold a
// nonzero value, packed in row-major pixel order into LDS, and lanes 0-4 of
// the first wavefront each sum one contribution over them, then add it to
// the Gaussian's gradient.  Leaving out an entry whose five values are all
// zero, and padding with +0.0, changes no sum that starts at +0.0, so the
// result equals the sequential sum over the whole footprint; most of a box
// lies outside the splat's support, and the sums are chains of dependent
// adds, so the chain is shortened to the entries that count.
template <int NW, int EPL>
__global__ void __launch_bounds__(NW * 64) k_splat_reduce(const int* __restrict__ itemi, int n,
                                                          const int* __restrict__ nkept,
                                                          const long long* __restrict__ area_off,
                                                          const double* __restrict__ contrib,
                                                          double* __restrict__ grads, int cap) {
    constexpr int NT = NW * 64, CH = NT * EPL, NG = NW * EPL;
    constexpr int ROW = CH / 2 + 1;  // double2 per contribution row (the +1 spreads the banks)
    __shared__ double2 buf[5 * ROW];
    __shared__ int gcnt[NG];
    const int r = blockIdx.x;
    const int m = *nkept;
    if (r >= m) return;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const long long total = area_off[m];
    const long long b = area_off[r], e = area_off[r + 1];
    double nx[EPL][5];
#pragma unroll
    for (int k = 0; k < EPL; ++k) {
        const long long sk = b + k * NT + t;
#pragma unroll
        for (int q = 0; q < 5; ++q) nx[k][q] = sk < e ? contrib[q * total + sk] : 0.0;
    }
    double* bufd = reinterpret_cast<double*>(buf);
    double acc = 0.0;
    for (long long s0 = b; s0 < e; s0 += CH) {
        // entry k * NT + t of the chunk is group k * NW + w, lane `lane`
        unsigned long long mk[EPL];
#pragma unroll
        for (int k = 0; k < EPL; ++k) {
            bool nz = false;
#pragma unroll
            for (int q = 0; q < 5; ++q) nz |= nx[k][q] != 0.0;
            mk[k] = __ballot(nz);
            if (lane == 0) gcnt[k * NW + w] = __popcll(mk[k]);
        }
        __syncthreads();
        int cnt = 0, base[EPL];
#pragma unroll
        for (int k = 0; k < EPL; ++k) base[k] = 0;
        for (int g = 0; g < NG; ++g) {
            const int c = gcnt[g];
#pragma unroll
            for (int k = 0; k < EPL; ++k) base[k] += g < k * NW + w ? c : 0;
            cnt += c;
        }
#pragma unroll
        for (int k = 0; k < EPL; ++k) {
            const int below = __builtin_amdgcn_mbcnt_hi(
                (unsigned)(mk[k] >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)mk[k], 0u));
            if ((mk[k] >> lane) & 1ull) {
                const int pos = base[k] + below;
#pragma unroll
                for (int q = 0; q < 5; ++q) bufd[q * 2 * ROW + pos] = nx[k][q];
            }
        }
        const int nb = (cnt + 31) >> 5;  // 32-entry batches, the tail padded with +0.0
        if (t < nb * 32 - cnt) {
#pragma unroll
            for (int q = 0; q < 5; ++q) bufd[q * 2 * ROW + cnt + t] = 0.0;
        }
        __syncthreads();
        if (s0 + CH < e) {
#pragma unroll
            for (int k = 0; k < EPL; ++k) {
                const long long sk = s0 + CH + k * NT + t;
#pragma unroll
                for (int q = 0; q < 5; ++q) nx[k][q] = sk < e ? contrib[q * total + sk] : 0.0;
            }
        }
        if (t < 5 && nb > 0) {
            const double2* v = buf + t * ROW;
            double2 wa[16], wb[16];
#pragma unroll
            for (int j = 0; j < 16; ++j) wa[j] = v[j];
            int k = 0;
            for (;;) {
                if (k + 1 < nb) {
#pragma unroll
                    for (int j = 0; j < 16; ++j) wb[j] = v[(k + 1) * 16 + j];
                }
#pragma unroll
                for (int j = 0; j < 16; ++j) {
                    acc += wa[j].x;
                    acc += wa[j].y;
                }
                if (++k >= nb) break;
                if (k + 1 < nb) {
#pragma unroll
                    for (int j = 0; j < 16; ++j) wa[j] = v[(k + 1) * 16 + j];
                }
#pragma unroll
                for (int j = 0; j < 16; ++j) {
                    acc += wb[j].x;
                    acc += wb[j].y;
                }
                if (++k >= nb) break;
            }
        }
    }
    if (t < 5) grads[(long long)t * cap + itemi[(long long)JG * n + r]] += acc;
}

__global__ void k_zero_f64(double* __restrict__ a, long long n) {
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += 256LL * gridDim.x) a[i] = 0.0;
}

// refine_gaussians' photometric loss: per-block partial sums of e^2 and
// the gradient 2 e / n
__global__ void __launch_bounds__(256) k_mse(const double* __restrict__ C,
                                             const double* __restrict__ target, long long n,
                                             double inv, double* __restrict__ gC,
                                             double* __restrict__ part) {
    __shared__ double s[256];
    double acc = 0.0;
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n; i += 256LL * gridDim.x) {
        const double e = C[i] - target[i];
        acc += e * e;
        gC[i] = 2.0 * e * inv;
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

__global__ void __launch_bounds__(256) k_mse_final(const double* __restrict__ part, int nb,
                                                   double inv, double* __restrict__ loss) {
    __shared__ double s[256];
    s[threadIdx.x] = threadIdx.x < nb ? part[threadIdx.x] : 0.0;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *loss = s[0] * inv;
}

// CloudAdam.step: groups cr, cg, cb, raw_op, raw_sc with their learning
// rates; gradients clipped to +-10; colours clamped to [0, 1] afterwards
__global__ void k_cloud_adam(double* __restrict__ cloud, const double* __restrict__ grads,
                             double* __restrict__ m, double* __restrict__ v, int n, int cap,
                             double c1, double c2, double lr_c, double lr_op, double lr_sc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;  // group
    if (i >= n) return;
    const double b1 = 0.9, b2 = 0.999, eps = 1.0e-8;
    const int pidx = k < 3 ? CR + k : (k == 3 ? ROP : RSC);
    const double lr = k < 3 ? lr_c : (k == 3 ? lr_op : lr_sc);
    double gi = grads[(long long)k * cap + i];
    if (gi > 10.0) gi = 10.0;
    else if (gi < -10.0) gi = -10.0;
    const double mi = b1 * m[(long long)k * cap + i] + (1.0 - b1) * gi;
    const double vi = b2 * v[(long long)k * cap + i] + (1.0 - b2) * gi * gi;
    m[(long long)k * cap + i] = mi;
    v[(long long)k * cap + i] = vi;
    double p = cloud[(long long)pidx * cap + i];
    p -= lr * (mi * c1) / (refm::fsqrt(vi * c2) + eps);
    if (k < 3) {
        if (p < 0.0) p = 0.0;
        else if (p > 1.0) p = 1.0;
    }
    cloud[(long long)pidx * cap + i] = p;
}

}  // namespace

extern "C" {

// gaussians_from_mesh: vertices xy [nv][2] (pixel coordinates), triangles
// [nt][3], per-pixel points [W*H][3] with valid flags, colour [W*H][3];
// writes the kept triangles' Gaussians to cloud [14][cap] in triangle order
// and their number to *count (device int).  raw_op / raw_sc are the initial
// inverse-sigmoid opacity and log-scale (host-computed, the reference's flog).
int sv_gaussians_from_mesh(const int* xy, int nv, const int* tris, int nt, const double* pts,
                           const uint8_t* pvalid, const double* rgb, int W, int H,
                           double thickness, double raw_op, double raw_sc, double* cloud, int cap,
                           int* count, void* stream) {
    SV_REQUIRE(count && cloud && cap >= 0 && nt >= 0 && nv >= 0 && W > 0 && H > 0 &&
                   (nt == 0 || (xy && tris && pts && pvalid && rgb)),
               "sv_gaussians_from_mesh: bad arguments");
    hipLaunchKernelGGL(k_gauss_from_mesh, dim3(1), dim3(1024), 0, (hipStream_t)stream, xy, nv,
                       tris, nt, pts, pvalid, rgb, W, thickness, raw_op, raw_sc, cloud, cap, count);
    SV_CHECK_LAUNCH();
    return SV_OK;
}

// Workspace of one projection of n Gaussians (items, ranks, offsets).
int sv_splat_workspace_bytes(int n, size_t* bytes) {
    SV_REQUIRE(bytes && n >= 0, "sv_splat_workspace_bytes: bad arguments");
    sv::Carver cv(nullptr);
    cv.take<double>((size_t)NID * n);  // raw items
    cv.take<int>((size_t)NJI * n);
    cv.take<uint8_t>((size_t)n);
    cv.take<double>((size_t)NID * n);  // ranked items
    cv.take<int>((size_t)NJI * n);
    cv.take<long long>((size_t)n + 1);
    cv.take<int>(1);
    *bytes = cv.used;
    return SV_OK;
}

namespace {
struct SplatWs {
    double *raw, *items;
    int *rawi, *itemi, *nkept;
    uint8_t* keep;
    long long* area_off;
};
SplatWs carve_splat(void* ws, int n) {
    sv::Carver cv(ws);
    SplatWs s;
    s.raw = cv.take<double>((size_t)NID * n);
    s.rawi = cv.take<int>((size_t)NJI * n);
    s.keep = cv.take<uint8_t>((size_t)n);
    s.items = cv.take<double>((size_t)NID * n);
    s.itemi = cv.take<int>((size_t)NJI * n);
    s.area_off = cv.take<long long>((size_t)n + 1);
    s.nkept = cv.take<int>(1);
    return s;
}
}  // namespace

// SplatRenderer._prepare: projects the cloud (cloud [14][cap], n Gaussians)
// with the world-to-camera rotation R (row-major 3x3) and translation t
// (host arrays), camera fx / cx / cy, and ranks the kept items by depth.
// The workspace then holds the item list and the footprint offsets; info
// (device, 2 entries, may be null) receives the number of kept items and
// their total footprint area, which sizes the backward buffers.
int sv_splat_prepare(const double* cloud, int n, int cap, const double* R, const double* t,
                     double fx, double cx, double cy, int W, int H, double blur, double cutoff,
                     void* ws, size_t ws_bytes, long long* info, void* stream) {
    SV_REQUIRE(cloud && R && t && ws && n >= 0 && cap >= n && W > 0 && H > 0,
               "sv_splat_prepare: bad arguments");
    size_t need = 0;
    sv_splat_workspace_bytes(n, &need);
    SV_REQUIRE(ws_bytes >= need, "sv_splat_prepare: workspace of %zu bytes needed", need);
    hipStream_t st = (hipStream_t)stream;
    SplatWs s = carve_splat(ws, n);
    Proj pj;
    for (int q = 0; q < 9; ++q) pj.R[q] = R[q];
    for (int q = 0; q < 3; ++q) pj.t[q] = t[q];
    pj.fx = fx;
    pj.cx = cx;
    pj.cy = cy;
    pj.blur = blur;
    pj.cutoff = cutoff;
    pj.W = W;
    pj.H = H;
    if (n > 0) {
        hipLaunchKernelGGL(k_splat_project, dim3((n + 255) / 256), dim3(256), 0, st, cloud, n, cap,
                           pj, s.raw, s.rawi, s.keep);
        hipLaunchKernelGGL(k_splat_rank, dim3((n + 3) / 4), dim3(256), 0, st, s.raw, s.rawi,
                           s.keep, n, s.items, s.itemi);
    }
    hipLaunchKernelGGL(k_count_kept, dim3(1), dim3(256), 0, st, s.keep, n, s.nkept);
    hipLaunchKernelGGL(k_splat_areas, dim3(1), dim3(1024), 0, st, s.itemi, n, s.nkept, s.area_off,
                       info);
    SV_CHECK_LAUNCH();
    return SV_OK;
}

// SplatRenderer.render after sv_splat_prepare: C [H*W][3] and the final
// transmittance T [H*W] (kept for the backward sweep); background (br, bg, bb).
int sv_splat_render(const void* ws, int n, int W, int H, double br, double bg, double bb,
                    double* C, double* T, void* stream) {
    SV_REQUIRE(ws && C && T && W > 0 && H > 0 && n >= 0, "sv_splat_render: bad arguments");
    SplatWs s = carve_splat(const_cast<void*>(ws), n);
    dim3 grid((W + kTile - 1) / kTile, (H + kTile - 1) / kTile);
    hipLaunchKernelGGL(k_splat_render, grid, dim3(256), 0, (hipStream_t)stream, s.items, s.itemi, n,
                       s.nkept, W, H, br, bg, bb, C, T);
    SV_CHECK_LAUNCH();
    return SV_OK;
}

// SplatRenderer.backward: adds the gradients of sum(gC * C) to grads
// [5][cap] (colour r, g, b, raw opacity, raw scale).  contrib holds
// 5 * total doubles, total = the footprint area of the prepared workspace
// (sv_splat_prepare's info[1]).
int sv_splat_backward(const void* ws, int n, int W, int H, const double* gC, const double* T,
                      double* contrib, long long total, double* grads, int cap, void* stream) {
    SV_REQUIRE(ws && gC && T && grads && W > 0 && H > 0 && n >= 0 && cap >= n && total >= 0 &&
                   (total == 0 || contrib),
               "sv_splat_backward: bad arguments");
    hipStream_t st = (hipStream_t)stream;
    SplatWs s = carve_splat(const_cast<void*>(ws), n);
    if (total > 0)
        hipLaunchKernelGGL(k_zero_f64, dim3((unsigned)std::min<long long>((5 * total + 255) / 256, 4096)),
                           dim3(256), 0, st, contrib, 5 * total);
    dim3 grid((W + kTile - 1) / kTile, (H + kTile - 1) / kTile);
    hipLaunchKernelGGL(k_splat_backward, grid, dim3(256), 0, st, s.items, s.itemi, n, s.nkept, W, H,
                       gC, T, s.area_off, contrib);
    if (n > 0)
        hipLaunchKernelGGL((k_splat_reduce<8, 4>), dim3(n), dim3(512), 0, st, s.itemi, n,
                           s.nkept, s.area_off, contrib, grads, cap);
    SV_CHECK_LAUNCH();
    return SV_OK;
}

// refine_gaussians' loss: gC = 2 (C - target) / n and loss = mean squared
// error (device double; partial sums in part, nb = 256 doubles).
int sv_photometric_mse(const double* C, const double* target, long long n, double* gC,
                       double* loss, double* part, void* stream) {
    SV_REQUIRE(C && target && gC && loss && part && n > 0, "sv_photometric_mse: bad arguments");
    hipStream_t st = (hipStream_t)stream;
    const int nb = 256;
    const double inv = 1.0 / (double)n;
    hipLaunchKernelGGL(k_mse, dim3(nb), dim3(256), 0, st, C, target, n, inv, gC, part);
    hipLaunchKernelGGL(k_mse_final, dim3(1), dim3(256), 0, st, part, nb, inv, loss);
    SV_CHECK_LAUNCH();
    return SV_OK;
}

// CloudAdam.step: m, v [5][cap]; c1 = 1 / (1 - b1^t), c2 = 1 / (1 - b2^t)
// from the caller's running products.
int sv_cloud_adam(double* cloud, const double* grads, double* m, double* v, int n, int cap,
                  double c1, double c2, double lr_colour, double lr_op, double lr_scale,
                  void* stream) {
    SV_REQUIRE(cloud && grads && m && v && n >= 0 && cap >= n, "sv_cloud_adam: bad arguments");
    if (n == 0) return SV_OK;
    hipLaunchKernelGGL(k_cloud_adam, dim3((n + 255) / 256, 5), dim3(256), 0, (hipStream_t)stream,
                       cloud, grads, m, v, n, cap, c1, c2, lr_colour, lr_op, lr_scale);
    SV_CHECK_LAUNCH();
    return SV_OK;
}

}  // extern "C"
