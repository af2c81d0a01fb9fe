/*
 * stereo_oracle.c -- CPU restatement of the reference's dense-disparity path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the GPU library is
 * compared against, and the CPU baseline that bench.py times.  Nothing in the
 * product (the HIP library and the Python package) links, loads or calls it.
 *
 * Every function restates one function of the reference (stereonet.py,
 * Ray-5106/From-Scratch-Stereo-Vision) in plain C with IEEE binary64
 * arithmetic, in the same operation order, so that integer results are
 * identical and floating-point results agree to the last few ulps.  The
 * reference builds its exp/log/sqrt by hand (stereonet.py:34-104); this file
 * uses libm instead, which differs by about 1e-16 relative.
 *
 * Compile with -ffp-contract=off so that no multiply-add is fused.
 *
 * How each part is pinned (tests/test_oracle.py):
 *   census, band costs, per-direction SGM, refine_fullres, left_to_right,
 *   lr_refine, median_filter: golden vectors written from the reference
 *   (tests/golden/*.npz) + closed forms (P1=P2=0 gives S=8C; constant centre
 *   equals a brute-force textbook SGM on tiny lines);
 *   network layers: torch float64 library routines (conv2d, conv3d,
 *   interpolate, avg_pool2d, softmax) + central differences;
 *   predict_auto: the reference demo page's embedded predictions;
 *   mirror_h, lr_consistency, fill_invalid, predict_consistent: golden
 *   vectors written from the reference (tests/golden/consistency.npz) +
 *   closed forms (a shift-consistent pair of maps is valid everywhere; a
 *   double flip is the identity);
 *   train_grads: a golden vector written from the reference;
 *   Adam: torch.optim.Adam.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define EXPORT __attribute__((visibility("default")))

/* ------------------------------------------------------------------ */
/* Section 25: census + band SGM + LR check (stereonet.py:5115-5391)   */
/* ------------------------------------------------------------------ */

/* census, stereonet.py:5115-5142.  Offsets dy outer, dx inner, (0,0)
 * skipped; each offset shifts the word left and sets bit 0 when the
 * clamped neighbour is strictly darker.  Returns -1 if the window has more
 * than 64 bits. */
EXPORT int or_census(const double *img, int W, int H, int rx, int ry,
                     uint64_t *out) {
    int nbits = (2 * rx + 1) * (2 * ry + 1) - 1;
    if (nbits > 64) return -1;
    for (int y = 0; y < H; y++) {
        for (int x = 0; x < W; x++) {
            double c = img[y * W + x];
            uint64_t s = 0;
            for (int dy = -ry; dy <= ry; dy++) {
                for (int dx = -rx; dx <= rx; dx++) {
                    if (!dx && !dy) continue;
                    int xx = x + dx, yy = y + dy;
                    if (xx < 0) xx = 0; else if (xx >= W) xx = W - 1;
                    if (yy < 0) yy = 0; else if (yy >= H) yy = H - 1;
                    s = s + s;
                    if (img[yy * W + xx] < c) s += 1;
                }
            }
            out[y * W + x] = s;
        }
    }
    return 0;
}

/* band_costs, stereonet.py:5156-5189.  C[p*K+k] for d = base[p]-R+k. */
EXPORT void or_band_costs(const uint64_t *cl, const uint64_t *cr,
                          const double *gl, const double *gr, const int *base,
                          int W, int H, int R, double ad_weight,
                          double ad_trunc, double invalid, const double *prior,
                          double prior_w, double prior_t, double *C) {
    int K = 2 * R + 1;
    for (long i = 0; i < (long)W * H * K; i++) C[i] = invalid;
    for (int y = 0; y < H; y++) {
        int row = y * W;
        for (int x = 0; x < W; x++) {
            int p = row + x;
            int b = base[p] - R;
            uint64_t a = cl[p];
            double il = gl[p];
            long o = (long)p * K;
            for (int k = 0; k < K; k++) {
                int xr = x - (b + k);
                if (b + k < 0 || xr < 0 || xr >= W) continue;
                int q = row + xr;
                double ad = il - gr[q];
                if (ad < 0) ad = -ad;
                ad = ad * 255.0;
                if (ad > ad_trunc) ad = ad_trunc;
                double c = (double)__builtin_popcountll(a ^ cr[q]) + ad_weight * ad;
                if (prior_w > 0.0) {
                    double e = (double)(b + k) - prior[p];
                    if (e < 0) e = -e;
                    if (e > prior_t) e = prior_t;
                    c += prior_w * e;
                }
                C[o + k] = c;
            }
        }
    }
}

/* _aggregate_path, stereonet.py:5192-5241: one path, added into S. */
EXPORT void or_aggregate_path(const double *C, const int *base, int W, int H,
                              int K, double P1, double P2, int dx, int dy,
                              double *S) {
    long N = (long)W * H;
    double *L = (double *)calloc((size_t)N * K, sizeof(double));
    double *mins = (double *)calloc((size_t)N, sizeof(double));
    for (int yi = 0; yi < H; yi++) {
        int y = dy >= 0 ? yi : H - 1 - yi;
        for (int xi = 0; xi < W; xi++) {
            int x = dx >= 0 ? xi : W - 1 - xi;
            long p = (long)y * W + x;
            long o = p * K;
            int qx = x - dx, qy = y - dy;
            if (qx < 0 || qx >= W || qy < 0 || qy >= H) {
                double m = 1.0e30;
                for (int k = 0; k < K; k++) {
                    double v = C[o + k];
                    L[o + k] = v;
                    if (v < m) m = v;
                }
                mins[p] = m;
                continue;
            }
            long q = (long)qy * W + qx;
            long oq = q * K;
            double mq = mins[q];
            int shift = base[p] - base[q];
            double m = 1.0e30;
            for (int k = 0; k < K; k++) {
                int kq = k + shift;
                double best = mq + P2;
                if (0 <= kq && kq < K) {
                    double v = L[oq + kq];
                    if (v < best) best = v;
                }
                if (0 <= kq - 1 && kq - 1 < K) {
                    double v = L[oq + kq - 1] + P1;
                    if (v < best) best = v;
                }
                if (0 <= kq + 1 && kq + 1 < K) {
                    double v = L[oq + kq + 1] + P1;
                    if (v < best) best = v;
                }
                double v = C[o + k] + best - mq;
                L[o + k] = v;
                if (v < m) m = v;
            }
            mins[p] = m;
        }
    }
    for (long i = 0; i < N * K; i++) S[i] += L[i];
    free(L);
    free(mins);
}

/* refine_fullres, stereonet.py:5244-5305.  conf may be NULL. */
EXPORT int or_refine_fullres(const double *left, const double *right,
                             const double *disp0, int W, int H, int R,
                             double P1, double P2, double ad_weight,
                             double ad_trunc, int paths, double uniq, int rx,
                             int ry, double prior_w, double prior_t,
                             double *out, double *conf) {
    int K = 2 * R + 1;
    long N = (long)W * H;
    int nbits = (2 * rx + 1) * (2 * ry + 1) - 1;
    if (nbits > 64) return -1;
    int *base = (int *)malloc(sizeof(int) * N);
    for (long i = 0; i < N; i++) {
        /* int() truncates toward zero (G1) */
        double t = disp0[i] + 0.5;
        int v = (int)t;
        base[i] = v > 0 ? v : 0;
    }
    uint64_t *cl = (uint64_t *)malloc(sizeof(uint64_t) * N);
    uint64_t *cr = (uint64_t *)malloc(sizeof(uint64_t) * N);
    or_census(left, W, H, rx, ry, cl);
    or_census(right, W, H, rx, ry, cr);
    double invalid = (double)nbits + ad_weight * ad_trunc + prior_w * prior_t + 4.0;
    double *C = (double *)malloc(sizeof(double) * N * K);
    or_band_costs(cl, cr, left, right, base, W, H, R, ad_weight, ad_trunc,
                  invalid, disp0, prior_w, prior_t, C);
    double *S = (double *)calloc((size_t)N * K, sizeof(double));
    static const int dirs[8][2] = {{1, 0}, {-1, 0}, {0, 1}, {0, -1},
                                   {1, 1}, {-1, 1}, {1, -1}, {-1, -1}};
    int nd = paths == 8 ? 8 : 4;
    for (int i = 0; i < nd; i++)
        or_aggregate_path(C, base, W, H, K, P1, P2, dirs[i][0], dirs[i][1], S);
    for (long p = 0; p < N; p++) {
        long o = p * K;
        int bk = 0;
        double bv = S[o];
        for (int k = 1; k < K; k++)
            if (S[o + k] < bv) { bv = S[o + k]; bk = k; }
        if (bk == 0 || bk == K - 1) {
            out[p] = disp0[p];
            if (conf) conf[p] = 0.0;
            continue;
        }
        double sv = 1.0e30;
        for (int k = 0; k < K; k++)
            if (k < bk - 1 || k > bk + 1)
                if (S[o + k] < sv) sv = S[o + k];
        double c = sv < 1.0e29 ? (sv - bv) / (bv + 1.0e-6) : 1.0;
        if (conf) conf[p] = c;
        if (c < uniq) { out[p] = disp0[p]; continue; }
        double a = S[o + bk - 1], b = S[o + bk + 1];
        double den = a - 2.0 * bv + b;
        double off = 0.0;
        if (den > 1.0e-9) off = 0.5 * (a - b) / den;
        out[p] = (double)(base[p] - R + bk) + off;
    }
    free(base); free(cl); free(cr); free(C); free(S);
    return 0;
}

/* _mirror1, stereonet.py:5308-5314 */
static void mirror1(const double *img, int W, int H, double *out) {
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) out[y * W + x] = img[y * W + W - 1 - x];
}

/* int(x - d + 0.5) in [0, W) <=> -1 < t < W (truncation toward zero) */
static int warp_col(int x, double d, int W) {
    double t = (double)x - d + 0.5;
    if (!(t > -1.0) || !(t < (double)W)) return -1;
    return (int)t;
}

/* left_to_right, stereonet.py:5317-5340 (the in-place fill is written out
 * literally). */
EXPORT void or_left_to_right(const double *dl, int W, int H, double *dr) {
    for (long i = 0; i < (long)W * H; i++) dr[i] = -1.0;
    for (int y = 0; y < H; y++) {
        long r = (long)y * W;
        for (int x = 0; x < W; x++) {
            double d = dl[r + x];
            int xr = warp_col(x, d, W);
            if (xr >= 0 && d > dr[r + xr]) dr[r + xr] = d;
        }
        for (int x = 0; x < W; x++) {
            if (dr[r + x] < 0.0) {
                int a = x - 1;
                while (a >= 0 && dr[r + a] < 0.0) a--;
                int b = x + 1;
                while (b < W && dr[r + b] < 0.0) b++;
                double va = a >= 0 ? dr[r + a] : 1.0e9;
                double vb = b < W ? dr[r + b] : 1.0e9;
                double v = va < vb ? va : vb;
                dr[r + x] = v < 1.0e8 ? v : 0.0;
            }
        }
    }
}

/* lr_refine, stereonet.py:5343-5391. */
EXPORT int or_lr_refine(const double *left, const double *right,
                        const double *disp0, int W, int H, double tol, int R,
                        double P1, double P2, double ad_weight,
                        double ad_trunc, int paths, double uniq, int rx,
                        int ry, double prior_w, double prior_t, double *out,
                        int *status) {
    long N = (long)W * H;
    double *dl = (double *)malloc(sizeof(double) * N);
    double *d0r = (double *)malloc(sizeof(double) * N);
    double *mr = (double *)malloc(sizeof(double) * N);
    double *ml = (double *)malloc(sizeof(double) * N);
    double *md = (double *)malloc(sizeof(double) * N);
    double *drm = (double *)malloc(sizeof(double) * N);
    double *dr = (double *)malloc(sizeof(double) * N);
    int rc = or_refine_fullres(left, right, disp0, W, H, R, P1, P2, ad_weight,
                               ad_trunc, paths, uniq, rx, ry, prior_w,
                               prior_t, dl, NULL);
    if (rc == 0) {
        or_left_to_right(disp0, W, H, d0r);
        mirror1(right, W, H, mr);
        mirror1(left, W, H, ml);
        mirror1(d0r, W, H, md);
        rc = or_refine_fullres(mr, ml, md, W, H, R, P1, P2, ad_weight,
                               ad_trunc, paths, uniq, rx, ry, prior_w,
                               prior_t, drm, NULL);
        mirror1(drm, W, H, dr);
    }
    if (rc == 0) {
        for (long i = 0; i < N; i++) { status[i] = 0; out[i] = dl[i]; }
        for (int y = 0; y < H; y++) {
            long r = (long)y * W;
            for (int x = 0; x < W; x++) {
                double d = dl[r + x];
                int xr = warp_col(x, d, W);
                if (xr < 0) { status[r + x] = 1; continue; }
                double e = dr[r + xr];
                if (fabs(d - e) <= tol) continue;
                status[r + x] = e > d ? 1 : 2;
            }
        }
        for (long i = 0; i < N; i++)
            if (status[i] == 2) out[i] = disp0[i];
        for (int y = 0; y < H; y++) {
            long r = (long)y * W;
            for (int x = 0; x < W; x++) {
                if (status[r + x] != 1) continue;
                int a = x - 1;
                while (a >= 0 && status[r + a] == 1) a--;
                int b = x + 1;
                while (b < W && status[r + b] == 1) b++;
                double va = a >= 0 ? out[r + a] : 1.0e9;
                double vb = b < W ? out[r + b] : 1.0e9;
                double v = va < vb ? va : vb;
                out[r + x] = v < 1.0e8 ? v : disp0[r + x];
            }
        }
    }
    free(dl); free(d0r); free(mr); free(ml); free(md); free(drm); free(dr);
    return rc;
}

static int cmp_double(const void *a, const void *b) {
    double x = *(const double *)a, y = *(const double *)b;
    return (x > y) - (x < y);
}

/* median_filter, stereonet.py:2231-2258: clamped window, ascending sort,
 * element len//2. */
EXPORT void or_median_filter(const double *disp, int W, int H, int radius,
                             int passes, double *out) {
    long N = (long)W * H;
    double *cur = (double *)malloc(sizeof(double) * N);
    double *nxt = (double *)malloc(sizeof(double) * N);
    int n2 = (2 * radius + 1) * (2 * radius + 1);
    double *vals = (double *)malloc(sizeof(double) * n2);
    memcpy(cur, disp, sizeof(double) * N);
    for (int it = 0; it < passes; it++) {
        for (int y = 0; y < H; y++) {
            int y0 = y - radius > 0 ? y - radius : 0;
            int y1 = y + radius < H - 1 ? y + radius : H - 1;
            for (int x = 0; x < W; x++) {
                int x0 = x - radius > 0 ? x - radius : 0;
                int x1 = x + radius < W - 1 ? x + radius : W - 1;
                int n = 0;
                for (int yy = y0; yy <= y1; yy++)
                    for (int xx = x0; xx <= x1; xx++) vals[n++] = cur[(long)yy * W + xx];
                qsort(vals, n, sizeof(double), cmp_double);
                nxt[(long)y * W + x] = vals[n / 2];
            }
        }
        double *t = cur; cur = nxt; nxt = t;
    }
    memcpy(out, cur, sizeof(double) * N);
    free(cur); free(nxt); free(vals);
}

/* refine_disparity with planes=False, stereonet.py:5865-5884. */
EXPORT int or_refine_disparity_plain(const double *left, const double *right,
                                     const double *disp0, int W, int H, int R,
                                     double P1, double P2, int rx, int ry,
                                     double tol, double *out, int *status) {
    long N = (long)W * H;
    double *d = (double *)malloc(sizeof(double) * N);
    int rc = or_lr_refine(left, right, disp0, W, H, tol, R, P1, P2, 0.0, 20.0,
                          8, 0.0, rx, ry, 0.0, 4.0, d, status);
    if (rc == 0) or_median_filter(d, W, H, 1, 1, out);
    free(d);
    return rc;
}

/* ------------------------------------------------------------------ */
/* Sections 5-8: layers, network, loss, Adam (stereonet.py:453-1388)   */
/* ------------------------------------------------------------------ */

static int conv_out(int n, int k, int s, int p, int d) {
    int eff = d * (k - 1) + 1;
    int num = n + 2 * p - eff;
    /* Python floor division */
    int q = num >= 0 ? num / s : -((-num + s - 1) / s);
    return q + 1;
}

/* Conv2d.forward, stereonet.py:489-526 */
EXPORT void or_conv2d_fwd(const double *x, int IC, int H, int W,
                          const double *w, const double *b, int OC, int k,
                          int s, int p, int d, double *y) {
    int OH = conv_out(H, k, s, p, d), OW = conv_out(W, k, s, p, d);
    long HW = (long)H * W;
    int kk = k * k, ickk = IC * kk;
    for (int o = 0; o < OC; o++) {
        long wo = (long)o * ickk;
        long ob = (long)o * OH * OW;
        for (int oy = 0; oy < OH; oy++) {
            int iy0 = oy * s - p;
            for (int ox = 0; ox < OW; ox++) {
                int ix0 = ox * s - p;
                double acc = b[o];
                for (int c = 0; c < IC; c++) {
                    long xb = c * HW, wc = wo + (long)c * kk;
                    for (int ky = 0; ky < k; ky++) {
                        int iy = iy0 + ky * d;
                        if (iy < 0 || iy >= H) continue;
                        for (int kx = 0; kx < k; kx++) {
                            int ix = ix0 + kx * d;
                            if (ix < 0 || ix >= W) continue;
                            acc += x[xb + (long)iy * W + ix] * w[wc + ky * k + kx];
                        }
                    }
                }
                y[ob + (long)oy * OW + ox] = acc;
            }
        }
    }
}

/* Conv2d.backward, stereonet.py:528-568.  gw, gb accumulate; gx is
 * overwritten (may be NULL). */
EXPORT void or_conv2d_bwd(const double *x, int IC, int H, int W,
                          const double *w, int OC, int k, int s, int p, int d,
                          const double *gy, double *gx, double *gw,
                          double *gb) {
    int OH = conv_out(H, k, s, p, d), OW = conv_out(W, k, s, p, d);
    long HW = (long)H * W;
    int kk = k * k, ickk = IC * kk;
    double *gxl = gx ? gx : (double *)malloc(sizeof(double) * IC * HW);
    memset(gxl, 0, sizeof(double) * IC * HW);
    for (int o = 0; o < OC; o++) {
        long wo = (long)o * ickk;
        long ob = (long)o * OH * OW;
        double gsum = 0.0;
        for (int oy = 0; oy < OH; oy++) {
            int iy0 = oy * s - p;
            for (int ox = 0; ox < OW; ox++) {
                double g = gy[ob + (long)oy * OW + ox];
                if (g == 0.0) continue;
                gsum += g;
                int ix0 = ox * s - p;
                for (int c = 0; c < IC; c++) {
                    long xb = c * HW, wc = wo + (long)c * kk;
                    for (int ky = 0; ky < k; ky++) {
                        int iy = iy0 + ky * d;
                        if (iy < 0 || iy >= H) continue;
                        for (int kx = 0; kx < k; kx++) {
                            int ix = ix0 + kx * d;
                            if (ix < 0 || ix >= W) continue;
                            long xi = xb + (long)iy * W + ix;
                            long wi = wc + ky * k + kx;
                            gw[wi] += x[xi] * g;
                            gxl[xi] += w[wi] * g;
                        }
                    }
                }
            }
        }
        gb[o] += gsum;
    }
    if (!gx) free(gxl);
}

/* Conv3d.forward, stereonet.py:596-641 (k x k x k, stride 1, pad k//2) */
EXPORT void or_conv3d_fwd(const double *x, int IC, int D, int H, int W,
                          const double *w, const double *b, int OC, int k,
                          double *y) {
    int p = k / 2;
    long HW = (long)H * W, DHW = (long)D * HW;
    int kk = k * k, kkk = k * kk;
    long ickkk = (long)IC * kkk;
    for (int o = 0; o < OC; o++) {
        long wo = o * ickkk, ob = o * DHW;
        for (int od = 0; od < D; od++)
            for (int oy = 0; oy < H; oy++)
                for (int ox = 0; ox < W; ox++) {
                    double acc = b[o];
                    for (int c = 0; c < IC; c++) {
                        long xb = c * DHW, wc = wo + (long)c * kkk;
                        for (int kd = 0; kd < k; kd++) {
                            int dd = od - p + kd;
                            if (dd < 0 || dd >= D) continue;
                            for (int ky = 0; ky < k; ky++) {
                                int yy = oy - p + ky;
                                if (yy < 0 || yy >= H) continue;
                                for (int kx = 0; kx < k; kx++) {
                                    int xx = ox - p + kx;
                                    if (xx < 0 || xx >= W) continue;
                                    acc += x[xb + dd * HW + (long)yy * W + xx] *
                                           w[wc + kd * kk + ky * k + kx];
                                }
                            }
                        }
                    }
                    y[ob + od * HW + (long)oy * W + ox] = acc;
                }
    }
}

/* Conv3d.backward, stereonet.py:643-692 */
EXPORT void or_conv3d_bwd(const double *x, int IC, int D, int H, int W,
                          const double *w, int OC, int k, const double *gy,
                          double *gx, double *gw, double *gb) {
    int p = k / 2;
    long HW = (long)H * W, DHW = (long)D * HW;
    int kk = k * k, kkk = k * kk;
    long ickkk = (long)IC * kkk;
    double *gxl = gx ? gx : (double *)malloc(sizeof(double) * IC * DHW);
    memset(gxl, 0, sizeof(double) * IC * DHW);
    for (int o = 0; o < OC; o++) {
        long wo = o * ickkk, ob = o * DHW;
        double gsum = 0.0;
        for (int od = 0; od < D; od++)
            for (int oy = 0; oy < H; oy++)
                for (int ox = 0; ox < W; ox++) {
                    double g = gy[ob + od * HW + (long)oy * W + ox];
                    if (g == 0.0) continue;
                    gsum += g;
                    for (int c = 0; c < IC; c++) {
                        long xb = c * DHW, wc = wo + (long)c * kkk;
                        for (int kd = 0; kd < k; kd++) {
                            int dd = od - p + kd;
                            if (dd < 0 || dd >= D) continue;
                            for (int ky = 0; ky < k; ky++) {
                                int yy = oy - p + ky;
                                if (yy < 0 || yy >= H) continue;
                                for (int kx = 0; kx < k; kx++) {
                                    int xx = ox - p + kx;
                                    if (xx < 0 || xx >= W) continue;
                                    long xi = xb + dd * HW + (long)yy * W + xx;
                                    long wi = wc + kd * kk + ky * k + kx;
                                    gw[wi] += x[xi] * g;
                                    gxl[xi] += w[wi] * g;
                                }
                            }
                        }
                    }
                }
        gb[o] += gsum;
    }
    if (!gx) free(gxl);
}

/* LeakyReLU, stereonet.py:698-723 (x > 0 strictly) */
static void lrelu_fwd(const double *x, long n, double s, double *y) {
    for (long i = 0; i < n; i++) y[i] = x[i] > 0.0 ? x[i] : x[i] * s;
}
static void lrelu_bwd(const double *x, long n, double s, double *g) {
    for (long i = 0; i < n; i++) g[i] = x[i] > 0.0 ? g[i] : g[i] * s;
}

/* CostVolume.forward, stereonet.py:771-799: |fl - fr(x-d)|, 2.0 if x < d */
EXPORT void or_costvol_fwd(const double *fl, const double *fr, int C, int D,
                           int H, int W, double *vol) {
    long HW = (long)H * W, DHW = (long)D * HW;
    for (int c = 0; c < C; c++)
        for (int d = 0; d < D; d++)
            for (int y = 0; y < H; y++)
                for (int x = 0; x < W; x++) {
                    long idx = c * DHW + d * HW + (long)y * W + x;
                    int xr = x - d;
                    if (xr < 0) { vol[idx] = 2.0; continue; }
                    double v = fl[c * HW + (long)y * W + x] - fr[c * HW + (long)y * W + xr];
                    vol[idx] = v < 0.0 ? -v : v;
                }
}

/* CostVolume.backward, stereonet.py:801-825 (sign -1 when diff < 0 else +1) */
EXPORT void or_costvol_bwd(const double *fl, const double *fr, int C, int D,
                           int H, int W, const double *gvol, double *gfl,
                           double *gfr) {
    long HW = (long)H * W, DHW = (long)D * HW;
    memset(gfl, 0, sizeof(double) * C * HW);
    memset(gfr, 0, sizeof(double) * C * HW);
    for (int c = 0; c < C; c++)
        for (int d = 0; d < D; d++)
            for (int y = 0; y < H; y++)
                for (int x = 0; x < W; x++) {
                    int xr = x - d;
                    if (xr < 0) continue;
                    long idx = c * DHW + d * HW + (long)y * W + x;
                    double v = fl[c * HW + (long)y * W + x] - fr[c * HW + (long)y * W + xr];
                    double sg = v < 0.0 ? -1.0 : 1.0;
                    double g = gvol[idx] * sg;
                    if (g == 0.0) continue;
                    gfl[c * HW + (long)y * W + x] += g;
                    gfr[c * HW + (long)y * W + xr] -= g;
                }
}

/* CostVolume.invalid_mask, stereonet.py:827-840 */
static void invalid_mask(int D, int H, int W, double penalty, double *m) {
    long HW = (long)H * W;
    memset(m, 0, sizeof(double) * D * HW);
    for (int d = 1; d < D; d++)
        for (int y = 0; y < H; y++)
            for (int x = 0; x < d && x < W; x++) m[d * HW + (long)y * W + x] = penalty;
}

/* SoftArgmin.forward, stereonet.py:861-889 (penalty may be NULL) */
EXPORT void or_softargmin_fwd(const double *cost, int D, int H, int W,
                              const double *penalty, double *disp,
                              double *probs) {
    long HW = (long)H * W;
    for (long i = 0; i < HW; i++) {
#define CST(dd) (penalty ? cost[(dd) * HW + i] + penalty[(dd) * HW + i] : cost[(dd) * HW + i])
        double best = -CST(0);
        for (int d = 1; d < D; d++) {
            double v = -CST(d);
            if (v > best) best = v;
        }
        double total = 0.0;
        for (int d = 0; d < D; d++) {
            double e = exp(-CST(d) - best);
            probs[d * HW + i] = e;
            total += e;
        }
#undef CST
        double inv = 1.0 / total, acc = 0.0;
        for (int d = 0; d < D; d++) {
            double p = probs[d * HW + i] * inv;
            probs[d * HW + i] = p;
            acc += d * p;
        }
        disp[i] = acc;
    }
}

/* SoftArgmin.backward, stereonet.py:891-901 */
EXPORT void or_softargmin_bwd(const double *probs, const double *disp, int D,
                              int H, int W, const double *gdisp,
                              double *gcost) {
    long HW = (long)H * W;
    memset(gcost, 0, sizeof(double) * D * HW);
    for (long i = 0; i < HW; i++) {
        double g = gdisp[i];
        if (g == 0.0) continue;
        for (int d = 0; d < D; d++)
            gcost[d * HW + i] = g * probs[d * HW + i] * (disp[i] - d);
    }
}

/* BilinearUpsample taps, stereonet.py:913-950 */
static void bl_tap(int o, int in, int out, int *i0, int *i1, double *w) {
    double sc = (double)in / (double)out;
    double f = (o + 0.5) * sc - 0.5;
    if (f < 0.0) f = 0.0;
    int a = (int)f;
    if (a > in - 1) a = in - 1;
    *i0 = a;
    *i1 = a + 1 < in ? a + 1 : a;
    *w = f - a;
}

/* BilinearUpsample.forward, stereonet.py:913-960 */
EXPORT void or_bilinear_fwd(const double *x, int C, int H, int W, int H2,
                            int W2, double *y) {
    long HW = (long)H * W;
    for (int c = 0; c < C; c++)
        for (int oy = 0; oy < H2; oy++) {
            int y0, y1; double wy;
            bl_tap(oy, H, H2, &y0, &y1, &wy);
            for (int ox = 0; ox < W2; ox++) {
                int x0, x1; double wx;
                bl_tap(ox, W, W2, &x0, &x1, &wx);
                double w00 = (1.0 - wy) * (1.0 - wx), w01 = (1.0 - wy) * wx;
                double w10 = wy * (1.0 - wx), w11 = wy * wx;
                const double *xc = x + c * HW;
                y[(long)c * H2 * W2 + (long)oy * W2 + ox] =
                    xc[y0 * W + x0] * w00 + xc[y0 * W + x1] * w01 +
                    xc[y1 * W + x0] * w10 + xc[y1 * W + x1] * w11;
            }
        }
}

/* BilinearUpsample.backward, stereonet.py:962-973 (scatter, output order) */
EXPORT void or_bilinear_bwd(const double *gy, int C, int H, int W, int H2,
                            int W2, double *gx) {
    long HW = (long)H * W;
    memset(gx, 0, sizeof(double) * C * HW);
    for (int c = 0; c < C; c++)
        for (int oy = 0; oy < H2; oy++) {
            int y0, y1; double wy;
            bl_tap(oy, H, H2, &y0, &y1, &wy);
            for (int ox = 0; ox < W2; ox++) {
                int x0, x1; double wx;
                bl_tap(ox, W, W2, &x0, &x1, &wx);
                double g = gy[(long)c * H2 * W2 + (long)oy * W2 + ox];
                if (g == 0.0) continue;
                double *gc = gx + c * HW;
                gc[y0 * W + x0] += g * ((1.0 - wy) * (1.0 - wx));
                gc[y0 * W + x1] += g * ((1.0 - wy) * wx);
                gc[y1 * W + x0] += g * (wy * (1.0 - wx));
                gc[y1 * W + x1] += g * (wy * wx);
            }
        }
}

/* avgpool2, stereonet.py:979-995 */
EXPORT void or_avgpool2(const double *img, int C, int H, int W, double *out) {
    int H2 = H / 2, W2 = W / 2;
    for (int c = 0; c < C; c++)
        for (int y = 0; y < H2; y++)
            for (int x = 0; x < W2; x++) {
                const double *r0 = img + (long)c * H * W + (long)(2 * y) * W;
                const double *r1 = r0 + W;
                out[(long)c * H2 * W2 + (long)y * W2 + x] =
                    0.25 * (r0[2 * x] + r0[2 * x + 1] + r1[2 * x] + r1[2 * x + 1]);
            }
}

/* pseudo_huber, stereonet.py:1326-1351 (valid may be NULL) */
EXPORT double or_pseudo_huber(const double *pred, const double *target,
                              const uint8_t *valid, long n, double alpha,
                              double *grad, long *count_out) {
    double a2 = alpha * alpha, total = 0.0;
    long count = 0;
    for (long i = 0; i < n; i++) {
        grad[i] = 0.0;
        if (valid && !valid[i]) continue;
        double x = pred[i] - target[i];
        double r = sqrt(x * x + a2);
        total += r - alpha;
        grad[i] = x / r;
        count++;
    }
    if (count_out) *count_out = count;
    if (count == 0) return 0.0;
    double inv = 1.0 / count;
    for (long i = 0; i < n; i++) grad[i] *= inv;
    return total * inv;
}

/* disparity_ce, stereonet.py:1274-1323 */
EXPORT double or_disparity_ce(const double *probs, const double *target,
                              const uint8_t *valid, int D, int H, int W,
                              double sigma, double *gcost, long *count_out) {
    long HW = (long)H * W;
    double inv2s2 = 1.0 / (2.0 * sigma * sigma), total = 0.0;
    long count = 0;
    double q[1024];
    memset(gcost, 0, sizeof(double) * D * HW);
    for (long i = 0; i < HW; i++) {
        if (!valid[i]) continue;
        double dg = target[i], qsum = 0.0;
        for (int d = 0; d < D; d++) {
            double e = d - dg;
            double qv = exp(-e * e * inv2s2);
            q[d] = qv;
            qsum += qv;
        }
        if (qsum <= 0.0) continue;
        double qinv = 1.0 / qsum;
        for (int d = 0; d < D; d++) {
            q[d] *= qinv;
            double pk = probs[d * HW + i];
            if (pk > 1.0e-12) total -= q[d] * log(pk);
            gcost[d * HW + i] = q[d] - pk;
        }
        count++;
    }
    if (count_out) *count_out = count;
    if (count == 0) return 0.0;
    double inv = 1.0 / count;
    for (long i = 0; i < D * HW; i++) gcost[i] *= inv;
    return total * inv;
}

/* Adam.step, stereonet.py:1368-1387.  b1t/b2t are the running products
 * after this step's update (the caller multiplies them in first). */
EXPORT void or_adam_step(double *p, const double *g, double *m, double *v,
                         long n, double lr, double b1, double b2, double eps,
                         double b1t, double b2t, double clip) {
    double c1 = 1.0 / (1.0 - b1t), c2 = 1.0 / (1.0 - b2t);
    for (long i = 0; i < n; i++) {
        double gi = g[i];
        if (gi > clip) gi = clip; else if (gi < -clip) gi = -clip;
        double mi = b1 * m[i] + (1.0 - b1) * gi;
        double vi = b2 * v[i] + (1.0 - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        p[i] -= lr * (mi * c1) / (sqrt(vi * c2) + eps);
    }
}

/* ---------------- StereoNet (stereonet.py:1022-1268) ---------------- */

typedef struct {
    int feat_C, cost_C, D, refine_C, ndil, dil[8];
} OrArch;

/* offsets of each conv's (w, b) in the flat parameter vector, in
 * all_params order (stereonet.py:1195-1197, :1662-1667) */
typedef struct {
    long w, b;
    int ic, oc, k, s, p, d; /* 2D convs; for 3D, s=p=d unused */
} ConvSlot;

typedef struct {
    ConvSlot feat[4], filt[3], ref[2][9];
    int nref; /* convs per refinement incl. the final 1-channel one */
    long nparams;
} Layout;

static void mk2(ConvSlot *c, long *off, int ic, int oc, int k, int s, int d) {
    c->ic = ic; c->oc = oc; c->k = k; c->s = s; c->d = d;
    c->p = (d * (k - 1)) / 2;
    c->w = *off; *off += (long)oc * ic * k * k;
    c->b = *off; *off += oc;
}
static void mk3(ConvSlot *c, long *off, int ic, int oc, int k) {
    c->ic = ic; c->oc = oc; c->k = k; c->s = 1; c->d = 1; c->p = k / 2;
    c->w = *off; *off += (long)oc * ic * k * k * k;
    c->b = *off; *off += oc;
}

static void make_layout(const OrArch *a, Layout *L) {
    long off = 0;
    int C = a->feat_C;
    mk2(&L->feat[0], &off, 1, C, 3, 2, 1);
    mk2(&L->feat[1], &off, C, C, 3, 2, 1);
    mk2(&L->feat[2], &off, C, C, 3, 1, 1);
    mk2(&L->feat[3], &off, C, C, 3, 1, 1);
    mk3(&L->filt[0], &off, C, a->cost_C, 3);
    mk3(&L->filt[1], &off, a->cost_C, a->cost_C, 3);
    mk3(&L->filt[2], &off, a->cost_C, 1, 3);
    L->nref = a->ndil + 1;
    for (int r = 0; r < 2; r++) {
        int prev = 2;
        for (int i = 0; i < a->ndil; i++) {
            mk2(&L->ref[r][i], &off, prev, a->refine_C, 3, 1, a->dil[i]);
            prev = a->refine_C;
        }
        mk2(&L->ref[r][a->ndil], &off, prev, 1, 3, 1, 1);
    }
    L->nparams = off;
}

EXPORT long or_net_num_params(int feat_C, int cost_C, int D, int refine_C,
                              int ndil, const int *dil) {
    OrArch a = {feat_C, cost_C, D, refine_C, ndil, {0}};
    for (int i = 0; i < ndil; i++) a.dil[i] = dil[i];
    Layout L;
    make_layout(&a, &L);
    return L.nparams;
}

/* Forward tape: every layer input (pre-activation values where the
 * backward needs them). */
typedef struct {
    int H, W, H2, W2, h, w;
    int fh[4], fw[4];            /* feature conv output dims */
    double *fin[2][4];           /* input of each feature conv, per image */
    double *fpre[2][4];          /* pre-activation of each feature conv */
    double *fl, *fr, *vol;
    double *cin[3], *cpre[3];    /* cost filter conv inputs / outputs */
    double *probs, *d0;
    double *d0up, *d1;           /* up1 output (x2), level-1 output */
    double *d1up, *d2;
    double *rin[2][9], *rpre[2][9]; /* refinement conv inputs / outputs */
    double *rsum[2];             /* disp + residual before the output ReLU */
} Tape;

static double *dalloc(long n) { return (double *)calloc((size_t)n, sizeof(double)); }

static void tape_free(Tape *t, const Layout *L) {
    for (int im = 0; im < 2; im++)
        for (int i = 0; i < 4; i++) { free(t->fin[im][i]); free(t->fpre[im][i]); }
    free(t->fl); free(t->fr); free(t->vol);
    for (int i = 0; i < 3; i++) { free(t->cin[i]); free(t->cpre[i]); }
    free(t->probs); free(t->d0); free(t->d0up); free(t->d1); free(t->d1up); free(t->d2);
    for (int r = 0; r < 2; r++) {
        for (int i = 0; i < L->nref; i++) { free(t->rin[r][i]); free(t->rpre[r][i]); }
        free(t->rsum[r]);
    }
}

static void feat_forward(const Layout *L, const double *P, const double *img,
                         int H, int W, Tape *t, int im, double **outp) {
    const double *x = img;
    int h = H, w = W, ic = 1;
    for (int i = 0; i < 4; i++) {
        const ConvSlot *c = &L->feat[i];
        int oh = conv_out(h, 3, c->s, c->p, c->d), ow = conv_out(w, 3, c->s, c->p, c->d);
        t->fin[im][i] = dalloc((long)ic * h * w);
        memcpy(t->fin[im][i], x, sizeof(double) * ic * h * w);
        double *y = dalloc((long)c->oc * oh * ow);
        or_conv2d_fwd(x, ic, h, w, P + c->w, P + c->b, c->oc, 3, c->s, c->p, c->d, y);
        t->fpre[im][i] = y;
        t->fh[i] = oh; t->fw[i] = ow;
        if (i < 3) {
            double *a = dalloc((long)c->oc * oh * ow);
            lrelu_fwd(y, (long)c->oc * oh * ow, 0.2, a);
            if (i > 0) free((void *)x == (void *)img ? NULL : (void *)x);
            x = a;
        } else {
            if (x != img) free((void *)x);
            double *o = dalloc((long)c->oc * oh * ow);
            memcpy(o, y, sizeof(double) * c->oc * oh * ow);
            *outp = o;
        }
        h = oh; w = ow; ic = c->oc;
    }
}

static double *refine_forward(const Layout *L, const double *P, int r,
                              const double *disp, const double *guide, int H,
                              int W, Tape *t) {
    long HW = (long)H * W;
    double *x = dalloc(2 * HW);
    memcpy(x, disp, sizeof(double) * HW);
    memcpy(x + HW, guide, sizeof(double) * HW);
    int ic = 2;
    for (int i = 0; i < L->nref; i++) {
        const ConvSlot *c = &L->ref[r][i];
        t->rin[r][i] = x;
        double *y = dalloc((long)c->oc * HW);
        or_conv2d_fwd(x, ic, H, W, P + c->w, P + c->b, c->oc, 3, 1, c->p, c->d, y);
        t->rpre[r][i] = y;
        if (i < L->nref - 1) {
            double *a = dalloc((long)c->oc * HW);
            lrelu_fwd(y, (long)c->oc * HW, 0.2, a);
            x = a;
        }
        ic = c->oc;
    }
    double *res = t->rpre[r][L->nref - 1];
    double *sum = dalloc(HW);
    for (long i = 0; i < HW; i++) sum[i] = disp[i] + res[i];
    t->rsum[r] = sum;
    double *out = dalloc(HW);
    lrelu_fwd(sum, HW, 0.0, out);
    return out;
}

static void net_forward(const OrArch *a, const Layout *L, const double *P,
                        const double *left, const double *right, int H, int W,
                        const double *guide_half, const double *guide_full,
                        Tape *t) {
    memset(t, 0, sizeof(*t));
    feat_forward(L, P, left, H, W, t, 0, &t->fl);
    feat_forward(L, P, right, H, W, t, 1, &t->fr);
    int h = t->fh[3], w = t->fw[3];
    int C = a->feat_C, D = a->D;
    long hw = (long)h * w;
    t->H = H; t->W = W; t->h = h; t->w = w;
    t->vol = dalloc((long)C * D * hw);
    or_costvol_fwd(t->fl, t->fr, C, D, h, w, t->vol);
    const double *x = t->vol;
    int ic = C;
    for (int i = 0; i < 3; i++) {
        const ConvSlot *c = &L->filt[i];
        t->cin[i] = dalloc((long)ic * D * hw);
        memcpy(t->cin[i], x, sizeof(double) * ic * D * hw);
        double *y = dalloc((long)c->oc * D * hw);
        or_conv3d_fwd(x, ic, D, h, w, P + c->w, P + c->b, c->oc, 3, y);
        t->cpre[i] = y;
        if (i > 0) free((void *)x);
        if (i < 2) {
            double *act = dalloc((long)c->oc * D * hw);
            lrelu_fwd(y, (long)c->oc * D * hw, 0.2, act);
            x = act;
        }
        ic = c->oc;
    }
    double *mask = dalloc((long)D * hw);
    invalid_mask(D, h, w, 8.0, mask);
    t->probs = dalloc((long)D * hw);
    t->d0 = dalloc(hw);
    or_softargmin_fwd(t->cpre[2], D, h, w, mask, t->d0, t->probs);
    free(mask);
    int H2 = H / 2, W2 = W / 2;
    t->H2 = H2; t->W2 = W2;
    t->d0up = dalloc((long)H2 * W2);
    or_bilinear_fwd(t->d0, 1, h, w, H2, W2, t->d0up);
    for (long i = 0; i < (long)H2 * W2; i++) t->d0up[i] *= 2.0;
    t->d1 = refine_forward(L, P, 0, t->d0up, guide_half, H2, W2, t);
    t->d1up = dalloc((long)H * W);
    or_bilinear_fwd(t->d1, 1, H2, W2, H, W, t->d1up);
    for (long i = 0; i < (long)H * W; i++) t->d1up[i] *= 2.0;
    t->d2 = refine_forward(L, P, 1, t->d1up, guide_full, H, W, t);
}

static void make_arch(OrArch *a, int feat_C, int cost_C, int D, int refine_C,
                      int ndil, const int *dil) {
    memset(a, 0, sizeof(*a));
    a->feat_C = feat_C; a->cost_C = cost_C; a->D = D; a->refine_C = refine_C;
    a->ndil = ndil;
    for (int i = 0; i < ndil; i++) a->dil[i] = dil[i];
}

/* StereoNet.forward, stereonet.py:1203-1233.  Outputs d0 (h4*w4),
 * d1 (H/2*W/2), d2 (H*W) and optionally the soft-argmin probabilities
 * (D*h4*w4).  dims_out (may be NULL) receives h4, w4. */
EXPORT void or_net_forward(int feat_C, int cost_C, int D, int refine_C,
                           int ndil, const int *dil, const double *P,
                           const double *left, const double *right, int H,
                           int W, const double *guide_half,
                           const double *guide_full, double *d0, double *d1,
                           double *d2, double *probs, int *dims_out) {
    OrArch a;
    make_arch(&a, feat_C, cost_C, D, refine_C, ndil, dil);
    Layout L;
    make_layout(&a, &L);
    Tape t;
    net_forward(&a, &L, P, left, right, H, W, guide_half, guide_full, &t);
    long hw = (long)t.h * t.w;
    if (d0) memcpy(d0, t.d0, sizeof(double) * hw);
    if (d1) memcpy(d1, t.d1, sizeof(double) * t.H2 * t.W2);
    if (d2) memcpy(d2, t.d2, sizeof(double) * H * W);
    if (probs) memcpy(probs, t.probs, sizeof(double) * D * hw);
    if (dims_out) { dims_out[0] = t.h; dims_out[1] = t.w; }
    tape_free(&t, &L);
}

static double *refine_backward(const Layout *L, const double *P, int r,
                               const double *gout, int H, int W, Tape *t,
                               double *G) {
    long HW = (long)H * W;
    double *g = dalloc(HW);
    memcpy(g, gout, sizeof(double) * HW);
    lrelu_bwd(t->rsum[r], HW, 0.0, g);
    double *gcur = dalloc(HW);
    memcpy(gcur, g, sizeof(double) * HW);
    for (int i = L->nref - 1; i >= 0; i--) {
        const ConvSlot *c = &L->ref[r][i];
        if (i < L->nref - 1) lrelu_bwd(t->rpre[r][i], (long)c->oc * HW, 0.2, gcur);
        double *gx = dalloc((long)c->ic * HW);
        or_conv2d_bwd(t->rin[r][i], c->ic, H, W, P + c->w, c->oc, 3, 1, c->p,
                      c->d, gcur, gx, G + c->w, G + c->b);
        free(gcur);
        gcur = gx;
    }
    double *gdisp = dalloc(HW);
    for (long i = 0; i < HW; i++) gdisp[i] = gcur[i] + g[i];
    free(gcur);
    free(g);
    return gdisp;
}

static void feat_backward(const Layout *L, const double *P, Tape *t, int im,
                          const double *gout, double *G) {
    double *gcur = dalloc((long)L->feat[3].oc * t->fh[3] * t->fw[3]);
    memcpy(gcur, gout, sizeof(double) * L->feat[3].oc * t->fh[3] * t->fw[3]);
    for (int i = 3; i >= 0; i--) {
        const ConvSlot *c = &L->feat[i];
        int ih = i == 0 ? t->H : t->fh[i - 1];
        int iw = i == 0 ? t->W : t->fw[i - 1];
        if (i < 3) lrelu_bwd(t->fpre[im][i], (long)c->oc * t->fh[i] * t->fw[i], 0.2, gcur);
        double *gx = dalloc((long)c->ic * ih * iw);
        or_conv2d_bwd(t->fin[im][i], c->ic, ih, iw, P + c->w, c->oc, 3, c->s,
                      c->p, c->d, gcur, gx, G + c->w, G + c->b);
        free(gcur);
        gcur = gx;
    }
    free(gcur);
}

/* StereoNet.backward, stereonet.py:1235-1267.  gcost_extra may be NULL.
 * Gradients accumulate into G. */
static void net_backward(const OrArch *a, const Layout *L, const double *P,
                         Tape *t, const double *g0, const double *g1,
                         const double *g2, const double *gcost_extra,
                         double *G) {
    int H = t->H, W = t->W, H2 = t->H2, W2 = t->W2, h = t->h, w = t->w;
    int C = a->feat_C, D = a->D;
    long hw = (long)h * w;
    double *gd1up = refine_backward(L, P, 1, g2, H, W, t, G);
    for (long i = 0; i < (long)H * W; i++) gd1up[i] *= 2.0;
    double *gd1 = dalloc((long)H2 * W2);
    or_bilinear_bwd(gd1up, 1, H2, W2, H, W, gd1);
    for (long i = 0; i < (long)H2 * W2; i++) gd1[i] += g1[i];
    double *gd0up = refine_backward(L, P, 0, gd1, H2, W2, t, G);
    for (long i = 0; i < (long)H2 * W2; i++) gd0up[i] *= 2.0;
    double *gd0 = dalloc(hw);
    or_bilinear_bwd(gd0up, 1, h, w, H2, W2, gd0);
    for (long i = 0; i < hw; i++) gd0[i] += g0[i];
    double *gcost = dalloc((long)D * hw);
    or_softargmin_bwd(t->probs, t->d0, D, h, w, gd0, gcost);
    if (gcost_extra)
        for (long i = 0; i < (long)D * hw; i++) gcost[i] += gcost_extra[i];
    double *gcur = gcost;
    for (int i = 2; i >= 0; i--) {
        const ConvSlot *c = &L->filt[i];
        if (i < 2) lrelu_bwd(t->cpre[i], (long)c->oc * D * hw, 0.2, gcur);
        double *gx = dalloc((long)c->ic * D * hw);
        or_conv3d_bwd(t->cin[i], c->ic, D, h, w, P + c->w, c->oc, 3, gcur, gx,
                      G + c->w, G + c->b);
        free(gcur);
        gcur = gx;
    }
    double *gfl = dalloc((long)C * hw), *gfr = dalloc((long)C * hw);
    or_costvol_bwd(t->fl, t->fr, C, D, h, w, gcur, gfl, gfr);
    feat_backward(L, P, t, 1, gfr, G);
    feat_backward(L, P, t, 0, gfl, G);
    free(gcur); free(gfl); free(gfr);
    free(gd1up); free(gd1); free(gd0up); free(gd0);
}

/* pool_disp / pool_valid / build_targets, stereonet.py:1572-1612 */
static void pool_valid(const uint8_t *v, int W, int H, uint8_t *out) {
    int H2 = H / 2, W2 = W / 2;
    for (int y = 0; y < H2; y++)
        for (int x = 0; x < W2; x++) {
            long r0 = (long)(2 * y) * W, r1 = r0 + W;
            out[(long)y * W2 + x] = v[r0 + 2 * x] && v[r0 + 2 * x + 1] &&
                                    v[r1 + 2 * x] && v[r1 + 2 * x + 1];
        }
}

/* train_grads, stereonet.py:1619-1653 with LEVEL_WEIGHTS = (.5, .7, 1),
 * CE_WEIGHT = 2 (:1615-1616).  G is overwritten with the gradient.
 * d2_out (may be NULL) receives the full-resolution prediction. */
EXPORT double or_train_grads(int feat_C, int cost_C, int D, int refine_C,
                             int ndil, const int *dil, const double *P,
                             const double *left, const double *right,
                             const double *disp, const uint8_t *valid, int W,
                             int H, double *G, double *d2_out) {
    static const double LW[3] = {0.5, 0.7, 1.0};
    const double CE_WEIGHT = 2.0;
    OrArch a;
    make_arch(&a, feat_C, cost_C, D, refine_C, ndil, dil);
    Layout L;
    make_layout(&a, &L);
    int W2 = W / 2, H2 = H / 2, W4 = W2 / 2, H4 = H2 / 2;
    double *dh = dalloc((long)W2 * H2), *dq = dalloc((long)W4 * H4);
    double *gh = dalloc((long)W2 * H2);
    uint8_t *vh = (uint8_t *)calloc((size_t)W2 * H2, 1);
    uint8_t *vq = (uint8_t *)calloc((size_t)W4 * H4, 1);
    or_avgpool2(disp, 1, H, W, dh);
    for (long i = 0; i < (long)W2 * H2; i++) dh[i] *= 0.5;
    pool_valid(valid, W, H, vh);
    or_avgpool2(left, 1, H, W, gh);
    or_avgpool2(dh, 1, H2, W2, dq);
    for (long i = 0; i < (long)W4 * H4; i++) dq[i] *= 0.5;
    pool_valid(vh, W2, H2, vq);

    memset(G, 0, sizeof(double) * L.nparams);
    Tape t;
    net_forward(&a, &L, P, left, right, H, W, gh, left, &t);
    const double *preds[3] = {t.d0, t.d1, t.d2};
    const double *tg[3] = {dq, dh, disp};
    const uint8_t *vv[3] = {vq, vh, valid};
    long ns[3] = {(long)t.h * t.w, (long)H2 * W2, (long)H * W};
    double *g[3];
    double total = 0.0;
    for (int l = 0; l < 3; l++) {
        g[l] = dalloc(ns[l]);
        double loss = or_pseudo_huber(preds[l], tg[l], vv[l], ns[l], 1.0, g[l], NULL);
        total += LW[l] * loss;
        for (long i = 0; i < ns[l]; i++) g[l][i] *= LW[l];
    }
    double *gce = dalloc((long)D * W4 * H4);
    double ce = or_disparity_ce(t.probs, dq, vq, D, H4, W4, 0.8, gce, NULL);
    total += CE_WEIGHT * ce;
    for (long i = 0; i < (long)D * W4 * H4; i++) gce[i] *= CE_WEIGHT;
    net_backward(&a, &L, P, &t, g[0], g[1], g[2], gce, G);
    if (d2_out) memcpy(d2_out, t.d2, sizeof(double) * H * W);
    for (int l = 0; l < 3; l++) free(g[l]);
    free(gce); free(dh); free(dq); free(gh); free(vh); free(vq);
    tape_free(&t, &L);
    return total;
}

/* predict, stereonet.py:5891-5897 */
EXPORT void or_predict(int feat_C, int cost_C, int D, int refine_C, int ndil,
                       const int *dil, const double *P, const double *left,
                       const double *right, int W, int H, double *out) {
    double *gh = dalloc((long)(W / 2) * (H / 2));
    or_avgpool2(left, 1, H, W, gh);
    or_net_forward(feat_C, cost_C, D, refine_C, ndil, dil, P, left, right, H,
                   W, gh, left, NULL, NULL, out, NULL, NULL);
    free(gh);
}

/* _peak, stereonet.py:3866-3869 */
EXPORT double or_peak(const double *disp, long n, double q) {
    double *s = (double *)malloc(sizeof(double) * n);
    memcpy(s, disp, sizeof(double) * n);
    qsort(s, n, sizeof(double), cmp_double);
    double v = s[(long)((double)n * q)];
    free(s);
    return v > 0.5 ? v : 0.5;
}

/* predict_at_scale, stereonet.py:3932-3960.  out_full is W*H, out_work
 * has room for W*H; dims receives (Ww, Hw). */
EXPORT void or_predict_at_scale(int feat_C, int cost_C, int D, int refine_C,
                                int ndil, const int *dil, const double *P,
                                const double *left, const double *right,
                                int W, int H, int work_width,
                                double *out_full, double *out_work,
                                int *dims) {
    if (work_width >= W) {
        or_predict(feat_C, cost_C, D, refine_C, ndil, dil, P, left, right, W, H, out_full);
        memcpy(out_work, out_full, sizeof(double) * W * H);
        dims[0] = W; dims[1] = H;
        return;
    }
    double scale = work_width / (double)W;
    int Ww = work_width;
    int Hw = (int)(H * scale + 0.5);
    if (Hw < 8) Hw = 8;
    double *lw = dalloc((long)Ww * Hw), *rw = dalloc((long)Ww * Hw);
    or_bilinear_fwd(left, 1, H, W, Hw, Ww, lw);
    or_bilinear_fwd(right, 1, H, W, Hw, Ww, rw);
    or_predict(feat_C, cost_C, D, refine_C, ndil, dil, P, lw, rw, Ww, Hw, out_work);
    or_bilinear_fwd(out_work, 1, Hw, Ww, H, W, out_full);
    double inv = 1.0 / scale;
    for (long i = 0; i < (long)W * H; i++) out_full[i] *= inv;
    dims[0] = Ww; dims[1] = Hw;
    free(lw); free(rw);
}

static int width_for(double pk, int W, double target_peak, int max_width) {
    double m = pk > 1.0e-6 ? pk : 1.0e-6;
    int w = (int)(W * target_peak / m + 0.5);
    if (w > W) w = W;
    if (w > max_width) w = max_width;
    if (w < 48) w = 48;
    return w;
}

/* predict_auto, stereonet.py:3872-3929.  trace (may be NULL, 8 ints)
 * receives [probe_w, probe_h, n_passes, work_1, work_2, work_3, ...]. */
EXPORT void or_predict_auto(int feat_C, int cost_C, int D, int refine_C,
                            int ndil, const int *dil, const double *P,
                            double target_peak, const double *left,
                            const double *right, int W, int H,
                            int probe_width, int max_width, int passes,
                            double *out_full, double *out_work, int *dims,
                            int *trace) {
    int pw = probe_width < W ? probe_width : W;
    int ph = (int)(H * pw / (double)W + 0.5);
    if (ph < 8) ph = 8;
    double *lp = dalloc((long)pw * ph), *rp = dalloc((long)pw * ph);
    double *dp = dalloc((long)pw * ph);
    or_bilinear_fwd(left, 1, H, W, ph, pw, lp);
    or_bilinear_fwd(right, 1, H, W, ph, pw, rp);
    or_predict(feat_C, cost_C, D, refine_C, ndil, dil, P, lp, rp, pw, ph, dp);
    double peak = or_peak(dp, (long)pw * ph, 0.95) * (W / (double)pw);
    int work = width_for(peak, W, target_peak, max_width);
    if (trace) { trace[0] = pw; trace[1] = ph; trace[2] = 1; trace[3] = work; }
    or_predict_at_scale(feat_C, cost_C, D, refine_C, ndil, dil, P, left, right,
                        W, H, work, out_full, out_work, dims);
    for (int it = 0; it < passes - 1; it++) {
        double np_ = or_peak(out_full, (long)W * H, 0.95);
        int nw = width_for(np_, W, target_peak, max_width);
        if (abs(nw - work) <= 0.08 * work) break;
        work = nw;
        if (trace && it + 4 < 8) { trace[2] += 1; trace[it + 4] = work; }
        or_predict_at_scale(feat_C, cost_C, D, refine_C, ndil, dil, P, left,
                            right, W, H, work, out_full, out_work, dims);
    }
    free(lp); free(rp); free(dp);
}

/* ------------------------------------------------------------------ */
/* Section 24: network left-right consistency (stereonet.py:4974-5076) */
/* ------------------------------------------------------------------ */

/* mirror_h, stereonet.py:4974-4983: every row of every channel reversed. */
EXPORT void or_mirror_h(const double *img, int C, int W, int H, double *out) {
    for (long r = 0; r < (long)C * H; r++) {
        const double *src = img + r * W;
        double *dst = out + r * W;
        for (int x = 0; x < W; x++) dst[x] = src[W - 1 - x];
    }
}

/* lr_consistency, stereonet.py:4986-5008.  valid[p] = 0 where x - d leaves
 * [0, W-1] or |d - dr(x - d)| > tol, dr linearly interpolated between
 * columns int(x - d) and the next (clamped).  A NaN disparity, on which the
 * reference's int() raises, is marked invalid here. */
EXPORT void or_lr_consistency(const double *dl, const double *dr, int W, int H, double tol,
                              uint8_t *valid) {
    for (int y = 0; y < H; y++) {
        const long row = (long)y * W;
        for (int x = 0; x < W; x++) {
            const double d = dl[row + x];
            const double xr = x - d;
            if (xr < 0.0 || xr > W - 1 || xr != xr) {
                valid[row + x] = 0;
                continue;
            }
            const int x0 = (int)xr;
            const int x1 = x0 + 1 < W ? x0 + 1 : x0;
            const double f = xr - x0;
            const double v = dr[row + x0] * (1.0 - f) + dr[row + x1] * f;
            double diff = d - v;
            if (diff < 0.0) diff = -diff;
            valid[row + x] = diff > tol ? 0 : 1;
        }
    }
}

/* fill_invalid, stereonet.py:5011-5046: each run of invalid pixels takes
 * the smaller of its nearest valid neighbours on the row (the one there is
 * when only one side has one); a row with no valid pixel is left as is. */
EXPORT void or_fill_invalid(const double *disp, const uint8_t *valid, int W, int H,
                            double *out) {
    memcpy(out, disp, sizeof(double) * (size_t)W * H);
    for (int y = 0; y < H; y++) {
        const long row = (long)y * W;
        int x = 0;
        while (x < W) {
            if (valid[row + x]) {
                x++;
                continue;
            }
            const int start = x;
            while (x < W && !valid[row + x]) x++;
            const int end = x;
            const int has_l = start > 0, has_r = end < W;
            if (!has_l && !has_r) continue;
            const double lv = has_l ? disp[row + start - 1] : 0.0;
            const double rv = has_r ? disp[row + end] : 0.0;
            const double fv = !has_l ? rv : (!has_r ? lv : (lv < rv ? lv : rv));
            for (int i = start; i < end; i++) out[row + i] = fv;
        }
    }
}

/* ------------------------------------------------------------------ */
/* Section 25 (planes): SLIC, plane fit, plane_snap (stereonet.py:5394-5563) */
/* ------------------------------------------------------------------ */

/* slic, stereonet.py:5394-5448.  rgb: 3 floats per pixel in [0,1],
 * interleaved.  lab receives a label per pixel (-1 where no centre ever
 * reached); returns the number of centres. */
EXPORT int or_slic(const double *rgb, int W, int H, int step, double compactness, int iters,
                   int *lab) {
    int K = 0;
    for (int y = step / 2; y < H; y += step)
        for (int x = step / 2; x < W; x += step) K++;
    double *cs = (double *)malloc(sizeof(double) * 5 * (K > 0 ? K : 1));
    int k = 0;
    for (int y = step / 2; y < H; y += step)
        for (int x = step / 2; x < W; x += step) {
            long i = (long)y * W + x;
            cs[5 * k + 0] = rgb[3 * i] * 255.0;
            cs[5 * k + 1] = rgb[3 * i + 1] * 255.0;
            cs[5 * k + 2] = rgb[3 * i + 2] * 255.0;
            cs[5 * k + 3] = (double)x;
            cs[5 * k + 4] = (double)y;
            k++;
        }
    const long N = (long)W * H;
    for (long i = 0; i < N; i++) lab[i] = -1;
    const double q = compactness / step;
    const double m2 = q * q;
    double *dist = (double *)malloc(sizeof(double) * N);
    double *acc = (double *)malloc(sizeof(double) * 6 * (K > 0 ? K : 1));
    for (int it = 0; it < iters; it++) {
        for (long i = 0; i < N; i++) dist[i] = 1.0e30;
        for (int ci = 0; ci < K; ci++) {
            const double cr = cs[5 * ci], cg = cs[5 * ci + 1], cb = cs[5 * ci + 2];
            const double cx = cs[5 * ci + 3], cy = cs[5 * ci + 4];
            int x0 = (int)cx - step, x1 = (int)cx + step, y0 = (int)cy - step, y1 = (int)cy + step;
            if (x0 < 0) x0 = 0;
            if (y0 < 0) y0 = 0;
            if (x1 >= W) x1 = W - 1;
            if (y1 >= H) y1 = H - 1;
            for (int yy = y0; yy <= y1; yy++) {
                const double dy = yy - cy;
                for (int xx = x0; xx <= x1; xx++) {
                    const long i = (long)yy * W + xx;
                    const double a = rgb[3 * i] * 255.0 - cr;
                    const double b = rgb[3 * i + 1] * 255.0 - cg;
                    const double c = rgb[3 * i + 2] * 255.0 - cb;
                    const double dx = xx - cx;
                    const double dd = a * a + b * b + c * c + m2 * (dx * dx + dy * dy);
                    if (dd < dist[i]) {
                        dist[i] = dd;
                        lab[i] = ci;
                    }
                }
            }
        }
        memset(acc, 0, sizeof(double) * 6 * (K > 0 ? K : 1));
        for (long i = 0; i < N; i++) {
            const int l = lab[i];
            if (l < 0) continue;
            double *a = acc + 6 * l;
            a[0] += rgb[3 * i];
            a[1] += rgb[3 * i + 1];
            a[2] += rgb[3 * i + 2];
            a[3] += (double)(i % W);
            a[4] += (double)(i / W);
            a[5] += 1.0;
        }
        for (int ci = 0; ci < K; ci++) {
            const double *a = acc + 6 * ci;
            if (a[5] != 0.0) {
                const double n = a[5];
                cs[5 * ci + 0] = a[0] / n * 255.0;
                cs[5 * ci + 1] = a[1] / n * 255.0;
                cs[5 * ci + 2] = a[2] / n * 255.0;
                cs[5 * ci + 3] = a[3] / n;
                cs[5 * ci + 4] = a[4] / n;
            }
        }
    }
    free(cs);
    free(dist);
    free(acc);
    return K;
}

/* _solve3, stereonet.py:5451-5463 (Cramer's rule); 0 when |det| < 1e-9. */
static int solve3(const double A[9], const double b[3], double out[3]) {
    const double a = A[0], bb = A[1], c = A[2], d = A[3], e = A[4], f = A[5], g = A[6],
                 h = A[7], k = A[8];
    const double det = a * (e * k - f * h) - bb * (d * k - f * g) + c * (d * h - e * g);
    if (fabs(det) < 1.0e-9) return 0;
    out[0] = (b[0] * (e * k - f * h) - bb * (b[1] * k - f * b[2]) + c * (b[1] * h - e * b[2])) / det;
    out[1] = (a * (b[1] * k - f * b[2]) - b[0] * (d * k - f * g) + c * (d * b[2] - b[1] * g)) / det;
    out[2] = (a * (e * b[2] - b[1] * h) - bb * (d * b[2] - b[1] * g) + b[0] * (d * h - e * g)) / det;
    return 1;
}

/* _lsq_plane, stereonet.py:5466-5495: least squares d = a x + b y + c over
 * the points idx[0..n) (x = i % W, y = i / W), sums taken in list order. */
static void lsq_plane(const long *idx, long n, int W, const double *disp, double out[3]) {
    const double nn = (double)n;
    double sx = 0.0, sy = 0.0, sd = 0.0;
    for (long j = 0; j < n; j++) sx += (double)(idx[j] % W);
    for (long j = 0; j < n; j++) sy += (double)(idx[j] / W);
    for (long j = 0; j < n; j++) sd += disp[idx[j]];
    const double mx = sx / nn, my = sy / nn, md = sd / nn;
    double sxx = 0.0, sxy = 0.0, syy = 0.0, sxd = 0.0, syd = 0.0;
    for (long j = 0; j < n; j++) {
        const double u = (double)(idx[j] % W) - mx;
        const double v = (double)(idx[j] / W) - my;
        const double w = disp[idx[j]] - md;
        sxx += u * u;
        sxy += u * v;
        syy += v * v;
        sxd += u * w;
        syd += v * w;
    }
    const double det = sxx * syy - sxy * sxy;
    if (fabs(det) < 1.0e-9) {
        out[0] = 0.0;
        out[1] = 0.0;
        out[2] = md;
        return;
    }
    out[0] = (sxd * syy - syd * sxy) / det;
    out[1] = (syd * sxx - sxd * sxy) / det;
    out[2] = md - out[0] * mx - out[1] * my;
}

static int on_plane(double a, double b, double c, long i, int W, const double *disp, double thr) {
    return fabs(a * (double)(i % W) + b * (double)(i / W) + c - disp[i]) <= thr;
}

/* plane_snap, stereonet.py:5498-5563 (SLIC with compactness 10 and 5
 * iterations).  valid may be NULL.  Returns the number of snapped pixels. */
EXPORT long or_plane_snap(const double *disp, const double *rgb, int W, int H,
                          const uint8_t *valid, int step, double thr, double min_inl,
                          int trials, uint32_t seed, double *out) {
    const long N = (long)W * H;
    int *lab = (int *)malloc(sizeof(int) * (N > 0 ? N : 1));
    const int K = or_slic(rgb, W, H, step, 10.0, 5, lab);
    /* groups in pixel order: counting sort by label */
    long *start = (long *)calloc((size_t)K + 1, sizeof(long));
    for (long i = 0; i < N; i++)
        if (lab[i] >= 0) start[lab[i] + 1]++;
    for (int g = 0; g < K; g++) start[g + 1] += start[g];
    long *members = (long *)malloc(sizeof(long) * (N > 0 ? N : 1));
    long *fill = (long *)malloc(sizeof(long) * ((size_t)K + 1));
    memcpy(fill, start, sizeof(long) * ((size_t)K + 1));
    for (long i = 0; i < N; i++)
        if (lab[i] >= 0) members[fill[lab[i]]++] = i;
    long *pts = (long *)malloc(sizeof(long) * (N > 0 ? N : 1));
    memcpy(out, disp, sizeof(double) * N);
    uint32_t s = seed;
    long snapped = 0;
    for (int g = 0; g < K; g++) {
        const long *gm = members + start[g];
        const long gn = start[g + 1] - start[g];
        long n = 0;
        for (long j = 0; j < gn; j++)
            if (!valid || valid[gm[j]]) pts[n++] = gm[j];
        if (n < 12) continue;
        int have = 0;
        long bestc = -1;
        double best[3] = {0.0, 0.0, 0.0};
        for (int t = 0; t < trials; t++) {
            long tri[3];
            for (int q = 0; q < 3; q++) {
                s ^= s << 13;
                s ^= s >> 17;
                s ^= s << 5;
                tri[q] = pts[s % (uint32_t)n];
            }
            double A[9], bv[3], pl[3];
            for (int q = 0; q < 3; q++) {
                A[3 * q] = (double)(tri[q] % W);
                A[3 * q + 1] = (double)(tri[q] / W);
                A[3 * q + 2] = 1.0;
                bv[q] = disp[tri[q]];
            }
            if (!solve3(A, bv, pl)) continue;
            long cnt = 0;
            for (long j = 0; j < n; j++) cnt += on_plane(pl[0], pl[1], pl[2], pts[j], W, disp, thr);
            if (cnt > bestc) {
                bestc = cnt;
                best[0] = pl[0];
                best[1] = pl[1];
                best[2] = pl[2];
                have = 1;
            }
        }
        if (!have || (double)bestc < min_inl * (double)n) continue;
        long ni = 0;
        for (long j = 0; j < n; j++)
            if (on_plane(best[0], best[1], best[2], pts[j], W, disp, thr)) pts[ni++] = pts[j];
        double fit[3];
        lsq_plane(pts, ni, W, disp, fit);
        for (long j = 0; j < gn; j++) {
            const long i = gm[j];
            const double v = fit[0] * (double)(i % W) + fit[1] * (double)(i / W) + fit[2];
            if (fabs(v - disp[i]) <= thr || (valid && !valid[i])) {
                out[i] = v > 0.0 ? v : 0.0;
                snapped++;
            }
        }
    }
    free(lab);
    free(start);
    free(members);
    free(fill);
    free(pts);
    return snapped;
}
