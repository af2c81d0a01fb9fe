// Mesh construction of the downstream 3D output on the host: Bowyer-Watson
// Delaunay triangulation (stereonet.py:2038-2141), the depth-ratio / edge
// and slant culls (:2144-2199, :2324-2372) and mesh-vertex selection
// (:2375-2482).  These are serial, branchy walks over a few hundred to a few
// thousand vertices (each Delaunay insertion depends on the triangulation
// left by the previous one; the vertex selection is a greedy walk with a
// look-back window), so they run as native host code next to the GPU
// kernels rather than as kernels; the per-pixel inputs come from the GPU
// (disparity_to_points, background_region) and the per-triangle outputs go
// back to it (the Gaussian construction and the splat renderer).
//
// Every function visits its inputs in the reference's order and evaluates
// the same float64 expressions (with the reference's own fsqrt), so the
// triangle lists -- including their order, which the later stages inherit
// -- are identical.
#include <algorithm>
#include <unordered_map>
#include <utility>
#include <vector>

#include "common.h"
#include "refmath.h"

namespace {

struct Tri {
    int i, j, k;
    double ux, uy, r2;
};

// _circumcircle, stereonet.py:2038-2049; false when collinear
bool circumcircle(double ax, double ay, double bx, double by, double cx, double cy, double& ux,
                  double& uy, double& r2) {
    const double d = 2.0 * (ax * (by - cy) + bx * (cy - ay) + cx * (ay - by));
    if (refm::fabs_(d) < 1.0e-12) return false;
    const double a2 = ax * ax + ay * ay;
    const double b2 = bx * bx + by * by;
    const double c2 = cx * cx + cy * cy;
    ux = (a2 * (by - cy) + b2 * (cy - ay) + c2 * (ay - by)) / d;
    uy = (a2 * (cx - bx) + b2 * (ax - cx) + c2 * (bx - ax)) / d;
    r2 = (ax - ux) * (ax - ux) + (ay - uy) * (ay - uy);
    return true;
}

}  // namespace

extern "C" {

// delaunay (stereonet.py:2052-2141): xy = n points (x, y); writes up to cap
// index triples to tris and their number to *count (the full count even
// when it exceeds cap, which is then an error).
int sv_delaunay(const double* xy, int n, int* tris, int cap, int* count) {
    SV_REQUIRE(count && n >= 0 && (n == 0 || xy) && cap >= 0 && (cap == 0 || tris),
               "sv_delaunay: bad arguments");
    *count = 0;
    if (n < 3) return SV_OK;
    std::vector<double> vx(n + 3), vy(n + 3);
    double minx = xy[0], maxx = xy[0], miny = xy[1], maxy = xy[1];
    for (int p = 0; p < n; ++p) {
        const double x = xy[2 * p], y = xy[2 * p + 1];
        vx[p] = x;
        vy[p] = y;
        if (x < minx) minx = x;
        if (x > maxx) maxx = x;
        if (y < miny) miny = y;
        if (y > maxy) maxy = y;
    }
    const double dx = maxx - minx, dy = maxy - miny;
    double span = dx > dy ? dx : dy;
    if (span <= 0.0) span = 1.0;
    const double midx = 0.5 * (minx + maxx), midy = 0.5 * (miny + maxy);
    const double m = span * 20.0;
    vx[n] = midx - m;
    vy[n] = midy - span;
    vx[n + 1] = midx;
    vy[n + 1] = midy + m;
    vx[n + 2] = midx + m;
    vy[n + 2] = midy - span;
    auto make_tri = [&](int i, int j, int k, Tri& t) {
        t.i = i;
        t.j = j;
        t.k = k;
        return circumcircle(vx[i], vy[i], vx[j], vy[j], vx[k], vy[k], t.ux, t.uy, t.r2);
    };
    std::vector<Tri> cur, keep, bad;
    {
        Tri t;
        make_tri(n, n + 1, n + 2, t);  // the reference keeps it even if degenerate (never is)
        cur.push_back(t);
    }
    // the cavity's edges in first-use order, counted (the reference's dict)
    std::vector<std::pair<long long, int>> edges;
    std::unordered_map<long long, int> where;
    for (int p = 0; p < n; ++p) {
        const double px = vx[p], py = vy[p];
        keep.clear();
        bad.clear();
        for (const Tri& t : cur) {
            const double ddx = px - t.ux, ddy = py - t.uy;
            if (ddx * ddx + ddy * ddy < t.r2 - 1.0e-12) bad.push_back(t);
            else keep.push_back(t);
        }
        if (bad.empty()) continue;
        edges.clear();
        where.clear();
        for (const Tri& t : bad) {
            const int e[3][2] = {{t.i, t.j}, {t.j, t.k}, {t.k, t.i}};
            for (const auto& ab : e) {
                const int a = ab[0] < ab[1] ? ab[0] : ab[1], b = ab[0] < ab[1] ? ab[1] : ab[0];
                const long long key = (long long)a * (long long)(n + 3) + b;
                auto it = where.find(key);
                if (it == where.end()) {
                    where.emplace(key, (int)edges.size());
                    edges.emplace_back(key, 1);
                } else {
                    edges[it->second].second += 1;
                }
            }
        }
        cur.swap(keep);
        for (const auto& e : edges) {
            if (e.second != 1) continue;
            const int a = (int)(e.first / (n + 3)), b = (int)(e.first % (n + 3));
            Tri t;
            if (make_tri(a, b, p, t)) cur.push_back(t);
        }
    }
    int c = 0;
    for (const Tri& t : cur) {
        if (t.i >= n || t.j >= n || t.k >= n) continue;
        if (c < cap) {
            tris[3 * c] = t.i;
            tris[3 * c + 1] = t.j;
            tris[3 * c + 2] = t.k;
        }
        ++c;
    }
    *count = c;
    SV_REQUIRE(c <= cap, "sv_delaunay: %d triangles, room for %d", c, cap);
    return SV_OK;
}

// cull_triangles (stereonet.py:2144-2199) on the vertices' 3D points vp
// ([nv][3], valid flags vv): keeps triangles whose three vertices are valid,
// whose depth ratio is within max_z_ratio and whose longest edge is at most
// max_edge_factor times the median longest edge, in their input order.
int sv_cull_triangles(const double* vp, const uint8_t* vv, int nv, const int* tris, int nt,
                      double max_z_ratio, double max_edge_factor, int* out, int* count) {
    SV_REQUIRE(count && nt >= 0 && nv >= 0 && (nt == 0 || (tris && out && vp && vv)),
               "sv_cull_triangles: bad arguments");
    std::vector<double> lengths;
    std::vector<std::pair<int, double>> prepared;
    for (int t = 0; t < nt; ++t) {
        const int id[3] = {tris[3 * t], tris[3 * t + 1], tris[3 * t + 2]};
        bool ok = true;
        for (int q = 0; q < 3; ++q) {
            SV_REQUIRE(id[q] >= 0 && id[q] < nv, "sv_cull_triangles: vertex %d of %d", id[q], nv);
            if (!vv[id[q]]) ok = false;
        }
        if (!ok) continue;
        const double* P[3] = {vp + 3 * id[0], vp + 3 * id[1], vp + 3 * id[2]};
        const double zmin = refm::fmin_(P[0][2], refm::fmin_(P[1][2], P[2][2]));
        const double zmax = refm::fmax_(P[0][2], refm::fmax_(P[1][2], P[2][2]));
        if (zmin <= 0.0 || zmax > zmin * max_z_ratio) continue;
        double e = 0.0;
        for (int a = 0; a < 3; ++a) {
            const int b = (a + 1) % 3;
            const double dx = P[a][0] - P[b][0], dy = P[a][1] - P[b][1], dz = P[a][2] - P[b][2];
            const double d = refm::fsqrt(dx * dx + dy * dy + dz * dz);
            if (d > e) e = d;
        }
        lengths.push_back(e);
        prepared.emplace_back(t, e);
    }
    *count = 0;
    if (prepared.empty()) return SV_OK;
    std::sort(lengths.begin(), lengths.end());
    const double limit = lengths[lengths.size() / 2] * max_edge_factor;
    int c = 0;
    for (const auto& pe : prepared) {
        if (pe.second > limit) continue;
        for (int q = 0; q < 3; ++q) out[3 * c + q] = tris[3 * pe.first + q];
        ++c;
    }
    *count = c;
    return SV_OK;
}

// cull_slanted (stereonet.py:2324-2372): drops triangles seen nearly
// edge-on from the camera at the origin (|cos| of normal vs. centroid ray
// below max_slant_cos) or degenerate, in input order.
int sv_cull_slanted(const double* vp, const uint8_t* vv, int nv, const int* tris, int nt,
                    double max_slant_cos, int* out, int* count) {
    SV_REQUIRE(count && nt >= 0 && nv >= 0 && (nt == 0 || (tris && out && vp && vv)),
               "sv_cull_slanted: bad arguments");
    int c = 0;
    for (int t = 0; t < nt; ++t) {
        const int id[3] = {tris[3 * t], tris[3 * t + 1], tris[3 * t + 2]};
        bool ok = true;
        for (int q = 0; q < 3; ++q) {
            SV_REQUIRE(id[q] >= 0 && id[q] < nv, "sv_cull_slanted: vertex %d of %d", id[q], nv);
            if (!vv[id[q]]) ok = false;
        }
        if (!ok) continue;
        const double* P[3] = {vp + 3 * id[0], vp + 3 * id[1], vp + 3 * id[2]};
        const double ux = P[1][0] - P[0][0], uy = P[1][1] - P[0][1], uz = P[1][2] - P[0][2];
        const double vx = P[2][0] - P[0][0], vy = P[2][1] - P[0][1], vz = P[2][2] - P[0][2];
        const double nx = uy * vz - uz * vy, ny = uz * vx - ux * vz, nz = ux * vy - uy * vx;
        const double nl = refm::fsqrt(nx * nx + ny * ny + nz * nz);
        if (nl < 1.0e-12) continue;
        const double cxx = (P[0][0] + P[1][0] + P[2][0]) / 3.0;
        const double cyy = (P[0][1] + P[1][1] + P[2][1]) / 3.0;
        const double czz = (P[0][2] + P[1][2] + P[2][2]) / 3.0;
        const double vl = refm::fsqrt(cxx * cxx + cyy * cyy + czz * czz);
        if (vl < 1.0e-12) continue;
        const double cosang = refm::fabs_((nx * cxx + ny * cyy + nz * czz) / (nl * vl));
        if (cosang < max_slant_cos) continue;
        for (int q = 0; q < 3; ++q) out[3 * c + q] = id[q];
        ++c;
    }
    *count = c;
    return SV_OK;
}

// select_points (stereonet.py:2375-2482) on a host disparity map: the four
// corners, a uniform grid walk, then the strongest depth gradients with a
// minimum spacing against the last 40 placed.  valid may be null.  Writes up
// to cap (x, y) pairs; *count is the number chosen.
int sv_select_points(const double* disp, int W, int H, const uint8_t* valid, int target,
                     double edge_fraction, int* xy, int cap, int* count) {
    SV_REQUIRE(disp && W > 0 && H > 0 && xy && count && cap >= 0, "sv_select_points: bad arguments");
    const long long N = (long long)W * H;
    std::vector<double> grad(N, 0.0);
    for (int y = 0; y < H - 1; ++y) {
        const long long row = (long long)y * W;
        for (int x = 0; x < W - 1; ++x) {
            const double gx = disp[row + x + 1] - disp[row + x];
            const double gy = disp[row + W + x] - disp[row + x];
            grad[row + x] = refm::fsqrt(gx * gx + gy * gy);
        }
    }
    const int n_edge = (int)((double)target * edge_fraction);
    const int n_grid = target - n_edge;
    std::vector<char> taken(N, 0);
    int c = 0;
    auto push = [&](int x, int y) {
        if (c < cap) {
            xy[2 * c] = x;
            xy[2 * c + 1] = y;
        }
        ++c;
    };
    auto add = [&](int x, int y) {
        const long long i = (long long)y * W + x;
        if (taken[i]) return false;
        if (valid && !valid[i]) return false;
        taken[i] = 1;
        push(x, y);
        return true;
    };
    const int corners[4][2] = {{0, 0}, {W - 1, 0}, {0, H - 1}, {W - 1, H - 1}};
    for (const auto& q : corners) {
        taken[(long long)q[1] * W + q[0]] = 1;
        push(q[0], q[1]);
    }
    const double step = refm::fsqrt((double)(W * H) / (double)(n_grid > 1 ? n_grid : 1));
    for (double gy = 0.0; gy < H; gy += step) {
        for (double gx = 0.0; gx < W; gx += step) {
            int x = (int)gx, y = (int)gy;
            if (x >= W) x = W - 1;
            if (y >= H) y = H - 1;
            add(x, y);
        }
    }
    // strongest gradients first; equal gradients larger index first (the
    // reference sorts (grad, index) tuples in reverse)
    std::vector<long long> ranked;
    for (long long i = 0; i < N; ++i)
        if (grad[i] > 0.0 && !taken[i]) ranked.push_back(i);
    std::sort(ranked.begin(), ranked.end(), [&](long long a, long long b) {
        if (grad[a] != grad[b]) return grad[a] > grad[b];
        return a > b;
    });
    const double min_sep = 1.5;
    std::vector<std::pair<int, int>> placed;
    for (long long i : ranked) {
        if ((int)placed.size() >= n_edge) break;
        const int x = (int)(i % W), y = (int)(i / W);
        bool ok = true;
        const size_t from = placed.size() > 40 ? placed.size() - 40 : 0;
        for (size_t q = from; q < placed.size(); ++q) {
            if (refm::fabs_((double)(placed[q].first - x)) < min_sep &&
                refm::fabs_((double)(placed[q].second - y)) < min_sep) {
                ok = false;
                break;
            }
        }
        if (ok && add(x, y)) placed.emplace_back(x, y);
    }
    *count = c;
    SV_REQUIRE(c <= cap, "sv_select_points: %d points, room for %d", c, cap);
    return SV_OK;
}

}  // extern "C"
