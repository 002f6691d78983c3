// geometry.cpp -- host-side panel geometry: Gauss-Legendre rules, best_grid and the
// graded box mesher of single_dielectric_box3d.
//
// Node coordinates feed 1e-12 operator parity, so every formula keeps the
// reference's arithmetic order and this file is compiled with
// -ffp-contract=off (no fused multiply-add where the reference has none).
#include <algorithm>
#include <array>
#include <cfloat>
#include <cmath>
#include <deque>
#include <memory>

#include "internal.hpp"

namespace dbie {

namespace {
inline double len3(double a, double b, double c) { return std::sqrt(a * a + b * b + c * c); }

void legendre_pd(int n, double x, double &p, double &dp) {
    double p0 = 1.0, p1 = x;
    for (int k = 2; k <= n; ++k) {
        const double pk = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = pk;
    }
    p = p1;
    dp = n * (x * p1 - p0) / (x * x - 1.0);
}
}  // namespace

// FastGaussQuadrature.gausslegendre: closed forms for n <= 5, Newton on the
// three-term recurrence for larger n (ascending nodes).
void gauss_legendre(int n, double *x, double *w) {
    if (n < 1 || n > 64) fail(DBIE_BAD_ARG, "gauss_legendre: n must be in 1..64");
    switch (n) {
    case 1: x[0] = 0.0; w[0] = 2.0; return;
    case 2:
        x[0] = -1.0 / std::sqrt(3.0); x[1] = 1.0 / std::sqrt(3.0);
        w[0] = w[1] = 1.0;
        return;
    case 3:
        x[0] = -std::sqrt(3.0 / 5.0); x[1] = 0.0; x[2] = std::sqrt(3.0 / 5.0);
        w[0] = 5.0 / 9.0; w[1] = 8.0 / 9.0; w[2] = 5.0 / 9.0;
        return;
    case 4: {
        const double a = 2.0 / 7.0 * std::sqrt(6.0 / 5.0);
        x[0] = -std::sqrt(3.0 / 7.0 + a); x[1] = -std::sqrt(3.0 / 7.0 - a);
        x[2] = std::sqrt(3.0 / 7.0 - a);  x[3] = std::sqrt(3.0 / 7.0 + a);
        const double wo = (18.0 - std::sqrt(30.0)) / 36.0, wi = (18.0 + std::sqrt(30.0)) / 36.0;
        w[0] = wo; w[1] = wi; w[2] = wi; w[3] = wo;
        return;
    }
    case 5: {
        const double b = 2.0 * std::sqrt(10.0 / 7.0);
        x[0] = -std::sqrt(5.0 + b) / 3.0; x[1] = -std::sqrt(5.0 - b) / 3.0; x[2] = 0.0;
        x[3] = std::sqrt(5.0 - b) / 3.0;  x[4] = std::sqrt(5.0 + b) / 3.0;
        const double wo = (322.0 - 13.0 * std::sqrt(70.0)) / 900.0, wi = (322.0 + 13.0 * std::sqrt(70.0)) / 900.0;
        w[0] = wo; w[1] = wi; w[2] = 128.0 / 225.0; w[3] = wi; w[4] = wo;
        return;
    }
    default: break;
    }
    const int half = (n + 1) / 2;
    for (int i = 0; i < half; ++i) {
        double z = std::cos(M_PI * (i + 0.75) / (n + 0.5)), p, dp;
        for (int it = 0; it < 100; ++it) {
            legendre_pd(n, z, p, dp);
            const double dz = p / dp;
            z -= dz;
            if (std::fabs(dz) < 1e-16) break;
        }
        legendre_pd(n, z, p, dp);
        z -= p / dp;
        legendre_pd(n, z, p, dp);
        const double wz = 2.0 / ((1.0 - z * z) * dp * dp);
        x[n - 1 - i] = z;  w[n - 1 - i] = wz;
        x[i] = -z;         w[i] = wz;
    }
    if (n % 2 == 1) x[n / 2] = 0.0;
}

// best_grid (src/utils/best_grid.jl:8-126): continued-fraction convergents and
// semiconvergents of W/H; smallest m*n whose cells have aspect <= alpha.
bool best_grid(double W, double H, double alpha, int *m_out, int *n_out) {
    constexpr int64_t kMaxDen = 1000000, kMaxIter = 10000;
    if (!(W > 0 && H > 0) || alpha < 1) return false;
    const bool swapped = W < H;
    if (swapped) std::swap(W, H);
    const double r = W / H;
    int64_t bm = 0, bn = 0, bN = INT64_MAX;
    double brho = INFINITY;
    auto consider = [&](int64_t m, int64_t n) {
        if (m <= 0 || n <= 0 || m > kMaxDen || n > kMaxDen) return;
        const double ratio = (W * (double)n) / (H * (double)m);
        const double rho = ratio >= 1 ? ratio : 1 / ratio;
        if (rho > alpha) return;
        const int64_t cells = m * n;
        if (cells < bN || (cells == bN && rho < brho)) { bm = m; bn = n; bN = cells; brho = rho; }
    };
    int64_t pmm = 0, pm = 1, qmm = 1, qm = 0;
    double x = r;
    if (!(x < 0x1p62)) return false;   // the first partial quotient would not fit an int64
    for (int64_t it = 0; it < kMaxIter; ++it) {
        // x < 2^62 here: r is checked above, and later x = 1/frac with frac > 1e-12
        const int64_t a = (int64_t)std::floor(x);
        const int64_t p = a * pm + pmm, q = a * qm + qmm;
        consider(p, q);
        if (a > 1) {
            int64_t tmax = a - 1;
            if (pm > 0) tmax = std::min(tmax, (kMaxDen - pmm) / pm);
            if (qm > 0) tmax = std::min(tmax, (kMaxDen - qmm) / qm);
            for (int64_t t = 1; t <= tmax; ++t) consider(t * pm + pmm, t * qm + qmm);
        }
        // p and q grow with every term (r >= 1), and so do the next semiconvergents: once either
        // passes kMaxDen no later candidate is admissible.  Stopping here also keeps a*pm and
        // a*qm (a < 1e12, pm, qm <= kMaxDen) inside int64.
        if (p > kMaxDen || q > kMaxDen) break;
        const double frac = x - (double)a;
        if (std::fabs(frac) <= 1e-12) break;
        pmm = pm; pm = p; qmm = qm; qm = q;
        x = 1 / frac;
    }
    if (bm == 0) return false;
    if (swapped) std::swap(bm, bn);
    *m_out = (int)bm;
    *n_out = (int)bn;
    return true;
}

// gl_barycentric_weights (src/utils/barycentric.jl:6-17)
static void barycentric_weights(int n, const double *x, const double *w, double *lam) {
    double mx = 0.0;
    for (int i = 0; i < n; ++i) {
        const double s = ((n - 1 - i) % 2 == 0) ? 1.0 : -1.0;
        lam[i] = s * std::sqrt((1 - x[i] * x[i]) * w[i] / 2);
        mx = std::max(mx, std::fabs(lam[i]));
    }
    for (int i = 0; i < n; ++i) lam[i] /= mx;
}

// barycentric_row! (src/utils/barycentric.jl:19-42); node snapping uses Julia's
// isapprox(xq, x_i) with rtol = sqrt(eps).
void barycentric_row(int n, const double *x, const double *lam, double xq, double *r) {
    const double rtol = std::sqrt(DBL_EPSILON);
    for (int i = 0; i < n; ++i) {
        if (xq == x[i] || std::fabs(xq - x[i]) <= rtol * std::max(std::fabs(xq), std::fabs(x[i]))) {
            for (int k = 0; k < n; ++k) r[k] = 0.0;
            r[i] = 1.0;
            return;
        }
    }
    double den = 0.0;
    for (int i = 0; i < n; ++i) {
        r[i] = lam[i] / (xq - x[i]);
        den += r[i];
    }
    const double inv = 1.0 / den;
    for (int i = 0; i < n; ++i) r[i] *= inv;
}

namespace {
using V3 = std::array<double, 3>;
struct Cell {  // TempPanel3D (src/shape/box3d.jl:2-19)
    V3 a, b, c, d;
    bool ca, cb, cc, cd;      // corner flags
    bool eab, ebc, ecd, eda;  // edge flags
    bool any() const { return ca || cb || cc || cd || eab || ebc || ecd || eda; }
};

// divide_temp_panel3d (src/shape/box3d.jl:66-97), row-major (i - 1) * ny + j
void split_cell(const Cell &t, int nx, int ny, std::vector<Cell> &out) {
    for (int i = 1; i <= nx; ++i)
        for (int j = 1; j <= ny; ++j) {
            const double u0 = (double)(i - 1) / nx, u1 = (double)i / nx;
            const double v0 = (double)(j - 1) / ny, v1 = (double)j / ny;
            Cell o;
            for (int k = 0; k < 3; ++k) {
                const double ba = t.b[k] - t.a[k], da = t.d[k] - t.a[k];
                o.a[k] = t.a[k] + ba * u0 + da * v0;
                o.b[k] = t.a[k] + ba * u1 + da * v0;
                o.c[k] = t.a[k] + ba * u1 + da * v1;
                o.d[k] = t.a[k] + ba * u0 + da * v1;
            }
            o.ca = (i == 1 && j == 1) && t.ca;
            o.cb = (i == nx && j == 1) && t.cb;
            o.cc = (i == nx && j == ny) && t.cc;
            o.cd = (i == 1 && j == ny) && t.cd;
            o.eab = (j == 1) && t.eab;
            o.ebc = (i == nx) && t.ebc;
            o.ecd = (j == ny) && t.ecd;
            o.eda = (i == 1) && t.eda;
            out.push_back(o);
        }
}

void init_rules(Geometry &g) {
    gauss_legendre(g.nq, g.glx, g.glw);
    barycentric_weights(g.nq, g.glx, g.glw, g.bary);
}

// rect_panel3d_discretize nodes and weights (src/shape/box3d.jl:22-64)
void fill_nodes(Geometry &g) {
    const int nq = g.nq, npp = nq * nq;
    g.N = g.P * npp;
    g.x.resize(g.N); g.y.resize(g.N); g.z.resize(g.N); g.w.resize(g.N);
    double *coord[3] = {g.x.data(), g.y.data(), g.z.data()};
    for (int64_t p = 0; p < g.P; ++p) {
        const double *A = &g.corners[12 * p], *B = A + 3, *C = A + 6, *D = A + 9;
        const double Lab = len3(B[0] - A[0], B[1] - A[1], B[2] - A[2]);
        const double Lbc = len3(C[0] - B[0], C[1] - B[1], C[2] - B[2]);
        double cc[3];
        for (int k = 0; k < 3; ++k) cc[k] = (A[k] + B[k] + C[k] + D[k]) / 4;
        for (int i = 0; i < nq; ++i)
            for (int j = 0; j < nq; ++j) {
                const int64_t n = p * npp + i * nq + j;
                for (int k = 0; k < 3; ++k)
                    coord[k][n] = cc[k] + (B[k] - A[k]) * (g.glx[i] / 2) + (D[k] - A[k]) * (g.glx[j] / 2);
                g.w[n] = g.glw[i] * g.glw[j] * Lab * Lbc / 4;
            }
    }
}
}  // namespace

namespace {
// rect_panel3d_adaptive_panels (src/shape/box3d.jl:99-133): a best_grid rough split of the root,
// then 2x2 splits (first in, first out) of every edge/corner panel longer than l_ec.  Appends the
// fine panels with normal nrm to g.
void append_adaptive(Geometry &g, const Cell &root, const V3 &nrm, double alpha, double l_ec) {
    const double Lab = len3(root.b[0] - root.a[0], root.b[1] - root.a[1], root.b[2] - root.a[2]);
    const double Lda = len3(root.a[0] - root.d[0], root.a[1] - root.d[1], root.a[2] - root.d[2]);
    int nx, ny;
    if (!best_grid(Lab, Lda, alpha, &nx, &ny)) fail(DBIE_BAD_ARG, "best_grid found no tiling");
    std::vector<Cell> scratch;
    split_cell(root, nx, ny, scratch);
    std::deque<Cell> queue(scratch.begin(), scratch.end());  // popfirst! / append!
    while (!queue.empty()) {
        const Cell t = queue.front();
        queue.pop_front();
        const double lab = len3(t.b[0] - t.a[0], t.b[1] - t.a[1], t.b[2] - t.a[2]);
        const double lda = len3(t.a[0] - t.d[0], t.a[1] - t.d[1], t.a[2] - t.d[2]);
        if (t.any() && std::max(lab, lda) > l_ec) {
            scratch.clear();
            split_cell(t, 2, 2, scratch);
            queue.insert(queue.end(), scratch.begin(), scratch.end());
        } else {
            for (const V3 *v : {&t.a, &t.b, &t.c, &t.d}) g.corners.insert(g.corners.end(), v->begin(), v->end());
            g.normal.insert(g.normal.end(), nrm.begin(), nrm.end());
            g.is_edge.push_back(t.any() ? 1 : 0);
        }
    }
}

void check_mesh_args(int nq, double l_ec) {
    if (!(l_ec > 0)) fail(DBIE_BAD_ARG, "l_ec must be positive");
    if (nq < 1 || nq > 8) fail(DBIE_BAD_ARG, "n_quad must be in 1..8");
}

// corner order of the six faces of a box and their outward normals (src/shape/box3d_single.jl:5-32,
// src/shape/box3d_multi.jl:11-27): vertices 0..7 = (+,+,+) (-,+,+) (-,-,+) (+,-,+) (+,+,-) (-,+,-)
// (-,-,-) (+,-,-) relative to the centre
constexpr int kFaceVtx[6][4] = {{0, 1, 2, 3}, {4, 7, 6, 5}, {7, 4, 0, 3}, {6, 2, 1, 5}, {5, 1, 0, 4}, {6, 7, 3, 2}};
const V3 kFaceNrm[6] = {{0, 0, 1}, {0, 0, -1}, {1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}};
}  // namespace

// single_dielectric_box3d (src/shape/box3d_single.jl:1-70) with the per-face
// rect_panel3d_adaptive_panels refinement (src/shape/box3d.jl:99-133).
Geometry *make_box(double Lx, double Ly, double Lz, int nq, double l_ec, double alpha) {
    if (!(Lx > 0 && Ly > 0 && Lz > 0)) fail(DBIE_BAD_ARG, "box sides must be positive");
    check_mesh_args(nq, l_ec);
    const V3 vtx[8] = {{Lx / 2, Ly / 2, Lz / 2},   {-Lx / 2, Ly / 2, Lz / 2},  {-Lx / 2, -Ly / 2, Lz / 2},
                       {Lx / 2, -Ly / 2, Lz / 2},  {Lx / 2, Ly / 2, -Lz / 2},  {-Lx / 2, Ly / 2, -Lz / 2},
                       {-Lx / 2, -Ly / 2, -Lz / 2}, {Lx / 2, -Ly / 2, -Lz / 2}};
    std::unique_ptr<Geometry> g(new Geometry);
    g->nq = nq;
    init_rules(*g);
    for (int f = 0; f < 6; ++f) {
        const Cell root{vtx[kFaceVtx[f][0]], vtx[kFaceVtx[f][1]], vtx[kFaceVtx[f][2]], vtx[kFaceVtx[f][3]],
                        true, true, true, true, true, true, true, true};
        append_adaptive(*g, root, kFaceNrm[f], alpha, l_ec);
    }
    g->P = (int64_t)g->is_edge.size();
    fill_nodes(*g);
    return g.release();
}

// ---- several touching boxes (src/shape/box3d_multi.jl) -----------------------------------------
namespace {
using Quad = std::array<V3, 4>;  // corners a, b, c, d
constexpr double kFaceTol = 1e-10;

// Julia's min and max on floats (-0.0 orders below 0.0)
double jmin(double x, double y) { return (y < x || (y == x && std::signbit(y))) ? y : x; }
double jmax(double x, double y) { return (y > x || (y == x && !std::signbit(y))) ? y : x; }

// in-plane axes (u, v) and normal axis of an axis-aligned rectangle
struct Axes {
    int u, v, n;
};
Axes plane_axes(const V3 &nrm) {
    const double a0 = std::fabs(nrm[0]), a1 = std::fabs(nrm[1]), a2 = std::fabs(nrm[2]);
    Axes ax = a0 > 0.5 ? Axes{1, 2, 0} : (a1 > 0.5 ? Axes{0, 2, 1} : Axes{0, 1, 2});
    if (!(a0 > 0.5 || a1 > 0.5 || a2 > 0.5)) fail(DBIE_BAD_ARG, "box faces must be axis-aligned");
    ax.n = a0 > 0.5 ? 0 : (a1 > 0.5 ? 1 : 2);
    return ax;
}

V3 in_plane(const Axes &ax, double u, double v, double w) {
    V3 p{0.0, 0.0, 0.0};
    p[ax.u] = u;
    p[ax.v] = v;
    p[ax.n] = w;
    return p;
}

// extent of a rectangle along one axis
void extent(const Quad &q, int k, double &lo, double &hi) {
    lo = jmin(jmin(jmin(q[0][k], q[1][k]), q[2][k]), q[3][k]);
    hi = jmax(jmax(jmax(q[0][k], q[1][k]), q[2][k]), q[3][k]);
}

// The six faces of a box (centre c, sides L) (_box3d_faces_at_center, src/shape/box3d_multi.jl:4-30).
void box_faces(const double *box, Quad faces[6]) {
    const double cx = box[0], cy = box[1], cz = box[2], hx = box[3] / 2, hy = box[4] / 2, hz = box[5] / 2;
    const V3 v[8] = {{cx + hx, cy + hy, cz + hz}, {cx - hx, cy + hy, cz + hz}, {cx - hx, cy - hy, cz + hz},
                     {cx + hx, cy - hy, cz + hz}, {cx + hx, cy + hy, cz - hz}, {cx - hx, cy + hy, cz - hz},
                     {cx - hx, cy - hy, cz - hz}, {cx + hx, cy - hy, cz - hz}};
    for (int f = 0; f < 6; ++f)
        for (int k = 0; k < 4; ++k) faces[f][k] = v[kFaceVtx[f][k]];
}

// Overlap of two coplanar rectangles with opposite normals, wound like the first
// (_rect_overlap_3d, src/shape/box3d_multi.jl:36-132).
bool rect_overlap(const Quad &f1, const V3 &n1, const Quad &f2, const V3 &n2, Quad &out) {
    if (len3(n1[0] + n2[0], n1[1] + n2[1], n1[2] + n2[2]) > kFaceTol) return false;
    for (const V3 &p : f2)
        if (std::fabs(n1[0] * (p[0] - f1[0][0]) + n1[1] * (p[1] - f1[0][1]) + n1[2] * (p[2] - f1[0][2])) > kFaceTol)
            return false;
    const Axes ax = plane_axes(n1);
    double u1lo, u1hi, v1lo, v1hi, u2lo, u2hi, v2lo, v2hi;
    extent(f1, ax.u, u1lo, u1hi);
    extent(f1, ax.v, v1lo, v1hi);
    extent(f2, ax.u, u2lo, u2hi);
    extent(f2, ax.v, v2lo, v2hi);
    const double lo_u = jmax(u1lo, u2lo), hi_u = jmin(u1hi, u2hi);
    const double lo_v = jmax(v1lo, v2lo), hi_v = jmin(v1hi, v2hi);
    if (hi_u - lo_u < kFaceTol || hi_v - lo_v < kFaceTol) return false;
    const double w = f1[0][ax.n];
    const double ab_u = f1[1][ax.u] - f1[0][ax.u], ab_v = f1[1][ax.v] - f1[0][ax.v];
    const double ad_u = f1[3][ax.u] - f1[0][ax.u], ad_v = f1[3][ax.v] - f1[0][ax.v];
    if (std::fabs(ab_u) > kFaceTol) {  // a -> b runs along u
        const double us = ab_u >= 0 ? lo_u : hi_u, ue = ab_u >= 0 ? hi_u : lo_u;
        const double vs = ad_v >= 0 ? lo_v : hi_v, ve = ad_v >= 0 ? hi_v : lo_v;
        out = {in_plane(ax, us, vs, w), in_plane(ax, ue, vs, w), in_plane(ax, ue, ve, w), in_plane(ax, us, ve, w)};
    } else {  // a -> b runs along v
        const double us = ad_u >= 0 ? lo_u : hi_u, ue = ad_u >= 0 ? hi_u : lo_u;
        const double vs = ab_v >= 0 ? lo_v : hi_v, ve = ab_v >= 0 ? hi_v : lo_v;
        out = {in_plane(ax, us, vs, w), in_plane(ax, us, ve, w), in_plane(ax, ue, ve, w), in_plane(ax, ue, vs, w)};
    }
    return true;
}

struct SharedFace {
    Quad region;
    int lo, hi;  // box ids, lo < hi
    V3 nrm;      // from box hi towards box lo
};

// _detect_shared_faces_3d (src/shape/box3d_multi.jl:139-162)
std::vector<SharedFace> shared_faces(int nb, const double *boxes) {
    std::vector<SharedFace> out;
    Quad fi[6], fj[6];
    for (int i = 0; i < nb; ++i) {
        box_faces(boxes + 6 * i, fi);
        for (int j = i + 1; j < nb; ++j) {
            box_faces(boxes + 6 * j, fj);
            for (int a = 0; a < 6; ++a)
                for (int b = 0; b < 6; ++b) {
                    Quad reg;
                    if (rect_overlap(fi[a], kFaceNrm[a], fj[b], kFaceNrm[b], reg))
                        out.push_back({reg, i, j, V3{-kFaceNrm[a][0], -kFaceNrm[a][1], -kFaceNrm[a][2]}});
                }
        }
    }
    return out;
}

// sort!(unique!(x)) under Julia's isequal / isless: -0.0 and 0.0 stay distinct, -0.0 first
void sort_unique(std::vector<double> &x) {
    auto less = [](double a, double b) { return a < b || (a == b && std::signbit(a) && !std::signbit(b)); };
    std::sort(x.begin(), x.end(), less);
    x.erase(std::unique(x.begin(), x.end(),
                        [](double a, double b) { return a == b && std::signbit(a) == std::signbit(b); }),
            x.end());
}

// A face minus the rectangles it shares with other boxes, as the cells of the grid their edges
// span, each with the flags of its physical edges (ab, bc, cd, da) and corners (a, b, c, d)
// (_subtract_rects_from_face_3d, src/shape/box3d_multi.jl:169-308).
void subtract_rects(const Quad &face, const V3 &nrm, const std::vector<Quad> &cut, std::vector<Cell> &out) {
    if (cut.empty()) {
        out.push_back(Cell{face[0], face[1], face[2], face[3], true, true, true, true, true, true, true, true});
        return;
    }
    const Axes ax = plane_axes(nrm);
    const double w = face[0][ax.n];
    double fu0, fu1, fv0, fv1;
    extent(face, ax.u, fu0, fu1);
    extent(face, ax.v, fv0, fv1);
    std::vector<double> us{fu0, fu1}, vs{fv0, fv1};
    for (const Quad &r : cut)
        for (const V3 &p : r) {
            us.push_back(p[ax.u]);
            vs.push_back(p[ax.v]);
        }
    sort_unique(us);
    sort_unique(vs);
    us.erase(std::remove_if(us.begin(), us.end(), [&](double u) { return !(fu0 - kFaceTol <= u && u <= fu1 + kFaceTol); }),
             us.end());
    vs.erase(std::remove_if(vs.begin(), vs.end(), [&](double v) { return !(fv0 - kFaceTol <= v && v <= fv1 + kFaceTol); }),
             vs.end());
    const int nu = (int)us.size() - 1, nv = (int)vs.size() - 1;
    if (nu < 1 || nv < 1) return;
    std::vector<char> sh((size_t)nu * nv);
    auto S = [&](int i, int j) -> char & { return sh[(size_t)i * nv + j]; };
    for (int i = 0; i < nu; ++i)
        for (int j = 0; j < nv; ++j) {
            if (us[i + 1] - us[i] < kFaceTol || vs[j + 1] - vs[j] < kFaceTol) {
                S(i, j) = 1;  // degenerate cell
                continue;
            }
            const double um = (us[i] + us[i + 1]) / 2, vm = (vs[j] + vs[j + 1]) / 2;
            char hit = 0;
            for (const Quad &r : cut) {
                double ru0, ru1, rv0, rv1;
                extent(r, ax.u, ru0, ru1);
                extent(r, ax.v, rv0, rv1);
                if (ru0 - kFaceTol <= um && um <= ru1 + kFaceTol && rv0 - kFaceTol <= vm && vm <= rv1 + kFaceTol) {
                    hit = 1;
                    break;
                }
            }
            S(i, j) = hit;
        }
    // a grid point is physical when any cell around it is shared or outside the face
    auto blocked = [&](int i, int j) { return i < 0 || i >= nu || j < 0 || j >= nv || S(i, j); };
    auto corner = [&](int gi, int gj) {
        return blocked(gi, gj) || blocked(gi - 1, gj) || blocked(gi, gj - 1) || blocked(gi - 1, gj - 1);
    };
    for (int i = 0; i < nu; ++i)
        for (int j = 0; j < nv; ++j) {
            if (S(i, j)) continue;
            Cell c;
            c.a = in_plane(ax, us[i], vs[j], w);
            c.b = in_plane(ax, us[i + 1], vs[j], w);
            c.c = in_plane(ax, us[i + 1], vs[j + 1], w);
            c.d = in_plane(ax, us[i], vs[j + 1], w);
            // an edge is physical unless the cell across it is another remaining cell
            c.eab = j == 0 || S(i, j - 1);
            c.ebc = i == nu - 1 || S(i + 1, j);
            c.ecd = j == nv - 1 || S(i, j + 1);
            c.eda = i == 0 || S(i - 1, j);
            c.ca = corner(i, j);
            c.cb = corner(i + 1, j);
            c.cc = corner(i + 1, j + 1);
            c.cd = corner(i, j + 1);
            out.push_back(c);
        }
}
}  // namespace

// multi_dielectric_box3d (src/shape/box3d_multi.jl:313-386): the external pieces of every box's
// faces (box order, face order, cell order) then the shared faces (detection order), each meshed by
// rect_panel3d_adaptive_panels.  eps_in / eps_out receive the permittivities of every panel.
Geometry *make_multi_box(int nq, double l_ec, int nb, const double *boxes, const double *epses, double eps_out,
                         double alpha, std::vector<double> &eps_in_p, std::vector<double> &eps_out_p) {
    if (nb < 1 || !boxes || !epses) fail(DBIE_BAD_ARG, "need at least one box and one permittivity per box");
    check_mesh_args(nq, l_ec);
    for (int b = 0; b < nb; ++b)
        if (!(boxes[6 * b + 3] > 0 && boxes[6 * b + 4] > 0 && boxes[6 * b + 5] > 0))
            fail(DBIE_BAD_ARG, "box sides must be positive");
    std::unique_ptr<Geometry> g(new Geometry);
    g->nq = nq;
    init_rules(*g);
    eps_in_p.clear();
    eps_out_p.clear();
    auto mesh = [&](const Cell &root, const V3 &nrm, double ei, double eo) {
        append_adaptive(*g, root, nrm, alpha, l_ec);
        eps_in_p.resize(g->is_edge.size(), ei);
        eps_out_p.resize(g->is_edge.size(), eo);
    };
    const std::vector<SharedFace> shared = shared_faces(nb, boxes);
    Quad faces[6];
    std::vector<Quad> cut;
    std::vector<Cell> cells;
    for (int b = 0; b < nb; ++b) {
        box_faces(boxes + 6 * b, faces);
        for (int f = 0; f < 6; ++f) {
            const V3 &fn = kFaceNrm[f];
            cut.clear();
            for (const SharedFace &sf : shared) {
                if (sf.lo != b && sf.hi != b) continue;
                Quad ov;
                if (rect_overlap(faces[f], fn, sf.region, V3{-fn[0], -fn[1], -fn[2]}, ov)) cut.push_back(ov);
            }
            cells.clear();
            subtract_rects(faces[f], fn, cut, cells);
            for (const Cell &c : cells) mesh(c, fn, epses[b], eps_out);
        }
    }
    for (const SharedFace &sf : shared)
        mesh(Cell{sf.region[0], sf.region[1], sf.region[2], sf.region[3], true, true, true, true, true, true, true,
                  true},
             sf.nrm, epses[sf.hi], epses[sf.lo]);
    g->P = (int64_t)g->is_edge.size();
    if (g->P == 0) fail(DBIE_BAD_ARG, "the boxes leave no surface");
    fill_nodes(*g);
    return g.release();
}

Geometry *make_from_panels(int64_t P, int nq, const double *corners, const double *normals, const int *is_edge) {
    if (P <= 0 || !corners || !normals) fail(DBIE_BAD_ARG, "from_panels: need P > 0, corners and normals");
    if (nq < 1 || nq > 8) fail(DBIE_BAD_ARG, "n_quad must be in 1..8");
    std::unique_ptr<Geometry> g(new Geometry);
    g->nq = nq;
    g->P = P;
    init_rules(*g);
    g->corners.assign(corners, corners + 12 * P);
    g->normal.assign(normals, normals + 3 * P);
    g->is_edge.assign(P, 0);
    if (is_edge)
        for (int64_t p = 0; p < P; ++p) g->is_edge[p] = is_edge[p] ? 1 : 0;
    fill_nodes(*g);
    return g.release();
}

Geometry *make_from_panel_nodes(int64_t P, int nq, const double *corners, const double *normals, const int *is_edge,
                                const double *points, const double *weights, const double *gl_xs, const double *gl_ws,
                                const double *bary) {
    if (!points || !weights || !gl_xs || !gl_ws)
        fail(DBIE_BAD_ARG, "from_panel_nodes: points, weights, gl_xs and gl_ws are required");
    std::unique_ptr<Geometry> g(make_from_panels(P, nq, corners, normals, is_edge));
    // the caller's rule and nodes, as given (FlatPanel.gl_xs / gl_ws / bary_weights / points / weights)
    for (int i = 0; i < nq; ++i) {
        g->glx[i] = gl_xs[i];
        g->glw[i] = gl_ws[i];
    }
    if (bary) {
        for (int i = 0; i < nq; ++i) g->bary[i] = bary[i];
    } else {
        barycentric_weights(nq, g->glx, g->glw, g->bary);
    }
    for (int64_t n = 0; n < g->N; ++n) {
        g->x[n] = points[3 * n];
        g->y[n] = points[3 * n + 1];
        g->z[n] = points[3 * n + 2];
        g->w[n] = weights[n];
    }
    return g.release();
}

namespace {
// psi_n(x) and psi_{n-1}(x) of the orthonormal Hermite functions
// psi_k = H_k(x) exp(-x^2/2) / sqrt(2^k k! sqrt(pi)), by their three-term recurrence.
void hermite_fn(int n, double x, double &pn, double &pn1) {
    double prev = 0.0, cur = std::pow(3.14159265358979311599796346854, -0.25) * std::exp(-x * x / 2);
    for (int k = 0; k < n; ++k) {
        const double next = std::sqrt(2.0 / (k + 1)) * x * cur - std::sqrt((double)k / (k + 1)) * prev;
        prev = cur;
        cur = next;
    }
    pn = cur;
    pn1 = prev;
}
}  // namespace

// Gauss-Hermite nodes (roots of H_n, bracketed on a grid and bisected to adjacent doubles)
// and weights w_i = exp(-x_i^2) / (n psi_{n-1}(x_i)^2); the rule is symmetrised.
void gauss_hermite(int n, double *x, double *w) {
    if (n < 1 || n > 200) fail(DBIE_BAD_ARG, "gauss_hermite: n must be in 1..200");
    const double L = std::sqrt(2.0 * n + 1) + 1;
    const int steps = 2000 * n;
    int found = 0;
    double xa = -L, fa, dummy;
    hermite_fn(n, xa, fa, dummy);
    for (int s = 1; s <= steps && found < n; ++s) {
        const double xb = -L + 2 * L * s / steps;
        double fb;
        hermite_fn(n, xb, fb, dummy);
        if (fb == 0.0) {
            x[found++] = xb;
        } else if (fa != 0.0 && (fa < 0) != (fb < 0)) {
            double a = xa, b = xb, f_a = fa;
            for (;;) {
                const double m = 0.5 * (a + b);
                if (m <= a || m >= b) break;
                double fm;
                hermite_fn(n, m, fm, dummy);
                if (fm == 0.0) { a = b = m; break; }
                if ((fm < 0) == (f_a < 0)) { a = m; f_a = fm; } else { b = m; }
            }
            x[found++] = 0.5 * (a + b);
        }
        xa = xb;
        fa = fb;
    }
    if (found != n) fail(DBIE_BAD_ARG, "gauss_hermite: root bracketing failed");
    for (int i = 0; i < n / 2; ++i) {
        const double r = 0.5 * (x[n - 1 - i] - x[i]);
        x[i] = -r;
        x[n - 1 - i] = r;
    }
    if (n % 2) x[n / 2] = 0.0;
    for (int i = 0; i < n; ++i) {
        double pn, pn1;
        hermite_fn(n, x[i], pn, pn1);
        w[i] = std::exp(-x[i] * x[i]) / (n * pn1 * pn1);
    }
    for (int i = 0; i < n / 2; ++i) {
        const double m = 0.5 * (w[i] + w[n - 1 - i]);
        w[i] = w[n - 1 - i] = m;
    }
}

}  // namespace dbie
