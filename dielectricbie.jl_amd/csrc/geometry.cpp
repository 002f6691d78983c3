// geometry.cpp -- host-side panel geometry: Gauss-Legendre rules, best_grid and the
// graded box mesher of single_dielectric_box3d.
//
// Node coordinates feed 1e-12 operator parity, so every formula keeps the
// reference's arithmetic order and this file is compiled with
// -ffp-contract=off (no fused multiply-add where the reference has none).
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
    for (int64_t it = 0; it < kMaxIter; ++it) {
        const int64_t a = (int64_t)std::floor(x);
        const int64_t p = a * pm + pmm, q = a * qm + qmm;
        consider(p, q);
        if (a > 1) {
            int64_t tmax = a - 1;
            if (pm > 0) tmax = std::min(tmax, (kMaxDen - pmm) / pm);
            if (qm > 0) tmax = std::min(tmax, (kMaxDen - qmm) / qm);
            for (int64_t t = 1; t <= tmax; ++t) consider(t * pm + pmm, t * qm + qmm);
        }
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

// single_dielectric_box3d (src/shape/box3d_single.jl:1-70) with the per-face
// rect_panel3d_adaptive_panels refinement (src/shape/box3d.jl:99-133).
Geometry *make_box(double Lx, double Ly, double Lz, int nq, double l_ec, double alpha) {
    if (!(Lx > 0 && Ly > 0 && Lz > 0)) fail(DBIE_BAD_ARG, "box sides must be positive");
    if (!(l_ec > 0)) fail(DBIE_BAD_ARG, "l_ec must be positive");
    if (nq < 1 || nq > 8) fail(DBIE_BAD_ARG, "n_quad must be in 1..8");
    const V3 vtx[8] = {{Lx / 2, Ly / 2, Lz / 2},   {-Lx / 2, Ly / 2, Lz / 2},  {-Lx / 2, -Ly / 2, Lz / 2},
                       {Lx / 2, -Ly / 2, Lz / 2},  {Lx / 2, Ly / 2, -Lz / 2},  {-Lx / 2, Ly / 2, -Lz / 2},
                       {-Lx / 2, -Ly / 2, -Lz / 2}, {Lx / 2, -Ly / 2, -Lz / 2}};
    const int faces[6][4] = {{0, 1, 2, 3}, {4, 7, 6, 5}, {7, 4, 0, 3}, {6, 2, 1, 5}, {5, 1, 0, 4}, {6, 7, 3, 2}};
    const V3 nrm[6] = {{0, 0, 1}, {0, 0, -1}, {1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}};

    std::unique_ptr<Geometry> g(new Geometry);
    g->nq = nq;
    init_rules(*g);
    std::vector<Cell> fine, scratch;
    for (int f = 0; f < 6; ++f) {
        Cell root{vtx[faces[f][0]], vtx[faces[f][1]], vtx[faces[f][2]], vtx[faces[f][3]],
                  true, true, true, true, true, true, true, true};
        const double Lab = len3(root.b[0] - root.a[0], root.b[1] - root.a[1], root.b[2] - root.a[2]);
        const double Lda = len3(root.a[0] - root.d[0], root.a[1] - root.d[1], root.a[2] - root.d[2]);
        int nx, ny;
        if (!best_grid(Lab, Lda, alpha, &nx, &ny)) fail(DBIE_BAD_ARG, "best_grid found no tiling");
        scratch.clear();
        split_cell(root, nx, ny, scratch);
        std::deque<Cell> queue(scratch.begin(), scratch.end());  // popfirst! / append!
        fine.clear();
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
                fine.push_back(t);
            }
        }
        for (const Cell &t : fine) {
            for (const V3 *v : {&t.a, &t.b, &t.c, &t.d}) g->corners.insert(g->corners.end(), v->begin(), v->end());
            g->normal.insert(g->normal.end(), nrm[f].begin(), nrm[f].end());
            g->is_edge.push_back(t.any() ? 1 : 0);
        }
    }
    g->P = (int64_t)g->is_edge.size();
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
