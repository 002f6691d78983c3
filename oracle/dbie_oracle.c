/*
 * dbie_oracle.c -- CPU FP64 restatement of DielectricBIE.jl's boundary-operator path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: it is compiled into
 * oracle/liboracle.so and loaded by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py.  The product library never links or calls it.
 *
 * Every function restates a reference function line by line (file:line given
 * in each comment, paths relative to the reference repository root).  Build
 * with -ffp-contract=off so that the arithmetic order matches the reference's
 * (Julia does not contract a*b+c into an FMA unless asked to).
 *
 * Third-party pieces the reference calls are restated from their published
 * algorithms:
 *   FastGaussQuadrature.gausslegendre (>=1.1.0): closed forms for n<=5,
 *     Newton on the three-term recurrence otherwise;
 *   Julia Base sqrt(::Complex) (Kahan's scaled algorithm) and hypot (the
 *     FMA-corrected branch);
 *   NearestNeighbors.inrange: replaced by a brute-force ball query (<= r).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_FOUR_PI (4.0 * 3.14159265358979311599796346854)

/* ------------------------------------------------------------------ */
/* Gauss-Legendre (FastGaussQuadrature.gausslegendre, ascending nodes) */
/* ------------------------------------------------------------------ */
static void legendre_eval(int n, double x, double *p, double *dp) {
    double p0 = 1.0, p1 = x;
    for (int k = 2; k <= n; ++k) {
        double pk = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = pk;
    }
    *p = p1;
    *dp = n * (x * p1 - p0) / (x * x - 1.0);
}

int orc_gauss_legendre(int n, double *x, double *w) {
    if (n < 1) return -1;
    if (n == 1) { x[0] = 0.0; w[0] = 2.0; return 0; }
    if (n == 2) {
        x[0] = -1.0 / sqrt(3.0); x[1] = 1.0 / sqrt(3.0);
        w[0] = 1.0; w[1] = 1.0; return 0;
    }
    if (n == 3) {
        x[0] = -sqrt(3.0 / 5.0); x[1] = 0.0; x[2] = sqrt(3.0 / 5.0);
        w[0] = 5.0 / 9.0; w[1] = 8.0 / 9.0; w[2] = 5.0 / 9.0; return 0;
    }
    if (n == 4) {
        double a = 2.0 / 7.0 * sqrt(6.0 / 5.0);
        x[0] = -sqrt(3.0 / 7.0 + a); x[1] = -sqrt(3.0 / 7.0 - a);
        x[2] = sqrt(3.0 / 7.0 - a);  x[3] = sqrt(3.0 / 7.0 + a);
        w[0] = (18.0 - sqrt(30.0)) / 36.0; w[1] = (18.0 + sqrt(30.0)) / 36.0;
        w[2] = (18.0 + sqrt(30.0)) / 36.0; w[3] = (18.0 - sqrt(30.0)) / 36.0;
        return 0;
    }
    if (n == 5) {
        double b = 2.0 * sqrt(10.0 / 7.0);
        x[0] = -sqrt(5.0 + b) / 3.0; x[1] = -sqrt(5.0 - b) / 3.0; x[2] = 0.0;
        x[3] = sqrt(5.0 - b) / 3.0;  x[4] = sqrt(5.0 + b) / 3.0;
        w[0] = (322.0 - 13.0 * sqrt(70.0)) / 900.0; w[1] = (322.0 + 13.0 * sqrt(70.0)) / 900.0;
        w[2] = 128.0 / 225.0;
        w[3] = (322.0 + 13.0 * sqrt(70.0)) / 900.0; w[4] = (322.0 - 13.0 * sqrt(70.0)) / 900.0;
        return 0;
    }
    /* Newton on P_n from Chebyshev-like guesses; positive half, mirrored. */
    int m = (n + 1) / 2;
    for (int i = 0; i < m; ++i) {
        double z = cos(M_PI * (i + 0.75) / (n + 0.5));
        double p, dp;
        for (int it = 0; it < 100; ++it) {
            legendre_eval(n, z, &p, &dp);
            double dz = p / dp;
            z -= dz;
            if (fabs(dz) < 1e-16) break;
        }
        legendre_eval(n, z, &p, &dp);
        z -= p / dp;
        legendre_eval(n, z, &p, &dp);
        double wi = 2.0 / ((1.0 - z * z) * dp * dp);
        x[n - 1 - i] = z; w[n - 1 - i] = wi;
        x[i] = -z;        w[i] = wi;
    }
    if (n % 2 == 1) x[n / 2] = 0.0;
    return 0;
}

/* ------------------------------------------------------------------ */
/* Julia Base scalar helpers                                          */
/* ------------------------------------------------------------------ */
/* hypot, FMA-corrected branch of Julia's Base.Math.hypot */
static double jl_hypot(double x, double y) {
    if (isinf(x) || isinf(y)) return INFINITY;
    double ax = fabs(x), ay = fabs(y);
    if (ay > ax) { double t = ax; ax = ay; ay = t; }
    if (ay <= ax * sqrt(DBL_EPSILON / 2.0)) return ax;
    double scale = DBL_EPSILON * sqrt(DBL_MIN);
    if (ax > sqrt(DBL_MAX / 2.0)) {
        ax = ax * scale; ay = ay * scale; scale = 1.0 / scale;
    } else if (ay < sqrt(DBL_MIN)) {
        ax = ax / scale; ay = ay / scale;
    } else {
        scale = 1.0;
    }
    double h = sqrt(fma(ax, ax, ay * ay));
    double hsq = h * h, axsq = ax * ax;
    h -= (fma(-ay, ay, hsq - axsq) + fma(h, h, -hsq) - fma(ax, ax, -axsq)) / (2.0 * h);
    return h * scale;
}

/* sqrt(::Complex{Float64}) of Julia Base (Kahan), normal-range branch plus scaling */
static void jl_csqrt(double x, double y, double *re, double *im) {
    if (x == 0.0 && y == 0.0) { *re = 0.0; *im = y; return; }
    int k = 0;
    double rho = x * x + y * y;
    if (!isfinite(rho) && (isinf(x) || isinf(y))) {
        rho = INFINITY;
    } else if (isinf(rho) || (rho == 0.0 && (x != 0.0 || y != 0.0)) ||
               rho < (4.9406564584124654e-324 / (2.0 * DBL_EPSILON * DBL_EPSILON))) {
        double mm = fmax(fabs(x), fabs(y));
        k = (mm == 0.0) ? 0 : ilogb(mm);
        double xk = ldexp(x, -k), yk = ldexp(y, -k);
        rho = xk * xk + yk * yk;
    }
    if (isfinite(x)) rho = ldexp(fabs(x), -k) + sqrt(rho);
    if (k % 2 != 0) {
        k = (k - 1) / 2;
    } else {
        k = k / 2 - 1;
        rho += rho;
    }
    rho = ldexp(sqrt(rho), k);
    double xi = rho, eta = y;
    if (rho != 0.0) {
        if (isfinite(eta)) eta = (eta / rho) / 2.0;
        if (x < 0.0) { xi = fabs(eta); eta = copysign(rho, y); }
    }
    *re = xi; *im = eta;
}

static inline double norm3(double a, double b, double c) { return sqrt(a * a + b * b + c * c); }

/* laplace3d_grad (src/kernel/laplace3d.jl:9-15): n.(t-s)/r^3/4pi */
static inline double k_grad(const double *s, const double *t, const double *n) {
    double d0 = s[0] - t[0], d1 = s[1] - t[1], d2 = s[2] - t[2];
    double r2 = d0 * d0 + d1 * d1 + d2 * d2;
    double r = sqrt(r2);
    double ir = 1.0 / r;
    double ir3 = ir * ir * ir;
    double v0 = ir3 * (t[0] - s[0]), v1 = ir3 * (t[1] - s[1]), v2 = ir3 * (t[2] - s[2]);
    return (n[0] * v0 + n[1] * v1 + n[2] * v2) / ORC_FOUR_PI;
}
/* laplace3d_pot (src/kernel/laplace3d.jl:1-7) */
static inline double k_pot(const double *s, const double *t) {
    double d0 = s[0] - t[0], d1 = s[1] - t[1], d2 = s[2] - t[2];
    double r2 = d0 * d0 + d1 * d1 + d2 * d2;
    double r = sqrt(r2);
    return (1.0 / r) / ORC_FOUR_PI;
}

double orc_laplace3d_grad(const double *s, const double *t, const double *n) { return k_grad(s, t, n); }
double orc_laplace3d_pot(const double *s, const double *t) { return k_pot(s, t); }

/* ------------------------------------------------------------------ */
/* best_grid (src/utils/best_grid.jl:8-126)                           */
/* ------------------------------------------------------------------ */
int orc_best_grid(double W, double H, double alpha, int *m_out, int *n_out) {
    const int64_t max_den = 1000000, max_iter = 10000;
    const double frac_tol = 1e-12;
    if (!(W > 0 && H > 0)) return -1;
    if (alpha < 1) return -2;
    int swapped = 0;
    if (W < H) { double t = W; W = H; H = t; swapped = 1; }
    double r = W / H;
    int64_t best_m = 0, best_n = 0, best_N = INT64_MAX;
    double best_rho = INFINITY;
#define CONSIDER(mm, nn) do { int64_t m_ = (mm), n_ = (nn);                         \
        if (m_ > 0 && n_ > 0 && m_ <= max_den && n_ <= max_den) {                   \
            double ratio = (W * (double)n_) / (H * (double)m_);                     \
            double rho_ = ratio >= 1 ? ratio : 1 / ratio;                           \
            if (rho_ <= alpha) { int64_t NN = m_ * n_;                              \
                if (NN < best_N || (NN == best_N && rho_ < best_rho)) {             \
                    best_m = m_; best_n = n_; best_N = NN; best_rho = rho_; } } } } while (0)
    int64_t pmm = 0, pm = 1, qmm = 1, qm = 0;
    double x = r;
    for (int64_t it = 0; it < max_iter; ++it) {
        int64_t a = (int64_t)floor(x);
        int64_t p = a * pm + pmm, q = a * qm + qmm;
        CONSIDER(p, q);
        if (a > 1) {
            int64_t t_max = a - 1;
            if (pm > 0) { int64_t v = (max_den - pmm) / pm; if (v < t_max) t_max = v; }
            if (qm > 0) { int64_t v = (max_den - qmm) / qm; if (v < t_max) t_max = v; }
            for (int64_t t = 1; t <= t_max; ++t) CONSIDER(t * pm + pmm, t * qm + qmm);
        }
        double frac = x - (double)a;
        if (fabs(frac) <= frac_tol) break;
        pmm = pm; pm = p; qmm = qm; qm = q;
        x = 1 / frac;
    }
#undef CONSIDER
    if (best_m == 0) return -3;
    if (swapped) { int64_t t = best_m; best_m = best_n; best_n = t; }
    *m_out = (int)best_m; *n_out = (int)best_n;
    return 0;
}

/* ------------------------------------------------------------------ */
/* Interface (DielectricInterface of FlatPanels)                      */
/* ------------------------------------------------------------------ */
typedef struct {
    int P, N, nq;
    double *corners;  /* P x 4 x 3 */
    double *normal;   /* P x 3 */
    int *is_edge;     /* P */
    double *eps_in, *eps_out; /* P */
    double gl_x[16], gl_w[16], bary[16];
    double *pt;       /* N x 3 (AoS) */
    double *wt;       /* N */
    int *pidx;        /* N: owning panel */
} orc_iface;

/* gl_barycentric_weights (src/utils/barycentric.jl:6-17) */
static void bary_weights(int n, const double *x, const double *w, double *lam) {
    double mx = 0.0;
    for (int i = 0; i < n; ++i) {
        double s = ((n - (i + 1)) % 2 == 0) ? 1.0 : -1.0;
        lam[i] = s * sqrt((1 - x[i] * x[i]) * w[i] / 2);
        if (fabs(lam[i]) > mx) mx = fabs(lam[i]);
    }
    for (int i = 0; i < n; ++i) lam[i] /= mx;
}

/* barycentric_row! (src/utils/barycentric.jl:19-42) with Julia isapprox snapping */
static void bary_row(int n, const double *x, const double *lam, double xq, double *r) {
    const double rtol = sqrt(DBL_EPSILON);
    for (int i = 0; i < n; ++i) {
        double a = fabs(xq), b = fabs(x[i]);
        if (xq == x[i] || fabs(xq - x[i]) <= rtol * (a > b ? a : b)) {
            for (int k = 0; k < n; ++k) r[k] = 0.0;
            r[i] = 1.0;
            return;
        }
    }
    double denom = 0.0;
    for (int i = 0; i < n; ++i) { double t = lam[i] / (xq - x[i]); r[i] = t; denom += t; }
    double inv = 1.0 / denom;
    for (int i = 0; i < n; ++i) r[i] *= inv;
}

/* rect_panel3d_discretize node/weight rules (src/shape/box3d.jl:22-64) */
static void panel_nodes(const orc_iface *f, int p, double *pt, double *wt) {
    const double *A = f->corners + 12 * p, *B = A + 3, *C = A + 6, *D = A + 9;
    int nq = f->nq;
    double Lab = norm3(B[0] - A[0], B[1] - A[1], B[2] - A[2]);
    double Lbc = norm3(C[0] - B[0], C[1] - B[1], C[2] - B[2]);
    double cc[3];
    for (int k = 0; k < 3; ++k) cc[k] = (A[k] + B[k] + C[k] + D[k]) / 4;
    int idx = 0;
    for (int i = 0; i < nq; ++i)
        for (int j = 0; j < nq; ++j, ++idx) {
            for (int k = 0; k < 3; ++k)
                pt[3 * idx + k] = cc[k] + (B[k] - A[k]) * (f->gl_x[i] / 2) + (D[k] - A[k]) * (f->gl_x[j] / 2);
            wt[idx] = f->gl_w[i] * f->gl_w[j] * Lab * Lbc / 4;
        }
}

static orc_iface *iface_alloc(int P, int nq) {
    orc_iface *f = (orc_iface *)calloc(1, sizeof(orc_iface));
    f->P = P; f->nq = nq; f->N = P * nq * nq;
    f->corners = (double *)malloc(sizeof(double) * 12 * (size_t)P);
    f->normal = (double *)malloc(sizeof(double) * 3 * (size_t)P);
    f->is_edge = (int *)malloc(sizeof(int) * (size_t)P);
    f->eps_in = (double *)malloc(sizeof(double) * (size_t)P);
    f->eps_out = (double *)malloc(sizeof(double) * (size_t)P);
    f->pt = (double *)malloc(sizeof(double) * 3 * (size_t)f->N);
    f->wt = (double *)malloc(sizeof(double) * (size_t)f->N);
    f->pidx = (int *)malloc(sizeof(int) * (size_t)f->N);
    orc_gauss_legendre(nq, f->gl_x, f->gl_w);
    bary_weights(nq, f->gl_x, f->gl_w, f->bary);
    return f;
}

static void iface_finish(orc_iface *f) {
    int npp = f->nq * f->nq;
    for (int p = 0; p < f->P; ++p) {
        panel_nodes(f, p, f->pt + 3 * (size_t)p * npp, f->wt + (size_t)p * npp);
        for (int k = 0; k < npp; ++k) f->pidx[p * npp + k] = p;
    }
}

void orc_iface_free(orc_iface *f) {
    if (!f) return;
    free(f->corners); free(f->normal); free(f->is_edge); free(f->eps_in); free(f->eps_out);
    free(f->pt); free(f->wt); free(f->pidx); free(f);
}

/* DielectricInterface from explicit panels (corners lb,rb,rt,lt) */
orc_iface *orc_iface_from_panels(int P, int nq, const double *corners, const double *normals,
                                 const int *is_edge, const double *eps_in, const double *eps_out) {
    orc_iface *f = iface_alloc(P, nq);
    memcpy(f->corners, corners, sizeof(double) * 12 * (size_t)P);
    memcpy(f->normal, normals, sizeof(double) * 3 * (size_t)P);
    for (int p = 0; p < P; ++p) {
        f->is_edge[p] = is_edge ? is_edge[p] : 0;
        f->eps_in[p] = eps_in[p]; f->eps_out[p] = eps_out[p];
    }
    iface_finish(f);
    return f;
}

/* ---- single_dielectric_box3d (src/shape/box3d_single.jl:1-70, src/shape/box3d.jl:66-133) ---- */
typedef struct {
    double a[3], b[3], c[3], d[3];
    int ca, cb, cc, cd;         /* corner flags */
    int eab, ebc, ecd, eda;     /* edge flags */
} tpanel;

static void divide_tpanel(const tpanel *t, int nx, int ny, tpanel *out) {
    for (int i = 1; i <= nx; ++i)
        for (int j = 1; j <= ny; ++j) {
            double u0 = (double)(i - 1) / nx, u1 = (double)i / nx;
            double v0 = (double)(j - 1) / ny, v1 = (double)j / ny;
            tpanel *o = &out[(i - 1) * ny + (j - 1)];
            for (int k = 0; k < 3; ++k) {
                double ba = t->b[k] - t->a[k], da = t->d[k] - t->a[k];
                o->a[k] = t->a[k] + ba * u0 + da * v0;
                o->b[k] = t->a[k] + ba * u1 + da * v0;
                o->c[k] = t->a[k] + ba * u1 + da * v1;
                o->d[k] = t->a[k] + ba * u0 + da * v1;
            }
            o->ca = (i == 1 && j == 1) ? t->ca : 0;
            o->cb = (i == nx && j == 1) ? t->cb : 0;
            o->cc = (i == nx && j == ny) ? t->cc : 0;
            o->cd = (i == 1 && j == ny) ? t->cd : 0;
            o->eab = (j == 1) ? t->eab : 0;
            o->ebc = (i == nx) ? t->ebc : 0;
            o->ecd = (j == ny) ? t->ecd : 0;
            o->eda = (i == 1) ? t->eda : 0;
        }
}

/* rect_panel3d_adaptive_panels (src/shape/box3d.jl:99-133): best_grid rough split of the root,
   then a FIFO of 2x2 splits of edge/corner panels longer than l_ec.  Appends the fine panels
   (with normal nrm) to the arrays fine and fn (count n, capacity cap).  Returns 0, or -1 when best_grid
   finds no tiling. */
static int rect_adaptive(const tpanel *root, const double nrm[3], double l_ec, double alpha, tpanel **fine,
                         double (**fn)[3], size_t *n, size_t *cap) {
    double Lab = norm3(root->b[0] - root->a[0], root->b[1] - root->a[1], root->b[2] - root->a[2]);
    double Lda = norm3(root->a[0] - root->d[0], root->a[1] - root->d[1], root->a[2] - root->d[2]);
    int nx, ny;
    if (orc_best_grid(Lab, Lda, alpha, &nx, &ny) != 0) return -1;
    /* FIFO queue (popfirst!/append!) */
    size_t qcap = (size_t)nx * ny + 4096, qh = 0, qt = 0;
    tpanel *q = (tpanel *)malloc(sizeof(tpanel) * qcap);
    divide_tpanel(root, nx, ny, q);
    qt = (size_t)nx * ny;
    while (qh < qt) {
        tpanel t = q[qh++];
        int has_ec = t.ca || t.cb || t.cc || t.cd || t.eab || t.ebc || t.ecd || t.eda;
        double L_ab = norm3(t.b[0] - t.a[0], t.b[1] - t.a[1], t.b[2] - t.a[2]);
        double L_da = norm3(t.a[0] - t.d[0], t.a[1] - t.d[1], t.a[2] - t.d[2]);
        double mx = L_ab > L_da ? L_ab : L_da;
        if (has_ec && mx > l_ec) {
            if (qt + 4 > qcap) {
                /* compact then grow */
                memmove(q, q + qh, sizeof(tpanel) * (qt - qh)); qt -= qh; qh = 0;
                if (qt + 4 > qcap) { qcap *= 2; q = (tpanel *)realloc(q, sizeof(tpanel) * qcap); }
            }
            divide_tpanel(&t, 2, 2, q + qt);
            qt += 4;
        } else {
            if (*n == *cap) {
                *cap *= 2;
                *fine = (tpanel *)realloc(*fine, sizeof(tpanel) * *cap);
                *fn = (double (*)[3])realloc(*fn, sizeof(double) * 3 * *cap);
            }
            (*fine)[*n] = t;
            memcpy((*fn)[*n], nrm, sizeof(double) * 3);
            ++*n;
        }
    }
    free(q);
    return 0;
}

/* Returns the panel count for a box, or <0 on error.  Panels are appended in
   face order; each face is meshed by rect_panel3d_adaptive_panels. */
static int box_panels(double Lx, double Ly, double Lz, double l_ec, double alpha, tpanel **out_panels,
                      double (**out_norm)[3]) {
    double V[8][3] = {
        { Lx / 2,  Ly / 2,  Lz / 2}, {-Lx / 2,  Ly / 2,  Lz / 2}, {-Lx / 2, -Ly / 2,  Lz / 2}, { Lx / 2, -Ly / 2,  Lz / 2},
        { Lx / 2,  Ly / 2, -Lz / 2}, {-Lx / 2,  Ly / 2, -Lz / 2}, {-Lx / 2, -Ly / 2, -Lz / 2}, { Lx / 2, -Ly / 2, -Lz / 2}};
    int F[6][4] = {{1, 2, 3, 4}, {5, 8, 7, 6}, {8, 5, 1, 4}, {7, 3, 2, 6}, {6, 2, 1, 5}, {7, 8, 4, 3}};
    double Nrm[6][3] = {{0, 0, 1}, {0, 0, -1}, {1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}};
    size_t cap = 1024, n = 0;
    tpanel *fine = (tpanel *)malloc(sizeof(tpanel) * cap);
    double (*fn)[3] = (double (*)[3])malloc(sizeof(double) * 3 * cap);
    for (int f = 0; f < 6; ++f) {
        tpanel root;
        memcpy(root.a, V[F[f][0] - 1], sizeof(double) * 3);
        memcpy(root.b, V[F[f][1] - 1], sizeof(double) * 3);
        memcpy(root.c, V[F[f][2] - 1], sizeof(double) * 3);
        memcpy(root.d, V[F[f][3] - 1], sizeof(double) * 3);
        root.ca = root.cb = root.cc = root.cd = 1;
        root.eab = root.ebc = root.ecd = root.eda = 1;
        if (rect_adaptive(&root, Nrm[f], l_ec, alpha, &fine, &fn, &n, &cap) != 0) { free(fine); free(fn); return -1; }
    }
    *out_panels = fine; *out_norm = fn;
    return (int)n;
}

/* rect_panel3d_adaptive_panels of one rectangle (corners a, b, c, d = abcd[0..11], normal nrm,
   edge flags ab, bc, cd, da and corner flags a, b, c, d).  Writes up to cap panels (corners
   12 each, is_edge) and returns the panel count (which may exceed cap: call again with a larger
   buffer), or -1 when best_grid finds no tiling. */
int orc_rect_adaptive_panels(const double *abcd, const double *nrm, const int *edge, const int *corner,
                             double l_ec, double alpha, int cap_out, double *corners, int *is_edge) {
    tpanel root;
    memcpy(root.a, abcd, 24); memcpy(root.b, abcd + 3, 24); memcpy(root.c, abcd + 6, 24); memcpy(root.d, abcd + 9, 24);
    root.eab = edge[0]; root.ebc = edge[1]; root.ecd = edge[2]; root.eda = edge[3];
    root.ca = corner[0]; root.cb = corner[1]; root.cc = corner[2]; root.cd = corner[3];
    size_t cap = 256, n = 0;
    tpanel *fine = (tpanel *)malloc(sizeof(tpanel) * cap);
    double (*fn)[3] = (double (*)[3])malloc(sizeof(double) * 3 * cap);
    int rc = rect_adaptive(&root, nrm, l_ec, alpha, &fine, &fn, &n, &cap);
    if (rc == 0) {
        for (size_t p = 0; p < n && (int)p < cap_out; ++p) {
            double *C = corners + 12 * p;
            memcpy(C, fine[p].a, 24); memcpy(C + 3, fine[p].b, 24); memcpy(C + 6, fine[p].c, 24);
            memcpy(C + 9, fine[p].d, 24);
            is_edge[p] = fine[p].eab || fine[p].ebc || fine[p].ecd || fine[p].eda || fine[p].ca || fine[p].cb ||
                         fine[p].cc || fine[p].cd;
        }
    }
    free(fine); free(fn);
    return rc == 0 ? (int)n : -1;
}

orc_iface *orc_iface_box(double Lx, double Ly, double Lz, int nq, double l_ec, double eps_in,
                         double eps_out, double alpha) {
    tpanel *tp; double (*tn)[3];
    int P = box_panels(Lx, Ly, Lz, l_ec, alpha, &tp, &tn);
    if (P <= 0) return NULL;
    orc_iface *f = iface_alloc(P, nq);
    for (int p = 0; p < P; ++p) {
        double *C = f->corners + 12 * p;
        memcpy(C, tp[p].a, 24); memcpy(C + 3, tp[p].b, 24); memcpy(C + 6, tp[p].c, 24); memcpy(C + 9, tp[p].d, 24);
        memcpy(f->normal + 3 * p, tn[p], 24);
        f->is_edge[p] = tp[p].eab || tp[p].ebc || tp[p].ecd || tp[p].eda || tp[p].ca || tp[p].cb || tp[p].cc || tp[p].cd;
        f->eps_in[p] = eps_in; f->eps_out[p] = eps_out;
    }
    free(tp); free(tn);
    iface_finish(f);
    return f;
}

void orc_iface_sizes(const orc_iface *f, int *P, int *N, int *nq) { *P = f->P; *N = f->N; *nq = f->nq; }

/* SoA node export: rows x,y,z,nx,ny,nz,w (7 x N) */
void orc_iface_nodes(const orc_iface *f, double *out) {
    int N = f->N;
    for (int i = 0; i < N; ++i) {
        int p = f->pidx[i];
        out[i] = f->pt[3 * i]; out[N + i] = f->pt[3 * i + 1]; out[2 * N + i] = f->pt[3 * i + 2];
        out[3 * N + i] = f->normal[3 * p]; out[4 * N + i] = f->normal[3 * p + 1]; out[5 * N + i] = f->normal[3 * p + 2];
        out[6 * N + i] = f->wt[i];
    }
}
void orc_iface_panels(const orc_iface *f, double *corners, double *normals, int *is_edge) {
    memcpy(corners, f->corners, sizeof(double) * 12 * (size_t)f->P);
    memcpy(normals, f->normal, sizeof(double) * 3 * (size_t)f->P);
    memcpy(is_edge, f->is_edge, sizeof(int) * (size_t)f->P);
}
void orc_iface_set_eps(orc_iface *f, const double *eps_in, const double *eps_out) {
    for (int p = 0; p < f->P; ++p) { f->eps_in[p] = eps_in[p]; f->eps_out[p] = eps_out[p]; }
}

/* diag t per node (src/solver/dielectric_box3d.jl:5-15, src/solver/multi_rhs.jl:86-96) */
void orc_diag(const orc_iface *f, double *t) {
    int npp = f->nq * f->nq;
    for (int p = 0; p < f->P; ++p) {
        double ei = f->eps_in[p], eo = f->eps_out[p];
        double tp = 0.5 * (eo + ei) / (eo - ei);
        for (int j = 0; j < npp; ++j) t[p * npp + j] = tp;
    }
}

/* ------------------------------------------------------------------ */
/* Dense operators (src/kernel/laplace3d.jl:18-52), column-major N x N */
/* ------------------------------------------------------------------ */
void orc_dense_DT(const orc_iface *f, double *M) {
    int N = f->N;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < N; ++j)
        for (int i = 0; i < N; ++i) {
            if (i == j) { M[(size_t)j * N + i] = 0.0; continue; }
            M[(size_t)j * N + i] = k_grad(f->pt + 3 * j, f->pt + 3 * i, f->normal + 3 * f->pidx[i]) * f->wt[j];
        }
}
void orc_dense_S(const orc_iface *f, double *M) {
    int N = f->N;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < N; ++j)
        for (int i = 0; i < N; ++i)
            M[(size_t)j * N + i] = (i == j) ? 0.0 : k_pot(f->pt + 3 * j, f->pt + 3 * i) * f->wt[j];
}
/* D[j,i] = grad(pt_j, pt_i, n_i) * w_i  (src/kernel/laplace3d.jl:42-52) */
void orc_dense_D(const orc_iface *f, double *M) {
    int N = f->N;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j)
            M[(size_t)i * N + j] = (i == j) ? 0.0 : k_grad(f->pt + 3 * j, f->pt + 3 * i, f->normal + 3 * f->pidx[i]) * f->wt[i];
}
/* lhs_dielectric_box3d (src/solver/dielectric_box3d.jl:1-17) */
void orc_lhs_dense(const orc_iface *f, double *M) {
    orc_dense_DT(f, M);
    int npp = f->nq * f->nq, N = f->N;
    for (int p = 0; p < f->P; ++p) {
        double ei = f->eps_in[p], eo = f->eps_out[p];
        double t = 0.5 * (eo + ei) / (eo - ei);
        for (int j = 0; j < npp; ++j) { size_t ii = (size_t)p * npp + j; M[ii * N + ii] += t; }
    }
}

/* Matrix-free direct sums.  mode 0: DT, 1: S, 2: D.  X, Y column-major N x K. */
void orc_apply_direct(const orc_iface *f, int mode, int K, const double *X, double *Y) {
    int N = f->N;
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < N; ++i) {
        const double *ti = f->pt + 3 * i;
        const double *ni = f->normal + 3 * f->pidx[i];
        for (int k = 0; k < K; ++k) Y[(size_t)k * N + i] = 0.0;
        for (int j = 0; j < N; ++j) {
            if (j == i) continue;
            double g;
            if (mode == 0) g = k_grad(f->pt + 3 * j, ti, ni) * f->wt[j];
            else if (mode == 1) g = k_pot(f->pt + 3 * j, ti) * f->wt[j];
            else g = k_grad(ti, f->pt + 3 * j, f->normal + 3 * f->pidx[j]) * f->wt[j];
            for (int k = 0; k < K; ++k) Y[(size_t)k * N + i] += g * X[(size_t)k * N + j];
        }
    }
}

/* Row subset of the DT direct sum (for sampled parity checks at large N) */
void orc_apply_DT_rows(const orc_iface *f, int nrows, const int *rows, const double *x, double *y) {
    int N = f->N;
#pragma omp parallel for schedule(dynamic, 4)
    for (int r = 0; r < nrows; ++r) {
        int i = rows[r];
        const double *ti = f->pt + 3 * i;
        const double *ni = f->normal + 3 * f->pidx[i];
        double acc = 0.0;
        for (int j = 0; j < N; ++j) {
            if (j == i) continue;
            acc += k_grad(f->pt + 3 * j, ti, ni) * f->wt[j] * x[j];
        }
        y[r] = acc;
    }
}

/* pottrg (src/kernel/laplace3d.jl:69-82): out[t,k] = sum_i w_i pot(x_i, trg_t) X[i,k] */
void orc_pottrg(const orc_iface *f, int M, const double *trg, int K, const double *X, double *out) {
    int N = f->N;
#pragma omp parallel for schedule(dynamic, 16)
    for (int t = 0; t < M; ++t) {
        for (int k = 0; k < K; ++k) out[(size_t)k * M + t] = 0.0;
        for (int i = 0; i < N; ++i) {
            double g = k_pot(f->pt + 3 * i, trg + 3 * t) * f->wt[i];
            for (int k = 0; k < K; ++k) out[(size_t)k * M + t] += g * X[(size_t)k * N + i];
        }
    }
}

/* ------------------------------------------------------------------ */
/* Right-hand sides (src/solver/dielectric_box3d.jl:95-124)           */
/* ------------------------------------------------------------------ */
/* one column per point charge: F[i,k] = -q_k grad(s_k, x_i, n_i) / eps_src */
void orc_rhs_points(const orc_iface *f, int K, const double *src, const double *q, double eps_src, double *F) {
    int N = f->N;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; ++i)
        for (int k = 0; k < K; ++k)
            F[(size_t)k * N + i] = -q[k] * k_grad(src + 3 * k, f->pt + 3 * i, f->normal + 3 * f->pidx[i]) / eps_src;
}
/* volume source: F[i] = -(sum_s w_s rho_s grad(y_s, x_i, n_i)) / eps_src */
void orc_rhs_volume(const orc_iface *f, int S, const double *pos, const double *w, const double *rho,
                    double eps_src, double *F) {
    int N = f->N;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; ++i) {
        double acc = 0.0;
        for (int s = 0; s < S; ++s) acc += w[s] * rho[s] * k_grad(pos + 3 * s, f->pt + 3 * i, f->normal + 3 * f->pidx[i]);
        F[i] = -acc / eps_src;
    }
}
/* SharpScreening via the interface bounds (src/core/sources.jl:283-406) */
int orc_screen_sharp(const orc_iface *f, int S, const double *pos, const double *rho, double *out) {
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int p = 0; p < f->P; ++p)
        for (int c = 0; c < 4; ++c)
            for (int k = 0; k < 3; ++k) {
                double v = f->corners[12 * p + 3 * c + k];
                mn[k] = fmin(mn[k], v); mx[k] = fmax(mx[k], v);
            }
    double ei = f->eps_in[0], eo = f->eps_out[0];
    for (int p = 0; p < f->P; ++p) if (f->eps_in[p] != ei || f->eps_out[p] != eo) return -1;
    double e0 = fabs(mx[0] - mn[0]), e1 = fabs(mx[1] - mn[1]), e2 = fabs(mx[2] - mn[2]);
    double tol = sqrt(DBL_EPSILON) * fmax(fmax(e0, e1), e2);
    for (int s = 0; s < S; ++s) {
        const double *y = pos + 3 * s;
        int in = (mn[0] - tol <= y[0] && y[0] <= mx[0] + tol) && (mn[1] - tol <= y[1] && y[1] <= mx[1] + tol) &&
                 (mn[2] - tol <= y[2] && y[2] <= mx[2] + tol);
        out[s] = rho[s] / (in ? ei : eo);
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* Near-pair list (src/kernel/laplace3d_near.jl:240-392)              */
/* ------------------------------------------------------------------ */
/* bernstein_rho_from_pole / bernstein_rho_panel (src/utils/bernstein.jl:12-62) */
static double rho_from_pole(double zr, double zi) {
    double ar = zr * zr - zi * zi - 1.0;   /* z*z - 1 */
    double ai = zr * zi + zi * zr;
    double sr, si;
    jl_csqrt(ar, ai, &sr, &si);
    double a1 = jl_hypot(zr + sr, zi + si), a2 = jl_hypot(zr - sr, zi - si);
    return a1 > a2 ? a1 : a2;
}
double orc_bernstein_rho_panel(const orc_iface *f, int p, const double *trg) {
    const double *A = f->corners + 12 * p, *B = A + 3, *C = A + 6, *D = A + 9;
    const double *n = f->normal + 3 * p;
    double cc[3], u1[3], u2[3], e1[3], e2[3], r[3];
    for (int k = 0; k < 3; ++k) { cc[k] = (A[k] + B[k] + C[k] + D[k]) / 4; u1[k] = B[k] - A[k]; u2[k] = D[k] - A[k]; }
    double len1 = norm3(u1[0], u1[1], u1[2]), len2 = norm3(u2[0], u2[1], u2[2]);
    double L1 = len1 / 2, L2 = len2 / 2;
    for (int k = 0; k < 3; ++k) { e1[k] = u1[k] / len1; e2[k] = u2[k] / len2; r[k] = trg[k] - cc[k]; }
    double xt = r[0] * e1[0] + r[1] * e1[1] + r[2] * e1[2];
    double yt = r[0] * e2[0] + r[1] * e2[1] + r[2] * e2[2];
    double zt = r[0] * n[0] + r[1] * n[1] + r[2] * n[2];
    double cy = yt < -L2 ? -L2 : (yt > L2 ? L2 : yt);
    double dy = yt - cy;
    double rho1 = rho_from_pole(xt / L1, sqrt(dy * dy + zt * zt) / L1);
    double cx = xt < -L1 ? -L1 : (xt > L1 ? L1 : xt);
    double dx = xt - cx;
    double rho2 = rho_from_pole(yt / L2, sqrt(dx * dx + zt * zt) / L2);
    return rho1 < rho2 ? rho1 : rho2;
}

typedef struct {
    int n_up, n_adp;
    int *up_i, *up_j, *up_p;   /* source panel, target panel, p_up (0-based panels) */
    int *adp_i, *adp_j;
} orc_nlist;

static int cmp_pair(const void *a, const void *b) {
    const int *x = (const int *)a, *y = (const int *)b;
    if (x[0] != y[0]) return x[0] < y[0] ? -1 : 1;
    if (x[1] != y[1]) return x[1] < y[1] ? -1 : 1;
    return 0;
}

orc_nlist *orc_neighbor_list(const orc_iface *f, int max_order, double atol, int correct_edges) {
    int P = f->P, N = f->N;
    double *centers = (double *)malloc(sizeof(double) * 3 * P), *lengths = (double *)malloc(sizeof(double) * P);
    double *offs = (double *)malloc(sizeof(double) * P);
    for (int p = 0; p < P; ++p) {
        const double *c = f->corners + 12 * p;
        for (int k = 0; k < 3; ++k) centers[3 * p + k] = (c[k] + c[3 + k] + c[6 + k] + c[9 + k]) / 4;
        double l12 = norm3(c[0] - c[3], c[1] - c[4], c[2] - c[5]);
        double l23 = norm3(c[3] - c[6], c[4] - c[7], c[5] - c[8]);
        lengths[p] = l12 > l23 ? l12 : l23;
        const double *n = f->normal + 3 * p;
        offs[p] = n[0] * c[0] + n[1] * c[1] + n[2] * c[2];
    }
    const double same_tol = sqrt(DBL_EPSILON);
    double L_max = 1.0;
    for (int p = 0; p < P; ++p) if (lengths[p] > L_max) L_max = lengths[p];
    double tol_corner = sqrt(DBL_EPSILON) * L_max;
    double log_atol = log(atol);

    /* per-source-panel output buffers: triples (i, j, p_up) and pairs (i, j) */
    int **ups = (int **)calloc(P, sizeof(int *)), *nups = (int *)calloc(P, sizeof(int));
    int **adps = (int **)calloc(P, sizeof(int *)), *nadps = (int *)calloc(P, sizeof(int));

#pragma omp parallel
    {
        double *rho_min = (double *)malloc(sizeof(double) * P);
        char *cand = (char *)malloc(P), *touch = (char *)malloc(P);
        int *list = (int *)malloc(sizeof(int) * P);
#pragma omp for schedule(dynamic, 4)
        for (int i = 0; i < P; ++i) {
            if (!correct_edges && f->is_edge[i]) continue;
            double l_i = lengths[i];
            int nq_i = f->nq;
            double rho_star = pow(atol, -1.0 / (2 * nq_i));
            double L_half = l_i / 2;
            double As = (rho_star + 1.0 / rho_star) / 2, Bs = (rho_star - 1.0 / rho_star) / 2;
            double r_i = 105.0 / 100.0 * L_half * sqrt(As * As + (1.0 + Bs) * (1.0 + Bs) + Bs * Bs);
            double r2 = r_i * r_i;
            memset(cand, 0, P); memset(touch, 0, P);
            for (int p = 0; p < P; ++p) rho_min[p] = INFINITY;
            int nl = 0;
            const double *ci = centers + 3 * i;
            for (int t = 0; t < N; ++t) {
                int j = f->pidx[t];
                if (j == i) continue;
                const double *x = f->pt + 3 * t;
                double d0 = x[0] - ci[0], d1 = x[1] - ci[1], d2 = x[2] - ci[2];
                if (d0 * d0 + d1 * d1 + d2 * d2 <= r2) {
                    if (!cand[j]) { cand[j] = 1; list[nl++] = j; }
                    double rho = orc_bernstein_rho_panel(f, i, x);
                    if (rho < rho_min[j]) rho_min[j] = rho;
                }
            }
            if (correct_edges) {
                double tc2 = tol_corner * tol_corner;
                for (int c = 0; c < 4; ++c) {
                    const double *cp = f->corners + 12 * i + 3 * c;
                    for (int j = 0; j < P; ++j) {
                        if (j == i || touch[j]) continue;
                        for (int c2 = 0; c2 < 4; ++c2) {
                            const double *q = f->corners + 12 * j + 3 * c2;
                            double d0 = q[0] - cp[0], d1 = q[1] - cp[1], d2 = q[2] - cp[2];
                            if (d0 * d0 + d1 * d1 + d2 * d2 <= tc2) { touch[j] = 1; break; }
                        }
                    }
                }
                for (int j = 0; j < P; ++j) if (touch[j] && !cand[j]) { cand[j] = 1; list[nl++] = j; }
            }
            int *u = (int *)malloc(sizeof(int) * 3 * (nl + 1)), nu = 0;
            int *a = (int *)malloc(sizeof(int) * 2 * (nl + 1)), na = 0;
            for (int q = 0; q < nl; ++q) {
                int j = list[q];
                if (!correct_edges && f->is_edge[j]) continue;
                const double *ni = f->normal + 3 * i, *nj = f->normal + 3 * j;
                double dn = ni[0] * nj[0] + ni[1] * nj[1] + ni[2] * nj[2];
                if (dn > 1 - same_tol && fabs(offs[i] - offs[j]) <= same_tol * (l_i > 1.0 ? l_i : 1.0)) continue;
                if (correct_edges && touch[j]) { a[2 * na] = i; a[2 * na + 1] = j; ++na; continue; }
                double rm = rho_min[j];
                if (!(rm <= rho_star)) continue;
                int p_up;
                if (rm > 1.0) p_up = (int)ceil(-log_atol / (2 * log(rm)));
                else p_up = max_order;
                if (p_up < nq_i) p_up = nq_i;
                if (p_up > max_order) p_up = max_order;
                u[3 * nu] = i; u[3 * nu + 1] = j; u[3 * nu + 2] = p_up; ++nu;
            }
            ups[i] = u; nups[i] = nu; adps[i] = a; nadps[i] = na;
        }
        free(rho_min); free(cand); free(touch); free(list);
    }
    orc_nlist *nlst = (orc_nlist *)calloc(1, sizeof(orc_nlist));
    int tu = 0, ta = 0;
    for (int i = 0; i < P; ++i) { tu += nups[i]; ta += nadps[i]; }
    int *U = (int *)malloc(sizeof(int) * 3 * (tu + 1)), *A = (int *)malloc(sizeof(int) * 2 * (ta + 1));
    int cu = 0, ca = 0;
    for (int i = 0; i < P; ++i) {
        if (nups[i]) memcpy(U + 3 * cu, ups[i], sizeof(int) * 3 * nups[i]);
        if (nadps[i]) memcpy(A + 2 * ca, adps[i], sizeof(int) * 2 * nadps[i]);
        cu += nups[i]; ca += nadps[i];
        free(ups[i]); free(adps[i]);
    }
    qsort(U, tu, sizeof(int) * 3, cmp_pair);
    qsort(A, ta, sizeof(int) * 2, cmp_pair);
    nlst->n_up = tu; nlst->n_adp = ta;
    nlst->up_i = (int *)malloc(sizeof(int) * (tu + 1)); nlst->up_j = (int *)malloc(sizeof(int) * (tu + 1));
    nlst->up_p = (int *)malloc(sizeof(int) * (tu + 1));
    nlst->adp_i = (int *)malloc(sizeof(int) * (ta + 1)); nlst->adp_j = (int *)malloc(sizeof(int) * (ta + 1));
    for (int k = 0; k < tu; ++k) { nlst->up_i[k] = U[3 * k]; nlst->up_j[k] = U[3 * k + 1]; nlst->up_p[k] = U[3 * k + 2]; }
    for (int k = 0; k < ta; ++k) { nlst->adp_i[k] = A[2 * k]; nlst->adp_j[k] = A[2 * k + 1]; }
    free(U); free(A); free(ups); free(nups); free(adps); free(nadps);
    free(centers); free(lengths); free(offs);
    return nlst;
}
void orc_nlist_sizes(const orc_nlist *l, int *n_up, int *n_adp) { *n_up = l->n_up; *n_adp = l->n_adp; }
void orc_nlist_get(const orc_nlist *l, int *up_i, int *up_j, int *up_p, int *adp_i, int *adp_j) {
    memcpy(up_i, l->up_i, sizeof(int) * l->n_up); memcpy(up_j, l->up_j, sizeof(int) * l->n_up);
    memcpy(up_p, l->up_p, sizeof(int) * l->n_up);
    memcpy(adp_i, l->adp_i, sizeof(int) * l->n_adp); memcpy(adp_j, l->adp_j, sizeof(int) * l->n_adp);
}
void orc_nlist_free(orc_nlist *l) {
    if (!l) return;
    free(l->up_i); free(l->up_j); free(l->up_p); free(l->adp_i); free(l->adp_j); free(l);
}

/* ------------------------------------------------------------------ */
/* Correction blocks (src/kernel/laplace3d_near_upsampling.jl:1-276,   */
/* src/kernel/laplace3d_near_adaptive.jl:119-224,308-387)             */
/* ------------------------------------------------------------------ */
typedef struct { double atol, rtol; int n_GL, max_depth, mode; } orc_adaptive_cfg;

/* Kernel of the correction blocks between a point y of the source panel and a target x:
   mode 0 (:DT) laplace3d_grad(y, x, n_trg), mode 1 (:D) laplace3d_grad(x, y, n_src)
   (src/kernel/laplace3d_near_upsampling.jl:183-197, src/kernel/laplace3d_near_adaptive.jl:169-175) */
static inline double corr_k(int mode, const double *y, const double *x, const double *tn, const double *sn) {
    return mode == 1 ? k_grad(x, y, sn) : k_grad(y, x, tn);
}

/* panel geometry helpers shared by the upsampled and adaptive rules */
typedef struct { double cc[3], bma[3], dma[3], scale; } pgeom;
static void panel_geom(const orc_iface *f, int p, pgeom *g) {
    const double *A = f->corners + 12 * p, *B = A + 3, *C = A + 6, *D = A + 9;
    for (int k = 0; k < 3; ++k) { g->cc[k] = (A[k] + B[k] + C[k] + D[k]) / 4; g->bma[k] = B[k] - A[k]; g->dma[k] = D[k] - A[k]; }
    double Lx = norm3(g->bma[0], g->bma[1], g->bma[2]), Ly = norm3(g->dma[0], g->dma[1], g->dma[2]);
    g->scale = Lx * Ly / 4;
}

/* laplace3d_DT_panel / laplace3d_D_panel (src/kernel/laplace3d_near.jl:25-45): Kd[t + c*ntrg] */
static void panel_direct(const orc_iface *f, int mode, int src, int trg, double *Kd) {
    int npp = f->nq * f->nq;
    const double *nt = f->normal + 3 * trg, *ns = f->normal + 3 * src;
    for (int c = 0; c < npp; ++c)
        for (int t = 0; t < npp; ++t) {
            int gs = src * npp + c, gt = trg * npp + t;
            Kd[t + c * npp] = corr_k(mode, f->pt + 3 * gs, f->pt + 3 * gt, nt, ns) * f->wt[gs];
        }
}

/* build_source_cache (src/kernel/laplace3d_near_upsampling.jl:8-50) */
static void source_cache(const orc_iface *f, int p, int n_up, double *pos /*n_up^2 x 3*/, double *Mt /*npp x n_up^2, col-major*/) {
    int nq = f->nq, npp = nq * nq, nu2 = n_up * n_up;
    double ns[16], ws[16], Ex[16 * 16], row[16];
    orc_gauss_legendre(n_up, ns, ws);
    for (int iu = 0; iu < n_up; ++iu) {
        bary_row(nq, f->gl_x, f->bary, ns[iu], row);
        for (int m = 0; m < nq; ++m) Ex[iu + m * n_up] = row[m];
    }
    pgeom g; panel_geom(f, p, &g);
    for (int iu = 0; iu < n_up; ++iu) {
        double x = ns[iu] / 2;
        for (int ju = 0; ju < n_up; ++ju) {
            double y = ns[ju] / 2;
            int al = iu * n_up + ju;
            for (int k = 0; k < 3; ++k) pos[3 * al + k] = g.cc[k] + g.bma[k] * x + g.dma[k] * y;
        }
    }
    for (int mx = 0; mx < nq; ++mx)
        for (int my = 0; my < nq; ++my) {
            int m = mx * nq + my;
            for (int iu = 0; iu < n_up; ++iu)
                for (int ju = 0; ju < n_up; ++ju) {
                    int al = iu * n_up + ju;
                    Mt[m + (size_t)al * npp] = g.scale * ws[iu] * ws[ju] * Ex[iu + mx * n_up] * Ex[ju + my * n_up];
                }
        }
    (void)nu2;
}

/* cell moments of one target over a (u,v) cell (walk_adaptive_tree's cell_moments_inplace!) */
static void cell_moments(const orc_iface *f, const pgeom *g, int nGL, const double *nsG, const double *wsG,
                         const double *trg, const double *tn, const double *sn, int mode, double ulo, double uhi,
                         double vlo, double vhi, double *out) {
    int nq = f->nq, npp = nq * nq;
    double rx[16], ry[16];
    for (int m = 0; m < npp; ++m) out[m] = 0.0;
    double hu = (uhi - ulo) / 2, hv = (vhi - vlo) / 2, mu = (uhi + ulo) / 2, mv = (vhi + vlo) / 2;
    double sc = hu * hv * g->scale;
    for (int gi = 0; gi < nGL; ++gi) {
        double u = mu + hu * nsG[gi], wu = wsG[gi];
        bary_row(nq, f->gl_x, f->bary, u, rx);
        for (int gj = 0; gj < nGL; ++gj) {
            double v = mv + hv * nsG[gj], wv = wsG[gj];
            bary_row(nq, f->gl_x, f->bary, v, ry);
            double y[3];
            for (int k = 0; k < 3; ++k) y[k] = g->cc[k] + g->bma[k] * (u / 2) + g->dma[k] * (v / 2);
            double kv = corr_k(mode, y, trg, tn, sn);
            double w = wu * wv * sc * kv;
            int m = 0;
            for (int mx = 0; mx < nq; ++mx)
                for (int my = 0; my < nq; ++my) { out[m] += w * rx[mx] * ry[my]; ++m; }
        }
    }
}

typedef struct { double *leaf; int n, cap; int hit; } leafbuf;
static void leaf_push(leafbuf *L, double a, double b, double c, double d) {
    if (L->n == L->cap) { L->cap = L->cap ? 2 * L->cap : 64; L->leaf = (double *)realloc(L->leaf, sizeof(double) * 4 * L->cap); }
    double *q = L->leaf + 4 * L->n++;
    q[0] = a; q[1] = b; q[2] = c; q[3] = d;
}
/* walk_adaptive_tree's recurse_walk! (src/kernel/laplace3d_near_adaptive.jl:191-220) */
static void walk(const orc_iface *f, const pgeom *g, int nGL, const double *nsG, const double *wsG, const double *trg,
                 const double *tn, const double *sn, const orc_adaptive_cfg *cfg, double ulo, double uhi, double vlo,
                 double vhi, int depth, leafbuf *L) {
    int npp = f->nq * f->nq;
    double par[64], c1[64], c2[64], c3[64], c4[64];
    cell_moments(f, g, nGL, nsG, wsG, trg, tn, sn, cfg->mode, ulo, uhi, vlo, vhi, par);
    double mu = (uhi + ulo) / 2, mv = (vhi + vlo) / 2;
    cell_moments(f, g, nGL, nsG, wsG, trg, tn, sn, cfg->mode, ulo, mu, vlo, mv, c1);
    cell_moments(f, g, nGL, nsG, wsG, trg, tn, sn, cfg->mode, mu, uhi, vlo, mv, c2);
    cell_moments(f, g, nGL, nsG, wsG, trg, tn, sn, cfg->mode, ulo, mu, mv, vhi, c3);
    cell_moments(f, g, nGL, nsG, wsG, trg, tn, sn, cfg->mode, mu, uhi, mv, vhi, c4);
    double err = 0.0, pmax = 0.0;
    for (int m = 0; m < npp; ++m) {
        double ch = c1[m] + c2[m] + c3[m] + c4[m];
        double e = fabs(par[m] - ch);
        if (m == 0 || e > err) err = e;
        double a = fabs(par[m]);
        if (m == 0 || a > pmax) pmax = a;
    }
    double tl = cfg->rtol * pmax;
    if (cfg->atol > tl) tl = cfg->atol;
    if (err < tl || depth == cfg->max_depth) {
        leaf_push(L, ulo, mu, vlo, mv);
        leaf_push(L, mu, uhi, vlo, mv);
        leaf_push(L, ulo, mu, mv, vhi);
        leaf_push(L, mu, uhi, mv, vhi);
        if (depth == cfg->max_depth && err >= tl) L->hit = 1;
        return;
    }
    walk(f, g, nGL, nsG, wsG, trg, tn, sn, cfg, ulo, mu, vlo, mv, depth + 1, L);
    walk(f, g, nGL, nsG, wsG, trg, tn, sn, cfg, mu, uhi, vlo, mv, depth + 1, L);
    walk(f, g, nGL, nsG, wsG, trg, tn, sn, cfg, ulo, mu, mv, vhi, depth + 1, L);
    walk(f, g, nGL, nsG, wsG, trg, tn, sn, cfg, mu, uhi, mv, vhi, depth + 1, L);
}

/* Adaptive block for one touching pair: Kb[t + m*npp] (adaptive_block_moments_inplace!) */
static int adaptive_block(const orc_iface *f, int src, int trg, const orc_adaptive_cfg *cfg, double *Kb, int *n_leaves) {
    int nq = f->nq, npp = nq * nq;
    int nGL = cfg->n_GL == 0 ? nq : cfg->n_GL;
    double nsG[32], wsG[32];
    orc_gauss_legendre(nGL, nsG, wsG);
    pgeom g; panel_geom(f, src, &g);
    const double *tn = f->normal + 3 * trg, *sn = f->normal + 3 * src;
    /* representative target: node of trg closest to first corner of src (first argmin) */
    const double *c1 = f->corners + 12 * src;
    int rep = 0; double best = INFINITY;
    for (int t = 0; t < npp; ++t) {
        const double *x = f->pt + 3 * (size_t)(trg * npp + t);
        double d = norm3(x[0] - c1[0], x[1] - c1[1], x[2] - c1[2]);
        if (d < best) { best = d; rep = t; }
    }
    leafbuf L = {0};
    walk(f, &g, nGL, nsG, wsG, f->pt + 3 * (size_t)(trg * npp + rep), tn, sn, cfg, -1.0, 1.0, -1.0, 1.0, 0, &L);
    for (int m = 0; m < npp * npp; ++m) Kb[m] = 0.0;
    double rx[16], ry[16], kv[64];
    for (int l = 0; l < L.n; ++l) {
        double ulo = L.leaf[4 * l], uhi = L.leaf[4 * l + 1], vlo = L.leaf[4 * l + 2], vhi = L.leaf[4 * l + 3];
        double hu = (uhi - ulo) / 2, hv = (vhi - vlo) / 2, mu = (uhi + ulo) / 2, mv = (vhi + vlo) / 2;
        double sc = hu * hv * g.scale;
        for (int gi = 0; gi < nGL; ++gi) {
            double u = mu + hu * nsG[gi], wu = wsG[gi];
            bary_row(nq, f->gl_x, f->bary, u, rx);
            for (int gj = 0; gj < nGL; ++gj) {
                double v = mv + hv * nsG[gj], wv = wsG[gj];
                bary_row(nq, f->gl_x, f->bary, v, ry);
                double y[3];
                for (int k = 0; k < 3; ++k) y[k] = g.cc[k] + g.bma[k] * (u / 2) + g.dma[k] * (v / 2);
                double wuv = wu * wv * sc;
                for (int t = 0; t < npp; ++t) kv[t] = corr_k(cfg->mode, y, f->pt + 3 * (size_t)(trg * npp + t), tn, sn);
                for (int mx = 0; mx < nq; ++mx)
                    for (int my = 0; my < nq; ++my) {
                        int m = mx * nq + my;
                        double wm = wuv * rx[mx] * ry[my];
                        for (int t = 0; t < npp; ++t) Kb[t + m * npp] += wm * kv[t];
                    }
            }
        }
    }
    *n_leaves = L.n;
    int hit = L.hit;
    free(L.leaf);
    return hit;
}

/* Correction blocks, one dense npp x npp block (column-major, rows = target
   nodes of panel j, cols = source nodes of panel i) per pair.  Upsampled pairs
   use the per-source-panel maximum n_up (laplace3d_near_upsampling.jl:165-183).
   out_blocks is (n_up + n_adp) x npp x npp, pairs in nlist order (upsampled
   first, then adaptive).  Returns the number of adaptive pairs that hit max_depth.
   mode 0: laplace3d_DT_corrections, mode 1: laplace3d_D_corrections
   (src/kernel/laplace3d_near_upsampling.jl:272-282). */
int orc_corrections_mode(const orc_iface *f, const orc_nlist *nl, int mode, double adp_atol, double adp_rtol,
                         int adp_nGL, int adp_max_depth, double *out_blocks) {
    int P = f->P, nq = f->nq, npp = nq * nq;
    int *nmax = (int *)calloc(P, sizeof(int));
    for (int k = 0; k < nl->n_up; ++k) if (nl->up_p[k] > nmax[nl->up_i[k]]) nmax[nl->up_i[k]] = nl->up_p[k];
    /* pairs of one source panel are contiguous (sorted) */
    int *start = (int *)malloc(sizeof(int) * (P + 1));
    for (int p = 0; p <= P; ++p) start[p] = 0;
    for (int k = 0; k < nl->n_up; ++k) start[nl->up_i[k] + 1]++;
    for (int p = 0; p < P; ++p) start[p + 1] += start[p];
#pragma omp parallel
    {
        double *pos = (double *)malloc(sizeof(double) * 3 * 256), *Mt = (double *)malloc(sizeof(double) * 64 * 256);
        double *Kmat = (double *)malloc(sizeof(double) * 256 * 64), *Kd = (double *)malloc(sizeof(double) * 64 * 64);
#pragma omp for schedule(dynamic, 2)
        for (int i = 0; i < P; ++i) {
            if (start[i] == start[i + 1]) continue;
            int n_up = nmax[i], nu2 = n_up * n_up;
            source_cache(f, i, n_up, pos, Mt);
            for (int k = start[i]; k < start[i + 1]; ++k) {
                int j = nl->up_j[k];
                const double *tn = f->normal + 3 * j, *sn = f->normal + 3 * i;
                for (int t = 0; t < npp; ++t)
                    for (int al = 0; al < nu2; ++al)
                        Kmat[al + t * nu2] = corr_k(mode, pos + 3 * al, f->pt + 3 * (size_t)(j * npp + t), tn, sn);
                panel_direct(f, mode, i, j, Kd);
                double *blk = out_blocks + (size_t)k * npp * npp;
                for (int c = 0; c < npp; ++c)
                    for (int t = 0; t < npp; ++t) {
                        double s = 0.0;
                        for (int al = 0; al < nu2; ++al) s += Mt[c + (size_t)al * npp] * Kmat[al + t * nu2];
                        blk[t + c * npp] = s - Kd[t + c * npp];
                    }
            }
        }
        free(pos); free(Mt); free(Kmat); free(Kd);
    }
    orc_adaptive_cfg cfg = {adp_atol, adp_rtol, adp_nGL, adp_max_depth, mode};
    int hits = 0;
#pragma omp parallel
    {
        double *Kb = (double *)malloc(sizeof(double) * 64 * 64), *Kd = (double *)malloc(sizeof(double) * 64 * 64);
#pragma omp for schedule(dynamic, 1) reduction(+ : hits)
        for (int k = 0; k < nl->n_adp; ++k) {
            int i = nl->adp_i[k], j = nl->adp_j[k], nlv;
            hits += adaptive_block(f, i, j, &cfg, Kb, &nlv);
            panel_direct(f, mode, i, j, Kd);
            double *blk = out_blocks + (size_t)(nl->n_up + k) * npp * npp;
            for (int m = 0; m < npp * npp; ++m) blk[m] = Kb[m] - Kd[m];
        }
        free(Kb); free(Kd);
    }
    free(nmax); free(start);
    return hits;
}

int orc_corrections(const orc_iface *f, const orc_nlist *nl, double adp_atol, double adp_rtol, int adp_nGL,
                    int adp_max_depth, double *out_blocks) {
    return orc_corrections_mode(f, nl, 0, adp_atol, adp_rtol, adp_nGL, adp_max_depth, out_blocks);
}

/* Apply C (as blocks) to X: Y[rows of j, k] += blk * X[cols of i, k] */
void orc_corrections_apply(const orc_iface *f, const orc_nlist *nl, const double *blocks, int K, const double *X, double *Y) {
    int N = f->N, npp = f->nq * f->nq, nb = nl->n_up + nl->n_adp;
    for (int b = 0; b < nb; ++b) {
        int i = b < nl->n_up ? nl->up_i[b] : nl->adp_i[b - nl->n_up];
        int j = b < nl->n_up ? nl->up_j[b] : nl->adp_j[b - nl->n_up];
        const double *blk = blocks + (size_t)b * npp * npp;
        for (int k = 0; k < K; ++k)
            for (int c = 0; c < npp; ++c) {
                double xv = X[(size_t)k * N + (size_t)i * npp + c];
                for (int t = 0; t < npp; ++t) Y[(size_t)k * N + (size_t)j * npp + t] += blk[t + c * npp] * xv;
            }
    }
}

/* GL-weight helper for tests of barycentric interpolation */
void orc_bary_row(int n, const double *x, const double *w, double xq, double *r) {
    double lam[16];
    bary_weights(n, x, w, lam);
    bary_row(n, x, lam, xq, r);
}
double orc_rho_from_pole(double zr, double zi) { return rho_from_pole(zr, zi); }
int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------ */
/* Off-surface targets: laplace3d_pottrg_fmm3d_corrected_hcubature      */
/* (src/kernel/laplace3d_near_hcubature.jl:337-375)                      */
/* ------------------------------------------------------------------ */
/* HCubature.hcubature(f, [-1,-1], [1,1]; atol) (HCubature >= 1.8.0, not in the reference
   tree) is restated from its published algorithm: the two-dimensional Genz-Malik degree-7
   rule with its embedded degree-5 rule (Genz & Malik 1980); the box with the largest error
   estimate is halved along the axis with the largest fourth divided difference
   |f(+-l3) + 12 f(0) - 7 f(+-l2)| (first axis on ties) until the summed error is <= atol.
   Boxes with equal errors are taken oldest first; the integral is re-summed over the final
   boxes in creation order. */
#define GM_MAXV 64
typedef struct { double a[2], b[2], I[GM_MAXV], E; int kdiv; long seq; } gm_box;

typedef struct {
    const double *cc, *bma, *dma, *t;
    double scale;
    const orc_iface *f;
} pot_integrand;

/* _pot_integrand_hcubature (src/kernel/laplace3d_near_hcubature.jl:59-84) */
static void pot_eval(const pot_integrand *pi, double u, double v, double *vals) {
    int nq = pi->f->nq;
    double rx[16], ry[16], p[3];
    bary_row(nq, pi->f->gl_x, pi->f->bary, u, rx);
    bary_row(nq, pi->f->gl_x, pi->f->bary, v, ry);
    for (int k = 0; k < 3; ++k) p[k] = pi->cc[k] + pi->bma[k] * (u / 2) + pi->dma[k] * (v / 2);
    double kv = k_pot(p, pi->t) * pi->scale;
    for (int ii = 0; ii < nq; ++ii)
        for (int jj = 0; jj < nq; ++jj) vals[ii * nq + jj] = kv * rx[ii] * ry[jj];
}

static void gm_rule(const pot_integrand *pi, int m, gm_box *bx) {
    const double l2 = sqrt(9.0 / 70.0), l3 = sqrt(9.0 / 10.0), l4 = sqrt(9.0 / 10.0), l5 = sqrt(9.0 / 19.0);
    const double w1 = 4.0 * (12824 - 9120 * 2 + 400 * 4) / 19683, w2 = 4.0 * 980 / 6561,
                 w3 = 4.0 * (1820 - 400 * 2) / 19683, w4 = 4.0 * 200 / 19683, w5 = 6859.0 / 19683;
    const double v1 = 4.0 * (729 - 950 * 2 + 50 * 4) / 729, v2 = 4.0 * 245 / 486, v3 = 4.0 * (265 - 100 * 2) / 1458,
                 v4 = 4.0 * 25 / 729;
    double c[2], h[2];
    for (int k = 0; k < 2; ++k) { c[k] = 0.5 * (bx->a[k] + bx->b[k]); h[k] = 0.5 * fabs(bx->b[k] - bx->a[k]); }
    double V = h[0] * h[1];
    double f1[GM_MAXV], f2[GM_MAXV] = {0}, f3[GM_MAXV] = {0}, f4[GM_MAXV] = {0}, f5[GM_MAXV] = {0};
    double ya[GM_MAXV], yb[GM_MAXV], s2[GM_MAXV], s3[GM_MAXV];
    pot_eval(pi, c[0], c[1], f1);
    double best = 0.0;
    bx->kdiv = 0;
    for (int ax = 0; ax < 2; ++ax) {
        double d2[2] = {0, 0}, d3[2] = {0, 0};
        d2[ax] = h[ax] * l2;
        d3[ax] = h[ax] * l3;
        pot_eval(pi, c[0] + d2[0], c[1] + d2[1], ya);
        pot_eval(pi, c[0] - d2[0], c[1] - d2[1], yb);
        for (int i = 0; i < m; ++i) s2[i] = ya[i] + yb[i];
        pot_eval(pi, c[0] + d3[0], c[1] + d3[1], ya);
        pot_eval(pi, c[0] - d3[0], c[1] - d3[1], yb);
        for (int i = 0; i < m; ++i) s3[i] = ya[i] + yb[i];
        double dd = 0.0;
        for (int i = 0; i < m; ++i) {
            f2[i] += s2[i];
            f3[i] += s3[i];
            double e = s3[i] + 12 * f1[i] - 7 * s2[i];
            dd += e * e;
        }
        dd = sqrt(dd);
        if (dd > best) { best = dd; bx->kdiv = ax; }
    }
    for (int sa = -1; sa <= 1; sa += 2)
        for (int sb = -1; sb <= 1; sb += 2) {
            pot_eval(pi, c[0] + sa * h[0] * l4, c[1] + sb * h[1] * l4, ya);
            pot_eval(pi, c[0] + sa * h[0] * l5, c[1] + sb * h[1] * l5, yb);
            for (int i = 0; i < m; ++i) { f4[i] += ya[i]; f5[i] += yb[i]; }
        }
    double e2 = 0.0;
    for (int i = 0; i < m; ++i) {
        bx->I[i] = V * (w1 * f1[i] + w2 * f2[i] + w3 * f3[i] + w4 * f4[i] + w5 * f5[i]);
        double lo = V * (v1 * f1[i] + v2 * f2[i] + v3 * f3[i] + v4 * f4[i]);
        e2 += (bx->I[i] - lo) * (bx->I[i] - lo);
    }
    bx->E = sqrt(e2);
}

/* returns 1 if atol was reached, 0 if the evaluation cap stopped it */
static int hcub2(const pot_integrand *pi, int m, double atol, long maxevals, double *I) {
    long cap = 64, n = 1, evals = 17, seq = 1;
    gm_box *bx = (gm_box *)malloc(sizeof(gm_box) * cap);
    bx[0].a[0] = bx[0].a[1] = -1; bx[0].b[0] = bx[0].b[1] = 1; bx[0].seq = 0;
    gm_rule(pi, m, &bx[0]);
    int ok;
    if (bx[0].E <= atol || evals >= maxevals) {
        memcpy(I, bx[0].I, sizeof(double) * m);
        ok = bx[0].E <= atol;
        free(bx);
        return ok;
    }
    double It[GM_MAXV], Et = bx[0].E;
    memcpy(It, bx[0].I, sizeof(double) * m);
    for (;;) {
        long k = 0;
        for (long q = 1; q < n; ++q)
            if (bx[q].E > bx[k].E || (bx[q].E == bx[k].E && bx[q].seq < bx[k].seq)) k = q;
        gm_box old = bx[k];
        if (n + 1 > cap) { cap *= 2; bx = (gm_box *)realloc(bx, sizeof(gm_box) * cap); }
        int d = old.kdiv;
        double mid = old.a[d] + (old.b[d] - old.a[d]) / 2;
        gm_box c1 = old, c2 = old;
        c1.b[d] = mid; c2.a[d] = mid;
        c1.seq = seq++; c2.seq = seq++;
        gm_rule(pi, m, &c1);
        gm_rule(pi, m, &c2);
        evals += 34;
        for (int i = 0; i < m; ++i) It[i] += c1.I[i] + c2.I[i] - old.I[i];
        Et += c1.E + c2.E - old.E;
        bx[k] = bx[n - 1];
        bx[n - 1] = c1;
        bx[n++] = c2;
        if (Et <= atol || evals >= maxevals) break;
    }
    /* re-sum in creation order */
    for (int i = 0; i < m; ++i) I[i] = 0.0;
    for (long s = 0; s < seq; ++s)
        for (long q = 0; q < n; ++q)
            if (bx[q].seq == s) { for (int i = 0; i < m; ++i) I[i] += bx[q].I[i]; break; }
    ok = Et <= atol;
    free(bx);
    return ok;
}

typedef struct {
    int M, N, Nr, nq;
    double *trg;
    orc_iface *ref;
    int *parent, *from_split;
    int *p_ptr, *p_col; double *p_val;     /* prolongation, Nr rows */
    long *c_ptr; int *c_col; double *c_val; /* corrections, M rows */
    long n_pairs, n_capped;
} orc_tmap;

static double panel_maxlen(const double *c) {
    double l1 = norm3(c[0] - c[3], c[1] - c[4], c[2] - c[5]);
    double l2 = norm3(c[3] - c[6], c[4] - c[7], c[5] - c[8]);
    return l1 > l2 ? l1 : l2;
}
static void panel_ctr(const double *c, double *cc) {
    for (int k = 0; k < 3; ++k) cc[k] = (c[k] + c[3 + k] + c[6 + k] + c[9 + k]) / 4;
}
/* inrange(tree, c, r) restated as a brute-force ball query |t - c|^2 <= r^2 */
static int any_in_ball(int M, const double *trg, const double *c, double r) {
    for (int t = 0; t < M; ++t) {
        double dx = trg[3 * t] - c[0], dy = trg[3 * t + 1] - c[1], dz = trg[3 * t + 2] - c[2];
        if (dx * dx + dy * dy + dz * dz <= r * r) return 1;
    }
    return 0;
}

typedef struct { int t, c; double v; } orc_trip;
static int cmp_trip(const void *a, const void *b) {
    const orc_trip *x = (const orc_trip *)a, *y = (const orc_trip *)b;
    if (x->t != y->t) return x->t < y->t ? -1 : 1;
    return (x->c > y->c) - (x->c < y->c);
}

orc_tmap *orc_tmap_build(const orc_iface *f, int M, const double *trg, double atol, double range_factor) {
    int nq = f->nq, npp = nq * nq;
    orc_tmap *tm = (orc_tmap *)calloc(1, sizeof(orc_tmap));
    tm->M = M; tm->N = f->N; tm->nq = nq;
    tm->trg = (double *)malloc(sizeof(double) * 3 * (size_t)(M > 0 ? M : 1));
    memcpy(tm->trg, trg, sizeof(double) * 3 * (size_t)M);
    /* _refine_interface_for_targets (src/kernel/laplace3d_near.jl:99-150): stack, pop! from the end */
    double limit = DBL_MAX;
    for (int p = 0; p < f->P; ++p) { double l = panel_maxlen(f->corners + 12 * p); if (l < limit) limit = l; }
    long scap = 2 * (long)f->P + 64, sn = 0, ocap = 2 * (long)f->P + 64, on = 0;
    double *sc = (double *)malloc(sizeof(double) * 12 * scap);
    int *spar = (int *)malloc(sizeof(int) * scap), *ssp = (int *)malloc(sizeof(int) * scap);
    double *oc = (double *)malloc(sizeof(double) * 12 * ocap);
    int *opar = (int *)malloc(sizeof(int) * ocap), *osp = (int *)malloc(sizeof(int) * ocap);
    for (int p = 0; p < f->P; ++p) {
        memcpy(sc + 12 * sn, f->corners + 12 * p, sizeof(double) * 12);
        spar[sn] = p; ssp[sn] = 0; ++sn;
    }
    while (sn > 0) {
        --sn;
        double c[12];
        memcpy(c, sc + 12 * sn, sizeof(double) * 12);
        int par = spar[sn], spl = ssp[sn];
        double l = panel_maxlen(c), cc[3];
        panel_ctr(c, cc);
        double r = range_factor * l / nq;
        if (any_in_ball(M, trg, cc, r) && l > limit) {
            /* _split_panel4 (src/kernel/laplace3d_near.jl:78-97) */
            const double *a = c, *b = c + 3, *cq = c + 6, *d = c + 9;
            double ab[3], bc[3], cd[3], da[3], m[3];
            for (int k = 0; k < 3; ++k) {
                ab[k] = (a[k] + b[k]) / 2; bc[k] = (b[k] + cq[k]) / 2; cd[k] = (cq[k] + d[k]) / 2;
                da[k] = (d[k] + a[k]) / 2; m[k] = (a[k] + b[k] + cq[k] + d[k]) / 4;
            }
            const double *ch[4][4] = {{a, ab, m, da}, {ab, b, bc, m}, {m, bc, cq, cd}, {da, m, cd, d}};
            if (sn + 4 > scap) {
                scap *= 2;
                sc = (double *)realloc(sc, sizeof(double) * 12 * scap);
                spar = (int *)realloc(spar, sizeof(int) * scap);
                ssp = (int *)realloc(ssp, sizeof(int) * scap);
            }
            for (int q = 0; q < 4; ++q) {
                for (int v = 0; v < 4; ++v) memcpy(sc + 12 * sn + 3 * v, ch[q][v], sizeof(double) * 3);
                spar[sn] = par; ssp[sn] = 1; ++sn;
            }
        } else {
            if (on == ocap) {
                ocap *= 2;
                oc = (double *)realloc(oc, sizeof(double) * 12 * ocap);
                opar = (int *)realloc(opar, sizeof(int) * ocap);
                osp = (int *)realloc(osp, sizeof(int) * ocap);
            }
            memcpy(oc + 12 * on, c, sizeof(double) * 12);
            opar[on] = par; osp[on] = spl; ++on;
        }
    }
    int Pr = (int)on;
    double *nr = (double *)malloc(sizeof(double) * 3 * Pr), *ei = (double *)malloc(sizeof(double) * Pr);
    int *ie = (int *)malloc(sizeof(int) * Pr);
    for (int q = 0; q < Pr; ++q) {
        memcpy(nr + 3 * q, f->normal + 3 * opar[q], sizeof(double) * 3);
        ie[q] = f->is_edge[opar[q]];
        ei[q] = 1.0;
    }
    tm->ref = orc_iface_from_panels(Pr, nq, oc, nr, ie, ei, ei);
    tm->parent = opar; tm->from_split = osp; tm->Nr = tm->ref->N;
    free(nr); free(ei); free(ie); free(sc); free(spar); free(ssp); free(oc);
    const orc_iface *R = tm->ref;
    /* _refined_interface_prolongation (src/kernel/laplace3d_near.jl:152-225) */
    tm->p_ptr = (int *)malloc(sizeof(int) * ((size_t)tm->Nr + 1));
    tm->p_col = (int *)malloc(sizeof(int) * (size_t)tm->Nr * npp);
    tm->p_val = (double *)malloc(sizeof(double) * (size_t)tm->Nr * npp);
    int e = 0;
    tm->p_ptr[0] = 0;
    for (int q = 0; q < Pr; ++q) {
        int par = tm->parent[q], col0 = par * npp;
        if (!tm->from_split[q]) {
            for (int k = 0; k < npp; ++k) {
                tm->p_col[e] = col0 + k; tm->p_val[e] = 1.0; ++e;
                tm->p_ptr[q * npp + k + 1] = e;
            }
            continue;
        }
        const double *A = f->corners + 12 * par, *B = A + 3, *D = A + 9;
        double cc[3], bma[3], dma[3];
        panel_ctr(A, cc);
        for (int k = 0; k < 3; ++k) { bma[k] = B[k] - A[k]; dma[k] = D[k] - A[k]; }
        double bb = bma[0] * bma[0] + bma[1] * bma[1] + bma[2] * bma[2];
        double bd = bma[0] * dma[0] + bma[1] * dma[1] + bma[2] * dma[2];
        double dd = dma[0] * dma[0] + dma[1] * dma[1] + dma[2] * dma[2];
        double det = bb * dd - bd * bd, i11 = dd / det, i12 = -bd / det, i22 = bb / det;
        for (int k = 0; k < npp; ++k) {
            const double *pt = R->pt + 3 * ((size_t)q * npp + k);
            double x0 = pt[0] - cc[0], x1 = pt[1] - cc[1], x2 = pt[2] - cc[2];
            double r1 = bma[0] * x0 + bma[1] * x1 + bma[2] * x2, r2 = dma[0] * x0 + dma[1] * x1 + dma[2] * x2;
            double s = i11 * r1 + i12 * r2, t = i12 * r1 + i22 * r2;
            double rx[16], ry[16];
            bary_row(nq, f->gl_x, f->bary, 2 * s, rx);
            bary_row(nq, f->gl_x, f->bary, 2 * t, ry);
            for (int ii = 0; ii < nq; ++ii)
                for (int jj = 0; jj < nq; ++jj) { tm->p_col[e] = col0 + ii * nq + jj; tm->p_val[e] = rx[ii] * ry[jj]; ++e; }
            tm->p_ptr[q * npp + k + 1] = e;
        }
    }
    /* build_target_neighbor_list (src/kernel/laplace3d_near.jl:47-70, edges excluded) and
       laplace3d_pottrg_corrections_hcubature (src/kernel/laplace3d_near_hcubature.jl:193-251) */
    long tcap = 1024, tn = 0;
    orc_trip *tr = (orc_trip *)malloc(sizeof(orc_trip) * tcap);
    for (int q = 0; q < Pr; ++q) {
        if (R->is_edge[q]) continue;
        const double *c = R->corners + 12 * q;
        double cc[3];
        panel_ctr(c, cc);
        double r = range_factor * panel_maxlen(c) / nq;
        pgeom pg;
        panel_geom(R, q, &pg);
        for (int t = 0; t < M; ++t) {
            double dx = trg[3 * t] - cc[0], dy = trg[3 * t + 1] - cc[1], dz = trg[3 * t + 2] - cc[2];
            if (!(dx * dx + dy * dy + dz * dz <= r * r)) continue;
            pot_integrand pi = {pg.cc, pg.bma, pg.dma, trg + 3 * t, pg.scale, R};
            double I[GM_MAXV];
            if (!hcub2(&pi, npp, atol, 2000000L, I)) ++tm->n_capped;
            ++tm->n_pairs;
            for (int k = 0; k < npp; ++k) {
                int n = q * npp + k;
                double direct = R->wt[n] * k_pot(R->pt + 3 * n, trg + 3 * t);
                double v = I[k] - direct;
                if (v == 0.0) continue;
                if (tn == tcap) { tcap *= 2; tr = (orc_trip *)realloc(tr, sizeof(orc_trip) * tcap); }
                tr[tn].t = t; tr[tn].c = n; tr[tn].v = v; ++tn;
            }
        }
    }
    qsort(tr, (size_t)tn, sizeof(orc_trip), cmp_trip);
    tm->c_ptr = (long *)calloc((size_t)M + 1, sizeof(long));
    tm->c_col = (int *)malloc(sizeof(int) * (size_t)(tn > 0 ? tn : 1));
    tm->c_val = (double *)malloc(sizeof(double) * (size_t)(tn > 0 ? tn : 1));
    for (long q = 0; q < tn; ++q) { tm->c_ptr[tr[q].t + 1]++; tm->c_col[q] = tr[q].c; tm->c_val[q] = tr[q].v; }
    for (int t = 0; t < M; ++t) tm->c_ptr[t + 1] += tm->c_ptr[t];
    free(tr);
    return tm;
}

void orc_tmap_sizes(const orc_tmap *tm, int *Nr, int *Pr, long *nnz_c, long *n_pairs, long *n_capped) {
    *Nr = tm->Nr; *Pr = tm->ref->P; *nnz_c = tm->c_ptr[tm->M]; *n_pairs = tm->n_pairs; *n_capped = tm->n_capped;
}
/* refined nodes (x; y; z; w), parents, split flags, prolongation and corrections (CSR) */
void orc_tmap_get(const orc_tmap *tm, double *xyzw, int *parent, int *from_split, int *p_ptr, int *p_col,
                  double *p_val, long *c_ptr, int *c_col, double *c_val) {
    int Nr = tm->Nr;
    for (int i = 0; i < Nr; ++i) {
        xyzw[i] = tm->ref->pt[3 * i]; xyzw[Nr + i] = tm->ref->pt[3 * i + 1];
        xyzw[2 * Nr + i] = tm->ref->pt[3 * i + 2]; xyzw[3 * Nr + i] = tm->ref->wt[i];
    }
    memcpy(parent, tm->parent, sizeof(int) * tm->ref->P);
    memcpy(from_split, tm->from_split, sizeof(int) * tm->ref->P);
    memcpy(p_ptr, tm->p_ptr, sizeof(int) * ((size_t)Nr + 1));
    memcpy(p_col, tm->p_col, sizeof(int) * (size_t)tm->p_ptr[Nr]);
    memcpy(p_val, tm->p_val, sizeof(double) * (size_t)tm->p_ptr[Nr]);
    memcpy(c_ptr, tm->c_ptr, sizeof(long) * ((size_t)tm->M + 1));
    memcpy(c_col, tm->c_col, sizeof(int) * (size_t)tm->c_ptr[tm->M]);
    memcpy(c_val, tm->c_val, sizeof(double) * (size_t)tm->c_ptr[tm->M]);
}

/* the LinearMap of laplace3d_pottrg_fmm3d_corrected_hcubature applied to S (N x K):
   U = pot_base(P S) + C (P S), with the direct sum in place of the FMM */
void orc_tmap_apply(const orc_tmap *tm, int K, const double *S, double *U) {
    int Nr = tm->Nr, M = tm->M, N = tm->N;
    double *Xr = (double *)malloc(sizeof(double) * (size_t)Nr * K);
    for (int k = 0; k < K; ++k)
        for (int r = 0; r < Nr; ++r) {
            double acc = 0.0;
            for (int e = tm->p_ptr[r]; e < tm->p_ptr[r + 1]; ++e) acc += tm->p_val[e] * S[(size_t)k * N + tm->p_col[e]];
            Xr[(size_t)k * Nr + r] = acc;
        }
    orc_pottrg(tm->ref, M, tm->trg, K, Xr, U);
    for (int k = 0; k < K; ++k)
        for (int t = 0; t < M; ++t)
            for (long e = tm->c_ptr[t]; e < tm->c_ptr[t + 1]; ++e)
                U[(size_t)k * M + t] += tm->c_val[e] * Xr[(size_t)k * Nr + tm->c_col[e]];
    free(Xr);
}

/* four_index_matrix (src/solver/multi_rhs.jl:299-321), induced term only:
   R[a,b] = sum_{s in group a} q_s U[s,b] with U from orc_tmap_apply; R is G x K */
void orc_reaction(const orc_tmap *tm, int G, const long *grp, const double *q, int K, const double *S, double *R) {
    int M = tm->M;
    double *U = (double *)malloc(sizeof(double) * (size_t)M * K);
    orc_tmap_apply(tm, K, S, U);
    for (int b = 0; b < K; ++b)
        for (int a = 0; a < G; ++a) {
            double acc = 0.0;
            for (long s = grp[a]; s < grp[a + 1]; ++s) acc += q[s] * U[(size_t)b * M + s];
            R[(size_t)b * G + a] = acc;
        }
    free(U);
}

void orc_tmap_free(orc_tmap *tm) {
    if (!tm) return;
    orc_iface_free(tm->ref);
    free(tm->trg); free(tm->parent); free(tm->from_split); free(tm->p_ptr); free(tm->p_col); free(tm->p_val);
    free(tm->c_ptr); free(tm->c_col); free(tm->c_val); free(tm);
}

/* cubature of the panel potential basis for one (panel, target) pair: I[npp] */
int orc_pot_panel_hcub(const orc_iface *f, int p, const double *trg, double atol, double *I) {
    pgeom pg;
    panel_geom(f, p, &pg);
    pot_integrand pi = {pg.cc, pg.bma, pg.dma, trg, pg.scale, f};
    return hcub2(&pi, f->nq * f->nq, atol, 2000000L, I);
}
