// nearfield.cpp -- near-pair classification and correction blocks (host setup).
//
// Restates build_neighbor_list (src/kernel/laplace3d_near.jl:240-392) and the
// upsampled / adaptive correction rules (src/kernel/laplace3d_near_upsampling.jl:1-276,
// src/kernel/laplace3d_near_adaptive.jl:119-387), emitting the correction
// matrix directly in the block-sparse layout the GPU SpMM consumes: one dense
// nq^2 x nq^2 block per (source panel, target panel) pair, block rows indexed by
// the target panel.  Decisions (rho <= rho*, the p_up ceiling, the quadtree
// stopping test) depend on exact floating-point comparisons, so the scalar
// helpers keep the reference's operation order and this file is compiled with
// -ffp-contract=off.
#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <thread>

#include "internal.hpp"
#include "kernels.hpp"

namespace dbie {

namespace {

inline double len3(double a, double b, double c) { return std::sqrt(a * a + b * b + c * c); }

// laplace3d_grad (src/kernel/laplace3d.jl:9-15)
inline double grad_kernel(const double *s, const double *t, const double *n) {
    const double d0 = s[0] - t[0], d1 = s[1] - t[1], d2 = s[2] - t[2];
    const double r = std::sqrt(d0 * d0 + d1 * d1 + d2 * d2);
    const double ir = 1.0 / r;
    const double ir3 = ir * ir * ir;
    return (n[0] * (ir3 * (t[0] - s[0])) + n[1] * (ir3 * (t[1] - s[1])) + n[2] * (ir3 * (t[2] - s[2]))) / kFourPi;
}

// Julia Base.Math.hypot, FMA-corrected (correctly rounded) branch.
double hypot_cr(double x, double y) {
    if (std::isinf(x) || std::isinf(y)) return INFINITY;
    double ax = std::fabs(x), ay = std::fabs(y);
    if (ay > ax) std::swap(ax, ay);
    if (ay <= ax * std::sqrt(DBL_EPSILON / 2.0)) return ax;
    double scale = DBL_EPSILON * std::sqrt(DBL_MIN);
    if (ax > std::sqrt(DBL_MAX / 2.0)) {
        ax *= scale; ay *= scale; scale = 1.0 / scale;
    } else if (ay < std::sqrt(DBL_MIN)) {
        ax /= scale; ay /= scale;
    } else {
        scale = 1.0;
    }
    double h = std::sqrt(std::fma(ax, ax, ay * ay));
    const double h2 = h * h, ax2 = ax * ax;
    h -= (std::fma(-ay, ay, h2 - ax2) + std::fma(h, h, -h2) - std::fma(ax, ax, -ax2)) / (2.0 * h);
    return h * scale;
}

// Julia Base sqrt(::Complex{Float64}) (Kahan's scaled algorithm).
void complex_sqrt(double x, double y, double &re, double &im) {
    if (x == 0.0 && y == 0.0) { re = 0.0; im = y; return; }
    int k = 0;
    double rho = x * x + y * y;
    if (!std::isfinite(rho) && (std::isinf(x) || std::isinf(y))) {
        rho = INFINITY;
    } else if (std::isinf(rho) || (rho == 0.0 && (x != 0.0 || y != 0.0)) ||
               rho < 4.9406564584124654e-324 / (2.0 * DBL_EPSILON * DBL_EPSILON)) {
        const double m = std::max(std::fabs(x), std::fabs(y));
        k = m == 0.0 ? 0 : std::ilogb(m);
        const double xk = std::ldexp(x, -k), yk = std::ldexp(y, -k);
        rho = xk * xk + yk * yk;
    }
    if (std::isfinite(x)) rho = std::ldexp(std::fabs(x), -k) + std::sqrt(rho);
    if (k % 2 != 0) {
        k = (k - 1) / 2;
    } else {
        k = k / 2 - 1;
        rho += rho;
    }
    rho = std::ldexp(std::sqrt(rho), k);
    double xi = rho, eta = y;
    if (rho != 0.0) {
        if (std::isfinite(eta)) eta = (eta / rho) / 2.0;
        if (x < 0.0) { xi = std::fabs(eta); eta = std::copysign(rho, y); }
    }
    re = xi; im = eta;
}

// bernstein_rho_from_pole (src/utils/bernstein.jl:12-17)
double rho_pole(double zr, double zi) {
    double sr, si;
    complex_sqrt(zr * zr - zi * zi - 1.0, zr * zi + zi * zr, sr, si);
    return std::max(hypot_cr(zr + sr, zi + si), hypot_cr(zr - sr, zi - si));
}

// Panel frame for bernstein_rho_panel (src/utils/bernstein.jl:33-62)
struct Frame {
    double cc[3], e1[3], e2[3], n[3], L1, L2;
};
Frame make_frame(const Geometry &g, int64_t p) {
    Frame f;
    const double *A = &g.corners[12 * p], *B = A + 3, *C = A + 6, *D = A + 9;
    double u1[3], u2[3];
    for (int k = 0; k < 3; ++k) {
        f.cc[k] = (A[k] + B[k] + C[k] + D[k]) / 4;
        u1[k] = B[k] - A[k];
        u2[k] = D[k] - A[k];
        f.n[k] = g.normal[3 * p + k];
    }
    const double l1 = len3(u1[0], u1[1], u1[2]), l2 = len3(u2[0], u2[1], u2[2]);
    f.L1 = l1 / 2;
    f.L2 = l2 / 2;
    for (int k = 0; k < 3; ++k) { f.e1[k] = u1[k] / l1; f.e2[k] = u2[k] / l2; }
    return f;
}
double bernstein_rho(const Frame &f, const double *t) {
    const double r0 = t[0] - f.cc[0], r1 = t[1] - f.cc[1], r2 = t[2] - f.cc[2];
    const double xt = r0 * f.e1[0] + r1 * f.e1[1] + r2 * f.e1[2];
    const double yt = r0 * f.e2[0] + r1 * f.e2[1] + r2 * f.e2[2];
    const double zt = r0 * f.n[0] + r1 * f.n[1] + r2 * f.n[2];
    const double dy = yt - std::min(std::max(yt, -f.L2), f.L2);
    const double rho1 = rho_pole(xt / f.L1, std::sqrt(dy * dy + zt * zt) / f.L1);
    const double dx = xt - std::min(std::max(xt, -f.L1), f.L1);
    const double rho2 = rho_pole(yt / f.L2, std::sqrt(dx * dx + zt * zt) / f.L2);
    return std::min(rho1, rho2);
}

inline int threads_of(int n) { return n > 0 ? n : (int)std::max(1u, std::thread::hardware_concurrency()); }

}  // namespace

PairPrep pair_prep(const Geometry &g, int max_order, double atol, bool correct_edges) {
    const int64_t P = g.P;
    PairPrep pp;
    pp.max_order = max_order;
    pp.atol = atol;
    pp.correct_edges = correct_edges;
    pp.len.resize(P);
    pp.off.resize(P);
    pp.pdata.resize((size_t)P * kPairData);
    for (int64_t p = 0; p < P; ++p) {
        const double *c = &g.corners[12 * p];
        pp.len[p] = std::max(len3(c[0] - c[3], c[1] - c[4], c[2] - c[5]), len3(c[3] - c[6], c[4] - c[7], c[5] - c[8]));
        const double *n = &g.normal[3 * p];
        pp.off[p] = n[0] * c[0] + n[1] * c[1] + n[2] * c[2];
    }
    double Lmax = 1.0;
    for (double l : pp.len) Lmax = std::max(Lmax, l);
    const double tol_corner = std::sqrt(DBL_EPSILON) * Lmax;
    pp.tc2 = tol_corner * tol_corner;
    pp.rho_star = std::pow(atol, -1.0 / (2 * g.nq));
    const double As = (pp.rho_star + 1.0 / pp.rho_star) / 2, Bs = (pp.rho_star - 1.0 / pp.rho_star) / 2;
    const int npp = g.npp();
    for (int64_t p = 0; p < P; ++p) {
        const double *c = &g.corners[12 * p];
        double *d = &pp.pdata[(size_t)p * kPairData];
        const Frame f = make_frame(g, p);
        for (int k = 0; k < 3; ++k) {
            d[k] = (c[k] + c[3 + k] + c[6 + k] + c[9 + k]) / 4;  // centre
            d[5 + k] = f.cc[k];
            d[8 + k] = f.e1[k];
            d[11 + k] = f.e2[k];
            d[14 + k] = f.n[k];
        }
        d[17] = f.L1;
        d[18] = f.L2;
        const double ri = 105.0 / 100.0 * (pp.len[p] / 2) * std::sqrt(As * As + (1.0 + Bs) * (1.0 + Bs) + Bs * Bs);
        d[3] = ri * ri;
        // largest distance from the centre to a node of the panel (target-side pruning)
        double rad = 0.0;
        for (int t = 0; t < npp; ++t) {
            const int64_t n = p * npp + t;
            rad = std::max(rad, len3(g.x[n] - d[0], g.y[n] - d[1], g.z[n] - d[2]));
        }
        d[4] = rad;
        d[19] = ri;
    }
    return pp;
}

namespace {
// candidate scan of one source panel on the host: target panels with a node in the ball (with
// the smallest Bernstein radius over those nodes) or a corner within tol_corner, ascending
void scan_source(const Geometry &g, const PairPrep &pp, int64_t i, std::vector<double> &rho_min,
                 std::vector<char> &seen, std::vector<char> &touch, std::vector<int> &cand, PairCandidates &out) {
    const int64_t P = g.P, N = g.N;
    const int npp = g.npp();
    cand.clear();
    const double *d = &pp.pdata[(size_t)i * kPairData];
    const double ri2 = d[3];
    const Frame fr = make_frame(g, i);
    const double cx = d[0], cy = d[1], cz = d[2];
    // ball query over target nodes (inrange on the node KD-tree, :322)
    for (int64_t t = 0; t < N; ++t) {
        const int64_t j = t / npp;
        if (j == i) continue;
        const double d0 = g.x[t] - cx, d1 = g.y[t] - cy, d2 = g.z[t] - cz;
        if (d0 * d0 + d1 * d1 + d2 * d2 > ri2) continue;
        const double tp[3] = {g.x[t], g.y[t], g.z[t]};
        const double rho = bernstein_rho(fr, tp);
        if (!seen[j]) { seen[j] = 1; rho_min[j] = rho; cand.push_back((int)j); }
        else if (rho < rho_min[j]) rho_min[j] = rho;
    }
    // corner-touching panels (Phase B, :335-352)
    if (pp.correct_edges) {
        for (int64_t j = 0; j < P; ++j) {
            if (j == i) continue;
            bool hit = false;
            for (int a = 0; a < 4 && !hit; ++a)
                for (int b = 0; b < 4 && !hit; ++b) {
                    const double *p = &g.corners[12 * i + 3 * a], *q = &g.corners[12 * j + 3 * b];
                    const double d0 = q[0] - p[0], d1 = q[1] - p[1], d2 = q[2] - p[2];
                    hit = d0 * d0 + d1 * d1 + d2 * d2 <= pp.tc2;
                }
            if (hit) {
                touch[j] = 1;
                if (!seen[j]) { seen[j] = 1; rho_min[j] = INFINITY; cand.push_back((int)j); }
            }
        }
    }
    std::sort(cand.begin(), cand.end());
    for (int j : cand) {
        out.trg.push_back(j);
        out.rho.push_back(rho_min[j]);
        out.touch.push_back(touch[j]);
        seen[j] = 0;
        touch[j] = 0;
    }
}
}  // namespace

PairList finish_pair_list(const Geometry &g, const PairPrep &pp, const PairCandidates &c) {
    const int64_t P = g.P;
    const double same_tol = std::sqrt(DBL_EPSILON);
    const double log_atol = std::log(pp.atol);
    PairList pl;
    for (int64_t i = 0; i < P; ++i) {
        if (!pp.correct_edges && g.is_edge[i]) continue;
        const double li = pp.len[i];
        for (int64_t k = c.ptr[i]; k < c.ptr[i + 1]; ++k) {
            const int j = c.trg[k];
            if (!pp.correct_edges && g.is_edge[j]) continue;
            const double *ni = &g.normal[3 * i], *nj = &g.normal[3 * j];
            const double dn = ni[0] * nj[0] + ni[1] * nj[1] + ni[2] * nj[2];
            if (dn > 1 - same_tol && std::fabs(pp.off[i] - pp.off[j]) <= same_tol * std::max(1.0, li)) continue;
            if (pp.correct_edges && c.touch[k]) {
                pl.adp_src.push_back((int)i);
                pl.adp_trg.push_back(j);
                continue;
            }
            const double rm = c.rho[k];
            if (!(rm <= pp.rho_star)) continue;
            int p_up = rm > 1.0 ? (int)std::ceil(-log_atol / (2 * std::log(rm))) : pp.max_order;
            p_up = std::min(std::max(p_up, g.nq), pp.max_order);
            pl.up_src.push_back((int)i);
            pl.up_trg.push_back(j);
            pl.up_order.push_back(p_up);
        }
    }
    return pl;
}

PairList build_pair_list(const Geometry &g, int max_order, double atol, bool correct_edges, int n_threads) {
    const int64_t P = g.P;
    const PairPrep pp = pair_prep(g, max_order, atol, correct_edges);
    const int nt = threads_of(n_threads);
    std::vector<std::vector<double>> rho_min_t(nt, std::vector<double>(P));
    std::vector<std::vector<char>> seen_t(nt, std::vector<char>(P)), touch_t(nt, std::vector<char>(P));
    std::vector<std::vector<int>> cand_t(nt);
    std::vector<PairCandidates> per(P);
    parallel_for(P, nt, [&](int64_t i, int tid) {
        if (!correct_edges && g.is_edge[i]) return;
        scan_source(g, pp, i, rho_min_t[tid], seen_t[tid], touch_t[tid], cand_t[tid], per[i]);
    });
    PairCandidates c;
    c.ptr.assign(P + 1, 0);
    for (int64_t i = 0; i < P; ++i) c.ptr[i + 1] = c.ptr[i] + (int64_t)per[i].trg.size();
    for (int64_t i = 0; i < P; ++i) {
        c.trg.insert(c.trg.end(), per[i].trg.begin(), per[i].trg.end());
        c.rho.insert(c.rho.end(), per[i].rho.begin(), per[i].rho.end());
        c.touch.insert(c.touch.end(), per[i].touch.begin(), per[i].touch.end());
    }
    return finish_pair_list(g, pp, c);
}

namespace {

struct PanelMap {  // affine map of a source panel (cc + bma*u/2 + dma*v/2)
    double cc[3], bma[3], dma[3], scale;
};
PanelMap panel_map(const Geometry &g, int64_t p) {
    PanelMap m;
    const double *A = &g.corners[12 * p], *B = A + 3, *C = A + 6, *D = A + 9;
    for (int k = 0; k < 3; ++k) {
        m.cc[k] = (A[k] + B[k] + C[k] + D[k]) / 4;
        m.bma[k] = B[k] - A[k];
        m.dma[k] = D[k] - A[k];
    }
    m.scale = len3(m.bma[0], m.bma[1], m.bma[2]) * len3(m.dma[0], m.dma[1], m.dma[2]) / 4;
    return m;
}

inline void node(const Geometry &g, int64_t n, double *p) { p[0] = g.x[n]; p[1] = g.y[n]; p[2] = g.z[n]; }

// Kernel of a correction block between a point y of the source panel and a target x: mode 0 (:DT)
// laplace3d_grad(y, x, n_trg), mode 1 (:D) laplace3d_grad(x, y, n_src)
// (src/kernel/laplace3d_near_upsampling.jl:183-197, src/kernel/laplace3d_near_adaptive.jl:169-175)
inline double corr_kernel(int mode, const double *y, const double *x, const double *tn, const double *sn) {
    return mode == 1 ? grad_kernel(x, y, sn) : grad_kernel(y, x, tn);
}

// Quadtree of the touching-pair rule (walk_adaptive_tree, src/kernel/laplace3d_near_adaptive.jl:119-224)
struct Quadtree {
    const Geometry &g;
    PanelMap m;
    int nGL;
    double ns[32], ws[32];
    const double *trg, *tn;
    double atol, rtol;
    int max_depth;
    std::vector<double> leaves;  // (ulo, uhi, vlo, vhi) per leaf
    bool hit = false;
    const double *sn = nullptr;  // source panel normal (mode 1)
    int mode = 0;

    void moments(double ulo, double uhi, double vlo, double vhi, double *out) const {
        const int nq = g.nq, npp = nq * nq;
        double rx[16], ry[16];
        std::fill(out, out + npp, 0.0);
        const double hu = (uhi - ulo) / 2, hv = (vhi - vlo) / 2, mu = (uhi + ulo) / 2, mv = (vhi + vlo) / 2;
        const double sc = hu * hv * m.scale;
        for (int a = 0; a < nGL; ++a) {
            const double u = mu + hu * ns[a], wu = ws[a];
            barycentric_row(nq, g.glx, g.bary, u, rx);
            for (int b = 0; b < nGL; ++b) {
                const double v = mv + hv * ns[b], wv = ws[b];
                barycentric_row(nq, g.glx, g.bary, v, ry);
                double y[3];
                for (int k = 0; k < 3; ++k) y[k] = m.cc[k] + m.bma[k] * (u / 2) + m.dma[k] * (v / 2);
                const double wk = wu * wv * sc * corr_kernel(mode, y, trg, tn, sn);
                int q = 0;
                for (int mx = 0; mx < nq; ++mx)
                    for (int my = 0; my < nq; ++my, ++q) out[q] += wk * rx[mx] * ry[my];
            }
        }
    }
    void walk(double ulo, double uhi, double vlo, double vhi, int depth) {
        const int npp = g.nq * g.nq;
        double par[64], c1[64], c2[64], c3[64], c4[64];
        const double mu = (uhi + ulo) / 2, mv = (vhi + vlo) / 2;
        moments(ulo, uhi, vlo, vhi, par);
        moments(ulo, mu, vlo, mv, c1);
        moments(mu, uhi, vlo, mv, c2);
        moments(ulo, mu, mv, vhi, c3);
        moments(mu, uhi, mv, vhi, c4);
        double err = 0.0, pmax = 0.0;
        for (int q = 0; q < npp; ++q) {
            const double e = std::fabs(par[q] - (c1[q] + c2[q] + c3[q] + c4[q]));
            if (q == 0 || e > err) err = e;
            const double a = std::fabs(par[q]);
            if (q == 0 || a > pmax) pmax = a;
        }
        const double tl = std::max(atol, rtol * pmax);
        if (err < tl || depth == max_depth) {
            const double cells[16] = {ulo, mu, vlo, mv, mu, uhi, vlo, mv, ulo, mu, mv, vhi, mu, uhi, mv, vhi};
            leaves.insert(leaves.end(), cells, cells + 16);
            if (depth == max_depth && err >= tl) hit = true;
            return;
        }
        walk(ulo, mu, vlo, mv, depth + 1);
        walk(mu, uhi, vlo, mv, depth + 1);
        walk(ulo, mu, mv, vhi, depth + 1);
        walk(mu, uhi, mv, vhi, depth + 1);
    }
};

}  // namespace

namespace {
// Slots of the merged pair list in BSR order (block rows = target panels, ascending source panel).
struct PairItems {
    struct Item { int src, trg, order; };  // order 0 = touching (adaptive)
    std::vector<Item> items;               // sorted by (src, trg)
    std::vector<int64_t> slot, first;      // BSR slot per item; items of source panel i: [first[i], first[i+1])
    std::vector<int> row_ptr, col;
};
PairItems order_pairs(const Geometry &g, const PairList &pl) {
    const int64_t P = g.P;
    PairItems pi;
    auto &items = pi.items;
    items.reserve(pl.up_src.size() + pl.adp_src.size());
    for (size_t k = 0; k < pl.up_src.size(); ++k) items.push_back({pl.up_src[k], pl.up_trg[k], pl.up_order[k]});
    for (size_t k = 0; k < pl.adp_src.size(); ++k) items.push_back({pl.adp_src[k], pl.adp_trg[k], 0});
    std::sort(items.begin(), items.end(), [](const PairItems::Item &a, const PairItems::Item &b) {
        return a.src != b.src ? a.src < b.src : a.trg < b.trg;
    });
    pi.row_ptr.assign(P + 1, 0);
    for (const auto &it : items) pi.row_ptr[it.trg + 1]++;
    for (int64_t p = 0; p < P; ++p) pi.row_ptr[p + 1] += pi.row_ptr[p];
    const int64_t nb = (int64_t)items.size();
    pi.col.resize(nb);
    pi.slot.resize(nb);
    std::vector<int> fill(pi.row_ptr.begin(), pi.row_ptr.end() - 1);
    for (int64_t k = 0; k < nb; ++k) {
        pi.slot[k] = fill[items[k].trg]++;
        pi.col[pi.slot[k]] = items[k].src;
    }
    pi.first.assign(P + 1, 0);
    for (const auto &it : items) pi.first[it.src + 1]++;
    for (int64_t p = 0; p < P; ++p) pi.first[p + 1] += pi.first[p];
    return pi;
}

// Source cache of panel i at order n_up (build_source_cache, src/kernel/laplace3d_near_upsampling.jl:8-50):
// upos (n_up^2 x 3) and Mt (bd x n_up^2, row = moment).
void source_cache(const Geometry &g, int64_t i, int n_up, double *upos, double *Mt) {
    const int nq = g.nq, bd = nq * nq;
    const PanelMap m = panel_map(g, i);
    double ns[16], ws[16], Ex[16][16];
    gauss_legendre(n_up, ns, ws);
    for (int a = 0; a < n_up; ++a) barycentric_row(nq, g.glx, g.bary, ns[a], Ex[a]);
    const int nu2 = n_up * n_up;
    for (int a = 0; a < n_up; ++a)
        for (int b = 0; b < n_up; ++b) {
            const int al = a * n_up + b;
            for (int k = 0; k < 3; ++k) upos[3 * al + k] = m.cc[k] + m.bma[k] * (ns[a] / 2) + m.dma[k] * (ns[b] / 2);
        }
    for (int mx = 0; mx < nq; ++mx)
        for (int my = 0; my < nq; ++my)
            for (int a = 0; a < n_up; ++a)
                for (int b = 0; b < n_up; ++b)
                    Mt[(size_t)(mx * nq + my) * nu2 + a * n_up + b] = m.scale * ws[a] * ws[b] * Ex[a][mx] * Ex[b][my];
}

// Touching pair: the adaptive block moments of source panel i at the targets of panel j, driven by the
// representative target (walk_adaptive_tree + adaptive_block_moments_inplace!,
// src/kernel/laplace3d_near_adaptive.jl:119-224,308-387).  blk: bd x bd row-major (K_up, no direct term).
bool adaptive_block(const Geometry &g, int64_t i, int64_t j, const NearOpts &o, double *blk) {
    const int nq = g.nq, bd = nq * nq;
    const int nGL = o.adaptive_n_GL == 0 ? nq : o.adaptive_n_GL;
    const PanelMap m = panel_map(g, i);
    const double *tn = &g.normal[3 * j], *sn = &g.normal[3 * i];
    double tpts[64][3];
    for (int t = 0; t < bd; ++t) node(g, j * bd + t, tpts[t]);
    Quadtree qt{g, m, nGL, {}, {}, nullptr, tn, o.adaptive_atol, o.adaptive_rtol, o.adaptive_max_depth, {}, false,
                sn, o.mode};
    gauss_legendre(nGL, qt.ns, qt.ws);
    const double *c1 = &g.corners[12 * i];
    int rep = 0;
    double best = INFINITY;
    for (int t = 0; t < bd; ++t) {
        const double d = len3(tpts[t][0] - c1[0], tpts[t][1] - c1[1], tpts[t][2] - c1[2]);
        if (d < best) { best = d; rep = t; }
    }
    qt.trg = tpts[rep];
    qt.walk(-1.0, 1.0, -1.0, 1.0, 0);
    std::fill(blk, blk + bd * bd, 0.0);
    double rx[16], ry[16], kt[64];
    for (size_t l = 0; l < qt.leaves.size(); l += 4) {
        const double ulo = qt.leaves[l], uhi = qt.leaves[l + 1], vlo = qt.leaves[l + 2], vhi = qt.leaves[l + 3];
        const double hu = (uhi - ulo) / 2, hv = (vhi - vlo) / 2, mu = (uhi + ulo) / 2, mv = (vhi + vlo) / 2;
        const double sc = hu * hv * m.scale;
        for (int a = 0; a < nGL; ++a) {
            const double u = mu + hu * qt.ns[a], wu = qt.ws[a];
            barycentric_row(nq, g.glx, g.bary, u, rx);
            for (int b = 0; b < nGL; ++b) {
                const double v = mv + hv * qt.ns[b], wv = qt.ws[b];
                barycentric_row(nq, g.glx, g.bary, v, ry);
                double y[3];
                for (int kk = 0; kk < 3; ++kk) y[kk] = m.cc[kk] + m.bma[kk] * (u / 2) + m.dma[kk] * (v / 2);
                const double wuv = wu * wv * sc;
                for (int t = 0; t < bd; ++t) kt[t] = corr_kernel(o.mode, y, tpts[t], tn, sn);
                for (int mx = 0; mx < nq; ++mx)
                    for (int my = 0; my < nq; ++my) {
                        const int c = mx * nq + my;
                        const double wm = wuv * rx[mx] * ry[my];
                        for (int t = 0; t < bd; ++t) blk[t * bd + c] += wm * kt[t];
                    }
            }
        }
    }
    return qt.hit;
}
}  // namespace

bool touching_block(const Geometry &g, int64_t i, int64_t j, const NearOpts &o, double *blk) {
    return adaptive_block(g, i, j, o, blk);
}

std::vector<double> quadtree_rules(const Geometry &g, int nGL) {
    std::vector<double> r(2 * g.nq + 2 * nGL);
    std::copy(g.glx, g.glx + g.nq, r.begin());
    std::copy(g.bary, g.bary + g.nq, r.begin() + g.nq);
    gauss_legendre(nGL, &r[2 * g.nq], &r[2 * g.nq + nGL]);
    return r;
}

CorrectionPlan plan_corrections(const Geometry &g, const PairList &pl, const NearOpts &o, bool host_quadtrees) {
    const int64_t P = g.P;
    const int bd = g.npp();
    PairItems pi = order_pairs(g, pl);
    CorrectionPlan cp;
    cp.bd = bd;
    cp.row_ptr = std::move(pi.row_ptr);
    cp.col = std::move(pi.col);
    cp.cache_off.assign(P, -1);
    cp.cache_nu.assign(P, 0);
    // cache layout and the item lists
    int64_t off = 0;
    for (int64_t i = 0; i < P; ++i) {
        int n_up = 0;
        for (int64_t k = pi.first[i]; k < pi.first[i + 1]; ++k) n_up = std::max(n_up, pi.items[k].order);
        if (n_up > 0) {
            cp.cache_off[i] = (int)off;
            cp.cache_nu[i] = n_up;
            off += (int64_t)n_up * n_up * (3 + bd);
        }
        for (int64_t k = pi.first[i]; k < pi.first[i + 1]; ++k) {
            const auto &it = pi.items[k];
            if (it.order > 0) {
                cp.up_src.push_back(it.src);
                cp.up_trg.push_back(it.trg);
                cp.up_slot.push_back((int)pi.slot[k]);
            } else {
                cp.adp_src.push_back(it.src);
                cp.adp_trg.push_back(it.trg);
                cp.adp_slot.push_back((int)pi.slot[k]);
            }
        }
    }
    if (off > INT32_MAX) fail(DBIE_BAD_ARG, "source caches too large");
    cp.caches.resize((size_t)off);
    parallel_for(P, o.n_threads, [&](int64_t i, int) {
        if (cp.cache_nu[i] == 0) return;
        const int nu = cp.cache_nu[i];
        double *c = &cp.caches[(size_t)cp.cache_off[i]];
        source_cache(g, i, nu, c, c + 3 * nu * nu);
    });
    if (!host_quadtrees) {
        // representative target: the target node nearest the source panel's first corner
        const size_t na = cp.adp_src.size();
        cp.adp_rep.resize(na);
        cp.adp_map.resize(na * 10);
        for (size_t k = 0; k < na; ++k) {
            const int i = cp.adp_src[k], j = cp.adp_trg[k];
            const PanelMap m = panel_map(g, i);
            double *mp = &cp.adp_map[k * 10];
            for (int c = 0; c < 3; ++c) { mp[c] = m.cc[c]; mp[3 + c] = m.bma[c]; mp[6 + c] = m.dma[c]; }
            mp[9] = m.scale;
            const double *c1 = &g.corners[12 * (size_t)i];
            int rep = 0;
            double best = INFINITY;
            for (int t = 0; t < bd; ++t) {
                const int64_t n = (int64_t)j * bd + t;
                const double d = len3(g.x[n] - c1[0], g.y[n] - c1[1], g.z[n] - c1[2]);
                if (d < best) { best = d; rep = t; }
            }
            cp.adp_rep[k] = rep;
        }
        return cp;
    }
    cp.adp_kup.resize(cp.adp_src.size() * (size_t)bd * bd);
    std::atomic<int> hits{0};
    parallel_for((int64_t)cp.adp_src.size(), o.n_threads, [&](int64_t k, int) {
        if (adaptive_block(g, cp.adp_src[k], cp.adp_trg[k], o, &cp.adp_kup[(size_t)k * bd * bd])) hits.fetch_add(1);
    });
    cp.adaptive_hits = hits.load();
    return cp;
}

BlockSparse build_corrections(const Geometry &g, const PairList &pl, const NearOpts &o) {
    const int64_t P = g.P;
    const int nq = g.nq, bd = nq * nq;
    BlockSparse bs;
    bs.bd = bd;
    // Merge both pair kinds; slot of a pair within its target block row follows ascending source panel.
    struct Item { int src, trg, order; };  // order 0 = touching (adaptive)
    std::vector<Item> items;
    items.reserve(pl.up_src.size() + pl.adp_src.size());
    for (size_t k = 0; k < pl.up_src.size(); ++k) items.push_back({pl.up_src[k], pl.up_trg[k], pl.up_order[k]});
    for (size_t k = 0; k < pl.adp_src.size(); ++k) items.push_back({pl.adp_src[k], pl.adp_trg[k], 0});
    std::sort(items.begin(), items.end(), [](const Item &a, const Item &b) {
        return a.src != b.src ? a.src < b.src : a.trg < b.trg;
    });
    bs.row_ptr.assign(P + 1, 0);
    for (const Item &it : items) bs.row_ptr[it.trg + 1]++;
    for (int64_t p = 0; p < P; ++p) bs.row_ptr[p + 1] += bs.row_ptr[p];
    const int64_t nb = (int64_t)items.size();
    bs.col.resize(nb);
    bs.val.assign((size_t)nb * bd * bd, 0.0);
    std::vector<int64_t> slot(nb);
    {
        std::vector<int> fill(bs.row_ptr.begin(), bs.row_ptr.end() - 1);
        for (int64_t k = 0; k < nb; ++k) {
            slot[k] = fill[items[k].trg]++;
            bs.col[slot[k]] = items[k].src;
        }
    }
    // items of one source panel are contiguous
    std::vector<int64_t> first(P + 1, 0);
    for (const Item &it : items) first[it.src + 1]++;
    for (int64_t p = 0; p < P; ++p) first[p + 1] += first[p];

    std::atomic<int> hits{0};
    const int nGL = o.adaptive_n_GL == 0 ? nq : o.adaptive_n_GL;
    parallel_for(P, o.n_threads, [&](int64_t i, int) {
        if (first[i] == first[i + 1]) return;
        const PanelMap m = panel_map(g, i);
        double src_pts[64][3];
        for (int c = 0; c < bd; ++c) node(g, i * bd + c, src_pts[c]);
        // Source cache at the per-source-panel maximum order (build_source_cache, :8-50)
        int n_up = 0;
        for (int64_t k = first[i]; k < first[i + 1]; ++k) n_up = std::max(n_up, items[k].order);
        std::vector<double> upos, Mt;  // upos: n_up^2 x 3; Mt: bd x n_up^2 (row = moment)
        if (n_up > 0) {
            double ns[16], ws[16], Ex[16][16];
            gauss_legendre(n_up, ns, ws);
            for (int a = 0; a < n_up; ++a) barycentric_row(nq, g.glx, g.bary, ns[a], Ex[a]);
            const int nu2 = n_up * n_up;
            upos.resize(3 * nu2);
            Mt.resize((size_t)bd * nu2);
            for (int a = 0; a < n_up; ++a)
                for (int b = 0; b < n_up; ++b) {
                    const int al = a * n_up + b;
                    for (int k = 0; k < 3; ++k)
                        upos[3 * al + k] = m.cc[k] + m.bma[k] * (ns[a] / 2) + m.dma[k] * (ns[b] / 2);
                }
            for (int mx = 0; mx < nq; ++mx)
                for (int my = 0; my < nq; ++my)
                    for (int a = 0; a < n_up; ++a)
                        for (int b = 0; b < n_up; ++b)
                            Mt[(size_t)(mx * nq + my) * nu2 + a * n_up + b] =
                                m.scale * ws[a] * ws[b] * Ex[a][mx] * Ex[b][my];
        }
        std::vector<double> kv, blk(bd * bd);
        for (int64_t k = first[i]; k < first[i + 1]; ++k) {
            const int j = items[k].trg;
            const double *tn = &g.normal[3 * j], *sn = &g.normal[3 * i];
            double tpts[64][3];
            for (int t = 0; t < bd; ++t) node(g, (int64_t)j * bd + t, tpts[t]);
            if (items[k].order > 0) {
                const int nu2 = n_up * n_up;
                kv.resize((size_t)nu2);
                for (int t = 0; t < bd; ++t) {
                    for (int al = 0; al < nu2; ++al) kv[al] = corr_kernel(o.mode, &upos[3 * al], tpts[t], tn, sn);
                    for (int c = 0; c < bd; ++c) {
                        const double *mrow = &Mt[(size_t)c * nu2];
                        double s = 0.0;
                        for (int al = 0; al < nu2; ++al) s += mrow[al] * kv[al];
                        blk[t * bd + c] = s;
                    }
                }
            } else {
                Quadtree qt{g, m, nGL, {}, {}, nullptr, tn, o.adaptive_atol, o.adaptive_rtol, o.adaptive_max_depth, {},
                            false, sn, o.mode};
                gauss_legendre(nGL, qt.ns, qt.ws);
                const double *c1 = &g.corners[12 * i];
                int rep = 0;
                double best = INFINITY;
                for (int t = 0; t < bd; ++t) {
                    const double d = len3(tpts[t][0] - c1[0], tpts[t][1] - c1[1], tpts[t][2] - c1[2]);
                    if (d < best) { best = d; rep = t; }
                }
                qt.trg = tpts[rep];
                qt.walk(-1.0, 1.0, -1.0, 1.0, 0);
                if (qt.hit) hits.fetch_add(1);
                // adaptive_block_moments_inplace! (:308-387)
                std::fill(blk.begin(), blk.end(), 0.0);
                double rx[16], ry[16], kt[64];
                for (size_t l = 0; l < qt.leaves.size(); l += 4) {
                    const double ulo = qt.leaves[l], uhi = qt.leaves[l + 1], vlo = qt.leaves[l + 2], vhi = qt.leaves[l + 3];
                    const double hu = (uhi - ulo) / 2, hv = (vhi - vlo) / 2, mu = (uhi + ulo) / 2, mv = (vhi + vlo) / 2;
                    const double sc = hu * hv * m.scale;
                    for (int a = 0; a < nGL; ++a) {
                        const double u = mu + hu * qt.ns[a], wu = qt.ws[a];
                        barycentric_row(nq, g.glx, g.bary, u, rx);
                        for (int b = 0; b < nGL; ++b) {
                            const double v = mv + hv * qt.ns[b], wv = qt.ws[b];
                            barycentric_row(nq, g.glx, g.bary, v, ry);
                            double y[3];
                            for (int kk = 0; kk < 3; ++kk) y[kk] = m.cc[kk] + m.bma[kk] * (u / 2) + m.dma[kk] * (v / 2);
                            const double wuv = wu * wv * sc;
                            for (int t = 0; t < bd; ++t) kt[t] = corr_kernel(o.mode, y, tpts[t], tn, sn);
                            for (int mx = 0; mx < nq; ++mx)
                                for (int my = 0; my < nq; ++my) {
                                    const int c = mx * nq + my;
                                    const double wm = wuv * rx[mx] * ry[my];
                                    for (int t = 0; t < bd; ++t) blk[t * bd + c] += wm * kt[t];
                                }
                        }
                    }
                }
            }
            // subtract the direct panel block (laplace3d_DT_panel / laplace3d_D_panel,
            // src/kernel/laplace3d_near.jl:25-45)
            double *dst = &bs.val[(size_t)slot[k] * bd * bd];
            for (int t = 0; t < bd; ++t)
                for (int c = 0; c < bd; ++c)
                    dst[t * bd + c] =
                        blk[t * bd + c] - corr_kernel(o.mode, src_pts[c], tpts[t], tn, sn) * g.w[i * bd + c];
        }
    });
    bs.adaptive_hits = hits.load();
    return bs;
}

}  // namespace dbie
